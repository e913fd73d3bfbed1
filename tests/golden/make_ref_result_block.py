"""Extracts a known-answer fixture for the Graph500 result block from a log of
the reference driver (mpi/main.cc:632-702, mpi/logfile.h:54-133).

Input: the stdout of the reference `runnable 16` at one rank (a log of one
such run is ref_s16.log).  Output: per-root BFS times (max precision
line), validate times, traversed edges, and the result block values.
usage: python make_ref_result_block.py ref_s16.log > ref_s16_result_block.json
"""
import json
import re
import sys

text = open(sys.argv[1]).read()
times = [float(m) for m in re.findall(r"Time for BFS \d+ is [0-9.]+ \(([0-9.e+-]+)\)", text)]
vtimes = [float(m) for m in re.findall(r"Validate time for BFS \d+ is ([0-9.]+)", text)]
nedge = [int(m) for m in re.findall(r"Number of traversed edges is (\d+)", text)]
block = {}
for k, v in re.findall(r"^([a-z_A-Z]+):\s+([-0-9.e+]+)\s*$", text.split("============= Result ==============")[1], re.M):
    block[k] = float(v)
json.dump({"_provenance": "reference driver stdout, SCALE 16, 1 rank, 16 roots",
           "bfs_times": times, "validate_times": vtimes, "nedge": nedge,
           "result_block": block}, sys.stdout, indent=1)
