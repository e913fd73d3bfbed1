"""Writes ref_autotune_traces.json: what the reference's own CPU driver
prints about every root's direction switches, run in this container.

The driver (oracle/_ref/ref_runnable, built from the reference's
mpi/main.cc and generator/splittable_mrg.c by `make -C oracle ref-bfs`) is
run as `mpiexec -n 1 ref_runnable SCALE -A [-C]`.  With -A it first draws the
roots by its rule (find_roots + redraw, mpi/main.cc:334-359), runs every root
once at alpha 10 / beta 14 and prints, per root index, the T2B level and the
frontier edges at the switch (mpi/main.cc:60-68: auto_tuning_data slots
T2B_LEVEL, GLOBAL_NQ_EDGES, NUM_GLOBAL_EDGES, PRE_GLOBAL_NQ_EDGES); after
tuning alpha it prints the B2T level and frontier sizes (mpi/main.cc:69-78)
measured at the tuned alpha, which it also prints.  Those lines, the tuned
alpha and each root's traversed-edge count are kept; timings are not (they
depend on the machine).

usage: python tests/golden/make_ref_traces.py   (needs /root/reference, MPICH)
"""
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_runnable")
MPIEXEC = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")
CASES = [(16, False), (18, False), (20, False), (16, True), (18, True)]

ALPHA = re.compile(r"\[(\d+)\] Perf\. = \S+ MTEPS, T2B = (-?\d+), (-?\d+) > (-?\d+)/\S+ \((-?\d+)\)")
BETA = re.compile(r"\[(\d+)\] Perf\. = \S+ MTEPS, B2T = (-?\d+), (-?\d+) < (-?\d+)/\(\S+ \* (\d+) \* 2\) : \((-?\d+)\)")


def run(scale, corebfs):
    cmd = [MPIEXEC, "-n", "1", EXE, str(scale), "-A"] + (["-C"] if corebfs else [])
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"))
    # the driver's log lines go to stderr, the result block to stdout
    p = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=1800,
                       cwd="/tmp", check=True)
    out = p.stderr + p.stdout
    block, alpha_rows, beta_rows = None, [], []
    tuned_alpha = None
    te = []
    for line in out.splitlines():
        if "START AUTO TUNING FOR ALPHA" in line:
            block = "alpha"
        elif "START AUTO TUNING FOR BETA" in line:
            block = "beta"
        elif "Alpha is determined to be" in line:
            tuned_alpha = line.split("determined to be")[1].strip()
        elif "Number of traversed edges is" in line:
            te.append(int(line.split()[-1]))
        m = ALPHA.search(line)
        if m and block == "alpha":
            alpha_rows.append([int(x) for x in m.groups()])
        m = BETA.search(line)
        if m and block == "beta":
            beta_rows.append([int(x) for x in m.groups()])
    # only the first ALPHA / BETA block of each kind (later rounds follow timing)
    n = len(te)
    return {"scale": scale, "corebfs": corebfs, "nroots": n,
            "alpha_block": alpha_rows[:n],
            "alpha_block_fields": ["index", "t2b_level", "global_nq_edges",
                                   "num_global_edges", "pre_global_nq_edges"],
            "tuned_alpha": tuned_alpha,
            "beta_block": beta_rows[:n],
            "beta_block_fields": ["index", "b2t_level", "global_nq_size", "num_global_verts",
                                  "edgefactor", "pre_global_nq_size"],
            "traversed_edges": te}


def main():
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} missing: make -C oracle ref-bfs")
    cases = []
    for scale, cb in CASES:
        print(f"SCALE {scale}{' -C' if cb else ''}", flush=True)
        cases.append(run(scale, cb))
    with open(os.path.join(HERE, "ref_autotune_traces.json"), "w") as f:
        json.dump({"source": "reference CPU driver (mpi/main.cc) built from its own sources, "
                             "mpiexec -n 1 ref_runnable SCALE -A [-C], MPICH 3.3.2",
                   "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
