"""Per-BFS summary of rocprofv3 CSV output for the plain BFS engine.

A BFS run is the span of kernels from one k_bfs_init to the k_pred_gather /
k_write_path that ends it.  The kernel-trace pass gives time per kernel per
run; the --pmc passes give FETCH_SIZE and WRITE_SIZE per kernel per run, each
scaled by what the same counter reads for a device-to-device copy of known
size (pass the calibration runs' CSV files).

For the batched engine a run is the span from one k_ms_init to the next;
`batched` groups consecutive runs (e.g. "1,5" = skip one, summarise five) and
reports the median time per level and kernel.

`times` also reports the device's idle time per run (the run's span minus
the union of its kernels' intervals on any queue): the time the GPU waits
for the host between levels.  `hitrate` reads one --pmc pass of
TCC_HIT_sum and TCC_MISS_sum (two of the four TCC counters one pass may
hold) and gives the L2 hit rate per kernel.

usage:
  python rocprof_summary.py times  KT_KERNEL_TRACE.csv [LAST_N | FIRST:END]
  python rocprof_summary.py traffic FETCH.csv WRITE.csv COPY_FETCH.csv COPY_WRITE.csv COPY_BYTES
  python rocprof_summary.py hitrate TCC.csv
  python rocprof_summary.py batched KT_KERNEL_TRACE.csv GROUPS
  python rocprof_summary.py batched-traffic FETCH.csv WRITE.csv FETCH_SCALE WRITE_SCALE GROUPS
"""
import collections
import csv
import statistics
import sys


def short(name):
    return name.replace("(anonymous namespace)::", "").split("(")[0].split("::")[-1]


def _end_of_run(names):
    return max(i for i, n in enumerate(names)
               if "k_pred_gather" in n or "k_pred_fixup" in n or "k_write_path" in n)


def runs(rows, name_of):
    """Split dispatch-ordered rows into BFS runs (drops the trailing segment)."""
    idx = [i for i, r in enumerate(rows) if "k_bfs_init" in name_of(r)]
    out = []
    for a, b in zip(idx, idx[1:]):
        seg = rows[a:b]
        out.append(seg[: _end_of_run([name_of(r) for r in seg]) + 1])
    return out


def _idle_us(seg):
    """Span of the run minus the union of its kernels' intervals."""
    iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in seg)
    busy, cur_a, cur_b = 0, iv[0][0], iv[0][1]
    gaps = []
    for a, b in iv[1:]:
        if a > cur_b:
            busy += cur_b - cur_a
            gaps.append((a - cur_b) / 1e3)
            cur_a, cur_b = a, b
        else:
            cur_b = max(cur_b, b)
    busy += cur_b - cur_a
    return (iv[-1][1] - iv[0][0] - busy) / 1e3, gaps


def times(path, last=0, window=None):
    """window = (first, end): summarise only BFS runs [first, end) of the trace."""
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    spans, per, idle, ngaps, launches = [], collections.defaultdict(list), [], [], \
        collections.defaultdict(list)
    segs = runs(rows, lambda r: r["Kernel_Name"])
    if window:
        print(f"BFS runs {window[0]}..{window[1] - 1} of the {len(segs)} in the trace")
        segs = segs[window[0]:window[1]]
    for seg in segs:
        spans.append((int(seg[-1]["End_Timestamp"]) - int(seg[0]["Start_Timestamp"])) / 1e3)
        c, n = collections.Counter(), collections.Counter()
        for r in seg:
            c[short(r["Kernel_Name"])] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            n[short(r["Kernel_Name"])] += 1
        for k in set(per) | set(c):
            per[k].append(c.get(k, 0.0))
            launches[k].append(n.get(k, 0))
        i, g = _idle_us(seg)
        idle.append(i)
        ngaps.append(sum(1 for x in g if x > 2.0))
    print(f"BFS runs: {len(spans)}; first-kernel-start to last-kernel-end per run: "
          f"mean {statistics.mean(spans):.1f} us, median {statistics.median(spans):.1f} us")
    if last:
        tail = spans[-last:]
        print(f"the last {len(tail)} runs (the timed ones): mean {statistics.mean(tail):.1f} us, "
              f"median {statistics.median(tail):.1f} us")
    print(f"device idle per run (no kernel on any queue): mean {statistics.mean(idle):.1f} us "
          f"in {statistics.mean(ngaps):.1f} gaps over 2 us")
    print(f"kernel-busy time per run: mean {sum(statistics.mean(v) for v in per.values()):.1f} us")
    for k, v in sorted(per.items(), key=lambda kv: -statistics.mean(kv[1])):
        nl = sum(launches[k])
        per_launch = sum(v) / nl if nl else 0.0
        print(f"  {k:26s} {statistics.mean(v):9.1f} us/run  {per_launch:9.1f} us/launch "
              f"({nl / len(v):.2f} launches/run)")


def _counter(path, name):
    out = []
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == name:
            out.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    return sorted(out)


def traffic(fetch, write, cfetch, cwrite, copy_bytes):
    kib = copy_bytes / 1024.0
    cf = statistics.mean(c[2] for c in _counter(cfetch, "FETCH_SIZE") if "copyBuffer" in c[1])
    cw = statistics.mean(c[2] for c in _counter(cwrite, "WRITE_SIZE") if "copyBuffer" in c[1])
    ff, fw = kib / cf, kib / cw
    print(f"calibration: a {copy_bytes} B copy reads FETCH_SIZE {cf:.1f} KiB (x{ff:.3f}), "
          f"WRITE_SIZE {cw:.1f} KiB (x{fw:.3f})")
    res = {}
    for label, path, cname, f in (("fetch", fetch, "FETCH_SIZE", ff), ("write", write, "WRITE_SIZE", fw)):
        tot, byk, one = [], collections.defaultdict(list), collections.defaultdict(list)
        for seg in runs(_counter(path, cname), lambda c: c[1]):
            t = collections.Counter()
            for c in seg:
                t[short(c[1])] += c[2] * 1024.0 * f
                one[short(c[1])].append(c[2] * 1024.0 * f)
            tot.append(sum(t.values()))
            for k in set(byk) | set(t):
                byk[k].append(t.get(k, 0.0))
        res[label] = (tot, byk, one)
    ft, fk, f1 = res["fetch"]
    wt, wk, w1 = res["write"]
    f_med, w_med = statistics.median(ft), statistics.median(wt)
    print(f"per run HBM bytes (median over {len(ft)} runs): fetch {f_med/1e9:.3f} GB, write {w_med/1e9:.3f} GB")
    print(f"traffic_bytes_per_bfs: {int(f_med + w_med)}")
    keys = sorted(set(f1) | set(w1), key=lambda k: -(statistics.median(fk.get(k, [0])) + statistics.median(wk.get(k, [0]))))
    print("per launch (median over launches):")
    for k in keys:
        print(f"  {k:26s} fetch {statistics.median(f1.get(k, [0]))/1e6:9.1f} MB  write "
              f"{statistics.median(w1.get(k, [0]))/1e6:9.1f} MB  per launch")
    print("per run (median over runs):")
    for k in keys:
        print(f"  run {k:22s} fetch {statistics.median(fk.get(k, [0]))/1e6:9.1f} MB  write "
              f"{statistics.median(wk.get(k, [0]))/1e6:9.1f} MB")


def hitrate(path):
    hit = collections.defaultdict(float)
    miss = collections.defaultdict(float)
    for name, d in (("TCC_HIT_sum", hit), ("TCC_MISS_sum", miss)):
        for _, k, v in _counter(path, name):
            d[short(k)] += v
    tot_h, tot_m = sum(hit.values()), sum(miss.values())
    print(f"all kernels: L2 hit rate {tot_h / max(tot_h + tot_m, 1):.3f} "
          f"({tot_h:.3e} hits, {tot_m:.3e} misses)")
    for k in sorted(set(hit) | set(miss), key=lambda k: -(hit[k] + miss[k])):
        h, m = hit[k], miss[k]
        print(f"  {k:26s} hit rate {h / max(h + m, 1):.3f}  requests {h + m:.3e}")


def batched(path, groups):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    rows = [r for r in rows if "k_ms_" in r["Kernel_Name"]]
    idx = [i for i, r in enumerate(rows) if "k_ms_init" in r["Kernel_Name"]] + [len(rows)]
    runs_ = [rows[a:b] for a, b in zip(idx, idx[1:])]
    pos = 0
    spec = [int(x) for x in groups.split(",")]
    for gi in range(0, len(spec), 2):
        skip, take = spec[gi], spec[gi + 1]
        sel = runs_[pos + skip:pos + skip + take]
        pos += skip + take
        spans = [(int(r[-1]["End_Timestamp"]) - int(r[0]["Start_Timestamp"])) / 1e3 for r in sel]
        per = collections.defaultdict(list)
        for run in sel:
            level, c = -1, collections.Counter()
            for r in run:
                k = short(r["Kernel_Name"])
                if k.startswith("k_ms_prep"):
                    level += 1
                tag = f"L{level}" if level >= 0 and not k.startswith(("k_ms_output", "k_ms_te", "k_ms_isolated")) else "out"
                c[(tag, k)] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            for key in set(per) | set(c):
                per[key].append(c.get(key, 0.0))
        print(f"runs {pos - take}..{pos - 1}: batch span median {statistics.median(spans):.1f} us")
        for (tag, k), v in sorted(per.items(), key=lambda kv: (kv[0][0] != "out", kv[0][0])):
            m = statistics.median(v)
            if m >= 20:
                print(f"  {tag:4s} {k:28s} {m:9.1f} us")


def batched_traffic(fetch, write, fscale, wscale, groups):
    res = {}
    for label, path, cname, f in (("fetch", fetch, "FETCH_SIZE", fscale),
                                  ("write", write, "WRITE_SIZE", wscale)):
        c = [x for x in _counter(path, cname) if "k_ms_" in x[1]]
        idx = [i for i, x in enumerate(c) if "k_ms_init" in x[1]] + [len(c)]
        res[label] = [sum(x[2] * 1024.0 * f for x in c[a:b]) for a, b in zip(idx, idx[1:])]
    spec = [int(x) for x in groups.split(",")]
    pos = 0
    for gi in range(0, len(spec), 2):
        skip, take = spec[gi], spec[gi + 1]
        fs = res["fetch"][pos + skip:pos + skip + take]
        ws = res["write"][pos + skip:pos + skip + take]
        pos += skip + take
        print(f"runs {pos - take}..{pos - 1}: fetch {statistics.median(fs)/1e9:.2f} GB, "
              f"write {statistics.median(ws)/1e9:.2f} GB per batch (median)")


if __name__ == "__main__":
    if sys.argv[1] == "times":
        arg = sys.argv[3] if len(sys.argv) > 3 else "0"
        if ":" in arg:
            a, b = arg.split(":")
            times(sys.argv[2], 0, (int(a), int(b)))
        else:
            times(sys.argv[2], int(arg))
    elif sys.argv[1] == "batched":
        batched(sys.argv[2], sys.argv[3])
    elif sys.argv[1] == "batched-traffic":
        batched_traffic(sys.argv[2], sys.argv[3], float(sys.argv[4]), float(sys.argv[5]),
                        sys.argv[6])
    elif sys.argv[1] == "hitrate":
        hitrate(sys.argv[2])
    else:
        traffic(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5], int(sys.argv[6]))
