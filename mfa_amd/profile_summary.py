"""Summaries of rocprofv3 csv output for the hot kernels.

--trace run_kernel_trace.csv : per-kernel average durations in the timed
    window of a bench.py run (the launches after the last mixed warm-up
    block); rebuild kernels are split into no-op launches and real rebuilds.
--pmc run_counter_collection.csv : FETCH_SIZE / WRITE_SIZE per dispatch of the
    force pass, scaled by the reading of the calibration copy of known size.
--counters run_counter_collection.csv : median of every counter per dispatch
    of each kernel.
--totals run_kernel_trace.csv : every kernel of the whole run by total time
    (launches, average and total duration), for runs without a timed window
    of the melt bench (the replica batch, the charged brush).
"""
import argparse
import csv
import re
import statistics
from collections import defaultdict

COPY_BYTES = 256 * 1024 * 1024
# kernel names of the mixed force pass without energies (any width, extra-terms
# flag and queue choice): MODE 0 (forces alone) and MODE 2 (the pass of a step,
# which also integrates)
MIXED_FORCE = re.compile(r"void k_force_mixed<\d+, [02]")
MIXED_STEP = re.compile(r"void k_force_mixed<\d+, 2")
FP64_FORCE = re.compile(r"void k_force64<\d+")


def _find(kern, pattern):
    for k, v in kern.items():
        if pattern.match(k):
            return v
    return None


def trace(path, n_timed):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    names = [r["Kernel_Name"] for r in rows]
    try:
        end = next(i for i, n in enumerate(names) if FP64_FORCE.match(n))
    except StopIteration:
        end = len(rows)
    idx = [i for i in range(end) if MIXED_STEP.match(names[i])] or \
        [i for i in range(end) if MIXED_FORCE.match(names[i])]
    seg = rows[idx[-n_timed]:end]
    d = defaultdict(list)
    for r in seg:
        d[r["Kernel_Name"].split("(")[0]].append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = []
    for k, v in sorted(d.items(), key=lambda kv: -sum(kv[1])):
        busy = [x for x in v if x > 8.0]
        out.append("%-28s launches %5d  avg %8.2f us  median %8.2f us  (launches > 8 us: %d, avg %.2f us)"
                   % (k, len(v), sum(v) / len(v), statistics.median(v), len(busy),
                      sum(busy) / len(busy) if busy else 0.0))
    t0 = int(seg[0]["Start_Timestamp"])
    t1 = int(seg[-1]["End_Timestamp"])
    busy = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in seg)
    out.append("window %.2f ms, kernels busy %.2f ms (%.0f%%)" % ((t1 - t0) / 1e6, busy / 1e6,
                                                                 100.0 * busy / (t1 - t0)))
    return "\n".join(out)


def pmc(path):
    rows = list(csv.DictReader(open(path)))
    by = defaultdict(lambda: defaultdict(list))
    for r in rows:
        by[r["Counter_Name"]][r["Kernel_Name"].split("(")[0]].append(float(r["Counter_Value"]))
    out = []
    for counter, kern in by.items():
        copy = [v for k, v in kern.items() if "copybuffer" in k.lower()]
        force = _find(kern, MIXED_FORCE) or _find(kern, FP64_FORCE) or []
        if not copy or not force:
            out.append("%s: calibration copy or force pass not found (%s)" % (counter, list(kern)))
            continue
        cal = max(max(v) for v in copy)  # KiB read or written by the 256 MiB copy (2nd dispatch)
        scale = COPY_BYTES / (cal * 1024.0)
        for k, v in sorted(kern.items()):
            if not (MIXED_FORCE.match(k) or FP64_FORCE.match(k)):
                continue
            f = statistics.median(v)
            out.append("%s: %s %.1f KiB/dispatch (median of %d), copy %.1f KiB for %d B "
                       "-> scale %.4f -> %.3f MB per dispatch"
                       % (counter, k, f, len(v), cal, COPY_BYTES, scale, f * 1024.0 * scale / 1e6))
    return "\n".join(out)


def counters(path):
    rows = list(csv.DictReader(open(path)))
    by = defaultdict(lambda: defaultdict(list))
    for r in rows:
        by[r["Kernel_Name"].split("(")[0]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = []
    for kern, cs in sorted(by.items()):
        n = max(len(v) for v in cs.values())
        out.append("%s (%d dispatches)" % (kern, n))
        for cname, v in sorted(cs.items()):
            out.append("    %-24s median %16.1f" % (cname, statistics.median(v)))
    return "\n".join(out)


def totals(path, top=14):
    rows = list(csv.DictReader(open(path)))
    d = defaultdict(list)
    for r in rows:
        d[r["Kernel_Name"].split("(")[0]].append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    tot = sum(sum(v) for v in d.values())
    out = []
    for k, v in sorted(d.items(), key=lambda kv: -sum(kv[1]))[:top]:
        out.append("%-34s launches %7d  avg %9.2f us  total %9.2f ms (%4.1f%%)"
                   % (k[:34], len(v), sum(v) / len(v), sum(v) / 1e3, 100.0 * sum(v) / tot))
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace")
    ap.add_argument("--timed", type=int, default=2400)
    ap.add_argument("--pmc")
    ap.add_argument("--counters")
    ap.add_argument("--totals")
    a = ap.parse_args()
    if a.trace:
        print(trace(a.trace, a.timed))
    if a.pmc:
        print(pmc(a.pmc))
    if a.counters:
        print(counters(a.counters))
    if a.totals:
        print(totals(a.totals))


if __name__ == "__main__":
    main()
