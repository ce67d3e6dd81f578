"""Writes tests/golden/dpd_brush_stats.json: per-run statistics of the
reference's own DPD brush runs (oracle/_ref/ref_md_brush, built from the
reference's Fortran sources by oracle/ref/build_ref.sh) at kT = 1 with its
ziggurat noise, on the seeded inputs of tests/dpd_stats.case(r),
r = 0 .. RUNS-1.

    python3 tests/golden/make_dpd_stats.py

The fixture holds each run's time averages and its input's checksum; the
tests regenerate the inputs, check the checksums, run the oracle or the HIP
path with their own counter-based noise and compare (tests/dpd_stats.check).
"""
import json
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mfa_amd import inputs as io  # noqa: E402
from tests import dpd_stats as ds  # noqa: E402

RUNS = 16


def one_run(run):
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_md_brush")
    x, p, r, v, t = ds.case(run)
    n = p.n_part
    with tempfile.TemporaryDirectory() as d:
        io.write_run_dir(d, x, t, r, v, np.zeros_like(r))
        t2, r2, v2, _ = io.read_conf(os.path.join(d, "conf_old"), n)
        steps = ds.N_EQ + ds.N_RUN
        res = subprocess.run([exe, "0", str(steps), "snap", str(ds.EVERY)], cwd=d,
                             capture_output=True, text=True)
        if res.returncode:
            raise RuntimeError(res.stdout[-2000:] + res.stderr[-2000:])
        raw = np.fromfile(os.path.join(d, "ref_snap.bin"), dtype=np.float64)
    k = steps // ds.EVERY
    raw = raw.reshape(k, 2, n, 3)
    snaps = [(raw[q, 0], raw[q, 1]) for q in range(ds.N_EQ // ds.EVERY, k)]
    out = ds.reduce_run(snaps, t2, list(p.boundary), p.n_mon, p.n_chain)
    out["input_checksum"] = ds.checksum(r2, v2)
    return out


def main():
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        runs = list(ex.map(one_run, range(RUNS)))
    out = os.path.join(ROOT, "tests", "golden", "dpd_brush_stats.json")
    with open(out, "w") as f:
        json.dump({"source": "oracle/_ref/ref_md_brush (reference Fortran), kT 1, ziggurat noise",
                   "runs": runs}, f, indent=0)
    t = [r["temperature"][0] for r in runs]
    print("wrote", out, "T = %.4f +- %.4f over %d runs" % (np.mean(t), np.std(t, ddof=1) /
                                                        np.sqrt(len(t)), len(t)))


if __name__ == "__main__":
    main()
