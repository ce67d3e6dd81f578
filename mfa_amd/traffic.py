"""Workload for the counter runs of the force pass and the list build.

Builds the seeded 100k-bead melt (density 0.85, the production cut-off, so
the neighbour structure of the timed runs), bins it `--rebuilds` times (each
a cell fill and a list build), runs the force pass `--reps` times with direct
launches, then a plain copy of a known size (the calibration for the
counters).  With --ewald it runs the k-space Ewald stage of the seeded
pe_brush27k start instead.  No trajectory is run, which keeps the number
of dispatches small for the counter runs.  Meant to run under rocprofv3 with one counter per run:

    MFA_GRAPH_STEPS=0 rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o run \
        -- python3 -m mfa_amd.traffic
    MFA_GRAPH_STEPS=0 rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT -o run \
        -- python3 -m mfa_amd.traffic

then `python3 -m mfa_amd.profile_summary --pmc OUT/run_counter_collection.csv`.
"""
import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

COPY_BYTES = 256 * 1024 * 1024  # bytes read and bytes written by the calibration copy


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--rebuilds", type=int, default=1)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--ewald", action="store_true",
                    help="run the k-space Ewald stage of the seeded pe_brush27k start --reps "
                         "times instead of the melt force pass")
    args = ap.parse_args()
    from mfa_amd import configs, lib
    if args.ewald:
        import numpy as np
        w = configs.WORKLOADS["pe_brush27k"]
        p, r, v, t = configs.build(w["make"](), w["kind"], min_dist=w["min_dist"])
        with lib.Context(p, lib.MFA_MIXED) as c:
            c.set_graph_steps(0)
            c.set_state(r, v, np.zeros_like(r), t)
            for _ in range(args.reps):
                c.ewald_k(1)
        print("ewald_k: %d calls" % args.reps, flush=True)
        return
    w = configs.WORKLOADS["melt100k"]
    import numpy as np
    p, r, v, t = configs.build(w["make"](), w["kind"], min_dist=w["min_dist"])
    prec = lib.MFA_MIXED if args.precision == "mixed" else lib.MFA_FP64
    with lib.Context(p, prec) as c:
        c.set_graph_steps(0)
        c.set_state(r, v, np.zeros_like(r), t)
        for _ in range(args.rebuilds):
            c.binning()
        ms = c.time_force(args.reps)
    cms = lib.copy_probe(COPY_BYTES)
    print("force pass %.4f ms; calibration copy of %d bytes in %.4f ms (%.0f GB/s read+write)"
          % (ms, COPY_BYTES, cms, 2 * COPY_BYTES / cms / 1e6), flush=True)


if __name__ == "__main__":
    main()
