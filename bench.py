"""Benchmark: particle-steps/s of the force-and-integrate loop on one MI355X.

Workload: the Kremer-Grest melt (2000 chains x 50 = 100,000 beads, density
0.85, periodic, LJ rc=2.5 shifted, FENE 30/1.5, DPD gamma=1 at kT=1,
dt=0.01), synthetic seeded input generated in-process.  A step is one pass of
md_main.f90:69-155: integrate, skin test (+ rebuild when due), force pass,
integrate.  `value` is the mixed FP32/FP64 form with the state resident in
HBM; the plain FP64 form, the mixed-vs-FP64 force error and the oracle's
single-thread CPU rate on a bounded sample are reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mfa_amd import configs, lib  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Input preparation (untimed): the seeded melt (overlaps down to 0.8 sigma) is
# relaxed on the GPU by the FORCE_SWITCH_ON ramp under WCA at dt = 0.002
# (configs.equilibrate); the warm-up and timed steps then run the production
# potential (rc = 2.5, dt = 0.01).
# Algorithmic HBM bytes per particle of the force pass: read r (24) and
# v (24) and the type (4), write F (24).
FORCE_BYTES_PER_PARTICLE = 76
STEP_BYTES_PER_PARTICLE = 196  # force pass 76 + fused integrate pass 120
# HBM bytes per force pass of the mixed form on this workload, from the
# FETCH_SIZE and WRITE_SIZE counter runs (profiles/force_pass_melt100k_hbm_traffic.md),
# each scaled by the same counter's reading for a copy of known size.
FORCE_TRAFFIC_BYTES = 79.433e6 + 2.443e6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(p, st0, precision, steps, warmup, first_step=1, graph_steps=16):
    """Warm-up, then the timed region: `steps` steps in one mfa_run_steps call
    (blocks of 16 steps captured in a hipGraph), state resident in HBM.  Then
    the force pass's average kernel duration at the final state, from a graph
    of 200 back-to-back launches between two HIP events on the same stream."""
    ctx = lib.Context(p, precision)
    ctx.set_graph_steps(graph_steps)
    ctx.set_state(st0["r"], st0["v"], st0["a"], st0["t"], st0["mic"])
    ctx.run_steps(first_step, warmup, want_obs=False)
    t0 = time.perf_counter()
    obs = ctx.run_steps(first_step + warmup, steps, want_obs=True)
    dt = time.perf_counter() - t0
    force_ms = ctx.time_force(200)
    st = ctx.get_state()
    ctx.close()
    return dt, obs, {"force_ms": force_ms}, st


def force_error_vs_fp64(p, st):
    """Mixed-form forces vs FP64 forces on the same state."""
    out = {}
    for prec in (lib.MFA_FP64, lib.MFA_MIXED):
        ctx = lib.Context(p, prec)
        ctx.set_state(st["r"], st["v"], st["a"], st["t"], st["mic"])
        ctx.energies(step=0)
        out[prec] = ctx.get_state()["f"]
        ctx.close()
    ref, f = out[lib.MFA_FP64], out[lib.MFA_MIXED]
    nref = np.linalg.norm(ref, axis=1)
    frms = np.sqrt(np.mean(nref ** 2))
    d = np.linalg.norm(f - ref, axis=1)
    return float((d / np.maximum(nref, 1e-3 * frms)).max()), float(np.sqrt(np.mean(d ** 2)) / frms)


def cpu_baseline(inputs, st0, n_steps):
    from oracle import oracle as orc
    s = orc.System(inputs)
    r = st0["r"]
    s.set_state(r, st0["v"], st0["a"], st0["t"], st0["mic"])
    s.run_steps(1, 2)  # bins the list and warms the caches
    t0 = time.perf_counter()
    s.run_steps(3, n_steps)
    dt = time.perf_counter() - t0
    return r.shape[0] * n_steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph-steps", type=int, default=16,
                    help="steps per captured hipGraph block (0: direct launches)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(each GPU box has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    w = configs.WORKLOADS["melt100k"]
    inputs = w["make"]()
    t0 = time.perf_counter()
    p, st0 = configs.equilibrate(inputs, w["kind"], min_dist=w["min_dist"])
    t = st0["t"]
    n = p.n_part
    log("melt100k: n=%d box=%.4f cells=%s, prepared in %.1f s" % (
        n, p.boundary[0], list(p.n_bin), time.perf_counter() - t0))

    dt_m, obs_m, tm_m, st_m = run_form(p, st0, lib.MFA_MIXED, args.steps, args.warmup,
                                       graph_steps=args.graph_steps)
    log("mixed: %.4f s for %d steps, force pass %.4f ms avg, rebuilds %d, T=%.3f" % (
        dt_m, args.steps, tm_m["force_ms"], obs_m["n_rebuilds"], 2.0 * obs_m["t_kin"] / (3.0 * n)))
    dt_d, obs_d, tm_d, _ = run_form(p, st0, lib.MFA_FP64, args.steps, args.warmup,
                                    graph_steps=args.graph_steps)
    log("fp64 : %.4f s for %d steps, force %.4f ms avg" % (dt_d, args.steps, tm_d["force_ms"]))

    st_m["t"] = t
    err_max, err_rms = force_error_vs_fp64(p, st_m)
    log("mixed vs fp64 force error: max %.3e rms %.3e" % (err_max, err_rms))

    value = n * args.steps / dt_m
    value_fp64 = n * args.steps / dt_d
    force_s = tm_m["force_ms"] * 1e-3
    achieved = FORCE_BYTES_PER_PARTICLE * n / force_s / 1e9 if force_s > 0 else None

    cpu = None
    if not args.no_cpu:
        rate, secs = cpu_baseline(inputs, st0, args.cpu_steps)
        cpu = {"value": rate, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": "oracle (plain-C FP64 restatement, 1 thread) on the same 100k-bead "
                         "melt, %d steps after 2 warm-up steps (%.1f s)" % (args.cpu_steps, secs)}
        log("cpu baseline: %.3e particle-steps/s (%.1f s)" % (rate, secs))

    line = {
        "metric": "particle-steps/s on 1 MI355X, FP32/FP64 mixed vs plain FP64 at <=1e-5 force err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_m / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair math, int64 fixed-point sums, fp64 state",
        "data": "synthetic (seeded random-walk melt, Maxwell-Boltzmann velocities, relaxed "
                "by the switch-on ramp before timing)",
        "config": {"workload": "melt100k", "n_particles": n, "chains": 2000, "chain_length": 50,
                   "density": 0.85, "lj_rc": 2.5, "dt": 0.01, "dpd_gamma": 1.0, "skin": 0.3,
                   "parallelism": "single GPU"},
        "fp64_value": value_fp64,
        "mixed_over_fp64": value / value_fp64,
        "force_err_max": err_max,
        "force_err_rms": err_rms,
        "rebuilds": obs_m["n_rebuilds"],
        "step_roofline_frac": value * STEP_BYTES_PER_PARTICLE / (HBM_PEAK_GBS * 1e9),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_force_mixed<8,0>",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": FORCE_TRAFFIC_BYTES,
            "force_ms_avg": tm_m["force_ms"],
            "force_ms_avg_fp64": tm_d["force_ms"],
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
