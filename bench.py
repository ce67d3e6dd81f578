"""Benchmark: particle-steps/s of the force-and-integrate loop on one MI355X.

Workload: the Kremer-Grest melt (2000 chains x 50 = 100,000 beads, density
0.85, periodic, LJ rc=2.5 shifted, FENE 30/1.5, DPD gamma=1 at kT=1,
dt=0.01), synthetic seeded input generated in-process.  A step is one pass of
md_main.f90:69-155: integrate, skin test (+ rebuild when due), force pass,
integrate.  `value` is the mixed FP32/FP64 form with the state resident in
HBM; the plain FP64 form, the mixed-vs-FP64 force error and the oracle's
single-thread CPU rate on a bounded sample are reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
    python bench.py --workload NAME --save-state S.npz     (then, under rocprofv3:)
    python bench.py --workload NAME --counters --state S.npz [--reps R]

Other workloads (one JSON line each, same keys): brush9k (DPD brush, 9,000
beads, dt 0.005), pe_brush27k (polyelectrolyte brush with Ewald, 27,000
beads) and batch256 (256 uneven brush replicas, ~6M beads, one launch per
kernel per step, against the same replicas run one after another).
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
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
# Input preparation (untimed): the seeded melt (overlaps down to 0.8 sigma) is
# relaxed on the GPU by the FORCE_SWITCH_ON ramp under WCA at dt = 0.002
# (configs.equilibrate); the warm-up and timed steps then run the production
# potential (rc = 2.5, dt = 0.01).
# Algorithmic HBM bytes per particle of a force pass alone: read r (24) and
# v (24) and the type (4), write F (24).
FORCE_BYTES_PER_PARTICLE = 76
# ... and of a whole step, which in the mixed form is one force pass that also
# integrates (verlet_velocities of the step, verlet_positions of the next):
# read r, v (48) and the type (4), write r and v (48).  The acceleration is
# formed and used in registers.  Everything else it moves (the FP32 view, the
# list, the skin reference, the rebuild) is this implementation's own.
STEP_BYTES_PER_PARTICLE = 100
# Counter figures of the mixed step's force pass (fused with the integration)
# on this workload, per (threads per particle, pair queue, cell grid) as the
# autotuners chose them, measured by `bench.py --counters` under rocprofv3 on the state
# the timed steps leave (profiles/force_pass_melt100k_hbm_traffic.md,
# profiles/force_and_list_melt100k_sq_counters.md): HBM bytes per launch
# (FETCH_SIZE and WRITE_SIZE, each scaled by its reading for a copy of known
# size) and VALU wave-instructions per launch (SQ_INSTS_VALU).  The issue
# peak: 1024 SIMDs, one wave64 VALU instruction per 4 cycles at 2.4 GHz.
FORCE_COUNTERS = {(4, 2, (15, 15, 15)): (96.917e6 + 9.603e6, 17107282)}
VALU_PEAK_GINSTS = 1024 * 2.4 / 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(p, st0, precision, steps, warmup, first_step=1, graph_steps=16, tune_skin=True):
    """Warm-up (the autotuners' choices of the force pass's width and list
    walk, of the cell width and, where the workload leaves it open, of the
    list skin, then the width and walk again on that grid; then `warmup`
    steps), then
    the timed region: `steps` steps in one mfa_run_steps call
    (blocks of 16 steps captured in a hipGraph), state resident in HBM.  Then
    the average kernel duration of the force pass at the final state, alone
    and as the step runs it (in the mixed form fused with the integration),
    each from a graph of 200 back-to-back launches between two HIP events on
    the same stream."""
    ctx = lib.Context(p, precision)
    ctx.set_graph_steps(graph_steps)
    ctx.set_state(st0["r"], st0["v"], st0["a"], st0["t"], st0["mic"])
    # untimed: the autotuner picks the threads per particle of the force pass
    tpp, _ = ctx.autotune(first_step, 200)
    first_step += autotune_trials(p, precision) * (max(graph_steps, 0) + 1 + 200)
    # untimed: the cell-size autotuner picks the cell width and, where the
    # workload leaves it open, the list skin (each pair over 200 whole steps)
    skins = list(configs.SKIN_CANDIDATES) if tune_skin else [p.skin]
    width, skin, _, first_step = ctx.autotune_cells(first_step, 200, configs.cell_width_candidates(p),
                                                    skins)
    # untimed: the width and list walk again, on the chosen grid and skin,
    # over longer trials (their rebuilds vary)
    tpp, _ = ctx.autotune(first_step, 500)
    first_step += autotune_trials(p, precision) * (max(graph_steps, 0) + 1 + 500)
    ctx.run_steps(first_step, warmup, want_obs=False)
    tuning = ctx.tuning()
    t0 = time.perf_counter()
    obs = ctx.run_steps(first_step + warmup, steps, want_obs=True)
    dt = time.perf_counter() - t0
    force_ms = ctx.time_force(200)
    step_force_ms = ctx.time_force_step(200)
    ewald_ms = ctx.time_ewald(200) if p.coulomb else None
    ewald_parts = ctx.time_ewald_parts(50) if p.coulomb else None
    rebuild_idle = ctx.time_rebuild(20, idle=True)
    st = ctx.get_state()
    rebuild = ctx.time_rebuild(20)  # resets the skin reference: after get_state
    ctx.close()
    return dt, obs, {"force_ms": force_ms, "step_force_ms": step_force_ms, "tpp": tpp, "skin": skin,
                     "cell_width": width, "n_bin": list(lib.with_cells(p, width).n_bin),
                     "pair_queue": tuning["pair_queue"], "half_list": tuning["half_list"],
                     "ewald_ms": ewald_ms, "ewald_parts_ms": ewald_parts,
                     "rebuild_ms": rebuild, "rebuild_idle_ms": rebuild_idle}, st


def kernel_xt(p):
    """The extra-terms bits of the mixed force kernel a system runs: 1 charges
    or bending, 2 the Langevin force, 4 DPD_CUT_OFF."""
    return ((1 if (p.coulomb or p.bending) else 0) | (2 if p.thermostat == 1 else 0) |
            (4 if p.dpd_cut_off_on else 0))


def autotune_trials(p, precision):
    """Trials mfa_autotune runs, two sweeps over its candidates (it advances
    the state by graph_steps + 1 + n steps per trial): widths 1, 2, 4, 8, 16 directly; in the mixed
    form also with the skin part queued and, for DPD_CUT_OFF, everything
    queued (systems without charges, bending or the Langevin force), and the
    half-list form at 4, 8 and 16 lanes (for DPD_CUT_OFF also queued)."""
    if precision != lib.MFA_MIXED:
        return 2 * 5
    qok = not p.coulomb and not p.bending and p.thermostat != 1
    dc = bool(p.dpd_cut_off_on)
    return 2 * (5 + (5 if qok else 0) + (5 if qok and dc else 0) + 3 + (3 if qok and dc else 0))


def _fits(p, skin):
    try:
        lib.with_skin(p, skin)
        return True
    except lib.MfaError:
        return False


def ewald_k_count(p, inputs):
    """Number of k vectors of the half space (ewald_k.f90:104-124): |k| <=
    (dk_x + dk_y) kmax / 2, without k = 0, the lower triangle of the plane
    kz = 0 and half of its diagonal; slab period 3 Lz with walls."""
    pi_f = float(np.float32(3.1415926536))
    lz = p.boundary[2] * (1.0 if p.symmetry == 1 else 3.0)
    dk = (2 * pi_f / p.boundary[0], 2 * pi_f / p.boundary[1], 2 * pi_f / lz)
    kc = (dk[0] + dk[1]) * inputs.n_kmax / 2.0
    m = [int(kc / d) for d in dk]
    nk = 0
    for z in range(m[2] + 1):
        for y in range(-m[1], m[1] + 1):
            for x in range(-m[0], m[0] + 1):
                if (x * dk[0]) ** 2 + (y * dk[1]) ** 2 + (z * dk[2]) ** 2 > kc * kc:
                    continue
                if z == 0 and (x == 0 and y == 0 or x + y < 0 or (-x == y and y > 0)):
                    continue
                nk += 1
    return nk


def ewald_roofline(p, inputs, types, ewald_ms):
    """FP64 roofline of the k-space stage: 24 flops per (k vector, charge),
    8 for its term of rho(k) and 16 for its two terms of the force sums."""
    nq = int(sum(1 for a in types if inputs.charge[int(a) - 1] != 0.0))
    nk = ewald_k_count(p, inputs)
    flops = 24.0 * nq * nk
    achieved = flops / (ewald_ms * 1e-3) / 1e12
    return {"bound": "fp64_valu", "kernels": "k_ewald_tables, k_ewald_rho, k_ewald_sum, "
            "k_ewald_force_rows, k_ewald_apply", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "flops_per_step": flops,
            "stage_ms_avg": ewald_ms, "charges": nq, "k_vectors": nk}


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


def host_cpu():
    """Model name and logical CPU count of the host running the bench."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


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


def cpu_rate(inputs, st0, n_steps):
    rate, secs = cpu_baseline(inputs, st0, n_steps)
    log("cpu baseline: %.3e particle-steps/s (%.1f s)" % (rate, secs))
    return rate, secs


def contract_line(args, workload, value, ms_per_step, dtype, data, config, extra):
    line = {
        "metric": "particle-steps/s on 1 MI355X, FP32/FP64 mixed vs plain FP64 at <=1e-5 force err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": data,
        "config": dict({"workload": workload}, **config),
    }
    line.update(extra)
    return line


def roofline(n, kernel_ms, traffic=None, kernel="k_force_mixed<8,0,0>",
             bytes_per_particle=FORCE_BYTES_PER_PARTICLE):
    s = kernel_ms * 1e-3
    achieved = bytes_per_particle * n / s / 1e9 if s > 0 else None
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": traffic, "kernel_ms_avg": kernel_ms,
            "algorithmic_bytes_per_particle": bytes_per_particle}


def bench_single(args, name):
    """melt100k, brush9k, pe_brush27k: mixed and FP64 forms on one system."""
    w = configs.WORKLOADS[name]
    inputs = w["make"]()
    t0 = time.perf_counter()
    if w["kind"] == 2:
        p, st0 = configs.equilibrate(inputs, w["kind"], min_dist=w["min_dist"])
    else:
        p, st0 = configs.prepare_brush(inputs, w["kind"], min_dist=w["min_dist"])
    t = st0["t"]
    n = p.n_part
    log("%s: n=%d box=%s cells=%s, prepared in %.1f s" % (
        name, n, [round(x, 4) for x in p.boundary], list(p.n_bin), time.perf_counter() - t0))
    tune = not w.get("skin_fixed", False)
    dt_m, obs_m, tm_m, st_m = run_form(p, st0, lib.MFA_MIXED, args.steps, args.warmup,
                                       graph_steps=args.graph_steps, tune_skin=tune)
    log("mixed: %.4f s for %d steps, force pass %.4f ms avg (%d threads per particle), "
        "rebuilds %d, T=%.3f" % (dt_m, args.steps, tm_m["force_ms"], tm_m["tpp"],
                                 obs_m["n_rebuilds"], 2.0 * obs_m["t_kin"] / (3.0 * n)))
    log("rebuild kernels (ms): %s" % tm_m["rebuild_ms"])
    dt_d, obs_d, tm_d, _ = run_form(p, st0, lib.MFA_FP64, args.steps, args.warmup,
                                    graph_steps=args.graph_steps, tune_skin=tune)
    log("fp64 : %.4f s for %d steps, force %.4f ms avg (%d threads per particle)" % (
        dt_d, args.steps, tm_d["force_ms"], tm_d["tpp"]))
    st_m["t"] = t
    if args.save_state:  # the state the timed steps left, for the counter runs
        np.savez(args.save_state, r=st_m["r"], v=st_m["v"], a=st_m["a"], mic=st_m["mic"], t=t,
                 tpp=tm_m["tpp"], skin=tm_m["skin"], cell_width=tm_m["cell_width"],
                 pair_queue=int(tm_m["pair_queue"]),
                 half_list=int(tm_m["half_list"]))
        log("saved the timed state to %s" % args.save_state)
    err_max, err_rms = force_error_vs_fp64(p, st_m)
    log("mixed vs fp64 force error: max %.3e rms %.3e" % (err_max, err_rms))
    value = n * args.steps / dt_m
    value_fp64 = n * args.steps / dt_d
    cpu = None
    if not args.no_cpu:
        rate, secs = cpu_rate(inputs, st0, args.cpu_steps)
        model, ncpu = host_cpu()
        cpu = {"value": rate, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "cpu_model": model, "host_logical_cpus": ncpu,
               "sample": "oracle (plain-C FP64 restatement of the reference's loop, gcc -O3 "
                         "-march=x86-64-v3, 1 thread) on the same %d-bead %s start, %d steps "
                         "after 2 warm-up steps (%.1f s)" % (n, name, args.cpu_steps, secs)}
    if name == "melt100k":
        config = {"n_particles": n, "chains": 2000, "chain_length": 50, "density": 0.85,
                  "lj_rc": 2.5, "dt": 0.01, "dpd_gamma": 1.0,
                  "parallelism": "single GPU"}
        data = ("synthetic (seeded random-walk melt, Maxwell-Boltzmann velocities, relaxed "
                "by the switch-on ramp before timing)")
        key = (tm_m["tpp"], int(tm_m["pair_queue"]), tuple(tm_m["n_bin"]))
        traffic, valu = (FORCE_COUNTERS.get(key, (None, None)) if not tm_m["half_list"]
                         else (None, None))
    else:
        config = {"n_particles": n, "chains": inputs.n_chain, "chain_length": inputs.n_mon,
                  "box": list(p.boundary), "dt": inputs.dt, "dpd_gamma": inputs.friction[0],
                  "parallelism": "single GPU"}
        if name == "pe_brush27k":
            config.update({"alpha_c": inputs.alpha_c, "n_kmax": inputs.n_kmax,
                           "rc_coul": inputs.rc_coul})
        data = ("synthetic (seeded grafted random walks and uniform solvent, Maxwell-Boltzmann "
                "velocities, relaxed by the 40-tau switch-on ramp before timing)")
        traffic = None
        valu = None
    xt = kernel_xt(p)
    # the dominant kernel of the mixed step: the force pass fused with the
    # integration (charged systems: and the k-space force's addition)
    roof = roofline(n, tm_m["step_force_ms"], traffic,
                    kernel="k_force_mixed<%d,2,%d,%d>" % (tm_m["tpp"], xt, tm_m["pair_queue"]),
                    bytes_per_particle=STEP_BYTES_PER_PARTICLE)
    roof["force_alone_ms_avg"] = tm_m["force_ms"]
    roof["force_ms_avg_fp64"] = tm_d["force_ms"]
    if valu:
        # the kernel is VALU-issue bound: its instruction rate against the issue peak
        g = valu / (tm_m["step_force_ms"] * 1e-3) / 1e9
        roof["valu_issue"] = {"achieved": g, "peak": VALU_PEAK_GINSTS,
                              "unit": "G wave-instructions/s", "frac": g / VALU_PEAK_GINSTS,
                              "instructions_per_launch": valu}
    config["threads_per_particle"] = {"mixed": tm_m["tpp"], "fp64": tm_d["tpp"]}
    config["list_skin"] = {"mixed": tm_m["skin"], "fp64": tm_d["skin"],
                           "chosen_by": "autotuner" if tune else "the workload's spec"}
    config["cells"] = {"mixed": tm_m["n_bin"], "fp64": tm_d["n_bin"], "reference": list(p.n_bin),
                       "chosen_by": "autotuner"}
    config["pair_queue"] = {"mixed": tm_m["pair_queue"], "chosen_by": "autotuner"}
    config["half_list"] = {"mixed": tm_m["half_list"], "chosen_by": "autotuner"}
    extra = {
        "fp64_value": value_fp64,
        "mixed_over_fp64": value / value_fp64,
        "force_err_max": err_max,
        "force_err_rms": err_rms,
        "rebuilds": obs_m["n_rebuilds"],
        # kernel durations from HIP events on the context's stream (the
        # force pass: 200 launches; a rebuild's three kernels: 20 rebuilds)
        "kernel_ms": {"force": tm_m["force_ms"], "force_in_step": tm_m["step_force_ms"],
                      **tm_m["rebuild_ms"],
                      "rebuild_idle": sum(tm_m["rebuild_idle_ms"].values()),
                      "rebuild_per_step": sum(tm_m["rebuild_ms"].values()) *
                      obs_m["n_rebuilds"] / args.steps},
        "step_roofline_frac": value * STEP_BYTES_PER_PARTICLE / (HBM_PEAK_GBS * 1e9),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if tm_m.get("ewald_ms"):
        extra["ewald_roofline"] = ewald_roofline(p, inputs, t, tm_m["ewald_ms"])
        extra["ewald_roofline"]["kernel_ms"] = tm_m["ewald_parts_ms"]
        log("k-space stage %.4f ms avg: %.2f TFLOP/s FP64" % (
            tm_m["ewald_ms"], extra["ewald_roofline"]["achieved"]))
    return contract_line(args, name, value, dt_m / args.steps * 1e3,
                         "fp32 pair math, exact fixed-point sums, fp64 state", data, config, extra)


def bench_batch(args, n_rep=256):
    """256 uneven brush replicas: one context, one launch of each kernel per
    step for all of them, against the same replicas advanced one context at a
    time.  Untimed: the autotuners (threads per particle, list skin), then the
    rest of the 40-tau switch-on ramp of the fresh starts; both timed runs
    start from the relaxed states the ramp leaves and run the production
    potential with the same width and skin."""
    t0 = time.perf_counter()
    reps = []
    for q in range(n_rep):
        x = configs.ensemble_replica(q)
        p, r, v, t = configs.build(x, 1, min_dist=0.7)
        reps.append((p, r, v, t))
    ntot = sum(p.n_part for p, *_ in reps)
    log("batch%d: %d beads (%d..%d per replica), generated in %.1f s" % (
        n_rep, ntot, min(p.n_part for p, *_ in reps), max(p.n_part for p, *_ in reps),
        time.perf_counter() - t0))
    prec = lib.MFA_MIXED
    with lib.Batch([s[0] for s in reps], prec) as b:
        b.set_graph_steps(args.graph_steps)
        for q, (p, r, v, t) in enumerate(reps):
            b.set_replica(q, r, v, np.zeros_like(r), t)
        tpp, ms_tune = b.autotune(1, 60)
        log("autotune: %d threads per particle, pair queue %d, half list %s (%.3f ms per step)" % (
            tpp, b.tuning()["pair_queue"], b.tuning()["half_list"], ms_tune))
        first = 1 + autotune_trials(reps[0][0], prec) * (args.graph_steps + 1 + 60)  # steps it ran
        # the reference grid (batches keep it); the skin autotuner over the
        # candidates those cells hold
        cand = [x for x in configs.SKIN_CANDIDATES if all(_fits(rp[0], x) for rp in reps)]
        skin, ms_skin = b.autotune_skin(first, 100, cand)
        first += len(cand) * (args.graph_steps + 1 + 100)
        width = None
        log("autotune: skin %.2f (%.3f ms per step)" % (skin, ms_skin))
        pair_queue = b.tuning()["pair_queue"]
        half_list = b.tuning()["half_list"]
        ramp = configs.ramp_steps(reps[0][0])  # every replica has dt = 0.005
        t2 = time.perf_counter()
        if first <= ramp:
            b.run_steps(first, ramp + 1 - first, want_obs=False)
            first = ramp + 1
        log("switch-on ramp done at step %d (%.1f s)" % (first - 1, time.perf_counter() - t2))
        relaxed = [] if args.no_seq else [b.get_replica(q) for q in range(n_rep)]
        b.run_steps(first, args.warmup, want_obs=False)
        t1 = time.perf_counter()
        b.run_steps(first + args.warmup, args.steps, want_obs=False)
        dt_b = time.perf_counter() - t1
        force_ms = b.time_force_step(50)
        rb_idle = b.time_rebuild(20, idle=True)
        b.get_replica(0)
        rb_full = b.time_rebuild(10)  # resets every skin reference: after the timed steps
    log("batched: %.3f s for %d steps; force pass %.3f ms; rebuild kernels idle %s, all replicas %s"
        % (dt_b, args.steps, force_ms, rb_idle, rb_full))
    dt_s = 0.0
    for q, st in enumerate(relaxed):
        p = reps[q][0]
        with lib.Context(p, prec) as c:
            c.set_graph_steps(args.graph_steps)
            c.set_tpp(tpp)
            c.set_skin(skin)
            c.set_pair_queue(pair_queue)
            c.set_half_list(half_list)
            c.set_state(st["r"], st["v"], st["a"], reps[q][3], st["mic"])
            c.run_steps(first, args.warmup, want_obs=False)
            t1 = time.perf_counter()
            c.run_steps(first + args.warmup, args.steps, want_obs=False)
            dt_s += time.perf_counter() - t1
    log("sequential: %.3f s for %d steps of each replica" % (dt_s, args.steps))
    value = ntot * args.steps / dt_b
    dt_s = dt_s if dt_s > 0 else float("nan")
    config = {"n_replicas": n_rep, "n_particles": ntot, "grafting_density": [0.05, 0.5],
              "chain_length": [20, 100], "lateral_edge": [8, 40], "lz": 12.5, "dt": 0.005,
              "solvent_density": 3.0, "threads_per_particle": tpp,
              "list_skin": {"value": skin, "chosen_by": "autotuner"},
              "cells": "the reference grid of each replica (batches are not cell-tuned)",
              "pair_queue": {"value": pair_queue, "chosen_by": "autotuner"},
              "half_list": {"value": half_list, "chosen_by": "autotuner"},
              "parallelism": "single GPU, replicas batched"}
    extra = {
        # (None with --no-seq: no sequential runs)
        "sequential_value": ntot * args.steps / dt_s if relaxed else None,
        "batched_over_sequential": dt_s / dt_b if relaxed else None,
        "kernel_ms": {"force_in_step": force_ms, "rebuild_idle": rb_idle, "rebuild_all": rb_full},
        "roofline": roofline(ntot, force_ms, None,
                             kernel="k_force_mixed<%d,2,%d,%d>" % (tpp, kernel_xt(reps[0][0]), pair_queue),
                             bytes_per_particle=STEP_BYTES_PER_PARTICLE),
        "cpu_baseline": None,
    }
    data = ("synthetic (seeded replicas relaxed by the 40-tau switch-on ramp of the reference's "
            "start-up, untimed; timed at the production potential from the relaxed states)")
    return contract_line(args, "batch256", value, dt_b / args.steps * 1e3,
                         "fp32 pair math, exact fixed-point sums, fp64 state", data, config, extra)


def bench_adjoint(args):
    """Gradient of the window-averaged brush height and top-wall pressure with
    respect to (epsilon_LJ, sigma, k_FENE, R0, gamma_DPD) over a window of
    `steps` steps with checkpoints every 20 steps: the forward pass in the
    mixed form (FP32 forces) and in the FP64 form, the adjoint in FP64; the
    FP64 gradient against central finite differences of the FP64 forward
    pass at relative steps 1e-4 (the workload's spec) and 1e-6.  Over a
    200-step window J(theta) is smooth only on scales below ~1e-4 (pairs
    cross the cut-offs that sigma moves), so the 1e-6 differences are the
    ones that resolve the derivative."""
    w = configs.WORKLOADS["adjoint50k"]
    inputs = w["make"]()
    t0 = time.perf_counter()
    p, st = configs.prepare_brush(inputs, w["kind"], min_dist=w["min_dist"])
    n = p.n_part
    log("adjoint50k: n=%d box=%s cells=%s, prepared in %.1f s" % (
        n, list(p.boundary), list(p.n_bin), time.perf_counter() - t0))
    theta = configs.theta_of(inputs)
    first = 2 + args.warmup
    # warm-up steps at the base parameters; every window (the finite
    # differences' included) starts from the state they reach
    with lib.Context(p, lib.MFA_FP64) as c:
        c.set_state(st["r"], st["v"], st["a"], st["t"], st["mic"])
        c.run_steps(2, args.warmup, want_obs=False)
        st0 = dict(c.get_state(), t=st["t"])

    def window(th, precision, want_grad):
        with lib.Context(lib.derive(configs.with_theta(inputs, th)), precision) as c:
            c.set_state(st0["r"], st0["v"], st0["a"], st0["t"], st0["mic"])
            c.adjoint_window(first, 1, 1, False)  # first list build and caches, untimed
            c.set_state(st0["r"], st0["v"], st0["a"], st0["t"], st0["mic"])
            t1 = time.perf_counter()
            J, G, nreb = c.adjoint_window(first, args.steps, 20, want_grad)
            return J, G, time.perf_counter() - t1

    J64, G64, t64 = window(theta, lib.MFA_FP64, True)
    Jm, Gm, tm = window(theta, lib.MFA_MIXED, True)
    _, _, tfwd = window(theta, lib.MFA_MIXED, False)
    log("adjoint window: mixed forward %.3f s (forward only %.3f s), fp64 forward %.3f s" % (
        tm, tfwd, t64))
    def fd(h):
        G = np.zeros((2, 5))
        for q in range(5):
            tp, tq = theta.copy(), theta.copy()
            tp[q] *= 1.0 + h
            tq[q] *= 1.0 - h
            G[:, q] = (window(tp, lib.MFA_FP64, False)[0] - window(tq, lib.MFA_FP64, False)[0]) / (
                tp[q] - tq[q])
        return G

    def relerr(G, ref):
        return np.abs(G - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max(axis=1, keepdims=True))

    Gfd, Gfd6 = fd(1e-4), fd(1e-6)
    rel, rel6, relm = relerr(G64, Gfd), relerr(G64, Gfd6), relerr(Gm, G64)
    log("J %s\nadjoint (fp64)\n%s\nFD h=1e-4 (fp64)\n%s\nFD h=1e-6 (fp64)\n%s\nmax rel err vs h=1e-4 "
        "%.3e, vs h=1e-6 %.3e; mixed-forward vs fp64 %.3e" % (
            J64, G64, Gfd, Gfd6, rel.max(), rel6.max(), relm.max()))
    value = n * args.steps / tm
    config = {"n_particles": n, "chains": inputs.n_chain, "chain_length": inputs.n_mon,
              "box": list(p.boundary), "dt": inputs.dt, "window": args.steps,
              "checkpoint_every": 20, "parameters": list(configs.THETA_NAMES),
              "parallelism": "single GPU"}
    extra = {
        "observables": {"height": J64[0], "p_wall": J64[1]},
        "grad_fp64": G64.tolist(), "grad_fd_fp64": Gfd.tolist(), "grad_mixed_forward": Gm.tolist(),
        "grad_vs_fd_max_rel": float(rel.max()),
        "grad_fd_h1e-6_fp64": Gfd6.tolist(), "grad_vs_fd_h1e-6_max_rel": float(rel6.max()),
        "grad_mixed_vs_fp64_max_rel": float(relm.max()),
        "window_s_mixed": tm, "window_s_fp64": t64, "forward_only_s_mixed": tfwd,
        "roofline": None, "cpu_baseline": None,
    }
    data = ("synthetic (seeded grafted random walks and uniform solvent, relaxed by the 40-tau "
            "switch-on ramp)")
    return contract_line(args, "adjoint50k", value, tm / args.steps * 1e3,
                         "fp32 forward forces, fp64 adjoint", data, config, extra)


COPY_BYTES = 256 * 1024 * 1024  # bytes read and bytes written by the calibration copy


def counter_workload(args):
    """The workload of the counter runs, meant to run under rocprofv3 with
    one counter set per run (few dispatches, so each pass takes seconds):
    loads the state a `--save-state` run left (the state the timed steps
    reached, and the width, list skin and pair-queue choice the autotuners
    made there), bins
    it `--rebuilds` times, runs the force pass `--reps` times with direct
    launches and then `--reps` times as the step runs it (melt100k,
    brush9k: fused with the integration), or the k-space Ewald stage
    `--reps` times (pe_brush27k), then a device copy of known size: the
    calibration of FETCH_SIZE and WRITE_SIZE."""
    w = configs.WORKLOADS[args.workload]
    p = lib.derive(w["make"]())
    st = np.load(args.state)
    tpp = int(st["tpp"])
    with lib.Context(p, lib.MFA_MIXED) as c:
        c.set_graph_steps(0)
        c.set_tpp(tpp)
        if "cell_width" in st:
            c.set_cells(float(st["cell_width"]))
        c.set_skin(float(st["skin"]))  # the autotuners' choices of the timed run
        c.set_pair_queue(int(st["pair_queue"]))
        c.set_half_list(bool(st["half_list"]) if "half_list" in st else False)
        c.set_state(st["r"], st["v"], st["a"], st["t"], st["mic"])
        for _ in range(args.rebuilds):
            c.binning()
        ms = ms_step = float("nan")
        if p.coulomb:
            for _ in range(args.reps):
                c.ewald_k(1)
        else:
            ms = c.time_force(args.reps)
            # the kernel the step runs (in the mixed form: the force pass that
            # also integrates), its writes sent to copies
            ms_step = c.time_force_step(args.reps)
    cms = lib.copy_probe(COPY_BYTES)
    log("%s: %d threads per particle; force pass %.4f ms, as the step runs it %.4f ms; "
        "calibration copy of %d bytes in %.4f ms"
        % (args.workload, tpp, ms, ms_step, COPY_BYTES, cms))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="melt100k",
                    choices=["melt100k", "brush9k", "pe_brush27k", "batch256", "adjoint50k"])
    ap.add_argument("--cpu-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-seq", action="store_true",
                    help="batch256: skip the sequential comparison (profiling runs)")
    ap.add_argument("--graph-steps", type=int, default=16,
                    help="steps per captured hipGraph block (0: direct launches)")
    ap.add_argument("--save-state", default=None,
                    help="write the state the timed mixed steps left (and the chosen threads "
                         "per particle) to this .npz, for --counters")
    ap.add_argument("--counters", action="store_true",
                    help="counter-run workload: load --state, bin, run --reps force passes "
                         "(or k-space stages), then a calibration copy; prints no JSON line")
    ap.add_argument("--state", default=None)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--rebuilds", type=int, default=1)
    args = ap.parse_args()
    if args.counters:
        if not args.state:
            ap.error("--counters needs --state (written by a --save-state run)")
        counter_workload(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(each GPU box has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    defaults = {"melt100k": (5000, 500, 40), "brush9k": (1000, 200, 1000),
                "pe_brush27k": (2000, 200, 3), "batch256": (500, 100, 0),
                "adjoint50k": (200, 20, 0)}
    d_steps, d_warm, d_cpu = defaults[args.workload]
    args.steps = d_steps if args.steps is None else args.steps
    args.warmup = d_warm if args.warmup is None else args.warmup
    args.cpu_steps = d_cpu if args.cpu_steps is None else args.cpu_steps
    if args.workload == "batch256":
        line = bench_batch(args)
    elif args.workload == "adjoint50k":
        line = bench_adjoint(args)
    else:
        line = bench_single(args, args.workload)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
