"""Cell grids other than the reference's (mfa_set_cells) and the cell-size
autotuner (mfa_autotune_cells).  The pairs within the list radius do not
depend on the grid, so the bar is the one of the reference grid: the
unordered pair set bit-exact against the oracle (whose grid is the
reference's), FP64 forces within 1e-12 of the oracle and FP64 trajectories
of 40 steps within 1e-9, mixed-form forces within an RMS error of 1e-5 (the
FP32 view holds positions relative to the cell origin, so the mixed form's
rounding, though not its pair set, depends on the grid); for single systems
and a batch."""
import numpy as np
import pytest

from mfa_amd import configs, lib
from tests.util import force_error, gpu_ctx, make_case, oracle_sys

pytestmark = pytest.mark.gpu


def _wider(p, extra):
    """A cell width `extra` beyond the reference grid's cell, and the grid
    it gives (at least 3 cells per axis)."""
    w = max(p.r_bin[d] for d in range(3)) + extra
    q = lib.with_cells(p, w)
    return w, [q.n_bin[d] for d in range(3)]


def _case(name):
    """Systems whose boxes hold a coarser grid than the reference's (at
    least 3 cells per axis): the brush9k workload and a 5,000-bead melt."""
    if name == "brush9k":
        w = configs.WORKLOADS["brush9k"]
        inputs, kind, md = w["make"](), w["kind"], w["min_dist"]
    elif name == "melt5k":
        inputs, kind, md = configs.melt(100, 50, 0.85, lj_rc=2.5, seed=41), 2, 0.8
    else:  # WCA melt: under the switch-on ramp an rc = 2.5 melt contracts
        inputs, kind, md = configs.melt(100, 50, 0.85, lj_rc=0.0, dt=0.005, seed=42), 2, 0.8
    p, r, v, t = configs.build(inputs, kind, min_dist=md)
    inputs.s_time = 1            # a fresh start: the reference's switch-on ramp
    inputs.force_switch_on = 1   # (init_force_switch_on.f90:15-26) keeps the seeded
    p = lib.derive(inputs)       # overlaps finite over the 40 steps
    return inputs, p, r, v, t, np.zeros_like(r)


@pytest.mark.parametrize("case", ["brush9k", "melt5k"])
def test_coarser_cells_same_pairs_and_forces(case):
    inputs, p, r, v, t, a = _case(case)
    w, nb = _wider(p, 0.35)
    print("%s: reference grid %s, width %.3f grid %s" % (case, list(p.n_bin), w, nb), flush=True)
    assert nb != list(p.n_bin)
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    op = s.get_pairs()
    fref, _ = s.forces(step=7)
    out = {}
    for coarse in (False, True):
        for prec in (lib.MFA_FP64, lib.MFA_MIXED):
            ctx = gpu_ctx(p, r, v, a, t, precision=prec)
            if coarse:
                ctx.set_cells(w)
            ctx.binning()
            gp = ctx.get_pairs()
            gp = gp[np.lexsort((gp[:, 1], gp[:, 0]))]
            assert gp.shape == op.shape and np.array_equal(gp, op), (coarse, prec)
            ctx.energies(step=7)
            f = ctx.get_state()["f"]
            per, rms = force_error(f, fref)
            if prec == lib.MFA_FP64:
                assert per < 1e-12, (coarse, per)
                ctx.set_state(r, v, a, t)  # the trajectory from the start state
                obs = ctx.run_steps(8, 40)
                out[coarse] = (ctx.get_state(), obs)
            else:
                assert rms < 1e-5, (coarse, rms)
            ctx.close()
    s.set_state(r, v, a, t)
    s.run_steps(8, 40)
    ref = s.get_state()
    for coarse in (False, True):
        st, obs = out[coarse]
        dr = np.abs(st["r"] - ref["r"]).max()
        print("coarse %s: 40 FP64 steps, max |dr| %.2e, %d rebuilds" % (coarse, dr, obs["n_rebuilds"]),
              flush=True)
        assert np.array_equal(st["mic"], ref["mic"]) and dr < 1e-9


def test_set_cells_refuses_grids_it_cannot_use():
    inputs, p, r, v, t, a = _case("wca5k")
    ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
    with pytest.raises(lib.MfaError):
        ctx.set_cells(0.5 * p.r_bin[0])  # narrower than the list radius
    with pytest.raises(lib.MfaError):
        ctx.set_cells(p.boundary[0])  # fewer than 3 cells
    # the context still runs on its grid
    ctx.run_steps(1, 10)
    ctx.close()


def test_batch_refuses_other_grids():
    """Replica batches keep the reference grid (mfa_set_cells refuses them)."""
    reps = [configs.build(configs.brush(20, 10, 900, 10.0, 10.0, 8.0, x_space=0.5, seed=31), 1),
            configs.build(configs.brush(12, 12, 700, 9.0, 9.0, 9.0, x_space=0.5, seed=32), 1)]
    with lib.Batch([q[0] for q in reps], lib.MFA_FP64) as b:
        with pytest.raises(lib.MfaError):
            b.set_cells(2.0)
        for q, (p, r, v, t) in enumerate(reps):
            b.set_replica(q, r, v, np.zeros_like(r), t)
        b.run_steps(1, 10)


@pytest.mark.skip(reason="replica batches refuse other cell grids: the cell-size autotuner on the "
                         "256-replica batch ended in a GPU memory fault whose cause is not found")
def test_batch_coarser_cells():
    # three brushes of different boxes, chain counts and seeds, each with
    # room for a coarser grid
    reps = [configs.build(configs.brush(20, 10, 900, 10.0, 10.0, 8.0, x_space=0.5, seed=31), 1),
            configs.build(configs.brush(12, 12, 700, 9.0, 9.0, 9.0, x_space=0.5, seed=32), 1),
            configs.build(configs.brush(16, 8, 800, 11.0, 10.0, 8.0, x_space=0.5, seed=33), 1)]
    w = max(max(q[0].r_bin[d] for d in range(3)) for q in reps) + 0.2
    for q in reps:
        assert list(lib.with_cells(q[0], w).n_bin) != list(q[0].n_bin)
    out = {}
    for coarse in (False, True):
        with lib.Batch([q[0] for q in reps], lib.MFA_FP64) as b:
            if coarse:
                b.set_cells(w)
            for q, (p, r, v, t) in enumerate(reps):
                b.set_replica(q, r, v, np.zeros_like(r), t)
            b.run_steps(1, 40)
            out[coarse] = [(b.get_replica(q), b.replica_obs(q)) for q in range(len(reps))]
    for q in range(len(reps)):
        (s0, o0), (s1, o1) = out[False][q], out[True][q]
        assert np.array_equal(s0["mic"], s1["mic"]), q
        assert np.abs(s0["r"] - s1["r"]).max() < 1e-9, q
        assert o0["n_rebuilds"] == o1["n_rebuilds"], q


def test_autotune_cells_keeps_a_candidate():
    inputs, p, r, v, t, a = _case("wca5k")
    ref_w = max(p.r_bin[d] for d in range(3))
    widths = [ref_w, ref_w + 0.25, ref_w + 0.5]
    skins = [0.3, 0.5, 0.7]
    assert len({tuple(lib.with_cells(p, w).n_bin) for w in widths}) == 3
    with lib.Context(p, lib.MFA_MIXED) as c:
        c.set_state(r, v, a, t)
        w, skin, ms, nxt = c.autotune_cells(1, 20, widths, skins)
        print("autotune_cells: width %.3f skin %.2f (%.4f ms per step), next step %d" % (
            w, skin, ms, nxt), flush=True)
        assert w in widths and skin in skins and ms > 0.0 and nxt > 1
        assert abs(c.tuning()["skin"] - skin) < 1e-12
        c.run_steps(nxt, 20)
