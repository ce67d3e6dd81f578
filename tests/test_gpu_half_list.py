"""Half-list form of the mixed force pass (mfa_set_half_list): each pair is
evaluated once, from the row of its lower sorted position, and the other
particle takes its term from the kept pair scalar and its own negated
separation.  Every pair term is odd in the separation and the fixed-point
sums are exact, so the bar is bit-identity with the full-list pass: forces,
40-step trajectories with rebuilds (the mirror slots found again after each
one), rebuild counts; a full-size melt; a batch of uneven replicas; and an
adjoint window, whose checkpoint restores bring back older lists."""
import numpy as np
import pytest

from mfa_amd import configs, lib
from tests.test_gpu_parity import _lgv_case
from tests.util import force_error, gpu_ctx, make_case, oracle_sys

pytestmark = pytest.mark.gpu

CASES = ["brush_small", "melt_small", "melt_switch_on", "pe_brush_small", "brush_bend"]


def _forces_and_run(p, r, v, a, t, tpp, half, queue=False, steps=40):
    ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED, tpp=tpp)
    ctx.set_pair_queue(queue)
    ctx.set_half_list(half)
    tun = ctx.tuning()
    assert tun["half_list"] == half and tun["pair_queue"] == queue
    ctx.energies(step=3)  # sets the step; a full-list pass with the energies
    f_pair = ctx.get_state()["f"]
    t_ms = ctx.time_force(5)  # force passes without energies at that step, the state kept
    f_all = ctx.get_state()["f"]
    if not p.coulomb:  # (the energy pass also adds the k-space forces)
        assert np.array_equal(f_pair, f_all)
    obs = ctx.run_steps(4, steps)
    st = ctx.get_state()
    ctx.close()
    return f_pair, f_all, st, obs, t_ms


def _same(a, b, what):
    fa, fb = a, b
    for k in range(2):
        assert np.array_equal(fa[k], fb[k]), (what, k)
    for key in ("r", "v", "a", "f", "mic"):
        assert np.array_equal(fa[2][key], fb[2][key]), (what, key)
    assert fa[3]["n_rebuilds"] == fb[3]["n_rebuilds"], what
    assert fa[3]["t_kin"] == fb[3]["t_kin"], what


@pytest.mark.parametrize("tpp", [4, 8, 16])
@pytest.mark.parametrize("case", CASES)
def test_half_list_bit_identical(case, tpp):
    inputs, p, r, v, t, a = make_case(case)
    full = _forces_and_run(p, r, v, a, t, tpp, False)
    half = _forces_and_run(p, r, v, a, t, tpp, True)
    print("%s tpp %d: %d rebuilds, force pass %.4f ms full, %.4f ms half" % (
        case, tpp, full[3]["n_rebuilds"], full[4], half[4]), flush=True)
    assert full[3]["n_rebuilds"] >= 1
    _same(full, half, case)


@pytest.mark.parametrize("tpp", [4, 16])
def test_half_list_with_pair_queue(tpp):
    """The pair queue (DPD_CUT_OFF brush) in the first pass: the queued
    pairs write their scalars from another lane."""
    inputs, p, r, v, t, a = make_case("brush_small")
    assert p.dpd_cut_off_on
    full = _forces_and_run(p, r, v, a, t, tpp, False, queue=False)
    half = _forces_and_run(p, r, v, a, t, tpp, True, queue=True)
    _same(full, half, "queue")


def test_half_list_langevin():
    inputs, p, r, v, t, a = _lgv_case("melt_small")
    _same(_forces_and_run(p, r, v, a, t, 8, False), _forces_and_run(p, r, v, a, t, 8, True),
          "langevin")


def test_half_list_refused_below_four_lanes_is_the_full_pass():
    """At 1 or 2 threads per particle the half-list form does not apply: the
    full-list pass runs, with the same forces."""
    inputs, p, r, v, t, a = make_case("melt_small")
    _same(_forces_and_run(p, r, v, a, t, 2, False), _forces_and_run(p, r, v, a, t, 2, True),
          "tpp 2")


def test_half_list_melt100k_forces_and_oracle():
    """Full-size melt: one force pass bit-identical to the full-list pass,
    and within the mixed form's bar of the oracle's FP64 forces."""
    w = configs.WORKLOADS["melt100k"]
    inputs = w["make"]()
    p, r, v, t = configs.build(inputs, w["kind"], min_dist=w["min_dist"])
    a = np.zeros_like(r)
    out = {}
    for half in (False, True):
        ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED, tpp=8)
        ctx.set_half_list(half)
        ctx.energies(step=9)  # the step of the passes below
        ctx.time_force(2)  # two passes on the same list (mirror slots reused)
        out[half] = ctx.get_state()["f"]
        ctx.close()
    assert np.array_equal(out[False], out[True])
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    fref, _ = s.forces(step=9)
    per, rms = force_error(out[True], fref)
    print("melt100k half-list vs oracle: max %.2e rms %.2e" % (per, rms), flush=True)
    assert rms < 1e-5


def test_half_list_batch():
    """Replicas of different boxes and sizes in one context: the half-list
    batch equals the full-list batch bit for bit over 40 steps."""
    cases = configs.small_cases()
    reps = []
    for name in ("brush_small", "brush_bend"):
        kind, inp = cases[name]
        reps.append(configs.build(inp, kind))
    reps.append(configs.build(configs.brush(6, 8, 150, 5.0, 5.0, 7.0, x_space=0.5, seed=21), 1))
    out = {}
    for half in (False, True):
        with lib.Batch([q[0] for q in reps], lib.MFA_MIXED) as b:
            b.set_tpp(8)
            b.set_half_list(half)
            for q, (p, r, v, t) in enumerate(reps):
                b.set_replica(q, r, v, np.zeros_like(r), t)
            b.run_steps(1, 40)
            out[half] = [(b.get_replica(q), b.replica_obs(q)) for q in range(len(reps))]
    for q in range(len(reps)):
        (s0, o0), (s1, o1) = out[False][q], out[True][q]
        for k in ("r", "v", "f", "mic"):
            assert np.array_equal(s0[k], s1[k]), (q, k)
        assert o0["n_rebuilds"] == o1["n_rebuilds"]


def test_half_list_adjoint_window():
    """Mixed forward pass of the adjoint window with the half-list form: its
    segments restart from checkpointed lists, for which the mirror slots are
    found again.  The observables equal the full-list window's bit for bit;
    the gradients to FP64 rounding (the adjoint pass sums a row's entries in
    FP64, and the half-list form orders the row's entries otherwise)."""
    kind, x = configs.small_cases()["brush_small"]
    p, st = configs.prepare_brush(x, kind, precision=lib.MFA_FP64)
    out = {}
    for half in (False, True):
        with lib.Context(p, lib.MFA_MIXED) as c:
            c.set_tpp(8)
            c.set_half_list(half)
            c.set_state(st["r"], st["v"], st["a"], st["t"], st["mic"])
            J, G, _ = c.adjoint_window(1, 40, 10, True)
            out[half] = (np.array(J), np.array(G))
    assert np.array_equal(out[False][0], out[True][0])
    g0, g1 = out[False][1], out[True][1]
    assert np.all(np.abs(g1 - g0) <= 1e-12 * np.abs(g0).max(axis=1, keepdims=True)), (g0, g1)


def test_autotune_may_choose_the_half_list():
    """The width autotuner tries the half-list form at 4, 8 and 16 lanes and
    reports its choice; steps after it keep running."""
    inputs, p, r, v, t, a = make_case("melt_small")
    with lib.Context(p, lib.MFA_MIXED) as c:
        c.set_state(r, v, a, t)
        tpp, ms = c.autotune(1, 10)
        tun = c.tuning()
        print("autotune: tpp %d half %s (%.4f ms)" % (tpp, tun["half_list"], ms), flush=True)
        assert tun["threads_per_particle"] == tpp and ms > 0.0
        assert not tun["half_list"] or tpp >= 4
        c.run_steps(200, 10)
