"""The HIP path against the oracle on the five benchmark workloads at their
own sizes (mfa_amd/configs.py), through the C ABI.

  brush9k      9,000 beads: cells and pair set bit-exact, FP64 forces and
               energies within 1e-12, mixed RMS within 1e-5, and a 40-step
               FP64 trajectory within 1e-9 (started under the reference's
               force switch-on, as a fresh run starts);
  melt100k     100,000 beads: FP64 forces and energies within 1e-12 and the
               mixed form's RMS error within 1e-5 of the oracle's forces;
  pe_brush27k  27,000 beads, 10,400 charges, 3,997 k vectors: FP64 forces
               and v_coul within 1e-12 in both Ewald forms (the reference's
               1/pi inside the LJ range, and erfc within rc_coul = 3), mixed
               RMS within 1e-5;
  batch256     all 256 replicas (2,400 .. 60,000 beads) in one context: the
               smallest, median and largest replicas bit-identical to
               single-context runs after 20 steps (both forms), the smallest
               within 1e-9 of the oracle;
  adjoint50k   50,000 beads: the FP64 adjoint of the 200-step window
               (checkpoints every 20 steps) within 1e-4 of central finite
               differences (h = 1e-6) of the FP64 forward pass, and of a
               40-step window within 1e-4 at h = 1e-4.
Per-particle force error: |dF| / max(|F|, 1e-3 F_rms); RMS: relative to
F_rms.  The start states are the seeded generator's (no relaxation), so the
forces include close contacts.
"""
import numpy as np
import pytest

from mfa_amd import configs, lib
from oracle import oracle as orc
from tests.util import force_error, rel

pytestmark = pytest.mark.gpu


def _start(name):
    w = configs.WORKLOADS[name]
    inputs = w["make"]()
    p, r, v, t = configs.build(inputs, w["kind"], min_dist=w["min_dist"])
    return inputs, p, r, v, t, np.zeros_like(r)


def _forces_vs_oracle(inputs, p, r, v, t, a, step=5, pairs=False):
    s = orc.System(inputs)
    s.set_state(r, v, a, t)
    s.binning()
    fref, eref = s.forces(step=step)
    out = {}
    for prec in (lib.MFA_FP64, lib.MFA_MIXED):
        with lib.Context(p, prec) as c:
            c.set_state(r, v, a, t)
            if pairs and prec == lib.MFA_FP64:
                c.binning()
                assert np.array_equal(c.get_cells(), s.get_cells())
                gp = c.get_pairs()
                gp = gp[np.lexsort((gp[:, 1], gp[:, 0]))]
                op = s.get_pairs()
                assert gp.shape == op.shape and np.array_equal(gp, op)
            obs = c.energies(step=step)
            out[prec] = (c.get_state()["f"], obs)
    f64, e64 = out[lib.MFA_FP64]
    per, rms = force_error(f64, fref)
    fm, em = out[lib.MFA_MIXED]
    per_m, rms_m = force_error(fm, fref)
    print("n=%d fp64 max %.2e rms %.2e; mixed max %.2e rms %.2e" % (p.n_part, per, rms, per_m, rms_m),
          flush=True)
    assert per < 1e-12, (per, rms)
    keys = ["v_fluid_fluid", "v_intra_molec", "virial"]
    if p.symmetry == 0:
        keys += ["v_fluid_wall", "wall_fz_bottom"]
    if p.coulomb:
        keys += ["v_coul"]
    for k in keys:
        assert rel(e64[k], eref[k]) < 1e-12, (k, e64[k], eref[k])
    assert rms_m < 1e-5, (per_m, rms_m)
    assert rel(em["v_fluid_fluid"], eref["v_fluid_fluid"]) < 1e-5
    return per, rms, per_m, rms_m


def test_brush9k_pairs_and_forces():
    print("brush9k", flush=True)
    _forces_vs_oracle(*_start("brush9k"), pairs=True)


def test_brush9k_trajectory_fp64():
    inputs, p, r, v, t, a = _start("brush9k")
    inputs.s_time = 1            # a fresh start: the reference's switch-on ramp
    inputs.force_switch_on = 1   # (init_force_switch_on.f90:15-26)
    p = lib.derive(inputs)
    s = orc.System(inputs)
    s.set_state(r, v, a, t)
    ref = s.run_steps(first_step=1, n=40)
    st_ref = s.get_state()
    with lib.Context(p, lib.MFA_FP64) as c:
        c.set_state(r, v, a, t)
        obs = c.run_steps(1, 40)
        st = c.get_state()
    dr = np.abs(st["r"] - st_ref["r"]).max()
    dv = np.abs(st["v"] - st_ref["v"]).max()
    print("brush9k 40 steps: dr %.2e dv %.2e rebuilds %d/%d" % (
        dr, dv, obs["n_rebuilds"], ref["n_rebuilds"]), flush=True)
    assert np.array_equal(st["mic"], st_ref["mic"])
    assert dr < 1e-9 and dv < 1e-8
    assert obs["n_rebuilds"] == ref["n_rebuilds"]
    for k in ("v_fluid_fluid", "v_intra_molec", "t_kin", "v_fluid_wall"):
        assert rel(obs[k], ref[k]) < 1e-9, (k, obs[k], ref[k])


def test_melt100k_forces():
    print("melt100k", flush=True)
    _forces_vs_oracle(*_start("melt100k"))


@pytest.mark.parametrize("compat", [False, True])
def test_pe_brush27k_ewald(compat):
    inputs, p, r, v, t, a = _start("pe_brush27k")
    if compat:  # the reference's form: 1/pi real-space factor inside the LJ range
        inputs.ewald_compat = 1
        inputs.rc_coul = 0.0
        p = lib.derive(inputs)
    nq = int(sum(1 for x in t if inputs.charge[x - 1] != 0.0))
    nk = len(orc.System(inputs).k_vectors())
    print("pe_brush27k compat=%d: %d charges, %d k vectors" % (compat, nq, nk), flush=True)
    assert nq == 10400 and nk == 3997
    _forces_vs_oracle(inputs, p, r, v, t, a, step=3)


@pytest.fixture(scope="module")
def ensemble():
    reps = []
    for q in range(256):
        x = configs.ensemble_replica(q)
        p, r, v, t = configs.build(x, 1, min_dist=0.7)
        reps.append((x, p, r, v, t))
    sizes = np.array([p.n_part for _, p, *_ in reps])
    order = np.argsort(sizes, kind="stable")
    pick = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    print("batch256: %d beads, replicas %s of sizes %s" % (
        sizes.sum(), pick, sizes[pick].tolist()), flush=True)
    return reps, pick


@pytest.mark.parametrize("precision", [lib.MFA_FP64, lib.MFA_MIXED])
def test_batch256_replicas_match_single_runs(ensemble, precision):
    reps, pick = ensemble
    n_steps = 20
    with lib.Batch([p for _, p, *_ in reps], precision) as b:
        b.set_tpp(8)
        for q, (_, p, r, v, t) in enumerate(reps):
            b.set_replica(q, r, v, np.zeros_like(r), t)
        b.run_steps(1, n_steps, want_obs=False)
        got = {q: b.get_replica(q) for q in pick}
    for q in pick:
        _, p, r, v, t = reps[q]
        with lib.Context(p, precision) as c:
            c.set_tpp(8)
            c.set_state(r, v, np.zeros_like(r), t)
            c.run_steps(1, n_steps, want_obs=False)
            st = c.get_state()
        for k in ("r", "v", "f", "a", "mic"):
            assert np.array_equal(got[q][k], st[k]), (q, k)
    if precision == lib.MFA_FP64:
        q = pick[0]
        x, p, r, v, t = reps[q]
        s = orc.System(x)
        s.set_state(r, v, np.zeros_like(r), t)
        s.run_steps(1, n_steps)
        ref = s.get_state()
        dr = np.abs(got[q]["r"] - ref["r"]).max()
        print("replica %d (n=%d) vs oracle: dr %.2e" % (q, p.n_part, dr), flush=True)
        assert dr < 1e-9 and np.array_equal(got[q]["mic"], ref["mic"])


# ------------------------------------------------------------- adjoint50k

def _grad_err(G, Gfd):
    """Per component |G - G_fd| / |G_fd|, components below 1e-3 of their
    row's largest against that largest value."""
    scale = np.abs(Gfd).max(axis=1, keepdims=True)
    return np.abs(G - Gfd) / np.maximum(np.abs(Gfd), 1e-3 * scale)


@pytest.fixture(scope="module")
def adjoint50k():
    w = configs.WORKLOADS["adjoint50k"]
    inputs = w["make"]()
    p, st = configs.prepare_brush(inputs, w["kind"], min_dist=w["min_dist"],
                                  precision=lib.MFA_FP64)
    print("adjoint50k relaxed: n=%d" % p.n_part, flush=True)
    th = configs.theta_of(inputs)

    def window(theta, n_steps, want_grad):
        with lib.Context(lib.derive(configs.with_theta(inputs, theta)), lib.MFA_FP64) as c:
            c.set_state(st["r"], st["v"], st["a"], st["t"], st["mic"])
            J, G, _ = c.adjoint_window(2, n_steps, 20, want_grad)
        return J, G
    return th, window


def _fd(window, th, n_steps, h, order):
    """Central differences with relative step h: the 3-point stencil
    (order 2) or the 5-point stencil (order 4, steps h and 2h)."""
    G = np.zeros((2, len(th)))
    for q in range(len(th)):
        def J(m):
            t = th.copy()
            t[q] = th[q] * (1.0 + m * h)
            return window(t, n_steps, False)[0]
        d = th[q] * h
        if order == 2:
            G[:, q] = (J(1) - J(-1)) / (2 * d)
        else:
            G[:, q] = (8 * (J(1) - J(-1)) - (J(2) - J(-2))) / (12 * d)
    return G


def test_adjoint50k_full_window(adjoint50k):
    """The 200-step window: the adjoint within 1e-4 of the central
    difference at h = 1e-6.  At h = 1e-4 the difference itself is off by up
    to ~1e-3 over this window: every pair that crosses a cut-off (the LJ
    force and the DPD weight have kinks there; sigma moves every cut-off)
    adds a kink to J(theta), so the window's J is only piecewise smooth on
    that scale and neither the 3-point nor the 5-point stencil converges
    there.  Those errors are printed for the record; the shorter window below
    meets the bar at h = 1e-4."""
    th, window = adjoint50k
    _, G = window(th, 200, True)
    G2s = _fd(window, th, 200, 1e-6, 2)
    G2 = _fd(window, th, 200, 1e-4, 2)
    G4 = _fd(window, th, 200, 1e-4, 4)
    e2s, e2, e4 = _grad_err(G, G2s).max(), _grad_err(G, G2).max(), _grad_err(G, G4).max()
    print("adjoint50k W=200\nadjoint\n%s\n3-point h=1e-6\n%s\nmax rel err: 3-point h=1e-6 %.2e, "
          "3-point h=1e-4 %.2e, 5-point h=1e-4 %.2e" % (G, G2s, e2s, e2, e4), flush=True)
    assert e2s < 1e-4


def test_adjoint50k_short_window(adjoint50k):
    """A 40-step window: the adjoint within 1e-4 of the plain 3-point central
    difference at h = 1e-4."""
    th, window = adjoint50k
    _, G = window(th, 40, True)
    G2 = _fd(window, th, 40, 1e-4, 2)
    e2 = _grad_err(G, G2).max()
    print("adjoint50k W=40: max rel err vs 3-point h=1e-4 %.2e" % e2, flush=True)
    assert e2 < 1e-4
