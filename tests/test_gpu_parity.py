"""HIP path vs the oracle on identical seeded inputs, through the C ABI.

Bars: cell assignment and the unordered pair set bit-exact; plain FP64 forces
and energies within 1e-12 (per-particle |dF| / max(|F|, 1e-3 F_rms)); the
mixed FP32/fixed-point form within an RMS force error of 1e-5 of the FP64
result; trajectories of the FP64 form within 1e-9 after 40 steps.
"""
import numpy as np
import pytest

from mfa_amd import configs, lib
from tests.util import force_error, gpu_ctx, make_case, oracle_sys, rel

pytestmark = pytest.mark.gpu

CASES = ["brush_small", "melt_small", "melt_wca", "melt_switch_on", "pe_brush_compat",
         "pe_brush_small", "brush_bend"]


@pytest.mark.parametrize("case", CASES)
def test_cells_and_pairs_bit_exact(case):
    inputs, p, r, v, t, a = make_case(case)
    print("case", case, "n", p.n_part, flush=True)
    ctx = gpu_ctx(p, r, v, a, t)
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    ctx.binning()
    assert np.array_equal(ctx.get_cells(), s.get_cells())
    gp = ctx.get_pairs()
    gp = gp[np.lexsort((gp[:, 1], gp[:, 0]))]
    op = s.get_pairs()
    assert gp.shape == op.shape and np.array_equal(gp, op)
    ctx.close()


@pytest.mark.parametrize("case", CASES)
def test_forces_fp64(case):
    inputs, p, r, v, t, a = make_case(case)
    print("case", case, flush=True)
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    fref, eref = s.forces(step=7)
    ctx = gpu_ctx(p, r, v, a, t)
    obs = ctx.energies(step=7)
    f = ctx.get_state()["f"]
    per, rms = force_error(f, fref)
    assert per < 1e-12, (per, rms)
    for k in ("v_fluid_fluid", "v_intra_molec", "virial"):
        assert rel(obs[k], eref[k]) < 1e-12, (k, obs[k], eref[k])
    if p.symmetry == 0:
        assert rel(obs["v_fluid_wall"], eref["v_fluid_wall"]) < 1e-12
        assert rel(obs["wall_fz_bottom"], eref["wall_fz_bottom"]) < 1e-12
    if p.coulomb:
        assert rel(obs["v_coul"], eref["v_coul"]) < 1e-12, (obs["v_coul"], eref["v_coul"])
    if p.bending:
        assert rel(obs["v_bend"], eref["v_bend"]) < 1e-12, (obs["v_bend"], eref["v_bend"])
    ctx.close()


@pytest.mark.parametrize("case", CASES)
def test_subroutine_entry_points_fp64(case):
    """The per-subroutine sequence of md_main.f90:88-106 (force = 0,
    fluid_fluid, fluid_wall, intra_molec) through the C ABI."""
    inputs, p, r, v, t, a = make_case(case)
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    fref, eref = s.forces(step=3)
    ctx = gpu_ctx(p, r, v, a, t)
    ctx.binning()
    ctx.zero_force()
    vff, vc = ctx.fluid_fluid(3)
    vfw = ctx.fluid_wall(2)
    vin = ctx.intra_molec()
    vb = ctx.bending(2)  # md_main.f90:108-111
    if p.bending:
        assert rel(vb, eref["v_bend"]) < 1e-12
    if p.coulomb:  # md_main.f90:128-132
        vc += ctx.ewald_k(1)
        if p.symmetry == 0:
            vc += ctx.dipolar_correction()
        assert rel(vc, eref["v_coul"]) < 1e-12, (vc, eref["v_coul"])
    f = ctx.get_state()["f"]
    per, _ = force_error(f, fref)
    assert per < 1e-12
    assert rel(vff, eref["v_fluid_fluid"]) < 1e-12
    assert rel(vin, eref["v_intra_molec"]) < 1e-12
    if p.symmetry == 0:
        assert rel(vfw, eref["v_fluid_wall"]) < 1e-12
    ctx.close()


@pytest.mark.parametrize("case", CASES)
def test_forces_mixed(case):
    inputs, p, r, v, t, a = make_case(case)
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    fref, eref = s.forces(step=5)
    ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
    obs = ctx.energies(step=5)
    f = ctx.get_state()["f"]
    per, rms = force_error(f, fref)
    print(case, "mixed max", per, "rms", rms)
    assert rms < 1e-5, (per, rms)
    assert per < 1e-3, (per, rms)
    assert rel(obs["v_fluid_fluid"], eref["v_fluid_fluid"]) < 1e-5
    ctx.close()


@pytest.mark.parametrize("case", ["brush_small", "brush_bend"])
def test_forces_mixed_switch_on_dpd_cut_off(case):
    """A DPD_CUT_OFF brush at a fresh start (FORCE_SWITCH_ON): early in the
    ramp the clamped r^2 of every pair inside the LJ range exceeds the DPD
    range, so those pairs have no DPD term (fluid_fluid.f90:94-96,201); the
    mixed form must decide as the reference (and the FP64 form) does."""
    inputs, p, r, v, t, a = make_case(case)
    inputs.s_time = 1
    inputs.force_switch_on = 1
    p = lib.derive(inputs)
    assert p.dpd_cut_off_on
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    for step in (3, 2000):  # clamp above the DPD range, then below it
        fref, _ = s.forces(step=step)
        ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
        ctx.energies(step=step)
        per, rms = force_error(ctx.get_state()["f"], fref)
        ctx.close()
        print("%s step %d: mixed max %.2e rms %.2e" % (case, step, per, rms), flush=True)
        assert rms < 1e-5, (step, per, rms)


@pytest.mark.parametrize("precision", [lib.MFA_FP64, lib.MFA_MIXED])
def test_mixed_threads_per_particle_invariant(precision):
    """Fixed-point sums do not depend on how the list is split across lanes."""
    inputs, p, r, v, t, a = make_case("melt_small")
    out = []
    for tpp in (1, 2, 4, 8, 16):
        ctx = gpu_ctx(p, r, v, a, t, precision=precision, tpp=tpp)
        ctx.energies(step=1)
        out.append(ctx.get_state()["f"])
        ctx.close()
    for f in out[1:]:
        if precision == lib.MFA_MIXED:
            assert np.array_equal(f, out[0])
        else:
            assert force_error(f, out[0])[0] < 1e-13


def test_mixed_momentum_exact():
    """Pair and bond forces are exactly antisymmetric in fixed point, so the
    total force of a periodic melt is exactly zero (DPD included)."""
    inputs, p, r, v, t, a = make_case("melt_small")
    ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
    ctx.energies(step=4)
    f = ctx.get_state()["f"]
    assert np.all(f.sum(axis=0) == 0.0)
    ctx.close()


@pytest.mark.parametrize("case", CASES)
def test_trajectory_fp64(case):
    inputs, p, r, v, t, a = make_case(case)
    s = oracle_sys(inputs, r, v, a, t)
    ref = s.run_steps(first_step=1, n=40)
    st_ref = s.get_state()
    ctx = gpu_ctx(p, r, v, a, t)
    obs = ctx.run_steps(first_step=1, n=40)
    st = ctx.get_state()
    dr = np.abs(st["r"] - st_ref["r"]).max()
    dv = np.abs(st["v"] - st_ref["v"]).max()
    print(case, "dr", dr, "dv", dv, "rebuilds", obs["n_rebuilds"], ref["n_rebuilds"])
    assert np.array_equal(st["mic"], st_ref["mic"])
    assert dr < 1e-9 and dv < 1e-8
    assert obs["n_rebuilds"] == ref["n_rebuilds"]
    for k in ("v_fluid_fluid", "v_intra_molec", "t_kin"):
        assert rel(obs[k], ref[k]) < 1e-9, (k, obs[k], ref[k])
    if p.coulomb:
        assert rel(obs["v_coul"], ref["v_coul"]) < 1e-9, (obs["v_coul"], ref["v_coul"])
    if p.bending:
        assert rel(obs["v_bend"], ref["v_bend"]) < 1e-9, (obs["v_bend"], ref["v_bend"])
    ctx.close()


def test_run_steps_bitwise_reproducible():
    inputs, p, r, v, t, a = make_case("brush_small")
    res = []
    for _ in range(2):
        ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
        ctx.run_steps(1, 30)
        res.append(ctx.get_state())
        ctx.close()
    for k in ("r", "v", "f"):
        assert np.array_equal(res[0][k], res[1][k])


def test_compat_loop_matches_oracle():
    """The driver's call sequence (md_main.f90:73-147) via the per-subroutine
    entry points, 10 steps, against the oracle's loop."""
    inputs, p, r, v, t, a = make_case("brush_small")
    s = oracle_sys(inputs, r, v, a, t)
    ref = s.run_steps(first_step=1, n=10)
    st_ref = s.get_state()
    ctx = gpu_ctx(p, r, v, a, t)
    ctx.binning()
    for it in range(10):
        ctx.verlet_positions()
        if ctx.check_skin():
            ctx.binning()
        ctx.zero_force()
        ctx.fluid_fluid(1 + it)
        ctx.fluid_wall(2)
        ctx.intra_molec()
        tk = ctx.verlet_velocities()
    st = ctx.get_state()
    assert np.abs(st["r"] - st_ref["r"]).max() < 1e-10
    assert rel(tk, ref["t_kin"]) < 1e-10
    ctx.close()


@pytest.mark.parametrize("precision", [lib.MFA_FP64, lib.MFA_MIXED])
def test_graph_replays_match_direct_launches(precision):
    """Captured step blocks give bit-identical results to direct launches; the
    graph is replayed many times with eager work and synchronisation between
    the replays, and every chunk is compared."""
    inputs, p, r, v, t, a = make_case("melt_small")
    eager = gpu_ctx(p, r, v, a, t, precision=precision)
    eager.set_graph_steps(0)
    graph = gpu_ctx(p, r, v, a, t, precision=precision)
    graph.set_graph_steps(4)
    step = 1
    for chunk in range(5):
        oe = eager.run_steps(step, 13)
        og = graph.run_steps(step, 13)
        step += 13
        se, sg = eager.get_state(), graph.get_state()
        for k in ("r", "v", "f", "a", "mic"):
            assert np.array_equal(se[k], sg[k]), (chunk, k)
        assert oe["n_rebuilds"] == og["n_rebuilds"]
        assert oe["v_fluid_fluid"] == og["v_fluid_fluid"]
    eager.close()
    graph.close()


def test_pairs_bit_exact_melt100k():
    """Full-size pair set (100k beads, 16^3 cells of about 24 and up to ~45
    particles) against the oracle's binning."""
    from mfa_amd import configs
    from oracle import oracle as orc
    w = configs.WORKLOADS["melt100k"]
    inputs = w["make"]()
    p, r, v, t = configs.build(inputs, w["kind"], min_dist=w["min_dist"])
    a = np.zeros_like(r)
    print("melt100k pairs", flush=True)
    ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
    ctx.binning()
    gp = ctx.get_pairs()
    gp = gp[np.lexsort((gp[:, 1], gp[:, 0]))]
    cells = ctx.get_cells()
    ctx.close()
    s = orc.System(inputs)
    s.set_state(r, v, a, t)
    s.binning()
    assert np.array_equal(cells, s.get_cells())
    op = s.get_pairs()
    assert gp.shape == op.shape and np.array_equal(gp, op)


def _vv_case(name):
    """A small case with DPD_VV switched on (new_dpd_fd after each
    verlet_velocities, md_main.f90:149-155)."""
    kind, inputs = configs.small_cases()[name]
    inputs.dpd_vv = 1
    p, r, v, t = configs.build(inputs, kind)
    return inputs, p, r, v, t, np.zeros_like(r)


@pytest.mark.parametrize("case", ["brush_small", "brush_bend"])
def test_dpd_vv_trajectory_fp64(case):
    inputs, p, r, v, t, a = _vv_case(case)
    s = oracle_sys(inputs, r, v, a, t)
    ref = s.run_steps(first_step=1, n=40)
    st_ref = s.get_state()
    ctx = gpu_ctx(p, r, v, a, t)
    obs = ctx.run_steps(first_step=1, n=40)
    st = ctx.get_state()
    dr = np.abs(st["r"] - st_ref["r"]).max()
    dv = np.abs(st["v"] - st_ref["v"]).max()
    print(case, "dr", dr, "dv", dv, "rebuilds", obs["n_rebuilds"], ref["n_rebuilds"])
    assert np.array_equal(st["mic"], st_ref["mic"])
    assert dr < 1e-9 and dv < 1e-8
    assert np.abs(st["a"] - st_ref["a"]).max() < 1e-8
    assert obs["n_rebuilds"] == ref["n_rebuilds"]
    for k in ("v_fluid_fluid", "v_intra_molec", "t_kin"):
        assert rel(obs[k], ref[k]) < 1e-9, (k, obs[k], ref[k])
    ctx.close()


def test_dpd_vv_entry_point_loop_matches_oracle():
    """The driver's call sequence with DPD_VV (md_main.f90:73-155): the
    per-subroutine entry points and new_dpd_fd, 10 steps."""
    inputs, p, r, v, t, a = _vv_case("brush_small")
    s = oracle_sys(inputs, r, v, a, t)
    ref = s.run_steps(first_step=1, n=10)
    st_ref = s.get_state()
    ctx = gpu_ctx(p, r, v, a, t)
    ctx.binning()
    for it in range(10):
        ctx.verlet_positions()
        if ctx.check_skin():
            ctx.binning()
        ctx.zero_force()
        ctx.fluid_fluid(1 + it)
        ctx.fluid_wall(2)
        ctx.intra_molec()
        ctx.verlet_velocities()
        tk = ctx.new_dpd_fd()
    st = ctx.get_state()
    assert np.abs(st["r"] - st_ref["r"]).max() < 1e-10
    assert np.abs(st["v"] - st_ref["v"]).max() < 1e-9
    assert rel(tk, ref["t_kin"]) < 1e-10
    ctx.close()


def test_new_dpd_fd_refused_without_dpd_vv():
    inputs, p, r, v, t, a = make_case("brush_small")
    ctx = gpu_ctx(p, r, v, a, t)
    with pytest.raises(lib.MfaError):
        ctx.new_dpd_fd()
    ctx.close()


def test_dpd_vv_graph_replays_match_direct_launches():
    inputs, p, r, v, t, a = _vv_case("brush_small")
    eager = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
    eager.set_graph_steps(0)
    graph = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
    graph.set_graph_steps(4)
    step = 1
    for chunk in range(4):
        oe = eager.run_steps(step, 11)
        og = graph.run_steps(step, 11)
        step += 11
        se, sg = eager.get_state(), graph.get_state()
        for k in ("r", "v", "f", "a", "mic"):
            assert np.array_equal(se[k], sg[k]), (chunk, k)
        assert oe["t_kin"] == og["t_kin"]
    eager.close()
    graph.close()


def _lgv_case(name):
    """A small case with the Langevin thermostat (THERMOSTAT 1)."""
    kind, inputs = configs.small_cases()[name]
    inputs.thermostat = 1
    p, r, v, t = configs.build(inputs, kind)
    return inputs, p, r, v, t, np.zeros_like(r)


@pytest.mark.parametrize("case", ["brush_small", "melt_small"])
def test_langevin_trajectory_fp64(case):
    """Langevin force (lgv_forces.f90:25-43) with the same counter deviates as
    the oracle: FP64 trajectories within 1e-9 after 40 steps."""
    inputs, p, r, v, t, a = _lgv_case(case)
    s = oracle_sys(inputs, r, v, a, t)
    ref = s.run_steps(first_step=1, n=40)
    st_ref = s.get_state()
    ctx = gpu_ctx(p, r, v, a, t)
    obs = ctx.run_steps(first_step=1, n=40)
    st = ctx.get_state()
    dr = np.abs(st["r"] - st_ref["r"]).max()
    dv = np.abs(st["v"] - st_ref["v"]).max()
    print(case, "dr", dr, "dv", dv)
    assert np.array_equal(st["mic"], st_ref["mic"])
    assert dr < 1e-9 and dv < 1e-8
    assert obs["n_rebuilds"] == ref["n_rebuilds"]
    for k in ("v_fluid_fluid", "v_intra_molec", "t_kin"):
        assert rel(obs[k], ref[k]) < 1e-9, (k, obs[k], ref[k])
    ctx.close()


def test_langevin_forces_mixed():
    inputs, p, r, v, t, a = _lgv_case("melt_small")
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    fref, _ = s.forces(step=5)
    ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
    ctx.energies(step=5)
    f = ctx.get_state()["f"]
    per, rms = force_error(f, fref)
    print("langevin mixed: max", per, "rms", rms)
    assert rms < 1e-5 and per < 1e-3
    ctx.close()


def test_langevin_holds_the_temperature():
    """The thermostat's fluctuation-dissipation balance: a melt started at
    twice the target temperature relaxes to it (mean kinetic temperature over
    the last 1500 of 2500 steps within 3% of T)."""
    inputs, p, r, v, t, a = _lgv_case("melt_small")
    ctx = gpu_ctx(p, r, 2.0 ** 0.5 * v, a, t, precision=lib.MFA_MIXED)
    step, temps = 1, []
    dof = 3.0 * p.n_part
    for chunk in range(50):
        obs = ctx.run_steps(step, 50)
        step += 50
        if chunk >= 20:
            temps.append(2.0 * obs["t_kin"] / dof)
    tm = float(np.mean(temps))
    print("Langevin: T target %.3f measured %.4f" % (inputs.temp, tm))
    assert abs(tm / inputs.temp - 1.0) < 0.03
    ctx.close()


@pytest.mark.parametrize("precision", [lib.MFA_MIXED, lib.MFA_FP64])
@pytest.mark.parametrize("case", ["melt_small", "brush_small", "pe_brush_small", "pe_brush_compat"])
def test_fused_step_matches_separate_kernels(case, precision):
    """The step whose force pass also integrates (either form;
    double-buffered positions and view; for charged systems with the k-space
    force of the step's positions added before the integration) against the same steps
    with the integration in its own kernel and the k-space force added by
    k_ewald_apply (kernel timing on switches the fusion off): bit-identical
    state and the same rebuilds, over 60 steps with several rebuilds."""
    inputs, p, r, v, t, a = make_case(case)
    fused = gpu_ctx(p, r, v, a, t, precision=precision)
    split = gpu_ctx(p, r, v, a, t, precision=precision)
    split.timing(True)
    of = fused.run_steps(1, 60)
    os_ = split.run_steps(1, 60)
    sf, ss = fused.get_state(), split.get_state()
    print("%s: %d rebuilds" % (case, of["n_rebuilds"]), flush=True)
    assert of["n_rebuilds"] == os_["n_rebuilds"] and of["n_rebuilds"] >= 2
    for k in ("r", "v", "f", "a", "mic"):
        assert np.array_equal(sf[k], ss[k]), k
    for k in ("v_fluid_fluid", "v_intra_molec", "t_kin"):
        assert of[k] == os_[k], k
    fused.close()
    split.close()


@pytest.mark.parametrize("tpp", [1, 2, 4, 16])
def test_fused_charged_step_every_width(tpp):
    """The fused step of a charged system at each lane-group width: the
    group sums the k-space force partials in the fixed order k_ewald_apply
    uses (each width has its own way of sharing the strands), so the state
    equals the separate kernels' bit for bit."""
    inputs, p, r, v, t, a = make_case("pe_brush_small")
    fused = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED, tpp=tpp)
    split = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED, tpp=tpp)
    split.timing(True)
    of = fused.run_steps(1, 30)
    os_ = split.run_steps(1, 30)
    sf, ss = fused.get_state(), split.get_state()
    assert of["n_rebuilds"] == os_["n_rebuilds"]
    for k in ("r", "v", "f", "a", "mic"):
        assert np.array_equal(sf[k], ss[k]), (tpp, k)
    fused.close()
    split.close()


def test_time_force_step_leaves_the_state():
    """mfa_time_force_step times the fused force kernel without changing the
    state: the steps that follow equal those of an untimed context."""
    inputs, p, r, v, t, a = make_case("melt_small")
    timed = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
    plain = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
    timed.run_steps(1, 20)
    plain.run_steps(1, 20)
    before = timed.get_state()
    ms = timed.time_force_step(10)
    after = timed.get_state()
    print("fused force pass %.4f ms" % ms, flush=True)
    assert ms > 0.0
    for k in ("r", "v", "f", "a", "mic"):
        assert np.array_equal(before[k], after[k]), k
    ot = timed.run_steps(21, 20)
    op = plain.run_steps(21, 20)
    st, sp = timed.get_state(), plain.get_state()
    for k in ("r", "v", "mic"):
        assert np.array_equal(st[k], sp[k]), k
    assert ot["n_rebuilds"] == op["n_rebuilds"]
    timed.close()
    plain.close()


@pytest.mark.parametrize("case", ["melt_small", "brush_small"])
def test_set_skin_pairs_and_trajectory(case):
    """mfa_set_skin on a running context: the pair set equals the oracle's
    at the new skin, bit for bit, and an FP64 trajectory that changes skin
    midway stays within 1e-9 of the oracle's at the old skin (the forces do
    not depend on the skin; only the list's order and length do)."""
    inputs, p, r, v, t, a = make_case(case)
    skin2 = 0.45
    y = lib.Inputs.from_buffer_copy(inputs)
    y.skin = skin2
    s2 = oracle_sys(y, r, v, a, t)
    s2.binning()
    ref_pairs = s2.get_pairs()
    ctx = gpu_ctx(p, r, v, a, t)
    ctx.set_skin(skin2)
    ctx.binning()
    gp = ctx.get_pairs()
    gp = gp[np.lexsort((gp[:, 1], gp[:, 0]))]
    assert gp.shape == ref_pairs.shape and np.array_equal(gp, ref_pairs)
    ctx.close()
    s = oracle_sys(inputs, r, v, a, t)
    ref = s.run_steps(first_step=1, n=30)
    st_ref = s.get_state()
    ctx = gpu_ctx(p, r, v, a, t)
    ctx.run_steps(1, 10)
    ctx.set_skin(skin2)
    obs = ctx.run_steps(11, 20)
    st = ctx.get_state()
    dr = np.abs(st["r"] - st_ref["r"]).max()
    print("%s: skin %.2f -> %.2f, dr %.2e" % (case, p.skin, skin2, dr), flush=True)
    assert dr < 1e-9 and np.array_equal(st["mic"], st_ref["mic"])
    assert rel(obs["t_kin"], ref["t_kin"]) < 1e-9
    ctx.close()


def test_autotune_skin_keeps_a_candidate():
    """The skin autotuner keeps one of the candidates that fit the cells
    (a candidate wider than the cells is skipped) and the context runs on."""
    inputs, p, r, v, t, a = make_case("melt_small")
    ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED)
    cand = [0.3, 0.4, 10.0]
    skin, ms = ctx.autotune_skin(1, 40, cand)
    print("autotune_skin: %.2f (%.4f ms per step)" % (skin, ms), flush=True)
    assert skin in cand[:2] and ms > 0.0
    first = 1 + 2 * (40 + 17)
    obs = ctx.run_steps(first, 20)
    assert np.isfinite(obs["t_kin"]) and obs["t_kin"] > 0.0
    ctx.close()


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("tpp", [1, 4, 16])
@pytest.mark.parametrize("case", ["brush_small", "melt_small", "melt_switch_on", "pe_brush_small"])
def test_pair_queue_forces_bit_identical(case, tpp, mode):
    """The mixed pass with the wave's interacting pairs queued in LDS (mode
    1: every entry; mode 2: near/shell rows, the skin part queued) gives the
    direct walk's forces bit for bit (the same fixed-point terms, summed
    exactly), and the energies to rounding; forces without energies and
    trajectories of 40 steps with rebuilds stay bit-identical.  (Charged
    systems have no queue: the mode is ignored there.)"""
    inputs, p, r, v, t, a = make_case(case)
    out = {}
    for q in (0, mode):
        ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED, tpp=tpp)
        ctx.set_pair_queue(q)
        assert ctx.tuning()["pair_queue"] == q
        e = ctx.energies(step=3)
        f = ctx.get_state()["f"]
        ctx.time_force(3)  # the passes without energies, at the same step and list
        fp = ctx.get_state()["f"]
        if not p.coulomb:
            assert np.array_equal(f, fp)
        obs = ctx.run_steps(4, 40)
        out[q] = (f, e, ctx.get_state(), obs, fp)
        ctx.close()
    (f0, e0, s0, o0, g0), (f1, e1, s1, o1, g1) = out[0], out[mode]
    print("%s tpp %d mode %d: %d rebuilds" % (case, tpp, mode, o0["n_rebuilds"]), flush=True)
    assert np.array_equal(f0, f1) and np.array_equal(g0, g1)
    for k in ("v_fluid_fluid", "virial", "v_coul"):
        assert rel(e0[k], e1[k]) < 1e-12 or abs(e0[k] - e1[k]) < 1e-12, (k, e0[k], e1[k])
    for k in ("r", "v", "f", "a", "mic"):
        assert np.array_equal(s0[k], s1[k]), k
    assert o0["n_rebuilds"] == o1["n_rebuilds"]


def test_near_shell_queue_melt100k():
    """The headline melt with near/shell rows and the skin part queued: one
    force pass bit-identical to the direct walk on plain rows, and within
    the mixed form's bar of the oracle."""
    w = configs.WORKLOADS["melt100k"]
    inputs = w["make"]()
    p, r, v, t = configs.build(inputs, w["kind"], min_dist=w["min_dist"])
    a = np.zeros_like(r)
    out = {}
    for mode in (0, 2):
        ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED, tpp=8)
        ctx.set_pair_queue(mode)
        ctx.energies(step=9)  # the step of the passes below
        ms = ctx.time_force(3)
        out[mode] = ctx.get_state()["f"]
        print("melt100k mode %d: force pass %.4f ms" % (mode, ms), flush=True)
        ctx.close()
    assert np.array_equal(out[0], out[2])
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    fref, _ = s.forces(step=9)
    per, rms = force_error(out[2], fref)
    assert rms < 1e-5, (per, rms)


def test_pair_queue_matches_oracle_brush9k():
    """Full-size brush (the case the queue is for) against the oracle."""
    w = configs.WORKLOADS["brush9k"]
    inputs = w["make"]()
    p, r, v, t = configs.build(inputs, w["kind"], min_dist=w["min_dist"])
    a = np.zeros_like(r)
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    fref, _ = s.forces(step=5)
    ctx = gpu_ctx(p, r, v, a, t, precision=lib.MFA_MIXED, tpp=4)
    ctx.set_pair_queue(True)
    ctx.energies(step=5)
    f = ctx.get_state()["f"]
    ctx.close()
    per, rms = force_error(f, fref)
    print("brush9k queued: max %.2e rms %.2e" % (per, rms), flush=True)
    assert rms < 1e-5
