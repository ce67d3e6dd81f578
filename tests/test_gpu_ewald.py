"""Ewald sum on the GPU (real space in the pair pass, k space and the slab
dipole correction as FP64 kernels) against the oracle and the Madelung
constant of rock salt."""
import numpy as np
import pytest

from mfa_amd import lib
from oracle import oracle as orc
from tests.test_oracle import _rock_salt
from tests.util import force_error, gpu_ctx, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [lib.MFA_FP64, lib.MFA_MIXED])
def test_madelung_rock_salt(precision):
    x, r, t = _rock_salt(8, 1.5, 1.96, 20, 3.0)
    p = lib.derive(x)
    print("rock salt", len(t), "ions, precision", precision, flush=True)
    ctx = gpu_ctx(p, r, np.zeros_like(r), np.zeros_like(r), t, precision=precision)
    obs = ctx.energies(step=1)
    ctx.close()
    expect = -(len(t) / 2) * 1.747564594633 / 1.5
    tol = 1e-6 if precision == lib.MFA_FP64 else 1e-5
    assert abs(obs["v_coul"] - expect) < tol * abs(expect), (obs["v_coul"], expect)


@pytest.mark.parametrize("precision", [lib.MFA_FP64, lib.MFA_MIXED])
def test_displaced_rock_salt_matches_oracle(precision):
    x, r, t = _rock_salt(8, 1.5, 1.44, 20, 3.0)
    rng = np.random.default_rng(5)
    r = r + rng.uniform(-0.2, 0.2, r.shape)
    s = orc.System(x)
    s.set_state(r, np.zeros_like(r), np.zeros_like(r), t)
    s.binning()
    fref, eref = s.forces(step=1)
    p = lib.derive(x)
    ctx = gpu_ctx(p, r, np.zeros_like(r), np.zeros_like(r), t, precision=precision)
    obs = ctx.energies(step=1)
    f = ctx.get_state()["f"]
    ctx.close()
    per, rms = force_error(f, fref)
    print("displaced rock salt precision", precision, "max", per, "rms", rms)
    if precision == lib.MFA_FP64:
        assert per < 1e-12 and rel(obs["v_coul"], eref["v_coul"]) < 1e-12
    else:
        assert rms < 1e-5 and rel(obs["v_coul"], eref["v_coul"]) < 1e-6


def test_time_ewald_keeps_forces():
    """mfa_time_ewald times the k-space stage and leaves F as it was; a
    system without charges is refused."""
    x, r, t = _rock_salt(8, 1.5, 1.44, 20, 3.0)
    r = r + np.random.default_rng(6).uniform(-0.2, 0.2, r.shape)
    p = lib.derive(x)
    print("time_ewald on displaced rock salt,", len(t), "ions", flush=True)
    ctx = gpu_ctx(p, r, np.zeros_like(r), np.zeros_like(r), t)
    ctx.energies(step=1)
    f0 = ctx.get_state()["f"].copy()
    for graph_steps in (0, 16):
        ctx.set_graph_steps(graph_steps)
        ms = ctx.time_ewald(20)
        assert 0.0 < ms < 100.0, ms
        assert np.array_equal(ctx.get_state()["f"], f0)
    ctx.close()
    y = lib.Inputs.from_buffer_copy(x)
    for q in range(4):
        y.charge[q] = 0.0
    ctx = gpu_ctx(lib.derive(y), r, np.zeros_like(r), np.zeros_like(r), t)
    with pytest.raises(lib.MfaError):
        ctx.time_ewald(5)
    ctx.close()


@pytest.mark.parametrize("n_kmax", list(range(1, 14)))
def test_k_space_every_row_width_matches_oracle(n_kmax):
    """The k-space force kernel is compiled once per row width nx_max (1 to
    12, the row kernel; 13 and wider take the segment kernel): a displaced
    rock salt with k_c = n_kmax dk runs each width, FP64 forces and v_coul
    within 1e-12 of the oracle."""
    x, r, t = _rock_salt(8, 1.5, 1.44, n_kmax, 3.0)
    r = r + np.random.default_rng(40 + n_kmax).uniform(-0.2, 0.2, r.shape)
    p = lib.derive(x)
    rc, plan, why = lib.ewald_plan(p, len(t))
    assert rc == 0, why
    assert plan["nx_max"] == n_kmax and plan["row_kernel"] == (1 if n_kmax <= 12 else 0), plan
    print("rock salt n_kmax", n_kmax, "k vectors", plan["k_vectors"], flush=True)
    s = orc.System(x)
    s.set_state(r, np.zeros_like(r), np.zeros_like(r), t)
    s.binning()
    fref, eref = s.forces(step=1)
    ctx = gpu_ctx(p, r, np.zeros_like(r), np.zeros_like(r), t, precision=lib.MFA_FP64)
    obs = ctx.energies(step=1)
    f = ctx.get_state()["f"]
    ctx.close()
    per, rms = force_error(f, fref)
    assert per < 1e-12 and rel(obs["v_coul"], eref["v_coul"]) < 1e-12, (per, obs["v_coul"], eref["v_coul"])
