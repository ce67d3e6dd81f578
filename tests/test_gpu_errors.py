"""Failure paths forced from the host (every kernel still exits normally: the
conditions are in the inputs, and the device code turns them into an error
flag that the C ABI returns as a status).

The reference prints and stops, or silently computes with the bad value
(SURVEY of init_system.f90:141-143, fluid_wall.f90:85-91 under NO_WARNS,
intra_molec.f90:36-44):
  - neighbour-list overflow (capacity too small)  -> MFA_ERR_LIST_OVERFLOW
  - more particles in a cell than its capacity     -> MFA_ERR_CELL_OVERFLOW
  - a particle outside the cell grid               -> MFA_ERR_CELL
  - two particles at one point, or on a wall       -> MFA_ERR_NAN
  - a particle slightly below the bottom wall and a bond stretched beyond R0
    compute finite forces as the reference does (NO_WARNS: no push-back;
    the FENE argument is clamped at 0.98), equal to the oracle's;
  - two ideal-solvent beads (epsilon 0, DPD forces only) 1e-3 apart, which
    such beads reach in long runs: finite forces equal to the oracle's.
After an error, a new state makes the context usable again.
"""
import numpy as np
import pytest

from mfa_amd import lib
from tests.util import force_error, gpu_ctx, make_case, oracle_sys

pytestmark = pytest.mark.gpu

ERR = {v: k for k, v in lib.ERRORS.items()}


def _expect(code, fn):
    with pytest.raises(lib.MfaError) as ei:
        fn()
    print(str(ei.value), flush=True)
    assert ei.value.code == ERR[code], str(ei.value)


def test_list_overflow():
    inputs, p, r, v, t, a = make_case("melt_small")
    print("list capacity 8 on melt_small", flush=True)
    with lib.Context(p, lib.MFA_MIXED, list_capacity=8) as c:
        c.set_state(r, v, a, t)
        _expect("MFA_ERR_LIST_OVERFLOW", c.binning)
        assert c.stats()["max_neigh"] > 8


@pytest.mark.parametrize("precision", [lib.MFA_FP64, lib.MFA_MIXED])
def test_cell_overflow_and_recovery(precision):
    inputs, p, r, v, t, a = make_case("melt_small")
    rng = np.random.default_rng(3)
    crowd = 0.1 + 0.8 * rng.random(r.shape) * np.array(p.r_bin)  # every bead in cell (0,0,0)
    print("all %d beads in one cell" % p.n_part, flush=True)
    with lib.Context(p, precision) as c:
        c.set_state(crowd, v, a, t)
        _expect("MFA_ERR_CELL_OVERFLOW", c.binning)
        c.set_state(r, v, a, t)  # a valid state clears the error
        c.energies(step=1)
        f = c.get_state()["f"]
    with lib.Context(p, precision) as c:
        c.set_state(r, v, a, t)
        c.energies(step=1)
        assert np.array_equal(f, c.get_state()["f"])


def test_particle_outside_the_grid():
    inputs, p, r, v, t, a = make_case("melt_small")
    r = r.copy()
    r[17, 0] = -0.6 * p.boundary[0]  # mod(int(x/r_bin), n) < 0: outside the grid
    with lib.Context(p, lib.MFA_FP64) as c:
        c.set_state(r, v, a, t)
        _expect("MFA_ERR_CELL", lambda: c.run_steps(1, 2))


@pytest.mark.parametrize("precision", [lib.MFA_FP64, lib.MFA_MIXED])
def test_coincident_particles_give_nan_status(precision):
    inputs, p, r, v, t, a = make_case("melt_small")
    r = r.copy()
    r[40] = r[300]
    with lib.Context(p, precision) as c:
        c.set_state(r, v, a, t)
        _expect("MFA_ERR_NAN", lambda: c.energies(step=1))
        _expect("MFA_ERR_NAN", lambda: c.run_steps(1, 3))


def test_particle_on_the_wall_gives_nan_status():
    inputs, p, r, v, t, a = make_case("brush_small")
    r = r.copy()
    r[-1, 2] = 0.0  # a solvent bead on the bottom wall: 1/z
    with lib.Context(p, lib.MFA_FP64) as c:
        c.set_state(r, v, a, t)
        _expect("MFA_ERR_NAN", lambda: c.energies(step=1))


def test_below_the_wall_and_stretched_bond_match_oracle():
    inputs, p, r, v, t, a = make_case("brush_small")
    r = r.copy()
    r[-1, 2] = -0.05                     # solvent bead just below z = 0
    r[5] = r[4] + np.array([0.0, 0.0, 1.6])  # bond 4-5 beyond R0 = 1.5
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    fref, eref = s.forces(step=2)
    assert np.all(np.isfinite(fref))
    ctx = gpu_ctx(p, r, v, a, t)
    obs = ctx.energies(step=2)
    f = ctx.get_state()["f"]
    ctx.close()
    per, _ = force_error(f, fref)
    print("below-wall bead and stretched bond: max err %.2e" % per, flush=True)
    assert per < 1e-12
    assert abs(obs["v_intra_molec"] - eref["v_intra_molec"]) <= 1e-12 * abs(eref["v_intra_molec"])


@pytest.mark.parametrize("precision", [lib.MFA_FP64, lib.MFA_MIXED])
def test_overlapping_ideal_solvent_beads_have_finite_forces(precision):
    """Ideal solvent beads interact only through DPD, so they may overlap;
    the LJ term with epsilon 0 is exactly zero in the reference's FP64.  In
    FP32, r^-12 r^-2 overflows below r ~ 0.002, so the mixed form must skip
    it rather than multiply by 0."""
    inputs, p, r, v, t, a = make_case("brush_small")
    solv = np.flatnonzero(t == 3)
    i, j = int(solv[0]), int(solv[1])
    r = r.copy()
    r[j] = r[i] + np.array([1e-3, 0.0, 0.0])
    s = oracle_sys(inputs, r, v, a, t)
    s.binning()
    fref, eref = s.forces(step=7)
    assert np.all(np.isfinite(fref))
    ctx = gpu_ctx(p, r, v, a, t, precision=precision)
    obs = ctx.energies(step=7)
    f = ctx.get_state()["f"]
    ctx.close()
    per, rms = force_error(f, fref)
    print("solvent pair at 1e-3: max %.2e rms %.2e" % (per, rms), flush=True)
    assert np.all(np.isfinite(f))
    if precision == lib.MFA_FP64:
        assert per < 1e-12
    else:
        assert rms < 1e-5 and per < 1e-3
    assert np.isfinite(obs["v_fluid_fluid"])
