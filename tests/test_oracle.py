"""CPU tests: pin the oracle against known answers and mathematics, and check
the host side of the library (ABI, derived constants).  No GPU needed.

The reference ships no tests, golden vectors or fixtures for this path, so the
oracle is pinned by:
  - Random123's published known-answer vectors for Threefry-2x32;
  - hand-evaluated reference expressions with float32 literals;
  - brute force (O(N^2) pair criterion) for the cell scan;
  - closed forms (LJ, FENE, 9-3 wall, free flight, constant force);
  - energy/force consistency by central finite differences;
  - invariants (Newton's third law, momentum, pinned heads).
Each check is built so that a plausible slip (dropped term, wrong sign or
index, transposed operand, float vs double literal) fails it.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mfa_amd import configs, lib
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C6 = float(np.float32(2.0) ** np.float32(np.float32(1.0) / np.float32(6.0)))  # 2.**(1./6.) in real(4)


# ---------------------------------------------------------------- RNG pins

# Random123 kat_vectors: threefry2x32 R ctr0 ctr1 key0 key1 -> out0 out1
THREEFRY_KAT = [
    (13, (0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x9d1c5ec6, 0x8bd50731)),
    (13, (0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xfd36d048, 0x2d17272c)),
    (13, (0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xba3e4725, 0xf27d669e)),
    (20, (0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x6b200159, 0x99ba4efe)),
    (20, (0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0x1cb996fc, 0xbb002be7)),
    (20, (0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xc4923a9c, 0x483df7a0)),
]


@pytest.mark.parametrize("rounds,inp,out", THREEFRY_KAT)
def test_threefry_known_answers(rounds, inp, out):
    assert orc.threefry2x32(*inp, rounds=rounds) == out


def test_pair_normal_moments_and_symmetry():
    xs = np.array([orc.pair_normal(1234, 7, i, i + 1 + (i % 13)) for i in range(200000)])
    assert abs(xs.mean()) < 0.01
    assert abs(xs.var() - 1.0) < 0.01
    k = ((xs - xs.mean()) ** 4).mean() / xs.var() ** 2
    assert abs(k - 3.0) < 0.05
    # same number from either side of the pair; new numbers per step
    assert orc.pair_normal(5, 3, 10, 20) == orc.pair_normal(5, 3, 20, 10)
    assert orc.pair_normal(5, 3, 10, 20) != orc.pair_normal(5, 4, 10, 20)


def test_particle_normal_moments_and_independence():
    """The Langevin deviates (one per particle and component): unit normal
    moments, no correlation between components, and a stream apart from the
    pair deviates (counter word 1 has its top bit set)."""
    n = 70000
    xs = np.array([[orc.particle_normal(1234, 7, i, d) for d in range(3)] for i in range(n)])
    for d in range(3):
        x = xs[:, d]
        assert abs(x.mean()) < 0.015
        assert abs(x.var() - 1.0) < 0.015
        k = ((x - x.mean()) ** 4).mean() / x.var() ** 2
        assert abs(k - 3.0) < 0.08
    c = np.corrcoef(xs.T)
    assert np.abs(c - np.eye(3)).max() < 0.015
    # new numbers per step and per seed; apart from the pair stream
    assert orc.particle_normal(5, 3, 10, 0) != orc.particle_normal(5, 4, 10, 0)
    assert orc.particle_normal(5, 3, 10, 0) != orc.particle_normal(6, 3, 10, 0)
    w = orc.threefry2x32(10, 0x80000000, 5, 3, rounds=13)
    u1 = (w[0] >> 8) * 2.0 ** -24 + 2.0 ** -25
    u2 = (w[1] >> 8) * 2.0 ** -24
    assert abs(orc.particle_normal(5, 3, 10, 0) -
               np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)) < 1e-14


# ---------------------------------------------------------------- constants

def test_derived_constants_float32_literals():
    inp = configs.brush(8, 10, 300, 6.0, 6.0, 8.0, x_space=0.5, skin=0.0)
    p = orc.derive(inp)
    # WCA range from the float32 2**(1/6) (init_params.f90:502), squared in FP64
    assert C6 == 1.1224620342254639
    assert p.range_2[1][1] == C6 * C6
    assert p.range_2[0][0] == 0.0            # head-head (init_params.f90:703-712)
    assert p.range_2[1][3] == 0.0            # no PARTICLE_4 (init_params.f90:604-607)
    assert p.skin == float(np.float32(0.4))  # skin = 0.4 (init_params.f90:800)
    x = (1.0 / (C6 * C6)) ** 3
    assert abs(p.e_shift[1][1] - x * (x - 1.0)) < 1e-15
    assert abs(p.e_shift[1][1] + 0.25) < 1e-6  # WCA shift is -1/4 (x^3 ~ 1/2)
    assert p.epsil[1][1] == 4.0
    assert p.k_chain == 30.0 * 1.5 ** 2       # k_chain*R0^2 (init_params.f90:398)
    # DPD cut-off macro literal and r_cut_max (init_params.f90:770-781)
    assert p.r_cut_dpd == 1.0 and p.r_cut_max == C6
    assert p.skin_2 == (C6 + p.skin) ** 2 - C6 ** 2
    # binning grid: int(L/(r_cut_max + 0.4f)) (make_binning_boxes.f90:44-46)
    assert list(p.n_bin) == [int(6.0 / (C6 + float(np.float32(0.4)))),
                             int(6.0 / (C6 + float(np.float32(0.4)))),
                             int(8.0 / (C6 + float(np.float32(0.4))))]
    assert abs(p.sig - np.sqrt(2 * 1.0 * 4.5 / 0.005)) < 1e-12


def test_poor_solvent_and_particle4_ranges():
    inp = configs.brush(8, 10, 300, 6.0, 6.0, 8.0, x_space=0.5)
    inp.solvent = 0
    p = orc.derive(inp)
    c2 = float(np.float32(2.0) * np.float32(C6))
    assert p.range_2[1][2] == c2 * c2 and p.range_2[3][1] == 0.0
    inp.solvent = 1
    inp.particle_4 = 1
    inp.inter4 = 2
    p = orc.derive(inp)
    assert p.range_2[3][3] == (2.0 * C6) ** 2 and p.range_2[1][3] == C6 * C6


@pytest.mark.parametrize("case", ["brush_small", "melt_small", "melt_wca", "melt_switch_on"])
def test_library_derive_matches_oracle(case):
    """Host-side derivation of libmfa_hip (mfa_derive) vs the oracle's."""
    kind, inp = configs.small_cases()[case]
    p = lib.derive(inp)
    q = orc.derive(inp)
    for f in ("range_2", "inv_range_2", "e_shift", "sigma_2", "epsil"):
        assert np.array_equal(np.array(getattr(p, f)), np.array(getattr(q, f))), f
    for f in ("r_cut_max", "skin", "skin_2", "sig", "k_chain", "inv_r_chain_2", "dt_2",
              "r_cut_dpd_2", "friction1", "z_space_wall", "a_w"):
        assert getattr(p, f) == getattr(q, f), f
    assert list(p.n_bin) == list(q.n_bin) and list(p.r_bin) == list(q.r_bin)
    assert list(p.sigma_wall) == list(q.sigma_wall) and list(p.a_wall) == list(q.a_wall)
    assert p.fso_r2_init == q.r_2_min_init and p.fso_time == q.r_2_min_time
    assert p.n_part == q.n_part and p.part_init_d == q.part_init_d


def test_derive_refuses_grids_that_miss_pairs():
    inp = configs.brush(8, 10, 300, 6.0, 6.0, 8.0, x_space=1.0)  # int(0.4*6)=2 -> delta_bin 0
    with pytest.raises(lib.MfaError):
        lib.derive(inp)


# ---------------------------------------------------------------- cell scan

@pytest.mark.parametrize("case", ["brush_small", "melt_small", "melt_wca"])
def test_oracle_pair_set_equals_brute_force(case):
    kind, inp = configs.small_cases()[case]
    p, r, v, t = configs.build(inp, kind)
    s = orc.System(inp)
    s.set_state(r, v, np.zeros_like(r), t)
    s.binning()
    assert np.array_equal(s.get_pairs(), s.brute_pairs())
    # cells by the closed form mod(int(r*inv_r_bin), n_bin)
    cells = s.get_cells()
    q = s.p
    exp = np.stack([np.mod((r[:, d] * q.inv_r_bin[d]).astype(np.int64), q.n_bin[d])
                    for d in range(3)], axis=1)
    assert np.array_equal(cells, exp)


# ---------------------------------------------------------------- closed forms

def _two_body(r12, symmetry=1, types=(3, 3), thermostat=-1, lj_rc=2.5, box=12.0):
    inp = configs.melt(1, 2, 0.001, lj_rc=lj_rc, thermostat=thermostat, n_cell=16)
    inp.x_space = inp.y_space = box / 16
    inp.z_space_wall = box
    inp.n_mon_d, inp.n_chain_d = 1, 2  # two free beads, no bond
    s = orc.System(inp)
    r = np.array([[5.0, 5.0, 5.0], [5.0 + r12, 5.0, 5.0]])
    z = np.zeros((2, 3))
    s.set_state(r, z, z, np.array(types, dtype=np.int32))
    return s


@pytest.mark.parametrize("rr", [0.95, 1.0, 1.12, 1.5, 2.2, 2.49])
def test_lj_closed_form(rr):
    s = _two_body(rr)
    s.binning()
    f, e = s.forces(step=0)
    fmag = 24.0 * (2.0 / rr ** 13 - 1.0 / rr ** 7)
    assert abs(f[1, 0] - fmag) < 1e-12 * max(1.0, abs(fmag))
    assert f[0, 0] == -f[1, 0]
    shift = 4.0 * (1 / 2.5 ** 12 - 1 / 2.5 ** 6)
    ev = 4.0 * (1 / rr ** 12 - 1 / rr ** 6) - shift
    assert abs(e["v_fluid_fluid"] - ev) < 1e-12


def test_lj_outside_cutoff_is_zero():
    s = _two_body(2.51)
    s.binning()
    f, e = s.forces(step=0)
    assert np.all(f == 0.0) and e["v_fluid_fluid"] == 0.0


def test_fene_closed_form():
    inp = configs.melt(1, 2, 0.001, lj_rc=2.5, thermostat=-1, n_cell=16)
    inp.x_space = inp.y_space = 12.0 / 16
    inp.z_space_wall = 12.0
    for t_eps in range(4):
        for u in range(4):
            inp.epsil[t_eps][u] = 0.0
    s = orc.System(inp)
    rr = 1.1
    r = np.array([[5.0, 5.0, 5.0], [5.0, 5.0 + rr, 5.0]])
    z = np.zeros((2, 3))
    s.set_state(r, z, z, np.array([3, 3], dtype=np.int32))
    s.binning()
    f, e = s.forces(step=0)
    k, R0 = 30.0, 1.5
    assert abs(f[0, 1] - k * rr / (1 - (rr / R0) ** 2)) < 1e-11   # pulled toward bead 2
    assert abs(e["v_intra_molec"] + 0.5 * k * R0 ** 2 * np.log(1 - (rr / R0) ** 2)) < 1e-11


# ---------------------------------------------------------------- virial
# The virial W = sum over LJ pairs and FENE bonds of r_ij . F_ij (F_ij the
# force on i from j, r_ij = r_i - r_j) = -sum r dU/dr: conservative pair and
# bond terms only (no DPD, Coulomb, wall or bending terms).  The reference
# computes no virial; these pins fix its definition: a two-body closed form
# for each term (sign and factor) and W = -dU(lambda)/dlambda at lambda = 1
# under uniform scaling of positions and box (every term present once).

@pytest.mark.parametrize("rr", [0.95, 1.0, 1.12, 1.5, 2.2, 2.49])
def test_virial_two_body_lj(rr):
    s = _two_body(rr)
    s.binning()
    f, e = s.forces(step=0)
    # r_01 = -rr x, F_0 = -F_1; W = r_01 . F_0 = rr * F_1x = -rr dU/dr
    w = rr * 24.0 * (2.0 / rr ** 13 - 1.0 / rr ** 7)
    assert abs(e["virial"] - w) < 1e-12 * max(1.0, abs(w)), (e["virial"], w)
    assert (e["virial"] > 0) == (rr < C6)  # positive while repulsive


def test_virial_two_body_fene():
    inp = configs.melt(1, 2, 0.001, lj_rc=2.5, thermostat=-1, n_cell=16)
    inp.x_space = inp.y_space = 12.0 / 16
    inp.z_space_wall = 12.0
    for t_eps in range(4):
        for u in range(4):
            inp.epsil[t_eps][u] = 0.0
    k, R0 = 30.0, 1.5
    for rr in (0.8, 0.97, 1.1, 1.3):
        s = orc.System(inp)
        r = np.array([[5.0, 5.0, 5.0], [5.0, 5.0 + rr, 5.0]])
        z = np.zeros((2, 3))
        s.set_state(r, z, z, np.array([3, 3], dtype=np.int32))
        s.binning()
        _, e = s.forces(step=0)
        w = -k * rr ** 2 / (1 - (rr / R0) ** 2)  # attractive bond: W < 0
        assert abs(e["virial"] - w) < 1e-11 * abs(w), (rr, e["virial"], w)


def test_virial_is_minus_scaling_derivative_of_the_energy():
    """W = -dU(lambda r, lambda L)/dlambda at lambda = 1 (central
    difference) on the periodic melt with LJ rc = 2.5 and FENE bonds."""
    kind, inp = configs.small_cases()["melt_small"]
    inp.thermostat = -1
    p, r, v, t = configs.build(inp, kind)
    s = orc.System(inp)
    s.set_state(r, v, np.zeros_like(r), t)
    s.binning()
    _, e0 = s.forces(step=0)

    def energy(lam):
        x = lib.Inputs.from_buffer_copy(inp)
        x.x_space *= lam
        x.y_space *= lam
        x.z_space_wall *= lam
        q = lib.derive(x)
        ratio = np.array([q.boundary[d] / p.boundary[d] for d in range(3)])
        assert np.allclose(ratio, lam, rtol=1e-12, atol=0.0)
        s2 = orc.System(x)
        s2.set_state(r * ratio, v, np.zeros_like(r), t)
        s2.binning()
        _, e = s2.forces(step=0)
        return e["v_fluid_fluid"] + e["v_intra_molec"]
    h = 1e-5
    dudl = (energy(1.0 + h) - energy(1.0 - h)) / (2 * h)
    w = e0["virial"]
    print("virial", w, "-dU/dlambda", -dudl)
    assert abs(w + dudl) < 1e-6 * abs(w), (w, -dudl)
    # both terms matter here: the bonds alone carry a sizeable share
    s.zero_force()
    assert abs(s.intra_molec()[1]) > 0.05 * abs(w)


def test_wall_closed_form_and_zero_force_point():
    inp = configs.brush(2, 1, 2, 6.0, 6.0, 8.0, x_space=0.5, thermostat=-1)
    s = orc.System(inp)
    r = np.array([[1.0, 1.0, 1.0], [2.0, 3.0, 2.0], [4.0, 1.0, 1.3], [3.0, 4.0, 4.0]])
    z = np.zeros((4, 3))
    s.set_state(r, z, z, np.array([1, 1, 3, 3], dtype=np.int32))
    s.zero_force()
    out = s.fluid_wall()
    f = s.get_state()["f"]
    A, sw, L = 2.0, 1.0, 8.0   # |a_w| with a_w = -2; sigma_wall = (1+1)/2

    def fz(zz):
        return 9 * A * sw ** 9 / zz ** 10 + 3 * A * sw ** 3 / zz ** 4 \
            - (9 * A * sw ** 9 / (L - zz) ** 10 + 3 * A * sw ** 3 / (L - zz) ** 4)

    def vz(zz):
        return A * (sw / zz) ** 9 + A * (sw / zz) ** 3 + A * (sw / (L - zz)) ** 9 + A * (sw / (L - zz)) ** 3
    for i, zz in enumerate(r[:, 2]):
        assert abs(f[i, 2] - fz(zz)) < 1e-11 * max(1, abs(fz(zz)))
    assert abs(out[0] - sum(vz(zz) for zz in r[:, 2])) < 1e-10
    # midplane: forces of the two walls cancel
    assert abs(fz(4.0)) < 1e-15 and abs(f[3, 2]) < 1e-15


def test_energy_force_consistency_fd():
    """F = -grad V by central differences for LJ + FENE + walls (no DPD)."""
    kind, inp = configs.small_cases()["brush_small"]
    inp.thermostat = -1
    p, r, v, t = configs.build(inp, kind)
    s = orc.System(inp)
    s.set_state(r, v, np.zeros_like(r), t)
    s.binning()
    f, _ = s.forces(step=0)

    def energy(rr):
        s2 = orc.System(inp)
        s2.set_state(rr, v, np.zeros_like(rr), t)
        s2.binning()
        _, e = s2.forces(step=0)
        return e["v_fluid_fluid"] + e["v_intra_molec"] + e["v_fluid_wall"]
    rng = np.random.default_rng(0)
    h = 1e-6
    for i in rng.choice(p.n_part, 12, replace=False):
        for d in range(3):
            rp, rm = r.copy(), r.copy()
            rp[i, d] += h
            rm[i, d] -= h
            g = -(energy(rp) - energy(rm)) / (2 * h)
            assert abs(g - f[i, d]) < 1e-5 * max(1.0, abs(f[i, d])), (i, d, g, f[i, d])


def test_newton_third_law_and_dpd_momentum():
    kind, inp = configs.small_cases()["melt_small"]
    p, r, v, t = configs.build(inp, kind)
    s = orc.System(inp)
    s.set_state(r, v, np.zeros_like(r), t)
    s.binning()
    f, _ = s.forces(step=3)
    frms = np.sqrt((f ** 2).sum(1).mean())
    assert np.abs(f.sum(0)).max() < 1e-10 * frms * p.n_part


def test_verlet_free_flight_and_constant_acceleration():
    inp = configs.melt(1, 2, 0.001, lj_rc=2.5, thermostat=-1, n_cell=16)
    inp.x_space = inp.y_space = 12.0 / 16
    inp.z_space_wall = 12.0
    inp.n_mon_d, inp.n_chain_d = 1, 2
    s = orc.System(inp)
    r = np.array([[1.0, 2.0, 3.0], [7.0, 8.0, 9.0]])
    v = np.array([[0.5, -0.25, 1.0], [-1.0, 0.0, 0.75]])
    a = np.array([[0.1, 0.0, -0.2], [0.0, 0.3, 0.0]])
    s.set_state(r, v, a, np.array([3, 3], dtype=np.int32))
    dt = s.p.dt
    s.verlet_positions()
    st = s.get_state()
    assert np.allclose(st["r"], r + dt * v + 0.5 * dt * dt * a, atol=1e-15)
    assert np.allclose(st["v"], v + 0.5 * dt * a, atol=1e-15)
    # periodic wrap with the image counter (verlet_positions.f90:243-249)
    s.set_state(np.array([[11.999, 2.0, 0.001], [7.0, 8.0, 9.0]]),
                np.array([[1.0, 0, -1.0], [0, 0, 0]]), np.zeros((2, 3)),
                np.array([3, 3], dtype=np.int32))
    s.verlet_positions()
    st = s.get_state()
    assert st["mic"][0].tolist() == [1, 0, -1]
    assert 0.0 <= st["r"][0, 0] < 0.01 and 11.99 < st["r"][0, 2] <= 12.0


def test_check_skin_threshold():
    """check_skin.f90:8-21: rebuild once a displacement exceeds skin/2."""
    inp = configs.melt(1, 2, 0.001, lj_rc=2.5, thermostat=-1, n_cell=16)
    inp.x_space = inp.y_space = 12.0 / 16
    inp.z_space_wall = 12.0
    inp.n_mon_d, inp.n_chain_d = 1, 2
    s = orc.System(inp)
    r = np.array([[1.0, 2.0, 3.0], [7.0, 8.0, 9.0]])
    z = np.zeros((2, 3))
    half = s.p.skin / 2
    v = np.array([[0.9 * half / s.p.dt, 0.0, 0.0], [0.0, 0.0, 0.0]])
    s.set_state(r, v, z, np.array([3, 3], dtype=np.int32))
    s.binning()
    s.verlet_positions()           # moved 0.9 * skin/2
    assert s.check_skin() == 0
    s.verlet_positions()           # moved 1.8 * skin/2
    assert s.check_skin() == 1


def test_pinned_heads():
    kind, inp = configs.small_cases()["brush_small"]
    p, r, v, t = configs.build(inp, kind)
    s = orc.System(inp)
    s.set_state(r, v, np.zeros_like(r), t)
    s.run_steps(1, 5)
    st = s.get_state()
    heads = np.arange(0, p.n_mon * p.n_chain, p.n_mon)
    assert np.array_equal(st["r"][heads], r[heads])
    assert np.all(st["v"][heads] == 0.0) and np.all(st["f"][heads] == 0.0)


# ---------------------------------------------------------------- library ABI

def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "mfa_hip.h")).read()
    names = sorted(set(re.findall(r"\b(mfa_[a-z_0-9]+)\s*\(", hdr)))
    dll = C.CDLL(lib.LIB_PATH)
    missing = [n for n in names if not hasattr(dll, n)]
    assert not missing, missing
    assert set(names) <= set(lib.SIGNATURES), set(names) - set(lib.SIGNATURES)


def test_struct_layouts_match_library():
    assert lib.abi_check()


def test_generators_are_seeded_and_valid():
    kind, inp = configs.small_cases()["melt_small"]
    p, r1, v1, t1 = configs.build(inp, kind)
    _, r2, v2, t2 = configs.build(inp, kind)
    assert np.array_equal(r1, r2) and np.array_equal(v1, v2)
    L = p.boundary[0]
    assert r1.min() >= 0 and r1.max() < L
    d = r1[1:] - r1[:-1]
    d -= L * np.rint(d / L)
    bonds = np.linalg.norm(d, axis=1).reshape(-1)
    chain_bonds = bonds[(np.arange(1, p.n_part) % p.n_mon_d) != 0]
    assert np.allclose(chain_bonds, 0.97, atol=1e-12)
    assert np.abs(v1.sum(0)).max() < 1e-9


def _rock_salt(n_side, a0, alpha_c, n_kmax, rc):
    """Rock-salt crystal of n_side^3 unit charges (+1 type 3, -1 type 4) at
    nearest-neighbour distance a0 in a periodic cube, Ewald in the corrected
    form (1/sqrt(pi), real space within rc), no LJ or DPD forces."""
    from mfa_amd import lib
    x = configs.base_inputs()
    x.system, x.symmetry = 2, 1
    x.dt = 0.005
    L = n_side * a0
    x.n_cell_w_x = x.n_cell_w_y = n_side
    x.x_space = x.y_space = a0
    x.z_space_wall = L
    x.n_mon = x.n_chain = 0
    x.n_mon_d, x.n_chain_d = 1, n_side ** 3
    x.particle_4, x.inter4 = 1, 1
    for t in range(lib.NTYPE):
        x.friction[t] = 0.0
        for u in range(lib.NTYPE):
            x.epsil[t][u] = 0.0
    x.charge[2], x.charge[3] = 1.0, -1.0
    x.alpha_c, x.n_kmax, x.rc_coul, x.ewald_compat = alpha_c, n_kmax, rc, 0
    idx = np.array([(i, j, k) for i in range(n_side) for j in range(n_side) for k in range(n_side)])
    r = (idx + 0.5) * a0
    t = np.where(idx.sum(axis=1) % 2 == 0, 3, 4).astype(np.int32)
    return x, r, t


def _coulomb_energy(x, r, t):
    s = orc.System(x)
    s.set_state(r, np.zeros_like(r), np.zeros_like(r), t)
    s.binning()
    f, e = s.forces(step=1)
    return e["v_coul"], f


def test_ewald_madelung_rock_salt():
    """Corrected form: the Ewald sum of a rock-salt crystal reproduces the
    Madelung constant 1.747564594633 (E = -N/2 M / a0), and the forces on
    every ion vanish by symmetry."""
    x, r, t = _rock_salt(8, 1.5, 1.96, 20, 3.0)
    e, f = _coulomb_energy(x, r, t)
    expect = -(len(t) / 2) * 1.747564594633 / 1.5
    # residual 3e-7: erfc(4.2) real-space truncation and the float32 pi
    assert abs(e - expect) < 1e-6 * abs(expect), (e, expect)
    assert np.abs(f).max() < 1e-6


def test_ewald_total_independent_of_alpha():
    """The sum of real, self and k-space terms does not depend on the
    splitting parameter (a displaced rock salt, two alpha values)."""
    x, r, t = _rock_salt(8, 1.5, 1.44, 20, 3.0)
    rng = np.random.default_rng(3)
    r = r + rng.uniform(-0.2, 0.2, r.shape)
    e1, f1 = _coulomb_energy(x, r, t)
    x.alpha_c = 1.96
    e2, f2 = _coulomb_energy(x, r, t)
    assert abs(e1 - e2) < 2e-6 * abs(e1), (e1, e2)
    assert np.abs(f1 - f2).max() < 1e-4 * np.abs(f1).max()


def test_ewald_k_count_matches_brute_enumeration():
    """The half-space k set of ewald_k.f90:104-124: every k in the sphere
    appears exactly once with its mirror image -k absent."""
    kind, inp = configs.small_cases()["pe_brush_compat"]
    p, r, v, t = configs.build(inp, kind)
    s = orc.System(inp)
    s.set_state(r, v, np.zeros_like(r), t)
    kv = s.k_vectors()
    keys = {tuple(np.round(k, 9)) for k in kv}
    assert len(keys) == len(kv)
    assert not any(tuple(np.round(-k, 9)) in keys for k in kv)
    # the full sphere (both halves, minus k = 0) has twice as many points
    dk = np.array([kv[:, 0][kv[:, 0] > 0].min(), kv[:, 1][kv[:, 1] > 0].min(),
                   kv[:, 2][kv[:, 2] > 0].min()])
    kc = (dk[0] + dk[1]) * inp.n_kmax / 2
    n = [int(kc / d) for d in dk]
    full = 0
    for a in range(-n[0], n[0] + 1):
        for b in range(-n[1], n[1] + 1):
            for c in range(-n[2], n[2] + 1):
                k2 = (a * dk[0]) ** 2 + (b * dk[1]) ** 2 + (c * dk[2]) ** 2
                full += 0 < k2 <= kc * kc
    assert full == 2 * len(kv)


def test_bench_ewald_flop_count_uses_the_oracle_k_set():
    """bench.py's k-vector count (the flop count of the k-space roofline) is
    the oracle's k set, for the slab pe brushes and a periodic box."""
    import bench
    cases = configs.small_cases()
    for name in ("pe_brush_compat", "pe_brush_small"):
        kind, inp = cases[name]
        p, r, v, t = configs.build(inp, kind)
        s = orc.System(inp)
        s.set_state(r, v, np.zeros_like(r), t)
        assert bench.ewald_k_count(p, inp) == len(s.k_vectors()), name
    x, r, t = _rock_salt(8, 1.5, 1.96, 20, 3.0)
    p = lib.derive(x)
    s = orc.System(x)
    s.set_state(r, np.zeros_like(r), np.zeros_like(r), t)
    assert p.symmetry == 1 and bench.ewald_k_count(p, x) == len(s.k_vectors())


# ---------------------------------------------------------------- Ewald plan
# The k-space stage's launch plan (mfa_ewald_plan, host only) and its
# checks: the k set equals the oracle's enumeration, the row table covers
# it exactly, and every launch shape stays inside its buffers.

def _plan_invariants(plan, n_charged):
    assert plan["lds_rho"] <= 65536 and plan["lds_force"] <= 65536
    assert plan["force_partials"] <= plan["force_partials_cap"]
    if plan["row_kernel"]:
        assert plan["force_grid"] % 8 == 0 and plan["force_grid"] >= plan["force_blocks"]
        pb = -(-n_charged // 256)
        assert plan["force_blocks"] == pb * plan["force_groups"]
        assert (plan["force_groups"] - 1) * plan["rows_per_group"] < plan["rows"]
        assert plan["force_groups"] * plan["rows_per_group"] >= plan["rows"]
        assert plan["lds_force"] == plan["rows_per_group"] * (2 * plan["nx_max"] + 1) * 16


@pytest.mark.parametrize("name", ["brush9k", "melt100k", "pe_brush27k"])
def test_params_set_skin_matches_derive(name):
    """mfa_params_set_skin on derived params gives the params mfa_derive
    gives with that skin, byte for byte (list criterion with and without
    DPD_CUT_OFF, and with a real-space Coulomb cut-off of its own)."""
    x = configs.WORKLOADS[name]["make"]()
    base = lib.derive(x)
    for skin in (0.2, 0.35, 0.5):
        y = lib.Inputs.from_buffer_copy(x)
        y.skin = skin
        try:
            want = lib.derive(y)
        except lib.MfaError:
            with pytest.raises(lib.MfaError):
                lib.with_skin(base, skin)
            continue
        got = lib.with_skin(base, skin)
        assert bytes(got) == bytes(want), (name, skin)


def test_params_set_skin_refuses_lists_wider_than_the_cells():
    x = configs.WORKLOADS["melt100k"]["make"]()
    p = lib.derive(x)
    reach = max(p.r_bin) - np.sqrt(max(max(row) for row in p.range_2))
    with pytest.raises(lib.MfaError) as ei:
        lib.with_skin(p, reach + 0.05)
    assert "narrower than the list radius" in str(ei.value)
    q = lib.with_skin(p, min(p.r_bin) - np.sqrt(max(max(row) for row in p.range_2)) - 0.01)
    assert max(max(row) for row in q.list_2) <= min(p.r_bin) ** 2


@pytest.mark.parametrize("name", ["brush9k", "melt100k", "pe_brush27k"])
def test_params_set_cells(name):
    """mfa_params_set_cells: at the reference's cell width (r_cut_max + 0.4,
    make_binning_boxes.f90:42-47) it gives mfa_derive's params byte for byte;
    a wider width gives n_bin = int(L / width) with r_bin = L / n_bin and
    leaves every other field alone; widths with fewer than 3 cells on an
    axis, or cells narrower than the list radius, are refused."""
    x = configs.WORKLOADS[name]["make"]()
    p = lib.derive(x)
    ref_w = p.r_cut_max + float(np.float32(0.4))
    assert bytes(lib.with_cells(p, ref_w)) == bytes(p)
    w = max(p.r_bin) + 0.3
    q = lib.with_cells(p, w)
    for d in range(3):
        assert q.n_bin[d] == int(p.boundary[d] / w) and q.n_bin[d] < p.n_bin[d]
        assert q.r_bin[d] == p.boundary[d] / q.n_bin[d] and q.inv_r_bin[d] == 1.0 / q.r_bin[d]
    r = lib.Params.from_buffer_copy(q)
    for d in range(3):
        r.n_bin[d], r.r_bin[d], r.inv_r_bin[d] = p.n_bin[d], p.r_bin[d], p.inv_r_bin[d]
    assert bytes(r) == bytes(p)
    with pytest.raises(lib.MfaError) as ei:
        lib.with_cells(p, 0.9 * np.sqrt(max(max(row) for row in p.list_2)))
    assert "narrower than the list radius" in str(ei.value)
    with pytest.raises(lib.MfaError) as ei:
        lib.with_cells(p, min(p.boundary) / 2.5)
    assert "fewer than 3" in str(ei.value)


def test_ewald_plan_pe_brush27k_grid():
    w = configs.WORKLOADS["pe_brush27k"]
    x = w["make"]()
    p = lib.derive(x)
    rc, plan, msg = lib.ewald_plan(p, 10400)
    assert rc == 0, msg
    assert plan["k_vectors"] == len(orc.System(x).k_vectors()) == 3997
    assert plan["row_kernel"] == 1 and plan["nx_max"] == 10
    _plan_invariants(plan, 10400)


@pytest.mark.parametrize("case", ["pe_brush_compat", "pe_brush_small"])
def test_ewald_plan_test_grids(case):
    kind, x = configs.small_cases()[case]
    p = lib.derive(x)
    nq = p.n_part // 3
    rc, plan, msg = lib.ewald_plan(p, nq)
    assert rc == 0, msg
    assert plan["k_vectors"] == len(orc.System(x).k_vectors())
    _plan_invariants(plan, nq)


def test_ewald_plan_wide_grid_takes_the_segment_kernel():
    """kmax = 16 in a cube (nx_max 16 > 12): no row kernel; the segment
    kernel's tiles still fit."""
    x = configs.melt(40, 10, 0.85, lj_rc=0.0, dt=0.005, seed=13, n_cell=16)
    x.charge[2] = 1.0
    x.alpha_c = 0.81
    x.n_kmax = 16
    p = lib.derive(x)
    rc, plan, msg = lib.ewald_plan(p, p.n_part)
    assert rc == 0, msg
    assert plan["row_kernel"] == 0 and plan["nx_max"] == 16
    _plan_invariants(plan, p.n_part)


@pytest.mark.parametrize("n_charged", [65600, 300000])
def test_ewald_plan_many_charges_on_the_kmax10_grid(n_charged):
    """Beyond ~65.5k charges the segment kernel's force tile would exceed
    64 KB; the row kernel only needs its rows of G, so such systems are
    accepted (rows per group bounded by the LDS limit)."""
    x = configs.pe_brush(400, 40, 5800, 63.0, 63.0, 40.0, n_kmax=10, alpha_c=0.81, rc_coul=3.0)
    p = lib.derive(x)
    p.n_part = max(p.n_part, n_charged)
    rc, plan, msg = lib.ewald_plan(p, n_charged)
    assert rc == 0, msg
    assert plan["row_kernel"] == 1
    _plan_invariants(plan, n_charged)


def test_error_codes_match_the_header():
    """Every MFA_ERR_* of include/mfa_hip.h has its name in lib.ERRORS."""
    text = open(os.path.join(ROOT, "include", "mfa_hip.h")).read()
    codes = {int(v): k for k, v in re.findall(r"(MFA_ERR_\w+)\s*=\s*(\d+)", text)}
    assert codes and all(lib.ERRORS.get(v) == k for v, k in codes.items()), codes


def test_pair_deviate_is_box_muller_of_threefry13():
    """The pair deviate is Box-Muller of Threefry-2x32-13 on counter (min,
    max) and key (seed, step), in the oracle and (by its round constant) in
    the HIP path."""
    w = orc.threefry2x32(10, 20, 5, 3, rounds=13)
    u1 = (w[0] >> 8) * 2.0 ** -24 + 2.0 ** -25
    u2 = (w[1] >> 8) * 2.0 ** -24
    assert abs(orc.pair_normal(5, 3, 20, 10) -
               np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)) < 1e-14
    dev = open(os.path.join(ROOT, "mfa_amd", "csrc", "mfa_rng.h")).read()
    orc_h = open(os.path.join(ROOT, "oracle", "mfa_oracle.h")).read()
    r_dev = int(re.search(r"#define MFA_TF_ROUNDS (\d+)", dev).group(1))
    r_orc = int(re.search(r"#define ORC_TF_ROUNDS (\d+)", orc_h).group(1))
    assert r_dev == r_orc == 13
