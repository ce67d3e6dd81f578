"""ctypes binding of the oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.  It never stands in for the HIP library.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
NT = 4

_d4 = C.c_double * NT
_d44 = (C.c_double * NT) * NT
_d3 = C.c_double * 3
_i3 = C.c_int * 3


class OrcParams(C.Structure):
    _fields_ = [
        ("n_part", C.c_int), ("n_mon", C.c_int), ("n_chain", C.c_int), ("n_mon_d", C.c_int),
        ("n_chain_d", C.c_int), ("n_mon_e", C.c_int), ("n_chain_e", C.c_int),
        ("part_init_d", C.c_int), ("part_init_e", C.c_int),
        ("system", C.c_int), ("symmetry", C.c_int), ("thermostat", C.c_int),
        ("dpd_cut_off_on", C.c_int), ("particle_4", C.c_int),
        ("boundary", _d3), ("inv_boundary", _d3),
        ("dt", C.c_double), ("dt_2", C.c_double), ("temp", C.c_double),
        ("mass_type", _d4),
        ("sigma", _d44), ("sigma_2", _d44), ("epsil", _d44),
        ("range_2", _d44), ("inv_range_2", _d44), ("e_shift", _d44),
        ("r_cut_max", C.c_double), ("r_cut_max_2", C.c_double),
        ("r_cut_dpd", C.c_double), ("r_cut_dpd_2", C.c_double),
        ("skin", C.c_double), ("skin_2", C.c_double),
        ("friction1", C.c_double), ("sig", C.c_double),
        ("k_chain", C.c_double), ("r_chain_2", C.c_double), ("inv_r_chain_2", C.c_double),
        ("a_w", C.c_double), ("sigma_wall", _d4), ("a_wall", _d4), ("z_space_wall", C.c_double),
        ("va_spring_twall", _d3),
        ("n_bin", _i3), ("delta_bin", _i3), ("r_bin", _d3), ("inv_r_bin", _d3),
        ("seed", C.c_uint32), ("r_2_min_init", C.c_double), ("r_2_min_time", C.c_double),
        ("coulomb", C.c_int), ("ewald_compat", C.c_int), ("n_kmax", C.c_int),
        ("charge", _d4), ("alpha_c", C.c_double), ("sqrt_alpha", C.c_double),
        ("inv_sqrt_pi", C.c_double), ("rc_coul_2", C.c_double),
        ("bending", C.c_int), ("k_bend", C.c_double), ("alpha_eq", C.c_double),
        ("dpd_vv", C.c_int),
    ]


_lib = None


class ObsSums(C.Structure):
    """The layout of mfa_obs_sums (include/mfa_hip.h)."""
    _fields_ = [
        ("count_obs", C.c_int64),
        ("v_total_1", C.c_double), ("v_total_2", C.c_double),
        ("t_total_1", C.c_double), ("t_total_2", C.c_double),
        ("e_total_1", C.c_double), ("e_total_2", C.c_double),
        ("v_fluid_fluid_1", C.c_double), ("v_fluid_fluid_2", C.c_double),
        ("t_fluid_1", C.c_double), ("t_fluid_2", C.c_double),
        ("e_fluid_fluid_1", C.c_double), ("e_fluid_fluid_2", C.c_double),
        ("v_tot_intra_1", C.c_double), ("v_tot_intra_2", C.c_double),
        ("tot_p", C.c_double * 3), ("r_g2_mean", (C.c_double * 3) * 3),
        ("r_par_per2", C.c_double * 2), ("count_hist", C.c_int64),
        ("hist_dim", C.c_int32), ("pad", C.c_int32),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    u32 = C.c_uint32
    lib.orc_threefry2x32.argtypes = [u32, u32, u32, u32, C.c_int, C.POINTER(u32), C.POINTER(u32)]
    lib.orc_threefry2x32.restype = None
    lib.orc_pair_normal.argtypes = [u32, C.c_int64, C.c_int, C.c_int]
    lib.orc_pair_normal.restype = C.c_double
    lib.orc_particle_normal.argtypes = [u32, C.c_int64, C.c_int, C.c_int]
    lib.orc_particle_normal.restype = C.c_double
    lib.orc_derive.argtypes = [C.c_void_p, C.POINTER(OrcParams)]
    lib.orc_derive.restype = C.c_int
    lib.orc_create.argtypes = [C.POINTER(OrcParams), C.c_int]
    lib.orc_create.restype = P
    lib.orc_destroy.argtypes = [P]
    lib.orc_set_state.argtypes = [P, dp, dp, dp, ip, ip]
    lib.orc_get_state.argtypes = [P, dp, dp, dp, dp, ip]
    lib.orc_verlet_positions.argtypes = [P]
    lib.orc_check_skin.argtypes = [P]
    lib.orc_check_skin.restype = C.c_int
    lib.orc_binning.argtypes = [P]
    lib.orc_binning.restype = C.c_int
    lib.orc_get_cells.argtypes = [P, ip]
    lib.orc_get_pairs.argtypes = [P, ip, C.c_int64]
    lib.orc_get_pairs.restype = C.c_int64
    lib.orc_brute_pairs.argtypes = [P, ip, C.c_int64]
    lib.orc_brute_pairs.restype = C.c_int64
    lib.orc_zero_force.argtypes = [P]
    lib.orc_fluid_fluid.argtypes = [P, C.c_int64, dp]
    lib.orc_fluid_wall.argtypes = [P, dp]
    lib.orc_intra_molec.argtypes = [P, dp]
    lib.orc_verlet_velocities.argtypes = [P]
    lib.orc_verlet_velocities.restype = C.c_double
    lib.orc_new_dpd_fd.argtypes = [P]
    lib.orc_new_dpd_fd.restype = C.c_double
    lib.orc_run_steps.argtypes = [P, C.c_int64, C.c_int, dp]
    lib.orc_run_steps.restype = C.c_int
    lib.orc_obs_init.argtypes = [P, C.c_int]
    lib.orc_obs_init.restype = None
    lib.orc_observe.argtypes = [P, dp, C.c_int]
    lib.orc_observe.restype = None
    lib.orc_obs_read.argtypes = [P, C.POINTER(ObsSums), dp, dp]
    lib.orc_obs_read.restype = None
    lib.orc_ewald_k.argtypes = [P]
    lib.orc_ewald_k.restype = C.c_double
    lib.orc_ewald_nk.argtypes = [P, dp]
    lib.orc_ewald_nk.restype = C.c_int
    lib.orc_dipolar.argtypes = [P]
    lib.orc_dipolar.restype = C.c_double
    lib.orc_bending.argtypes = [P]
    lib.orc_bending.restype = C.c_double
    _lib = lib
    return lib


def threefry2x32(c0, c1, k0, k1, rounds=13):
    lib = load()
    a, b = C.c_uint32(), C.c_uint32()
    lib.orc_threefry2x32(c0, c1, k0, k1, rounds, C.byref(a), C.byref(b))
    return a.value, b.value


def pair_normal(seed, step, i, j):
    return load().orc_pair_normal(seed, step, i, j)


def particle_normal(seed, step, i, d):
    return load().orc_particle_normal(seed, step, i, d)


def derive(inputs):
    lib = load()
    p = OrcParams()
    rc = lib.orc_derive(C.byref(inputs), C.byref(p))
    if rc:
        raise RuntimeError("orc_derive failed")
    return p


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


class System:
    """The reference's state and routines on the CPU (FP64, one thread)."""

    def __init__(self, inputs, n_neigh=256):
        self.lib = load()
        self.p = derive(inputs)
        self.n = self.p.n_part
        self.h = self.lib.orc_create(C.byref(self.p), n_neigh)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def set_state(self, r, v, a, t, mic=None):
        self._keep = [np.ascontiguousarray(r, dtype=np.float64),
                      np.ascontiguousarray(v, dtype=np.float64),
                      np.ascontiguousarray(a, dtype=np.float64),
                      np.ascontiguousarray(t, dtype=np.int32),
                      None if mic is None else np.ascontiguousarray(mic, dtype=np.int32)]
        r, v, a, t, m = self._keep
        n = self.n
        assert r.shape == (n, 3) and v.shape == (n, 3) and a.shape == (n, 3) and t.shape == (n,)
        assert m is None or m.shape == (n, 3)
        self.lib.orc_set_state(self.h, _dp(r), _dp(v), _dp(a), _ip(t), _ip(m))

    def get_state(self):
        n = self.n
        r, v, a, f = (np.zeros((n, 3)) for _ in range(4))
        mic = np.zeros((n, 3), dtype=np.int32)
        self.lib.orc_get_state(self.h, _dp(r), _dp(v), _dp(a), _dp(f), _ip(mic))
        return {"r": r, "v": v, "a": a, "f": f, "mic": mic}

    def verlet_positions(self):
        self.lib.orc_verlet_positions(self.h)

    def check_skin(self):
        return self.lib.orc_check_skin(self.h)

    def binning(self):
        rc = self.lib.orc_binning(self.h)
        if rc:
            raise RuntimeError("orc_binning error %d" % rc)

    def get_cells(self):
        c = np.zeros((self.n, 3), dtype=np.int32)
        self.lib.orc_get_cells(self.h, _ip(c))
        return c

    def get_pairs(self):
        m = self.lib.orc_get_pairs(self.h, None, 0)
        p = np.zeros((m, 2), dtype=np.int32)
        self.lib.orc_get_pairs(self.h, _ip(p), m)
        return p

    def brute_pairs(self):
        m = self.lib.orc_brute_pairs(self.h, None, 0)
        p = np.zeros((m, 2), dtype=np.int32)
        self.lib.orc_brute_pairs(self.h, _ip(p), m)
        return p

    def zero_force(self):
        self.lib.orc_zero_force(self.h)

    def fluid_fluid(self, step):
        out = np.zeros(3)
        self.lib.orc_fluid_fluid(self.h, step, _dp(out))
        return out

    def fluid_wall(self):
        out = np.zeros(3)
        self.lib.orc_fluid_wall(self.h, _dp(out))
        return out

    def intra_molec(self):
        out = np.zeros(2)
        self.lib.orc_intra_molec(self.h, _dp(out))
        return out

    def verlet_velocities(self):
        return self.lib.orc_verlet_velocities(self.h)

    def new_dpd_fd(self):
        return self.lib.orc_new_dpd_fd(self.h)

    def run_steps(self, first_step, n):
        obs = np.zeros(9)
        rc = self.lib.orc_run_steps(self.h, first_step, n, _dp(obs))
        if rc:
            raise RuntimeError("orc_run_steps error %d" % rc)
        return {"v_fluid_fluid": obs[0], "v_fluid_wall": obs[1], "v_intra_molec": obs[2],
                "t_kin": obs[3], "virial": obs[4], "n_rebuilds": obs[5], "wall_fz_bottom": obs[6],
                "v_coul": obs[7], "v_bend": obs[8]}

    def obs_init(self, hist_dim=200):
        self.lib.orc_obs_init(self.h, hist_dim)
        self.hist_dim = hist_dim

    def observe(self, obs, coul_in_total=False):
        """One observation after a step; obs as returned by run_steps."""
        e = np.array([obs["v_fluid_fluid"], obs["v_fluid_wall"], obs["v_intra_molec"], obs["t_kin"],
                      obs["v_coul"]])
        self.lib.orc_observe(self.h, _dp(e), 1 if coul_in_total else 0)

    def obs_read(self):
        s = ObsSums()
        hd = max(getattr(self, "hist_dim", 0), 0)
        dens = np.zeros((3, max(hd, 1)))
        heads = np.zeros((max(self.p.n_chain, 1), 3))
        self.lib.orc_obs_read(self.h, C.byref(s), _dp(dens), _dp(heads))
        d = {}
        for f, _ in s._fields_:
            x = getattr(s, f)
            d[f] = np.ctypeslib.as_array(x).copy() if hasattr(x, "_length_") else x
        d["dens_z"] = dens[:, :hd]
        d["f_on_heads"] = heads[:self.p.n_chain]
        return d

    def bending(self):
        return self.lib.orc_bending(self.h)

    def ewald_k(self):
        return self.lib.orc_ewald_k(self.h)

    def dipolar(self):
        return self.lib.orc_dipolar(self.h)

    def k_vectors(self):
        nk = self.lib.orc_ewald_nk(self.h, None)
        kv = np.zeros((nk, 3))
        if nk:
            self.lib.orc_ewald_nk(self.h, _dp(kv))
        return kv

    def forces(self, step):
        """Total force of one step's force pass at the current state
        (md_main.f90:88-106 without integrating)."""
        self.zero_force()
        e_ff = self.fluid_fluid(step)
        e_fw = self.fluid_wall() if self.p.symmetry == 0 else np.zeros(3)
        e_in = self.intra_molec()
        v_bend = self.bending()  # md_main.f90:110-111
        v_coul = e_ff[2]
        if self.p.coulomb:  # md_main.f90:128-132
            v_coul += self.ewald_k()
            if self.p.symmetry == 0:
                v_coul += self.dipolar()
        return self.get_state()["f"], {"v_fluid_fluid": e_ff[0], "v_fluid_wall": e_fw[0],
                                       "v_intra_molec": e_in[0], "virial": e_ff[1] + e_in[1],
                                       "wall_fz_bottom": e_fw[1], "wall_fz_top": e_fw[2],
                                       "v_coul": v_coul, "v_bend": v_bend}
