"""ctypes binding of libmfa_hip (include/mfa_hip.h).

This is the Python side of the C ABI that the Fortran driver also binds
(mfa_amd/fortran/mfa_hip_bind.f90).  The structures below mirror the C
structs field by field; `abi_check()` compares their sizes with the library's.
Loading fails loudly when the library is missing: there is no fallback.
"""
import ctypes as C
import os

import numpy as np

NTYPE = 4
HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfa_hip.so")

MFA_FP64 = 0
MFA_MIXED = 1

ERRORS = {
    0: "OK", 1: "MFA_ERR_ARG", 2: "MFA_ERR_HIP", 3: "MFA_ERR_LIST_OVERFLOW",
    4: "MFA_ERR_CELL", 5: "MFA_ERR_NOMEM", 6: "MFA_ERR_NAN", 7: "MFA_ERR_IO",
    8: "MFA_ERR_CELL_OVERFLOW", 9: "MFA_ERR_INTERNAL",
}

_d4 = C.c_double * NTYPE
_d44 = (C.c_double * NTYPE) * NTYPE
_d3 = C.c_double * 3
_i3 = C.c_int32 * 3


class Inputs(C.Structure):
    _fields_ = [
        ("n_mon", C.c_int32), ("n_chain", C.c_int32), ("n_cell_w_x", C.c_int32),
        ("n_cell_w_y", C.c_int32), ("n_mon_d", C.c_int32), ("n_chain_d", C.c_int32),
        ("n_mon_e", C.c_int32), ("n_chain_e", C.c_int32), ("inter4", C.c_int32),
        ("pad0", C.c_int32),
        ("a_w", C.c_double), ("sigma_w", C.c_double), ("a_w4", C.c_double),
        ("sigma_w4", C.c_double), ("charge", _d4), ("alpha_c", C.c_double),
        ("n_kmax", C.c_int32), ("pad1", C.c_int32),
        ("s_time", C.c_int32), ("n_relax", C.c_int32), ("n_obser", C.c_int32),
        ("n_safe", C.c_int32), ("dt", C.c_double), ("temp", C.c_double),
        ("x_space", C.c_double), ("y_space", C.c_double), ("z_space_wall", C.c_double),
        ("mass_type", _d4), ("epsil", _d44), ("sigma_ii", _d4), ("friction", _d4),
        ("k_bend", C.c_double), ("alpha_eq", C.c_double), ("iseed", C.c_int64),
        ("k_chain", C.c_double), ("r_chain", C.c_double), ("va_spring_twall", _d3),
        ("system", C.c_int32), ("symmetry", C.c_int32), ("solvent", C.c_int32),
        ("thermostat", C.c_int32), ("particle_4", C.c_int32), ("bending", C.c_int32),
        ("dpd_cut_off", C.c_double), ("skin", C.c_double), ("lj_rc", C.c_double),
        ("rc_coul", C.c_double), ("ewald_compat", C.c_int32), ("force_switch_on", C.c_int32),
        ("dpd_vv", C.c_int32), ("pad2", C.c_int32),
    ]


class Params(C.Structure):
    _fields_ = [
        ("n_part", C.c_int32), ("n_mon_tot", C.c_int32),
        ("n_mon", C.c_int32), ("n_chain", C.c_int32), ("n_mon_d", C.c_int32),
        ("n_chain_d", C.c_int32), ("n_mon_e", C.c_int32), ("n_chain_e", C.c_int32),
        ("part_init_d", C.c_int32), ("part_init_e", C.c_int32),
        ("system", C.c_int32), ("symmetry", C.c_int32), ("thermostat", C.c_int32),
        ("dpd_weight", C.c_int32), ("dpd_cut_off_on", C.c_int32), ("coulomb", C.c_int32),
        ("bending", C.c_int32), ("ewald_compat", C.c_int32),
        ("boundary", _d3), ("inv_boundary", _d3),
        ("dt", C.c_double), ("dt_2", C.c_double), ("temp", C.c_double),
        ("mass_type", _d4), ("inv_mass_type", _d4),
        ("epsil", _d44), ("sigma", _d44), ("sigma_2", _d44), ("range_2", _d44),
        ("inv_range_2", _d44), ("e_shift", _d44), ("list_2", _d44),
        ("r_cut_max", C.c_double), ("r_cut_max_2", C.c_double),
        ("r_cut_dpd", C.c_double), ("r_cut_dpd_2", C.c_double),
        ("skin", C.c_double), ("skin_2", C.c_double), ("skin_half_2", C.c_double),
        ("friction1", C.c_double), ("sig", C.c_double),
        ("k_chain", C.c_double), ("r_chain_2", C.c_double), ("inv_r_chain_2", C.c_double),
        ("a_w", C.c_double), ("sigma_wall", _d4), ("a_wall", _d4), ("z_space_wall", C.c_double),
        ("va_spring_twall", _d3), ("k_bend", C.c_double), ("alpha_eq", C.c_double),
        ("n_bin", _i3), ("delta_bin", _i3), ("r_bin", _d3), ("inv_r_bin", _d3),
        ("charge", _d4), ("alpha_c", C.c_double), ("sqrt_alpha", C.c_double),
        ("inv_sqrt_pi_real", C.c_double), ("rc_coul_2", C.c_double),
        ("v_coul_self", C.c_double), ("n_kmax", C.c_int32), ("pad0", C.c_int32),
        ("seed", C.c_uint64), ("fso_r2_init", C.c_double), ("fso_time", C.c_double),
        ("dpd_vv", C.c_int32), ("pad1", C.c_int32),
    ]


class Obs(C.Structure):
    _fields_ = [
        ("v_fluid_fluid", C.c_double), ("v_fluid_wall", C.c_double),
        ("v_intra_molec", C.c_double), ("v_bend", C.c_double), ("v_coul", C.c_double),
        ("t_kin", C.c_double), ("virial", C.c_double), ("wall_fz_bottom", C.c_double),
        ("wall_fz_top", C.c_double), ("z_moment", C.c_double), ("n_rebuilds", C.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ObsSums(C.Structure):
    """mfa_obs_sums: running sums of the observation stage."""
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

    def as_dict(self):
        d = {}
        for f, _ in self._fields_:
            x = getattr(self, f)
            d[f] = np.ctypeslib.as_array(x).copy() if hasattr(x, "_length_") else x
        return d


class AdjointOut(C.Structure):
    _fields_ = [
        ("height", C.c_double), ("p_wall", C.c_double),
        ("d_height", C.c_double * 5), ("d_p_wall", C.c_double * 5),
        ("n_rebuilds", C.c_int64),
    ]


# Every function include/mfa_hip.h declares, with its ctypes signature.
_P = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
SIGNATURES = {
    "mfa_abi_sizes": (C.c_int32, [_ip, _ip, _ip]),
    "mfa_derive": (C.c_int32, [C.POINTER(Inputs), C.POINTER(Params)]),
    "mfa_create": (C.c_int32, [C.POINTER(Params), C.c_int32, C.c_int32, C.POINTER(_P)]),
    "mfa_destroy": (C.c_int32, [_P]),
    "mfa_set_precision": (C.c_int32, [_P, C.c_int32]),
    "mfa_set_threads_per_particle": (C.c_int32, [_P, C.c_int32]),
    "mfa_last_error": (C.c_char_p, [_P]),
    "mfa_set_state": (C.c_int32, [_P, _dp, _dp, _dp, _ip, _ip]),
    "mfa_get_state": (C.c_int32, [_P, _dp, _dp, _dp, _dp, _ip]),
    "mfa_verlet_positions": (C.c_int32, [_P]),
    "mfa_check_skin": (C.c_int32, [_P, _ip]),
    "mfa_binning": (C.c_int32, [_P]),
    "mfa_get_cells": (C.c_int32, [_P, _ip]),
    "mfa_get_pairs": (C.c_int32, [_P, _ip, C.c_int64, _lp]),
    "mfa_zero_force": (C.c_int32, [_P]),
    "mfa_fluid_fluid": (C.c_int32, [_P, C.c_int64, _dp, _dp]),
    "mfa_fluid_wall": (C.c_int32, [_P, C.c_int32, _dp]),
    "mfa_intra_molec": (C.c_int32, [_P, _dp]),
    "mfa_ewald_k": (C.c_int32, [_P, C.c_int32, _dp]),
    "mfa_ewald_plan": (C.c_int32, [C.POINTER(Params), C.c_int32, _lp, C.c_int32]),
    "mfa_dipolar_correction": (C.c_int32, [_P, _dp]),
    "mfa_bending": (C.c_int32, [_P, C.c_int32, _dp]),
    "mfa_create_batch": (C.c_int32, [C.POINTER(Params), C.c_int32, C.c_int32, C.c_int32,
                                     C.POINTER(_P)]),
    "mfa_set_state_replica": (C.c_int32, [_P, C.c_int32, _dp, _dp, _dp, _ip, _ip]),
    "mfa_get_state_replica": (C.c_int32, [_P, C.c_int32, _dp, _dp, _dp, _dp, _ip]),
    "mfa_replica_obs": (C.c_int32, [_P, C.c_int32, C.POINTER(Obs)]),
    "mfa_autotune": (C.c_int32, [_P, C.c_int64, C.c_int32, C.POINTER(C.c_int32), _dp]),
    "mfa_obs_init": (C.c_int32, [_P, C.c_int32]),
    "mfa_observe": (C.c_int32, [_P, C.POINTER(Obs), C.c_int32]),
    "mfa_obs_read": (C.c_int32, [_P, C.POINTER(ObsSums), _dp, _dp]),
    "mfa_verlet_velocities": (C.c_int32, [_P, _dp]),
    "mfa_new_dpd_fd": (C.c_int32, [_P, _dp]),
    "mfa_verlet_velocities_noalloc": (C.c_int32, [C.POINTER(Params), C.c_void_p, C.c_void_p,
                                                  C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
    "mfa_run_steps": (C.c_int32, [_P, C.c_int64, C.c_int32, C.POINTER(Obs)]),
    "mfa_energies": (C.c_int32, [_P, C.c_int64, C.POINTER(Obs)]),
    "mfa_timing": (C.c_int32, [_P, C.c_int32, _dp, _lp, _dp, _lp]),
    "mfa_stats": (C.c_int32, [_P, _lp]),
    "mfa_set_graph_steps": (C.c_int32, [_P, C.c_int32]),
    "mfa_time_force": (C.c_int32, [_P, C.c_int32, _dp]),
    "mfa_time_force_step": (C.c_int32, [_P, C.c_int32, _dp]),
    "mfa_time_ewald_parts": (C.c_int32, [_P, C.c_int32, _dp]),
    "mfa_time_rebuild_idle": (C.c_int32, [_P, C.c_int32, _dp]),
    "mfa_params_set_skin": (C.c_int32, [C.c_void_p, C.c_double]),
    "mfa_set_skin": (C.c_int32, [_P, C.c_double]),
    "mfa_set_pair_queue": (C.c_int32, [_P, C.c_int32]),
    "mfa_set_half_list": (C.c_int32, [_P, C.c_int32]),
    "mfa_get_tuning": (C.c_int32, [_P, _ip, _ip, _dp, _ip]),
    "mfa_autotune_skin": (C.c_int32, [_P, C.c_int64, C.c_int32, C.c_int32, _dp, _dp, _dp]),
    "mfa_params_set_cells": (C.c_int32, [C.c_void_p, C.c_double]),
    "mfa_set_cells": (C.c_int32, [_P, C.c_double]),
    "mfa_autotune_cells": (C.c_int32, [_P, C.c_int64, C.c_int32, C.c_int32, _dp, C.c_int32, _dp,
                                       _dp, _dp, _dp, C.POINTER(C.c_int64)]),
    "mfa_time_ewald": (C.c_int32, [_P, C.c_int32, _dp]),
    "mfa_time_rebuild": (C.c_int32, [_P, C.c_int32, _dp]),
    "mfa_copy_probe": (C.c_int32, [C.c_int64, _dp]),
    "mfa_generate": (C.c_int32, [C.POINTER(Params), C.c_int32, C.c_uint64, C.c_double,
                                 C.c_double, C.c_double, _dp, _dp, _ip]),
    "mfa_adjoint_window": (C.c_int32, [_P, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                       C.POINTER(AdjointOut)]),
}

_lib = None


def load(path=LIB_PATH):
    """Load libmfa_hip; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            "libmfa_hip.so not found at %s: build it with `python -m mfa_amd.build` "
            "(there is no CPU fallback)" % path)
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def abi_check():
    lib = load()
    a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
    lib.mfa_abi_sizes(C.byref(a), C.byref(b), C.byref(c))
    return (a.value, b.value, c.value) == (C.sizeof(Inputs), C.sizeof(Params), C.sizeof(Obs))


class MfaError(RuntimeError):
    """A nonzero status of the C ABI; `code` is the MFA_ERR_* value."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def copy_probe(nbytes):
    """Time a device-to-device copy of nbytes; returns milliseconds."""
    l = load()
    ms = C.c_double()
    rc = l.mfa_copy_probe(nbytes, C.byref(ms))
    if rc:
        raise MfaError("mfa_copy_probe: %s" % l.mfa_last_error(None).decode())
    return ms.value


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def derive(inputs):
    lib = load()
    p = Params()
    rc = lib.mfa_derive(C.byref(inputs), C.byref(p))
    if rc:
        raise MfaError("mfa_derive: %s (%s)" % (lib.mfa_last_error(None).decode(), ERRORS[rc]))
    return p


def with_cells(params, width):
    """A copy of derived params with another cell grid (mfa_params_set_cells)."""
    lib = load()
    q = Params.from_buffer_copy(params)
    rc = lib.mfa_params_set_cells(C.byref(q), width)
    if rc:
        raise MfaError("mfa_params_set_cells: %s (%s)" % (lib.mfa_last_error(None).decode(), ERRORS[rc]),
                       rc)
    return q


def with_skin(params, skin):
    """A copy of derived params with another list skin (mfa_params_set_skin)."""
    lib = load()
    q = Params.from_buffer_copy(params)
    rc = lib.mfa_params_set_skin(C.byref(q), skin)
    if rc:
        raise MfaError("mfa_params_set_skin: %s (%s)" % (lib.mfa_last_error(None).decode(), ERRORS[rc]),
                       rc)
    return q


EWPLAN_FIELDS = ("k_vectors", "segments", "rows", "nx_max", "ny_max", "nz_max", "row_kernel",
                 "rows_per_group", "force_groups", "force_blocks", "force_grid", "lds_rho",
                 "lds_force", "force_partials", "force_partials_cap", "chunks")


def ewald_plan(params, n_charged):
    """The checked launch plan of the k-space stage (host only).  Returns
    (status, dict of EWPLAN_FIELDS, message)."""
    lib = load()
    out = np.zeros(len(EWPLAN_FIELDS), dtype=np.int64)
    rc = lib.mfa_ewald_plan(C.byref(params), n_charged, out.ctypes.data_as(_lp), len(out))
    msg = lib.mfa_last_error(None).decode() if rc else ""
    return rc, dict(zip(EWPLAN_FIELDS, (int(x) for x in out))), msg


def generate(params, kind, seed, head_z=1.0, bond=0.97, min_dist=0.8):
    lib = load()
    n = params.n_part
    r = np.zeros((n, 3))
    v = np.zeros((n, 3))
    t = np.zeros(n, dtype=np.int32)
    rc = lib.mfa_generate(C.byref(params), kind, seed, head_z, bond, min_dist,
                          _ptr(r, C.c_double), _ptr(v, C.c_double), _ptr(t, C.c_int32))
    if rc:
        raise MfaError("mfa_generate failed (%s)" % ERRORS[rc])
    return r, v, t


class Context:
    """A device context: the state of one system in HBM."""

    def autotune(self, first_step, n_steps):
        t = C.c_int32()
        ms = C.c_double()
        self._chk(self.lib.mfa_autotune(self.h, first_step, n_steps, C.byref(t), C.byref(ms)),
                  "mfa_autotune")
        return t.value, ms.value

    def set_skin(self, skin):
        self._chk(self.lib.mfa_set_skin(self.h, skin), "mfa_set_skin")

    def set_pair_queue(self, mode):
        """0 (or False): walk the list directly; 1 (or True): every entry
        through the pair queue; 2: the near part directly and the skin part
        through the queue (near/shell rows)."""
        self._chk(self.lib.mfa_set_pair_queue(self.h, int(mode)), "mfa_set_pair_queue")

    def set_half_list(self, on):
        self._chk(self.lib.mfa_set_half_list(self.h, 1 if on else 0), "mfa_set_half_list")

    def tuning(self):
        """The current choices: threads per particle, pair queue, list skin,
        half-list form."""
        t, q, sk, hl = C.c_int32(), C.c_int32(), C.c_double(), C.c_int32()
        self._chk(self.lib.mfa_get_tuning(self.h, C.byref(t), C.byref(q), C.byref(sk), C.byref(hl)),
                  "mfa_get_tuning")
        return {"threads_per_particle": t.value, "pair_queue": q.value, "skin": sk.value,
                "half_list": bool(hl.value)}

    def autotune_skin(self, first_step, n_steps, skins):
        """The fastest of `skins` over n_steps whole steps each (advances the
        state); the context keeps it.  Returns (skin, ms per step)."""
        cand = np.ascontiguousarray(skins, dtype=np.float64)
        best, ms = C.c_double(), C.c_double()
        self._chk(self.lib.mfa_autotune_skin(self.h, first_step, n_steps, len(cand),
                                             cand.ctypes.data_as(_dp), C.byref(best), C.byref(ms)),
                  "mfa_autotune_skin")
        return best.value, ms.value

    def set_cells(self, width):
        self._chk(self.lib.mfa_set_cells(self.h, width), "mfa_set_cells")

    def autotune_cells(self, first_step, n_steps, widths, skins):
        """The fastest (cell width, skin) pair over n_steps whole steps each
        (advances the state); the context keeps it.  Returns (width, skin,
        ms per step, the step after the last one run)."""
        w = np.ascontiguousarray(widths, dtype=np.float64)
        sk = np.ascontiguousarray(skins, dtype=np.float64)
        bw, bs, ms, nxt = C.c_double(), C.c_double(), C.c_double(), C.c_int64()
        self._chk(self.lib.mfa_autotune_cells(self.h, first_step, n_steps, len(w),
                                              w.ctypes.data_as(_dp), len(sk), sk.ctypes.data_as(_dp),
                                              C.byref(bw), C.byref(bs), C.byref(ms), C.byref(nxt)),
                  "mfa_autotune_cells")
        return bw.value, bs.value, ms.value, nxt.value

    def __init__(self, params, precision=MFA_FP64, list_capacity=0):
        self.lib = load()
        self.params = params
        self.n = params.n_part
        h = C.c_void_p()
        rc = self.lib.mfa_create(C.byref(params), precision, list_capacity, C.byref(h))
        if rc:
            raise MfaError("mfa_create: %s (%s)" % (
                self.lib.mfa_last_error(None).decode(), ERRORS.get(rc, rc)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.mfa_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, what):
        if rc:
            raise MfaError("%s: %s (%s)" % (what, self.lib.mfa_last_error(self.h).decode(),
                                            ERRORS.get(rc, rc)), rc)

    def set_precision(self, precision):
        self._chk(self.lib.mfa_set_precision(self.h, precision), "mfa_set_precision")

    def set_tpp(self, tpp):
        self._chk(self.lib.mfa_set_threads_per_particle(self.h, tpp), "set_threads_per_particle")

    def set_graph_steps(self, steps):
        self._chk(self.lib.mfa_set_graph_steps(self.h, steps), "mfa_set_graph_steps")

    def time_force(self, reps=100):
        ms = C.c_double()
        self._chk(self.lib.mfa_time_force(self.h, reps, C.byref(ms)), "mfa_time_force")
        return ms.value

    def time_force_step(self, reps=100):
        """Average ms of the force kernel the step runs (fused with the
        integration in the mixed form), the state left as it was."""
        ms = C.c_double()
        self._chk(self.lib.mfa_time_force_step(self.h, reps, C.byref(ms)), "mfa_time_force_step")
        return ms.value

    def time_ewald_parts(self, reps=20):
        """Average ms of the k-space stage's kernels."""
        out = np.zeros(5)
        self._chk(self.lib.mfa_time_ewald_parts(self.h, reps, out.ctypes.data_as(_dp)),
                  "mfa_time_ewald_parts")
        return dict(zip(("tables", "rho", "rho_sum", "force", "apply"), map(float, out)))

    def time_rebuild(self, reps=20, idle=False):
        """Average ms of the cell fill, cell sort and list build of a rebuild
        (idle: of the same kernels with no rebuild due)."""
        out = np.zeros(3)
        fn = self.lib.mfa_time_rebuild_idle if idle else self.lib.mfa_time_rebuild
        self._chk(fn(self.h, reps, out.ctypes.data_as(_dp)), "mfa_time_rebuild")
        return {"cell_fill": float(out[0]), "cell_sort": float(out[1]), "list_build": float(out[2])}

    def time_ewald(self, reps=100):
        """Average duration (ms) of the k-space Ewald stage; F is kept."""
        ms = C.c_double()
        self._chk(self.lib.mfa_time_ewald(self.h, reps, C.byref(ms)), "mfa_time_ewald")
        return ms.value

    def set_state(self, r, v, a, a_type, mic=None):
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        a = np.ascontiguousarray(a, dtype=np.float64)
        t = np.ascontiguousarray(a_type, dtype=np.int32)
        assert r.shape == (self.n, 3) and v.shape == (self.n, 3) and a.shape == (self.n, 3)
        assert t.shape == (self.n,)
        m = None if mic is None else np.ascontiguousarray(mic, dtype=np.int32)
        if m is not None:
            assert m.shape == (self.n, 3)
        self._chk(self.lib.mfa_set_state(self.h, _ptr(r, C.c_double), _ptr(v, C.c_double),
                                         _ptr(a, C.c_double), _ptr(t, C.c_int32),
                                         _ptr(m, C.c_int32)), "mfa_set_state")

    def get_state(self):
        n = self.n
        r, v, a, f = (np.zeros((n, 3)) for _ in range(4))
        mic = np.zeros((n, 3), dtype=np.int32)
        self._chk(self.lib.mfa_get_state(self.h, _ptr(r, C.c_double), _ptr(v, C.c_double),
                                         _ptr(a, C.c_double), _ptr(f, C.c_double),
                                         _ptr(mic, C.c_int32)), "mfa_get_state")
        return {"r": r, "v": v, "a": a, "f": f, "mic": mic}

    def verlet_positions(self):
        self._chk(self.lib.mfa_verlet_positions(self.h), "mfa_verlet_positions")

    def check_skin(self):
        f = C.c_int32()
        self._chk(self.lib.mfa_check_skin(self.h, C.byref(f)), "mfa_check_skin")
        return f.value

    def binning(self):
        self._chk(self.lib.mfa_binning(self.h), "mfa_binning")

    def get_cells(self):
        c = np.zeros((self.n, 3), dtype=np.int32)
        self._chk(self.lib.mfa_get_cells(self.h, _ptr(c, C.c_int32)), "mfa_get_cells")
        return c

    def get_pairs(self):
        m = C.c_int64()
        self._chk(self.lib.mfa_get_pairs(self.h, None, 0, C.byref(m)), "mfa_get_pairs")
        p = np.zeros((m.value, 2), dtype=np.int32)
        self._chk(self.lib.mfa_get_pairs(self.h, _ptr(p, C.c_int32), m.value, C.byref(m)),
                  "mfa_get_pairs")
        return p

    def zero_force(self):
        self._chk(self.lib.mfa_zero_force(self.h), "mfa_zero_force")

    def fluid_fluid(self, step):
        e, q = C.c_double(), C.c_double()
        self._chk(self.lib.mfa_fluid_fluid(self.h, step, C.byref(e), C.byref(q)), "mfa_fluid_fluid")
        return e.value, q.value

    def fluid_wall(self, inter_type=2):
        e = C.c_double()
        self._chk(self.lib.mfa_fluid_wall(self.h, inter_type, C.byref(e)), "mfa_fluid_wall")
        return e.value

    def intra_molec(self):
        e = C.c_double()
        self._chk(self.lib.mfa_intra_molec(self.h, C.byref(e)), "mfa_intra_molec")
        return e.value

    def ewald_k(self, mode=1):
        e = C.c_double()
        self._chk(self.lib.mfa_ewald_k(self.h, mode, C.byref(e)), "mfa_ewald_k")
        return e.value

    def bending(self, mode=2):
        e = C.c_double()
        self._chk(self.lib.mfa_bending(self.h, mode, C.byref(e)), "mfa_bending")
        return e.value

    def dipolar_correction(self):
        e = C.c_double()
        self._chk(self.lib.mfa_dipolar_correction(self.h, C.byref(e)), "mfa_dipolar_correction")
        return e.value

    def verlet_velocities(self):
        t = C.c_double()
        self._chk(self.lib.mfa_verlet_velocities(self.h, C.byref(t)), "mfa_verlet_velocities")
        return t.value

    def obs_init(self, hist_dim=200):
        self._chk(self.lib.mfa_obs_init(self.h, hist_dim), "mfa_obs_init")
        self.hist_dim = hist_dim

    def observe(self, obs, coul_in_total=False):
        """One observation of the current state; obs: the step's Obs (or its
        dict from run_steps)."""
        if isinstance(obs, dict):
            o = Obs()
            for f, _ in Obs._fields_:
                setattr(o, f, obs[f])
            obs = o
        self._chk(self.lib.mfa_observe(self.h, C.byref(obs), 1 if coul_in_total else 0), "mfa_observe")

    def obs_read(self):
        s = ObsSums()
        hd = max(getattr(self, "hist_dim", 0), 0)
        dens = np.zeros((3, max(hd, 1)))
        heads = np.zeros((max(self.params.n_chain, 1), 3))
        self._chk(self.lib.mfa_obs_read(self.h, C.byref(s), _ptr(dens, C.c_double) if hd else None,
                                        _ptr(heads, C.c_double)), "mfa_obs_read")
        d = s.as_dict()
        d["dens_z"] = dens[:, :hd]
        d["f_on_heads"] = heads[:self.params.n_chain]
        return d

    def new_dpd_fd(self):
        t = C.c_double()
        self._chk(self.lib.mfa_new_dpd_fd(self.h, C.byref(t)), "mfa_new_dpd_fd")
        return t.value

    def run_steps(self, first_step, n, want_obs=True):
        obs = Obs()
        self._chk(self.lib.mfa_run_steps(self.h, first_step, n, C.byref(obs) if want_obs else None),
                  "mfa_run_steps")
        return obs.as_dict() if want_obs else None

    def adjoint_window(self, first_step, n_steps, checkpoint_every=20, want_grad=True):
        """Brush height and top-wall pressure averaged over the window and
        (want_grad) their gradients with respect to (epsilon_LJ, sigma,
        k_FENE, R0, gamma_DPD).  Returns J (2,) and G (2, 5) as numpy arrays
        plus the forward pass's rebuild count."""
        out = AdjointOut()
        self._chk(self.lib.mfa_adjoint_window(self.h, first_step, n_steps, checkpoint_every,
                                              1 if want_grad else 0, C.byref(out)),
                  "mfa_adjoint_window")
        J = np.array([out.height, out.p_wall])
        G = np.array([list(out.d_height), list(out.d_p_wall)])
        return J, G, out.n_rebuilds

    def energies(self, step):
        obs = Obs()
        self._chk(self.lib.mfa_energies(self.h, step, C.byref(obs)), "mfa_energies")
        return obs.as_dict()

    def stats(self):
        out = np.zeros(4, dtype=np.int64)
        self._chk(self.lib.mfa_stats(self.h, out.ctypes.data_as(_lp)), "mfa_stats")
        return {"max_neigh": int(out[0]), "capacity": int(out[1]), "error": int(out[2]),
                "rebuilds": int(out[3])}

    def timing(self, enable):
        fm, im = C.c_double(), C.c_double()
        fn, inn = C.c_int64(), C.c_int64()
        self._chk(self.lib.mfa_timing(self.h, int(enable), C.byref(fm), C.byref(fn),
                                      C.byref(im), C.byref(inn)), "mfa_timing")
        return {"force_ms": fm.value, "force_n": fn.value, "integrate_ms": im.value,
                "integrate_n": inn.value}


class Batch(Context):
    """Independent replicas (each its own box, chains and seed) advanced
    together: one launch of each kernel per step covers every replica."""

    def __init__(self, params_list, precision=MFA_FP64, list_capacity=0):
        self.lib = load()
        self.params_list = list(params_list)
        self.params = self.params_list[0]
        self.n = self.params.n_part
        arr = (Params * len(self.params_list))(*self.params_list)
        h = C.c_void_p()
        rc = self.lib.mfa_create_batch(arr, len(self.params_list), precision, list_capacity,
                                       C.byref(h))
        if rc:
            raise MfaError("mfa_create_batch: %s (%s)" % (
                self.lib.mfa_last_error(None).decode(), ERRORS.get(rc, rc)))
        self.h = h

    def set_replica(self, q, r, v, a, a_type, mic=None):
        n = self.params_list[q].n_part
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        a = np.ascontiguousarray(a, dtype=np.float64)
        t = np.ascontiguousarray(a_type, dtype=np.int32)
        assert r.shape == (n, 3) and v.shape == (n, 3) and a.shape == (n, 3) and t.shape == (n,)
        m = None if mic is None else np.ascontiguousarray(mic, dtype=np.int32)
        self._chk(self.lib.mfa_set_state_replica(self.h, q, _ptr(r, C.c_double), _ptr(v, C.c_double),
                                                 _ptr(a, C.c_double), _ptr(t, C.c_int32),
                                                 _ptr(m, C.c_int32)), "mfa_set_state_replica")

    def get_replica(self, q):
        n = self.params_list[q].n_part
        r, v, a, f = (np.zeros((n, 3)) for _ in range(4))
        mic = np.zeros((n, 3), dtype=np.int32)
        self._chk(self.lib.mfa_get_state_replica(self.h, q, _ptr(r, C.c_double), _ptr(v, C.c_double),
                                                 _ptr(a, C.c_double), _ptr(f, C.c_double),
                                                 _ptr(mic, C.c_int32)), "mfa_get_state_replica")
        return {"r": r, "v": v, "a": a, "f": f, "mic": mic}

    def replica_obs(self, q):
        obs = Obs()
        self._chk(self.lib.mfa_replica_obs(self.h, q, C.byref(obs)), "mfa_replica_obs")
        return obs.as_dict()

    def autotune(self, first_step, n_steps):
        t = C.c_int32()
        ms = C.c_double()
        self._chk(self.lib.mfa_autotune(self.h, first_step, n_steps, C.byref(t), C.byref(ms)),
                  "mfa_autotune")
        return t.value, ms.value
