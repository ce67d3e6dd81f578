"""Writers and readers of MFA_PROG's run-time files, line-compatible with the
reference's list-directed readers:
  system_input  init_system.f90:12-144
  mfa_input     init_params.f90:137-244
  conf_old/new  init_config.f90:110-209, store_config.f90:19-61
and `hip_options`, the run-time form of control_simulation.h read by the
Fortran driver (mfa_amd/fortran/mfa_inputs_mod.f90).
"""
import os

import numpy as np

NT = 4


def _f(x):
    return repr(float(x))


def write_system_input(inp, path):
    L = ["# system_input written by mfa_amd.inputs",
         "# line 5: n_mon n_chain n_cell_w_x n_cell_w_y",
         "# the number of lines is fixed",
         "#",
         "%d %d %d %d" % (inp.n_mon, inp.n_chain, inp.n_cell_w_x, inp.n_cell_w_y),
         "# melt/solvent: n_mon_d n_chain_d drop_cell_x drop_cell_y z_skin",
         "%d %d 0 0 0.0" % (inp.n_mon_d, inp.n_chain_d)]
    if inp.particle_4:
        L += ["# particle 4: n_mon_e n_chain_e inter4",
              "%d %d %d" % (inp.n_mon_e, inp.n_chain_e, inp.inter4)]
    L.append("# wall: a_w sigma_w" + (" a_w4 sigma_w4" if inp.particle_4 else ""))
    if inp.particle_4:
        L.append("%s %s %s %s" % (_f(inp.a_w), _f(inp.sigma_w), _f(inp.a_w4), _f(inp.sigma_w4)))
    else:
        L.append("%s %s" % (_f(inp.a_w), _f(inp.sigma_w)))
    L += ["# charges of types 1..4", " ".join(_f(q) for q in inp.charge),
          "# Ewald: alpha_c n_kmax", "%s %d" % (_f(inp.alpha_c), inp.n_kmax)]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_mfa_input(inp, path):
    L = ["%d initial integer starting time" % inp.s_time,
         "%d number of relaxation steps" % inp.n_relax,
         "%d number of observation steps" % inp.n_obser,
         "%d steps between configuration storage" % inp.n_safe,
         "%s time step" % _f(inp.dt),
         "line",
         "%s obsolete" % _f(inp.temp),
         "%s temperature" % _f(inp.temp),
         "%s obsolete" % _f(inp.temp),
         "line",
         "%s spacing in x" % _f(inp.x_space),
         "%s spacing in y" % _f(inp.y_space),
         "%s inter-wall spacing" % _f(inp.z_space_wall),
         "line"]
    for t in range(NT):
        L.append("%s mass" % _f(inp.mass_type[t]))
        for u in range(NT):
            L.append("%s Lennard Jones epsilon" % _f(inp.epsil[t][u]))
        L.append("%s Lennard Jones sigma" % _f(inp.sigma_ii[t]))
        L.append("%s friction" % _f(inp.friction[t]))
    L.append("0.0 spring")
    if inp.bending:
        L += ["%s bending constant" % _f(inp.k_bend), "%s bending angle" % _f(inp.alpha_eq)]
    L += ["line", "%d seed" % inp.iseed, "line",
          "%s FENE k" % _f(inp.k_chain), "%s FENE R0" % _f(inp.r_chain),
          "line", "line", "line", "line", "0 0 0",
          " ".join(_f(x) for x in inp.va_spring_twall), "0.0 0.0 0.0"]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_hip_options(inp, path, precision="fp64", mode="compat"):
    keys = ["system", "symmetry", "solvent", "thermostat", "particle_4", "bending",
            "dpd_cut_off", "skin", "lj_rc", "rc_coul", "ewald_compat", "force_switch_on", "dpd_vv"]
    L = ["# control_simulation.h switches as run-time options"]
    for k in keys:
        val = getattr(inp, k)
        L.append("%s %s" % (k, _f(val) if isinstance(val, float) else str(val)))
    L += ["precision %s" % precision, "mode %s" % mode]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_conf(path, inp, a_type, r, v, a):
    """conf_old / conf_new text (formats init_config.f90:561-568)."""
    n = len(a_type)
    L = ["1", "%10d%10d%10d%10d" % (inp.n_mon, inp.n_chain, inp.n_mon_d, inp.n_chain_d),
         "%10d%10d" % (inp.n_cell_w_x, inp.n_cell_w_y), "%3d" % 2]
    for i in range(n):
        L.append("%3d%13.7f%13.7f%13.7f" % (a_type[i], r[i, 0], r[i, 1], r[i, 2]))
    for i in range(n):
        L.append("%15.7f%15.7f%15.7f" % tuple(v[i]))
    for i in range(n):
        L.append("%15.7f%15.7f%15.7f" % tuple(a[i]))
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def read_conf(path, n):
    """Parse conf text with the fixed-width fields of (i3,3f13.7) / (3f15.7)."""
    with open(path) as f:
        lines = f.read().split("\n")
    t = np.zeros(n, dtype=np.int32)
    r = np.zeros((n, 3))
    v = np.zeros((n, 3))
    a = np.zeros((n, 3))
    k = 4
    for i in range(n):
        s = lines[k + i]
        t[i] = int(s[0:3])
        r[i] = [float(s[3 + 13 * d:16 + 13 * d]) for d in range(3)]
    k += n
    for arr in (v, a):
        for i in range(n):
            s = lines[k + i]
            arr[i] = [float(s[15 * d:15 * (d + 1)]) for d in range(3)]
        k += n
    return t, r, v, a


def read_state_sidecar(path):
    """conf_new.bin: int32 n, then r0, v, a (double 3n each), mic (int32 3n)."""
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    off = 4
    out = {}
    for k in ("r", "v", "a"):
        out[k] = np.frombuffer(raw[off:off + 24 * n], dtype=np.float64).reshape(n, 3).copy()
        off += 24 * n
    out["mic"] = np.frombuffer(raw[off:off + 12 * n], dtype=np.int32).reshape(n, 3).copy()
    return out


def write_run_dir(d, inp, a_type, r, v, a, precision="fp64", mode="compat"):
    os.makedirs(d, exist_ok=True)
    write_system_input(inp, os.path.join(d, "system_input"))
    write_mfa_input(inp, os.path.join(d, "mfa_input"))
    write_hip_options(inp, os.path.join(d, "hip_options"), precision, mode)
    write_conf(os.path.join(d, "conf_old"), inp, a_type, r, v, a)
