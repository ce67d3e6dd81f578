"""Writes the reference-run fixtures the GPU tests compare against, so that
no GPU test has to run the reference itself:

  tests/golden/adjoint_ref_fd.npz
      start state (the small brush relaxed by the oracle's switch-on ramp,
      then passed through the conf_old text, 7 decimals), the parameters
      theta, and the reference's own window observables J(theta) and their
      central finite differences (relative step 1e-4), each from a run of
      oracle/_ref/ref_md_brush (the reference's Fortran, built from its
      sources by oracle/ref/build_ref.sh) at temperature 0 and list skin 0.4;
  tests/golden/conf_new_brush_small/conf_new
      the reference's restart file (store_config(2), oracle/_ref/ref_md_brush
      with `store`) after 20 steps of the seeded small brush at temperature 0,
      with the sha256 of the conf_old it started from;
  tests/golden/conf_new_brush_small/conf_new_hip
      the same restart file written by the Fortran driver on the GPU
      (mfa_amd/mfa_hip_prog, the same run directory): made only with
      --driver, on a machine with a GPU and the library built;
  tests/golden/obs_brush_small/
      the reference's observation files (fort.40-43, fort.80,
      dens_prof.mide) of 20 observed steps of the seeded small brush at
      temperature 0 (oracle/_ref/ref_md_brush_obs), plus the sha256 of the
      conf_old it read, which the test checks before it compares.

    python3 tests/golden/make_ref_fixtures.py            (reference fixtures)
    python3 tests/golden/make_ref_fixtures.py --driver   (conf_new_hip, on a GPU)
"""
import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mfa_amd import configs, inputs as io, lib  # noqa: E402
from oracle import oracle as orc  # noqa: E402

GOLD = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(ROOT, "oracle", "_ref")
ADJ_STEPS = 20
OBS_FILES = ("fort.40", "fort.41", "fort.42", "fort.43", "fort.80", "dens_prof.mide")


def adjoint_inputs():
    """The small brush of the adjoint tests at temperature 0 (no random
    force) with the reference's own list skin."""
    kind, x = configs.small_cases()["brush_small"]
    x.temp = 0.0
    x.skin = 0.0
    return kind, x


def relaxed_start(x, kind):
    """The seeded brush after the switch-on ramp (s_time = 1), run by the
    oracle on the CPU, as the GPU's configs.prepare_brush runs it."""
    p, r, v, t = configs.build(x, kind)
    w = lib.Inputs.from_buffer_copy(x)
    w.s_time = 1
    w.force_switch_on = 1
    w.temp = 1.0
    wp = lib.derive(w)
    s = orc.System(w)
    s.set_state(r, v, np.zeros_like(r), t)
    s.run_steps(1, configs.ramp_steps(wp))
    st = s.get_state()
    return st["r"], st["v"], st["a"], t


def ref_window(exe, x, st, n, workdir):
    """J = (mean z of the type-2 beads, -sum F_z(top)/(Lx Ly)) averaged over
    n steps of the reference's own build."""
    os.makedirs(workdir)
    io.write_run_dir(workdir, x, st["t"], st["r"], st["v"], st["a"])
    res = subprocess.run([exe, "0", str(n), "trace"], cwd=workdir, capture_output=True, text=True,
                         timeout=600)
    if res.returncode != 0:
        raise RuntimeError(res.stdout[-2000:] + res.stderr[-2000:])
    tr = np.fromfile(os.path.join(workdir, "ref_trace.bin"), dtype=np.float64).reshape(n, 2)
    p = lib.derive(x)
    n2 = int((st["t"] == 2).sum())
    return np.array([tr[:, 0].mean() / n2, -tr[:, 1].mean() / (p.boundary[0] * p.boundary[1])])


def make_adjoint_fixture(tmp):
    exe = os.path.join(REF, "ref_md_brush")
    kind, x = adjoint_inputs()
    r, v, a, t = relaxed_start(x, kind)
    d0 = os.path.join(tmp, "conf")
    io.write_run_dir(d0, x, t, r, v, a)
    t2, r2, v2, a2 = io.read_conf(os.path.join(d0, "conf_old"), len(t))
    st = {"r": r2, "v": v2, "a": a2, "t": t2}
    th = configs.theta_of(x)
    J = ref_window(exe, x, st, ADJ_STEPS, os.path.join(tmp, "J"))
    h = 1e-4
    G = np.zeros((2, len(th)))
    for q in range(len(th)):
        tp, tm = th.copy(), th.copy()
        tp[q] = th[q] * (1.0 + h)
        tm[q] = th[q] * (1.0 - h)
        jp = ref_window(exe, configs.with_theta(x, tp), st, ADJ_STEPS, os.path.join(tmp, "p%d" % q))
        jm = ref_window(exe, configs.with_theta(x, tm), st, ADJ_STEPS, os.path.join(tmp, "m%d" % q))
        G[:, q] = (jp - jm) / (tp[q] - tm[q])
    out = os.path.join(GOLD, "adjoint_ref_fd.npz")
    np.savez(out, t=t2, r=r2, v=v2, a=a2, theta=th, J=J, G_fd=G, n_steps=ADJ_STEPS,
             s_time=x.s_time, h=h)
    print("wrote", out, "J", J, "\nG_fd\n", G)


def obs_inputs():
    kind, inp = configs.small_cases()["brush_small"]
    inp.skin = 0.0
    inp.temp = 0.0
    inp.n_relax, inp.n_obser, inp.n_safe = 0, 20, 1000
    return kind, inp


def conf_digest(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def make_obs_fixture(tmp):
    exe = os.path.join(REF, "ref_md_brush_obs")
    kind, inp = obs_inputs()
    p, r, v, t = configs.build(inp, kind)
    d = os.path.join(tmp, "obs")
    io.write_run_dir(d, inp, t, r, v, np.zeros_like(r), precision="fp64", mode="compat")
    res = subprocess.run([exe, "3", "20", "obs"], cwd=d, capture_output=True, text=True, timeout=600)
    if res.returncode != 0:
        raise RuntimeError(res.stdout[-2000:] + res.stderr[-2000:])
    out = os.path.join(GOLD, "obs_brush_small")
    os.makedirs(out, exist_ok=True)
    for f in OBS_FILES:
        shutil.copy(os.path.join(d, f), os.path.join(out, f))
    with open(os.path.join(out, "conf_old.sha256"), "w") as fh:
        fh.write(conf_digest(os.path.join(d, "conf_old")) + "\n")
    print("wrote", out)


def conf_new_inputs():
    kind, inp = configs.small_cases()["brush_small"]
    inp.skin = 0.0
    inp.temp = 0.0
    inp.n_relax, inp.n_obser, inp.n_safe = 0, 20, 20
    return kind, inp


def make_conf_new_fixture(tmp):
    exe = os.path.join(REF, "ref_md_brush")
    kind, inp = conf_new_inputs()
    p, r, v, t = configs.build(inp, kind)
    d = os.path.join(tmp, "store")
    io.write_run_dir(d, inp, t, r, v, np.zeros_like(r), precision="fp64", mode="compat")
    res = subprocess.run([exe, "3", "20", "store"], cwd=d, capture_output=True, text=True,
                         timeout=600)
    if res.returncode != 0:
        raise RuntimeError(res.stdout[-2000:] + res.stderr[-2000:])
    out = os.path.join(GOLD, "conf_new_brush_small")
    os.makedirs(out, exist_ok=True)
    shutil.copy(os.path.join(d, "conf_new"), os.path.join(out, "conf_new"))
    with open(os.path.join(out, "conf_old.sha256"), "w") as fh:
        fh.write(conf_digest(os.path.join(d, "conf_old")) + "\n")
    print("wrote", out)


def make_driver_conf_new(tmp):
    exe = os.path.join(ROOT, "mfa_amd", "mfa_hip_prog")
    kind, inp = conf_new_inputs()
    p, r, v, t = configs.build(inp, kind)
    d = os.path.join(tmp, "driver")
    io.write_run_dir(d, inp, t, r, v, np.zeros_like(r), precision="fp64", mode="compat")
    res = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=600)
    if res.returncode != 0:
        raise RuntimeError(res.stdout[-2000:] + res.stderr[-2000:])
    out = os.path.join(GOLD, "conf_new_brush_small")
    want = open(os.path.join(out, "conf_old.sha256")).read().split()[0]
    if conf_digest(os.path.join(d, "conf_old")) != want:
        raise RuntimeError("the driver's conf_old differs from the reference run's")
    shutil.copy(os.path.join(d, "conf_new"), os.path.join(out, "conf_new_hip"))
    print("wrote", os.path.join(out, "conf_new_hip"))


def main():
    if "--driver" in sys.argv[1:]:
        with tempfile.TemporaryDirectory() as tmp:
            make_driver_conf_new(tmp)
        return
    for prog in ("ref_md_brush", "ref_md_brush_obs"):
        if not os.path.exists(os.path.join(REF, prog)):
            sys.exit("%s missing: run oracle/ref/build_ref.sh first" % prog)
    with tempfile.TemporaryDirectory() as tmp:
        make_adjoint_fixture(tmp)
        make_obs_fixture(tmp)
        make_conf_new_fixture(tmp)


if __name__ == "__main__":
    main()
