"""The observation stage on the device (mfa_obs_init / mfa_observe /
mfa_obs_read) against the oracle's restatement of observation.f90, and the
Fortran driver's observation files against the reference's own."""
import os
import subprocess

import numpy as np
import pytest

from mfa_amd import configs, inputs as io, lib
from oracle import oracle as orc
from tests.test_ref_pin import compare_obs, read_obs_files
from tests.util import gpu_ctx, make_case, oracle_sys, rel

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "mfa_amd", "mfa_hip_prog")
SUMS = ["v_total_1", "v_total_2", "t_total_1", "t_total_2", "e_total_1", "e_total_2",
        "v_fluid_fluid_1", "t_fluid_1", "e_fluid_fluid_1", "v_tot_intra_1"]


def _observe_both(case, precision, n_steps=20, hist_dim=50):
    inputs, p, r, v, t, a = make_case(case)
    s = oracle_sys(inputs, r, v, a, t)
    s.obs_init(hist_dim)
    ctx = gpu_ctx(p, r, v, a, t, precision=precision)
    ctx.obs_init(hist_dim)
    for it in range(1, n_steps + 1):
        o_ref = s.run_steps(first_step=it, n=1)
        o = ctx.run_steps(it, 1)
        coul = it % 10 == 0 and p.system in (2, 3)
        s.observe(o_ref, coul_in_total=coul)
        ctx.observe(o, coul_in_total=coul)
    out = ctx.obs_read(), s.obs_read(), p
    ctx.close()
    return out


@pytest.mark.parametrize("case", ["brush_small", "melt_small", "pe_brush_small"])
def test_observe_matches_oracle_fp64(case):
    g, o, p = _observe_both(case, lib.MFA_FP64)
    assert g["count_obs"] == o["count_obs"] == 20 and g["count_hist"] == o["count_hist"] == 20
    for k in SUMS:
        assert rel(g[k], o[k]) < 1e-9, (k, g[k], o[k])
    assert np.allclose(g["tot_p"], o["tot_p"], rtol=0, atol=1e-9 * max(1.0, np.abs(o["tot_p"]).max()))
    assert np.array_equal(np.isnan(g["r_g2_mean"]), np.isnan(o["r_g2_mean"]))
    ok = ~np.isnan(o["r_g2_mean"])
    assert np.allclose(g["r_g2_mean"][ok], o["r_g2_mean"][ok], rtol=1e-10, atol=1e-12)
    ok = ~np.isnan(o["r_par_per2"])
    assert np.array_equal(np.isnan(g["r_par_per2"]), np.isnan(o["r_par_per2"]))
    assert np.allclose(g["r_par_per2"][ok], o["r_par_per2"][ok], rtol=1e-10, atol=1e-12)
    assert np.array_equal(g["dens_z"], o["dens_z"])
    assert g["dens_z"].sum() == 20 * p.n_part
    if p.n_chain:
        assert np.allclose(g["f_on_heads"], o["f_on_heads"], rtol=1e-9, atol=1e-9)
    if p.system == 3:
        assert np.all(o["r_g2_mean"][0] > 0)


def test_observe_mixed_form():
    g, o, p = _observe_both("brush_small", lib.MFA_MIXED)
    for k in ("t_total_1", "e_total_1", "v_tot_intra_1"):
        assert rel(g[k], o[k]) < 1e-3, (k, g[k], o[k])
    assert np.allclose(g["r_g2_mean"][0], o["r_g2_mean"][0], rtol=1e-3)
    assert np.abs(g["dens_z"] - o["dens_z"]).sum() <= 0.01 * o["dens_z"].sum()


def test_observe_refused_for_batches():
    inputs, p, r, v, t, a = make_case("brush_small")
    with lib.Batch([p, p], lib.MFA_FP64) as b:
        with pytest.raises(lib.MfaError):
            b.obs_init(10)


@pytest.mark.parametrize("mode", ["compat", "fused"])
def test_driver_observation_files_match_reference(tmp_path, mode):
    """The Fortran driver's fort.40-43, fort.80 and dens_prof.mide against
    the reference's own observation stage on the same input (T = 0), as
    recorded by tests/golden/make_ref_fixtures.py."""
    import hashlib
    gold = os.path.join(ROOT, "tests", "golden", "obs_brush_small")
    assert os.path.exists(PROG), "Fortran driver not built (python -m mfa_amd.build)"
    kind, inp = configs.small_cases()["brush_small"]
    inp.skin = 0.0
    inp.temp = 0.0
    inp.n_relax, inp.n_obser, inp.n_safe = 0, 20, 1000
    p, r, v, t = configs.build(inp, kind)
    dme = str(tmp_path / "hip")
    io.write_run_dir(dme, inp, t, r, v, np.zeros_like(r), precision="fp64", mode=mode)
    digest = hashlib.sha256(open(os.path.join(dme, "conf_old"), "rb").read()).hexdigest()
    assert digest == open(os.path.join(gold, "conf_old.sha256")).read().strip()
    print("driver observation", mode, flush=True)
    res = subprocess.run([PROG], cwd=dme, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    compare_obs(read_obs_files(dme, p.n_chain), read_obs_files(gold, p.n_chain))
