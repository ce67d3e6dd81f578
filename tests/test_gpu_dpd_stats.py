"""DPD with kT = 1 against the reference's own runs: the noise streams differ
(the reference draws a sequential ziggurat stream in pair-visit order, the
HIP path a counter-based deviate per pair and step), so parity is
statistical.  tests/golden/dpd_brush_stats.json holds the time averages of 16
independent reference runs (tests/golden/make_dpd_stats.py); this test makes
8 independent runs of the same inputs on the GPU in each precision form and
applies the bars of tests/dpd_stats.check: the kinetic temperature within two
standard errors, and for the monomer and solvent density profiles and the
monomer pair-distance distribution at most 25% of the bins beyond two
standard errors and a mean chi^2 per bin below 2."""
import json
import os

import numpy as np
import pytest

from mfa_amd import inputs as io, lib
from tests import dpd_stats as ds

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dpd_brush_stats.json")


@pytest.mark.parametrize("precision", [lib.MFA_FP64, lib.MFA_MIXED])
def test_dpd_statistics_match_reference(precision, tmp_path):
    ref = json.load(open(GOLD))["runs"]
    runs = []
    for run in range(8):
        x, p, r, v, t = ds.case(run)
        d = str(tmp_path / ("r%d" % run))
        io.write_run_dir(d, x, t, r, v, np.zeros_like(r))
        t2, r2, v2, a2 = io.read_conf(os.path.join(d, "conf_old"), p.n_part)
        chk = ref[run]["input_checksum"]
        assert abs(ds.checksum(r2, v2) - chk) < 1e-9 * abs(chk)
        snaps = []
        with lib.Context(p, precision) as c:
            c.set_state(r2, v2, a2, t2)
            step = x.s_time
            c.run_steps(step, ds.N_EQ, want_obs=False)
            step += ds.N_EQ
            for _ in range(ds.N_RUN // ds.EVERY):
                c.run_steps(step, ds.EVERY, want_obs=False)
                step += ds.EVERY
                st = c.get_state()
                snaps.append((st["r"], st["v"]))
        runs.append(ds.reduce_run(snaps, t2, list(p.boundary), p.n_mon, p.n_chain))
    assert ds.check(runs, ref, "gpu precision %d" % precision)
