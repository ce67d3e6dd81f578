"""The oracle's DPD (counter-based noise, Box-Muller deviates) against the
reference's own kT = 1 runs (tests/golden/dpd_brush_stats.json, 16 runs):
two independent oracle runs, the bars of tests/dpd_stats.check.  Pins the
random DPD force of the oracle, which no closed form fixes."""
import json
import os

import numpy as np

from mfa_amd import inputs as io
from oracle import oracle as orc
from tests import dpd_stats as ds

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dpd_brush_stats.json")


def test_oracle_dpd_statistics_match_reference(tmp_path):
    ref = json.load(open(GOLD))["runs"]
    runs = []
    for run in range(2):
        x, p, r, v, t = ds.case(run)
        d = str(tmp_path / ("r%d" % run))
        io.write_run_dir(d, x, t, r, v, np.zeros_like(r))
        t2, r2, v2, a2 = io.read_conf(os.path.join(d, "conf_old"), p.n_part)
        chk = ref[run]["input_checksum"]
        assert abs(ds.checksum(r2, v2) - chk) < 1e-9 * abs(chk)
        s = orc.System(x)
        s.set_state(r2, v2, a2, t2)
        step = x.s_time
        s.run_steps(step, ds.N_EQ)
        step += ds.N_EQ
        snaps = []
        for _ in range(ds.N_RUN // ds.EVERY):
            s.run_steps(step, ds.EVERY)
            step += ds.EVERY
            st = s.get_state()
            snaps.append((st["r"], st["v"]))
        runs.append(ds.reduce_run(snaps, t2, list(p.boundary), p.n_mon, p.n_chain))
    assert ds.check(runs, ref, "oracle")
