"""Calibrates the CPU baseline: the oracle (oracle/mfa_oracle.c, the plain-C
restatement that bench.py times as `cpu_baseline`) against the reference's
own Fortran brush path built for speed (oracle/_ref/ref_md_brush_o3, built by
oracle/ref/build_ref.sh with -O3 -march=native), both on one core, on the
same 9,000-bead brush (the brush9k workload of mfa_amd/configs.py: 100
chains of 30 beads, DPD solvent at density 3, 10 x 10 x 20 box, DPD cut-off
1.0, list skin 0.3) from the same start, relaxed first.

The reference program reads the start from conf_old and runs `ref_md
<force_step> <n>` (oracle/ref/ref_md.F90: the loop of md_main.f90:68-147
without observation); its start-up is timed with n = 0 and subtracted.

Test infrastructure only (a measurement of the CPU paths); run in the build
container, where the reference's sources are:

    python3 oracle/calibrate_cpu.py [n_steps]
"""
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from mfa_amd import configs, inputs as io, lib  # noqa: E402
from oracle import oracle as orc  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "ref_md_brush_o3")


def start():
    """The seeded brush relaxed by the reference's switch-on ramp
    (init_force_switch_on.f90:15-26, s_time = 1), run by the oracle, as the
    bench relaxes it on the GPU before timing: the seeded start itself has
    overlaps that the reference's binning cannot survive."""
    w = configs.WORKLOADS["brush9k"]
    inputs = w["make"]()
    p, r, v, t = configs.build(inputs, w["kind"], min_dist=w["min_dist"])
    ramp = lib.Inputs.from_buffer_copy(inputs)
    ramp.s_time = 1
    ramp.force_switch_on = 1
    s = orc.System(ramp)
    s.set_state(r, v, np.zeros_like(r), t)
    s.run_steps(1, configs.ramp_steps(lib.derive(ramp)))
    st = s.get_state()
    return inputs, p, st["r"], st["v"], t


def time_reference(inputs, r, v, t, n_steps):
    with tempfile.TemporaryDirectory() as d:
        io.write_run_dir(d, inputs, t, r, v, np.zeros_like(r))
        secs = []
        for n in (0, n_steps):
            t0 = time.perf_counter()
            res = subprocess.run(["taskset", "-c", "0", EXE, "1", str(n)], cwd=d,
                                 capture_output=True, text=True, timeout=3600)
            secs.append(time.perf_counter() - t0)
            if res.returncode != 0:
                raise RuntimeError(res.stdout[-2000:] + res.stderr[-2000:])
    return secs[1] - secs[0]


def time_oracle(inputs, r, v, t, n_steps):
    os.sched_setaffinity(0, {0})
    s = orc.System(inputs)
    s.set_state(r, v, np.zeros_like(r), t)
    t0 = time.perf_counter()
    s.run_steps(1, n_steps)
    return time.perf_counter() - t0


def main():
    n_steps = int(sys.argv[1]) if len(sys.argv) > 1 else 200
    if not os.path.exists(EXE):
        sys.exit("%s missing: run oracle/ref/build_ref.sh first" % EXE)
    inputs, p, r, v, t = start()
    tr = time_reference(inputs, r, v, t, n_steps)
    to = time_oracle(inputs, r, v, t, n_steps)
    n = p.n_part
    print("brush9k, %d beads, %d steps, one core" % (n, n_steps))
    print("reference (-O3 -march=native): %.3f s, %.3e particle-steps/s" % (tr, n * n_steps / tr))
    print("oracle    (-O3 -march=x86-64-v3): %.3f s, %.3e particle-steps/s" % (to, n * n_steps / to))
    print("oracle / reference throughput: %.3f" % (tr / to))


if __name__ == "__main__":
    main()
