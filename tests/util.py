"""Shared helpers of the parity tests: the same seeded state goes to the HIP
path (through the C ABI) and to the oracle."""
import numpy as np

from mfa_amd import configs, lib
from oracle import oracle as orc


def make_case(name):
    cases = configs.small_cases()
    kind, inputs = cases[name]
    p, r, v, t = configs.build(inputs, kind)
    a = np.zeros_like(r)
    return inputs, p, r, v, t, a


def gpu_ctx(p, r, v, a, t, precision=lib.MFA_FP64, tpp=8):
    ctx = lib.Context(p, precision)
    ctx.set_tpp(tpp)
    ctx.set_state(r, v, a, t)
    return ctx


def oracle_sys(inputs, r, v, a, t):
    s = orc.System(inputs)
    s.set_state(r, v, a, t)
    return s


def force_error(f, ref):
    """Per-particle |dF| / max(|F_ref|, 1e-3 F_rms) (max) and the RMS error
    relative to the RMS force."""
    nref = np.linalg.norm(ref, axis=1)
    frms = np.sqrt(np.mean(nref ** 2))
    d = np.linalg.norm(f - ref, axis=1)
    per = d / np.maximum(nref, 1e-3 * frms)
    rms = np.sqrt(np.mean(d ** 2)) / frms
    return float(per.max()), float(rms)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def remove_vcm(v, t, mass_type):
    """Centre-of-mass velocity removal of SYSTEM 2/3 bulk runs
    (init_config.f90:336-356), in the reference's summation order."""
    vc = [0.0, 0.0, 0.0]
    mt = 0.0
    for i in range(len(t)):
        m = mass_type[t[i] - 1]
        for d in range(3):
            vc[d] = vc[d] + m * v[i, d]
        mt = mt + m
    vc = [x / mt for x in vc]
    v = v.copy()
    for d in range(3):
        v[:, d] = v[:, d] - vc[d]
    return v
