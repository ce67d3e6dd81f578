"""mfa_verlet_velocities_noalloc (no allocation, caller-owned device arrays)
against mfa_verlet_velocities on the same state, through the C ABI: two
calls in a row, the second on a smaller system, each bit-identical in v, a
and F and in the kinetic energy (the same kernels and the same partial sums).
Also: sizes beyond the scratch limit are refused before any device work."""
import ctypes as C

import numpy as np
import pytest
import torch

from mfa_amd import lib
from tests.util import gpu_ctx, make_case

pytestmark = pytest.mark.gpu


def _reference(case):
    """F of a force pass, then verlet_velocities through the context."""
    inputs, p, r, v, t, a = make_case(case)
    ctx = gpu_ctx(p, r, v, a, t)
    ctx.energies(step=3)
    st0 = ctx.get_state()
    tk = ctx.verlet_velocities()
    st1 = ctx.get_state()
    ctx.close()
    return p, t, st0, st1, tk


def _noalloc(p, t, st0):
    dev = torch.device("cuda:0")
    f = torch.from_numpy(st0["f"].copy()).to(dev)
    v = torch.from_numpy(st0["v"].copy()).to(dev)
    a = torch.from_numpy(st0["a"].copy()).to(dev)
    ty = torch.from_numpy(np.ascontiguousarray(t, dtype=np.int32)).to(dev)
    torch.cuda.synchronize()
    tk = C.c_double()
    L = lib.load()
    rc = L.mfa_verlet_velocities_noalloc(C.byref(p), f.data_ptr(), v.data_ptr(), a.data_ptr(),
                                         ty.data_ptr(), None, C.byref(tk))
    assert rc == 0, L.mfa_last_error(None).decode()
    torch.cuda.synchronize()
    return f.cpu().numpy(), v.cpu().numpy(), a.cpu().numpy(), tk.value


def test_noalloc_matches_verlet_velocities_twice():
    for case in ("melt_small", "brush_small"):  # 600 beads, then 380 (pinned heads)
        p, t, st0, st1, tk_ref = _reference(case)
        print("no-alloc verlet_velocities: %s, n=%d" % (case, p.n_part), flush=True)
        f, v, a, tk = _noalloc(p, t, st0)
        assert np.array_equal(v, st1["v"]) and np.array_equal(a, st1["a"])
        assert np.array_equal(f, st1["f"])
        assert tk == tk_ref, (tk, tk_ref)


def test_noalloc_refuses_sizes_beyond_the_scratch():
    inputs, p, r, v, t, a = make_case("melt_small")
    q = lib.Params.from_buffer_copy(p)
    q.n_part = (1 << 20) + 1
    L = lib.load()
    tk = C.c_double()
    dummy = C.c_void_p(1)  # never dereferenced: the size check comes first
    rc = L.mfa_verlet_velocities_noalloc(C.byref(q), dummy, dummy, dummy, dummy, None, C.byref(tk))
    assert rc == 1 and "scratch limit" in L.mfa_last_error(None).decode()
