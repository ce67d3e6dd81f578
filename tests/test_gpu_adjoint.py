"""The adjoint window (mfa_adjoint_window) on the GPU: gradients of the
time-averaged brush height and top-wall pressure with respect to
(epsilon_LJ, sigma, k_FENE, R0, gamma_DPD) against central finite
differences (relative step 1e-4) of the same forward pass, of the oracle's
trajectories and of the reference's own Fortran build (a committed fixture); through the
torch.autograd wrapper; and invariance under the checkpoint interval.

Bar: every gradient component within 1e-4 relative of the finite
difference (components below 1e-3 of their row's largest count against that
largest value); the GPU and oracle differences use h = 1e-6 (see
test_adjoint_matches_finite_differences), the reference's fixture h = 1e-4
over a 20-step window."""
import os

import numpy as np
import pytest
import torch

from mfa_amd import configs, lib
from oracle import oracle as orc
from tests.adjoint_torch import WindowObservables, central_fd, make_runner

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
W, K = 40, 10


def _grad_close(G, Gfd, tol=1e-4):
    for o in range(2):
        scale = np.abs(Gfd[o]).max()
        for q in range(G.shape[1]):
            ref = Gfd[o, q]
            den = abs(ref) if abs(ref) > 1e-3 * scale else scale
            assert abs(G[o, q] - ref) <= tol * den, (o, q, G[o, q], ref)


@pytest.fixture(scope="module")
def relaxed():
    """The small brush after the switch-on ramp (a start without overlaps)."""
    kind, x = configs.small_cases()["brush_small"]
    p, st = configs.prepare_brush(x, kind, precision=lib.MFA_FP64)
    print("relaxed brush_small: n=%d" % p.n_part, flush=True)
    return x, st


def test_adjoint_matches_finite_differences(relaxed):
    """Central differences at h = 1e-6 within 1e-4, and at h = 1e-4 within
    1e-3: J(theta) is only once differentiable where a pair crosses a
    cut-off inside the step (the WCA force and the DPD weight vanish there
    with a kink), so the h = 1e-4 difference carries an O(h) error that can
    reach a few 1e-4 on the sigma column."""
    x, st = relaxed
    run = make_runner(x, st, 1, W, K)
    th = configs.theta_of(x)
    J, G = run(th)
    fwd = make_runner(x, st, 1, W, K, want_grad=False)
    Gfd = central_fd(fwd, th, h=1e-6)
    Gfd4 = central_fd(fwd, th, h=1e-4)
    print("J", J, "\nadjoint\n", G, "\nFD h=1e-6\n", Gfd, "\nFD h=1e-4\n", Gfd4, flush=True)
    _grad_close(G, Gfd)
    _grad_close(G, Gfd4, tol=1e-3)


def test_checkpoint_interval_and_forward_state(relaxed):
    x, st = relaxed
    th = configs.theta_of(x)
    p = lib.derive(x)
    outs = []
    for k in (W, 7, 1):
        with lib.Context(p, lib.MFA_FP64) as c:
            c.set_state(st["r"], st["v"], st["a"], st["t"], st["mic"])
            J, G, _ = c.adjoint_window(1, W, k, True)
            outs.append((J, G, c.get_state()))
    with lib.Context(p, lib.MFA_FP64) as c:
        c.set_state(st["r"], st["v"], st["a"], st["t"], st["mic"])
        c.run_steps(1, W, want_obs=False)
        ref = c.get_state()
    # the forward pass and the recomputed segments repeat run_steps bit for
    # bit; the adjoint's FP64 sums follow the order of the list rebuilt in the
    # backward sweep, so the gradients agree to rounding
    for J, G, s in outs:
        assert np.array_equal(J, outs[0][0])
        assert np.allclose(G, outs[0][1], rtol=1e-11, atol=1e-14 * np.abs(G).max())
        for key in ("r", "v", "a", "mic"):
            assert np.array_equal(s[key], ref[key]), key


def test_torch_autograd_wrapper(relaxed):
    x, st = relaxed
    run = make_runner(x, st, 1, W, K)
    th0 = configs.theta_of(x)
    theta = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    J = WindowObservables.apply(theta, run)
    loss = 2.0 * J[0] - 0.5 * J[1]
    (g,) = torch.autograd.grad(loss, theta)
    _, G = run(th0)
    assert np.allclose(g.numpy(), 2.0 * G[0] - 0.5 * G[1], rtol=1e-12, atol=0.0)


def test_mixed_forward_gradient_close_to_fp64(relaxed):
    x, st = relaxed
    th = configs.theta_of(x)
    _, G64 = make_runner(x, st, 1, W, K)(th)
    Jm, Gm = make_runner(x, st, 1, W, K, precision=lib.MFA_MIXED)(th)
    print("mixed-forward gradient\n", Gm, flush=True)
    _grad_close(Gm, G64, tol=1e-3)


def _oracle_window(x, st, theta, first, n):
    """The observables from the oracle's trajectory: mean z of the type-2
    beads and -sum F_z(top wall)/(Lx Ly) after each step, averaged."""
    xi = configs.with_theta(x, theta)
    s = orc.System(xi)
    s.set_state(st["r"], st["v"], st["a"], st["t"], st["mic"])
    p = lib.derive(xi)
    t2 = st["t"] == 2
    hs, ps = 0.0, 0.0
    for k in range(n):
        s.run_steps(first + k, 1)
        r = s.get_state()["r"]
        hs += r[t2, 2].mean()
        s.zero_force()
        ps += -s.fluid_wall()[2] / (p.boundary[0] * p.boundary[1])
    return np.array([hs / n, ps / n]), None


def test_adjoint_matches_oracle_finite_differences(relaxed):
    x, st = relaxed
    th = configs.theta_of(x)
    _, G = make_runner(x, st, 1, W, K)(th)
    J_orc, _ = _oracle_window(x, st, th, 1, W)
    J_gpu, _ = make_runner(x, st, 1, W, K, want_grad=False)(th)
    assert np.allclose(J_orc, J_gpu, rtol=1e-10, atol=0.0), (J_orc, J_gpu)
    Gfd = central_fd(lambda t: _oracle_window(x, st, t, 1, W), th, h=1e-6)
    print("oracle FD (h = 1e-6)\n", Gfd, flush=True)
    _grad_close(G, Gfd)


def test_adjoint_matches_reference_finite_differences():
    """FD of the reference's own build (ref_md_brush) at temperature 0 (no
    random force, so its ziggurat stream drops out) and its list skin 0.4,
    recorded by tests/golden/make_ref_fixtures.py from the same start state
    (the relaxed small brush through the conf_old text)."""
    fx = np.load(os.path.join(ROOT, "tests", "golden", "adjoint_ref_fd.npz"))
    kind, x = configs.small_cases()["brush_small"]
    x.temp = 0.0
    x.skin = 0.0
    th = configs.theta_of(x)
    assert np.array_equal(th, fx["theta"])
    n, s_time = int(fx["n_steps"]), int(fx["s_time"])
    st = {"r": fx["r"], "v": fx["v"], "a": fx["a"], "t": fx["t"], "mic": None}
    print("reference FD fixture: n=%d, %d steps" % (len(st["t"]), n), flush=True)
    J_gpu, G = make_runner(x, st, s_time, n, 5)(th)
    print("J ref", fx["J"], "gpu", J_gpu, flush=True)
    assert np.allclose(fx["J"], J_gpu, rtol=1e-8, atol=0.0)
    print("adjoint\n", G, "\nreference FD\n", fx["G_fd"], flush=True)
    _grad_close(G, fx["G_fd"])
