"""torch.autograd wrapper of the adjoint window (test harness).

`WindowObservables.apply(theta, run)` returns the two window observables
(brush height, top-wall pressure) as a float64 tensor and, on backward,
contracts the incoming gradient with the Jacobian mfa_adjoint_window
computed on the GPU.  `run(theta_numpy) -> (J (2,), G (2, 5))` does the
call through the C ABI.
"""
import numpy as np
import torch

from mfa_amd import configs, lib


class WindowObservables(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, run):
        J, G = run(theta.detach().cpu().numpy().astype(np.float64))
        ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(G)))
        return torch.from_numpy(np.ascontiguousarray(J))

    @staticmethod
    def backward(ctx, grad_out):
        (G,) = ctx.saved_tensors
        return grad_out.to(G.dtype) @ G, None


def make_runner(inputs, state, first_step, n_steps, checkpoint_every,
                precision=lib.MFA_FP64, want_grad=True):
    """A `run` for WindowObservables: a fresh context with the parameters
    theta, the given start state, one adjoint window."""
    def run(theta):
        p = lib.derive(configs.with_theta(inputs, theta))
        with lib.Context(p, precision) as c:
            c.set_state(state["r"], state["v"], state["a"], state["t"], state.get("mic"))
            J, G, _ = c.adjoint_window(first_step, n_steps, checkpoint_every, want_grad)
        return J, G
    return run


def central_fd(run, theta, h=1e-4):
    """Central differences of the observables, relative step h."""
    G = np.zeros((2, len(theta)))
    for q in range(len(theta)):
        tp = theta.copy()
        tm = theta.copy()
        tp[q] = theta[q] * (1.0 + h)
        tm[q] = theta[q] * (1.0 - h)
        G[:, q] = (run(tp)[0] - run(tm)[0]) / (tp[q] - tm[q])
    return G
