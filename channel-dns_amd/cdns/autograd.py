"""torch.autograd bridge to the backend's discrete adjoint.

`TerminalObjective.apply(u0, force, channel, nsteps, k_ckpt, objective)` runs `nsteps` steps of
the channel from the initial velocity `u0` with the constant body force `force` and returns the
terminal objective (0: kinetic energy, 1: mean wall shear) as a 0-d tensor.  Its backward pass
calls the backend's checkpointed adjoint (`cdns_adjoint_run`, which recomputes the forward
from u0 with checkpoints), so
`torch.autograd.grad(J, (u0, force))` gives the discrete gradients.  The complex spectral
arrays travel as real views (..., 2) of (re, im), so the gradients are plain real gradients
with respect to both parts.
"""
import numpy as np
import torch


def _to_complex(t):
    return np.ascontiguousarray(t.detach().cpu().numpy()).view(np.complex128)[..., 0]


def _to_real(a, like):
    r = np.ascontiguousarray(a).view(np.float64).reshape(a.shape + (2,))
    return torch.as_tensor(r, dtype=torch.float64, device=like.device)


def forward_objective(channel, u, f, nsteps, objective):
    """J(u(T)) by the forward steps alone (zero history, as at a fresh start)."""
    channel.set_body_force(f)
    channel.set_state("velocity", u)
    channel.set_state("velocity_prev", channel.vector())
    channel.set_state("nonlinear_prev", channel.vector())
    channel.set_state("pressure_bc_prev", channel.boundary())
    channel.run(nsteps, it_begin=1)
    J, _ = channel.objective(objective, channel.get_state("velocity"))
    return J


class TerminalObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, force, channel, nsteps, k_ckpt, objective):
        ctx.channel, ctx.nsteps, ctx.k_ckpt, ctx.objective = channel, nsteps, k_ckpt, objective
        ctx.save_for_backward(u0, force)
        J = forward_objective(channel, _to_complex(u0), _to_complex(force), nsteps, objective)
        return torch.tensor(J, dtype=torch.float64, device=u0.device)

    @staticmethod
    def backward(ctx, gJ):
        u0, force = ctx.saved_tensors
        _, gu, gf = ctx.channel.adjoint_run(_to_complex(u0), _to_complex(force), ctx.nsteps, ctx.k_ckpt,
                                            ctx.objective)
        return gJ * _to_real(gu, u0), gJ * _to_real(gf, force), None, None, None, None


def terminal_objective(channel, u0, force, nsteps, k_ckpt=20, objective=0):
    """J(u(T)) as a differentiable function of (u0, force), both real views (..., 3, 2)."""
    return TerminalObjective.apply(u0, force, channel, nsteps, k_ckpt, objective)
