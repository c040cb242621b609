"""torch.autograd binding of the reaction-diffusion objective.

`RDObjective.apply(x, solver, mode)` evaluates J(x) for every ensemble member with the HIP forward
and adjoint solves (DerivativeOperatorsRD::evaluateObjectiveAndGradient,
src/grad/DerivativeOperatorsRD.cpp:130-226) and keeps dJ for the backward pass, so that
`J.sum().backward()` leaves the adjoint gradient in `x.grad`.  x is a float64 tensor of shape
(nmembers, np + 2) holding [p | kappa | rho] per member (all members need the same number of bases;
use RDSolver.objective_and_gradient directly otherwise).  The gradient is the exact gradient of the
discrete objective, computed by one forward and one adjoint solve; no autograd graph is built
through the PDE, and the checkpoint policy of the solver (store-all or binomial checkpointing)
decides the memory it uses.
"""
import numpy as np
import torch

from . import _lib


class RDObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, solver, mode=_lib.MODE_ALL):
        if x.dim() != 2 or x.shape[0] != solver.nmembers:
            raise ValueError("x must have shape (nmembers, np + 2) = (%d, np + 2)" % solver.nmembers)
        xs = x.detach().to(device="cpu", dtype=torch.float64).numpy()
        J, g = solver.objective_and_gradient([xs[m] for m in range(xs.shape[0])], mode)
        ctx.save_for_backward(torch.from_numpy(np.stack(g)).to(device=x.device, dtype=x.dtype))
        return torch.from_numpy(J).to(device=x.device, dtype=x.dtype)

    @staticmethod
    def backward(ctx, grad_j):
        (g,) = ctx.saved_tensors
        return grad_j[:, None] * g, None, None


def rd_objective(x, solver, mode=_lib.MODE_ALL):
    """J(x) per member as a differentiable torch tensor (see RDObjective)."""
    return RDObjective.apply(x, solver, mode)
