"""torch.autograd face of the design-gradient adjoint.

    J = DesignObjective.apply(rho, solver, nsteps, omega, monitors, origin, eps_bg, eps_scale)

maps a permittivity density rho (a torch tensor shaped like the design box)
to eps = eps_bg + eps_scale * rho, runs the library's forward pass and its
hand-written adjoint (dtm_adjoint_design_gradient) on the GPU, and returns
J = sum_m |DFT_omega Ez_m|^2 as a float64 scalar tensor.  J.backward() then
fills rho.grad with dJ/drho = dJ/d(1/eps) * (-eps_scale / eps^2).  PyTorch is
only the Python-side plumbing: the forward and adjoint sweeps are the HIP
kernels.  The solver is left at step 0 with zero fields after each call.
"""
import numpy as np
import torch


class DesignObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, solver, nsteps, omega, monitors, origin, eps_bg=1.0, eps_scale=11.0):
        r = rho.detach().to("cpu", torch.float64).numpy()
        eps = eps_bg + eps_scale * r
        solver.set_design_eps(*origin, eps)
        J, g_inv = solver.design_gradient(nsteps, omega, monitors)
        grad = torch.from_numpy(g_inv * (-eps_scale / eps ** 2)).to(rho.device, rho.dtype)
        ctx.save_for_backward(grad)
        return torch.tensor(J, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, gJ):
        (grad,) = ctx.saved_tensors
        return (gJ.to(grad.dtype) * grad, None, None, None, None, None, None, None)


def design_objective(rho, solver, nsteps, omega, monitors, origin, eps_bg=1.0, eps_scale=11.0):
    """Functional form of DesignObjective.apply."""
    return DesignObjective.apply(rho, solver, nsteps, omega, np.asarray(monitors), tuple(origin), eps_bg, eps_scale)
