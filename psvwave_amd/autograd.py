"""``torch.autograd`` over the device adjoint.

``misfit(vp, vs, rho, model)`` returns the L2 misfit of `model`'s observed
data as a torch scalar; its backward runs adjoint sources, the adjoint of
every shot and the kernel mapping (src/fdModel.cpp:789-969, 1147-1183) and
returns the velocity kernels (vp, vs, rho) scaled by the incoming gradient.
This is the reference's gradient -- the adjoint of the continuous problem
sampled every snapshot interval, not the exact discrete adjoint -- so it
agrees with finite differences of the misfit to a few percent.

With ``model.set_checkpointing(k)`` the forward keeps every k-th snapshot
and the backward recomputes the rest (revolve-style, one level), with
identical gradients.
"""
import numpy as np
import torch


class _Misfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, vs, rho, model):
        model.set_parameter_fields(*(t.detach().to("cpu", torch.float64).numpy()
                                     for t in (vp, vs, rho)))
        # forward of every shot with snapshots, then the misfit
        # (src/fdModel.cpp:1291-1297); two HIP streams, bitwise the per-shot calls
        model.run_model(False, False)
        ctx.model = model
        return torch.tensor(model.misfit, dtype=torch.float64, device=vp.device)

    @staticmethod
    def backward(ctx, grad):
        m = ctx.model
        m.calculate_l2_adjoint_sources()
        m.reset_kernels()
        for s in range(m.n_shots):
            m.adjoint_simulate(s)
        m.map_kernels_to_velocity()
        g = float(grad)
        out = [torch.from_numpy(np.ascontiguousarray(k * g)).to(grad.device)
               for k in m.get_kernels()]
        return out[0], out[1], out[2], None


def misfit(vp, vs, rho, model):
    """L2 misfit of fdModel `model` at (vp, vs, rho) (each (nx, nz) float64
    tensors) w.r.t. its observed data, differentiable in vp, vs, rho."""
    return _Misfit.apply(vp, vs, rho, model)
