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

The backward uses the synthetics and snapshots the model holds.  If the
model has run anything since this misfit's forward (another ``misfit``, a
forward of other parameters), the backward restores the forward's
parameters and recomputes them first; if the observed data changed, it
raises ``RuntimeError``.
"""
import numpy as np
import torch


def _host(t):
    return t.detach().to("cpu", torch.float64).numpy()


class _Misfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, vs, rho, model):
        model.set_parameter_fields(*(_host(t) for t in (vp, vs, rho)))
        # forward of every shot with snapshots, then the misfit
        # (src/fdModel.cpp:1291-1297); two HIP streams, bitwise the per-shot calls
        model.run_model(False, False)
        ctx.model = model
        # the state the backward needs -- parameters, synthetics, snapshots
        # and observed data -- is identified by the model's change counters
        ctx.generation = model._generation
        ctx.obs_generation = model._obs_generation
        ctx.save_for_backward(vp, vs, rho)
        return torch.tensor(model.misfit, dtype=torch.float64, device=vp.device)

    @staticmethod
    def backward(ctx, grad):
        m = ctx.model
        if m._obs_generation != ctx.obs_generation:
            raise RuntimeError("the model's observed data changed after this misfit "
                               "was computed; its gradient is no longer defined")
        if m._generation != ctx.generation:
            # the model ran something else since the forward (a line-search
            # trial, another misfit): restore this forward's parameters and
            # recompute its synthetics and snapshots (deterministic, so
            # bitwise those of the forward)
            vp, vs, rho = ctx.saved_tensors
            m.set_parameter_fields(*(_host(t) for t in (vp, vs, rho)))
            m.run_model(False, False)
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
