"""torch.autograd front end of the differentiable rollout.

``rollout(compliance, velocity, r)`` runs ``r`` (a ``xpbd.Rollout``) forward on
the GPU and returns the final body-local vertex positions (V x 3); its
backward pass is the library's hand-written adjoint.  Gradients flow to the
per-edge compliance (E,) and the initial velocities (V, 3).  Tensors live on
the device the context was created on, in float32; the kernels run on torch's
current stream.
"""
from __future__ import annotations

import torch


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _check(t, n, name):
    if t.device.type != "cuda" or t.dtype != torch.float32 or t.numel() != n:
        raise ValueError(f"{name}: expected a float32 device tensor of {n} elements, "
                         f"got {t.dtype} {tuple(t.shape)} on {t.device}")


class RolloutFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, compliance, velocity, r):
        _check(compliance, r.num_edges, "compliance")
        _check(velocity, 3 * r.num_vertices, "velocity")
        compliance = compliance.detach().contiguous()
        velocity = velocity.detach().contiguous()
        positions = torch.empty((r.num_vertices, 3), dtype=torch.float32, device=compliance.device)
        r.forward(compliance.data_ptr(), velocity.data_ptr(), positions.data_ptr(), _stream())
        ctx.r = r
        ctx.generation = r.generation
        ctx.shapes = (compliance.shape, velocity.shape)
        ctx.keep = (compliance, velocity)
        return positions

    @staticmethod
    def backward(ctx, grad_positions):
        r = ctx.r
        grad_positions = grad_positions.detach().to(torch.float32).contiguous()
        device = grad_positions.device
        if r.generation != ctx.generation:
            # the rollout ran again since: restore this pass's checkpoints first
            compliance, velocity = ctx.keep
            scratch = torch.empty((r.num_vertices, 3), dtype=torch.float32, device=device)
            r.forward(compliance.data_ptr(), velocity.data_ptr(), scratch.data_ptr(), _stream())
            ctx.generation = r.generation
        grad_compliance = torch.empty(r.num_edges, dtype=torch.float32, device=device)
        grad_velocity = torch.empty(3 * r.num_vertices, dtype=torch.float32, device=device)
        r.backward(grad_positions.data_ptr(), grad_compliance.data_ptr(), grad_velocity.data_ptr(), _stream())
        return grad_compliance.view(ctx.shapes[0]), grad_velocity.view(ctx.shapes[1]), None


def rollout(compliance: torch.Tensor, velocity: torch.Tensor, r) -> torch.Tensor:
    """Final positions (V, 3) of the rollout ``r``; differentiable in both arguments."""
    return RolloutFunction.apply(compliance, velocity, r)
