"""Differentiable SAXS profile: torch.autograd.Function over the engine's
forward (saxs_forward_frames) and adjoint (saxs_backward) entry points.

    I, C = saxs_profile(xyz, engine, elem)      # xyz [frames, atoms, 3]
    loss = chi2(I.mean(0), target, sigma, bins)
    loss.backward()                              # xyz.grad = dL/dx

The reference computes I(q) only; the gradient is an addition that lets a
chi^2 restraint against a measured profile drive coordinate refinement.
"""
from __future__ import annotations

import torch


class _SaxsHistogram(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, engine, elem):
        H, C = engine.forward_frames(xyz.detach(), elem)
        ctx.engine = engine
        ctx.elem = elem
        ctx.generation = engine.generation
        ctx.save_for_backward(xyz.detach())
        ctx.mark_non_differentiable(C)
        return H, C

    @staticmethod
    def backward(ctx, gH, gC):
        (xyz,) = ctx.saved_tensors
        if gH is None:
            return None, None, None
        grad = ctx.engine.backward(xyz, ctx.elem, gH.contiguous(), generation=ctx.generation)
        return grad.to(xyz.dtype), None, None


def saxs_histogram(xyz: torch.Tensor, engine, elem: torch.Tensor):
    """Per-frame (H, C) [frames, num_bins] float64; H is differentiable."""
    if xyz.dim() == 2:
        xyz = xyz.unsqueeze(0)
    return _SaxsHistogram.apply(xyz, engine, elem)


def saxs_profile(xyz: torch.Tensor, engine, elem: torch.Tensor):
    """Per-frame mean intensity per bin I_b = H_b / C_b (0 where C_b = 0)."""
    H, C = saxs_histogram(xyz, engine, elem)
    safe = torch.where(C > 0, C, torch.ones_like(C))
    return torch.where(C > 0, H / safe, torch.zeros_like(H)), C


def chi2(I: torch.Tensor, target: torch.Tensor, sigma: torch.Tensor, bins=None) -> torch.Tensor:
    """chi^2 = sum_b ((I_b - T_b) / s_b)^2 over the selected bins (a boolean
    mask or an index list).  A mask is applied with torch.where, so the loss
    keeps static shapes (no host synchronisation, an elementwise backward);
    unselected bins get s = 1 so that their zero gradient stays finite."""
    if bins is None:
        return (((I - target) / sigma) ** 2).sum()
    if isinstance(bins, torch.Tensor) and bins.dtype == torch.bool:
        s = torch.where(bins, sigma, torch.ones_like(sigma))
        r = torch.where(bins, (I - target) / s, torch.zeros_like(I))
        return (r ** 2).sum()
    I, target, sigma = I[..., bins], target[..., bins], sigma[..., bins]
    return (((I - target) / sigma) ** 2).sum()


class RestraintStep:
    """One restraint step on fixed shapes, captured once as a graph.

        step = RestraintStep(engine, elem, x0, lambda I: chi2(I[0], T, s, sel))
        loss, grad = step(x)        # x [frames, atoms, 3] on the engine's GPU

    Each call copies x into the captured input and replays the engine's
    forward (saxs_forward_frames), the loss, its backward and the engine's
    adjoint (saxs_backward on the forward's spectra): no allocation, no host
    synchronisation and one graph launch per step, where the eager step
    issues some forty launches through autograd.  The returned tensors are
    the graph's own and are overwritten by the next call (clone to keep
    them).  The loss must keep static shapes and synchronise nothing (chi2
    with a boolean bin mask does).  Construction runs `warmup` eager steps
    on a side stream first (the first backward allocates the engine's
    partial-gradient slots); the engine and the loss's tensors must outlive
    the step."""

    def __init__(self, engine, elem: torch.Tensor, x_example: torch.Tensor, loss_fn,
                 warmup: int = 2):
        x = x_example.detach()
        if x.dim() == 2:
            x = x.unsqueeze(0)
        self.x = x.to(device=engine.device, dtype=torch.float32).contiguous().clone()
        self.x.requires_grad_(True)
        self.elem = elem.to(device=engine.device, dtype=torch.int32).contiguous()
        self.engine = engine

        def run():
            I, _ = saxs_profile(self.x, engine, self.elem)
            loss = loss_fn(I)
            loss.backward()
            return loss

        side = torch.cuda.Stream(device=engine.device)
        side.wait_stream(torch.cuda.current_stream(engine.device))
        with torch.cuda.stream(side):
            for _ in range(max(1, warmup)):
                self.x.grad = None
                run()
        torch.cuda.current_stream(engine.device).wait_stream(side)
        self.x.grad = None
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.loss = run()
        self.grad = self.x.grad

    def __call__(self, x: torch.Tensor):
        if x.dim() == 2:
            x = x.unsqueeze(0)
        with torch.no_grad():
            self.x.copy_(x)
        self.graph.replay()
        return self.loss, self.grad
