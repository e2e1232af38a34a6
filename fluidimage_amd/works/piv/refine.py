"""Differentiable refinement of the final PIV field (an extension: the
reference stops after the last multipass, src/fluidimage/works/piv/multipass.py:185-188).

The displacements (dx, dy) at the nodes of the final vector grid give the
loss of a pair: the sum over its interrogation windows (n0 x n1, centred on
the nodes) of 1 - NCC between im0(p - d/2) and im1(p + d/2), where d is the
bilinear interpolation of the nodes at pixel p and the images are resampled
by Keys cubic convolution (the windowed normalised SSD is 2 (1 - NCC)).
``RefineLoss`` is a ``torch.autograd.Function`` whose forward and backward
are the hand-written kernels ``piv_refine_forward`` / ``piv_refine_backward``
(FP64, per-window recomputation in the backward).

With a ``TpsNodeField`` the unknowns are TPS node values instead: the node
displacements are their ThinPlateSplineSubdom smoothing (the multipass's
TPS without outlier iterations, a linear map T of the values), and
``RefineTpsLoss`` chains T (``piv_tps_node_field``) before the loss and T^T
after its backward.

``refine`` runs gradient steps from a starting field (the multipass result)
and keeps a step for a pair only when it lowers that pair's loss; the
multipass field is the fallback.
"""

import ctypes

import numpy as np

from ... import _native
from .deform import DeformedPass


class RefineProblem:
    """Images of npairs pairs and the final pass's grid and window shape."""

    def __init__(self, work, imageset, npairs):
        self.work = work
        self.imageset = imageset
        self.npairs = int(npairs)
        self.n0, self.n1 = (int(n) for n in work.shape_crop_im0)
        self.nvec = work.nb_vectors
        self.tables = DeformedPass(work, imageset.device)  # grid record + cell tables

    def forward(self, dx, dy, stream=None):
        import torch

        dev = self.imageset.device
        loss = torch.zeros(self.npairs, dtype=torch.float64, device=dev)
        stats = torch.empty(self.npairs * self.nvec * 5, dtype=torch.float64, device=dev)
        lib = _native.load()
        _native.check(
            lib.piv_refine_forward(int(self.imageset.dtype),
                                   ctypes.c_void_p(self.imageset.table.data_ptr()), self.npairs,
                                   ctypes.byref(self.tables.grid), self.n0, self.n1,
                                   ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()),
                                   ctypes.c_void_p(stats.data_ptr()),
                                   ctypes.c_void_p(loss.data_ptr()), _native.stream_ptr(stream)),
            "piv_refine_forward")
        return loss, stats

    def backward(self, dx, dy, stats, grad_out, stream=None):
        import torch

        gdx = torch.zeros_like(dx)
        gdy = torch.zeros_like(dy)
        lib = _native.load()
        _native.check(
            lib.piv_refine_backward(int(self.imageset.dtype),
                                    ctypes.c_void_p(self.imageset.table.data_ptr()), self.npairs,
                                    ctypes.byref(self.tables.grid), self.n0, self.n1,
                                    ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()),
                                    ctypes.c_void_p(stats.data_ptr()),
                                    ctypes.c_void_p(grad_out.data_ptr()),
                                    ctypes.c_void_p(gdx.data_ptr()),
                                    ctypes.c_void_p(gdy.data_ptr()), _native.stream_ptr(stream)),
            "piv_refine_backward")
        return gdx, gdy


def _refine_loss_class():
    import torch

    class RefineLoss(torch.autograd.Function):
        """loss[k] (float64, one per pair) of the node displacements
        dx, dy (npairs x nvec float64 device tensors)."""

        @staticmethod
        def forward(ctx, dx, dy, problem):
            dx = dx.contiguous()
            dy = dy.contiguous()
            loss, stats = problem.forward(dx, dy)
            ctx.save_for_backward(dx, dy, stats)
            ctx.problem = problem
            return loss

        @staticmethod
        def backward(ctx, grad_out):
            dx, dy, stats = ctx.saved_tensors
            g = grad_out.to(dtype=dx.dtype).contiguous()
            gdx, gdy = ctx.problem.backward(dx, dy, stats, g)
            return gdx, gdy, None

    return RefineLoss


def _refine_tps_loss_class():
    import torch

    class RefineTpsLoss(torch.autograd.Function):
        """loss[k] of the TPS node values vx, vy (npairs x nvec float64):
        the loss of the node displacements T (vx, vy)."""

        @staticmethod
        def forward(ctx, vx, vy, problem, field):
            d = field.forward(torch.stack([vx.reshape(-1), vy.reshape(-1)]))
            dx = d[0].view(vx.shape)
            dy = d[1].view(vy.shape)
            loss, stats = problem.forward(dx, dy)
            ctx.save_for_backward(dx, dy, stats)
            ctx.problem = problem
            ctx.field = field
            return loss

        @staticmethod
        def backward(ctx, grad_out):
            dx, dy, stats = ctx.saved_tensors
            g = grad_out.to(dtype=dx.dtype).contiguous()
            gdx, gdy = ctx.problem.backward(dx, dy, stats, g)
            gv = ctx.field.adjoint(torch.stack([gdx.reshape(-1), gdy.reshape(-1)]))
            return gv[0].view(dx.shape), gv[1].view(dy.shape), None, None

    return RefineTpsLoss


_RefineLoss = None
_RefineTpsLoss = None


def refine_loss(dx, dy, problem, field=None):
    """Differentiable loss per pair (autograd over the HIP kernels): of the
    node displacements (dx, dy), or with ``field`` (a TpsNodeField on the
    grid nodes) of the TPS node values (dx, dy)."""
    global _RefineLoss, _RefineTpsLoss
    if field is not None:
        if _RefineTpsLoss is None:
            _RefineTpsLoss = _refine_tps_loss_class()
        return _RefineTpsLoss.apply(dx, dy, problem, field)
    if _RefineLoss is None:
        _RefineLoss = _refine_loss_class()
    return _RefineLoss.apply(dx, dy, problem)


def node_field(problem, subdom_size, smoothing_coef):
    """The TpsNodeField of the problem's grid nodes."""
    from ...calcul.tps_field import TpsNodeField

    w = problem.work
    return TpsNodeField(w.ixvecs_grid, w.iyvecs_grid, subdom_size, smoothing_coef,
                        problem.imageset.device)


def field_of(field, vx, vy):
    """Node displacements (dx, dy) of the node values (vx, vy)."""
    import torch

    d = field.forward(torch.stack([vx.reshape(-1), vy.reshape(-1)]))
    return d[0].view(vx.shape), d[1].view(vy.shape)


def refine(problem, dx0, dy0, steps=20, step_px=0.05, field=None):
    """Gradient steps from the multipass field (dx0, dy0) (npairs x nvec
    float64). Per pair, the step moves the unknown of largest gradient by
    step_px; it is kept only if the pair's loss drops (then the step grows
    by 1.2), else it is discarded and halved.

    Without ``field`` the unknowns are the node displacements. With a
    TpsNodeField they are TPS node values starting from (dx0, dy0), the node
    displacements being T (values); a pair's result is the field of the
    best values only where its loss is below that of the multipass field
    (dx0, dy0) itself, which stays the fallback.

    Returns (dx, dy, loss0, loss, accepted): the node displacements kept,
    the loss of the multipass field and of the kept field per pair, and the
    number of steps that lowered the kept loss."""
    import torch

    dx = dx0.detach().clone().double()
    dy = dy0.detach().clone().double()
    n = problem.npairs
    eta = torch.full((n, 1), float(step_px), dtype=torch.float64, device=dx.device)
    accepted = torch.zeros(n, dtype=torch.int64, device=dx.device)
    with torch.no_grad():
        loss0, _ = problem.forward(dx, dy)
        if field is None:
            loss = loss0.clone()
        else:
            loss, _ = problem.forward(*field_of(field, dx, dy))
    best = loss0.clone()
    bx, by = dx.clone(), dy.clone()
    for _ in range(steps):
        x = dx.clone().requires_grad_(True)
        y = dy.clone().requires_grad_(True)
        cur = refine_loss(x, y, problem, field)
        gx, gy = torch.autograd.grad(cur.sum(), (x, y))
        gmax = torch.maximum(gx.abs().amax(1, keepdim=True), gy.abs().amax(1, keepdim=True))
        scale = eta / torch.clamp(gmax, min=1e-30)
        tx = dx - scale * gx
        ty = dy - scale * gy
        with torch.no_grad():
            fx, fy = (tx, ty) if field is None else field_of(field, tx, ty)
            trial, _ = problem.forward(fx, fy)
        better = (trial < loss)
        b1 = better[:, None]
        dx = torch.where(b1, tx, dx)
        dy = torch.where(b1, ty, dy)
        loss = torch.where(better, trial, loss)
        eta = torch.where(b1, eta * 1.2, eta * 0.5)
        kept = trial < best
        k1 = kept[:, None]
        bx = torch.where(k1, fx, bx)
        by = torch.where(k1, fy, by)
        best = torch.where(kept, trial, best)
        accepted += kept.to(torch.int64)
    return bx, by, loss0, best, accepted
