"""Differentiable projectors and unrolled SQS-TV for PyTorch.

The reference exposes its projectors to TensorFlow as ops whose gradients are
registered in Python: grad(FP) = BP and grad(BP) = FP
(src/ct_projector/projector/tensorflow/cone.py:193-276,
circular_2d.py:140-191,392-443).  Here each projector is a
``torch.autograd.Function`` whose backward is a hand-written HIP kernel:

* Siddon cone beam: the backprojector is the exact transpose of the forward
  projector, so fp / bp are an exact adjoint pair;
* distance-driven fan / parallel beam: the reference's backprojector is not the
  transpose of its forward projector.  ``adjoint='exact'`` (default) uses the
  library's exact transpose (``ctpDistance*ProjectionAdjoint``) as the gradient
  of ``fp``; ``adjoint='reference'`` uses the reference's backprojector, as its
  TensorFlow op does.  ``op.adjoint(y)`` is the exact transpose as a
  differentiable operator (its gradient is ``fp``);
* tomosynthesis: the reference's convention (gradient of fp = bp).

Projector backward passes are themselves built from these Functions, so
higher-order gradients work.  ``tv_sqs`` is the TV surrogate of the HIP prior
kernel with a hand-written HIP vector-Jacobian product (first order only);
``tv_sqs_torch`` is the same surrogate in torch arithmetic.
``sqs_tv_unrolled`` runs K SQS-TV iterations as a differentiable network with
per-iteration checkpointing where memory requires it.
"""

from typing import Callable, Optional, Sequence, Tuple

import torch
import torch.utils.checkpoint
from torch.autograd.function import once_differentiable

from .ddplan import DDPlan
from .prior import cupy as prior_dev
from .projector.cupy import cone, fan_equiangular, parallel, tomo


class _Linear(torch.autograd.Function):
    """y = A x for a linear map given as a pair (A, A^T) of tensor functions."""

    @staticmethod
    def forward(ctx, x, fwd: Callable, adj: Callable):
        ctx.fwd, ctx.adj = fwd, adj
        return fwd(x)

    @staticmethod
    def backward(ctx, g):
        return _Linear.apply(g.contiguous(), ctx.adj, ctx.fwd), None, None


def linear(x: torch.Tensor, fwd: Callable, adj: Callable) -> torch.Tensor:
    """Apply the linear map ``fwd`` with gradient ``adj`` (its transpose)."""
    return _Linear.apply(x, fwd, adj)


class _Op:
    """fp / bp of one geometry as differentiable tensor functions."""

    def _fp(self, x):
        raise NotImplementedError

    def _bp(self, y):
        raise NotImplementedError

    def _adj(self, y):
        return self._bp(y)

    def fp(self, x: torch.Tensor) -> torch.Tensor:
        """Forward projection; gradient = the transpose (exact where available)."""
        return linear(x, self._fp, self._adj)

    def bp(self, y: torch.Tensor) -> torch.Tensor:
        """The reference's backprojection; gradient = forward projection (the
        reference's TensorFlow convention)."""
        return linear(y, self._bp, self._fp)

    def adjoint(self, y: torch.Tensor) -> torch.Tensor:
        """Exact transpose of fp (= bp where bp is exact); gradient = fp."""
        return linear(y, self._adj, self._fp)


class ConeSiddon(_Op):
    """Siddon cone beam, arbitrary trajectory (geometry rows (x, y, z))."""

    def __init__(self, projector, det_center, det_u, det_v, src):
        self.p = projector
        self.geo = (det_center, det_u, det_v, src)

    def _fp(self, x):
        return cone.siddon_fp_arbitrary(self.p, x, *self.geo)

    def _bp(self, y):
        return cone.siddon_bp_arbitrary(self.p, y, *self.geo)


class _DD(_Op):
    """``plan=True`` (default): device tensors go through persistent plans
    (``DDPlan``, one per batch size and device) built on first use, so the
    geometry (projector attributes, angles) is fixed from then on; the calls
    are asynchronous.  ``plan=False`` calls the per-call entry points."""
    mod = None
    fan = True

    def __init__(self, projector, angles, branchless: bool = False, adjoint: str = 'exact', plan: bool = True):
        if adjoint not in ('exact', 'reference'):
            raise ValueError("adjoint must be 'exact' or 'reference'")
        self.p, self.angles, self.branchless, self.mode = projector, angles, branchless, adjoint
        self.use_plan = plan
        self._plans = {}

    def _plan(self, t):
        if not (self.use_plan and t.is_cuda):
            return None
        key = (int(t.shape[0]), t.device)
        pl = self._plans.get(key)
        if pl is None:
            pl = DDPlan(self.p, self.angles, batch=t.shape[0], fan=self.fan, branchless=self.branchless,
                        device=t.device)
            self._plans[key] = pl
        return pl

    def _fp(self, x):
        pl = self._plan(x)
        if pl is not None:
            return pl.fp(x)
        return self.mod.distance_driven_fp(self.p, x, self.angles, branchless=self.branchless)

    def _bp(self, y):
        pl = self._plan(y)
        if pl is not None:
            return pl.bp(y)
        return self.mod.distance_driven_bp(self.p, y, self.angles, branchless=self.branchless)

    def _exact(self, y):
        pl = self._plan(y)
        if pl is not None:
            return pl.adjoint(y)
        return self.mod.distance_driven_fp_adjoint(self.p, y, self.angles, branchless=self.branchless)

    def fp(self, x):
        return linear(x, self._fp, self._exact if self.mode == 'exact' else self._bp)

    def adjoint(self, y):
        return linear(y, self._exact, self._fp)


class FanDD(_DD):
    """Distance-driven equiangular fan beam (projector.du is the arc pitch at dsd)."""
    mod = fan_equiangular


class ParallelDD(_DD):
    """Distance-driven parallel beam."""
    mod = parallel
    fan = False


class TomoDD(_Op):
    """Distance-driven tomosynthesis (the reference's convention: grad fp = bp)."""

    def __init__(self, projector, det_center, src, branchless: bool = False):
        self.p, self.dc, self.src, self.branchless = projector, det_center, src, branchless

    def _fp(self, x):
        return tomo.distance_driven_fp(self.p, x, self.dc, self.src, branchless=self.branchless)

    def _bp(self, y):
        return tomo.distance_driven_bp(self.p, y, self.dc, self.src, branchless=self.branchless)


# --------------------------------------------------------------------- TV
def tv_sqs_torch(img: torch.Tensor, weights: Sequence[float] = (1.0, 1.0, 1.0), eps: float = 1e-8
                 ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """The anisotropic TV surrogate (s1, s2, var) of prior-cuda/TotalVariation.cu
    (the HIP kernel behind ``prior.cupy.tv_sqs``) in torch arithmetic, so that
    autograd differentiates through it.  weights along (z, y, x)."""
    wz, wy, wx = (float(w) for w in weights)

    def back(a, dim):
        return torch.cat([torch.zeros_like(a.narrow(dim, 0, 1)), a.narrow(dim, 1, a.shape[dim] - 1)
                          - a.narrow(dim, 0, a.shape[dim] - 1)], dim)

    dx0, dy0, dz0 = back(img, 3), back(img, 2), back(img, 1)
    var = torch.sqrt(wx * dx0 * dx0 + wy * dy0 * dy0 + wz * dz0 * dz0 + eps)
    s1 = (wx * dx0 + wy * dy0 + wz * dz0) / var
    s2 = (wx + wy + wz) / var
    for dim, w in ((3, wx), (2, wy), (1, wz)):
        n = img.shape[dim]
        d1 = torch.cat([img.narrow(dim, 0, n - 1) - img.narrow(dim, 1, n - 1),
                        torch.zeros_like(img.narrow(dim, 0, 1))], dim)
        vn = torch.cat([var.narrow(dim, 1, n - 1), var.narrow(dim, n - 1, 1)], dim)
        s1 = s1 + w * d1 / vn
        s2 = s2 + w / vn
    return s1, s2, var


class _TVSurrogate(torch.autograd.Function):
    """(s1, s2, var) of the HIP TV kernel; backward = the HIP vector-Jacobian
    product (ctpTVSQSBackward3D)."""

    @staticmethod
    def forward(ctx, img, weights, eps):
        ctx.set_materialize_grads(False)
        ctx.weights, ctx.eps = tuple(float(w) for w in weights), float(eps)
        ctx.save_for_backward(img)
        s1, s2, var = prior_dev.tv_sqs(img, ctx.weights, ctx.eps)
        return s1, s2, var

    @staticmethod
    @once_differentiable
    def backward(ctx, g1, g2, gv):
        img, = ctx.saved_tensors
        if g1 is None and g2 is None and gv is None:
            return None, None, None
        z = None
        if g1 is None or g2 is None:
            z = torch.zeros_like(img)
        d = prior_dev.tv_sqs_backward(img, z if g1 is None else g1, z if g2 is None else g2, ctx.weights, ctx.eps,
                                      gvar=gv)
        return d.to(img.dtype), None, None


def tv_sqs(img: torch.Tensor, weights: Sequence[float] = (1.0, 1.0, 1.0), eps: float = 1e-8
           ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Differentiable TV surrogate (s1, s2, var) of img [batch, nz, ny, nx] on
    the device, weights along (z, y, x): forward and gradient are HIP kernels
    (two passes over the volume each)."""
    return _TVSurrogate.apply(img, weights, eps)


def unrolled_needs_checkpoint(prj: torch.Tensor, x0: torch.Tensor, niter: int) -> bool:
    """Whether sqs_tv_unrolled checkpoints by default: when keeping every
    iteration's intermediates (two sinograms and four volumes, counted
    generously) would take more than a quarter of the device's free memory."""
    if not x0.is_cuda:
        return False
    free, _ = torch.cuda.mem_get_info(x0.device)
    need = niter * (2 * prj.numel() + 4 * x0.numel()) * x0.element_size()
    return need > free // 4


def sqs_tv_unrolled(op: _Op, prj: torch.Tensor, x0: torch.Tensor, norm_img: torch.Tensor, projector_norm: float,
                    beta, niter: int, weights: Sequence[float] = (1.0, 1.0, 1.0), eps: float = 1e-8,
                    weight: Optional[torch.Tensor] = None, checkpoint: Optional[bool] = None,
                    tv: str = 'kernel') -> torch.Tensor:
    """K differentiable SQS-TV iterations (recon/cupy/sqs_algorithms.py:22-57 with
    prior_func = tv_sqs): x <- x - (A^T w (A x - p) / n^2 + beta s1) / (norm_img + beta s2).
    ``op.adjoint`` supplies A^T so that the network's gradient is exact.
    ``checkpoint`` True recomputes each iteration in the backward pass instead
    of keeping its intermediates (two more projector passes per iteration;
    memory per iteration: one volume); False keeps them (about one sinogram
    and four volumes per iteration); None (default) keeps them unless that
    would take more than a quarter of the device's free memory.  ``tv``
    selects the surrogate: 'kernel' (HIP forward and gradient, first-order
    gradients) or 'torch' (torch arithmetic, any order)."""
    if tv not in ('kernel', 'torch'):
        raise ValueError("tv must be 'kernel' or 'torch'")
    if checkpoint is None:
        checkpoint = unrolled_needs_checkpoint(prj, x0, niter)
    w = 1.0 if weight is None else weight
    pn = float(projector_norm)
    surrogate = tv_sqs if tv == 'kernel' else tv_sqs_torch

    def step(x, p):
        r = op.fp(x) / pn - p / pn
        g = op.adjoint(r * w) / pn
        s1, s2, _ = surrogate(x, weights, eps)
        return x - (g + beta * s1) / (norm_img + beta * s2)

    x = x0
    for _ in range(niter):
        if checkpoint and torch.is_grad_enabled():
            x = torch.utils.checkpoint.checkpoint(step, x, prj, use_reentrant=False)
        else:
            x = step(x, prj)
    return x
