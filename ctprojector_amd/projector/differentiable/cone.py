"""Cone-beam Siddon projection with per-sample geometry, differentiable: the
reference's TensorFlow op pair SiddonConeFP / SiddonConeBP
(src/ct_projector/projector/tensorflow/cone.py:19-276,
kernel/projector/tensorflow/siddonCone.cpp:30-340,
kernel/projector/tensorflow/projectorBase.h:34-220) with torch tensors.

Layouts and semantics follow the reference op:

* img [batch, nz, ny, nx, channel] and prj [batch, nview, nv, nu, channel];
  channels share their sample's geometry;
* geometry [batch, nview * 4, 3]: per sample the rows of det_center, det_u,
  det_v and src, each [nview, 3] in (x, y, z);
* grid [batch, 6] or [6]: (dx, dy, dz, cx, cy, cz); detector [batch, 4] or
  [4]: (du, dv, off_u, off_v); both default to the projector's attributes;
* the output shape is default_shape where its entries are positive, else the
  first row of output_shape (projectorBase.h:159-190);
* the gradient of the forward projection is the backprojection with the same
  per-sample parameters and vice versa (cone.py:193-276); here both are exact
  transposes of each other.

The reference's Python wrapper of the backprojection defaults its output shape
to (nx, ny, nz) (cone.py:157) while the op allocates (nz, ny, nx), which
disagree for non-cubic grids; here the default is (nz, ny, nx).  Each sample
is one call of the library's cone-beam entry points (the reference also loops
over the batch, siddonCone.cpp:142).
"""

from types import SimpleNamespace
from typing import List, Optional, Sequence

import numpy as np
import torch

from ..cupy import cone as _cone


def make_geometry(det_center, det_u, det_v, src) -> np.ndarray:
    """[nview * 4, 3] geometry of one sample (cone.py:19-38)."""
    return np.concatenate((det_center, det_u, det_v, src), 0).astype(np.float32)


def _rows(x, n: int, batch: int, default: Sequence[float]) -> List[List[float]]:
    """Per-sample parameter rows: None -> default, [n] -> tiled, [batch, n]."""
    if x is None:
        return [list(map(float, default))] * batch
    a = x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        if a.shape[0] != n:
            raise ValueError('expected %d values, got %s' % (n, a.shape))
        return [list(a)] * batch
    if a.shape != (batch, n):
        raise ValueError('expected shape [%d, %d], got %s' % (batch, n, a.shape))
    return [list(r) for r in a]


def _final_shape(output_shape, default_shape) -> List[int]:
    """default_shape where positive, else the first row of output_shape
    (projectorBase.h:159-190)."""
    first = None
    if output_shape is not None:
        o = output_shape.detach().cpu().numpy() if isinstance(output_shape, torch.Tensor) else np.asarray(output_shape)
        first = [int(v) for v in np.asarray(o).reshape(-1, 3)[0]]
    shape = []
    for i in range(3):
        dflt = int(default_shape[i]) if default_shape is not None else -1
        v = dflt if dflt > 0 else (first[i] if first is not None else -1)
        if v <= 0:
            raise ValueError('either output_shape or default_shape must give a positive size for every dimension')
        shape.append(v)
    return shape


def _sample_projector(grid_row, det_row, img_shape, prj_shape) -> SimpleNamespace:
    nz, ny, nx = img_shape
    nview, nv, nu = prj_shape
    dx, dy, dz, cx, cy, cz = grid_row
    du, dv, off_u, off_v = det_row
    return SimpleNamespace(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, cx=cx, cy=cy, cz=cz, nu=nu, nv=nv, nview=nview,
                           du=du, dv=dv, off_u=off_u, off_v=off_v)


def _split(geo: torch.Tensor, nview: int):
    return geo[:nview], geo[nview:2 * nview], geo[2 * nview:3 * nview], geo[3 * nview:4 * nview]


def _project(x: torch.Tensor, geometry: torch.Tensor, grids, dets, img_shape, prj_shape, forward: bool):
    """Per-sample FP ([B, nz, ny, nx, C] -> [B, nview, nv, nu, C]) or BP."""
    B, C = x.shape[0], x.shape[-1]
    nview = prj_shape[0]
    dev = x.device
    geometry = geometry.to(device=dev, dtype=torch.float32)
    outs = []
    for i in range(B):
        p = _sample_projector(grids[i], dets[i], img_shape, prj_shape)
        g = _split(geometry[i], nview)
        xi = x[i].permute(3, 0, 1, 2).contiguous()                # [C, ...] channels as the library's batch
        yi = _cone.siddon_fp_arbitrary(p, xi, *g) if forward else _cone.siddon_bp_arbitrary(p, xi, *g)
        outs.append(yi.permute(1, 2, 3, 0))
    return torch.stack(outs, 0).contiguous() if B else x.new_zeros((0,) + tuple(prj_shape if forward else img_shape)
                                                                    + (C,))


class _FP(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, geometry, grids, dets, img_shape, prj_shape):
        ctx.save_for_backward(geometry)
        ctx.params = (grids, dets, img_shape, prj_shape)
        return _project(img, geometry, grids, dets, img_shape, prj_shape, True)

    @staticmethod
    def backward(ctx, g):
        geometry, = ctx.saved_tensors
        grids, dets, img_shape, prj_shape = ctx.params
        return _BP.apply(g.contiguous(), geometry, grids, dets, img_shape, prj_shape), None, None, None, None, None


class _BP(torch.autograd.Function):
    @staticmethod
    def forward(ctx, prj, geometry, grids, dets, img_shape, prj_shape):
        ctx.save_for_backward(geometry)
        ctx.params = (grids, dets, img_shape, prj_shape)
        return _project(prj, geometry, grids, dets, img_shape, prj_shape, False)

    @staticmethod
    def backward(ctx, g):
        geometry, = ctx.saved_tensors
        grids, dets, img_shape, prj_shape = ctx.params
        return _FP.apply(g.contiguous(), geometry, grids, dets, img_shape, prj_shape), None, None, None, None, None


def _defaults(projector):
    return ([projector.dx, projector.dy, projector.dz, projector.cx, projector.cy, projector.cz],
            [projector.du, projector.dv, projector.off_u, projector.off_v])


def siddon_fp_arbitrary(projector, img: torch.Tensor, geometry: torch.Tensor, grid=None, detector=None,
                        output_shape=None, default_shape: Optional[Sequence[int]] = None) -> torch.Tensor:
    """Forward projection img [batch, nz, ny, nx, channel] -> prj [batch, nview,
    nv, nu, channel] with per-sample geometry, grid and detector (cone.py:41-108).
    default_shape defaults to (nview from geometry, projector.nv, projector.nu)."""
    if img.dim() != 5 or geometry.dim() != 3 or geometry.shape[0] != img.shape[0] or geometry.shape[2] != 3:
        raise ValueError('img must be [batch, nz, ny, nx, channel] and geometry [batch, nview * 4, 3]')
    B = img.shape[0]
    g0, d0 = _defaults(projector)
    if default_shape is None:
        default_shape = [geometry.shape[1] // 4, projector.nv, projector.nu]
    prj_shape = tuple(_final_shape(output_shape, default_shape))
    if geometry.shape[1] != 4 * prj_shape[0]:
        raise ValueError('geometry must have shape [batch, nview * 4, 3]')
    return _FP.apply(img, geometry, _rows(grid, 6, B, g0), _rows(detector, 4, B, d0), tuple(img.shape[1:4]),
                     prj_shape)


def siddon_bp_arbitrary(projector, prj: torch.Tensor, geometry: torch.Tensor, grid=None, detector=None,
                        output_shape=None, default_shape: Optional[Sequence[int]] = None) -> torch.Tensor:
    """Backprojection prj [batch, nview, nv, nu, channel] -> img [batch, nz, ny,
    nx, channel] (cone.py:111-179); shapes are (nz, ny, nx), default_shape
    defaults to (projector.nz, projector.ny, projector.nx)."""
    if prj.dim() != 5 or geometry.dim() != 3 or geometry.shape[0] != prj.shape[0] or geometry.shape[2] != 3:
        raise ValueError('prj must be [batch, nview, nv, nu, channel] and geometry [batch, nview * 4, 3]')
    if geometry.shape[1] != 4 * prj.shape[1]:
        raise ValueError('geometry must have shape [batch, nview * 4, 3]')
    B = prj.shape[0]
    g0, d0 = _defaults(projector)
    if default_shape is None:
        default_shape = [projector.nz, projector.ny, projector.nx]
    img_shape = tuple(_final_shape(output_shape, default_shape))
    return _BP.apply(prj, geometry, _rows(grid, 6, B, g0), _rows(detector, 4, B, d0), img_shape,
                     tuple(prj.shape[1:4]))


class SiddonFPArbitrary(torch.nn.Module):
    """Module form of siddon_fp_arbitrary (the reference's Keras layer,
    cone.py:280-363)."""

    def __init__(self, projector, default_shape: Optional[Sequence[int]] = None):
        super().__init__()
        self.projector, self.default_shape = projector, default_shape

    def forward(self, img, geometry, grid=None, detector=None, output_shape=None):
        return siddon_fp_arbitrary(self.projector, img, geometry, grid, detector, output_shape, self.default_shape)


class SiddonBPArbitrary(torch.nn.Module):
    """Module form of siddon_bp_arbitrary (cone.py:366-400)."""

    def __init__(self, projector, default_shape: Optional[Sequence[int]] = None):
        super().__init__()
        self.projector, self.default_shape = projector, default_shape

    def forward(self, prj, geometry, grid=None, detector=None, output_shape=None):
        return siddon_bp_arbitrary(self.projector, prj, geometry, grid, detector, output_shape, self.default_shape)
