"""Distance-driven fan / parallel projection with per-sample geometry,
differentiable: the reference's TensorFlow op pair DistanceDriven_2D_FP /
DistanceDriven_2D_BP (src/ct_projector/projector/tensorflow/circular_2d.py,
kernel/projector/tensorflow/distanceDriven2D.cpp:90-520) with torch tensors.

* img [batch, nz, ny, nx, channel], prj [batch, nview, nv, nu, channel];
* angles [batch, nview] (or [nview]), dsd / dso [batch, 1] (fan beam; the
  parallel beam ignores them), grid [batch, 6] (dx, dy, dz, cx, cy, cz),
  detector [batch, 4] (du, dv, off_u, off_v; for the fan beam du is the
  angular pitch, the projector's du / dsd by default);
* the box-integral projector with type_projector 0 (IR) or 4 (force-FBP
  weighting, distanceDriven2D.cpp:145);
* output shape as in the cone op (default_shape where positive, else the first
  row of output_shape);
* gradients as registered by the reference (circular_2d.py:140-192, 392-444):
  the gradient of the forward projection is the reference's backprojection
  with the same parameters, and that of the backprojection is the forward
  projection.  For the exact transpose of the forward projector use
  ctprojector_amd.autograd (FanDD / ParallelDD with adjoint='exact').
"""

from ctypes import c_float, c_int, c_ulong
from enum import IntEnum
from typing import Optional, Sequence

import numpy as np
import torch

from ... import _lib
from ..cupy._dev import on_device, ptr
from .cone import _final_shape, _rows


class TypeGeometry(IntEnum):
    PARALLEL = 0
    FAN_EQUIANGULAR = 1


class TypeProjector(IntEnum):
    IR = 0
    FORCE_FBP = 4


def _angles(angles, batch: int, dev) -> torch.Tensor:
    a = angles if isinstance(angles, torch.Tensor) else torch.as_tensor(np.asarray(angles, np.float32))
    a = a.to(device=dev, dtype=torch.float32)
    if a.dim() == 1:
        a = a[None].expand(batch, -1)
    if a.dim() != 2 or a.shape[0] != batch:
        raise ValueError('angles must be [batch, nview] or [nview]')
    return a.contiguous()


def _call(forward: bool, fan: bool, out, inp, ang, grid_row, det_row, dsd, dso, img_shape, prj_shape, typ):
    nz, ny, nx = img_shape
    nview, nv, nu = prj_shape
    C = inp.shape[0]
    args = [c_ulong(C), c_ulong(nx), c_ulong(ny), c_ulong(nz)] + [c_float(v) for v in grid_row] + \
        [c_ulong(nu), c_ulong(nv), c_ulong(nview)] + [c_float(v) for v in det_row]
    if fan:
        args += [c_float(dsd), c_float(dso)]
    name = ('cupyDistanceDriven%s%s' % ('Fan' if fan else 'Parallel', 'Projection' if forward else 'Backprojection'))
    with on_device(out.device):
        err = getattr(_lib.module(), name)(ptr(out), ptr(inp), ptr(ang), *args, c_int(typ))
    if err != 0:
        raise RuntimeError('%s failed (%d): %s' % (name, err, _lib.last_error_message()))


def _project(x, angles, dsd, dso, grids, dets, img_shape, prj_shape, fan, typ, forward):
    B = x.shape[0]
    outs = []
    for i in range(B):
        xi = x[i].permute(3, 0, 1, 2).contiguous()            # channels as the library's batch
        C = xi.shape[0]
        out = torch.zeros((C,) + (tuple(prj_shape) if forward else tuple(img_shape)), dtype=torch.float32,
                          device=x.device)
        _call(forward, fan, out, xi, angles[i].contiguous(), grids[i], dets[i], dsd[i], dso[i], img_shape,
              prj_shape, typ)
        outs.append(out.permute(1, 2, 3, 0))
    return torch.stack(outs, 0).contiguous()


class _FP(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, angles, dsd, dso, grids, dets, img_shape, prj_shape, fan, typ):
        ctx.save_for_backward(angles)
        ctx.params = (dsd, dso, grids, dets, img_shape, prj_shape, fan, typ)
        return _project(img, angles, dsd, dso, grids, dets, img_shape, prj_shape, fan, typ, True)

    @staticmethod
    def backward(ctx, g):
        angles, = ctx.saved_tensors
        return (_BP.apply(g.contiguous(), angles, *ctx.params),) + (None,) * 9


class _BP(torch.autograd.Function):
    @staticmethod
    def forward(ctx, prj, angles, dsd, dso, grids, dets, img_shape, prj_shape, fan, typ):
        ctx.save_for_backward(angles)
        ctx.params = (dsd, dso, grids, dets, img_shape, prj_shape, fan, typ)
        return _project(prj, angles, dsd, dso, grids, dets, img_shape, prj_shape, fan, typ, False)

    @staticmethod
    def backward(ctx, g):
        angles, = ctx.saved_tensors
        return (_FP.apply(g.contiguous(), angles, *ctx.params),) + (None,) * 9


def _scalars(x, batch: int, default: float):
    """[batch, 1], [batch], a scalar or None (the projector's value)."""
    if x is None:
        return [float(default)] * batch
    a = np.asarray(x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x, np.float64).reshape(-1)
    if a.size == 1:
        return [float(a[0])] * batch
    if a.size != batch:
        raise ValueError('dsd / dso must have shape [batch, 1]')
    return [float(v) for v in a]


def _common(projector, x, angles, dsd, dso, grid, detector, type_geometry, type_projector):
    if x.dim() != 5:
        raise ValueError('expected a [batch, ..., channel] tensor of rank 5')
    if int(type_projector) not in (0, 4):
        raise ValueError('type_projector must be 0 (IR) or 4 (force FBP weight)')
    B = x.shape[0]
    fan = TypeGeometry(type_geometry) == TypeGeometry.FAN_EQUIANGULAR
    ang = _angles(angles, B, x.device)
    if fan:
        dsd_r, dso_r = _scalars(dsd, B, projector.dsd), _scalars(dso, B, projector.dso)
        d0 = [projector.du / projector.dsd, projector.dv, projector.off_u, projector.off_v]
    else:
        dsd_r, dso_r = [1000.0] * B, [500.0] * B
        d0 = [projector.du, projector.dv, projector.off_u, projector.off_v]
    g0 = [projector.dx, projector.dy, projector.dz, projector.cx, projector.cy, projector.cz]
    return ang, dsd_r, dso_r, _rows(grid, 6, B, g0), _rows(detector, 4, B, d0), fan


def distance_driven_2d_fp(projector, img: torch.Tensor, angles, dsd=None, dso=None, grid=None, detector=None,
                          output_shape=None, default_shape: Optional[Sequence[int]] = None,
                          type_geometry: int = TypeGeometry.PARALLEL,
                          type_projector: int = TypeProjector.IR) -> torch.Tensor:
    """img [batch, nz, ny, nx, channel] -> prj [batch, nview, nv, nu, channel]
    (circular_2d.py:36-138); default_shape defaults to (nview, nv, nu)."""
    ang, dsd_r, dso_r, grids, dets, fan = _common(projector, img, angles, dsd, dso, grid, detector, type_geometry,
                                                  type_projector)
    if default_shape is None:
        default_shape = [ang.shape[1], projector.nv, projector.nu]
    prj_shape = tuple(_final_shape(output_shape, default_shape))
    if prj_shape[0] != ang.shape[1]:
        raise ValueError('angles must have shape [batch, nview]')
    return _FP.apply(img, ang, dsd_r, dso_r, grids, dets, tuple(img.shape[1:4]), prj_shape, fan, int(type_projector))


def distance_driven_2d_bp(projector, prj: torch.Tensor, angles, dsd=None, dso=None, grid=None, detector=None,
                          output_shape=None, default_shape: Optional[Sequence[int]] = None,
                          type_geometry: int = TypeGeometry.PARALLEL,
                          type_projector: int = TypeProjector.IR) -> torch.Tensor:
    """prj [batch, nview, nv, nu, channel] -> img [batch, nz, ny, nx, channel]
    (circular_2d.py:288-389); default_shape defaults to (nz, ny, nx)."""
    ang, dsd_r, dso_r, grids, dets, fan = _common(projector, prj, angles, dsd, dso, grid, detector, type_geometry,
                                                  type_projector)
    if ang.shape[1] != prj.shape[1]:
        raise ValueError('angles must have shape [batch, nview]')
    if default_shape is None:
        default_shape = [projector.nz, projector.ny, projector.nx]
    img_shape = tuple(_final_shape(output_shape, default_shape))
    return _BP.apply(prj, ang, dsd_r, dso_r, grids, dets, img_shape, tuple(prj.shape[1:4]), fan,
                     int(type_projector))


class DistanceDriven2DFP(torch.nn.Module):
    """Module form of distance_driven_2d_fp (circular_2d.py:194-285)."""

    def __init__(self, projector, angles=None, type_geometry: int = TypeGeometry.PARALLEL,
                 type_projector: int = TypeProjector.IR, default_shape: Optional[Sequence[int]] = None):
        super().__init__()
        self.projector, self.angles, self.default_shape = projector, angles, default_shape
        self.type_geometry, self.type_projector = type_geometry, type_projector

    def forward(self, img, angles=None, dsd=None, dso=None, grid=None, detector=None, output_shape=None):
        return distance_driven_2d_fp(self.projector, img, self.angles if angles is None else angles, dsd, dso, grid,
                                     detector, output_shape, self.default_shape, self.type_geometry,
                                     self.type_projector)


class DistanceDriven2DBP(torch.nn.Module):
    """Module form of distance_driven_2d_bp (circular_2d.py:446-536)."""

    def __init__(self, projector, angles=None, type_geometry: int = TypeGeometry.PARALLEL,
                 type_projector: int = TypeProjector.IR, default_shape: Optional[Sequence[int]] = None):
        super().__init__()
        self.projector, self.angles, self.default_shape = projector, angles, default_shape
        self.type_geometry, self.type_projector = type_geometry, type_projector

    def forward(self, prj, angles=None, dsd=None, dso=None, grid=None, detector=None, output_shape=None):
        return distance_driven_2d_bp(self.projector, prj, self.angles if angles is None else angles, dsd, dso, grid,
                                     detector, output_shape, self.default_shape, self.type_geometry,
                                     self.type_projector)
