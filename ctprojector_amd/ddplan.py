"""Persistent distance-driven plans: one fan-beam (equiangular) or parallel-beam
geometry and batch count whose tables are built once (``ctpDDPlan*`` of
libprojector.so), so that every later projection launches kernels only, with
no host work, no device-to-host copy of the angles and no synchronisation.

The reference rebuilds its projector object, copies its inputs and
synchronises on every call (cuda/distanceDriven2D.cu:109-634); iterative
reconstructions and unrolled networks call the same geometry thousands of
times, which is what a plan is for.

    plan = DDPlan(projector, angles, batch=8, fan=True)
    prj = plan.fp(img)           # = fan_equiangular.distance_driven_fp(projector, img, angles)
    img = plan.bp(prj)           # = fan_equiangular.distance_driven_bp(...)
    img = plan.adjoint(prj)      # = fan_equiangular.distance_driven_fp_adjoint(...)

Calls are asynchronous on torch's current stream of the plan's device.  Each
direction keeps its scratch in the plan, so the calls can be captured in a
HIP graph (torch.cuda.graph); calls of one direction on one plan must not
overlap in time (use one stream, or order the streams).  The projector's
attributes (nx..cz, nu, nv, du, dv, off_u, off_v, dsd, dso) are read once, at
construction.
"""

import ctypes
from ctypes import c_float, c_int, c_ulong, c_void_p
from typing import Optional, Union

import numpy as np
import torch

from . import _lib


def _host_angles(angles) -> np.ndarray:
    if isinstance(angles, torch.Tensor):
        angles = angles.detach().to('cpu', torch.float32).numpy()
    return np.ascontiguousarray(np.asarray(angles, dtype=np.float32).ravel())


class DDPlan:
    def __init__(self, projector, angles, batch: int = 1, fan: bool = True, branchless: Union[bool, int] = False,
                 is_fbp: Union[bool, int] = False, device=None):
        self.device = torch.device('cuda', torch.cuda.current_device()) if device is None else torch.device(device)
        p = projector
        self.batch, self.fan = int(batch), bool(fan)
        self._angles = _host_angles(angles)
        self.nview = int(self._angles.shape[0])
        self.img_shape = (self.batch, int(p.nz), int(p.ny), int(p.nx))
        self.prj_shape = (self.batch, self.nview, int(p.nv), int(p.nu))
        da = p.du / p.dsd if self.fan else p.du
        dsd, dso = (p.dsd, p.dso) if self.fan else (0.0, 0.0)
        typ = (1 if is_fbp else 0) + (2 if branchless else 0)
        self._plan = c_void_p()
        with torch.cuda.device(self.device):
            err = _lib.module().ctpDDPlanCreate(
                ctypes.byref(self._plan), c_int(1 if self.fan else 0), self._angles.ctypes.data_as(c_void_p),
                c_ulong(self.batch), c_ulong(p.nx), c_ulong(p.ny), c_ulong(p.nz), c_float(p.dx), c_float(p.dy),
                c_float(p.dz), c_float(p.cx), c_float(p.cy), c_float(p.cz), c_ulong(p.nu), c_ulong(p.nv),
                c_ulong(self.nview), c_float(da), c_float(p.dv), c_float(p.off_u), c_float(p.off_v), c_float(dsd),
                c_float(dso), c_int(typ))
        if err != 0:
            raise RuntimeError('ctpDDPlanCreate failed (%d): %s' % (err, _lib.last_error_message()))

    def _stream(self):
        return c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _input(self, x: torch.Tensor, shape, what: str) -> torch.Tensor:
        if tuple(x.shape) != tuple(shape):
            raise ValueError('%s: expected shape %s, got %s' % (what, tuple(shape), tuple(x.shape)))
        return x.to(device=self.device, dtype=torch.float32).contiguous()

    def _output(self, out: Optional[torch.Tensor], shape, zero: bool) -> torch.Tensor:
        if out is None:
            return (torch.zeros if zero else torch.empty)(shape, dtype=torch.float32, device=self.device)
        if tuple(out.shape) != tuple(shape) or out.dtype != torch.float32 or not out.is_contiguous() \
                or out.device != self.device:
            raise ValueError('out: expected a contiguous float32 tensor of shape %s on %s' % (tuple(shape),
                                                                                            self.device))
        if zero:
            out.zero_()
        return out

    def _call(self, name, out, inp):
        with torch.cuda.device(self.device):
            err = getattr(_lib.module(), name)(self._plan, c_void_p(out.data_ptr()), c_void_p(inp.data_ptr()),
                                               self._stream())
        if err != 0:
            raise RuntimeError('%s failed (%d): %s' % (name, err, _lib.last_error_message()))

    def fp(self, img: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Forward projection [batch, nview, nv, nu] of img [batch, nz, ny, nx]
        (every element of the result is written)."""
        img = self._input(img, self.img_shape, 'img')
        prj = self._output(out, self.prj_shape, zero=False)
        self._call('ctpDDPlanForward', prj, img)
        return prj

    def bp(self, prj: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """The reference's backprojection of prj (into a zeroed image)."""
        prj = self._input(prj, self.prj_shape, 'prj')
        img = self._output(out, self.img_shape, zero=True)
        self._call('ctpDDPlanBackward', img, prj)
        return img

    def adjoint(self, prj: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """The exact transpose of ``fp`` applied to prj (into a zeroed image)."""
        prj = self._input(prj, self.prj_shape, 'prj')
        img = self._output(out, self.img_shape, zero=True)
        self._call('ctpDDPlanAdjoint', img, prj)
        return img

    def close(self) -> None:
        if getattr(self, '_plan', None) is not None and self._plan.value:
            _lib.module().ctpDDPlanDestroy(self._plan)
            self._plan = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
