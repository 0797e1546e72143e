"""Equiangular fan-beam geometry on device tensors (the reference's
src/ct_projector/projector/cupy/fan_equiangular.py).  The kernels receive the
angular pitch du/dsd (:59,114,177,244)."""

from ctypes import c_float, c_int, c_ulong
from typing import Union

import torch

from ... import _lib
from ..geometry import filter_code
from ._dev import dev_f32, device_of, on_device, ptr


def _grid(img_shape, p):
    return (c_ulong(img_shape[0]), c_ulong(img_shape[3]), c_ulong(img_shape[2]), c_ulong(img_shape[1]),
            c_float(p.dx), c_float(p.dy), c_float(p.dz), c_float(p.cx), c_float(p.cy), c_float(p.cz))


def _det(prj_shape, p):
    return (c_ulong(prj_shape[3]), c_ulong(prj_shape[2]), c_ulong(prj_shape[1]), c_float(p.du / p.dsd),
            c_float(p.dv), c_float(p.off_u), c_float(p.off_v), c_float(p.dsd), c_float(p.dso))


def ramp_filter(projector, prj, filter_type: str = 'hann') -> torch.Tensor:
    """Equiangular ramp filtering (cupy/fan_equiangular.py:19-71)."""
    dev = device_of(prj)
    prj = dev_f32(prj, dev)
    fprj = torch.zeros(prj.shape, dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().cupyfbpFanFilter(
            ptr(fprj), ptr(prj), c_ulong(prj.shape[0]), c_ulong(prj.shape[3]), c_ulong(prj.shape[2]),
            c_ulong(prj.shape[1]), c_float(projector.du / projector.dsd), c_float(projector.dv),
            c_float(projector.off_u), c_float(projector.off_v), c_float(projector.dsd), c_float(projector.dso),
            c_int(filter_code(filter_type)))
    if err != 0:
        print(err)
    return fprj


def fbp_bp(projector, prj, angles) -> torch.Tensor:
    """Pixel-driven FBP backprojection (cupy/fan_equiangular.py:74-125).  The
    reference passes device pointers to its host-staging entry point here
    (:95-97); this front end calls the device entry point."""
    dev = device_of(prj, angles)
    prj = dev_f32(prj, dev)
    angles = dev_f32(angles, dev)
    img = torch.zeros([prj.shape[0], projector.nz, projector.ny, projector.nx], dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().cupyfbpFanBackprojection(ptr(img), ptr(prj), ptr(angles), *_grid(img.shape, projector),
                                                     *_det(prj.shape, projector))
    if err != 0:
        print(err)
    return img


def distance_driven_fp(projector, img, angles, branchless: Union[bool, int] = False) -> torch.Tensor:
    """Distance-driven fan-beam forward projection (cupy/fan_equiangular.py:128-189)."""
    dev = device_of(img, angles)
    img = dev_f32(img, dev)
    angles = dev_f32(angles, dev)
    prj = torch.zeros([img.shape[0], angles.shape[0], projector.nv, projector.nu], dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().cupyDistanceDrivenFanProjection(ptr(prj), ptr(img), ptr(angles),
                                                            *_grid(img.shape, projector), *_det(prj.shape, projector),
                                                            c_int(2 if branchless else 0))
    if err != 0:
        print(err)
    return prj


def distance_driven_bp(projector, prj, angles, is_fbp: Union[bool, int] = False,
                       branchless: Union[bool, int] = False) -> torch.Tensor:
    """Distance-driven fan-beam backprojection (cupy/fan_equiangular.py:192-256)."""
    dev = device_of(prj, angles)
    prj = dev_f32(prj, dev)
    angles = dev_f32(angles, dev)
    img = torch.zeros([prj.shape[0], projector.nz, projector.ny, projector.nx], dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().cupyDistanceDrivenFanBackprojection(
            ptr(img), ptr(prj), ptr(angles), *_grid(img.shape, projector), *_det(prj.shape, projector),
            c_int((1 if is_fbp else 0) + (2 if branchless else 0)))
    if err != 0:
        print(err)
    return img


def distance_driven_fp_adjoint(projector, prj, angles, branchless: Union[bool, int] = False) -> torch.Tensor:
    """Exact transpose of ``distance_driven_fp`` (the gradient of the forward
    projection; the reference's backprojector is not this transpose).  No
    counterpart in the reference."""
    dev = device_of(prj, angles)
    prj = dev_f32(prj, dev)
    angles = dev_f32(angles, dev)
    img = torch.zeros([prj.shape[0], projector.nz, projector.ny, projector.nx], dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().ctpDistanceDrivenFanProjectionAdjoint(
            ptr(img), ptr(prj), ptr(angles), *_grid(img.shape, projector), *_det(prj.shape, projector),
            c_int(2 if branchless else 0))
    if err != 0:
        raise RuntimeError('ctpDistanceDrivenFanProjectionAdjoint failed (%d): %s' % (err, _lib.last_error_message()))
    return img
