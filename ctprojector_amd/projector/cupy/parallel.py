"""Parallel-beam geometry on device tensors (the reference's
src/ct_projector/projector/cupy/parallel.py)."""

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
    return (c_ulong(prj_shape[3]), c_ulong(prj_shape[2]), c_ulong(prj_shape[1]),
            c_float(p.du), c_float(p.dv), c_float(p.off_u), c_float(p.off_v))


def ramp_filter(projector, prj, filter_type: str = 'hann') -> torch.Tensor:
    """Equispaced ramp filtering (cupy/parallel.py:19-70)."""
    dev = device_of(prj)
    prj = dev_f32(prj, dev)
    fprj = torch.zeros(prj.shape, dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().cupyfbpParallelFilter(
            ptr(fprj), ptr(prj), c_ulong(prj.shape[0]), c_ulong(prj.shape[3]), c_ulong(prj.shape[2]),
            c_ulong(prj.shape[1]), c_float(projector.du), c_float(projector.dv), c_float(projector.off_u),
            c_float(projector.off_v), c_int(filter_code(filter_type)))
    if err != 0:
        print(err)
    return fprj


def distance_driven_fp(projector, img, angles, branchless: Union[bool, int] = False) -> torch.Tensor:
    """Distance-driven parallel forward projection (cupy/parallel.py:73-131)."""
    dev = device_of(img, angles)
    img = dev_f32(img, dev)
    angles = dev_f32(angles, dev)
    prj = torch.zeros([img.shape[0], angles.shape[0], projector.nv, projector.nu], dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().cupyDistanceDrivenParallelProjection(
            ptr(prj), ptr(img), ptr(angles), *_grid(img.shape, projector), *_det(prj.shape, projector),
            c_int(2 if branchless else 0))
    if err != 0:
        print(err)
    return prj


def distance_driven_bp(projector, prj, angles, is_fbp: Union[bool, int] = False,
                       branchless: Union[bool, int] = False) -> torch.Tensor:
    """Distance-driven parallel backprojection (cupy/parallel.py:134-194)."""
    dev = device_of(prj, angles)
    prj = dev_f32(prj, dev)
    angles = dev_f32(angles, dev)
    img = torch.zeros([prj.shape[0], projector.nz, projector.ny, projector.nx], dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().cupyDistanceDrivenParallelBackprojection(
            ptr(img), ptr(prj), ptr(angles), *_grid(img.shape, projector), *_det(prj.shape, projector),
            c_int((1 if is_fbp else 0) + (2 if branchless else 0)))
    if err != 0:
        print(err)
    return img


def distance_driven_fp_adjoint(projector, prj, angles, branchless: Union[bool, int] = False) -> torch.Tensor:
    """Exact transpose of ``distance_driven_fp`` (the gradient of the forward
    projection).  No counterpart in the reference."""
    dev = device_of(prj, angles)
    prj = dev_f32(prj, dev)
    angles = dev_f32(angles, dev)
    img = torch.zeros([prj.shape[0], projector.nz, projector.ny, projector.nx], dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().ctpDistanceDrivenParallelProjectionAdjoint(
            ptr(img), ptr(prj), ptr(angles), *_grid(img.shape, projector), *_det(prj.shape, projector),
            c_int(2 if branchless else 0))
    if err != 0:
        raise RuntimeError('ctpDistanceDrivenParallelProjectionAdjoint failed (%d): %s'
                           % (err, _lib.last_error_message()))
    return img
