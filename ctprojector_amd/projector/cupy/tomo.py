"""Tomosynthesis distance-driven projector on device tensors (the reference's
src/ct_projector/projector/cupy/tomo.py:15-150)."""

from ctypes import c_float, c_int, c_ulong

import torch

from ... import _lib
from ._dev import dev_f32, device_of, on_device, ptr


def _args(p, img_shape, prj_shape, branchless):
    return (c_ulong(img_shape[0]), c_ulong(img_shape[3]), c_ulong(img_shape[2]), c_ulong(img_shape[1]),
            c_float(p.dx), c_float(p.dy), c_float(p.dz), c_float(p.cx), c_float(p.cy), c_float(p.cz),
            c_ulong(prj_shape[3]), c_ulong(prj_shape[2]), c_ulong(prj_shape[1]),
            c_float(p.du), c_float(p.dv), c_float(p.off_u), c_float(p.off_v), c_int(1 if branchless else 0))


def distance_driven_fp(projector, img, det_center, src, branchless: bool = False) -> torch.Tensor:
    dev = device_of(img)
    img = dev_f32(img, dev)
    det_center = dev_f32(det_center, dev)
    src = dev_f32(src, dev)
    prj = torch.zeros([img.shape[0], det_center.shape[0], projector.nv, projector.nu], dtype=torch.float32,
                      device=dev)
    with on_device(dev):
        err = _lib.module().cupyDistanceDrivenTomoProjection(ptr(prj), ptr(img), ptr(det_center), ptr(src),
                                                             *_args(projector, img.shape, prj.shape, branchless))
    if err != 0:
        print(err)
    return prj


def distance_driven_bp(projector, prj, det_center, src, branchless: bool = False) -> torch.Tensor:
    dev = device_of(prj)
    prj = dev_f32(prj, dev)
    det_center = dev_f32(det_center, dev)
    src = dev_f32(src, dev)
    img = torch.zeros([prj.shape[0], projector.nz, projector.ny, projector.nx], dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().cupyDistanceDrivenTomoBackprojection(ptr(img), ptr(prj), ptr(det_center), ptr(src),
                                                                 *_args(projector, img.shape, prj.shape, branchless))
    if err != 0:
        print(err)
    return img


def filter_fbp(projector, prj, det_center, src, filter_type: str = 'hann', cutoff_x: float = 1,
               cutoff_y: float = 1) -> torch.Tensor:
    """Filter of the projections for tomosynthesis FBP (cupy/tomo.py:157-220):
    prj [batch, nview, nv, nu]; geometry rows (x, y, z); 'hamming', 'hann' or
    Ram-Lak; cutoffs in (0, 1] along u (x) and v (z)."""
    dev = device_of(prj)
    prj = dev_f32(prj, dev)
    det_center = dev_f32(det_center, dev)
    src = dev_f32(src, dev)
    code = {'hamming': 1, 'hann': 2}.get(filter_type.lower(), 0)
    fprj = torch.zeros(prj.shape, dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().cupyFbpTomoFilter(
            ptr(fprj), ptr(prj), ptr(det_center), ptr(src), c_int(prj.shape[0]), c_ulong(prj.shape[3]),
            c_ulong(prj.shape[2]), c_ulong(prj.shape[1]), c_float(projector.du), c_float(projector.dx),
            c_float(projector.dz), c_int(code), c_float(cutoff_x), c_float(cutoff_y))
    if err != 0:
        print(err)
    return fprj
