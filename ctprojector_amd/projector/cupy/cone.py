"""Cone-beam arbitrary-trajectory Siddon projector on device tensors (the
reference's src/ct_projector/projector/cupy/cone.py).  Geometry rows are
(x, y, z); see the numpy front end."""

from ctypes import c_float, c_ulong

import torch

from ... import _lib
from ._dev import dev_f32, device_of, on_device, ptr


def _args(p, img_shape, prj_shape):
    return (c_ulong(img_shape[0]), c_ulong(img_shape[3]), c_ulong(img_shape[2]), c_ulong(img_shape[1]),
            c_float(p.dx), c_float(p.dy), c_float(p.dz), c_float(p.cx), c_float(p.cy), c_float(p.cz),
            c_ulong(prj_shape[3]), c_ulong(prj_shape[2]), c_ulong(prj_shape[1]),
            c_float(p.du), c_float(p.dv), c_float(p.off_u), c_float(p.off_v))


def siddon_fp_arbitrary(projector, img, det_center, det_u, det_v, src) -> torch.Tensor:
    """img [batch, nz, ny, nx] -> prj [batch, nview, projector.nv, projector.nu] (cupy/cone.py:15-85)."""
    dev = device_of(img)
    img = dev_f32(img, dev)
    g = [dev_f32(a, dev) for a in (det_center, det_u, det_v, src)]
    prj = torch.zeros([img.shape[0], g[0].shape[0], projector.nv, projector.nu], dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().cupySiddonConeProjectionArbitrary(ptr(prj), ptr(img), *[ptr(a) for a in g],
                                                              *_args(projector, img.shape, prj.shape))
    if err != 0:
        print(err)
    return prj


def siddon_bp_arbitrary(projector, prj, det_center, det_u, det_v, src) -> torch.Tensor:
    """prj [batch, nview, nv, nu] -> img [batch, projector.nz, projector.ny, projector.nx], the exact
    transpose of siddon_fp_arbitrary (cupy/cone.py:88-157)."""
    dev = device_of(prj)
    prj = dev_f32(prj, dev)
    g = [dev_f32(a, dev) for a in (det_center, det_u, det_v, src)]
    img = torch.zeros([prj.shape[0], projector.nz, projector.ny, projector.nx], dtype=torch.float32, device=dev)
    with on_device(dev):
        err = _lib.module().cupySiddonConeBackprojectionArbitrary(ptr(img), ptr(prj), *[ptr(a) for a in g],
                                                                  *_args(projector, img.shape, prj.shape))
    if err != 0:
        print(err)
    return img
