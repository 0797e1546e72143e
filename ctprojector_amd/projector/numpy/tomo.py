"""Tomosynthesis geometry, distance driven with z as the main axis and detector
axes u = (1,0,0), v = (0,1,0) (the reference's src/ct_projector/projector/numpy/tomo.py)."""

from ctypes import POINTER, c_float, c_int, c_ulong

import numpy as np

from ... import _lib
from .ct_projector import ct_projector


def _p(a):
    return a.ctypes.data_as(POINTER(c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _args(p, img_shape, prj_shape, branchless):
    return (c_ulong(img_shape[0]), c_ulong(img_shape[3]), c_ulong(img_shape[2]), c_ulong(img_shape[1]),
            c_float(p.dx), c_float(p.dy), c_float(p.dz), c_float(p.cx), c_float(p.cy), c_float(p.cz),
            c_ulong(prj_shape[3]), c_ulong(prj_shape[2]), c_ulong(prj_shape[1]),
            c_float(p.du), c_float(p.dv), c_float(p.off_u), c_float(p.off_v), c_int(1 if branchless else 0))


def distance_driven_fp(projector: ct_projector, img: np.array, det_center: np.array, src: np.array,
                       branchless: bool = False) -> np.array:
    """Forward projection; returns [batch, nview, projector.nv, projector.nu] (tomo.py:19-90)."""
    img = _f32(img)
    det_center = _f32(det_center)
    src = _f32(src)
    prj = np.zeros([img.shape[0], det_center.shape[0], projector.nv, projector.nu], np.float32)
    err = _lib.module().cDistanceDrivenTomoProjection(_p(prj), _p(img), _p(det_center), _p(src),
                                                      *_args(projector, img.shape, prj.shape, branchless))
    if err != 0:
        print(err)
    return prj


def distance_driven_bp(projector: ct_projector, prj: np.array, det_center: np.array, src: np.array,
                       branchless: bool = False) -> np.array:
    """Backprojection; returns [batch, projector.nz, projector.ny, projector.nx] (tomo.py:93-165)."""
    prj = _f32(prj)
    det_center = _f32(det_center)
    src = _f32(src)
    img = np.zeros([prj.shape[0], projector.nz, projector.ny, projector.nx], np.float32)
    err = _lib.module().cDistanceDrivenTomoBackprojection(_p(img), _p(prj), _p(det_center), _p(src),
                                                          *_args(projector, img.shape, prj.shape, branchless))
    if err != 0:
        print(err)
    return img
