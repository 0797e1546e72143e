"""Cone-beam geometry with arbitrary trajectory and flat panel, Siddon ray
tracing (the reference's src/ct_projector/projector/numpy/cone.py).

Geometry rows are (x, y, z) with x on the fastest image axis: the reference's
kernels read each [nview, 3] row as a float3 (cuda/siddonCone.cu:267) even
though its docstrings say (z, y, x).
"""

from ctypes import POINTER, c_float, c_ulong

import numpy as np

from ... import _lib
from .ct_projector import ct_projector


def _p(a):
    return a.ctypes.data_as(POINTER(c_float))


def siddon_fp_arbitrary(projector: ct_projector, img: np.array, det_center: np.array, det_u: np.array,
                        det_v: np.array, src: np.array) -> np.array:
    """Forward projection; img [batch, nz, ny, nx] (its size overrides projector.n*),
    geometry [nview, 3]; returns [batch, nview, projector.nv, projector.nu]
    (numpy/cone.py:15-88)."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    det_center = np.ascontiguousarray(det_center, dtype=np.float32)
    det_u = np.ascontiguousarray(det_u, dtype=np.float32)
    det_v = np.ascontiguousarray(det_v, dtype=np.float32)
    src = np.ascontiguousarray(src, dtype=np.float32)
    prj = np.zeros([img.shape[0], det_center.shape[0], projector.nv, projector.nu], np.float32)
    err = _lib.module().cSiddonConeProjectionArbitrary(
        _p(prj), _p(img), _p(det_center), _p(det_u), _p(det_v), _p(src),
        c_ulong(img.shape[0]), c_ulong(img.shape[3]), c_ulong(img.shape[2]), c_ulong(img.shape[1]),
        c_float(projector.dx), c_float(projector.dy), c_float(projector.dz),
        c_float(projector.cx), c_float(projector.cy), c_float(projector.cz),
        c_ulong(prj.shape[3]), c_ulong(prj.shape[2]), c_ulong(prj.shape[1]),
        c_float(projector.du), c_float(projector.dv), c_float(projector.off_u), c_float(projector.off_v))
    if err != 0:
        print(err)
    return prj


def siddon_bp_arbitrary(projector: ct_projector, prj: np.array, det_center: np.array, det_u: np.array,
                        det_v: np.array, src: np.array) -> np.array:
    """Backprojection (exact transpose of siddon_fp_arbitrary); prj [batch, nview, nv, nu];
    returns [batch, projector.nz, projector.ny, projector.nx] (numpy/cone.py:91-170)."""
    prj = np.ascontiguousarray(prj, dtype=np.float32)
    det_center = np.ascontiguousarray(det_center, dtype=np.float32)
    det_u = np.ascontiguousarray(det_u, dtype=np.float32)
    det_v = np.ascontiguousarray(det_v, dtype=np.float32)
    src = np.ascontiguousarray(src, dtype=np.float32)
    img = np.zeros([prj.shape[0], projector.nz, projector.ny, projector.nx], np.float32)
    err = _lib.module().cSiddonConeBackprojectionArbitrary(
        _p(img), _p(prj), _p(det_center), _p(det_u), _p(det_v), _p(src),
        c_ulong(img.shape[0]), c_ulong(img.shape[3]), c_ulong(img.shape[2]), c_ulong(img.shape[1]),
        c_float(projector.dx), c_float(projector.dy), c_float(projector.dz),
        c_float(projector.cx), c_float(projector.cy), c_float(projector.cz),
        c_ulong(prj.shape[3]), c_ulong(prj.shape[2]), c_ulong(prj.shape[1]),
        c_float(projector.du), c_float(projector.dv), c_float(projector.off_u), c_float(projector.off_v))
    if err != 0:
        print(err)
    return img
