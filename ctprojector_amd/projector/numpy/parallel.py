"""Parallel-beam geometry (the reference's src/ct_projector/projector/numpy/parallel.py)."""

from ctypes import POINTER, c_float, c_int, c_ulong
from typing import Union

import numpy as np

from ... import _lib
from ..geometry import filter_code
from .ct_projector import ct_projector


def _p(a):
    return a.ctypes.data_as(POINTER(c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _grid(img_shape, p):
    return (c_ulong(img_shape[0]), c_ulong(img_shape[3]), c_ulong(img_shape[2]), c_ulong(img_shape[1]),
            c_float(p.dx), c_float(p.dy), c_float(p.dz), c_float(p.cx), c_float(p.cy), c_float(p.cz))


def _det(prj_shape, p):
    return (c_ulong(prj_shape[3]), c_ulong(prj_shape[2]), c_ulong(prj_shape[1]),
            c_float(p.du), c_float(p.dv), c_float(p.off_u), c_float(p.off_v))


def ramp_filter(projector: ct_projector, prj: np.array, filter_type: str = 'hann') -> np.array:
    """Ramp filtering for the equispaced parallel detector (parallel.py:17-72)."""
    prj = _f32(prj)
    fprj = np.zeros(prj.shape, np.float32)
    err = _lib.module().cfbpParallelFilter(
        _p(fprj), _p(prj), c_ulong(prj.shape[0]), c_ulong(prj.shape[3]), c_ulong(prj.shape[2]),
        c_ulong(prj.shape[1]), c_float(projector.du), c_float(projector.dv), c_float(projector.off_u),
        c_float(projector.off_v), c_int(filter_code(filter_type)))
    if err != 0:
        print(err)
    return fprj


def pixel_driven_bp(projector: ct_projector, prj: np.array, angles: np.array) -> np.array:
    """Pixel-driven parallel-beam backprojection (parallel.py:75-129):
    prj [batch, nview, nv, nu] -> img [batch, projector.nz, projector.ny, projector.nx]."""
    prj = _f32(prj)
    angles = _f32(angles)
    img = np.zeros([prj.shape[0], projector.nz, projector.ny, projector.nx], np.float32)
    err = _lib.module().cfbpParallelBackprojection(_p(img), _p(prj), _p(angles), *_grid(img.shape, projector),
                                                   *_det(prj.shape, projector), c_int(0))
    if err != 0:
        print(err)
    return img


def distance_driven_fp(projector: ct_projector, img: np.array, angles: np.array,
                       branchless: Union[bool, int] = False) -> np.array:
    """Distance-driven parallel forward projection (parallel.py:131-190)."""
    type_projector = 2 if branchless else 0
    img = _f32(img)
    angles = _f32(angles)
    prj = np.zeros([img.shape[0], len(angles), projector.nv, projector.nu], np.float32)
    err = _lib.module().cDistanceDrivenParallelProjection(_p(prj), _p(img), _p(angles), *_grid(img.shape, projector),
                                                          *_det(prj.shape, projector), c_int(type_projector))
    if err != 0:
        print(err)
    return prj


def distance_driven_bp(projector: ct_projector, prj: np.array, angles: np.array, is_fbp: Union[bool, int] = False,
                       branchless: Union[bool, int] = False) -> np.array:
    """Distance-driven parallel backprojection (parallel.py:193-254)."""
    type_projector = (1 if is_fbp else 0) + (2 if branchless else 0)
    prj = _f32(prj)
    angles = _f32(angles)
    img = np.zeros([prj.shape[0], projector.nz, projector.ny, projector.nx], np.float32)
    err = _lib.module().cDistanceDrivenParallelBackprojection(_p(img), _p(prj), _p(angles),
                                                              *_grid(img.shape, projector),
                                                              *_det(prj.shape, projector), c_int(type_projector))
    if err != 0:
        print(err)
    return img
