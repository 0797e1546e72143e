"""TV prior on host arrays: the reference's
src/ct_projector/prior/numpy/recon_prior.py:166-214 (``tv_sqs``), computed by
the HIP library (``cTVSQS3D``).  The NLM priors of that module are not on the
projector path and are not provided."""

from ctypes import POINTER, c_float, c_ulong
from typing import Tuple

import numpy as np

from ... import _lib


def tv_sqs(img: np.ndarray, weights: Tuple[float, float, float], eps: float = 1e-8):
    """Anisotropic TV surrogate of img [batch, nz, ny, nx]; weights along (z, y, x).

    Returns (s1, s2, var): first derivative (numerator), second derivative
    (denominator) and variation map, each [batch, nz, ny, nx] float32."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    s1 = np.zeros(img.shape, np.float32)
    s2 = np.zeros(img.shape, np.float32)
    var = np.zeros(img.shape, np.float32)
    fp = POINTER(c_float)
    err = _lib.module().cTVSQS3D(
        s1.ctypes.data_as(fp), s2.ctypes.data_as(fp), var.ctypes.data_as(fp), img.ctypes.data_as(fp),
        c_float(weights[2]), c_float(weights[1]), c_float(weights[0]),
        c_ulong(img.shape[0]), c_ulong(img.shape[3]), c_ulong(img.shape[2]), c_ulong(img.shape[1]), c_float(eps))
    if err != 0:
        print(err)
    return s1, s2, var
