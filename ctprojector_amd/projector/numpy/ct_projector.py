"""Host-array projector wrapper: the reference's
src/ct_projector/projector/numpy/ct_projector.py.  Arrays are staged through
the GPU on every call (the ``c*`` entry points of libprojector.so)."""

from ctypes import c_int

import numpy as np

from ... import _lib
from ..geometry import GeometryBase


def set_device(device: int) -> int:
    """Set the computing device and return any error code (SetDevice)."""
    return _lib.module().SetDevice(c_int(device))


class ct_projector(GeometryBase):
    """CT projector wrapper (numpy arrays)."""

    def _uniform(self, shape, seed):
        return np.random.default_rng(seed).uniform(size=shape).astype(np.float32)

    def _ones(self, shape):
        return np.ones(shape, dtype=np.float32)

    def _norm(self, x):
        return np.linalg.norm(x)

    def _sqrt(self, x):
        return np.sqrt(x)
