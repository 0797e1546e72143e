"""Device-tensor projector wrapper (the reference's
src/ct_projector/projector/cupy/ct_projector.py, torch tensors for cupy arrays)."""

from ctypes import c_int

import numpy as np
import torch

from ... import _lib
from ..geometry import GeometryBase


def set_device(device: int) -> int:
    """Set the computing device (for the library and torch) and return any error code."""
    err = _lib.module().SetDevice(c_int(device))
    if err == 0:
        torch.cuda.set_device(device)
    return err


class ct_projector(GeometryBase):
    """CT projector wrapper (torch tensors on a HIP device)."""

    def _dev(self):
        return torch.device('cuda', torch.cuda.current_device())

    def _uniform(self, shape, seed):
        x = np.random.default_rng(seed).uniform(size=shape).astype(np.float32)
        return torch.from_numpy(x).to(self._dev())

    def _ones(self, shape):
        return torch.ones(shape, dtype=torch.float32, device=self._dev())

    def _norm(self, x):
        return float(torch.linalg.vector_norm(x.double()))

    def _sqrt(self, x):
        return torch.sqrt(x)
