"""Scan-geometry holder shared by the numpy and device front ends.

Mirrors the ``ct_projector`` class of the reference
(src/ct_projector/projector/numpy/ct_projector.py:27-195 and
src/ct_projector/projector/cupy/ct_projector.py:27-195): the same attribute
names and defaults, the INI loader (keys ending in ``_mm`` lose the suffix,
every value is cast to float32, then the ``n*`` sizes to int), the projector /
backprojector binding through ``set_projector`` / ``set_backprojector`` and the
power-method norm.  The array backend is supplied by the subclass.

Difference from the reference: ``calc_projector_norm`` draws its start vector
from ``numpy.random.default_rng(seed)`` (seed 0 by default) so that the norm is
reproducible; the reference uses the unseeded global generator
(src/ct_projector/projector/cupy/ct_projector.py:157).
"""

import configparser
from typing import Callable

import numpy as np


class GeometryBase:
    def __init__(self):
        self.nview = 720
        self.rotview = 720
        self.nu = 512
        self.nv = 512
        self.nx = 512
        self.ny = 512
        self.nz = 512
        self.dx = 1
        self.dy = 1
        self.dz = 1
        self.cx = 0
        self.cy = 0
        self.cz = 0
        self.dsd = 1085.6
        self.dso = 595
        self.du = 1.2
        self.dv = 1.2
        self.off_u = 0
        self.off_v = 0
        self.fp_kwargs = {}
        self.bp_kwargs = {}

    # ------------------------------------------------------------ config
    def from_file(self, filename: str) -> None:
        """Load the [geometry] section of an INI file (example/projector_fan.cfg)."""
        self.geometry = configparser.ConfigParser()
        _ = self.geometry.read(filename)
        for o in self.geometry['geometry']:
            if '_mm' in o:
                setattr(self, o[:-3], np.float32(self.geometry['geometry'][o]))
            else:
                setattr(self, o, np.float32(self.geometry['geometry'][o]))
        self.nview = int(self.nview)
        self.nu = int(self.nu)
        self.nv = int(self.nv)
        self.nx = int(self.nx)
        self.ny = int(self.ny)
        self.nz = int(self.nz)

    # --------------------------------------------------------- binding
    def set_projector(self, projector: Callable, **kwargs) -> None:
        """Bind ``projector(self, img, **kwargs)`` as ``self.fp``."""
        self.projector = projector
        self.fp_kwargs = kwargs

    def set_backprojector(self, backprojector: Callable, **kwargs) -> None:
        """Bind ``backprojector(self, prj, **kwargs)`` as ``self.bp``."""
        self.backprojector = backprojector
        self.bp_kwargs = kwargs

    def fp(self, img, **kwargs):
        for k in self.fp_kwargs:
            if k not in kwargs:
                kwargs[k] = self.fp_kwargs[k]
        return self.projector(self, img, **kwargs)

    def bp(self, prj, **kwargs):
        for k in self.bp_kwargs:
            if k not in kwargs:
                kwargs[k] = self.bp_kwargs[k]
        return self.backprojector(self, prj, **kwargs)

    # ---------------------------------------------------------- norms
    def _uniform(self, shape, seed):
        raise NotImplementedError

    def _ones(self, shape):
        raise NotImplementedError

    def _norm(self, x):
        raise NotImplementedError

    def _sqrt(self, x):
        raise NotImplementedError

    def calc_projector_norm(self, weight=None, niter: int = 10, seed: int = 0, verbose: bool = True) -> float:
        """Power method for ||A|| (or ||sqrt(w) A||)."""
        weight = self._sqrt(weight) if weight is not None else 1
        x = self._uniform([1, self.nz, self.ny, self.nx], seed)
        x = x / self._norm(x)
        norm = 0.0
        for i in range(niter):
            if verbose:
                print(i, end=',', flush=True)
            fp = self.fp(x)
            norm = self._norm(fp)
            x = self.bp(fp * weight)
            x = x / self._norm(x)
        if verbose:
            print('')
        return norm

    def calc_norm_img(self, weight=None):
        """norm_img = A^T w A 1."""
        if weight is None:
            weight = 1
        x = self._ones([1, self.nz, self.ny, self.nx])
        return self.bp(self.fp(x) * weight)

    def get_angles(self) -> np.ndarray:
        """Angles of a full circular scan of ``nview`` views."""
        return np.arange(0, self.nview, dtype=np.float32) * 2 * np.pi / self.nview


FILTER_CODES = {'hamming': 1, 'hann': 2, 'cosine': 3}


def filter_code(filter_type: str) -> int:
    """Filter-name map of the reference (numpy/fan_equiangluar.py:39-46): unknown names give RL."""
    return FILTER_CODES.get(filter_type.lower(), 0)
