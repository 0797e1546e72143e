"""Shared helpers of the test suite: small seeded geometries and error metrics."""

import numpy as np

from ctprojector_amd.phantom import (Workload, circular_cone_geometry, ellipse_phantom, ellipsoid_phantom)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


class Geo:
    """A ct_projector-like attribute holder for the oracle."""

    def __init__(self, **kw):
        self.nx = self.ny = self.nz = 1
        self.dx = self.dy = self.dz = 1.0
        self.cx = self.cy = self.cz = 0.0
        self.nu = self.nv = 1
        self.du = self.dv = 1.0
        self.off_u = self.off_v = 0.0
        self.dsd = 1085.6
        self.dso = 595.0
        self.__dict__.update(kw)

    def to(self, proj):
        for k, v in self.__dict__.items():
            setattr(proj, k, v)
        return proj


def small_cone(seed=7, n=24, nview=12, nu=40, nv=32, d=1.0, du=1.6, dv=1.6, off=(0.0, 0.0), c=(0.0, 0.0, 0.0)):
    g = Geo(nx=n, ny=n, nz=n, dx=d, dy=d, dz=d, cx=c[0], cy=c[1], cz=c[2], nu=nu, nv=nv, du=du, dv=dv,
            off_u=off[0], off_v=off[1], dsd=300.0, dso=180.0)
    geo = circular_cone_geometry(nview, 180.0, 300.0, -4.0, 4.0, jitter=1.0, seed=seed, angle_range=2 * np.pi)
    img = ellipsoid_phantom(n, n, n, d, d, d, seed=seed)[None]
    return g, geo, img


def small_fan(n=48, nz=3, nview=40, nu=72, nv=3, dx=1.0, dz=1.0, du=1.5, off=(0.0, 0.0), c=(0.0, 0.0, 0.0), seed=3,
              B=1):
    g = Geo(nx=n, ny=n, nz=nz, dx=dx, dy=dx, dz=dz, cx=c[0], cy=c[1], cz=c[2], nu=nu, nv=nv, du=du, dv=dz,
            off_u=off[0], off_v=off[1], dsd=200.0, dso=120.0)
    # include the 45-degree multiples, where the main-axis switches happen
    angles = (np.arange(nview) * 2 * np.pi / nview).astype(np.float32)
    img = np.stack([ellipsoid_phantom(nz, n, n, dz, dx, dx, seed=seed + b) for b in range(B)], 0)
    return g, angles, img


def small_parallel(n=48, nz=2, nview=36, nu=70, nv=2, dx=1.0, du=1.0, off=(0.0, 0.0), seed=5, B=1):
    g = Geo(nx=n, ny=n, nz=nz, dx=dx, dy=dx, dz=1.0, nu=nu, nv=nv, du=du, dv=1.0, off_u=off[0], off_v=off[1])
    angles = (np.arange(nview) * np.pi / nview).astype(np.float32)
    img = np.stack([ellipsoid_phantom(nz, n, n, 1.0, dx, dx, seed=seed + b) for b in range(B)], 0)
    return g, angles, img
