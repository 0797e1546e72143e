"""Seeded synthetic phantoms and scan geometries for the benchmark workloads.

There is no network and the reference ships no data (its notebooks read
``forbild512.mat``, ``3.nii.gz`` ... which are not in the tree), so every
workload here is synthetic:

* ``ellipse_phantom``   - 2D ellipses rasterised at pixel centres, later ones
  overwrite earlier ones;
* ``ellipsoid_phantom`` - the same rule in 3D with a rotation about z;
* geometry builders for the parallel, fan and cone workloads.

Seeds follow ``numpy.random.default_rng(1000 + N)`` for workload N and
``1000 + 10 N + j`` for sample j of a batched workload.
"""

from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np


@dataclass
class Ellipse:
    cx: float
    cy: float
    a: float
    b: float
    phi: float
    value: float


@dataclass
class Ellipsoid:
    cx: float
    cy: float
    cz: float
    a: float
    b: float
    c: float
    phi: float
    value: float


def random_ellipses(rng: np.random.Generator, n: int, radius: float, vmax: float = 0.02) -> List[Ellipse]:
    """Centres U(-0.6R, 0.6R), semi-axes U(0.05R, 0.4R), angle U(0, pi), value U(0, vmax)."""
    out = []
    for _ in range(n):
        cx, cy = rng.uniform(-0.6 * radius, 0.6 * radius, 2)
        a, b = rng.uniform(0.05 * radius, 0.4 * radius, 2)
        phi = rng.uniform(0, np.pi)
        val = rng.uniform(0, vmax)
        out.append(Ellipse(cx, cy, a, b, phi, val))
    return out


def random_ellipsoids(rng: np.random.Generator, n: int, radius: float, zradius: float,
                      vmax: float = 0.02) -> List[Ellipsoid]:
    out = []
    for _ in range(n):
        cx, cy = rng.uniform(-0.6 * radius, 0.6 * radius, 2)
        cz = rng.uniform(-0.6 * zradius, 0.6 * zradius)
        a, b = rng.uniform(0.05 * radius, 0.4 * radius, 2)
        c = rng.uniform(0.05 * zradius, 0.4 * zradius) if zradius > 0 else 1.0
        phi = rng.uniform(0, np.pi)
        val = rng.uniform(0, vmax)
        out.append(Ellipsoid(cx, cy, cz, a, b, c, phi, val))
    return out


def _centres(n, d, c=0.0):
    # voxel centre x = (ix - (n-1)/2) * d + c   (reference example/readme.md:18-22)
    return (np.arange(n) - (n - 1) / 2.0) * d + c


def rasterize_ellipses(ells: List[Ellipse], nx: int, ny: int, dx: float, dy: float,
                       cx: float = 0.0, cy: float = 0.0) -> np.ndarray:
    x = _centres(nx, dx, cx)[None, :]
    y = _centres(ny, dy, cy)[:, None]
    img = np.zeros((ny, nx), np.float32)
    for e in ells:
        c, s = np.cos(e.phi), np.sin(e.phi)
        xr = (x - e.cx) * c + (y - e.cy) * s
        yr = -(x - e.cx) * s + (y - e.cy) * c
        img[(xr / e.a) ** 2 + (yr / e.b) ** 2 <= 1] = e.value
    return img


def rasterize_ellipsoids(ells: List[Ellipsoid], nx: int, ny: int, nz: int, dx: float, dy: float, dz: float,
                         cx: float = 0.0, cy: float = 0.0, cz: float = 0.0) -> np.ndarray:
    x = _centres(nx, dx, cx)[None, None, :]
    y = _centres(ny, dy, cy)[None, :, None]
    z = _centres(nz, dz, cz)[:, None, None]
    img = np.zeros((nz, ny, nx), np.float32)
    for e in ells:
        c, s = np.cos(e.phi), np.sin(e.phi)
        xr = (x - e.cx) * c + (y - e.cy) * s
        yr = -(x - e.cx) * s + (y - e.cy) * c
        img[(xr / e.a) ** 2 + (yr / e.b) ** 2 + ((z - e.cz) / e.c) ** 2 <= 1] = e.value
    return img


def ellipse_phantom(n: int, d: float, seed: int, nellipse: int = 10) -> np.ndarray:
    """[1, 1, n, n] float32 phantom of ``nellipse`` random ellipses (radius R = n*d/2)."""
    rng = np.random.default_rng(seed)
    ells = random_ellipses(rng, nellipse, n * d / 2)
    return rasterize_ellipses(ells, n, n, d, d)[None, None]


def ellipsoid_phantom(nz: int, ny: int, nx: int, dz: float, dy: float, dx: float, seed: int,
                      nellipsoid: int = 10) -> np.ndarray:
    """[nz, ny, nx] float32 phantom of random ellipsoids."""
    rng = np.random.default_rng(seed)
    ells = random_ellipsoids(rng, nellipsoid, min(nx * dx, ny * dy) / 2, nz * dz / 2)
    return rasterize_ellipsoids(ells, nx, ny, nz, dx, dy, dz)


def ellipse_line_integral(e: Ellipse, px: np.ndarray, py: np.ndarray, dirx: np.ndarray, diry: np.ndarray):
    """Exact integral of an ellipse's indicator (times its value) along lines
    p + t*dir (dir unit).  Used to pin projectors against closed forms."""
    c, s = np.cos(e.phi), np.sin(e.phi)
    ox = (px - e.cx) * c + (py - e.cy) * s
    oy = -(px - e.cx) * s + (py - e.cy) * c
    ux = dirx * c + diry * s
    uy = -dirx * s + diry * c
    A = (ux / e.a) ** 2 + (uy / e.b) ** 2
    B = 2 * (ox * ux / e.a ** 2 + oy * uy / e.b ** 2)
    C = (ox / e.a) ** 2 + (oy / e.b) ** 2 - 1
    disc = B * B - 4 * A * C
    chord = np.where(disc > 0, np.sqrt(np.maximum(disc, 0)) / A, 0.0)
    return chord * e.value


# --------------------------------------------------------------- geometries
@dataclass
class ConeGeometry:
    det_center: np.ndarray
    det_u: np.ndarray
    det_v: np.ndarray
    src: np.ndarray


def circular_cone_geometry(nview: int, dso: float, dsd: float, z0: float = 0.0, z1: float = 0.0,
                           jitter: float = 0.0, seed: int = 0, angle_range: float = 2 * np.pi) -> ConeGeometry:
    """Helical-like cone trajectory: source at (dso cos t, dso sin t, z_k), detector
    centre at (-(dsd-dso) cos t, -(dsd-dso) sin t, z_k), det_u = (-sin t, cos t, 0),
    det_v = (0, 0, 1); rows are (x, y, z).  ``jitter`` adds U(-j, j) to every
    component of source and detector centre."""
    t = np.arange(nview) * angle_range / nview
    zk = np.linspace(z0, z1, nview) if nview > 1 else np.array([z0])
    src = np.stack([dso * np.cos(t), dso * np.sin(t), zk], 1)
    dc = np.stack([-(dsd - dso) * np.cos(t), -(dsd - dso) * np.sin(t), zk], 1)
    du = np.stack([-np.sin(t), np.cos(t), np.zeros_like(t)], 1)
    dv = np.tile(np.array([[0.0, 0.0, 1.0]]), (nview, 1))
    if jitter > 0:
        rng = np.random.default_rng(seed)
        src = src + rng.uniform(-jitter, jitter, src.shape)
        dc = dc + rng.uniform(-jitter, jitter, dc.shape)
    f = np.float32
    return ConeGeometry(dc.astype(f), du.astype(f), dv.astype(f), src.astype(f))


@dataclass
class Workload:
    """Scan parameters in the attribute names of the reference ct_projector class."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    nu: int
    nv: int
    du: float
    dv: float
    nview: int
    dsd: float = 1085.6
    dso: float = 595.0
    cx: float = 0.0
    cy: float = 0.0
    cz: float = 0.0
    off_u: float = 0.0
    off_v: float = 0.0
    angles: np.ndarray = field(default=None, repr=False)


def parallel_256_workload() -> Tuple[Workload, np.ndarray]:
    """Parallel-beam DD, image 1x1x256x256 @1 mm, 360 views over 180 deg, 367 columns."""
    w = Workload(256, 256, 1, 1.0, 1.0, 1.0, 367, 1, 1.0, 1.0, 360)
    w.angles = (np.arange(360) * np.pi / 360).astype(np.float32)
    img = ellipse_phantom(256, 1.0, seed=1001)
    return w, img


def fan_512x32_workload() -> Tuple[Workload, np.ndarray]:
    """Fan-beam equiangular DD, volume 1x32x512x512 @0.7 mm (1 mm slices), 984 views
    over 360 deg, 888 columns of 1.2858 mm at dsd, 32 rows of 1 mm, dso 595, dsd 1085.6."""
    w = Workload(512, 512, 32, 0.7, 0.7, 1.0, 888, 32, 1.2858, 1.0, 984, dsd=1085.6, dso=595.0)
    w.angles = (np.arange(984) * 2 * np.pi / 984).astype(np.float32)
    img = ellipsoid_phantom(32, 512, 512, 1.0, 0.7, 0.7, seed=1002)[None]
    return w, img


def fan_slices_256_workload(batch: int = 8) -> Tuple[Workload, np.ndarray]:
    """Fan-beam DD, batch of 1x256x256 slices @1.4 mm, 492 views, 444 columns of 2.5716 mm."""
    w = Workload(256, 256, 1, 1.4, 1.4, 1.0, 444, 1, 2.5716, 1.0, 492, dsd=1085.6, dso=595.0)
    w.angles = (np.arange(492) * 2 * np.pi / 492).astype(np.float32)
    img = np.concatenate([ellipse_phantom(256, 1.4, seed=1040 + j) for j in range(batch)], 0)
    return w, img


CONE_SAMPLES = [  # (nz, ny, nx, dz, dy, dx)
    (128, 128, 128, 1.0, 1.0, 1.0),
    (192, 192, 192, 0.75, 0.75, 0.75),
    (256, 256, 256, 0.5, 0.5, 0.5),
    (128, 256, 256, 1.0, 0.75, 0.75),
]


def cone_workload(sample: int = 2, nview: int = 720) -> Tuple[Workload, np.ndarray, ConeGeometry]:
    """Cone-beam Siddon sample j of the batched workload: helical-like trajectory
    (z from -20 to +20 mm), +-2 mm jitter of source and detector centre seeded per
    sample (1030 + j), flat panel 512x384 @1 mm, ellipsoid phantom (seed 1130 + j)."""
    nz, ny, nx, dz, dy, dx = CONE_SAMPLES[sample]
    w = Workload(nx, ny, nz, dx, dy, dz, 512, 384, 1.0, 1.0, nview, dsd=1085.6, dso=595.0)
    geo = circular_cone_geometry(nview, 595.0, 1085.6, -20.0, 20.0, jitter=2.0, seed=1030 + sample)
    img = ellipsoid_phantom(nz, ny, nx, dz, dy, dx, seed=1130 + sample)[None]
    return w, img, geo


DD_BATCH_SAMPLES = [  # (beam, n, nz, nview, nu, pixel mm): uneven DD jobs of a mixed stream
    ('fan', 512, 1, 984, 888, 0.7),
    ('fan', 256, 1, 492, 444, 1.4),
    ('fan', 384, 4, 720, 666, 0.9),
    ('fan', 320, 2, 600, 555, 1.1),
    ('fan', 256, 8, 492, 444, 1.4),
    ('fan', 448, 1, 820, 777, 0.8),
    ('fan', 192, 1, 360, 333, 1.8),
    ('fan', 512, 3, 984, 888, 0.7),
    ('parallel', 256, 1, 360, 367, 1.0),
    ('parallel', 128, 1, 180, 183, 2.0),
    ('parallel', 320, 2, 400, 453, 0.8),
    ('parallel', 200, 1, 300, 283, 1.25),
    ('parallel', 384, 1, 512, 543, 0.67),
    ('parallel', 160, 4, 240, 227, 1.6),
    ('parallel', 96, 1, 128, 137, 2.6),
    ('parallel', 256, 2, 360, 367, 1.0),
]


def dd_batch_workload() -> List[Tuple[bool, Workload, np.ndarray]]:
    """The sixteen uneven distance-driven jobs of DD_BATCH_SAMPLES: (fan beam?,
    geometry, image [1, nz, n, n]).  Fan beams: dso 595, dsd 1085.6, the
    fan's angular span of the 888-column detector at 1.2858 mm (du = 1.2858 x
    888 / nu), views over 360 deg; parallel beams: du = 1.4336 x n x pixel / nu,
    views over 180 deg; slices of 1 mm; ellipsoid phantoms seeded 1200 + j."""
    out = []
    for j, (beam, n, nz, nview, nu, pix) in enumerate(DD_BATCH_SAMPLES):
        fan = beam == 'fan'
        du = 1.2858 * 888 / nu if fan else 1.4336 * n * pix / nu
        w = Workload(n, n, nz, pix, pix, 1.0, nu, nz, du, 1.0, nview, dsd=1085.6, dso=595.0)
        span = 2 * np.pi if fan else np.pi
        w.angles = (np.arange(nview) * span / nview).astype(np.float32)
        img = ellipsoid_phantom(nz, n, n, 1.0, pix, pix, seed=1200 + j)[None]
        out.append((fan, w, img))
    return out

