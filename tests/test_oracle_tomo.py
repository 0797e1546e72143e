"""Pins the FP64 oracle's distance-driven tomosynthesis projector (oracle/oracle.cpp
following cuda/distanceDriven3DBoxInt.cu, cuda/distanceDriven3DBranchless.cu and
cuda/distanceDriven3DUtils.cu) against computations that do not reuse its code:

* brute force in physical coordinates: each detector cell's footprint in every
  slice plane is found by intersecting the source->cell-edge lines with the
  plane, and the image is averaged over it from exact pixel/rectangle overlaps;
* the box-integral mode (IntegralBoxXY) and the branchless mode (summed-area
  tables) are two independent algorithms and agree;
* closed forms: a constant slab projects to value x thickness x secant, and a
  projection of c / (pre-weight) backprojects to nview x c;
* a brute-force backprojection with the reference's conventions (voxel centre
  projected with ProjectConeToDetCart, boxes over detector cells [i, i+1)).

The reference samples slice iz in the plane through its lower face (the FP
loop uses the corner-indexed image coordinate iz,
cuda/distanceDriven3DBoxInt.cu:96) while the BP projects voxel centres
(cuda/distanceDriven3DBoxInt.cu:215-217); both conventions are kept.
"""

import numpy as np
import pytest

from helpers import Geo, rel_l2


def f32(x):
    return float(np.float32(x))


def tomo_case(seed=0, nview=3, nu=36, nv=34):
    rng = np.random.default_rng(seed)
    g = Geo(nx=20, ny=18, nz=6, dx=1.1, dy=0.9, dz=1.3, cx=0.4, cy=-0.3, cz=20.0, nu=nu, nv=nv, du=1.0, dv=0.95,
            off_u=0.3, off_v=-0.2)
    t = np.linspace(-0.35, 0.35, nview)
    src = np.stack([300 * np.sin(t), 6 + 0 * t, 300 * np.cos(t) + 20], 1).astype(np.float32)
    dc = np.stack([rng.uniform(-1, 1, nview), rng.uniform(-1, 1, nview), -10 + 0 * t], 1).astype(np.float32)
    img = rng.uniform(size=[1, g.nz, g.ny, g.nx]).astype(np.float32)
    return g, dc, src, img


def _overlap(a1, a2, edges):
    """Overlap lengths of intervals [a1, a2] (shape [n]) with cells between
    consecutive edges (shape [m+1]) -> [n, m]."""
    lo = np.maximum(a1[:, None], edges[None, :-1])
    hi = np.minimum(a2[:, None], edges[None, 1:])
    return np.maximum(hi - lo, 0.0)


def brute_fp(g, img, dc, src):
    nu, nv, nx, ny, nz = g.nu, g.nv, g.nx, g.ny, g.nz
    du, dv, dx, dy, dz = (f32(getattr(g, k)) for k in ('du', 'dv', 'dx', 'dy', 'dz'))
    xe = (np.arange(nx + 1) - nx / 2) * dx + f32(g.cx)
    ye = (np.arange(ny + 1) - ny / 2) * dy + f32(g.cy)
    out = np.zeros([len(src), nv, nu])
    for k in range(len(src)):
        S = src[k].astype(np.float64)
        C = dc[k].astype(np.float64)
        u = (np.arange(nu) - f32(g.off_u) - (nu - 1) / 2) * du
        v = (np.arange(nv) - f32(g.off_v) - (nv - 1) / 2) * dv
        for iz in range(nz):
            z = (iz - nz / 2) * dz + f32(g.cz)          # lower face of slice iz
            r = (z - S[2]) / (C[2] - S[2])              # detector plane z = C.z
            xa = S[0] + r * (C[0] + u - du / 2 - S[0])
            xb = S[0] + r * (C[0] + u + du / 2 - S[0])
            ya = S[1] + r * (C[1] + v - dv / 2 - S[1])
            yb = S[1] + r * (C[1] + v + dv / 2 - S[1])
            x1, x2 = np.minimum(xa, xb), np.maximum(xa, xb)
            y1, y2 = np.minimum(ya, yb), np.maximum(ya, yb)
            ox = _overlap(x1, x2, xe)                   # [nu, nx]
            oy = _overlap(y1, y2, ye)                   # [nv, ny]
            s = oy @ img[0, iz].astype(np.float64) @ ox.T
            out[k] += s / np.outer(y2 - y1, x2 - x1)
        Dx = C[0] + u[None, :]
        Dy = C[1] + v[:, None]
        L = np.sqrt((S[0] - Dx) ** 2 + (S[1] - Dy) ** 2 + (S[2] - C[2]) ** 2)
        out[k] *= dz * L / abs(S[2] - C[2])
    return out[None]


def preweight(g, dc, src):
    du, dv, dz = f32(g.du), f32(g.dv), f32(g.dz)
    w = np.zeros([len(src), g.nv, g.nu])
    for k in range(len(src)):
        S = src[k].astype(np.float64)
        C = dc[k].astype(np.float64)
        u = (np.arange(g.nu) - f32(g.off_u) - (g.nu - 1) / 2) * du
        v = (np.arange(g.nv) - f32(g.off_v) - (g.nv - 1) / 2) * dv
        L = np.sqrt((S[0] - C[0] - u[None, :]) ** 2 + (S[1] - C[1] - v[:, None]) ** 2 + (S[2] - C[2]) ** 2)
        w[k] = dz * L / abs(S[2] - C[2])
    return w


def brute_bp(g, prj, dc, src):
    nu, nv, nx, ny, nz = g.nu, g.nv, g.nx, g.ny, g.nz
    du, dv, dx, dy, dz = (f32(getattr(g, k)) for k in ('du', 'dv', 'dx', 'dy', 'dz'))
    wp = prj[0].astype(np.float64) * preweight(g, dc, src)
    x = (np.arange(nx) - (nx - 1) / 2) * dx + f32(g.cx)
    y = (np.arange(ny) - (ny - 1) / 2) * dy + f32(g.cy)
    ue = np.arange(nu + 1.0)
    ve = np.arange(nv + 1.0)
    img = np.zeros([nz, ny, nx])
    for iz in range(nz):
        z = (iz - (nz - 1) / 2) * dz + f32(g.cz)        # voxel centre
        for k in range(len(src)):
            S = src[k].astype(np.float64)
            C = dc[k].astype(np.float64)
            r = (C[2] - S[2]) / (z - S[2])
            ua = (S[0] - C[0] + r * (x - dx / 2 - S[0])) / du + f32(g.off_u) + (nu - 1) / 2
            ub = (S[0] - C[0] + r * (x + dx / 2 - S[0])) / du + f32(g.off_u) + (nu - 1) / 2
            va = (S[1] - C[1] + r * (y - dy / 2 - S[1])) / dv + f32(g.off_v) + (nv - 1) / 2
            vb = (S[1] - C[1] + r * (y + dy / 2 - S[1])) / dv + f32(g.off_v) + (nv - 1) / 2
            ou = _overlap(ua, ub, ue)                    # [nx, nu]
            ov = _overlap(va, vb, ve)                    # [ny, nv]
            img[iz] += (ov @ wp[k] @ ou.T) / np.outer(vb - va, ub - ua)
    return img[None]


@pytest.mark.parametrize('branchless', [False, True])
def test_fp_brute_force(oracle, branchless):
    g, dc, src, img = tomo_case(1)
    fp = oracle.dd_tomo_fp(g, img, dc, src, branchless=branchless)
    ref = brute_fp(g, img, dc, src)
    assert ref.max() > 1
    assert rel_l2(fp, ref) < 1e-12


def test_fp_brute_force_detects_centre_plane(oracle):
    """A plausible mistake (sampling slices at their centres) is caught."""
    g, dc, src, img = tomo_case(1)
    fp = oracle.dd_tomo_fp(g, img, dc, src)
    g2 = Geo(**{**g.__dict__, 'cz': g.cz + g.dz / 2})
    assert rel_l2(fp, brute_fp(g2, img, dc, src)) > 1e-3


@pytest.mark.parametrize('branchless', [False, True])
def test_bp_brute_force(oracle, branchless):
    g, dc, src, _ = tomo_case(2)
    y = np.random.default_rng(5).uniform(size=[1, 3, g.nv, g.nu]).astype(np.float32)
    bp = oracle.dd_tomo_bp(g, y, dc, src, branchless=branchless)
    ref = brute_bp(g, y, dc, src)
    assert rel_l2(bp, ref) < 1e-12


def test_boxint_equals_branchless(oracle):
    g, dc, src, img = tomo_case(3, nview=5)
    a = oracle.dd_tomo_fp(g, img, dc, src, branchless=False)
    b = oracle.dd_tomo_fp(g, img, dc, src, branchless=True)
    assert rel_l2(a, b) < 1e-12
    y = np.random.default_rng(1).uniform(size=a.shape).astype(np.float32)
    assert rel_l2(oracle.dd_tomo_bp(g, y, dc, src, branchless=False),
                  oracle.dd_tomo_bp(g, y, dc, src, branchless=True)) < 1e-12


def _cells_inside(g, dc, src):
    """[nview, nv, nu] mask of cells whose footprint lies inside the grid's
    x/y extent in every slice plane (geometry only)."""
    du, dv, dz = f32(g.du), f32(g.dv), f32(g.dz)
    xl, xh = f32(g.cx) - g.nx * f32(g.dx) / 2, f32(g.cx) + g.nx * f32(g.dx) / 2
    yl, yh = f32(g.cy) - g.ny * f32(g.dy) / 2, f32(g.cy) + g.ny * f32(g.dy) / 2
    m = np.ones([len(src), g.nv, g.nu], bool)
    u = (np.arange(g.nu) - f32(g.off_u) - (g.nu - 1) / 2) * du
    v = (np.arange(g.nv) - f32(g.off_v) - (g.nv - 1) / 2) * dv
    for k in range(len(src)):
        S, C = src[k].astype(np.float64), dc[k].astype(np.float64)
        for iz in range(g.nz + 1):
            r = ((iz - g.nz / 2) * dz + f32(g.cz) - S[2]) / (C[2] - S[2])
            xa, xb = S[0] + r * (C[0] + u - du / 2 - S[0]), S[0] + r * (C[0] + u + du / 2 - S[0])
            ya, yb = S[1] + r * (C[1] + v - dv / 2 - S[1]), S[1] + r * (C[1] + v + dv / 2 - S[1])
            m[k] &= ((np.minimum(ya, yb) >= yl) & (np.maximum(ya, yb) <= yh))[:, None]
            m[k] &= ((np.minimum(xa, xb) >= xl) & (np.maximum(xa, xb) <= xh))[None, :]
    return m


def test_constant_slab_closed_form(oracle):
    """Cells whose footprints stay inside the grid in every slice see
    value x nz dz x |S - D| / |S_z - D_z|; cells that miss the grid see 0."""
    g, dc, src, _ = tomo_case(4, nu=64, nv=48)
    img = np.full([1, g.nz, g.ny, g.nx], 0.25, np.float32)
    fp = oracle.dd_tomo_fp(g, img, dc, src)[0]
    ref = 0.25 * g.nz * preweight(g, dc, src)
    m = _cells_inside(g, dc, src)
    assert m.sum() > 300
    assert np.abs(fp - ref)[m].max() < 1e-12
    assert (fp[:, :, 0] == 0).all() and (fp[:, :, -1] == 0).all()


@pytest.mark.parametrize('branchless', [False, True])
def test_weighted_constant_backprojects_to_nview(oracle, branchless):
    """With a detector that catches every voxel's shadow, a projection equal to
    c / pre-weight backprojects to nview x c everywhere."""
    g, dc, src, _ = tomo_case(5, nview=4, nu=64, nv=48)
    y = (0.5 / preweight(g, dc, src))[None].astype(np.float32)
    bp = oracle.dd_tomo_bp(g, y, dc, src, branchless=branchless)[0]
    assert np.abs(bp - 4 * 0.5).max() < 1e-6


def test_fp32_mirror_close_to_fp64(oracle):
    g, dc, src, img = tomo_case(6)
    for br in (False, True):
        assert rel_l2(oracle.dd_tomo_fp(g, img, dc, src, branchless=br, prec='f'),
                      oracle.dd_tomo_fp(g, img, dc, src, branchless=br)) < 1e-5
