"""Pins the FP64 oracle's distance-driven fan/parallel projector (oracle/oracle.cpp
following cuda/distanceDriven2DBoxInt.cu, cuda/distanceDriven2DBranchless.cu,
cuda/distanceDriven2DUtils.cu, cuda/cudaMath.cu:533-602) against facts that do
not come from the oracle's own formulas:

* the known-answer case of the reference notebook ex_dd_fbp_forbild.ipynb
  (cells 2-3): a 2x2 block of ones, dx = du = 0.25, angle 0 -> 0.5 at iu = 7, 8;
* constant images, whose projections are path lengths;
* two independent algorithms agree: the box-integral path (IntegralBoxX) and
  the branchless path (interpolated prefix sums), outside the +-0.12 blend band;
* the exact relation BP = A^T D of the parallel projector (D per view);
* analytic line integrals of ellipses (distance driven ~ line integral);
* rotation covariance (angle convention and direction of rotation);
* fan beam -> parallel beam as the source distance grows;
* z handling: stacked identical slices, the truncated row of the BoxInt BP.
"""

import numpy as np
import pytest

from ctprojector_amd.phantom import Ellipse, ellipse_line_integral, rasterize_ellipses
from helpers import Geo, rel_l2

MARGIN = float(np.float32(0.12))


def par_geo(n=32, nu=48, du=1.0, dx=1.0, nz=1, nv=1, **kw):
    return Geo(nx=n, ny=n, nz=nz, dx=dx, dy=dx, dz=1.0, nu=nu, nv=nv, du=du, dv=1.0, **kw)


def fan_geo(n=32, nu=64, du=1.5, dx=1.0, nz=1, nv=1, dsd=200.0, dso=120.0, **kw):
    return Geo(nx=n, ny=n, nz=nz, dx=dx, dy=dx, dz=1.0, nu=nu, nv=nv, du=du, dv=1.0, dsd=dsd, dso=dso, **kw)


def test_notebook_known_answer(oracle):
    g = par_geo(n=16, nu=16, du=0.25, dx=0.25)
    img = np.zeros([1, 1, 16, 16], np.float32)
    img[0, 0, 7:9, 7:9] = 1
    for branchless in (False, True):
        fp = oracle.dd_par_fp(g, img, np.zeros(1, np.float32), branchless=branchless)[0, 0, 0]
        ref = np.zeros(16)
        ref[7:9] = 0.5
        assert np.abs(fp - ref).max() < 1e-12


def test_constant_image_path_length(oracle):
    # parallel rays along y (angle 0), du = dx: every fully covered cell sees ny*dy
    g = par_geo(n=20, nu=30, du=1.0, dx=1.0)
    img = np.full([1, 1, 20, 20], 0.5, np.float32)
    fp = oracle.dd_par_fp(g, img, np.zeros(1, np.float32))[0, 0, 0]
    inner = fp[6:24]
    assert np.abs(inner - 0.5 * 20).max() < 1e-12
    assert fp[0] == 0 and fp[-1] == 0


def _outside_band(angles):
    c, s = np.abs(np.cos(angles.astype(np.float64))), np.abs(np.sin(angles.astype(np.float64)))
    return np.abs(c - s) > MARGIN + 1e-6


@pytest.mark.parametrize('fan', [False, True])
def test_boxint_equals_branchless_outside_band(oracle, fan):
    rng = np.random.default_rng(3)
    g = fan_geo() if fan else par_geo()
    img = rng.uniform(size=[1, 1, 32, 32]).astype(np.float32)
    ang = rng.uniform(0, 2 * np.pi, 40).astype(np.float32)
    ang = ang[_outside_band(ang)]
    f = oracle.dd_fan_fp if fan else oracle.dd_par_fp
    a = f(g, img, ang, branchless=False)
    b = f(g, img, ang, branchless=True)
    assert rel_l2(a, b) < 1e-12
    bpf = oracle.dd_fan_bp if fan else oracle.dd_par_bp
    y = rng.uniform(size=a.shape).astype(np.float32)
    for fbp in (False, True):
        assert rel_l2(bpf(g, y, ang, is_fbp=fbp, branchless=False), bpf(g, y, ang, is_fbp=fbp, branchless=True)) < 1e-12


def test_parallel_bp_is_transpose_times_view_weight(oracle):
    """<A x, D y> = <x, BP y> with D = du / (dx |cos|) (y-main) or du / (dy |sin|)
    (x-main) per view, for views outside the blend band (exact for parallel beams)."""
    rng = np.random.default_rng(7)
    g = par_geo(n=24, nu=40, du=0.8, dx=1.1)
    ang = rng.uniform(0, np.pi, 30).astype(np.float32)
    ang = ang[_outside_band(ang)]
    x = rng.uniform(size=[1, 1, 24, 24]).astype(np.float32)
    y = rng.uniform(size=[1, len(ang), 1, 40]).astype(np.float32)
    ax = oracle.dd_par_fp(g, x, ang)
    bpy = oracle.dd_par_bp(g, y, ang)
    c, s = np.abs(np.cos(ang.astype(np.float64))), np.abs(np.sin(ang.astype(np.float64)))
    du, dx = float(np.float32(0.8)), float(np.float32(1.1))
    D = np.where(c > s, du / (dx * c), du / (dx * s))
    lhs = float((ax * (D[None, :, None, None] * y)).sum())
    rhs = float((x * bpy).sum())
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


@pytest.mark.parametrize('fan', [False, True])
def test_line_integrals_of_ellipses(oracle, fan):
    """The distance-driven FP of a finely sampled ellipse approaches its analytic line integrals."""
    n, d = 128, 0.5
    ells = [Ellipse(3.0, -4.0, 14.0, 8.0, 0.6, 1.0), Ellipse(-9.0, 6.0, 6.0, 10.0, 2.2, 0.5)]
    img = rasterize_ellipses(ells, n, n, d, d)[None, None]
    ang = np.linspace(0, np.pi, 12, endpoint=False).astype(np.float32)
    if fan:
        g = fan_geo(n=n, dx=d, nu=120, du=1.0, dsd=400.0, dso=250.0)
        fp = oracle.dd_fan_fp(g, img, ang)[0, :, 0]
        da = np.float32(g.du / g.dsd)
        a = (-(np.arange(g.nu) - (g.nu - 1) / 2.0)) * float(da)
        ref = np.zeros(fp.shape)
        for k, t in enumerate(ang.astype(np.float64)):
            sx, sy = g.dso * np.sin(t), -g.dso * np.cos(t)
            # ray direction towards the detector cell centre, rotated by the view angle
            ex, ey = g.dsd * np.sin(a), -g.dso + g.dsd * np.cos(a)
            dxr, dyr = ex * np.cos(t) - ey * np.sin(t) - sx, ex * np.sin(t) + ey * np.cos(t) - sy
            nrm = np.hypot(dxr, dyr)
            ref[k] = sum(ellipse_line_integral(e, sx, sy, dxr / nrm, dyr / nrm) for e in ells)
    else:
        g = par_geo(n=n, dx=d, nu=100, du=0.7)
        fp = oracle.dd_par_fp(g, img, ang)[0, :, 0]
        u = (-(np.arange(g.nu) - (g.nu - 1) / 2.0)) * float(np.float32(g.du))
        ref = np.zeros(fp.shape)
        for k, t in enumerate(ang.astype(np.float64)):
            # point on the ray: (u, 0) rotated; direction (0, 1) rotated
            px, py = u * np.cos(t), u * np.sin(t)
            dxr, dyr = -np.sin(t), np.cos(t)
            ref[k] = sum(ellipse_line_integral(e, px, py, np.full_like(px, dxr), np.full_like(px, dyr)) for e in ells)
    assert ref.max() > 10
    assert rel_l2(fp, ref) < 0.02


@pytest.mark.parametrize('fan', [False, True])
def test_rotation_covariance(oracle, fan):
    """The source turns counter-clockwise with the angle (at 0 it sits at (0, -dso),
    cuda/distanceDriven2DBoxInt.cu:217): projecting the image turned by -90 degrees
    at t - 90 degrees equals projecting the image at t."""
    rng = np.random.default_rng(11)
    g = fan_geo() if fan else par_geo()
    img = rng.uniform(size=[1, 1, 32, 32]).astype(np.float32)
    rot = np.ascontiguousarray(np.rot90(img, k=1, axes=(2, 3)))   # centred (X, Y) -> (Y, -X): -90 degrees
    t = np.array([1.9, 2.6, 4.0], np.float32)
    f = oracle.dd_fan_fp if fan else oracle.dd_par_fp
    half_pi = np.float32(np.pi / 2)
    a = f(g, rot, (t - half_pi).astype(np.float32))
    b = f(g, img, t)
    assert rel_l2(a, b) < 1e-5
    wrong = f(g, rot, (t + half_pi).astype(np.float32))
    assert rel_l2(wrong, b) > 1e-2   # the opposite sense of rotation is detected


def test_fan_tends_to_parallel(oracle):
    rng = np.random.default_rng(2)
    img = rng.uniform(size=[1, 1, 32, 32]).astype(np.float32)
    ang = np.array([0.2, 0.9, 2.0], np.float32)
    gp = par_geo(nu=48, du=1.0)
    far = 1.0e5
    gf = fan_geo(nu=48, du=1.0 * (far + 20) / far, dsd=far + 20, dso=far)   # cell width 1 mm at the centre
    pf = oracle.dd_par_fp(gp, img, ang)
    ff = oracle.dd_fan_fp(gf, img, ang)
    assert rel_l2(ff, pf) < 1e-3


def test_stacked_identical_slices(oracle):
    """With identical slices every detector row equals the single-slice projection."""
    rng = np.random.default_rng(4)
    s = rng.uniform(size=[1, 1, 24, 24]).astype(np.float32)
    img = np.repeat(s, 5, axis=1)
    ang = np.array([0.1, 0.8, 1.7], np.float32)
    g1 = fan_geo(n=24, nu=40)
    g5 = fan_geo(n=24, nu=40, nz=5, nv=7)
    a = oracle.dd_fan_fp(g1, s, ang)
    b = oracle.dd_fan_fp(g5, img, ang)
    for iv in range(7):
        assert rel_l2(b[:, :, iv], a[:, :, 0]) < 1e-12


def test_boxint_bp_row_truncation(oracle):
    """BoxInt BP: slice z reads detector row trunc(z/dv + off_v + nv/2 - 0.5)
    (cuda/distanceDriven2DBoxInt.cu:350,383); rows outside [0, nv) read 0."""
    g = fan_geo(n=16, nu=32, nz=4, nv=4)
    g.dz = 1.0
    g.dv = 1.6                     # slice centres z = -1.5, -0.5, 0.5, 1.5 -> rows 0.56, 1.19, 1.81, 2.44
    ang = np.array([0.4], np.float32)
    for r in range(4):
        y = np.zeros([1, 1, 4, 32], np.float32)
        y[0, 0, r] = 1.0
        bp = oracle.dd_fan_bp(g, y, ang)[0]
        lit = [z for z in range(4) if np.abs(bp[z]).max() > 0]
        expect = [z for z in range(4) if int((z - 1.5) / float(np.float32(1.6)) + 2 - 0.5) == r]
        assert lit == expect


def test_fp32_mirror_differs_only_at_45_degree_switch(oracle):
    """The reference decides the 45-degree main-axis switch of the BP in FP32; the
    FP64 oracle in FP64.  Away from exact 45-degree views the two agree to rounding."""
    rng = np.random.default_rng(9)
    g = fan_geo(nz=2, nv=2)
    ang = rng.uniform(0, 2 * np.pi, 25).astype(np.float32)
    y = rng.uniform(size=[1, 25, 2, 64]).astype(np.float32)
    for fbp in (False, True):
        assert rel_l2(oracle.dd_fan_bp(g, y, ang, is_fbp=fbp, prec='f'), oracle.dd_fan_bp(g, y, ang, is_fbp=fbp)) < 1e-5


@pytest.mark.parametrize('beam,typ', [('fan', 0), ('parallel', 0), ('parallel', 4)])
def test_fp_adjoint_is_exact_transpose(oracle, beam, typ):
    """The oracle's A^T (dd_{fan,par}_fp_adjoint, the forward loops transposed)
    satisfies <A x, y> = <x, A^T y> against the pinned forward projector, with
    views inside the +-0.12 blend band, several slices (z interpolation),
    offsets and two batch items; and equals A^T built column by column."""
    rng = np.random.default_rng(21)
    if beam == 'fan':
        g = fan_geo(n=20, nu=40, nz=3, nv=4, off_u=0.3, off_v=0.2, cx=1.0, cy=-0.5)
        fp, adj = oracle.dd_fan_fp, oracle.dd_fan_fp_adjoint
    else:
        g = par_geo(n=20, nu=36, nz=3, nv=4, du=0.8, off_u=-0.4, off_v=0.3, cx=0.5)
        fp, adj = oracle.dd_par_fp, oracle.dd_par_fp_adjoint
    g.dz, g.dv = 1.0, 0.9
    ang = np.concatenate([rng.uniform(0, 2 * np.pi, 12), [np.pi / 4 + 0.05, 3 * np.pi / 4 - 0.02]]).astype(np.float32)
    x = rng.uniform(size=[2, 3, 20, 20]).astype(np.float32)
    y = rng.uniform(size=[2, len(ang), 4, g.nu]).astype(np.float32)
    ax = fp(g, x, ang, type_projector=typ)
    aty = adj(g, y, ang, type_projector=typ)
    lhs = float((ax * y).sum())
    rhs = float((x * aty).sum())
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # column by column on a small case
    gs = fan_geo(n=6, nu=10, nz=1, nv=1) if beam == 'fan' else par_geo(n=6, nu=10)
    angs = ang[:5]
    eye = np.eye(36, dtype=np.float32).reshape(36, 1, 6, 6)
    A = fp(gs, eye, angs, type_projector=typ).reshape(36, -1).T
    ys = rng.uniform(size=[1, 5, 1, 10]).astype(np.float32)
    assert np.abs(adj(gs, ys, angs, type_projector=typ).ravel() - A.T @ ys.ravel().astype(np.float64)).max() < 1e-12
