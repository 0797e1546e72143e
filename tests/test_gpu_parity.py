"""GPU parity: every hot-path entry point of libprojector.so against the FP64
oracle on small seeded cases (rel-L2 tolerance 1e-4, FP32 arithmetic), plus the
adjoint dot-product test of the Siddon pair.  Calls go through the C ABI (the
host-array ``c*`` entry points via the numpy front end)."""

import numpy as np
import pytest

from ctprojector_amd.projector.numpy import cone, fan_equiangluar as fan, parallel, tomo
from ctprojector_amd.projector.numpy.ct_projector import ct_projector

from helpers import rel_l2, small_cone, small_fan, small_parallel, Geo

pytestmark = pytest.mark.gpu
TOL = 1e-4


def _proj(g):
    return g.to(ct_projector())


# ----------------------------------------------------------------- Siddon cone
@pytest.mark.parametrize('case', ['plain', 'offsets'])
def test_siddon_cone_fp_bp(oracle, case):
    if case == 'plain':
        g, geo, img = small_cone()
    else:
        g, geo, img = small_cone(seed=11, off=(0.3, -0.7), c=(1.5, -2.0, 0.5))
    p = _proj(g)
    args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
    fp = cone.siddon_fp_arbitrary(p, img, *args)
    ref = oracle.siddon_cone_fp(g, img, *args)
    e = rel_l2(fp, ref)
    print('siddon fp rel-L2', e)
    assert e < TOL
    y = np.random.default_rng(1).uniform(size=ref.shape).astype(np.float32)
    bp = cone.siddon_bp_arbitrary(p, y, *args)
    refb = oracle.siddon_cone_bp(g, y, *args)
    e = rel_l2(bp, refb)
    print('siddon bp rel-L2', e)
    assert e < TOL


def test_siddon_cone_adjoint():
    g, geo, img = small_cone(seed=5)
    p = _proj(g)
    args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
    rng = np.random.default_rng(2)
    x = rng.uniform(size=img.shape).astype(np.float32)
    fpx = cone.siddon_fp_arbitrary(p, x, *args)
    y = rng.uniform(size=fpx.shape).astype(np.float32)
    bty = cone.siddon_bp_arbitrary(p, y, *args)
    lhs = np.dot(fpx.ravel().astype(np.float64), y.ravel().astype(np.float64))
    rhs = np.dot(x.ravel().astype(np.float64), bty.ravel().astype(np.float64))
    rel = abs(lhs - rhs) / (np.linalg.norm(fpx) * np.linalg.norm(y))
    print('adjoint', lhs, rhs, rel)
    assert rel < 1e-6


@pytest.mark.parametrize('nb', [2, 3, 5])
def test_siddon_cone_batch(nb):
    """Batch items sharing the geometry (projected in groups of up to 4 volumes
    by the column-driven FP) equal single-volume calls."""
    g, geo, img = small_cone(seed=9, n=16, nview=6, nu=24, nv=20)
    rng = np.random.default_rng(nb)
    imgs = np.concatenate([img] + [rng.uniform(size=img.shape).astype(np.float32) for _ in range(nb - 1)], 0)
    p = _proj(g)
    args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
    fp = cone.siddon_fp_arbitrary(p, imgs, *args)
    for b in range(nb):
        assert rel_l2(fp[b:b + 1], cone.siddon_fp_arbitrary(p, imgs[b:b + 1], *args)) < 1e-6
    y = rng.uniform(size=fp.shape).astype(np.float32)
    bp = cone.siddon_bp_arbitrary(p, y, *args)
    for b in range(nb):
        assert rel_l2(bp[b:b + 1], cone.siddon_bp_arbitrary(p, y[b:b + 1], *args)) < 1e-6


def test_siddon_cone_batch_of_four_tall_detector(oracle):
    """Four volumes sharing the geometry on a 300-row detector: the four-volume
    column FP follows up to five rays per lane (300 reaching rows in one
    traversal); equal to single-volume calls and to the oracle."""
    from helpers import Geo
    from ctprojector_amd.phantom import circular_cone_geometry, ellipsoid_phantom
    g = Geo(nx=24, ny=24, nz=40, dx=1.0, dy=1.0, dz=1.0, nu=32, nv=300, du=1.6, dv=0.25, dsd=300.0, dso=180.0)
    geo = circular_cone_geometry(5, 180.0, 300.0, -1.0, 1.0, jitter=0.5, seed=17)
    rng = np.random.default_rng(17)
    base = ellipsoid_phantom(40, 24, 24, 1.0, 1.0, 1.0, seed=17)[None]
    imgs = np.concatenate([base] + [rng.uniform(size=base.shape).astype(np.float32) for _ in range(3)], 0)
    p = _proj(g)
    args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
    print('cone batch of four, 300-row detector', flush=True)
    fp = cone.siddon_fp_arbitrary(p, imgs, *args)
    for b in range(4):
        assert rel_l2(fp[b:b + 1], cone.siddon_fp_arbitrary(p, imgs[b:b + 1], *args)) < 1e-6
    assert rel_l2(fp, oracle.siddon_cone_fp(g, imgs, *args)) < 1e-4


@pytest.fixture
def tuning():
    """Set library tuning keys for one test; restores the defaults after."""
    from ctprojector_amd import _lib
    keys = []

    def set_(k, v):
        keys.append(k)
        _lib.set_tuning(k, v)
    yield set_
    for k in keys:
        _lib.set_tuning(k, 0)


def _siddon_check(oracle, g, geo, img, tol=TOL):
    p = _proj(g)
    args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
    fp = cone.siddon_fp_arbitrary(p, img, *args)
    assert rel_l2(fp, oracle.siddon_cone_fp(g, img, *args)) < tol
    y = np.random.default_rng(4).uniform(size=fp.shape).astype(np.float32)
    bp = cone.siddon_bp_arbitrary(p, y, *args)
    assert rel_l2(bp, oracle.siddon_cone_bp(g, y, *args)) < tol
    lhs = np.dot(fp.ravel().astype(np.float64), y.ravel().astype(np.float64))
    rhs = np.dot(img.ravel().astype(np.float64), bp.ravel().astype(np.float64))
    assert abs(lhs - rhs) / (np.linalg.norm(fp) * np.linalg.norm(y)) < 1e-6
    return bp


def test_siddon_bp_column_and_general_modes(oracle, tuning):
    """The column-geometry backprojector (det_v along z, det_u horizontal) and
    the general one agree with the oracle and with each other."""
    g, geo, img = small_cone(seed=13)
    a = _siddon_check(oracle, g, geo, img)
    tuning('cone_bp_mode', 1)
    b = _siddon_check(oracle, g, geo, img)
    assert rel_l2(a, b) < 1e-5


def test_siddon_tilted_detector(oracle):
    """A detector rotated about its normal and tilted: the general path."""
    from ctprojector_amd.phantom import ConeGeometry
    g, geo, img = small_cone(seed=17)
    rng = np.random.default_rng(3)
    du = geo.det_u.astype(np.float64) + rng.uniform(-0.2, 0.2, geo.det_u.shape) * [0, 0, 1]
    du /= np.linalg.norm(du, axis=1, keepdims=True)
    w = (geo.det_center - geo.src).astype(np.float64)
    dv = np.cross(w, du)
    dv /= np.linalg.norm(dv, axis=1, keepdims=True)
    geo2 = ConeGeometry(geo.det_center, du.astype(np.float32), dv.astype(np.float32), geo.src)
    _siddon_check(oracle, g, geo2, img)


def test_siddon_near_axis_parallel_rows(oracle):
    """Rows whose rays have a z extent in (0, 1e-6] voxels are axis-parallel
    for the reference (cuda/siddon.cu:15) but not exactly horizontal: the
    column-mode backprojector hands them to its exact ray-driven path."""
    from ctprojector_amd.phantom import ConeGeometry
    g = Geo(nx=16, ny=16, nz=16, dx=1.0, dy=1.0, dz=1.0, nu=30, nv=31, du=1.5, dv=1.5, dsd=300.0, dso=180.0)
    nview = 6
    t = np.arange(nview) * 2 * np.pi / nview
    src = np.stack([180 * np.cos(t), 180 * np.sin(t), np.zeros_like(t)], 1)
    dc = np.stack([-120 * np.cos(t), -120 * np.sin(t), np.full_like(t, 1e-7)], 1)   # centre row: w_z = 1e-7 mm
    du = np.stack([-np.sin(t), np.cos(t), np.zeros_like(t)], 1)
    dv = np.tile([[0.0, 0.0, 1.0]], (nview, 1))
    geo = ConeGeometry(*(a.astype(np.float32) for a in (dc, du, dv, src)))
    assert geo.det_center[0, 2] != 0
    img = np.random.default_rng(8).uniform(size=[1, 16, 16, 16]).astype(np.float32)
    _siddon_check(oracle, g, geo, img)


@pytest.mark.parametrize('case', ['default', 'steep', 'axis'])
def test_siddon_fp_column_and_ray_modes(oracle, tuning, case):
    """The column-driven forward projector (rays of a detector column share
    their xy line) and the ray-per-lane one agree with the oracle and with each
    other.  'steep': slices thin enough that the outer rows exceed the column
    form's slope bound and are traced ray by ray; 'axis': views at multiples
    of 90 degrees with an odd detector width, so one column per view runs
    exactly along x or y (traced ray by ray)."""
    from ctprojector_amd.phantom import circular_cone_geometry
    if case == 'default':
        g, geo, img = small_cone(seed=21)
    else:
        nz, dz = (60, 0.2) if case == 'steep' else (12, 1.0)
        g = Geo(nx=14, ny=12, nz=nz, dx=1.5, dy=1.6, dz=dz, cx=0.3, cy=-0.2, nu=25, nv=40, du=1.3, dv=0.9,
                dsd=110.0, dso=60.0)
        if case == 'axis':
            geo = circular_cone_geometry(4, 60.0, 110.0, angle_range=2 * np.pi)
        else:
            geo = circular_cone_geometry(5, 60.0, 110.0, -1.0, 1.0, jitter=0.5, seed=4)
        img = np.random.default_rng(22).uniform(size=[1, g.nz, g.ny, g.nx]).astype(np.float32)
    p = _proj(g)
    args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
    ref = oracle.siddon_cone_fp(g, img, *args)
    a = cone.siddon_fp_arbitrary(p, img, *args)
    tuning('cone_fp_mode', 1)
    b = cone.siddon_fp_arbitrary(p, img, *args)
    print(case, 'column', rel_l2(a, ref), 'ray', rel_l2(b, ref))
    assert rel_l2(a, ref) < TOL
    assert rel_l2(b, ref) < TOL
    assert rel_l2(a, b) < TOL


# ------------------------------------------------------------ DD fan / parallel
@pytest.mark.parametrize('branchless', [False, True])
@pytest.mark.parametrize('variant', ['plain', 'offsets', 'batch'])
def test_dd_fan_fp(oracle, branchless, variant):
    kw = {}
    if variant == 'offsets':
        kw = dict(off=(0.4, 0.3), c=(2.0, -1.0, 0.6))
    if variant == 'batch':
        kw = dict(B=2)
    g, ang, img = small_fan(**kw)
    p = _proj(g)
    fp = fan.distance_driven_fp(p, img, ang, branchless=branchless)
    ref = oracle.dd_fan_fp(g, img, ang, branchless=branchless)
    e = rel_l2(fp, ref)
    print('dd fan fp', branchless, variant, e)
    assert e < TOL


@pytest.mark.parametrize('branchless', [False, True])
@pytest.mark.parametrize('is_fbp', [False, True])
@pytest.mark.parametrize('variant', ['plain', 'offsets', 'batch'])
def test_dd_fan_bp(oracle, branchless, is_fbp, variant):
    kw = {}
    if variant == 'offsets':
        kw = dict(off=(0.4, 0.3), c=(2.0, -1.0, 0.6))
    if variant == 'batch':
        kw = dict(B=2)
    g, ang, img = small_fan(**kw)
    p = _proj(g)
    y = np.random.default_rng(4).uniform(size=[img.shape[0], len(ang), g.nv, g.nu]).astype(np.float32)
    bp = fan.distance_driven_bp(p, y, ang, is_fbp=is_fbp, branchless=branchless)
    ref = oracle.dd_fan_bp(g, y, ang, is_fbp=is_fbp, branchless=branchless)
    e = rel_l2(bp, ref)
    print('dd fan bp', branchless, is_fbp, variant, e)
    assert e < TOL


@pytest.mark.parametrize('branchless', [False, True])
@pytest.mark.parametrize('variant', ['plain', 'offsets', 'batch'])
def test_dd_parallel_fp_bp(oracle, branchless, variant):
    kw = {}
    if variant == 'offsets':
        kw = dict(off=(0.4, -0.2))
    if variant == 'batch':
        kw = dict(B=2)
    g, ang, img = small_parallel(**kw)
    p = _proj(g)
    fp = parallel.distance_driven_fp(p, img, ang, branchless=branchless)
    ref = oracle.dd_par_fp(g, img, ang, branchless=branchless)
    e = rel_l2(fp, ref)
    print('dd par fp', branchless, variant, e)
    assert e < TOL
    for is_fbp in (False, True):
        y = np.random.default_rng(6).uniform(size=ref.shape).astype(np.float32)
        bp = parallel.distance_driven_bp(p, y, ang, is_fbp=is_fbp, branchless=branchless)
        refb = oracle.dd_par_bp(g, y, ang, is_fbp=is_fbp, branchless=branchless)
        e = rel_l2(bp, refb)
        print('dd par bp', branchless, is_fbp, variant, e)
        assert e < TOL


# --------------------------------------------------------------- FBP pieces
@pytest.mark.parametrize('filt', ['hann', 'hamming', 'cosine', 'ramp'])
def test_fan_ramp_filter(oracle, filt):
    g, ang, img = small_fan(off=(0.25, 0.0))
    p = _proj(g)
    prj = oracle.dd_fan_fp(g, img, ang).astype(np.float32)
    out = fan.ramp_filter(p, prj, filt)
    ref = oracle.fan_ramp_filter(g, prj, filt)
    e = rel_l2(out, ref)
    print('fan filter', filt, e)
    assert e < TOL


@pytest.mark.parametrize('filt', ['hann', 'hamming', 'cosine', 'ramp'])
def test_parallel_ramp_filter(oracle, filt):
    g, ang, img = small_parallel()
    p = _proj(g)
    prj = oracle.dd_par_fp(g, img, ang).astype(np.float32)
    out = parallel.ramp_filter(p, prj, filt)
    ref = oracle.par_ramp_filter(g, prj, filt)
    e = rel_l2(out, ref)
    print('par filter', filt, e)
    assert e < TOL


@pytest.mark.parametrize('variant', ['plain', 'offsets', 'batch'])
def test_fan_fbp_bp(oracle, variant):
    kw = dict(nz=7, nv=6)
    if variant == 'offsets':
        kw.update(off=(0.4, 0.3), c=(2.0, -1.0, 0.6))
    if variant == 'batch':
        kw.update(B=2)
    g, ang, img = small_fan(**kw)
    p = _proj(g)
    y = np.random.default_rng(8).uniform(size=[img.shape[0], len(ang), g.nv, g.nu]).astype(np.float32)
    bp = fan.fbp_bp(p, y, ang)
    ref = oracle.fan_fbp_bp(g, y, ang)
    e = rel_l2(bp, ref)
    print('fan fbp bp', variant, e)
    assert e < TOL


@pytest.mark.parametrize('case', ['fan_32_lpc4', 'fan_32_lpc2', 'fan_2x20_lpc4', 'fan_2x8_lpc4', 'fan_5_lpc2',
                                  'parallel_32_lpc4', 'parallel_2x8_lpc2'])
def test_fbp_bp_split_lanes(oracle, tuning, case):
    """The fan FBP / parallel pixel-driven backprojection with 2 or 4 lanes per
    pixel (views' geometry spread over the pixel's lanes; 41 views, not a
    multiple of the lane count) against the oracle and one lane per pixel."""
    kind, lpc = case.split('_lpc')
    beam, dims = kind.split('_', 1)
    nz, B = (int(dims.split('x')[1]), int(dims.split('x')[0])) if 'x' in dims else (int(dims), 1)
    if beam == 'fan':
        g, ang, img = small_fan(nz=nz, nv=nz + 2, nview=41, off=(0.4, 0.3), c=(1.0, -0.5, 0.3), B=B)
        run, ref_fn = fan.fbp_bp, oracle.fan_fbp_bp
    else:
        g, ang, img = small_parallel(nz=nz, nv=nz, nview=41, off=(0.4, -0.2), B=B)
        run, ref_fn = parallel.pixel_driven_bp, oracle.par_pixel_bp
    y = np.random.default_rng(9).uniform(size=[B, len(ang), g.nv, g.nu]).astype(np.float32)
    tuning('fbp_bp_lanes_per_pixel', int(lpc))
    a = run(_proj(g), y, ang)
    tuning('fbp_bp_lanes_per_pixel', 1)
    b = run(_proj(g), y, ang)
    e = rel_l2(a, ref_fn(g, y, ang))
    print('fbp bp split', case, e, rel_l2(a, b))
    assert e < TOL
    assert rel_l2(a, b) < 1e-6


@pytest.mark.parametrize('case', ['fan_32', 'fan_8', 'fan_2x20', 'fan_ragged_32', 'fan_mixed_32', 'fan_wide_32',
                                  'fan_wide_8', 'parallel_32', 'parallel_2x8', 'parallel_wide_32'])
def test_fbp_bp_staged_cells_equal_global(oracle, tuning, case):
    """The split FBP / pixel-driven backprojection with each wavefront's
    detector cells staged in LDS (fbp_bp_stage 1; the default at 32 channels
    with 4 lanes per pixel) performs the arithmetic of the
    same kernel reading global memory in the same order: equal results, and
    both within the tolerance of the oracle.  'wide' geometries (4 mm pixels on
    0.3 mm cells) give cell ranges larger than a wavefront's 4 KiB of LDS, so
    every view group takes the global-memory reads; 'mixed' (1 mm pixels on
    0.3 mm cells) sits at the limit, some groups staged and some not; 'ragged'
    is a 45 x 43 image (partial tiles on two sides).  Staging needs detectors
    of at least four copy instructions' cells (256 / CH per chunk of CH
    channels): the 8-slice case has 160 cells of half the size."""
    parts = case.split('_')
    beam, nzs = parts[0], parts[-1]
    shape = parts[1] if len(parts) == 3 else 'plain'
    nz, B = (int(nzs.split('x')[1]), int(nzs.split('x')[0])) if 'x' in nzs else (int(nzs), 1)
    nview = 41
    if beam == 'fan':
        kw = dict(nx=48, ny=48, nz=nz, dx=1.0, dy=1.0, dz=1.0, cx=1.0, cy=-0.5, cz=0.3, nu=80, nv=nz + 2, du=1.5,
                  dv=1.0, off_u=0.4, off_v=0.3, dsd=200.0, dso=120.0)
        if nz == 8 and shape == 'plain':
            kw.update(nu=160, du=0.75)
        if shape == 'ragged':
            kw.update(nx=45, ny=43)
        if shape == 'mixed':
            kw.update(nu=420, du=0.3)
        if shape == 'wide':
            kw.update(nx=16, ny=16, dx=4.0, dy=4.0, nu=1000, du=0.3)
        ang = (np.arange(nview) * 2 * np.pi / nview).astype(np.float32)
        run, ref_fn = fan.fbp_bp, oracle.fan_fbp_bp
    else:
        kw = dict(nx=48, ny=48, nz=nz, dx=1.0, dy=1.0, dz=1.0, nu=70, nv=nz, du=1.0, dv=1.0, off_u=0.4, off_v=-0.2)
        if shape == 'wide':
            kw.update(nx=16, ny=16, dx=4.0, dy=4.0, nu=320, du=0.3)
        ang = (np.arange(nview) * np.pi / nview).astype(np.float32)
        run, ref_fn = parallel.pixel_driven_bp, oracle.par_pixel_bp
    g = Geo(**kw)
    y = np.random.default_rng(10).uniform(size=[B, nview, g.nv, g.nu]).astype(np.float32)
    tuning('fbp_bp_stage', 1)
    a = run(_proj(g), y, ang)
    tuning('fbp_bp_stage', 2)
    b = run(_proj(g), y, ang)
    e = rel_l2(a, ref_fn(g, y, ang))
    print('fbp bp staged', case, e, rel_l2(a, b))
    assert np.array_equal(a, b), rel_l2(a, b)
    assert e < TOL


# ----------------------------------------------------------------- tomo
def _tomo_case(B=1):
    g = Geo(nx=40, ny=36, nz=12, dx=1.0, dy=1.0, dz=1.0, cz=20.0, nu=56, nv=52, du=1.0, dv=1.0)
    nview = 7
    t = np.linspace(-0.4, 0.4, nview)
    src = np.stack([300 * np.sin(t), 5 * np.ones_like(t), 300 * np.cos(t) + 20], 1).astype(np.float32)
    dc = np.stack([np.zeros_like(t), np.zeros_like(t), -10 * np.ones_like(t)], 1).astype(np.float32)
    from ctprojector_amd.phantom import ellipsoid_phantom
    img = np.stack([ellipsoid_phantom(12, 36, 40, 1.0, 1.0, 1.0, seed=21 + b) for b in range(B)], 0)
    return g, dc, src, img


@pytest.mark.parametrize('branchless', [False, True])
def test_tomo_fp_bp(oracle, branchless):
    g, dc, src, img = _tomo_case(B=2)
    p = _proj(g)
    fp = tomo.distance_driven_fp(p, img, dc, src, branchless=branchless)
    ref = oracle.dd_tomo_fp(g, img, dc, src, branchless=branchless)
    e = rel_l2(fp, ref)
    print('tomo fp', branchless, e)
    assert e < TOL
    y = np.random.default_rng(3).uniform(size=ref.shape).astype(np.float32)
    bp = tomo.distance_driven_bp(p, y, dc, src, branchless=branchless)
    refb = oracle.dd_tomo_bp(g, y, dc, src, branchless=branchless)
    e = rel_l2(bp, refb)
    print('tomo bp', branchless, e)
    assert e < TOL


@pytest.mark.parametrize('fbp', [False, True])
def test_bp_view_groups_match_single_group(tuning, fbp):
    """Splitting the view loop of the pixel-driven BPs changes only the order
    of the final sums."""
    g, ang, img = small_fan(nview=64)
    p = _proj(g)
    y = np.random.default_rng(6).uniform(size=[1, 64, g.nv, g.nu]).astype(np.float32)
    run = (lambda: fan.fbp_bp(p, y, ang)) if fbp else (lambda: fan.distance_driven_bp(p, y, ang))
    a = run()
    tuning('bp_view_groups', 1)
    b = run()
    tuning('bp_view_groups', 7)
    c = run()
    assert rel_l2(a, b) < 1e-6 and rel_l2(c, b) < 1e-6


@pytest.mark.parametrize('fan_beam', [True, False])
def test_dd_fp_staged_rows_match_direct(oracle, tuning, fan_beam):
    """The LDS-staged forward projector (dd_fp_mode 2) against the oracle."""
    tuning('dd_fp_mode', 2)
    if fan_beam:
        g, ang, img = small_fan(nz=3, nv=4, off=(0.4, 0.3))
        fp = fan.distance_driven_fp(_proj(g), img, ang)
        ref = oracle.dd_fan_fp(g, img, ang)
    else:
        g, ang, img = small_parallel(nz=2, nv=2)
        fp = parallel.distance_driven_fp(_proj(g), img, ang)
        ref = oracle.dd_par_fp(g, img, ang)
    assert rel_l2(fp, ref) < TOL


@pytest.mark.parametrize('case', ['fan_3', 'fan_2x20', 'parallel_2', 'fan_1'])
def test_dd_fp_staged_rows_equal_direct(tuning, case):
    """The LDS-staged forward projector performs the direct kernel's arithmetic
    in the same order (rows copied to LDS by DMA): equal results, including
    several channel chunks (2 volumes x 20 slices = 40 channels)."""
    if case.startswith('fan'):
        nz, B = {'fan_3': (3, 1), 'fan_2x20': (20, 2), 'fan_1': (1, 1)}[case]
        g, ang, img = small_fan(n=56, nz=nz, nv=nz, nview=48, nu=80, off=(0.4, 0.3), B=B)
        run = lambda: fan.distance_driven_fp(_proj(g), img, ang)
    else:
        g, ang, img = small_parallel(nz=2, nv=2, B=2)
        run = lambda: parallel.distance_driven_fp(_proj(g), img, ang)
    tuning('dd_fp_mode', 1)
    a = run()
    tuning('dd_fp_mode', 2)
    b = run()
    assert np.array_equal(a, b), rel_l2(a, b)


@pytest.mark.parametrize('branchless', [False, True])
@pytest.mark.parametrize('case', ['fan_32_lpc4', 'fan_20_lpc8', 'fan_2x8_lpc2', 'fan_2x8_lpc4', 'fan_5_lpc2',
                                  'parallel_32_lpc4', 'parallel_2x8_lpc2'])
def test_dd_fp_split_lanes(oracle, tuning, branchless, case):
    """The forward projector with 2, 4 or 8 lanes per detector column (each
    lane a part of the channel vector) against the oracle and against one lane
    per column: 32 slices, 20 slices in a 32-channel chunk (unused channels),
    2 x 8 slices (16 channels), 5 slices (8 channels); 41 views (a block's
    last view group partly filled)."""
    kind, lpc = case.split('_lpc')
    beam, dims = kind.split('_', 1)
    nz, B = (int(dims.split('x')[1]), int(dims.split('x')[0])) if 'x' in dims else (int(dims), 1)
    if beam == 'fan':
        g, ang, img = small_fan(nz=nz, nv=nz + 2, nview=41, off=(0.4, 0.3), c=(1.0, -0.5, 0.3), B=B)
        run = lambda: fan.distance_driven_fp(_proj(g), img, ang, branchless=branchless)
        ref = oracle.dd_fan_fp(g, img, ang, branchless=branchless)
    else:
        g, ang, img = small_parallel(nz=nz, nv=nz, nview=41, off=(0.4, -0.2), B=B)
        run = lambda: parallel.distance_driven_fp(_proj(g), img, ang, branchless=branchless)
        ref = oracle.dd_par_fp(g, img, ang, branchless=branchless)
    tuning('dd_fp_lanes_per_column', int(lpc))
    a = run()
    tuning('dd_fp_lanes_per_column', 1)
    b = run()
    e = rel_l2(a, ref)
    print('dd fp split', case, branchless, e, rel_l2(a, b))
    assert e < TOL
    assert rel_l2(a, b) < 1e-6


@pytest.mark.parametrize('branchless', [False, True])
@pytest.mark.parametrize('case', ['fan_32_lpc2', 'fan_32_lpc4', 'fan_2x20_lpc2', 'fan_2x20_lpc4', 'fan_2x8_lpc4',
                                  'fan_5_lpc2', 'parallel_32_lpc2', 'parallel_2x8_lpc4'])
def test_dd_bp_split_lanes(oracle, tuning, branchless, case):
    """The pixel-driven backprojector with 2 or 4 lanes per pixel against the
    oracle and against one lane per pixel (FBP and plain weights): 32 slices,
    2 x 20 slices (two 32-channel chunks), 2 x 8 slices, 5 slices; 41 views
    (view groups that do not fill the pixel's lanes)."""
    kind, lpc = case.split('_lpc')
    beam, dims = kind.split('_', 1)
    nz, B = (int(dims.split('x')[1]), int(dims.split('x')[0])) if 'x' in dims else (int(dims), 1)
    if beam == 'fan':
        g, ang, img = small_fan(nz=nz, nv=nz + 2, nview=41, off=(0.4, 0.3), c=(1.0, -0.5, 0.3), B=B)
        mod, ref_fn = fan, oracle.dd_fan_bp
    else:
        g, ang, img = small_parallel(nz=nz, nv=nz, nview=41, off=(0.4, -0.2), B=B)
        mod, ref_fn = parallel, oracle.dd_par_bp
    y = np.random.default_rng(8).uniform(size=[B, len(ang), g.nv, g.nu]).astype(np.float32)
    for is_fbp in (False, True):
        tuning('dd_bp_lanes_per_pixel', int(lpc))
        a = mod.distance_driven_bp(_proj(g), y, ang, is_fbp=is_fbp, branchless=branchless)
        tuning('dd_bp_lanes_per_pixel', 1)
        b = mod.distance_driven_bp(_proj(g), y, ang, is_fbp=is_fbp, branchless=branchless)
        ref = ref_fn(g, y, ang, is_fbp=is_fbp, branchless=branchless)
        e = rel_l2(a, ref)
        print('dd bp split', case, branchless, is_fbp, e, rel_l2(a, b))
        assert e < TOL
        assert rel_l2(a, b) < 1e-6


# ------------------------------------------------------------ Siddon fan beam
@pytest.mark.parametrize('variant', ['plain', 'offsets_batch'])
def test_siddon_fan(oracle, variant):
    kw = {} if variant == 'plain' else dict(off=(0.4, 0.3), c=(1.5, -1.0, 0.4), B=2)
    g, ang, img = small_fan(nz=3, nv=3, **kw)
    p = _proj(g)
    fp = fan.siddon_fp(p, img, ang)
    assert rel_l2(fp, oracle.siddon_fan_fp(g, img, ang)) < TOL
    y = np.random.default_rng(4).uniform(size=fp.shape).astype(np.float32)
    bp = fan.siddon_bp(p, y, ang)
    assert rel_l2(bp, oracle.siddon_fan_bp(g, y, ang)) < TOL
    lhs = np.dot(fp.ravel().astype(np.float64), y.ravel().astype(np.float64))
    rhs = np.dot(img.ravel().astype(np.float64), bp.ravel().astype(np.float64))
    assert abs(lhs - rhs) / (np.linalg.norm(fp) * np.linalg.norm(y)) < 1e-6


@pytest.mark.parametrize('variant', ['plain', 'offsets_batch'])
def test_parallel_pixel_bp(oracle, variant):
    kw = {} if variant == 'plain' else dict(off=(0.3, -0.2), B=2)
    g, ang, img = small_parallel(nz=2, nv=2, **kw)
    y = np.random.default_rng(5).uniform(size=[img.shape[0], len(ang), g.nv, g.nu]).astype(np.float32)
    bp = parallel.pixel_driven_bp(_proj(g), y, ang)
    assert rel_l2(bp, oracle.par_pixel_bp(g, y, ang)) < TOL
