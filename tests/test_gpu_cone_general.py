"""GPU: cone Siddon geometries outside the voxel-driven window backprojector's
range, and volumes of 2^29 voxels or more, against the FP64 oracle.

* the ray-driven backprojector (atomic scatter, the reference's transpose;
  cone_bp_mode 2) on the default small case: oracle, dot test, equality with
  the window backprojector;
* the detector of the reference's example/projector.cfg (0.25 mm cells, dsd
  3000, dso 2700) with 1 mm voxels: a voxel shadow of more than 4 cells, so
  the backprojection takes the ray-driven path by itself;
* sources inside and just outside the volume's corner radius;
* z-slabs: small volumes forced into slabs (cone_slab_slices) equal the
  oracle and the unslabbed call; a 1024 x 1024 x 513 volume (more than 2^29
  voxels) on 2 views (the BP there against an FP32 bound: 1e-3, and a tenth
  of the reference arithmetic's own FP32 error).

Each case prints its name before its GPU work."""

import numpy as np
import pytest

from helpers import Geo, rel_l2, small_cone

pytestmark = pytest.mark.gpu
TOL = 1e-4


@pytest.fixture
def tuning():
    from ctprojector_amd import _lib
    keys = []

    def set_(k, v):
        keys.append(k)
        _lib.set_tuning(k, v)
    yield set_
    for k in keys:
        _lib.set_tuning(k, 0)


def _say(*a):
    print('[cone-general]', *a, flush=True)


def _check(oracle, g, geo, img, seed=4, dot=True):
    from ctprojector_amd.projector.numpy import cone
    from ctprojector_amd.projector.numpy.ct_projector import ct_projector
    p = g.to(ct_projector())
    args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
    fp = cone.siddon_fp_arbitrary(p, img, *args)
    e_fp = rel_l2(fp, oracle.siddon_cone_fp(g, img, *args))
    y = np.random.default_rng(seed).uniform(size=fp.shape).astype(np.float32)
    bp = cone.siddon_bp_arbitrary(p, y, *args)
    e_bp = rel_l2(bp, oracle.siddon_cone_bp(g, y, *args))
    _say('fp %.2e bp %.2e' % (e_fp, e_bp))
    assert e_fp < TOL and e_bp < TOL
    if dot:
        lhs = np.dot(fp.ravel().astype(np.float64), y.ravel().astype(np.float64))
        rhs = np.dot(img.ravel().astype(np.float64), bp.ravel().astype(np.float64))
        assert abs(lhs - rhs) / (np.linalg.norm(fp) * np.linalg.norm(y)) < 1e-6
    return fp, bp


def test_scatter_backprojection(oracle, tuning):
    g, geo, img = small_cone(seed=31)
    _say('default geometry, window backprojector')
    _, a = _check(oracle, g, geo, img)
    tuning('cone_bp_mode', 2)
    _say('default geometry, ray-driven scatter')
    _, b = _check(oracle, g, geo, img)
    assert rel_l2(b, a) < 1e-5


def test_reference_example_detector(oracle):
    """example/projector.cfg: du = dv = 0.25 mm, dsd 3000, dso 2700 (nu, nv
    reduced to the volume's shadow); 48 x 48 x 32 voxels of 1 mm."""
    from ctprojector_amd.phantom import circular_cone_geometry, ellipsoid_phantom
    g = Geo(nx=48, ny=48, nz=32, dx=1.0, dy=1.0, dz=1.0, cx=0.4, cy=-0.3, cz=0.2, nu=256, nv=176, du=0.25, dv=0.25,
            off_u=0.3, off_v=-0.2, dsd=3000.0, dso=2700.0)
    geo = circular_cone_geometry(6, 2700.0, 3000.0, -1.0, 1.0, jitter=0.5, seed=33)
    img = ellipsoid_phantom(32, 48, 48, 1.0, 1.0, 1.0, seed=33)[None]
    _say('example detector 256x176 @0.25 mm, dsd 3000, dso 2700, 48x48x32 @1 mm')
    _check(oracle, g, geo, img)


def test_reference_example_detector_large_volume(oracle):
    """The example detector on a 256 x 256 x 96 volume of 1 mm voxels (ray-driven
    backprojection on a grid where accumulated crossings would drift): FP and
    BP against the FP64 oracle at 1e-4; the reference's own FP32 arithmetic is
    printed beside it."""
    from ctprojector_amd.phantom import circular_cone_geometry, ellipsoid_phantom
    from ctprojector_amd.projector.numpy import cone
    from ctprojector_amd.projector.numpy.ct_projector import ct_projector
    g = Geo(nx=256, ny=256, nz=96, dx=1.0, dy=1.0, dz=1.0, cx=0.7, cy=-0.4, nu=1200, nv=480, du=0.25, dv=0.25,
            off_u=0.3, off_v=-0.2, dsd=3000.0, dso=2700.0)
    geo = circular_cone_geometry(2, 2700.0, 3000.0, -1.0, 1.0, jitter=0.5, seed=41, angle_range=np.pi / 2)
    img = ellipsoid_phantom(96, 256, 256, 1.0, 1.0, 1.0, seed=41)[None]
    p = g.to(ct_projector())
    args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
    _say('example detector 1200x480 @0.25 mm, 256x256x96 @1 mm: FP')
    fp = cone.siddon_fp_arbitrary(p, img, *args)
    e = rel_l2(fp, oracle.siddon_cone_fp(g, img, *args))
    _say('fp rel-L2 %.2e' % e)
    assert e < TOL
    y = np.random.default_rng(6).uniform(size=fp.shape).astype(np.float32)
    _say('BP (ray-driven)')
    bp = cone.siddon_bp_arbitrary(p, y, *args)
    ref = oracle.siddon_cone_bp(g, y, *args)
    e = rel_l2(bp, ref)
    e32 = rel_l2(oracle.siddon_cone_bp(g, y, *args, prec='f'), ref)
    _say('bp rel-L2 %.2e (the reference arithmetic in FP32: %.2e)' % (e, e32))
    assert e < TOL


@pytest.mark.parametrize('dso', [8.0, 15.0])
def test_source_inside_or_near_volume(oracle, dso):
    """A 24^3 volume of 1 mm voxels (half width 12 mm, corner radius 20.8 mm)
    with the source at 8 mm (inside) or 15 mm (between the faces and the
    corners) from the centre."""
    from ctprojector_amd.phantom import circular_cone_geometry, ellipsoid_phantom
    g = Geo(nx=24, ny=24, nz=24, dx=1.0, dy=1.0, dz=1.0, nu=64, nv=48, du=2.0, dv=2.0, dsd=dso + 40.0, dso=dso)
    geo = circular_cone_geometry(5, dso, dso + 40.0, -2.0, 2.0, jitter=0.3, seed=35)
    img = ellipsoid_phantom(24, 24, 24, 1.0, 1.0, 1.0, seed=35)[None]
    _say('source at %.0f mm from the centre of a 24 mm cube' % dso)
    _check(oracle, g, geo, img)


@pytest.mark.parametrize('slices', [7, 24])
def test_forced_slabs(oracle, tuning, slices):
    from ctprojector_amd.projector.numpy import cone
    from ctprojector_amd.projector.numpy.ct_projector import ct_projector
    g, geo, img = small_cone(seed=37, c=(0.5, -1.0, 1.5))
    img2 = np.concatenate([img, 0.5 * img[:, ::-1].copy()], 0)
    p = g.to(ct_projector())
    args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
    whole_fp = cone.siddon_fp_arbitrary(p, img2, *args)
    y = np.random.default_rng(3).uniform(size=whole_fp.shape).astype(np.float32)
    whole_bp = cone.siddon_bp_arbitrary(p, y, *args)
    tuning('cone_slab_slices', slices)
    _say('slabs of %d slices, batch of 2' % slices)
    fp = cone.siddon_fp_arbitrary(p, img2, *args)
    bp = cone.siddon_bp_arbitrary(p, y, *args)
    assert rel_l2(fp, oracle.siddon_cone_fp(g, img2, *args)) < TOL
    assert rel_l2(bp, oracle.siddon_cone_bp(g, y, *args)) < TOL
    assert rel_l2(fp, whole_fp) < 1e-5 and rel_l2(bp, whole_bp) < 1e-5


def test_volume_of_more_than_2pow29_voxels(oracle):
    import torch
    from ctprojector_amd.phantom import circular_cone_geometry
    from ctprojector_amd.projector.cupy import cone, ct_projector
    nx, ny, nz = 1024, 1024, 513
    assert nx * ny * nz > 1 << 29
    g = Geo(nx=nx, ny=ny, nz=nz, dx=0.25, dy=0.25, dz=0.25, cz=0.3, nu=96, nv=64, du=5.0, dv=5.0, off_u=0.2,
            dsd=1000.0, dso=600.0)
    geo = circular_cone_geometry(2, 600.0, 1000.0, -3.0, 3.0, jitter=0.5, seed=39, angle_range=np.pi / 2)
    dev = torch.device('cuda', 0)
    x = torch.linspace(0, 6.0, nx, device=dev)
    yv = torch.linspace(0, 4.0, ny, device=dev)
    z = torch.linspace(0, 1.0, nz, device=dev)
    img = (0.01 * (1 + torch.sin(x)[None, None, :] * torch.cos(yv)[None, :, None]) * (1 + z[:, None, None]))[None]
    img = img.contiguous()
    p = g.to(ct_projector())
    gd = [torch.from_numpy(a).to(dev) for a in (geo.det_center, geo.det_u, geo.det_v, geo.src)]
    _say('1024x1024x513 volume (%d voxels), 2 views, 96x64 detector: FP' % (nx * ny * nz))
    fp = cone.siddon_fp_arbitrary(p, img, *gd).cpu().numpy()
    img_h = img.cpu().numpy()
    del img
    hg = (geo.det_center, geo.det_u, geo.det_v, geo.src)
    e = rel_l2(fp, oracle.siddon_cone_fp(g, img_h, *hg))
    _say('fp rel-L2 %.2e' % e)
    assert e < TOL
    del img_h
    y = np.random.default_rng(5).uniform(size=fp.shape).astype(np.float32)
    _say('BP')
    bp = cone.siddon_bp_arbitrary(p, torch.from_numpy(y).to(dev), *gd).cpu().numpy()
    ref = oracle.siddon_cone_bp(g, y, *hg)
    e = rel_l2(bp, ref)
    # FP32 bound: a 0.25 mm voxel seen from 600 mm is ~2400 voxels from the
    # source, so each chord, a difference of ray parameters of that size, is
    # good to ~1e-4 in FP32.  The reference's own FP32 traversal (the oracle's
    # FP32 mirror: crossings accumulated step by step) is 1.6% off here.
    e32 = rel_l2(oracle.siddon_cone_bp(g, y, *hg, prec='f'), ref)
    _say('bp rel-L2 %.2e (the reference arithmetic in FP32: %.2e)' % (e, e32))
    assert e < 1e-3 and e < 0.1 * e32
