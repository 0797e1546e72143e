"""GPU parity of the distance-driven modes the numpy front end does not reach:
the force-FBP bit (typeProjector 4, the mode of the reference's TensorFlow 2D
ops, cuda/distanceDriven2D.cu:22-25, cuda/distanceDriven2DBoxInt.cu:283-287,501)
in every combination with the other bits, called through the C ABI on batches
of two, and the split-lane backprojectors on images whose sides are not
multiples of the pixel tiles (45 x 43: ragged tiles at both image edges)."""

from ctypes import POINTER, c_float, c_int, c_ulong

import numpy as np
import pytest

from ctprojector_amd import _lib
from ctprojector_amd.phantom import ellipsoid_phantom
from ctprojector_amd.projector.numpy import fan_equiangluar as fan, parallel
from ctprojector_amd.projector.numpy.ct_projector import ct_projector

from helpers import Geo, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-4


def _p(a):
    return a.ctypes.data_as(POINTER(c_float))


def _call(beam, direction, g, out, inp, ang, typ):
    """cDistanceDriven{Fan,Parallel}{Projection,Backprojection} with an explicit
    typeProjector (argument lists of cuda/distanceDriven2D.cu:109-205,384-482)."""
    img_shape = out.shape if direction == 'bp' else inp.shape
    prj_shape = inp.shape if direction == 'bp' else out.shape
    grid = (c_ulong(img_shape[0]), c_ulong(img_shape[3]), c_ulong(img_shape[2]), c_ulong(img_shape[1]),
            c_float(g.dx), c_float(g.dy), c_float(g.dz), c_float(g.cx), c_float(g.cy), c_float(g.cz))
    da = g.du / g.dsd if beam == 'fan' else g.du
    det = (c_ulong(prj_shape[3]), c_ulong(prj_shape[2]), c_ulong(prj_shape[1]), c_float(da), c_float(g.dv),
           c_float(g.off_u), c_float(g.off_v))
    if beam == 'fan':
        det = det + (c_float(g.dsd), c_float(g.dso))
    name = 'cDistanceDriven%s%s' % ('Fan' if beam == 'fan' else 'Parallel',
                                    'Projection' if direction == 'fp' else 'Backprojection')
    err = getattr(_lib.module(), name)(_p(out), _p(inp), _p(ang), *grid, *det, c_int(typ))
    assert err == 0, (name, err)
    return out


def _geo(beam, n=(40, 36), nz=3, nv=3):
    kw = dict(nx=n[0], ny=n[1], nz=nz, dx=1.1, dy=0.9, dz=1.0, cx=1.5, cy=-0.8, cz=0.3, nv=nv, dv=1.0,
              off_u=0.35, off_v=0.2)
    if beam == 'fan':
        kw.update(nu=80, du=1.5, dsd=200.0, dso=120.0)
    else:
        kw.update(nu=72, du=0.9)
    return Geo(**kw)


@pytest.mark.parametrize('beam', ['fan', 'parallel'])
@pytest.mark.parametrize('typ', [4, 5, 6, 7])
def test_dd_force_fbp_fp_bp(oracle, beam, typ):
    g = _geo(beam)
    nview = 37                                            # includes views at 45-degree multiples +- rounding
    span = 2 * np.pi if beam == 'fan' else np.pi
    ang = (np.arange(nview) * span / nview).astype(np.float32)
    B = 2
    img = np.stack([ellipsoid_phantom(g.nz, g.ny, g.nx, g.dz, g.dy, g.dx, seed=40 + b) for b in range(B)], 0)
    fp = _call(beam, 'fp', g, np.zeros([B, nview, g.nv, g.nu], np.float32), np.ascontiguousarray(img), ang, typ)
    ref_fp = (oracle.dd_fan_fp if beam == 'fan' else oracle.dd_par_fp)(g, img, ang, type_projector=typ)
    e = rel_l2(fp, ref_fp)
    print('force-fbp fp', beam, typ, e)
    assert e < TOL
    y = np.random.default_rng(typ).uniform(size=fp.shape).astype(np.float32)
    bp = _call(beam, 'bp', g, np.zeros([B, g.nz, g.ny, g.nx], np.float32), y, ang, typ)
    ref_bp = (oracle.dd_fan_bp if beam == 'fan' else oracle.dd_par_bp)(g, y, ang, type_projector=typ)
    e = rel_l2(bp, ref_bp)
    print('force-fbp bp', beam, typ, e)
    assert e < TOL


@pytest.fixture
def tuning():
    keys = []

    def set_(k, v):
        keys.append(k)
        _lib.set_tuning(k, v)
    yield set_
    for k in keys:
        _lib.set_tuning(k, 0)


@pytest.mark.parametrize('beam', ['fan', 'parallel'])
@pytest.mark.parametrize('nz,lanes', [(8, 2), (32, 4), (32, 2)])
def test_split_bp_ragged_image(oracle, tuning, beam, nz, lanes):
    """Split-lane DD BP (plain and FBP weights) and split-lane FBP / pixel BP on
    a 45 x 43 image: the 16 x 8 and 8 x 8 pixel tiles leave partial tiles on
    two sides, whose lanes must neither write nor disturb their neighbours."""
    g = _geo(beam, n=(45, 43), nz=nz, nv=nz)
    nview = 29
    span = 2 * np.pi if beam == 'fan' else np.pi
    ang = (np.arange(nview) * span / nview + 0.01).astype(np.float32)
    y = np.random.default_rng(nz + lanes).uniform(size=[1, nview, g.nv, g.nu]).astype(np.float32)
    p = g.to(ct_projector())
    mod = fan if beam == 'fan' else parallel
    ref_fn = oracle.dd_fan_bp if beam == 'fan' else oracle.dd_par_bp
    tuning('dd_bp_lanes_per_pixel', lanes)
    for is_fbp in (False, True):
        a = mod.distance_driven_bp(p, y, ang, is_fbp=is_fbp)
        e = rel_l2(a, ref_fn(g, y, ang, is_fbp=is_fbp))
        print('dd bp split ragged', beam, nz, lanes, is_fbp, e)
        assert e < TOL
    tuning('fbp_bp_lanes_per_pixel', lanes)
    if beam == 'fan':
        a, ref = fan.fbp_bp(p, y, ang), oracle.fan_fbp_bp(g, y, ang)
    else:
        a, ref = parallel.pixel_driven_bp(p, y, ang), oracle.par_pixel_bp(g, y, ang)
    e = rel_l2(a, ref)
    print('fbp bp split ragged', beam, nz, lanes, e)
    assert e < TOL
