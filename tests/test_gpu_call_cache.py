"""GPU: the per-device cache of the DD entry points' geometry tables.

Repeated calls with the same geometry and angles reuse the tables; a call
with other angles (same count), another batch count or after a tuning change
gets tables of its own; more geometries than the cache holds evict the least
recently used ones.  Every result equals a call that built its tables fresh,
and the FP64 oracle."""

import numpy as np
import pytest

from helpers import rel_l2, small_fan

pytestmark = pytest.mark.gpu


def _proj(g):
    from ctprojector_amd.projector.numpy.ct_projector import ct_projector
    return g.to(ct_projector())


def _say(*a):
    print('[call-cache]', *a, flush=True)


def test_repeated_and_changed_calls(oracle):
    from ctprojector_amd import _lib
    from ctprojector_amd.projector.numpy import fan_equiangluar as fan
    g, ang, img = small_fan(n=48, nz=16, nv=16, nview=60, nu=80, off=(0.3, 0.2))
    p = _proj(g)
    _say('first call, repeated call')
    a1 = fan.distance_driven_fp(p, img, ang)
    a2 = fan.distance_driven_fp(p, img, ang)
    assert np.array_equal(a1, a2)
    assert rel_l2(a1, oracle.dd_fan_fp(g, img, ang)) < 1e-4
    _say('same count, other angles')
    ang_b = (ang + np.float32(0.013)).astype(np.float32)
    b = fan.distance_driven_fp(p, img, ang_b)
    assert rel_l2(b, oracle.dd_fan_fp(g, img, ang_b)) < 1e-4
    assert not np.array_equal(a1, b)
    _say('backprojection of both, twice')
    y = np.random.default_rng(3).uniform(size=a1.shape).astype(np.float32)
    for _ in range(2):
        bp_a = fan.distance_driven_bp(p, y, ang)
        bp_b = fan.distance_driven_bp(p, y, ang_b)
        assert rel_l2(bp_a, oracle.dd_fan_bp(g, y, ang)) < 1e-4
        assert rel_l2(bp_b, oracle.dd_fan_bp(g, y, ang_b)) < 1e-4
    _say('a tuning change between calls')
    try:
        _lib.set_tuning('dd_fp_lanes_per_column', 2)
        c = fan.distance_driven_fp(p, img, ang)
    finally:
        _lib.set_tuning('dd_fp_lanes_per_column', 0)
    assert rel_l2(c, a1) < 1e-6
    _say('batch of two with the same geometry')
    img2 = np.concatenate([img, 0.5 * img], 0)
    d2 = fan.distance_driven_fp(p, img2, ang)
    assert rel_l2(d2[0], a1[0]) < 1e-6
    assert rel_l2(d2[1], 0.5 * a1[0]) < 1e-6


def test_eviction_keeps_results(oracle):
    from ctprojector_amd.projector.numpy import fan_equiangluar as fan
    g, ang, img = small_fan(n=32, nz=2, nv=2, nview=24, nu=48)
    p = _proj(g)
    first = fan.distance_driven_fp(p, img, ang)
    _say('20 geometries (view counts 25 .. 44), more than the cache holds')
    for k in range(20):
        a = (np.arange(25 + k) * 2 * np.pi / (25 + k)).astype(np.float32)
        out = fan.distance_driven_fp(p, img, a)
        if k % 7 == 0:
            assert rel_l2(out, oracle.dd_fan_fp(g, img, a)) < 1e-4
    _say('the first geometry again (evicted and rebuilt)')
    again = fan.distance_driven_fp(p, img, ang)
    assert np.array_equal(first, again)
    from ctprojector_amd import _lib
    _say('after ctpReleaseWorkspace (scratch and cached tables freed)')
    assert _lib.release_workspace() == 0
    assert np.array_equal(first, fan.distance_driven_fp(p, img, ang))
