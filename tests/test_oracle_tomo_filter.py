"""Pins the oracle's tomosynthesis filter (oracle.tomo_filter, the reference's
cuda/fbpTomo.cu procedure with numpy.fft):

* with the Ram-Lak code it equals the pinned parallel-beam Ram-Lak filter at the
  view's virtual pitch du cos(b) (same kernel, same scale), for oblique views;
* for a source straight above the detector (b = 0) with dx = du and cutoffs 1
  the Hann window is the standard one except at one clipped frequency bin, so
  it agrees with the parallel-beam Hann filter to ~1e-4;
* cutoffs below 1 change the result (the window is really applied).
"""

import numpy as np

from helpers import Geo, rel_l2


def _geo(angles, dist=600.0):
    t = np.asarray(angles, np.float64)
    src = np.stack([dist * np.sin(t), 0 * t, dist * np.cos(t)], 1).astype(np.float32)
    dc = np.zeros_like(src)
    return dc, src


def test_ramlak_equals_parallel_filter_at_virtual_pitch(oracle):
    rng = np.random.default_rng(0)
    prj = rng.uniform(size=[1, 3, 4, 33]).astype(np.float32)
    ang = [-0.4, 0.1, 0.35]
    dc, src = _geo(ang)
    got = oracle.tomo_filter(prj, dc, src, du=0.8, dx=1.0, dz=1.0, filter_type='ramp')
    for v, t in enumerate(ang):
        ex, ez = float(dc[v, 0] - src[v, 0]), float(dc[v, 2] - src[v, 2])
        vdu = float(np.float32(0.8)) * (-ez / np.hypot(ex, ez))
        ref = oracle.par_ramp_filter(Geo(nu=33, nv=4, du=vdu), prj[:, v:v + 1], 'ramp')
        # par_ramp_filter uses nview = 1 in its scale; the tomo scale has nview = 3
        assert rel_l2(got[:, v:v + 1] * 3, ref) < 1e-6


def test_vertical_hann_matches_parallel_hann(oracle):
    rng = np.random.default_rng(1)
    prj = rng.uniform(size=[2, 1, 3, 40]).astype(np.float32)
    dc, src = _geo([0.0])
    got = oracle.tomo_filter(prj, dc, src, du=0.5, dx=0.5, dz=1.0, filter_type='hann')
    ref = oracle.par_ramp_filter(Geo(nu=40, nv=3, du=0.5), prj, 'hann')
    assert rel_l2(got, ref) < 3e-4


def test_cutoff_changes_result(oracle):
    prj = np.random.default_rng(2).uniform(size=[1, 2, 2, 30]).astype(np.float32)
    dc, src = _geo([0.2, -0.3])
    a = oracle.tomo_filter(prj, dc, src, 0.7, 1.0, 1.0, 'hamming', 1.0, 1.0)
    b = oracle.tomo_filter(prj, dc, src, 0.7, 1.0, 1.0, 'hamming', 0.5, 0.5)
    assert rel_l2(a, b) > 1e-2
