"""Pins the oracle's fan-beam Siddon projector (oracle.siddon_fan_fp/bp,
following cuda/siddonFan.cu:12-80):

* closed form: a uniform box phantom projects to value x |ray ∩ box| in every
  row whose plane lies in the volume (independent numpy slab clipping of the
  in-plane ray from the source through the arc detector point), 0 elsewhere;
* the BP is the exact transpose of the FP (dot-product test).
"""

import numpy as np

from helpers import Geo


def f32(x):
    return float(np.float32(x))


def test_uniform_box_closed_form(oracle):
    g = Geo(nx=20, ny=14, nz=3, dx=1.1, dy=0.9, dz=1.3, cx=0.5, cy=-0.4, cz=0.2, nu=48, nv=5, du=1.7, dv=1.2,
            dsd=300.0, dso=170.0, off_u=0.3, off_v=0.4)
    img = np.full([1, 3, 14, 20], 0.5, np.float32)
    ang = np.array([0.2, 1.3, 2.9, 4.4], np.float32)
    fp = oracle.siddon_fan_fp(g, img, ang)[0]
    da = f32(g.du / g.dsd)
    a = (-(np.arange(g.nu) - (g.nu - 1) / 2.0) - f32(g.off_u)) * da
    lo = np.array([f32(g.cx) - g.nx * f32(g.dx) / 2, f32(g.cy) - g.ny * f32(g.dy) / 2])
    hi = lo + np.array([g.nx * f32(g.dx), g.ny * f32(g.dy)])
    zlo, zhi = f32(g.cz) - g.nz * f32(g.dz) / 2, f32(g.cz) + g.nz * f32(g.dz) / 2
    for k, t in enumerate(ang.astype(np.float64)):
        S = np.array([g.dso * np.sin(t), -g.dso * np.cos(t)])
        rx, ry = g.dsd * np.sin(a), -g.dso + g.dsd * np.cos(a)
        Q = np.stack([rx * np.cos(t) - ry * np.sin(t), rx * np.sin(t) + ry * np.cos(t)], 1)
        d = Q - S
        with np.errstate(divide='ignore'):
            t0, t1 = (lo - S) / d, (hi - S) / d
        tin = np.maximum(np.max(np.minimum(t0, t1), 1), 0)
        tout = np.min(np.maximum(t0, t1), 1)
        chord = np.maximum(tout - tin, 0) * np.linalg.norm(d, axis=1)
        for iv in range(g.nv):
            z = (iv - (g.nv - 1) / 2.0 + f32(g.off_v)) * f32(g.dv)
            want = 0.5 * chord if zlo <= z < zhi else 0 * chord
            assert np.abs(fp[k, iv] - want).max() < 1e-9 * max(chord.max(), 1)
    assert fp.max() > 5


def test_bp_is_transpose(oracle):
    rng = np.random.default_rng(3)
    g = Geo(nx=18, ny=18, nz=2, dx=1.0, dy=1.0, dz=1.0, nu=36, nv=2, du=1.5, dv=1.0, dsd=250.0, dso=150.0)
    ang = rng.uniform(0, 2 * np.pi, 7).astype(np.float32)
    x = rng.uniform(size=[1, 2, 18, 18]).astype(np.float32)
    y = rng.uniform(size=[1, 7, 2, 36]).astype(np.float32)
    lhs = float((oracle.siddon_fan_fp(g, x, ang) * y).sum())
    rhs = float((x * oracle.siddon_fan_bp(g, y, ang)).sum())
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
