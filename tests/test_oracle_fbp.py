"""Pins the FP64 oracle's ramp filter and fan FBP backprojection
(oracle/oracle.cpp following cuda/fbpMath.cu:90-268, cuda/fbp.cu:15-162,264-359,
cuda/fbpParallel.cu:14-109):

* the windowed kernel equals the reference's frequency-domain procedure (spatial
  Ram-Lak kernel -> DFT -> window -> inverse DFT / L) evaluated with numpy.fft;
* the filtered projections equal the reference's padded-FFT filtering
  (zero-pad to L = 2nu-1, multiply by the DFT of h_w, inverse, read at nu-1)
  evaluated with numpy.fft;
* the Ram-Lak kernel's closed-form values;
* a full fan FBP (DD projection -> ramp filter -> fbp_bp) of a uniform disk
  reconstructs the disk's attenuation inside it (textbook FBP).
"""

import numpy as np
import pytest

from ctprojector_amd.phantom import Ellipse, rasterize_ellipses
from helpers import Geo, rel_l2

PI = float(np.float32(3.141592657))


def spatial_ramp(nu, da, equispace):
    L = 2 * nu - 1
    k = np.arange(L) - (nu - 1)
    h = np.zeros(L)
    h[k == 0] = 1.0 / (4 * da * da)
    odd = k % 2 != 0
    if equispace:
        h[odd] = -1.0 / (PI ** 2 * k[odd].astype(float) ** 2 * da * da)
    else:
        h[odd] = -1.0 / (PI ** 2 * np.sin(k[odd] * da) ** 2)
    return h


def window(name, L, pi=PI):
    """cuda/fbpMath.cu:180-198.  The reference evaluates the cosines in FP32 with
    PI = 3.141592657f; the oracle's 3-tap form of Hamming/Hann is the exact
    identity for the window with the true pi (relative difference < 1e-7, below
    FP32 rounding of the reference's own window)."""
    f = np.arange(L)
    if name == 'hamming':
        return float(np.float32(0.54)) + float(np.float32(0.46)) * np.cos(2 * pi * f / L)
    if name == 'hann':
        return 0.5 + 0.5 * np.cos(2 * pi * f / L)
    if name == 'cosine':
        return np.abs(np.cos(pi * f / L))
    return np.ones(L)


@pytest.mark.parametrize('filt', ['hann', 'hamming', 'cosine', 'ramp'])
@pytest.mark.parametrize('equispace', [False, True])
def test_windowed_kernel_matches_fft_procedure(oracle, filt, equispace):
    nu, da = 37, 0.013 if not equispace else 0.9
    h = spatial_ramp(nu, da, equispace)
    L = 2 * nu - 1
    exact_pi = filt in ('hann', 'hamming')
    ref = np.real(np.fft.ifft(np.fft.fft(h) * window(filt, L, np.pi if exact_pi else PI)))
    got = oracle.ramp_kernel(nu, da, filt, equispace)
    # FFT roundoff scales with the spatial kernel's magnitude (the windowed
    # kernel's odd taps cancel down to ~1e-8 of it)
    assert np.abs(got - ref).max() < 1e-13 * np.abs(h).max()
    # the reference's FP32 pi moves the window by less than FP32 rounding
    ref_f = np.real(np.fft.ifft(np.fft.fft(h) * window(filt, L, PI)))
    assert np.abs(got - ref_f).max() < 1e-7 * np.abs(h).max()
    # a plausible mistake: windowing the spatial kernel directly
    if filt != 'ramp':
        assert np.abs(h * window(filt, 2 * nu - 1) - ref).max() > 1e-6 * np.abs(h).max()


def test_ram_lak_closed_form(oracle):
    h = oracle.ramp_kernel(5, 0.5, 'ramp', True)
    assert h[4] == pytest.approx(1 / (4 * 0.25))
    assert h[5] == pytest.approx(-1 / (PI ** 2 * 0.25))
    assert h[6] == 0
    assert h[7] == pytest.approx(-1 / (PI ** 2 * 9 * 0.25))


def _fft_filter(prj, h, pw, scale):
    """The reference's procedure in FP64: zero-pad each row to L, multiply the
    real DFTs, inverse, take [nu-1, 2nu-1) and scale."""
    nu = prj.shape[-1]
    L = 2 * nu - 1
    x = np.zeros(prj.shape[:-1] + (L,))
    x[..., :nu] = prj.astype(np.float64) * pw
    y = np.fft.irfft(np.fft.rfft(x, axis=-1) * np.fft.rfft(h), n=L, axis=-1)
    return y[..., nu - 1:] * scale


@pytest.mark.parametrize('filt', ['hann', 'cosine'])
def test_fan_filter_matches_padded_fft(oracle, filt):
    rng = np.random.default_rng(1)
    g = Geo(nu=41, nv=3, du=1.3, dsd=900.0, dso=500.0, off_u=0.3)
    prj = rng.uniform(size=[2, 5, 3, 41]).astype(np.float32)
    da = float(np.float32(g.du / g.dsd))
    h = oracle.ramp_kernel(41, da, filt, False)
    off = float(np.float32(g.off_u))
    pw = np.cos(((np.arange(41) - 20.0) - off) * da)
    ref = _fft_filter(prj, h, pw, PI / 5 * da / 500.0)
    assert rel_l2(oracle.fan_ramp_filter(g, prj, filt), ref) < 1e-12


def test_parallel_filter_matches_padded_fft(oracle):
    rng = np.random.default_rng(2)
    g = Geo(nu=30, nv=2, du=0.8)
    prj = rng.uniform(size=[1, 7, 2, 30]).astype(np.float32)
    du = float(np.float32(0.8))
    h = oracle.ramp_kernel(30, du, 'hamming', True)
    ref = _fft_filter(prj, h, 1.0, PI / 7 * du)
    assert rel_l2(oracle.par_ramp_filter(g, prj, 'hamming'), ref) < 1e-12


def test_fan_fbp_reconstructs_uniform_disk(oracle):
    """DD fan projection -> Ram-Lak/Hann filter -> fbp_bp of a uniform disk."""
    n, d, mu = 96, 0.5, 0.02
    img = rasterize_ellipses([Ellipse(0.0, 0.0, 15.0, 15.0, 0.0, mu)], n, n, d, d)[None, None]
    g = Geo(nx=n, ny=n, nz=1, dx=d, dy=d, dz=1.0, nu=160, nv=1, du=0.6, dv=1.0, dsd=300.0, dso=200.0)
    ang = (np.arange(360) * 2 * np.pi / 360).astype(np.float32)
    prj = oracle.dd_fan_fp(g, img, ang).astype(np.float32)
    rec = oracle.fan_fbp_bp(g, oracle.fan_ramp_filter(g, prj, 'hann').astype(np.float32), ang)[0, 0]
    yy, xx = np.mgrid[:n, :n]
    r = np.hypot((xx - (n - 1) / 2) * d, (yy - (n - 1) / 2) * d)
    inside = rec[r < 10.0]
    outside = rec[r > 20.0]
    assert abs(inside.mean() - mu) < 0.02 * mu
    assert inside.std() < 0.05 * mu
    assert abs(outside.mean()) < 0.02 * mu


def test_fbp_bp_close_to_dd_bp_with_fbp_weight(oracle):
    """Pixel-driven (bilinear) and distance-driven (box) FBP-weighted fan BPs of a
    smooth sinogram agree closely."""
    g = Geo(nx=48, ny=48, nz=1, dx=1.0, dy=1.0, dz=1.0, nu=90, nv=1, du=1.5, dv=1.0, dsd=300.0, dso=200.0)
    ang = (np.arange(60) * 2 * np.pi / 60).astype(np.float32)
    u = np.linspace(-1, 1, 90)
    y = np.broadcast_to(np.exp(-4 * u ** 2), [1, 60, 1, 90]).astype(np.float32)
    a = oracle.fan_fbp_bp(g, y, ang)
    b = oracle.dd_fan_bp(g, y, ang, is_fbp=True)
    assert rel_l2(a, b) < 0.02


def test_parallel_pixel_bp_closed_forms(oracle):
    """Bilinear interpolation reproduces a linear function exactly: a
    projection equal to its own column index backprojects (one view) to
    u(x, y) = -((x cos + y sin)/du + off_u) + (nu-1)/2 (pixel centres), which pins
    the mirrored detector axis, the offset and the angle convention
    (cuda/fbpParallel.cu:224-240); constant data over nview views gives nview."""
    g = Geo(nx=12, ny=10, nz=2, dx=0.8, dy=1.1, dz=1.0, cx=0.4, cy=-0.3, nu=64, nv=2, du=0.7, dv=1.0, off_u=0.6)
    t = 0.7
    y = np.broadcast_to(np.arange(64, dtype=np.float32), [1, 1, 2, 64]).copy()
    bp = oracle.par_pixel_bp(g, y, np.array([t], np.float32))[0]
    f = lambda v: float(np.float32(v))
    xs = (np.arange(12) + 0.5 - 6) * f(0.8) + f(0.4)
    ys = (np.arange(10) + 0.5 - 5) * f(1.1) + f(-0.3)
    c, s = np.cos(f(t)), np.sin(f(t))
    u = -((xs[None, :] * c + ys[:, None] * s) / f(0.7) + f(0.6)) + 63 / 2.0
    assert np.abs(bp[0] - u).max() < 1e-9
    ones = np.ones([1, 9, 2, 64], np.float32)
    assert np.abs(oracle.par_pixel_bp(g, ones, np.linspace(0, 3, 9).astype(np.float32)) - 9).max() < 1e-9
