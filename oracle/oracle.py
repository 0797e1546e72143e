"""Python front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` (built from ``oracle/oracle.cpp`` by
``oracle/Makefile``) and exposes the reference's projector functions as numpy
calls.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import this module; the product never does.

Every function takes a geometry object ``p`` with the attributes of the
reference's ``ct_projector`` class (``nx ny nz dx dy dz cx cy cz nu nv du dv
off_u off_v dsd dso``; see src/ct_projector/projector/numpy/ct_projector.py:31-50)
and arrays laid out as the reference lays them out: images ``[B, nz, ny, nx]``,
projections ``[B, nview, nv, nu]``, geometry rows ``[nview, 3]`` as (x, y, z).

``prec='d'`` selects the FP64 oracle, ``prec='f'`` the FP32 mirror of the
reference arithmetic (OpenMP).  The argument conventions follow the numpy
wrappers of the reference:
  * cone: src/ct_projector/projector/numpy/cone.py:15-170
  * fan:  src/ct_projector/projector/numpy/fan_equiangluar.py:19-366 (du/dsd is
    passed as the angular pitch, :60,117,287,354)
  * parallel: src/ct_projector/projector/numpy/parallel.py:17-254
  * tomo: src/ct_projector/projector/numpy/tomo.py
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CTP_ORACLE_LIB selects another build of oracle.cpp (the benchmark's CPU
# baseline loads the -march=native build made on the host it runs on)
_LIB_PATH = os.environ.get('CTP_ORACLE_LIB') or os.path.join(_HERE, 'liboracle.so')
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle library with make (g++ + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(['make', '-s', '-C', _HERE, '-B' if force else 'liboracle.so'])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_num_threads.restype = ctypes.c_int
    return _lib


def num_threads() -> int:
    return lib().orc_num_threads()


def set_threads(n: int) -> None:
    lib().orc_set_threads(ctypes.c_int(int(n)))


_u = ctypes.c_size_t
_f = ctypes.c_float


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _out_t(prec):
    return np.float64 if prec == 'd' else np.float32


def _grid(p, img_shape):
    return (_u(img_shape[3]), _u(img_shape[2]), _u(img_shape[1]),
            _f(p.dx), _f(p.dy), _f(p.dz), _f(p.cx), _f(p.cy), _f(p.cz))


def _det(p, nu, nv, nview, du=None):
    return (_u(nu), _u(nv), _u(nview), _f(p.du if du is None else du), _f(p.dv), _f(p.off_u), _f(p.off_v))


# ---------------------------------------------------------------- Siddon cone
def siddon_cone_fp(p, img, det_center, det_u, det_v, src, prec='d'):
    img = _f32(img)
    nview = det_center.shape[0]
    prj = np.zeros([img.shape[0], nview, p.nv, p.nu], _out_t(prec))
    fn = getattr(lib(), 'orc_siddon_cone_fp_' + prec)
    fn(_ptr(prj), _ptr(img), _ptr(_f32(det_center)), _ptr(_f32(det_u)), _ptr(_f32(det_v)), _ptr(_f32(src)),
       _u(img.shape[0]), *_grid(p, img.shape), *_det(p, p.nu, p.nv, nview))
    return prj


def siddon_cone_bp(p, prj, det_center, det_u, det_v, src, prec='d'):
    prj = _f32(prj)
    img = np.zeros([prj.shape[0], p.nz, p.ny, p.nx], _out_t(prec))
    fn = getattr(lib(), 'orc_siddon_cone_bp_' + prec)
    fn(_ptr(img), _ptr(prj), _ptr(_f32(det_center)), _ptr(_f32(det_u)), _ptr(_f32(det_v)), _ptr(_f32(src)),
       _u(prj.shape[0]), *_grid(p, img.shape), *_det(p, prj.shape[3], prj.shape[2], prj.shape[1]))
    return img


# ------------------------------------------------------------ DD fan/parallel
def _dd_type(is_fbp=False, branchless=False, force_fbp=False):
    return int(bool(is_fbp)) + 2 * int(bool(branchless)) + 4 * int(bool(force_fbp))


def siddon_fan_fp(p, img, angles, prec='d'):
    """Siddon fan beam (cuda/siddonFan.cu:12-46), prj += A img."""
    img = _f32(img)
    angles = _f32(angles)
    prj = np.zeros([img.shape[0], len(angles), p.nv, p.nu], _out_t(prec))
    getattr(lib(), 'orc_siddon_fan_fp_' + prec)(
        _ptr(prj), _ptr(img), _ptr(angles), _u(img.shape[0]), *_grid(p, img.shape),
        *_det(p, p.nu, p.nv, len(angles), du=np.float32(p.du / p.dsd)), _f(p.dsd), _f(p.dso))
    return prj


def siddon_fan_bp(p, prj, angles, prec='d'):
    """Siddon fan beam backprojection (cuda/siddonFan.cu:48-80, atomic scatter)."""
    prj = _f32(prj)
    angles = _f32(angles)
    img = np.zeros([prj.shape[0], p.nz, p.ny, p.nx], _out_t(prec))
    getattr(lib(), 'orc_siddon_fan_bp_' + prec)(
        _ptr(img), _ptr(prj), _ptr(angles), _u(prj.shape[0]), *_grid(p, img.shape),
        *_det(p, prj.shape[3], prj.shape[2], prj.shape[1], du=np.float32(p.du / p.dsd)), _f(p.dsd), _f(p.dso))
    return img


def dd_fan_fp(p, img, angles, branchless=False, type_projector=None, prec='d'):
    img = _f32(img)
    angles = _f32(angles)
    t = _dd_type(False, branchless) if type_projector is None else type_projector
    prj = np.zeros([img.shape[0], len(angles), p.nv, p.nu], _out_t(prec))
    getattr(lib(), 'orc_dd_fan_fp_' + prec)(
        _ptr(prj), _ptr(img), _ptr(angles), _u(img.shape[0]), *_grid(p, img.shape),
        *_det(p, p.nu, p.nv, len(angles), du=np.float32(p.du / p.dsd)), _f(p.dsd), _f(p.dso), ctypes.c_int(t))
    return prj


def dd_fan_bp(p, prj, angles, is_fbp=False, branchless=False, type_projector=None, prec='d'):
    prj = _f32(prj)
    angles = _f32(angles)
    t = _dd_type(is_fbp, branchless) if type_projector is None else type_projector
    img = np.zeros([prj.shape[0], p.nz, p.ny, p.nx], _out_t(prec))
    getattr(lib(), 'orc_dd_fan_bp_' + prec)(
        _ptr(img), _ptr(prj), _ptr(angles), _u(prj.shape[0]), *_grid(p, img.shape),
        *_det(p, prj.shape[3], prj.shape[2], prj.shape[1], du=np.float32(p.du / p.dsd)), _f(p.dsd), _f(p.dso),
        ctypes.c_int(t))
    return img


def dd_fan_fp_adjoint(p, prj, angles, type_projector=0, prec='d'):
    """img = A^T prj for the BoxInt fan-beam DD forward projector (exact transpose,
    no reference counterpart; the loops of dd_fan_fp transposed)."""
    if type_projector & 2:
        raise ValueError('the exact transpose is restated for the BoxInt mode only')
    prj = _f32(prj)
    angles = _f32(angles)
    img = np.zeros([prj.shape[0], p.nz, p.ny, p.nx], _out_t(prec))
    getattr(lib(), 'orc_dd_fan_fp_adjoint_' + prec)(
        _ptr(img), _ptr(prj), _ptr(angles), _u(prj.shape[0]), *_grid(p, img.shape),
        *_det(p, prj.shape[3], prj.shape[2], prj.shape[1], du=np.float32(p.du / p.dsd)), _f(p.dsd), _f(p.dso),
        ctypes.c_int(type_projector))
    return img


def dd_par_fp_adjoint(p, prj, angles, type_projector=0, prec='d'):
    """img = A^T prj for the BoxInt parallel-beam DD forward projector."""
    if type_projector & 2:
        raise ValueError('the exact transpose is restated for the BoxInt mode only')
    prj = _f32(prj)
    angles = _f32(angles)
    img = np.zeros([prj.shape[0], p.nz, p.ny, p.nx], _out_t(prec))
    getattr(lib(), 'orc_dd_par_fp_adjoint_' + prec)(
        _ptr(img), _ptr(prj), _ptr(angles), _u(prj.shape[0]), *_grid(p, img.shape),
        *_det(p, prj.shape[3], prj.shape[2], prj.shape[1]), ctypes.c_int(type_projector))
    return img


def dd_par_fp(p, img, angles, branchless=False, type_projector=None, prec='d'):
    img = _f32(img)
    angles = _f32(angles)
    t = _dd_type(False, branchless) if type_projector is None else type_projector
    prj = np.zeros([img.shape[0], len(angles), p.nv, p.nu], _out_t(prec))
    getattr(lib(), 'orc_dd_par_fp_' + prec)(
        _ptr(prj), _ptr(img), _ptr(angles), _u(img.shape[0]), *_grid(p, img.shape),
        *_det(p, p.nu, p.nv, len(angles)), ctypes.c_int(t))
    return prj


def dd_par_bp(p, prj, angles, is_fbp=False, branchless=False, type_projector=None, prec='d'):
    prj = _f32(prj)
    angles = _f32(angles)
    t = _dd_type(is_fbp, branchless) if type_projector is None else type_projector
    img = np.zeros([prj.shape[0], p.nz, p.ny, p.nx], _out_t(prec))
    getattr(lib(), 'orc_dd_par_bp_' + prec)(
        _ptr(img), _ptr(prj), _ptr(angles), _u(prj.shape[0]), *_grid(p, img.shape),
        *_det(p, prj.shape[3], prj.shape[2], prj.shape[1]), ctypes.c_int(t))
    return img


# ------------------------------------------------------------------- DD tomo
def dd_tomo_fp(p, img, det_center, src, branchless=False, prec='d'):
    img = _f32(img)
    nview = det_center.shape[0]
    prj = np.zeros([img.shape[0], nview, p.nv, p.nu], _out_t(prec))
    getattr(lib(), 'orc_dd_tomo_fp_' + prec)(
        _ptr(prj), _ptr(img), _ptr(_f32(det_center)), _ptr(_f32(src)), _u(img.shape[0]), *_grid(p, img.shape),
        *_det(p, p.nu, p.nv, nview), ctypes.c_int(1 if branchless else 0))
    return prj


def dd_tomo_bp(p, prj, det_center, src, branchless=False, prec='d'):
    prj = _f32(prj)
    img = np.zeros([prj.shape[0], p.nz, p.ny, p.nx], _out_t(prec))
    getattr(lib(), 'orc_dd_tomo_bp_' + prec)(
        _ptr(img), _ptr(prj), _ptr(_f32(det_center)), _ptr(_f32(src)), _u(prj.shape[0]), *_grid(p, img.shape),
        *_det(p, prj.shape[3], prj.shape[2], prj.shape[1]), ctypes.c_int(1 if branchless else 0))
    return img


# ----------------------------------------------------------------- filters
FILTERS = {'hamming': 1, 'hann': 2, 'cosine': 3}


def filter_code(name: str) -> int:
    """Filter-name map of src/ct_projector/projector/numpy/fan_equiangluar.py:39-46."""
    return FILTERS.get(name.lower(), 0)


def ramp_kernel(nu, da, filter_type='hann', equispace=False):
    """Windowed ramp kernel h_w of length 2nu-1 in FP64 (cuda/fbpMath.cu:90-268)."""
    h = np.zeros(2 * nu - 1, np.float64)
    lib().orc_ramp_kernel(_ptr(h), _u(nu), ctypes.c_double(float(da)), ctypes.c_int(filter_code(filter_type)),
                          ctypes.c_int(1 if equispace else 0))
    return h


def fan_ramp_filter(p, prj, filter_type='hann', prec='d'):
    prj = _f32(prj)
    out = np.zeros(prj.shape, _out_t(prec))
    getattr(lib(), 'orc_fan_filter_' + prec)(
        _ptr(out), _ptr(prj), _u(prj.shape[0]), _u(prj.shape[3]), _u(prj.shape[2]), _u(prj.shape[1]),
        _f(np.float32(p.du / p.dsd)), _f(p.dv), _f(p.off_u), _f(p.off_v), _f(p.dsd), _f(p.dso),
        ctypes.c_int(filter_code(filter_type)))
    return out


def par_ramp_filter(p, prj, filter_type='hann', prec='d'):
    prj = _f32(prj)
    out = np.zeros(prj.shape, _out_t(prec))
    getattr(lib(), 'orc_par_filter_' + prec)(
        _ptr(out), _ptr(prj), _u(prj.shape[0]), _u(prj.shape[3]), _u(prj.shape[2]), _u(prj.shape[1]),
        _f(p.du), _f(p.dv), _f(p.off_u), _f(p.off_v), ctypes.c_int(filter_code(filter_type)))
    return out


def fan_fbp_bp(p, prj, angles, prec='d'):
    prj = _f32(prj)
    angles = _f32(angles)
    img = np.zeros([prj.shape[0], p.nz, p.ny, p.nx], _out_t(prec))
    getattr(lib(), 'orc_fan_fbp_bp_' + prec)(
        _ptr(img), _ptr(prj), _ptr(angles), _u(prj.shape[0]), *_grid(p, img.shape),
        *_det(p, prj.shape[3], prj.shape[2], prj.shape[1], du=np.float32(p.du / p.dsd)), _f(p.dsd), _f(p.dso))
    return img


# ------------------------------------------------------------------- TV prior
def tv_sqs(img, weights, eps=1e-8):
    """Anisotropic TV surrogate of the reference's prior-cuda/TotalVariation.cu
    (Variation3DKernel :21-55, TVSQS3DKernel :57-139), in FP64 numpy.

    img [B, nz, ny, nx]; weights (wz, wy, wx) as in
    src/ct_projector/prior/numpy/recon_prior.py:166-206 (the wrapper passes
    weights[2], [1], [0] as wx, wy, wz).  Returns (s1, s2, var)."""
    x = np.asarray(img, np.float64)
    wz, wy, wx = (float(np.float32(w)) for w in weights)
    eps = float(np.float32(eps))

    def back(a, axis):
        # a - a[i-1] along axis, 0 at the first index (the lower edge)
        d = np.zeros_like(a)
        sl = [slice(None)] * a.ndim
        sl[axis] = slice(1, None)
        sp = [slice(None)] * a.ndim
        sp[axis] = slice(None, -1)
        d[tuple(sl)] = a[tuple(sl)] - a[tuple(sp)]
        return d

    def fwd(a, axis):
        # (a - a[i+1], value of b at i+1) along axis, (0, b[i]) at the last index
        d = np.zeros_like(a)
        sl = [slice(None)] * a.ndim
        sl[axis] = slice(None, -1)
        sn = [slice(None)] * a.ndim
        sn[axis] = slice(1, None)
        d[tuple(sl)] = a[tuple(sl)] - a[tuple(sn)]
        return d, tuple(sl), tuple(sn)

    dx0, dy0, dz0 = back(x, 3), back(x, 2), back(x, 1)
    var = np.sqrt(wx * dx0 ** 2 + wy * dy0 ** 2 + wz * dz0 ** 2 + eps)
    s1 = (wx * dx0 + wy * dy0 + wz * dz0) / var
    s2 = (wx + wy + wz) / var
    for axis, w in ((3, wx), (2, wy), (1, wz)):
        d1, sl, sn = fwd(x, axis)
        vn = var.copy()                  # the voxel's own variation at the upper edge
        vn[sl] = var[sn]
        s1 = s1 + w * d1 / vn
        s2 = s2 + w / vn
    return s1, s2, var


def par_pixel_bp(p, prj, angles, prec='d'):
    """Parallel-beam pixel-driven backprojection (cuda/fbpParallel.cu:200-255)."""
    prj = _f32(prj)
    angles = _f32(angles)
    img = np.zeros([prj.shape[0], p.nz, p.ny, p.nx], _out_t(prec))
    getattr(lib(), 'orc_par_pixel_bp_' + prec)(
        _ptr(img), _ptr(prj), _ptr(angles), _u(prj.shape[0]), *_grid(p, img.shape),
        *_det(p, prj.shape[3], prj.shape[2], prj.shape[1]))
    return img


# ------------------------------------------------------------ tomo FDK filter
def tomo_filter(prj, det_center, src, du, dx, dz, filter_type='hann', cutoff_x=1.0, cutoff_z=1.0):
    """Tomosynthesis FDK filter, the reference's procedure of cuda/fbpTomo.cu:67-368
    restated literally in FP64 with numpy.fft: per view the equispace Ram-Lak
    kernel for vdu = du cos(b), its DFT times the (asymmetric) window W(f) of
    GetFrequency (:47-65), the real part of the inverse DFT; rows zero-padded to
    L = 2 nu - 1, multiplied by the kernel's real DFT, inverted, read at nu - 1,
    scaled by pi / nview * vdu.  Filter codes as src/ct_projector/projector/cupy/tomo.py:190-195."""
    prj = np.asarray(prj, np.float64)
    B, nview, nv, nu = prj.shape
    L = 2 * nu - 1
    PI = float(np.float32(3.141592657))
    f32 = lambda v: float(np.float32(v))
    code = {'hamming': 1, 'hann': 2}.get(filter_type.lower(), 0)
    out = np.zeros_like(prj)
    for v in range(nview):
        ex = float(det_center[v, 0]) - float(src[v, 0])
        ez = float(det_center[v, 2]) - float(src[v, 2])
        r = np.hypot(ex, ez)
        sinb, cosb = ex / r, -ez / r
        vdu = f32(du) * cosb
        k = np.arange(L) - (nu - 1)
        ramp = np.zeros(L)
        ramp[k == 0] = 1 / (4 * vdu * vdu)
        odd = k % 2 != 0
        ramp[odd] = -1 / (PI * PI * k[odd].astype(float) ** 2 * vdu * vdu)
        i = np.arange(L)
        fu = np.where(i > (L + 1) // 2, (i - L) / L / vdu, i / L / vdu)
        if code in (1, 2):
            A, Bw = (0.54, 0.46) if code == 1 else (0.5, 0.5)
            wx = np.minimum(fu * cosb * f32(dx) * 2, f32(cutoff_x))
            wz = np.minimum(fu * sinb * f32(dz) * 2, f32(cutoff_z))
            W = (A + Bw * np.cos(PI * wx / f32(cutoff_x))) * (A + Bw * np.cos(PI * wz / f32(cutoff_z)))
        else:
            W = np.ones(L)
        h = np.fft.ifft(np.fft.fft(ramp) * W).real
        x = np.zeros([B, nv, L])
        x[..., :nu] = prj[:, v]
        y = np.fft.irfft(np.fft.rfft(x, axis=-1) * np.fft.rfft(h), n=L, axis=-1)
        out[:, v] = y[..., nu - 1:] * (PI / nview * vdu)
    return out
