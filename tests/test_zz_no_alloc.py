"""The allocation-free distance-driven fan forward projection
(ctpDistanceDrivenFanProjectionNoAlloc: scratch in a fixed device array)
against the allocating entry point cupyDistanceDrivenFanProjection, through
the C ABI on device tensors: two calls in a row, the second smaller, must give
the same projections bit for bit (same kernels, same arithmetic); a call that
needs more than the array holds is refused with an error and leaves the output
untouched."""

from ctypes import c_float, c_int, c_ulong, c_void_p

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 80 << 20


def _args(prj, img, ang, p):
    B, nz, ny, nx = img.shape
    return (c_void_p(prj.data_ptr()), c_void_p(img.data_ptr()), c_void_p(ang.data_ptr()),
            c_ulong(B), c_ulong(nx), c_ulong(ny), c_ulong(nz), c_float(p['dx']), c_float(p['dx']), c_float(1.0),
            c_float(p['cx']), c_float(p['cy']), c_float(0.0), c_ulong(prj.shape[3]), c_ulong(prj.shape[2]),
            c_ulong(prj.shape[1]), c_float(p['du'] / p['dsd']), c_float(1.0), c_float(p['off_u']), c_float(0.1),
            c_float(p['dsd']), c_float(p['dso']), c_int(p['type']))


def _case(torch, seed, B, nz, n, nview, nu, typ):
    rng = np.random.default_rng(seed)
    dev = torch.device('cuda', 0)
    img = torch.from_numpy(rng.uniform(size=[B, nz, n, n]).astype(np.float32)).to(dev)
    ang = torch.from_numpy((np.arange(nview) * 2 * np.pi / nview + 0.01 * seed).astype(np.float32)).to(dev)
    p = dict(dx=0.9 * 256 / n, cx=1.5, cy=-0.7, du=1.3, dsd=1085.6, dso=595.0, off_u=0.25, type=typ)
    return img, ang, p, (B, nview, nz, nu)


def test_no_alloc_matches_allocating_entry_point():
    import torch
    from ctprojector_amd import _lib
    lib = _lib.module()
    dev = torch.device('cuda', 0)
    for seed, shape in ((1, (1, 32, 512, 180, 888, 0)), (2, (2, 5, 200, 96, 300, 2))):
        img, ang, p, oshape = _case(torch, seed, *shape)
        a = torch.zeros(oshape, device=dev)
        b = torch.full(oshape, 7.0, device=dev)        # the FP overwrites every cell
        assert lib.cupyDistanceDrivenFanProjection(*_args(a, img, ang, p)) == 0
        assert lib.ctpDistanceDrivenFanProjectionNoAlloc(*_args(b, img, ang, p)) == 0
        torch.cuda.synchronize()
        assert torch.equal(a, b), float((a - b).abs().max())
        assert float(a.abs().max()) > 0


def test_no_alloc_refuses_oversized_call():
    import torch
    from ctprojector_amd import _lib
    lib = _lib.module()
    dev = torch.device('cuda', 0)
    # 64 slices of 512^2: two channel-last copies need 128 MiB > 80 MiB
    img = torch.ones([1, 64, 512, 512], device=dev)
    ang = torch.zeros([4], device=dev)
    prj = torch.full([1, 4, 64, 64], 3.0, device=dev)
    p = dict(dx=0.7, cx=0.0, cy=0.0, du=1.3, dsd=1085.6, dso=595.0, off_u=0.0, type=0)
    err = lib.ctpDistanceDrivenFanProjectionNoAlloc(*_args(prj, img, ang, p))
    assert err != 0
    assert 'allocation-free' in _lib.last_error_message()
    assert bool((prj == 3.0).all())
    # the next call still works
    img2, ang2, p2, oshape = _case(torch, 3, 1, 8, 128, 40, 200, 0)
    a = torch.zeros(oshape, device=dev)
    b = torch.zeros(oshape, device=dev)
    assert lib.cupyDistanceDrivenFanProjection(*_args(a, img2, ang2, p2)) == 0
    assert lib.ctpDistanceDrivenFanProjectionNoAlloc(*_args(b, img2, ang2, p2)) == 0
    torch.cuda.synchronize()
    assert torch.equal(a, b)
