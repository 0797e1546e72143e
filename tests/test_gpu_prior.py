"""GPU: the TV prior (cupyTVSQS3D / cTVSQS3D) against the FP64 oracle, the
fused SQS-TV step against the step composed from tv_sqs, and an SQS-TV loop
through the reference-style ``sqs_one_step`` driver."""

import numpy as np
import pytest

from helpers import rel_l2, small_fan

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope='module')
def torch():
    import torch as T
    assert T.cuda.is_available(), 'torch sees no HIP device'
    return T


def _img(seed=0, shape=(2, 5, 17, 23)):
    rng = np.random.default_rng(seed)
    return (rng.uniform(size=shape) * (rng.uniform(size=shape) > 0.3)).astype(np.float32)


def test_tv_sqs_device_and_host_vs_oracle(torch, oracle):
    from ctprojector_amd.prior import cupy as PC, numpy as PN
    x = _img(1)
    w = (0.5, 1.0, 2.0)
    ref = oracle.tv_sqs(x, w, 1e-4)
    dev = [t.cpu().numpy() for t in PC.tv_sqs(torch.from_numpy(x).cuda(), w, 1e-4)]
    host = PN.tv_sqs(x, w, 1e-4)
    for a, b, r in zip(dev, host, ref):
        assert rel_l2(a, r) < TOL
        assert np.array_equal(a, b)


def test_fused_update_equals_composed_step(torch):
    from ctprojector_amd.prior.cupy import tv_sqs, tv_sqs_update
    rng = np.random.default_rng(2)
    x = torch.from_numpy(_img(3)).cuda()
    bp = torch.from_numpy(rng.normal(size=x.shape).astype(np.float32)).cuda()
    norm = torch.from_numpy(rng.uniform(1, 2, size=x.shape).astype(np.float32)).cuda()
    w = (1.0, 1.0, 1.0)
    s1, s2, var = tv_sqs(x, w, 1e-6)
    want = x - (0.7 * bp + 0.3 * s1) / (norm + 0.3 * s2)
    got, v = tv_sqs_update(x, bp, norm, 0.3, w, 1e-6, acc_fac=0.7, return_var=True)
    assert rel_l2(got.cpu().numpy(), want.cpu().numpy()) < 1e-6
    assert rel_l2(v.cpu().numpy(), var.cpu().numpy()) < 1e-7


def test_sqs_tv_loop_decreases_loss(torch):
    """Fan DD projector bound to a ct_projector, reference-style sqs_one_step with
    tv_sqs, against the fused sqs_tv_one_step."""
    from ctprojector_amd.projector import cupy as C
    from ctprojector_amd.projector.cupy import fan_equiangular as fan
    from ctprojector_amd.prior.cupy import tv_sqs
    from ctprojector_amd.recon.cupy.sqs_algorithms import sqs_one_step, sqs_tv_one_step
    g, ang, img = small_fan(n=32, nz=1, nv=1, nview=60, nu=56)
    p = g.to(C.ct_projector())
    dev = torch.device('cuda', 0)
    ang_d = torch.from_numpy(ang).to(dev)
    p.set_projector(fan.distance_driven_fp, angles=ang_d)
    p.set_backprojector(fan.distance_driven_bp, angles=ang_d)
    truth = torch.from_numpy(img).to(dev)
    prj = p.fp(truth)
    pnorm = p.calc_projector_norm(niter=5, verbose=False)
    norm_img = p.calc_norm_img() / pnorm / pnorm
    x1 = torch.zeros_like(truth)
    x2 = torch.zeros_like(truth)
    losses = []
    for _ in range(6):
        x1, dl, pl = sqs_one_step(p, x1, prj, norm_img, pnorm, 1e-3, tv_sqs,
                                  {'weights': (1.0, 1.0, 1.0), 'eps': 1e-8}, return_loss=True)
        x2 = sqs_tv_one_step(p, x2, prj, norm_img, pnorm, 1e-3, (1.0, 1.0, 1.0), 1e-8)
        losses.append(float(dl) + 1e-3 * float(pl))
    assert losses[-1] < losses[0]
    assert rel_l2(x1.cpu().numpy(), x2.cpu().numpy()) < 1e-5


def test_async_entry_points_equal_waiting_ones(torch):
    """ctpTVSQS3DAsync / ctpTVSQSBackward3DAsync (queued on a stream, no wait;
    the backward with caller scratch through the torch wrapper and with the
    library's workspace) give the bits of cupyTVSQS3D / ctpTVSQSBackward3D, on
    a side stream too."""
    from ctypes import c_float, c_ulong, c_void_p
    from ctprojector_amd import _lib
    from ctprojector_amd.prior.cupy import tv_sqs_backward
    x = torch.from_numpy(_img(3)).cuda()
    B, nz, ny, nx = x.shape
    w = (0.5, 1.0, 2.0)
    dims = (c_ulong(B), c_ulong(nx), c_ulong(ny), c_ulong(nz), c_float(1e-4))
    ws = (c_float(w[2]), c_float(w[1]), c_float(w[0]))
    p = lambda t: c_void_p(t.data_ptr()) if t is not None else None
    sync = [torch.zeros_like(x) for _ in range(3)]
    assert _lib.module().cupyTVSQS3D(*[p(t) for t in sync], p(x), *ws, *dims) == 0
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        asy = [torch.empty_like(x) for _ in range(3)]
        assert _lib.module().ctpTVSQS3DAsync(*[p(t) for t in asy], p(x), *ws, *dims,
                                             c_void_p(side.cuda_stream)) == 0
    torch.cuda.current_stream().wait_stream(side)
    for a, b in zip(sync, asy):
        assert torch.equal(a, b)
    g1, g2, gv = (torch.rand_like(x) for _ in range(3))
    for gvar in (None, gv):
        d_sync = torch.empty_like(x)
        assert _lib.module().ctpTVSQSBackward3D(p(d_sync), p(x), p(g1), p(g2), p(gvar), *ws, *dims) == 0
        d_async = tv_sqs_backward(x, g1, g2, w, 1e-4, gvar=gvar)
        assert torch.equal(d_sync, d_async)
        d_ws = torch.empty_like(x)
        assert _lib.module().ctpTVSQSBackward3DAsync(p(d_ws), p(x), p(g1), p(g2), p(gvar), None, *ws, *dims,
                                                     c_void_p(torch.cuda.current_stream().cuda_stream)) == 0
        assert torch.equal(d_sync, d_ws)
