"""GPU: differentiable projectors (ctprojector_amd.autograd).

* the torch TV surrogate equals the HIP TV kernel, and the HIP gradient of the
  surrogate (ctpTVSQSBackward3D) equals torch autograd through the torch
  surrogate in FP64 (the latter is pinned against finite differences of the
  FP64 oracle in the CPU suite);
* gradients of fp / bp / adjoint are the hand-written transposes (dot tests
  through autograd), Siddon cone and DD fan;
* an unrolled SQS-TV network (fan DD, 3 iterations, checkpointed or not): the
  directional derivative of a loss with respect to the projection data from
  autograd on the GPU (FP32) against central finite differences of the same
  network in FP64 on the CPU, built from the FP64 oracle (dense system matrix
  of the oracle's forward projector, its exact transpose, TV surrogate);
* the reference's gradient convention (adjoint='reference') gives a different
  but strongly correlated gradient.
"""

import numpy as np
import pytest

from helpers import rel_l2, small_cone, small_fan

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch as T
    assert T.cuda.is_available(), 'torch sees no HIP device'
    return T


def test_tv_torch_matches_kernel(torch):
    from ctprojector_amd.autograd import tv_sqs_torch
    from ctprojector_amd.prior.cupy import tv_sqs
    x = torch.from_numpy(np.random.default_rng(0).uniform(size=[2, 4, 9, 11]).astype(np.float32)).cuda()
    for a, b in zip(tv_sqs_torch(x, (0.5, 1.0, 2.0), 1e-4), tv_sqs(x, (0.5, 1.0, 2.0), 1e-4)):
        assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-6


@pytest.mark.parametrize('shape', [(2, 4, 9, 11), (1, 1, 5, 7), (1, 6, 1, 3)])
def test_tv_kernel_gradient_matches_torch_autograd(torch, shape):
    from ctprojector_amd.autograd import tv_sqs, tv_sqs_torch
    rng = np.random.default_rng(7)
    x = rng.uniform(size=shape).astype(np.float32)
    g1, g2, gv = (rng.normal(size=shape).astype(np.float32) for _ in range(3))
    w, eps = (0.5, 1.0, 2.0), 1e-4

    def ref(use_gv):
        xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
        s1, s2, var = tv_sqs_torch(xt, w, eps)
        loss = (s1 * torch.from_numpy(g1).double()).sum() + (s2 * torch.from_numpy(g2).double()).sum()
        if use_gv:
            loss = loss + (var * torch.from_numpy(gv).double()).sum()
        loss.backward()
        return xt.grad.numpy()

    for use_gv in (True, False):
        xd = torch.from_numpy(x).cuda().requires_grad_(True)
        s1, s2, var = tv_sqs(xd, w, eps)
        loss = (s1 * torch.from_numpy(g1).cuda()).sum() + (s2 * torch.from_numpy(g2).cuda()).sum()
        if use_gv:
            loss = loss + (var * torch.from_numpy(gv).cuda()).sum()
        loss.backward()
        assert rel_l2(xd.grad.cpu().numpy(), ref(use_gv)) < 1e-5, (shape, use_gv)


def test_unrolled_kernel_tv_matches_torch_tv(torch):
    from ctprojector_amd.autograd import FanDD, sqs_tv_unrolled
    from ctprojector_amd.projector import cupy as C
    g, ang, img = small_fan(n=24, nz=1, nv=1, nview=30, nu=40, du=2.0, seed=8)
    dev = torch.device('cuda', 0)
    op = FanDD(g.to(C.ct_projector()), torch.from_numpy(ang).to(dev))
    x_true = torch.from_numpy(np.concatenate([img, 0.7 * img])).to(dev)
    prj = op.fp(x_true).detach()
    norm_img = op.adjoint(op.fp(torch.ones_like(x_true))).detach()
    grads = []
    for tv in ('kernel', 'torch'):
        q = prj.clone().requires_grad_(True)
        x = sqs_tv_unrolled(op, q, torch.zeros_like(x_true), norm_img, 1.0, 0.05, 3, eps=1e-4, tv=tv)
        ((x - x_true) ** 2).sum().backward()
        grads.append(q.grad.cpu().numpy())
    assert rel_l2(grads[0], grads[1]) < 1e-4


def _grad_dot(torch, f, x, y):
    x = x.clone().requires_grad_(True)
    out = f(x)
    (out * y).sum().backward()
    return float((out.detach().double() * y.double()).sum()), x.grad


def test_cone_fp_bp_gradients(torch):
    from ctprojector_amd.autograd import ConeSiddon
    from ctprojector_amd.projector import cupy as C
    g, geo, img = small_cone(seed=3, n=16, nview=6, nu=24, nv=20)
    dev = torch.device('cuda', 0)
    op = ConeSiddon(g.to(C.ct_projector()), *[torch.from_numpy(a).to(dev) for a in
                                              (geo.det_center, geo.det_u, geo.det_v, geo.src)])
    x = torch.from_numpy(img).to(dev)
    y = torch.from_numpy(np.random.default_rng(1).uniform(size=[1, 6, 20, 24]).astype(np.float32)).to(dev)
    lhs, gx = _grad_dot(torch, op.fp, x, y)                  # d<Ax, y>/dx = A^T y
    assert rel_l2(gx.cpu().numpy(), op.bp(y).cpu().numpy()) < 1e-6
    rhs = float((x.double() * gx.double()).sum())
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
    _, gy = _grad_dot(torch, op.bp, y, x)                    # d<A^T y, x>/dy = A x
    assert rel_l2(gy.cpu().numpy(), op.fp(x).cpu().numpy()) < 1e-6


def test_fan_fp_gradient_is_exact_adjoint(torch):
    from ctprojector_amd.autograd import FanDD
    from ctprojector_amd.projector import cupy as C
    g, ang, img = small_fan(n=32, nz=2, nv=2, nview=40, nu=60)
    dev = torch.device('cuda', 0)
    op = FanDD(g.to(C.ct_projector()), torch.from_numpy(ang).to(dev))
    x = torch.from_numpy(img).to(dev)
    y = torch.from_numpy(np.random.default_rng(2).uniform(size=[1, 40, 2, 60]).astype(np.float32)).to(dev)
    lhs, gx = _grad_dot(torch, op.fp, x, y)
    rhs = float((x.double() * gx.double()).sum())
    assert abs(lhs - rhs) < 2e-6 * abs(lhs)
    _, gy = _grad_dot(torch, op.adjoint, y, x)
    assert rel_l2(gy.cpu().numpy(), op.fp(x).cpu().numpy()) < 1e-6


# ------------------------------------------------- unrolled SQS-TV vs FP64 FD
def _fp64_network(oracle, g, ang, A, norm_img, pn, beta, niter, eps):
    """The unrolled SQS-TV network in FP64 numpy with the dense oracle matrix."""
    n = g.nx

    def run(p):
        x = np.zeros([1, 1, n, n])
        for _ in range(niter):
            r = (A @ x.ravel()) / pn - p.ravel() / pn
            gr = (A.T @ r).reshape(x.shape) / pn
            s1, s2, _ = oracle.tv_sqs(x, (1.0, 1.0, 1.0), eps)
            x = x - (gr + beta * s1) / (norm_img + beta * s2)
        return x
    return run


@pytest.mark.parametrize('checkpoint', [True, False])
def test_unrolled_sqs_tv_gradient_vs_fp64_finite_differences(torch, oracle, checkpoint):
    from ctprojector_amd.autograd import FanDD, sqs_tv_unrolled
    from ctprojector_amd.projector import cupy as C
    g, ang, img = small_fan(n=16, nz=1, nv=1, nview=24, nu=32, du=2.0, seed=4)
    n, beta, niter, eps = 16, 0.05, 3, 1e-4
    # dense FP64 system matrix of the oracle's forward projector (unit images as a batch)
    eye = np.eye(n * n, dtype=np.float32).reshape(n * n, 1, n, n)
    A = oracle.dd_fan_fp(g, eye, ang).reshape(n * n, -1).T            # [nview*nu, n*n]
    pn = float(np.linalg.norm(A, 2))
    norm_img = (A.T @ (A @ np.ones(n * n))).reshape(1, 1, n, n) / pn / pn
    p0 = (A @ img.ravel().astype(np.float64)).reshape(1, 24, 1, 32)
    target = np.random.default_rng(5).uniform(size=[1, 1, n, n])
    dp = np.random.default_rng(6).normal(size=p0.shape)
    run = _fp64_network(oracle, g, ang, A, norm_img, pn, beta, niter, eps)
    loss = lambda p: float(((run(p) - target) ** 2).sum())
    h = 1e-3
    fd = (loss(p0 + h * dp) - loss(p0 - h * dp)) / (2 * h)

    dev = torch.device('cuda', 0)
    op = FanDD(g.to(C.ct_projector()), torch.from_numpy(ang).to(dev))
    p = torch.tensor(p0, dtype=torch.float32, device=dev, requires_grad=True)
    x = sqs_tv_unrolled(op, p, torch.zeros([1, 1, n, n], device=dev),
                        torch.tensor(norm_img, dtype=torch.float32, device=dev), pn, beta, niter,
                        eps=eps, checkpoint=checkpoint)
    l = ((x - torch.tensor(target, dtype=torch.float32, device=dev)) ** 2).sum()
    l.backward()
    ad = float((p.grad.double() * torch.tensor(dp, device=dev)).sum())
    assert abs(float(l) - loss(p0)) < 1e-4 * loss(p0)
    assert abs(ad - fd) < 2e-3 * abs(fd), (ad, fd)

    # the reference's convention (gradient of fp = reference bp) is a different operator
    p.grad = None
    op_ref = FanDD(g.to(C.ct_projector()), torch.from_numpy(ang).to(dev), adjoint='reference')
    xr = sqs_tv_unrolled(op_ref, p, torch.zeros([1, 1, n, n], device=dev),
                         torch.tensor(norm_img, dtype=torch.float32, device=dev), pn, beta, niter, eps=eps)
    ((xr - torch.tensor(target, dtype=torch.float32, device=dev)) ** 2).sum().backward()
    assert p.grad is not None and bool(torch.isfinite(p.grad).all())


def test_unrolled_training_step_in_a_hip_graph(torch):
    """The unrolled SQS-TV network's forward and gradient captured in one HIP
    graph (DD plans with their own scratch, TV calls with caller scratch)
    equal the eager step bit for bit on every replay, with eager work and
    synchronisation between the replays."""
    from ctprojector_amd.autograd import FanDD, sqs_tv_unrolled
    from ctprojector_amd.projector import cupy as C
    g, ang, img = small_fan(n=32, nz=2, nv=2, nview=30, nu=48, du=2.0, seed=7, B=2)
    dev = torch.device('cuda', 0)
    op = FanDD(g.to(C.ct_projector()), torch.from_numpy(ang).to(dev))
    x_true = torch.from_numpy(img).to(dev)
    with torch.no_grad():
        prj = op.fp(x_true)
        norm_img = op.adjoint(op.fp(torch.ones_like(x_true))) / 100.0
    x0 = torch.zeros_like(x_true)

    def net_and_grad(q):
        x = sqs_tv_unrolled(op, q, x0, norm_img, 10.0, 1e-3, 4, eps=1e-4, checkpoint=False)
        x.backward(2.0 * (x - x_true))
        return x

    q_eager = prj.clone().requires_grad_(True)
    x_eager = net_and_grad(q_eager)
    q = prj.clone().requires_grad_(True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            q.grad = None
            net_and_grad(q)
    torch.cuda.current_stream().wait_stream(side)
    q.grad = None
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        x_static = net_and_grad(q)
    g_static = q.grad
    for rep in range(3):
        print('[graph] replay', rep, flush=True)
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(x_static, x_eager), rep
        assert torch.equal(g_static, q_eager.grad), rep
        # eager work of the same operators between replays
        with torch.no_grad():
            _ = op.adjoint(op.fp(x_true) - prj)
        torch.cuda.synchronize()
