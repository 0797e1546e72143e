"""GPU: the differentiable unrolled SQS-TV network at its benchmark size
(8 fan-beam slices of 256^2 @1.4 mm, 492 views, 444 columns of 2.5716 mm,
Poisson data at 1e5 photons, 10 iterations, beta 2e-3), its loss and its
whole gradient with respect to the sinogram against the FP64 oracle network
(tests/oracle_network.py, pinned by central differences): with the exact
adjoint (the default) and with the reference's convention (the gradient of
the forward projection is the reference's backprojector)."""

import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def setup(oracle):
    import torch
    from ctprojector_amd.autograd import FanDD
    from ctprojector_amd.phantom import fan_slices_256_workload
    from ctprojector_amd.projector.cupy import ct_projector
    w, img = fan_slices_256_workload(8)
    line = oracle.dd_fan_fp(w, img, w.angles)
    counts = np.random.default_rng(1004).poisson(1e5 * np.exp(-line))
    prj = (-np.log(np.maximum(counts, 1) / 1e5)).astype(np.float32)
    wgt = (counts / 1e5).astype(np.float32)
    p = ct_projector()
    for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v',
              'dsd', 'dso'):
        setattr(p, k, getattr(w, k))
    dev = torch.device('cuda', 0)
    ang_d = torch.from_numpy(w.angles).to(dev)
    op = FanDD(p, ang_d)
    with torch.no_grad():
        v = torch.from_numpy(np.random.default_rng(7).uniform(size=img.shape).astype(np.float32)).to(dev)
        for _ in range(10):
            v = op.adjoint(op.fp(v))
            v = v / v.norm()
        pn = float(op.fp(v).norm())
        wd = torch.from_numpy(wgt).to(dev)
        norm_img = (op.adjoint(wd * op.fp(torch.ones_like(v))) / pn / pn).cpu().numpy()
    return dict(w=w, img=img, prj=prj, wgt=wgt, p=p, ang_d=ang_d, pn=pn, norm_img=norm_img)


@pytest.mark.parametrize('convention', ['exact', 'reference'])
def test_unrolled_gradient_vs_fp64_network(oracle, setup, convention):
    import torch
    from ctprojector_amd.autograd import FanDD, sqs_tv_unrolled
    from oracle_network import make_network
    S = setup
    dev = torch.device('cuda', 0)
    beta, niter, eps = 2e-3, 10, 1e-8
    op = FanDD(S['p'], S['ang_d'], adjoint=convention)
    print('[unrolled] %s convention: GPU network' % convention, flush=True)
    q = torch.from_numpy(S['prj']).to(dev).requires_grad_(True)
    x = sqs_tv_unrolled(op, q, torch.zeros(S['norm_img'].shape, device=dev), torch.from_numpy(S['norm_img']).to(dev),
                        S['pn'], beta, niter, eps=eps, weight=torch.from_numpy(S['wgt']).to(dev))
    loss = ((x - torch.from_numpy(S['img']).to(dev)) ** 2).sum()
    loss.backward()
    g_gpu = q.grad.cpu().numpy()
    print('[unrolled] FP64 oracle network', flush=True)
    run = make_network(oracle, S['w'], S['w'].angles, S['norm_img'], S['pn'], beta, niter, S['wgt'], eps, convention)
    p64 = torch.tensor(S['prj'].astype(np.float64), requires_grad=True)
    l64 = ((run(p64) - torch.as_tensor(S['img'].astype(np.float64))) ** 2).sum()
    l64.backward()
    g64 = p64.grad.numpy()
    el = abs(float(loss) - float(l64)) / float(l64)
    eg = rel_l2(g_gpu, g64)
    print('[unrolled] %s: loss rel %.2e, gradient rel-L2 %.2e' % (convention, el, eg), flush=True)
    assert el < 1e-4
    assert eg < 1e-3
