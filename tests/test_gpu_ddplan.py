"""GPU: persistent distance-driven plans (ctprojector_amd.ddplan, ctpDDPlan*).

A plan launches the same kernels as the per-call entry points from tables it
built once, so its results equal theirs exactly: forward (every element
written: the output is pre-filled with NaN), the reference's backprojection
(box-integral and FBP weighting, branchless), the exact adjoint; fan and
parallel beam, batches of several multi-slice volumes, and a single 2D slice
(whose backprojection splits the views into groups).  Repeated asynchronous
calls give the same result; wrong shapes are rejected.
"""

import numpy as np
import pytest

from helpers import small_fan, small_parallel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch as T
    assert T.cuda.is_available(), 'torch sees no HIP device'
    return T


def _case(kind, B):
    from ctprojector_amd.projector import cupy as C
    from ctprojector_amd.projector.cupy import fan_equiangular, parallel
    if kind == 'fan':
        g, ang, img = small_fan(n=40, nz=3, nv=3, nview=44, nu=64, B=B, seed=21)
        mod = fan_equiangular
    else:
        g, ang, img = small_parallel(n=40, nz=2, nv=2, nview=36, nu=60, B=B, seed=22)
        mod = parallel
    return g.to(C.ct_projector()), ang, img, mod


@pytest.mark.parametrize('kind', ['fan', 'parallel'])
@pytest.mark.parametrize('branchless', [False, True])
def test_plan_equals_per_call_entry_points(torch, kind, branchless):
    from ctprojector_amd.ddplan import DDPlan
    p, ang, img, mod = _case(kind, B=2)
    dev = torch.device('cuda', 0)
    a = torch.from_numpy(ang).to(dev)
    x = torch.from_numpy(img).to(dev)
    plan = DDPlan(p, a, batch=2, fan=(kind == 'fan'), branchless=branchless)
    ref = mod.distance_driven_fp(p, x, a, branchless=branchless)
    out = torch.full(ref.shape, float('nan'), device=dev)
    got = plan.fp(x, out=out)
    assert not bool(torch.isnan(got).any())
    assert torch.equal(got, ref)
    y = ref + 0.01
    assert torch.equal(plan.adjoint(y), mod.distance_driven_fp_adjoint(p, y, a, branchless=branchless))
    assert torch.equal(plan.bp(y), mod.distance_driven_bp(p, y, a, branchless=branchless))
    fplan = DDPlan(p, a, batch=2, fan=(kind == 'fan'), branchless=branchless, is_fbp=True)
    assert torch.equal(fplan.bp(y), mod.distance_driven_bp(p, y, a, is_fbp=True, branchless=branchless))
    plan.close()
    fplan.close()


def test_plan_single_slice_view_groups_and_repeats(torch):
    from ctprojector_amd.ddplan import DDPlan
    from ctprojector_amd.projector import cupy as C
    from ctprojector_amd.projector.cupy import fan_equiangular as F
    g, ang, img = small_fan(n=96, nz=1, nv=1, nview=120, nu=140, du=2.0, seed=23)
    p = g.to(C.ct_projector())
    dev = torch.device('cuda', 0)
    a = torch.from_numpy(ang).to(dev)
    x = torch.from_numpy(img).to(dev)
    plan = DDPlan(p, a, batch=1)
    ref_fp = F.distance_driven_fp(p, x, a)
    ref_adj = F.distance_driven_fp_adjoint(p, ref_fp, a)
    ref_bp = F.distance_driven_bp(p, ref_fp, a)
    outs = []
    for _ in range(4):                       # queued without host synchronisation
        y = plan.fp(x)
        outs.append((y, plan.adjoint(y), plan.bp(y)))
    torch.cuda.synchronize()
    for y, adj, bp in outs:
        assert torch.equal(y, ref_fp)
        assert torch.equal(adj, ref_adj)
        assert torch.equal(bp, ref_bp)
    with pytest.raises(ValueError):
        plan.fp(torch.zeros([2, 1, 96, 96], device=dev))
    plan.close()


def test_autograd_ops_with_and_without_plans_agree(torch):
    from ctprojector_amd.autograd import FanDD, ParallelDD
    for kind, Op in (('fan', FanDD), ('parallel', ParallelDD)):
        p, ang, img, _ = _case(kind, B=3)
        dev = torch.device('cuda', 0)
        a = torch.from_numpy(ang).to(dev)
        x = torch.from_numpy(img).to(dev).requires_grad_(True)
        res = []
        for plan in (True, False):
            op = Op(p, a, plan=plan)
            y = op.fp(x)
            (y * y).sum().backward()
            res.append((y.detach().clone(), x.grad.clone(), op.bp(y.detach()), op.adjoint(y.detach())))
            x.grad = None
        for u, v in zip(*res):
            assert torch.equal(u, v), kind
