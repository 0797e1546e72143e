"""GPU: the exact transpose of the distance-driven forward projector
(ctpDistance*ProjectionAdjoint) passes the dot-product test <A x, y> = <x, A^T y>
for fan and parallel beams, BoxInt and branchless modes, views in the 45-degree
blend band, multi-row detectors with z interpolation, offsets and batches."""

import numpy as np
import pytest

from helpers import small_fan, small_parallel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch as T
    assert T.cuda.is_available(), 'torch sees no HIP device'
    return T


def _dot_test(torch, mod, g, ang, img, branchless):
    from ctprojector_amd.projector import cupy as C
    p = g.to(C.ct_projector())
    dev = torch.device('cuda', 0)
    x = torch.from_numpy(img).to(dev)
    a = torch.from_numpy(ang).to(dev)
    ax = mod.distance_driven_fp(p, x, a, branchless=branchless)
    y = torch.from_numpy(np.random.default_rng(1).uniform(size=tuple(ax.shape)).astype(np.float32)).to(dev)
    aty = mod.distance_driven_fp_adjoint(p, y, a, branchless=branchless)
    lhs = float((ax.double() * y.double()).sum())
    rhs = float((x.double() * aty.double()).sum())
    rel = abs(lhs - rhs) / (float(ax.double().norm()) * float(y.double().norm()))
    assert rel < 2e-6, (lhs, rhs, rel)
    return aty


@pytest.mark.parametrize('branchless', [False, True])
@pytest.mark.parametrize('variant', ['plain', 'offsets_rows', 'batch'])
def test_fan_adjoint(torch, branchless, variant):
    from ctprojector_amd.projector.cupy import fan_equiangular as fan
    kw = {'plain': {}, 'offsets_rows': dict(nz=3, nv=5, dz=1.3, off=(0.4, 0.3), c=(2.0, -1.0, 0.6)),
          'batch': dict(B=2)}[variant]
    g, ang, img = small_fan(**kw)
    _dot_test(torch, fan, g, ang, img, branchless)


@pytest.mark.parametrize('branchless', [False, True])
def test_parallel_adjoint(torch, branchless):
    from ctprojector_amd.projector.cupy import parallel
    g, ang, img = small_parallel(nz=2, nv=3, off=(0.3, -0.2))
    _dot_test(torch, parallel, g, ang, img, branchless)


def test_adjoint_differs_from_reference_bp(torch):
    """The reference's backprojector is a different operator (so it is not the
    gradient); the exact adjoint correlates with it strongly."""
    from ctprojector_amd.projector import cupy as C
    from ctprojector_amd.projector.cupy import fan_equiangular as fan
    g, ang, img = small_fan()
    p = g.to(C.ct_projector())
    dev = torch.device('cuda', 0)
    a = torch.from_numpy(ang).to(dev)
    y = torch.from_numpy(np.random.default_rng(2).uniform(size=[1, len(ang), g.nv, g.nu]).astype(np.float32)).to(dev)
    at = fan.distance_driven_fp_adjoint(p, y, a).double()
    bp = fan.distance_driven_bp(p, y, a).double()
    cos = float((at * bp).sum() / (at.norm() * bp.norm()))
    assert cos > 0.9
    assert float((at - bp).norm() / bp.norm()) > 1e-3
