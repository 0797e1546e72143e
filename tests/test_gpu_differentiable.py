"""GPU: per-sample-geometry differentiable cone projector
(ctprojector_amd.projector.differentiable.cone, the reference's TensorFlow op
pair SiddonConeFP / SiddonConeBP).

* every sample uses its own geometry, grid and detector, every channel its
  sample's: the result equals the library's single-volume calls with those
  parameters;
* gradients: d<FP(x), y>/dx = BP(y) and d<BP(p), z>/dp = FP(z), per sample;
* output shape: positive default_shape entries win over output_shape's first
  row (kernel/projector/tensorflow/projectorBase.h:159-190);
* wrong geometry shapes are rejected.
"""

import numpy as np
import pytest

from helpers import small_cone

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch as T
    assert T.cuda.is_available(), 'torch sees no HIP device'
    return T


def _setup(torch):
    from ctprojector_amd.projector.cupy import ct_projector
    from ctprojector_amd.projector.differentiable import cone as D
    g1, geo1, img1 = small_cone(seed=31, n=20, nview=7, nu=30, nv=26)
    g2, geo2, img2 = small_cone(seed=32, n=20, nview=7, nu=30, nv=26, d=0.8, c=(0.5, -0.3, 0.2))
    p = g1.to(ct_projector())
    geometry = torch.from_numpy(np.stack([D.make_geometry(*(getattr(g, k) for k in ('det_center', 'det_u', 'det_v',
                                                                                    'src'))) for g in (geo1, geo2)]))
    grid = torch.tensor([[g.dx, g.dy, g.dz, g.cx, g.cy, g.cz] for g in (g1, g2)], dtype=torch.float32)
    det = torch.tensor([[g.du, g.dv, g.off_u, g.off_v] for g in (g1, g2)], dtype=torch.float32)
    det[1, 2] = 0.35                                    # a detector offset on the second sample only
    x = np.stack([img1[0], img2[0]], 0)                 # [2, nz, ny, nx]
    x = np.stack([x, 0.5 * x[:, ::-1]], -1).copy()      # two channels
    return D, p, (g1, g2), (geo1, geo2), geometry.cuda(), grid, det, torch.from_numpy(x.astype(np.float32)).cuda()


def test_per_sample_parameters_equal_single_calls(torch):
    from types import SimpleNamespace
    from ctprojector_amd.projector.cupy import cone
    D, p, gs, geos, geometry, grid, det, x = _setup(torch)
    y = D.siddon_fp_arbitrary(p, x, geometry, grid, det)
    assert tuple(y.shape) == (2, 7, 26, 30, 2)
    for i in range(2):
        q = SimpleNamespace(**{k: getattr(p, k) for k in ('nx', 'ny', 'nz', 'nu', 'nv')})
        for k, v in zip(('dx', 'dy', 'dz', 'cx', 'cy', 'cz'), grid[i].tolist()):
            setattr(q, k, v)
        for k, v in zip(('du', 'dv', 'off_u', 'off_v'), det[i].tolist()):
            setattr(q, k, v)
        g = [torch.from_numpy(a).cuda() for a in (geos[i].det_center, geos[i].det_u, geos[i].det_v, geos[i].src)]
        for c in range(2):
            ref = cone.siddon_fp_arbitrary(q, x[i, ..., c][None].contiguous(), *g)[0]
            assert torch.equal(y[i, ..., c], ref)
        ref_bp = cone.siddon_bp_arbitrary(q, y[i, ..., 0][None].contiguous(), *g)[0]
        bp = D.siddon_bp_arbitrary(p, y, geometry, grid, det)
        assert torch.equal(bp[i, ..., 0], ref_bp)


def test_gradients_are_the_transposes(torch):
    D, p, gs, geos, geometry, grid, det, x = _setup(torch)
    xr = x.clone().requires_grad_(True)
    y = D.siddon_fp_arbitrary(p, xr, geometry, grid, det)
    w = torch.rand(y.shape, device='cuda', generator=torch.Generator(device='cuda').manual_seed(3))
    (y * w).sum().backward()
    assert torch.equal(xr.grad, D.siddon_bp_arbitrary(p, w, geometry, grid, det))
    wr = w.clone().requires_grad_(True)
    b = D.siddon_bp_arbitrary(p, wr, geometry, grid, det)
    (b * x).sum().backward()
    assert torch.equal(wr.grad, D.siddon_fp_arbitrary(p, x, geometry, grid, det))
    # <A x, w> = <x, A^T w> per sample
    lhs = (y.detach().double() * w.double()).sum()
    rhs = (x.double() * xr.grad.double()).sum()
    assert abs(float(lhs - rhs)) < 1e-5 * abs(float(lhs))


def test_output_shape_rules_and_modules(torch):
    D, p, gs, geos, geometry, grid, det, x = _setup(torch)
    a = D.siddon_fp_arbitrary(p, x, geometry, grid, det)
    b = D.siddon_fp_arbitrary(p, x, geometry, grid, det, output_shape=[[7, 99, 99], [1, 1, 1]],
                              default_shape=[-1, 26, 30])
    assert torch.equal(a, b)
    m = D.SiddonFPArbitrary(p)
    assert torch.equal(m(x, geometry, grid, det), a)
    mb = D.SiddonBPArbitrary(p, default_shape=[-1, -1, -1])
    bp = mb(a, geometry, grid, det, output_shape=torch.tensor([[20, 20, 20]], dtype=torch.int32))
    assert torch.equal(bp, D.siddon_bp_arbitrary(p, a, geometry, grid, det))
    with pytest.raises(ValueError):
        D.siddon_fp_arbitrary(p, x, geometry[:, :-1], grid, det)
    with pytest.raises(ValueError):
        D.siddon_fp_arbitrary(p, x, geometry, grid[:, :5], det)
