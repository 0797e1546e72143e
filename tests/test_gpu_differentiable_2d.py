"""GPU: per-sample-geometry differentiable distance-driven fan / parallel ops
(ctprojector_amd.projector.differentiable.circular_2d, the reference's
TensorFlow op pair DistanceDriven_2D_FP / _BP).

* per-sample angles, distances, grid and detector: each sample and channel
  equals the library's single calls with those parameters (fan beam with the
  angular pitch in the detector row, parallel beam), IR and force-FBP types;
* the registered gradients: grad of FP = the reference's BP with the same
  parameters, grad of BP = FP;
* module forms.
"""

import numpy as np
import pytest

from helpers import small_fan

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch as T
    assert T.cuda.is_available(), 'torch sees no HIP device'
    return T


def _case(torch, seed=41):
    g, ang, img = small_fan(n=32, nz=2, nv=2, nview=24, nu=48, B=2, seed=seed)
    from ctprojector_amd.projector.cupy import ct_projector
    p = g.to(ct_projector())
    angles = np.stack([ang, ang + 0.05], 0).astype(np.float32)          # per-sample angles
    x = np.stack([img, 0.3 * img], -1).astype(np.float32)               # [2, nz, ny, nx, 2]
    return p, angles, torch.from_numpy(x).cuda()


@pytest.mark.parametrize('fan', [True, False])
@pytest.mark.parametrize('typ', [0, 4])
def test_per_sample_parameters_equal_single_calls(torch, fan, typ):
    from types import SimpleNamespace
    from ctprojector_amd import _lib
    from ctprojector_amd.projector.differentiable import circular_2d as C2
    p, angles, x = _case(torch)
    kind = C2.TypeGeometry.FAN_EQUIANGULAR if fan else C2.TypeGeometry.PARALLEL
    grid = np.array([[p.dx, p.dy, p.dz, p.cx, p.cy, p.cz], [1.1, 1.1, 1.0, 0.5, -0.5, 0.0]], np.float32)
    da = p.du / p.dsd if fan else p.du
    det = np.array([[da, p.dv, 0.0, 0.0], [da * 1.05, p.dv, 0.3, 0.0]], np.float32)
    dsd = np.array([[p.dsd], [p.dsd * 1.1]], np.float32)
    dso = np.array([[p.dso], [p.dso * 0.95]], np.float32)
    a = torch.from_numpy(angles).cuda()
    y = C2.distance_driven_2d_fp(p, x, a, dsd, dso, grid, det, type_geometry=kind, type_projector=typ)
    b = C2.distance_driven_2d_bp(p, y, a, dsd, dso, grid, det, type_geometry=kind, type_projector=typ)
    from ctprojector_amd.projector.cupy import fan_equiangular, parallel
    for i in range(2):
        q = SimpleNamespace(nx=p.nx, ny=p.ny, nz=p.nz, nu=p.nu, nv=p.nv, dx=float(grid[i, 0]), dy=float(grid[i, 1]),
                            dz=float(grid[i, 2]), cx=float(grid[i, 3]), cy=float(grid[i, 4]), cz=float(grid[i, 5]),
                            dv=float(det[i, 1]), off_u=float(det[i, 2]), off_v=float(det[i, 3]),
                            dsd=float(dsd[i, 0]), dso=float(dso[i, 0]))
        q.du = float(det[i, 0]) * q.dsd if fan else float(det[i, 0])
        xi = x[i].permute(3, 0, 1, 2).contiguous()
        if typ == 0:
            mod = fan_equiangular if fan else parallel
            ref = mod.distance_driven_fp(q, xi, a[i])
            assert torch.allclose(y[i].permute(3, 0, 1, 2), ref, rtol=1e-6, atol=1e-7)
            refb = mod.distance_driven_bp(q, y[i].permute(3, 0, 1, 2).contiguous(), a[i])
            assert torch.allclose(b[i].permute(3, 0, 1, 2), refb, rtol=1e-5, atol=1e-6)
    assert bool(torch.isfinite(y).all()) and bool(torch.isfinite(b).all())


def test_registered_gradients(torch):
    from ctprojector_amd.projector.differentiable import circular_2d as C2
    p, angles, x = _case(torch, seed=42)
    a = torch.from_numpy(angles).cuda()
    kind = C2.TypeGeometry.FAN_EQUIANGULAR
    xr = x.clone().requires_grad_(True)
    y = C2.distance_driven_2d_fp(p, xr, a, type_geometry=kind)
    w = torch.rand(y.shape, device='cuda', generator=torch.Generator(device='cuda').manual_seed(5))
    (y * w).sum().backward()
    assert torch.equal(xr.grad, C2.distance_driven_2d_bp(p, w, a, type_geometry=kind))
    wr = w.clone().requires_grad_(True)
    (C2.distance_driven_2d_bp(p, wr, a, type_geometry=kind) * x).sum().backward()
    assert torch.equal(wr.grad, C2.distance_driven_2d_fp(p, x, a, type_geometry=kind))
    m = C2.DistanceDriven2DFP(p, a, type_geometry=kind)
    assert torch.equal(m(x), y.detach())
    mb = C2.DistanceDriven2DBP(p, a, type_geometry=kind)
    assert torch.equal(mb(w), xr.grad)
    with pytest.raises(ValueError):        # 23 angles for 24 views
        C2.distance_driven_2d_fp(p, x, a[:, :-1], type_geometry=kind, default_shape=[24, 2, 48])
    with pytest.raises(ValueError):
        C2.distance_driven_2d_fp(p, x, a, type_geometry=kind, type_projector=2)
