"""GPU: the device-tensor front end (``cupy*`` entry points) agrees with the
host-array front end (``c*`` entry points) and with the FP64 oracle, keeps the
reference's shape rules, and runs on torch tensors without host round trips."""

import numpy as np
import pytest

from helpers import rel_l2, small_cone, small_fan, small_parallel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch as T
    # a GPU test: the device must be visible to torch in this process
    assert T.cuda.is_available(), 'torch sees no HIP device'
    return T


def _mk(mod, g):
    p = mod.ct_projector()
    return g.to(p)


def test_cone_torch_vs_numpy(torch, oracle):
    from ctprojector_amd.projector import cupy as C, numpy as N
    from ctprojector_amd.projector.cupy import cone as cc
    from ctprojector_amd.projector.numpy import cone as nc
    g, geo, img = small_cone(seed=3)
    dev = torch.device('cuda', 0)
    args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
    dargs = [torch.from_numpy(a).to(dev) for a in args]
    fp_t = cc.siddon_fp_arbitrary(_mk(C, g), torch.from_numpy(img).to(dev), *dargs)
    fp_n = nc.siddon_fp_arbitrary(_mk(N, g), img, *args)
    assert fp_t.shape == (1, 12, g.nv, g.nu) and fp_t.is_cuda
    assert np.array_equal(fp_t.cpu().numpy(), fp_n)
    bp_t = cc.siddon_bp_arbitrary(_mk(C, g), fp_t, *dargs)
    bp_n = nc.siddon_bp_arbitrary(_mk(N, g), fp_n, *args)
    assert bp_t.shape == (1, g.nz, g.ny, g.nx)
    assert np.array_equal(bp_t.cpu().numpy(), bp_n)
    assert rel_l2(bp_n, oracle.siddon_cone_bp(g, fp_n, *args)) < 1e-4


def test_fan_torch_vs_numpy(torch):
    from ctprojector_amd.projector import cupy as C, numpy as N
    from ctprojector_amd.projector.cupy import fan_equiangular as cf
    from ctprojector_amd.projector.numpy import fan_equiangluar as nf
    g, ang, img = small_fan(B=2)
    dev = torch.device('cuda', 0)
    img_d, ang_d = torch.from_numpy(img).to(dev), torch.from_numpy(ang).to(dev)
    for br in (False, True):
        fp_t = cf.distance_driven_fp(_mk(C, g), img_d, ang_d, branchless=br)
        fp_n = nf.distance_driven_fp(_mk(N, g), img, ang, branchless=br)
        assert np.array_equal(fp_t.cpu().numpy(), fp_n)
        for fbp in (False, True):
            bp_t = cf.distance_driven_bp(_mk(C, g), fp_t, ang_d, is_fbp=fbp, branchless=br)
            bp_n = nf.distance_driven_bp(_mk(N, g), fp_n, ang, is_fbp=fbp, branchless=br)
            assert np.array_equal(bp_t.cpu().numpy(), bp_n)
    fl_t = cf.ramp_filter(_mk(C, g), fp_t, 'hann')
    fl_n = nf.ramp_filter(_mk(N, g), fp_n, 'hann')
    assert np.array_equal(fl_t.cpu().numpy(), fl_n)
    r_t = cf.fbp_bp(_mk(C, g), fl_t, ang_d)
    r_n = nf.fbp_bp(_mk(N, g), fl_n, ang)
    assert np.array_equal(r_t.cpu().numpy(), r_n)


def test_parallel_torch_vs_numpy(torch):
    from ctprojector_amd.projector import cupy as C, numpy as N
    from ctprojector_amd.projector.cupy import parallel as cp_
    from ctprojector_amd.projector.numpy import parallel as np_
    g, ang, img = small_parallel()
    dev = torch.device('cuda', 0)
    fp_t = cp_.distance_driven_fp(_mk(C, g), torch.from_numpy(img).to(dev), torch.from_numpy(ang).to(dev))
    fp_n = np_.distance_driven_fp(_mk(N, g), img, ang)
    assert np.array_equal(fp_t.cpu().numpy(), fp_n)
    fl_t = cp_.ramp_filter(_mk(C, g), fp_t, 'hamming')
    assert np.array_equal(fl_t.cpu().numpy(), np_.ramp_filter(_mk(N, g), fp_n, 'hamming'))


def test_projector_binding_and_norm(torch):
    """set_projector / fp / bp binding and the seeded power method (ct_projector.py:72-189)."""
    from ctprojector_amd.projector import cupy as C
    from ctprojector_amd.projector.cupy import fan_equiangular as cf
    g, ang, img = small_fan(nz=1, nv=1)
    p = _mk(C, g)
    dev = torch.device('cuda', 0)
    ang_d = torch.from_numpy(ang).to(dev)
    p.set_projector(cf.distance_driven_fp, angles=ang_d)
    p.set_backprojector(cf.distance_driven_bp, angles=ang_d, is_fbp=False)
    n1 = p.calc_projector_norm(niter=5, verbose=False)
    n2 = p.calc_projector_norm(niter=5, verbose=False)
    assert n1 == n2 and n1 > 0
    ni = p.calc_norm_img()
    assert ni.shape == (1, 1, g.ny, g.nx) and float(ni.min()) >= 0


@pytest.mark.parametrize('filt', ['hann', 'hamming', 'ramp'])
def test_tomo_filter_vs_oracle(torch, oracle, filt):
    """The HIP tomosynthesis filter (per-view real kernels with the symmetrised
    window, direct convolution) against the reference's FFT procedure."""
    from ctprojector_amd.projector.cupy import tomo
    from helpers import Geo
    rng = np.random.default_rng(7)
    prj = rng.uniform(size=[2, 4, 5, 48]).astype(np.float32)
    t = np.array([-0.45, -0.1, 0.2, 0.4])
    src = np.stack([600 * np.sin(t), 3 + 0 * t, 600 * np.cos(t) + 10], 1).astype(np.float32)
    dc = np.stack([rng.uniform(-2, 2, 4), 0 * t, -20 + 0 * t], 1).astype(np.float32)
    g = Geo(du=0.9, dx=0.7, dz=1.5)
    dev = torch.device('cuda', 0)
    got = tomo.filter_fbp(g, torch.from_numpy(prj).to(dev), torch.from_numpy(dc).to(dev),
                          torch.from_numpy(src).to(dev), filt, 0.8, 0.6).cpu().numpy()
    ref = oracle.tomo_filter(prj, dc, src, 0.9, 0.7, 1.5, filt, 0.8, 0.6)
    assert rel_l2(got, ref) < 1e-4


def test_projector_norm_hip_dd_adjoint_pair(torch, oracle):
    """The power method over the HIP fan DD forward projector and its exact
    adjoint converges to the largest singular value of the oracle's dense
    system matrix of the same forward projector."""
    from ctprojector_amd.projector import cupy as C
    from ctprojector_amd.projector.cupy import fan_equiangular as cf
    from helpers import Geo
    g = Geo(nx=10, ny=9, nz=1, dx=1.0, dy=1.2, nu=24, nv=1, du=1.5, dv=1.0, dsd=200.0, dso=120.0, off_u=0.3)
    ang = (np.arange(15) * 2 * np.pi / 15).astype(np.float32)
    A = np.zeros([15 * 24, 90])
    for j in range(90):
        e = np.zeros([1, 1, 9, 10], np.float32)
        e.flat[j] = 1
        A[:, j] = oracle.dd_fan_fp(g, e, ang).ravel()
    smax = np.linalg.svd(A, compute_uv=False)[0]
    p = _mk(C, g)
    ang_d = torch.from_numpy(ang).to(torch.device('cuda', 0))
    p.set_projector(cf.distance_driven_fp, angles=ang_d)
    p.set_backprojector(cf.distance_driven_fp_adjoint, angles=ang_d)
    n = p.calc_projector_norm(niter=200, verbose=False)
    assert abs(n - smax) < 1e-4 * smax, (n, smax)
