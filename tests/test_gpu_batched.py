"""GPU: batched cone-beam plans (ctpConePlan*, ctprojector_amd.batched).

* uneven jobs (different volume sizes, voxel sizes, detectors, trajectories)
  in one plan give the same projections as the single-volume entry points and
  the FP64 oracle; backprojections likewise;
* near-axis-parallel rays go through the plan's exact ray path;
* the autotuner picks a valid configuration and the results do not change;
* persistent blocks (work counters) and one hardware-dispatched block per tile
  give the same results, and so do the backward's single launch over every
  job and its launches per window shape;
* HIP-graph capture: forward and backward graphs replayed several times on new
  inputs, with eager work in between, each replay checked against eager runs.
"""

import numpy as np
import pytest

from helpers import Geo, rel_l2, small_cone

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch as T
    assert T.cuda.is_available(), 'torch sees no HIP device'
    return T


def _jobs():
    a = small_cone(seed=21, n=20, nview=8, nu=36, nv=30, d=1.0)
    b = small_cone(seed=22, n=28, nview=10, nu=40, nv=34, d=0.75, c=(1.0, -0.5, 0.3))
    g, geo, _ = small_cone(seed=23, n=24, nview=6, nu=32, nv=28, d=2.0)
    g.nz = 12                                    # a flat volume of 2 mm voxels: a larger candidate window
    from ctprojector_amd.phantom import ellipsoid_phantom
    c = (g, geo, ellipsoid_phantom(12, 24, 24, 2.0, 2.0, 2.0, seed=24)[None])
    return [a, b, c]


def _batch(torch, cases, autotune=False):
    from ctprojector_amd.batched import ConeBatch, ConeJob
    from ctprojector_amd.projector.cupy import ct_projector
    return ConeBatch([ConeJob(g.to(ct_projector()), geo) for g, geo, _ in cases], autotune=autotune)


def test_batch_matches_single_and_oracle(torch, oracle):
    from ctprojector_amd.projector.numpy import cone
    from ctprojector_amd.projector.numpy.ct_projector import ct_projector
    cases = _jobs()
    bt = _batch(torch, cases)
    rng = np.random.default_rng(0)
    for j, (g, geo, img) in enumerate(cases):
        bt.img[j].copy_(torch.from_numpy(img[0]))
    bt.forward()
    torch.cuda.synchronize()
    ys = []
    for j, (g, geo, img) in enumerate(cases):
        args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
        single = cone.siddon_fp_arbitrary(g.to(ct_projector()), img, *args)[0]
        assert rel_l2(bt.prj[j].cpu().numpy(), single) < 1e-7
        assert rel_l2(bt.prj[j].cpu().numpy(), oracle.siddon_cone_fp(g, img, *args)[0]) < 1e-4
        y = rng.uniform(size=single.shape).astype(np.float32)
        ys.append(y)
        bt.prj[j].copy_(torch.from_numpy(y))
    bt.backward()
    torch.cuda.synchronize()
    for j, (g, geo, img) in enumerate(cases):
        args = (geo.det_center, geo.det_u, geo.det_v, geo.src)
        single = cone.siddon_bp_arbitrary(g.to(ct_projector()), ys[j][None], *args)[0]
        assert rel_l2(bt.img[j].cpu().numpy(), single) < 1e-6
        assert rel_l2(bt.img[j].cpu().numpy(), oracle.siddon_cone_bp(g, ys[j][None], *args)[0]) < 1e-4
    bt.close()


def test_batch_near_axis_parallel_rows(torch, oracle):
    from ctprojector_amd.phantom import ConeGeometry
    g = Geo(nx=16, ny=16, nz=16, dx=1.0, dy=1.0, dz=1.0, nu=30, nv=31, du=1.5, dv=1.5, dsd=300.0, dso=180.0)
    t = np.arange(6) * 2 * np.pi / 6
    geo = ConeGeometry(*(a.astype(np.float32) for a in (
        np.stack([-120 * np.cos(t), -120 * np.sin(t), np.full_like(t, 1e-7)], 1),
        np.stack([-np.sin(t), np.cos(t), np.zeros_like(t)], 1),
        np.tile([[0.0, 0.0, 1.0]], (6, 1)),
        np.stack([180 * np.cos(t), 180 * np.sin(t), np.zeros_like(t)], 1))))
    bt = _batch(torch, [(g, geo, None)])
    y = np.random.default_rng(3).uniform(size=[6, 31, 30]).astype(np.float32)
    bt.prj[0].copy_(torch.from_numpy(y))
    bt.backward()
    torch.cuda.synchronize()
    ref = oracle.siddon_cone_bp(g, y[None], geo.det_center, geo.det_u, geo.det_v, geo.src)[0]
    assert rel_l2(bt.img[0].cpu().numpy(), ref) < 1e-4
    bt.close()


def test_persistent_and_dispatched_plan_launches_agree(torch):
    from ctprojector_amd import _lib
    cases = _jobs()
    out = []
    try:
        for mode in (1, 2):
            _lib.set_tuning('cone_plan_persistent', mode)
            bt = _batch(torch, cases)
            for j, (g, geo, img) in enumerate(cases):
                bt.img[j].copy_(torch.from_numpy(img[0]))
            bt.forward()
            prj = [p.clone() for p in bt.prj]
            bt.backward()
            torch.cuda.synchronize()
            out.append((prj, [i.clone() for i in bt.img]))
            bt.close()
    finally:
        _lib.set_tuning('cone_plan_persistent', 0)
    for a, b in zip(out[0][0], out[1][0]):         # forward: written once per ray
        assert torch.equal(a, b)
    for a, b in zip(out[0][1], out[1][1]):         # backward: exception rays are scattered with atomics
        assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-6


def test_single_and_per_window_backward_launches_agree(torch):
    """The plan's backward as one launch over every job (each block running
    its job's window) and as one launch per window shape (the third job's
    2 mm voxels have a larger candidate window) give the same images."""
    from ctprojector_amd import _lib
    cases = _jobs()
    out = []
    try:
        for mode in (1, 2):
            _lib.set_tuning('cone_plan_bp_launch', mode)
            bt = _batch(torch, cases)
            for j, (g, geo, img) in enumerate(cases):
                bt.img[j].copy_(torch.from_numpy(img[0]))
            bt.forward()
            bt.backward()
            torch.cuda.synchronize()
            out.append([i.clone() for i in bt.img])
            bt.close()
    finally:
        _lib.set_tuning('cone_plan_bp_launch', 0)
    for a, b in zip(out[0], out[1]):                 # exception rays are scattered with atomics
        assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-6


def test_autotune_keeps_results(torch):
    cases = _jobs()
    plain = _batch(torch, cases)
    tuned = _batch(torch, cases, autotune=True)
    cfg = tuned.config()
    assert cfg['cone_fp_tile_u'] in (0, 16, 32, 64) and cfg['cone_bp_zrun'] in (2, 4, 8)
    for j, (g, geo, img) in enumerate(cases):
        plain.img[j].copy_(torch.from_numpy(img[0]))
        tuned.img[j].copy_(torch.from_numpy(img[0]))
    plain.forward()
    tuned.forward()
    torch.cuda.synchronize()
    # the same forward form gives the same sums; the column-driven and the
    # ray-per-lane forms are two FP32 restatements of them
    print('autotuned', cfg)
    tol = 1e-7 if cfg['cone_fp_tile_u'] == 0 else 1e-5
    for j in range(3):
        assert rel_l2(tuned.prj[j].cpu().numpy(), plain.prj[j].cpu().numpy()) < tol
    plain.backward()
    tuned.backward()
    torch.cuda.synchronize()
    for j in range(3):
        assert rel_l2(tuned.img[j].cpu().numpy(), plain.img[j].cpu().numpy()) < 10 * tol
    plain.close()
    tuned.close()


def test_graph_replays_match_eager(torch):
    cases = _jobs()
    bt = _batch(torch, cases)
    gf = bt.capture('forward')
    gb = bt.capture('backward')
    rng = np.random.default_rng(7)
    for it in range(3):
        xs = [rng.uniform(size=tuple(t.shape)).astype(np.float32) for t in bt.img]
        for t, x in zip(bt.img, xs):
            t.copy_(torch.from_numpy(x))
        gf.replay()
        torch.cuda.synchronize()
        got = [t.clone() for t in bt.prj]
        for t, x in zip(bt.img, xs):
            t.copy_(torch.from_numpy(x))
        bt.forward()                           # eager, same inputs
        torch.cuda.synchronize()
        for a, b in zip(got, bt.prj):
            assert torch.equal(a, b), it
        ys = [rng.uniform(size=tuple(t.shape)).astype(np.float32) for t in bt.prj]
        for t, y in zip(bt.prj, ys):
            t.copy_(torch.from_numpy(y))
        gb.replay()
        torch.cuda.synchronize()
        got = [t.clone() for t in bt.img]
        for t, y in zip(bt.prj, ys):
            t.copy_(torch.from_numpy(y))
        bt.backward()
        torch.cuda.synchronize()
        for a, b in zip(got, bt.img):
            assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-7, it
    bt.close()
