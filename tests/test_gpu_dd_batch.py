"""GPU: batched plans of uneven distance-driven jobs (ctpDDBatch*,
ctprojector_amd.batched.DDBatch).

* eight fan- and parallel-beam jobs of different grids (square and not,
  off-centre), detectors, view counts, batch counts and typeProjector bits
  (FBP, branchless, force-FBP), covering every kernel class: one-lane and
  split-lane forward projectors with one and two row segments, one and two
  channel chunks, one-lane and split-lane backprojectors with and without
  view groups: each job's forward and backward equal its single call through
  the C ABI, and two jobs are checked against the FP64 oracle;
* the autotuner, both block distributions (hardware dispatch, the default,
  and persistent work counters) and the kernel classes on four concurrent
  streams (the default) or one give the same results;
* HIP-graph capture (one stream, and four streams forked and joined inside
  the capture): forward and backward replayed three times on new inputs, with
  eager work in between, each replay equal to the eager run.

Every case prints what names it before its GPU work."""

from ctypes import POINTER, c_float, c_int, c_ulong

import numpy as np
import pytest

from helpers import Geo, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch as T
    assert T.cuda.is_available(), 'torch sees no HIP device'
    return T


def _say(*a):
    print('[dd-batch]', *a, flush=True)


def _job(fan, n, nz, nview, nu, B=1, typ=0, c=(0.0, 0.0), off=0.0, seed=0):
    from ctprojector_amd.phantom import ellipsoid_phantom
    nx, ny = n if isinstance(n, tuple) else (n, n)
    kw = dict(nx=nx, ny=ny, nz=nz, dx=1.0, dy=0.9, dz=1.0, cx=c[0], cy=c[1], cz=0.0, nv=nz, dv=1.0, off_u=off,
              off_v=0.0)
    if fan:
        kw.update(nu=nu, du=1.4, dsd=200.0, dso=120.0)
    else:
        kw.update(nu=nu, du=1.0)
    g = Geo(**kw)
    span = 2 * np.pi if fan else np.pi
    ang = (np.arange(nview) * span / nview + 0.01 * seed).astype(np.float32)
    img = np.stack([ellipsoid_phantom(nz, ny, nx, 1.0, 0.9, 1.0, seed=seed + b) for b in range(B)], 0)
    return dict(fan=fan, g=g, ang=ang, B=B, typ=typ, img=img.astype(np.float32))


def _jobs():
    return [
        _job(True, 48, 3, 41, 72, B=2, seed=1),                        # 6 channels: one-lane FP, 2-lane BP
        _job(False, (40, 34), 1, 30, 57, seed=3, c=(1.5, -0.7), off=0.3),   # 1 channel, view groups
        _job(True, 48, 16, 50, 80, seed=5, typ=1),                     # split-lane FP and BP, FBP weights
        _job(True, 64, 16, 60, 96, seed=7, typ=2),                     # two row segments, branchless
        _job(False, 44, 5, 33, 65, B=3, seed=9, typ=4),                # 15 channels, force-FBP
        _job(True, 40, 40, 24, 64, seed=11),                           # two channel chunks
        _job(True, (56, 40), 2, 37, 70, seed=13, typ=3, c=(2.0, 1.0), off=-0.4),
        _job(True, 64, 32, 700, 300, seed=15),                         # split-lane FP <32, 4>, one row segment
    ]


def _p(a):
    return a.ctypes.data_as(POINTER(c_float))


def _single(J, direction, inp):
    """The job's single call through the C ABI (host arrays)."""
    from ctprojector_amd import _lib
    g, ang, B = J['g'], J['ang'], J['B']
    nview = ang.shape[0]
    if direction == 'fp':
        out = np.zeros([B, nview, g.nv, g.nu], np.float32)
    else:
        out = np.zeros([B, g.nz, g.ny, g.nx], np.float32)
    grid = (c_ulong(B), c_ulong(g.nx), c_ulong(g.ny), c_ulong(g.nz), c_float(g.dx), c_float(g.dy), c_float(g.dz),
            c_float(g.cx), c_float(g.cy), c_float(g.cz))
    da = g.du / g.dsd if J['fan'] else g.du
    det = (c_ulong(g.nu), c_ulong(g.nv), c_ulong(nview), c_float(da), c_float(g.dv), c_float(g.off_u),
           c_float(g.off_v))
    if J['fan']:
        det = det + (c_float(g.dsd), c_float(g.dso))
    name = 'cDistanceDriven%s%s' % ('Fan' if J['fan'] else 'Parallel',
                                    'Projection' if direction == 'fp' else 'Backprojection')
    err = getattr(_lib.module(), name)(_p(out), _p(np.ascontiguousarray(inp)), _p(ang), *grid, *det,
                                        c_int(J['typ']))
    assert err == 0, name
    return out


def _batch(jobs, autotune=False):
    from ctprojector_amd.batched import DDBatch, DDJob
    from ctprojector_amd.projector.cupy import ct_projector
    dj = [DDJob(J['g'].to(ct_projector()), J['ang'], fan=J['fan'], batch=J['B'], is_fbp=bool(J['typ'] & 1),
                branchless=bool(J['typ'] & 2), force_fbp=bool(J['typ'] & 4)) for J in jobs]
    return DDBatch(dj, autotune=autotune)


def _run(torch, bt, jobs, seed=0):
    rng = np.random.default_rng(seed)
    ys = []
    for j, J in enumerate(jobs):
        bt.img[j].copy_(torch.from_numpy(J['img']))
    bt.forward()
    torch.cuda.synchronize()
    fps = [bt.prj[j].cpu().numpy() for j in range(len(jobs))]
    for j in range(len(jobs)):
        y = rng.uniform(size=fps[j].shape).astype(np.float32)
        ys.append(y)
        bt.prj[j].copy_(torch.from_numpy(y))
    bt.backward()
    torch.cuda.synchronize()
    bps = [bt.img[j].cpu().numpy() for j in range(len(jobs))]
    return fps, ys, bps


def test_batch_equals_single_calls(torch, oracle):
    jobs = _jobs()
    _say('plan of %d jobs' % len(jobs))
    bt = _batch(jobs)
    cfg = bt.config()
    _say('config', cfg)
    assert cfg['fp_launches'] >= 3 and cfg['bp_launches'] >= 4
    fps, ys, bps = _run(torch, bt, jobs)
    for j, J in enumerate(jobs):
        _say('job %d: single-call FP and BP' % j)
        ref = _single(J, 'fp', J['img'])
        e = rel_l2(fps[j], ref)
        same = np.array_equal(fps[j], ref)
        ref_b = _single(J, 'bp', ys[j])
        eb = rel_l2(bps[j], ref_b)
        same_b = np.array_equal(bps[j], ref_b)
        _say('job %d fp %.1e (identical %s) bp %.1e (identical %s)' % (j, e, same, eb, same_b))
        assert e < 1e-6 and eb < 1e-6
    for j in (0, 4):
        J = jobs[j]
        f = oracle.dd_fan_fp if J['fan'] else oracle.dd_par_fp
        b = oracle.dd_fan_bp if J['fan'] else oracle.dd_par_bp
        e = rel_l2(fps[j], f(J['g'], J['img'], J['ang'], type_projector=J['typ']))
        eb = rel_l2(bps[j], b(J['g'], ys[j], J['ang'], type_projector=J['typ']))
        _say('job %d against the oracle: fp %.2e bp %.2e' % (j, e, eb))
        assert e < 1e-4 and eb < 1e-4
    bt.close()


def test_autotune_and_dispatch_forms(torch):
    from ctprojector_amd import _lib
    jobs = _jobs()
    _say('reference run (defaults)')
    bt = _batch(jobs)
    ref = _run(torch, bt, jobs)
    bt.close()
    _say('autotuned plan')
    bt = _batch(jobs, autotune=True)
    _say('config', bt.config())
    got = _run(torch, bt, jobs)
    bt.close()
    for a, b in zip(ref[0] + ref[2], got[0] + got[2]):
        assert rel_l2(b, a) < 1e-6
    try:
        _lib.set_tuning('dd_batch_persistent', 1)
        _say('persistent blocks with work counters')
        bt = _batch(jobs)
        assert bt.config()['fp_persistent'] == 1
        got = _run(torch, bt, jobs)
        bt.close()
    finally:
        _lib.set_tuning('dd_batch_persistent', 0)
    for a, b in zip(ref[0] + ref[2], got[0] + got[2]):
        assert rel_l2(b, a) < 1e-6
    try:
        _lib.set_tuning('dd_batch_streams', 1)
        _say('kernel classes on one stream')
        bt = _batch(jobs)
        assert bt.config()['fp_streams'] == 1
        got = _run(torch, bt, jobs)
        bt.close()
    finally:
        _lib.set_tuning('dd_batch_streams', 0)
    for a, b in zip(ref[0] + ref[2], got[0] + got[2]):
        assert rel_l2(b, a) < 1e-6


@pytest.mark.parametrize('streams', [1, 4])
def test_graph_replays(torch, streams):
    from ctprojector_amd import _lib
    jobs = _jobs()[:5]
    _lib.set_tuning('dd_batch_streams', streams)
    try:
        bt = _batch(jobs)
    finally:
        _lib.set_tuning('dd_batch_streams', 0)
    assert bt.config()['bp_streams'] == streams
    rng = np.random.default_rng(4)
    _say('capture forward and backward graphs (%d streams)' % streams)
    gf = bt.capture('forward')
    gb = bt.capture('backward')
    for rep in range(3):
        imgs = [torch.from_numpy(rng.uniform(size=tuple(t.shape)).astype(np.float32)).cuda() for t in bt.img]
        prjs = [torch.from_numpy(rng.uniform(size=tuple(t.shape)).astype(np.float32)).cuda() for t in bt.prj]
        _say('replay %d' % rep)
        for t, x in zip(bt.img, imgs):
            t.copy_(x)
        gf.replay()
        torch.cuda.synchronize()
        got_f = [t.clone() for t in bt.prj]
        for t, x in zip(bt.prj, prjs):
            t.copy_(x)
        gb.replay()
        torch.cuda.synchronize()
        got_b = [t.clone() for t in bt.img]
        (torch.randn(1 << 20, device='cuda') * 2).sum().item()     # eager work between replays
        for t, x in zip(bt.img, imgs):
            t.copy_(x)
        bt.forward()
        torch.cuda.synchronize()
        for a, b in zip(got_f, bt.prj):
            assert torch.equal(a, b)
        for t, x in zip(bt.prj, prjs):
            t.copy_(x)
        bt.backward()
        torch.cuda.synchronize()
        for a, b in zip(got_b, bt.img):
            assert torch.equal(a, b)
    bt.close()
