"""GPU parity at the benchmark configurations' own geometry and size, against
the FP64 oracle (relative L2 <= 1e-4 in FP32 arithmetic):

* parallel DD, 1x1x256x256 @1 mm, 360 views over 180 degrees, 367 columns, one
  row, through the numpy (host-array) path: FP, BP (plain and FBP weights),
  Hann ramp filter;
* fan DD, the full 1x32x512x512 volume @0.7 mm (1 mm slices), 888x32 detector,
  dso 595, dsd 1085.6, on 130 of its 984 views (every 8th view plus the eight
  views on multiples of 45 degrees: the diagonal ones are where the
  backprojector's main axis switches): FP, BP, Hann ramp filter and fan FBP backprojection, through the
  device path the benchmark times;
* batched cone Siddon, the four uneven volumes (128^3 @1, 192^3 @0.75, 256^3
  @0.5, 128x256x256 @1/0.75 mm) with the full 512x384 panel, each on 24 views
  of its own jittered 720-view trajectory (every 30th), in one batched plan;
* the served stream of the benchmark (16 requests, seed 1005): every result
  is the same in the mix as served alone.

Every case prints what names it before its GPU work."""

import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-4


@pytest.fixture(scope='module')
def torch():
    import torch as T
    assert T.cuda.is_available(), 'torch sees no HIP device'
    return T


def _say(*a):
    print('[config]', *a, flush=True)


def test_parallel_256_numpy_path(oracle):
    from ctprojector_amd.phantom import parallel_256_workload
    from ctprojector_amd.projector.numpy import parallel
    from ctprojector_amd.projector.numpy.ct_projector import ct_projector
    w, img = parallel_256_workload()
    p = ct_projector()
    for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v'):
        setattr(p, k, getattr(w, k))
    img = img[None] if img.ndim == 3 else img
    _say('parallel 256^2 / 360 views / 367 columns: FP')
    fp = parallel.distance_driven_fp(p, img, w.angles)
    e = rel_l2(fp, oracle.dd_par_fp(w, img, w.angles))
    _say('fp rel-L2 %.2e' % e)
    assert e < TOL
    for fbp in (False, True):
        _say('BP is_fbp=%s' % fbp)
        bp = parallel.distance_driven_bp(p, fp, w.angles, is_fbp=fbp)
        e = rel_l2(bp, oracle.dd_par_bp(w, fp, w.angles, is_fbp=fbp))
        _say('bp rel-L2 %.2e' % e)
        assert e < TOL
    _say('Hann ramp filter')
    fl = parallel.ramp_filter(p, fp, 'hann')
    e = rel_l2(fl, oracle.par_ramp_filter(w, fp, 'hann'))
    _say('filter rel-L2 %.2e' % e)
    assert e < TOL


def _fan_subset():
    idx = sorted(set(range(0, 984, 8)) | {123 * k for k in range(8)})
    assert len(idx) == 130
    return np.array(idx)


def test_fan_512x32_full_volume(torch, oracle):
    from ctprojector_amd.phantom import fan_512x32_workload
    from ctprojector_amd.projector.cupy import ct_projector, fan_equiangular as fan
    w, img = fan_512x32_workload()
    idx = _fan_subset()
    ang = np.ascontiguousarray(w.angles[idx])
    # the four diagonal views (odd multiples of 45 degrees) are where |cos| = |sin|
    diag = np.abs(np.abs(np.cos(ang.astype(np.float64))) - np.abs(np.sin(ang.astype(np.float64)))) < 1e-6
    assert diag.sum() == 4
    p = ct_projector()
    for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v',
              'dsd', 'dso'):
        setattr(p, k, getattr(w, k))
    dev = torch.device('cuda', 0)
    ang_d = torch.from_numpy(ang).to(dev)
    _say('fan 32x512^2, 130 of 984 views, 888x32: FP')
    fp = fan.distance_driven_fp(p, torch.from_numpy(img).to(dev), ang_d)
    fp_h = fp.cpu().numpy()
    ref = oracle.dd_fan_fp(w, img, ang)
    e = rel_l2(fp_h, ref)
    _say('fp rel-L2 %.2e' % e)
    assert e < TOL
    _say('BP')
    bp = fan.distance_driven_bp(p, fp, ang_d).cpu().numpy()
    e = rel_l2(bp, oracle.dd_fan_bp(w, fp_h, ang))
    _say('bp rel-L2 %.2e' % e)
    assert e < TOL
    _say('Hann ramp filter')
    fl = fan.ramp_filter(p, fp, 'hann')
    fl_h = fl.cpu().numpy()
    e = rel_l2(fl_h, oracle.fan_ramp_filter(w, fp_h, 'hann'))
    _say('filter rel-L2 %.2e' % e)
    assert e < TOL
    _say('fan FBP backprojection')
    rec = fan.fbp_bp(p, fl, ang_d).cpu().numpy()
    e = rel_l2(rec, oracle.fan_fbp_bp(w, fl_h, ang))
    _say('fbp_bp rel-L2 %.2e' % e)
    assert e < TOL


def test_cone_batch_four_volumes(torch, oracle):
    from ctprojector_amd.batched import ConeBatch, ConeJob
    from ctprojector_amd.phantom import CONE_SAMPLES, ConeGeometry, cone_workload
    from ctprojector_amd.projector.cupy import ct_projector
    sel = np.arange(0, 720, 30)
    cases = []
    for j in range(len(CONE_SAMPLES)):
        w, img, geo = cone_workload(j, 720)
        sub = ConeGeometry(*[np.ascontiguousarray(a[sel]) for a in (geo.det_center, geo.det_u, geo.det_v, geo.src)])
        p = ct_projector()
        for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v'):
            setattr(p, k, getattr(w, k))
        cases.append((w, img, sub, p))
    _say('cone batch of four uneven volumes, 24 views each, 512x384 panel: plan')
    bt = ConeBatch([ConeJob(p, sub) for _, _, sub, p in cases])
    for j, (_, img, _, _) in enumerate(cases):
        bt.img[j].copy_(torch.from_numpy(img[0]))
    bt.forward()
    torch.cuda.synchronize()
    prjs = []
    for j, (w, img, sub, _) in enumerate(cases):
        args = (sub.det_center, sub.det_u, sub.det_v, sub.src)
        got = bt.prj[j].cpu().numpy()
        e = rel_l2(got, oracle.siddon_cone_fp(w, img, *args)[0])
        _say('volume %d fp rel-L2 %.2e' % (j, e))
        assert e < TOL
        prjs.append(got)
    bt.backward()
    torch.cuda.synchronize()
    for j, (w, img, sub, _) in enumerate(cases):
        args = (sub.det_center, sub.det_u, sub.det_v, sub.src)
        got = bt.img[j].cpu().numpy()
        e = rel_l2(got, oracle.siddon_cone_bp(w, prjs[j][None], *args)[0])
        _say('volume %d bp rel-L2 %.2e' % (j, e))
        assert e < TOL
    bt.close()


def test_served_stream_alone_equals_mix(torch):
    from ctprojector_amd.serving import Server, make_stream
    reqs = make_stream(16, seed=1005)
    free, _ = torch.cuda.mem_get_info()
    _say('served stream of 16 requests (seed 1005): mixed run')
    srv = Server(hbm_budget=int(0.5 * free), workers=4)
    srv.warm(reqs)
    st = srv.serve(reqs)
    assert st.completed == len(reqs)
    mixed = {r.rid: r.result.cpu().numpy() for r in reqs}
    for r in make_stream(16, seed=1005):
        _say('alone: request %d (%s)' % (r.rid, r.kind))
        Server(hbm_budget=int(0.5 * free), workers=1).serve([r])
        e = rel_l2(r.result.cpu().numpy(), mixed[r.rid])
        assert e < 1e-6, (r.rid, r.kind, e)
