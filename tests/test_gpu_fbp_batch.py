"""GPU: batched ramp filters and filtered backprojection of uneven jobs
(ctpFilterBatch*, ctprojector_amd.batched.RampFilterBatch / FBPBatch).

* fan- and parallel-beam filter jobs of different detectors (transform
  sizes 256 to 4096, the generic and the three-pass FFT kernels, two jobs
  sharing one size), view counts, batch counts, offsets and windows: each
  job's filtered sinogram equals its single ramp_filter call bit for bit;
* FBP of uneven jobs (filter batch into a DD batch with FBP weights) equals
  ramp_filter followed by distance_driven_bp(is_fbp=True) per job, and the
  FP64 oracle's filter and FBP-weighted backprojection within 1e-4 rel-L2;
* HIP-graph capture of the FBP batch: replayed three times on new inputs,
  with eager work in between, each replay equal to the eager run;
* many jobs of one transform size (seven generic, three three-pass), odd
  row counts: each equals its single call bit for bit;
* jobs of 2049 and 3000 columns (the direct convolution) beside an FFT job
  equal their single calls; more than 5000 columns are refused.

Every case prints what names it before its GPU work."""

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
    print('[fbp-batch]', *a, flush=True)


def _geo(fan, n, nz, nu, off=0.0, c=(0.0, 0.0)):
    kw = dict(nx=n, ny=n, nz=nz, dx=1.0, dy=1.0, dz=1.0, cx=c[0], cy=c[1], cz=0.0, nu=nu, nv=nz, dv=1.0,
              off_u=off, off_v=0.0)
    if fan:
        kw.update(du=1.4 * 64 / nu * 2, dsd=200.0, dso=120.0)
    else:
        kw.update(du=1.3 * n / nu)
    return Geo(**kw)


def _proj(g):
    from ctprojector_amd.projector.cupy import ct_projector
    return g.to(ct_projector())


# (fan, n, nz, nview, nu, batch, filter, offset): transform sizes 256, 1024
# (two jobs), 2048, 4096
FILTER_JOBS = [
    (True, 32, 2, 20, 100, 1, 'hann', 0.0),
    (True, 48, 3, 24, 300, 2, 'hamming', 0.4),
    (False, 40, 1, 18, 450, 1, 'hann', -0.3),
    (False, 64, 2, 12, 700, 1, 'ramlak', 0.0),
    (True, 64, 1, 6, 2000, 1, 'cosine', 0.2),
]


def test_filter_batch_equals_single_calls(torch):
    from ctprojector_amd.batched import FilterJob, RampFilterBatch
    from ctprojector_amd.projector.cupy import fan_equiangular as fan, parallel
    jobs, data = [], []
    rng = np.random.default_rng(5)
    for (f, n, nz, nview, nu, B, ft, off) in FILTER_JOBS:
        p = _proj(_geo(f, n, nz, nu, off=off))
        jobs.append(FilterJob(p, nview, fan=f, batch=B, filter_type=ft))
        data.append(rng.random((B, nview, nz, nu), dtype=np.float32))
    _say('filter batch of %d jobs, nu %s' % (len(jobs), [j[4] for j in FILTER_JOBS]))
    fb = RampFilterBatch(jobs)
    for t, x in zip(fb.prj, data):
        t.copy_(torch.from_numpy(x))
    fb.run()
    torch.cuda.synchronize()
    for k, (job, x) in enumerate(zip(jobs, data)):
        m = fan if job.fan else parallel
        single = m.ramp_filter(job.projector, torch.from_numpy(x).cuda(), job.filter_type)
        assert torch.equal(fb.fprj[k], single), k
    fb.close()



def test_filter_batch_many_jobs_per_size(torch):
    """Seven jobs of one generic-kernel size (N = 256) and three of the
    three-pass size N = 1024, odd row counts (a job's last pair has one row):
    each complex row finds its job by the binary search over the group."""
    from ctprojector_amd.batched import FilterJob, RampFilterBatch
    from ctprojector_amd.projector.cupy import fan_equiangular as fan, parallel
    specs = [(k % 2 == 0, 70 + 9 * k, 1 + k % 3, 5 + 2 * k, 1) for k in range(7)]
    specs += [(True, 300, 3, 7, 1), (False, 420, 1, 9, 3), (True, 511, 5, 3, 1)]
    jobs, data = [], []
    rng = np.random.default_rng(17)
    for k, (f, nu, nv, nview, B) in enumerate(specs):
        p = _proj(_geo(f, 32, nv, nu, off=0.1 * (k % 3) - 0.1))
        ft = ('hann', 'hamming', 'cosine', 'ramlak')[k % 4]
        jobs.append(FilterJob(p, nview, fan=f, batch=B, filter_type=ft))
        data.append(rng.random((B, nview, nv, nu), dtype=np.float32))
    _say('filter batch of %d jobs in two transform sizes' % len(jobs))
    fb = RampFilterBatch(jobs)
    for t, x in zip(fb.prj, data):
        t.copy_(torch.from_numpy(x))
    fb.run()
    torch.cuda.synchronize()
    for k, (job, x) in enumerate(zip(jobs, data)):
        m = fan if job.fan else parallel
        single = m.ramp_filter(job.projector, torch.from_numpy(x).cuda(), job.filter_type)
        assert torch.equal(fb.fprj[k], single), k
    fb.close()

def _fbp_jobs():
    # (fan, n, nz, nview, nu, batch, offset, centre)
    return [(True, 48, 3, 40, 90, 2, 0.3, (1.0, -0.5)), (False, 40, 1, 36, 70, 1, -0.2, (0.0, 0.0)),
            (True, 56, 8, 60, 300, 1, 0.0, (0.0, 0.0))]


def _fbp_batch(torch, filt='hann'):
    from ctprojector_amd.batched import DDJob, FBPBatch
    specs = _fbp_jobs()
    jobs = []
    for (f, n, nz, nview, nu, B, off, c) in specs:
        p = _proj(_geo(f, n, nz, nu, off=off, c=c))
        span = 2 * np.pi if f else np.pi
        ang = (np.arange(nview) * span / nview).astype(np.float32)
        jobs.append(DDJob(p, ang, fan=f, batch=B))
    return FBPBatch(jobs, filter_type=filt), jobs


def test_fbp_batch_equals_filter_then_bp(torch):
    from ctprojector_amd.projector.cupy import fan_equiangular as fan, parallel
    _say('fbp batch of %d uneven jobs' % len(_fbp_jobs()))
    fb, jobs = _fbp_batch(torch)
    rng = np.random.default_rng(8)
    data = [rng.random(tuple(t.shape), dtype=np.float32) for t in fb.prj]
    for t, x in zip(fb.prj, data):
        t.copy_(torch.from_numpy(x))
    fb.run()
    torch.cuda.synchronize()
    for k, (job, x) in enumerate(zip(jobs, data)):
        m = fan if job.fan else parallel
        ang = torch.from_numpy(np.asarray(job.angles)).cuda()
        fprj = m.ramp_filter(job.projector, torch.from_numpy(x).cuda(), 'hann')
        ref = m.distance_driven_bp(job.projector, fprj, ang, is_fbp=True)
        assert rel_l2(fb.img[k].cpu().numpy(), ref.cpu().numpy()) < 1e-6, k
    fb.close()



def test_fbp_batch_against_oracle(torch, oracle):
    """Each job of the uneven FBP batch against the FP64 oracle's ramp filter
    followed by its FBP-weighted DD backprojection, on forward projections of
    an ellipse phantom (the filter on random data would make the relative
    error meaningless): rel-L2 below 1e-4."""
    from ctprojector_amd.phantom import ellipsoid_phantom
    _say('fbp batch against the oracle')
    fb, jobs = _fbp_batch(torch)
    refs = []
    for k, ((f, n, nz, nview, nu, B, off, c), job) in enumerate(zip(_fbp_jobs(), jobs)):
        g = _geo(f, n, nz, nu, off=off, c=c)
        img = np.stack([ellipsoid_phantom(nz, n, n, 1.0, 1.0, 1.0, seed=60 + 7 * k + b) for b in range(B)])
        ang = np.asarray(job.angles, dtype=np.float32)
        if f:
            prj = oracle.dd_fan_fp(g, img, ang).astype(np.float32)
            refs.append(oracle.dd_fan_bp(g, oracle.fan_ramp_filter(g, prj, 'hann').astype(np.float32), ang,
                                         is_fbp=True))
        else:
            prj = oracle.dd_par_fp(g, img, ang).astype(np.float32)
            refs.append(oracle.dd_par_bp(g, oracle.par_ramp_filter(g, prj, 'hann').astype(np.float32), ang,
                                         is_fbp=True))
        fb.prj[k].copy_(torch.from_numpy(prj))
    fb.run()
    torch.cuda.synchronize()
    for k, ref in enumerate(refs):
        e = rel_l2(fb.img[k].cpu().numpy(), ref)
        _say('job %d rel-L2 %.3g' % (k, e))
        assert e < 1e-4, k
    fb.close()

def test_fbp_batch_graph_replays_match_eager(torch):
    _say('fbp batch graph capture')
    fb, _ = _fbp_batch(torch)
    g = fb.capture()
    rng = np.random.default_rng(11)
    for rep in range(3):
        data = [rng.random(tuple(t.shape), dtype=np.float32) for t in fb.prj]
        for t, x in zip(fb.prj, data):
            t.copy_(torch.from_numpy(x))
        for t in fb.img:
            t.zero_()
        g.replay()
        torch.cuda.synchronize()
        replay = [t.clone() for t in fb.img]
        # eager work in between, then the eager result on the same inputs
        _ = torch.ones(1000, device='cuda').sum().item()
        fb.run()
        torch.cuda.synchronize()
        for a, b in zip(replay, fb.img):
            assert torch.equal(a, b), rep
    fb.close()


def test_filter_batch_wide_rows(torch):
    """Jobs of 2049 to 5000 columns take the single call's direct
    convolution inside the batch (equal to it bit for bit, beside an FFT
    job); more than 5000 columns are refused, as by the single call."""
    from ctprojector_amd.batched import FilterJob, RampFilterBatch
    from ctprojector_amd.projector.cupy import fan_equiangular as fan, parallel
    specs = [(True, 2049, 2, 3, 'hann'), (False, 3000, 1, 2, 'ramlak'), (True, 500, 2, 5, 'hamming')]
    jobs, data = [], []
    rng = np.random.default_rng(23)
    for f, nu, nv, nview, ft in specs:
        jobs.append(FilterJob(_proj(_geo(f, 32, nv, nu, off=0.2)), nview, fan=f, filter_type=ft))
        data.append(rng.random((1, nview, nv, nu), dtype=np.float32))
    _say('filter batch with 2049- and 3000-column jobs')
    fb = RampFilterBatch(jobs)
    for t, x in zip(fb.prj, data):
        t.copy_(torch.from_numpy(x))
    fb.run()
    torch.cuda.synchronize()
    for k, (job, x) in enumerate(zip(jobs, data)):
        m = fan if job.fan else parallel
        single = m.ramp_filter(job.projector, torch.from_numpy(x).cuda(), job.filter_type)
        assert torch.equal(fb.fprj[k], single), k
    fb.close()
    _say('filter batch: a 5001-column detector is refused')
    with pytest.raises(RuntimeError, match='5000'):
        RampFilterBatch([FilterJob(_proj(_geo(True, 32, 1, 5001)), 2)])
