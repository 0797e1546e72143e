#!/usr/bin/env python3
"""Projector hot-path benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-extra]
                    [--no-serving] [--no-parity] [--calibrate-copy]
    python bench.py --summarize TRACE_DIR [--fetch DIR --write DIR [--hits DIR]]

One step = one forward projection plus one backprojection of the
fan-beam distance-driven workload: volume 1x32x512x512 at 0.7 mm (1 mm slices),
984 views over 360 degrees, 888 columns x 32 rows, dso 595 mm, dsd 1085.6 mm,
seeded ellipsoid phantom (synthetic).  Inputs are resident on the GPU when the
timed region starts; outputs are produced by the library's device entry points
(``cupy*`` symbols of libprojector.so) through the tensor front end.

The JSON line reports
  value        FP+BP voxel-updates per second (1 voxel-update = one (voxel, view)
               pair in one pass; a step counts 2 * B*nx*ny*nz*nview), in GVU/s;
  roofline     for the dominant kernel: its HBM traffic per launch (rocprofv3
               FETCH_SIZE scaled by a device copy's calibration + WRITE_SIZE,
               from the committed counter summary under profiles/) over its
               mean launch time, measured here with HIP events on the
               launching stream, against the 8 TB/s HBM peak; beside it the
               4 B per voxel-update throughput ratio (the voxel operand
               streamed once per view, which the caches mostly serve) and the
               compulsory-traffic fraction (|img| + |prj| per launch);
  cpu_baseline the FP32 OpenMP restatement of the reference algorithm (the
               oracle's FP32 mirror) built -march=native for this host, on a
               view subset of the same workload, median of 5 passes after one
               warm-up, OMP_PROC_BIND=close OMP_PLACES=cores, CPU model stated
               (oracle/cpu_baseline.py);
  parity_rel_l2  relative L2 of the GPU outputs against the FP64 oracle over the
               whole fan workload (FP, BP, Hann ramp filter, FBP BP) and the
               whole cone workload (FP, BP), computed on the host after the
               GPU timing (oracle/cpu_baseline.py --check);
  extra        the cone-beam Siddon FP+BP (256^3 volume at 0.5 mm, 720 views,
               512x384 panel), the batched cone plan, the fan ramp filter + FBP
               backprojection, the unrolled SQS-TV step, the TV prior and the
               served stream, measured the same way.

--calibrate-copy runs three 256 MiB device-to-device copies first, so that a
rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE pass over the same command can be
scaled by what the counters read for a copy of known size.  --summarize turns
rocprofv3 output directories (kernel trace with --stats, and optionally the
FETCH_SIZE, WRITE_SIZE and TCC hit/miss passes) into the markdown tables kept
under profiles/.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3      # MI355X published FP32 (the vector FMA rate with packed instructions)
COPY_BYTES = 256 << 20
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'hbm_traffic_fan512x32_cone256.md')


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_stats(times, name):
    ms, n = times.get(name, (0.0, 0))
    return (ms / n if n else 0.0), n


def fan_workload_bench(steps, warmup):
    import torch
    from ctprojector_amd import _lib
    from ctprojector_amd.phantom import fan_512x32_workload
    from ctprojector_amd.projector.cupy import ct_projector, fan_equiangular as fan

    w, img = fan_512x32_workload()
    p = ct_projector()
    for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v',
              'dsd', 'dso', 'nview'):
        setattr(p, k, getattr(w, k))
    dev = torch.device('cuda', 0)
    img_d = torch.from_numpy(img).to(dev)
    ang_d = torch.from_numpy(w.angles).to(dev)
    nvu = img.shape[0] * w.nx * w.ny * w.nz * w.nview

    log('[bench] fan DD warmup x%d' % warmup)
    for _ in range(warmup):
        prj = fan.distance_driven_fp(p, img_d, ang_d)
        _ = fan.distance_driven_bp(p, prj, ang_d)
    torch.cuda.synchronize()

    _lib.set_kernel_timing(True)
    _lib.reset_kernel_times()
    log('[bench] fan DD timed x%d' % steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        prj = fan.distance_driven_fp(p, img_d, ang_d)
        bp = fan.distance_driven_bp(p, prj, ang_d)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    times = _lib.kernel_times()
    _lib.set_kernel_timing(False)
    ms_step = (t1 - t0) / steps * 1e3
    finite = bool(torch.isfinite(bp).all().item()) and bool(torch.isfinite(prj).all().item())
    return dict(w=w, img=img, nvu=nvu, ms_step=ms_step, times=times, finite=finite,
                outputs={'prj': prj.cpu().numpy(), 'bp': bp.cpu().numpy()})


def cone_extra(steps, cpu=False, outputs=None):
    import torch
    from ctprojector_amd import _lib
    from ctprojector_amd.phantom import cone_workload
    from ctprojector_amd.projector.cupy import ct_projector, cone

    w, img, geo = cone_workload(2, 720)
    p = ct_projector()
    for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v'):
        setattr(p, k, getattr(w, k))
    dev = torch.device('cuda', 0)
    img_d = torch.from_numpy(img).to(dev)
    g = [torch.from_numpy(a).to(dev) for a in (geo.det_center, geo.det_u, geo.det_v, geo.src)]
    nvu = w.nx * w.ny * w.nz * w.nview
    log('[bench] cone Siddon warmup')
    prj = cone.siddon_fp_arbitrary(p, img_d, *g)
    _ = cone.siddon_bp_arbitrary(p, prj, *g)
    torch.cuda.synchronize()
    _lib.set_kernel_timing(True)
    _lib.reset_kernel_times()
    log('[bench] cone Siddon timed x%d' % steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        prj = cone.siddon_fp_arbitrary(p, img_d, *g)
        bp = cone.siddon_bp_arbitrary(p, prj, *g)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    times = _lib.kernel_times()
    _lib.set_kernel_timing(False)
    ms = (t1 - t0) / steps * 1e3
    fp_ms, _ = kernel_stats(times, 'siddon_cone_fp')
    bp_ms, _ = kernel_stats(times, 'siddon_cone_bp')
    tab_ms, _ = kernel_stats(times, 'siddon_bp_table')
    tr_ms, _ = kernel_stats(times, 'siddon_transpose')
    out = {
        'workload': 'cone Siddon FP+BP, 256^3 @0.5 mm, 720 views, 512x384 panel, jittered helical',
        'ms_per_step': round(ms, 3),
        'value_gvu_s': round(2 * nvu / (ms / 1e3) / 1e9, 2),
        'kernel_ms': {'siddon_transpose': round(tr_ms, 3), 'siddon_cone_fp': round(fp_ms, 3),
                      'siddon_bp_table': round(tab_ms, 3), 'siddon_cone_bp': round(bp_ms, 3)},
        # 4 B per voxel update and pass, over the summed kernel time of the step
        'roofline_frac_fp_bp_kernels': round(2 * 4 * nvu / ((fp_ms + bp_ms + tab_ms + tr_ms) / 1e3) / 1e9 /
                                             PEAK_HBM_GBS, 4) if fp_ms > 0 else None,
        'traffic_bytes': {'siddon_cone_fp': read_traffic('siddon_cone_fp'),
                          'siddon_cone_bp': read_traffic('siddon_cone_bp')},
    }
    # the same geometry for a batch of two volumes per call (column FP and BP
    # share the traversal and the chords across batch items)
    img2 = torch.cat([img_d, 0.5 * img_d], 0).contiguous()
    prj = cone.siddon_fp_arbitrary(p, img2, *g)
    _ = cone.siddon_bp_arbitrary(p, prj, *g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        prj = cone.siddon_fp_arbitrary(p, img2, *g)
        bp = cone.siddon_bp_arbitrary(p, prj, *g)
    torch.cuda.synchronize()
    ms2 = (time.perf_counter() - t0) / steps * 1e3
    out['batch2_ms_per_step'] = round(ms2, 3)
    out['batch2_value_gvu_s'] = round(2 * 2 * nvu / (ms2 / 1e3) / 1e9, 2)
    # and four volumes per call (the most the multi-volume kernels take per launch)
    img4 = torch.cat([img2, 0.25 * img2], 0).contiguous()
    prj = cone.siddon_fp_arbitrary(p, img4, *g)
    _ = cone.siddon_bp_arbitrary(p, prj, *g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        prj = cone.siddon_fp_arbitrary(p, img4, *g)
        bp = cone.siddon_bp_arbitrary(p, prj, *g)
    torch.cuda.synchronize()
    ms4 = (time.perf_counter() - t0) / steps * 1e3
    out['batch4_ms_per_step'] = round(ms4, 3)
    out['batch4_value_gvu_s'] = round(2 * 4 * nvu / (ms4 / 1e3) / 1e9, 2)
    del prj, bp, img2, img4
    if outputs is not None:              # single-volume outputs for the oracle check
        prj = cone.siddon_fp_arbitrary(p, img_d, *g)
        bp = cone.siddon_bp_arbitrary(p, prj, *g)
        outputs['prj'] = prj[0].cpu().numpy()[None]
        outputs['bp'] = bp.cpu().numpy()
        del prj, bp
    if cpu:
        out['cpu_baseline'] = run_cpu_baseline('cone')
    return out


def run_cpu_baseline(workload):
    """The oracle's FP32 OpenMP restatement of the reference kernels, built
    -march=native for this host and timed in a child process with
    OMP_PROC_BIND=close OMP_PLACES=cores (oracle/cpu_baseline.py): median of 5
    passes after 1 warm-up over a view subset of about 2.5 s per pass."""
    import subprocess
    env = dict(os.environ, OMP_PROC_BIND='close', OMP_PLACES='cores')
    log('[bench] cpu baseline (%s): building for this host and timing' % workload)
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'oracle', 'cpu_baseline.py'), '--workload', workload],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env, cwd=ROOT, timeout=600)
    if r.returncode != 0:
        return {'error': 'cpu baseline failed (%d): %s' % (r.returncode, r.stderr.strip()[-500:])}
    return json.loads(r.stdout.strip().splitlines()[-1])


def unrolled_sqs_extra(steps):
    """Differentiable unrolled SQS-TV (10 iterations, exact adjoint; the
    iterations' intermediates are kept, as the default does where they fit in
    a quarter of the free HBM) on 8 fan-beam 256^2 slices @1.4 mm, 492 views,
    444 columns, Poisson data at 1e5 photons: one training step = the network
    forward and the gradient with respect to the sinogram, timed eagerly and
    as a HIP graph replay of the whole step (the value reported)."""
    import torch
    from ctprojector_amd.autograd import FanDD, sqs_tv_unrolled, unrolled_needs_checkpoint
    from ctprojector_amd.phantom import fan_slices_256_workload
    from ctprojector_amd.projector.cupy import ct_projector

    w, img = fan_slices_256_workload(8)
    p = ct_projector()
    for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v',
              'dsd', 'dso', 'nview'):
        setattr(p, k, getattr(w, k))
    dev = torch.device('cuda', 0)
    op = FanDD(p, torch.from_numpy(w.angles).to(dev))
    x_true = torch.from_numpy(img).to(dev)
    gen = torch.Generator(device=dev).manual_seed(1004)
    with torch.no_grad():
        line = op.fp(x_true)
        counts = torch.poisson(1e5 * torch.exp(-line), generator=gen)
        prj = -torch.log(torch.clamp(counts, min=1.0) / 1e5)
        wgt = counts / 1e5
        v = torch.rand(x_true.shape, device=dev, generator=gen)
        for _ in range(10):                  # power method for the projector norm
            v = op.adjoint(op.fp(v))
            v = v / v.norm()
        pn = float(op.fp(v).norm())
        norm_img = op.adjoint(wgt * op.fp(torch.ones_like(x_true))) / pn / pn
    niter, beta = 10, 2e-3
    ckpt = unrolled_needs_checkpoint(prj, x_true, niter)

    x0 = torch.zeros_like(x_true)

    def net_and_grad(q):
        # loss = sum((x - x_true)^2); its gradient 2 (x - x_true) is handed to
        # backward directly, so the step holds no reduction
        x = sqs_tv_unrolled(op, q, x0, norm_img, pn, beta, niter, weight=wgt, checkpoint=ckpt)
        x.backward(2.0 * (x - x_true))
        return x

    def step():
        q = prj.clone().requires_grad_(True)
        x = net_and_grad(q)
        return x, q.grad

    log('[bench] unrolled SQS-TV warmup')
    step()
    torch.cuda.synchronize()
    log('[bench] unrolled SQS-TV eager x%d' % steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        x, grad = step()
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) / steps * 1e3
    # the same training step captured once in a HIP graph and replayed: every
    # kernel of the forward and the gradient runs in each replay, without the
    # host's per-launch work (warm-up on a side stream first, as graph capture
    # requires; the gradient buffer is written, not accumulated, by each replay)
    graphed = not ckpt
    ms = eager_ms
    same = None
    if graphed:
        log('[bench] unrolled SQS-TV graph capture')
        q_static = prj.clone().requires_grad_(True)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):
                q_static.grad = None
                net_and_grad(q_static)
        torch.cuda.current_stream().wait_stream(side)
        q_static.grad = None
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            x_static = net_and_grad(q_static)
        g_static = q_static.grad
        log('[bench] unrolled SQS-TV graph replay x%d' % steps)
        graph.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            graph.replay()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        same = bool(torch.equal(g_static, grad)) and bool(torch.equal(x_static, x))
        x, grad = x_static, g_static
    loss = ((x - x_true) ** 2).sum()
    nvu = img.shape[0] * w.nx * w.ny * w.nz * w.nview
    # FP and A^T per iteration forward, again in a checkpointed recomputation,
    # and the two products of the backward pass: 4 or 6 passes per iteration
    passes = (6 if ckpt else 4) * niter
    return {'workload': 'unrolled SQS-TV, 10 iterations, 8 fan slices 256^2 @1.4 mm, 492 views, 444 columns, '
                        'Poisson 1e5; forward + gradient w.r.t. the sinogram',
            'ms_per_step': round(ms, 3), 'ms_per_iteration': round(ms / niter, 3), 'checkpoint': ckpt,
            'hip_graph': graphed, 'eager_ms_per_step': round(eager_ms, 3), 'graph_equals_eager': same,
            'projector_passes_per_step': passes, 'value_gvu_s': round(passes * nvu / (ms / 1e3) / 1e9, 2),
            'finite': bool(torch.isfinite(grad).all().item()) and bool(torch.isfinite(loss).item())}


def tv_prior_extra(steps):
    """TV-SQS surrogate (s1, s2, var) and the fused SQS-TV update on a 256^3 volume:
    kernel time and bandwidth against the fused kernels' 16 B per voxel."""
    import torch
    from ctprojector_amd import _lib
    from ctprojector_amd.prior.cupy import tv_sqs, tv_sqs_update

    dev = torch.device('cuda', 0)
    gen = torch.Generator(device=dev).manual_seed(1006)
    x = torch.rand([1, 256, 256, 256], device=dev, generator=gen)
    bp = torch.rand_like(x)
    nrm = torch.rand_like(x) + 1.0
    tv_sqs(x, (1.0, 1.0, 1.0))
    tv_sqs_update(x, bp, nrm, 1e-3, (1.0, 1.0, 1.0))
    torch.cuda.synchronize()
    _lib.set_kernel_timing(True)
    _lib.reset_kernel_times()
    for _ in range(steps):
        tv_sqs(x, (1.0, 1.0, 1.0))
        tv_sqs_update(x, bp, nrm, 1e-3, (1.0, 1.0, 1.0))
    torch.cuda.synchronize()
    times = _lib.kernel_times()
    _lib.set_kernel_timing(False)
    a, _ = kernel_stats(times, 'tv_sqs')
    b, _ = kernel_stats(times, 'tv_sqs_update')
    nbytes = 16 * x.numel()
    return {'workload': 'TV-SQS prior on 256^3 (s1, s2, var) and the fused SQS-TV update',
            'kernel_ms': {'tv_sqs': round(a, 4), 'tv_sqs_update': round(b, 4)},
            'hbm_frac': {'tv_sqs': round(nbytes / (a / 1e3) / 1e9 / PEAK_HBM_GBS, 3) if a > 0 else None,
                         'tv_sqs_update': round(nbytes / (b / 1e3) / 1e9 / PEAK_HBM_GBS, 3) if b > 0 else None}}


def cone_batch_extra(steps):
    """The four uneven cone-beam volumes (128^3 @1, 192^3 @0.75, 256^3 @0.5,
    128x256x256 @1/0.75 mm; 720 views each, own jittered trajectories) in one
    batched plan: autotuned, forward and backward each one persistent launch,
    replayed from HIP graphs.  Compared with the same work through the
    single-volume device entry points, one volume after another."""
    import torch
    from ctprojector_amd.batched import ConeBatch, ConeJob
    from ctprojector_amd.phantom import CONE_SAMPLES, cone_workload
    from ctprojector_amd.projector.cupy import ct_projector, cone

    jobs, imgs, geos, projs = [], [], [], []
    for j in range(len(CONE_SAMPLES)):
        w, img, geo = cone_workload(j, 720)
        p = ct_projector()
        for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v'):
            setattr(p, k, getattr(w, k))
        jobs.append(ConeJob(p, geo))
        imgs.append(img)
        geos.append(geo)
        projs.append(p)
    nvu = sum(int(np.prod(im.shape)) * 720 for im in imgs)
    log('[bench] cone batch: plan + autotune')
    t0 = time.perf_counter()
    bt = ConeBatch(jobs, autotune=True)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    saved = [torch.from_numpy(im[0]).cuda() for im in imgs]
    for t, x in zip(bt.img, saved):
        t.copy_(x)
    gf = bt.capture('forward')
    gb = bt.capture('backward')
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    total = fp_t = bp_t = 0.0
    log('[bench] cone batch timed x%d' % steps)
    for _ in range(steps + 1):
        for t, x in zip(bt.img, saved):
            t.copy_(x)
        ev[0].record()
        gf.replay()
        ev[1].record()
        gb.replay()
        ev[2].record()
        torch.cuda.synchronize()
        if _ > 0:
            total += ev[0].elapsed_time(ev[2])
            fp_t += ev[0].elapsed_time(ev[1])
            bp_t += ev[1].elapsed_time(ev[2])
    ms = total / steps
    finite = all(bool(torch.isfinite(t).all()) for t in bt.img)
    cfg = bt.config()
    bt.close()
    del bt
    # sequential single-volume calls (device entry points, one sync per call)
    log('[bench] cone batch: sequential single-volume calls')
    dg = [[torch.from_numpy(a).cuda() for a in (g.det_center, g.det_u, g.det_v, g.src)] for g in geos]
    seq = 0.0
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for p, x, g in zip(projs, saved, dg):
            y = cone.siddon_fp_arbitrary(p, x[None], *g)
            _ = cone.siddon_bp_arbitrary(p, y, *g)
        torch.cuda.synchronize()
        seq = time.perf_counter() - t0
    return {
        'workload': 'batched cone Siddon FP+BP: 128^3 @1, 192^3 @0.75, 256^3 @0.5, 128x256x256 @1/0.75/0.75 mm, '
                    '720 views each (own jittered helical trajectory), 512x384 panel; one persistent launch per '
                    'direction, autotuned, HIP-graph replay',
        'ms_per_step': round(ms, 3),
        'fp_ms': round(fp_t / steps, 3),
        'bp_ms': round(bp_t / steps, 3),
        'value_gvu_s': round(2 * nvu / (ms / 1e3) / 1e9, 2),
        'sequential_ms_per_step': round(seq * 1e3, 3),
        'speedup_vs_sequential': round(seq * 1e3 / ms, 3),
        'plan_and_autotune_s': round(plan_s, 3),
        'autotuned': cfg,
        'finite': finite,
    }


SMALL_CONE_JOBS = [(n, dx, nview) for (n, dx) in ((48, 2.0), (64, 1.5), (80, 1.25), (96, 1.0))
                   for nview in (90, 120, 180)]


def cone_small_batch_extra(steps):
    """Twelve small uneven cone-beam jobs (48^3 @2 mm to 96^3 @1 mm, 90 to
    180 views each, own jittered trajectories, 64x48 panel at 4 mm): the
    regime where one job does not fill the GPU, in one batched plan
    (autotuned, one persistent launch per direction, HIP-graph replay),
    against the same work through the single-volume device entry points, one
    job after another."""
    import torch
    from ctprojector_amd.batched import ConeBatch, ConeJob
    from ctprojector_amd.phantom import Workload, circular_cone_geometry, ellipsoid_phantom
    from ctprojector_amd.projector.cupy import ct_projector, cone

    jobs, imgs, geos, projs = [], [], [], []
    for j, (n, dx, nview) in enumerate(SMALL_CONE_JOBS):
        w = Workload(n, n, n, dx, dx, dx, 64, 48, 4.0, 4.0, nview, dsd=1085.6, dso=595.0)
        geo = circular_cone_geometry(nview, 595.0, 1085.6, -10.0, 10.0, jitter=2.0, seed=1400 + j)
        img = ellipsoid_phantom(n, n, n, dx, dx, dx, seed=1500 + j)
        p = ct_projector()
        for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v'):
            setattr(p, k, getattr(w, k))
        jobs.append(ConeJob(p, geo))
        imgs.append(img)
        geos.append(geo)
        projs.append(p)
    nvu = sum(int(np.prod(im.shape)) * nv for im, (_, _, nv) in zip(imgs, SMALL_CONE_JOBS))
    log('[bench] small cone batch: plan + autotune (%d jobs)' % len(jobs))
    t0 = time.perf_counter()
    bt = ConeBatch(jobs, autotune=True)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    saved = [torch.from_numpy(im).cuda() for im in imgs]
    for t, x in zip(bt.img, saved):
        t.copy_(x)
    gf = bt.capture('forward')
    gb = bt.capture('backward')
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    best = fp_b = bp_b = 1e30
    log('[bench] small cone batch timed x%d' % steps)
    for _ in range(steps + 1):
        for t, x in zip(bt.img, saved):
            t.copy_(x)
        ev[0].record()
        gf.replay()
        ev[1].record()
        gb.replay()
        ev[2].record()
        torch.cuda.synchronize()
        best = min(best, ev[0].elapsed_time(ev[2]))
        fp_b = min(fp_b, ev[0].elapsed_time(ev[1]))
        bp_b = min(bp_b, ev[1].elapsed_time(ev[2]))
    batch_img = [t.clone() for t in bt.img]
    cfg = bt.config()
    bt.close()
    del bt
    log('[bench] small cone batch: sequential single-volume calls')
    dg = [[torch.from_numpy(a).cuda() for a in (g.det_center, g.det_u, g.det_v, g.src)] for g in geos]
    seq = 1e30
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs = []
        for p, x, g in zip(projs, saved, dg):
            y = cone.siddon_fp_arbitrary(p, x[None], *g)
            outs.append(cone.siddon_bp_arbitrary(p, y, *g))
        torch.cuda.synchronize()
        seq = min(seq, (time.perf_counter() - t0) * 1e3)
    worst = 0.0
    for a, b in zip(batch_img, outs):
        b = b.reshape(a.shape)
        worst = max(worst, float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b).clamp_min(1e-30)))
    return {
        'workload': 'batched cone Siddon FP+BP of %d small uneven jobs: 48^3 @2, 64^3 @1.5, 80^3 @1.25, 96^3 @1 mm '
                    'x 90/120/180 views (own jittered trajectories), 64x48 panel @4 mm; one persistent launch per '
                    'direction, autotuned, HIP-graph replay' % len(jobs),
        'ms_per_step': round(best, 3),
        'fp_ms': round(fp_b, 3),
        'bp_ms': round(bp_b, 3),
        'value_gvu_s': round(2 * nvu / (best / 1e3) / 1e9, 2),
        'sequential_ms_per_step': round(seq, 3),
        'speedup_vs_sequential': round(seq / best, 3),
        'max_rel_l2_vs_single_calls': worst,
        'plan_and_autotune_s': round(plan_s, 3),
        'autotuned': cfg,
    }

def dd_batch_extra(steps):
    """The sixteen uneven distance-driven jobs of phantom.DD_BATCH_SAMPLES
    (fan and parallel beams, 96^2 to 512^2, 1 to 8 slices, 128 to 984 views)
    in one DD batch: autotuned, a few launches per direction (one per kernel
    class), replayed from HIP graphs.  Compared with the same work as one
    persistent DD plan per job called one job after another (the same kernels,
    launched per job, no host work), and as the per-call device entry points
    (host tables and a synchronisation per call, as the reference's calls)."""
    import torch
    from ctprojector_amd.batched import DDBatch, DDJob
    from ctprojector_amd.ddplan import DDPlan
    from ctprojector_amd.phantom import dd_batch_workload
    from ctprojector_amd.projector.cupy import ct_projector, fan_equiangular as fan, parallel

    work = dd_batch_workload()
    projs = []
    for _, w, _ in work:
        p = ct_projector()
        for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v',
                  'dsd', 'dso'):
            setattr(p, k, getattr(w, k))
        projs.append(p)
    nvu = sum(int(np.prod(img.shape)) * w.nview for _, w, img in work)
    log('[bench] dd batch: plan + autotune (%d jobs)' % len(work))
    t0 = time.perf_counter()
    bt = DDBatch([DDJob(p, w.angles, fan=f) for p, (f, w, _) in zip(projs, work)], autotune=True)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    saved = [torch.from_numpy(img).cuda() for _, _, img in work]
    for t, x in zip(bt.img, saved):
        t.copy_(x)
    gf = bt.capture('forward')
    gb = bt.capture('backward')
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fp_t = bp_t = 0.0
    log('[bench] dd batch timed x%d' % steps)
    for i in range(steps + 1):
        for t, x in zip(bt.img, saved):
            t.copy_(x)
        ev[0].record()
        gf.replay()
        ev[1].record()
        gb.replay()
        ev[2].record()
        torch.cuda.synchronize()
        if i > 0:
            fp_t += ev[0].elapsed_time(ev[1])
            bp_t += ev[1].elapsed_time(ev[2])
    ms = (fp_t + bp_t) / steps
    finite = all(bool(torch.isfinite(t).all()) for t in bt.img)
    cfg = bt.config()
    bt.close()
    del bt
    log('[bench] dd batch: sequential plans')
    plans = [DDPlan(p, w.angles, batch=1, fan=f) for p, (f, w, _) in zip(projs, work)]
    prjs = [torch.empty(pl.prj_shape, device='cuda') for pl in plans]
    outs = [torch.empty(pl.img_shape, device='cuda') for pl in plans]
    seq_plans = 1e30
    for _ in range(steps + 1):
        torch.cuda.synchronize()
        ev[0].record()
        for pl, x, y, o in zip(plans, saved, prjs, outs):
            pl.fp(x, out=y)
            pl.bp(y, out=o)
        ev[1].record()
        torch.cuda.synchronize()
        seq_plans = min(seq_plans, ev[0].elapsed_time(ev[1]))
    for pl in plans:
        pl.close()
    log('[bench] dd batch: sequential per-call entry points')
    dang = [torch.from_numpy(w.angles).cuda() for _, w, _ in work]
    seq_calls = 1e30
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for p, (f, _, _), x, a in zip(projs, work, saved, dang):
            m = fan if f else parallel
            y = m.distance_driven_fp(p, x, a)
            _ = m.distance_driven_bp(p, y, a)
        torch.cuda.synchronize()
        seq_calls = min(seq_calls, (time.perf_counter() - t0) * 1e3)
    return {
        'workload': 'uneven DD batch: %d fan/parallel jobs (phantom.DD_BATCH_SAMPLES: 96^2 to 512^2, 1 to 8 slices, '
                    '128 to 984 views, 137 to 888 columns), FP then BP of each; autotuned, one launch per kernel '
                    'class and step, HIP-graph replay' % len(work),
        'ms_per_step': round(ms, 3),
        'fp_ms': round(fp_t / steps, 3),
        'bp_ms': round(bp_t / steps, 3),
        'value_gvu_s': round(2 * nvu / (ms / 1e3) / 1e9, 2),
        'sequential_plans_ms_per_step': round(seq_plans, 3),
        'speedup_vs_sequential_plans': round(seq_plans / ms, 3),
        'sequential_calls_ms_per_step': round(seq_calls, 3),
        'speedup_vs_sequential_calls': round(seq_calls / ms, 3),
        'plan_and_autotune_s': round(plan_s, 3),
        'autotuned': cfg,
        'finite': finite,
    }


def fbp_batch_extra(steps):
    """Filtered backprojection of the sixteen uneven jobs of
    phantom.DD_BATCH_SAMPLES (Hann window): one FBP batch (the ramp filters in
    one launch per FFT size, then the DD batch's FBP-weighted backprojection,
    a few launches per kernel class), replayed from a HIP graph; against the
    per-call entry points (ramp_filter then distance_driven_bp(is_fbp=True)
    per job, host tables and a synchronisation per call, as the reference's
    calls).  The sinograms are the jobs' forward projections."""
    import torch
    from ctprojector_amd.batched import DDJob, FBPBatch
    from ctprojector_amd.phantom import dd_batch_workload
    from ctprojector_amd.projector.cupy import ct_projector, fan_equiangular as fan, parallel

    work = dd_batch_workload()
    projs = []
    for _, w, _ in work:
        p = ct_projector()
        for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v',
                  'dsd', 'dso'):
            setattr(p, k, getattr(w, k))
        projs.append(p)
    nvu = sum(int(np.prod(img.shape)) * w.nview for _, w, img in work)
    dang = [torch.from_numpy(w.angles).cuda() for _, w, _ in work]
    sinos = []
    for p, (f, _, img), a in zip(projs, work, dang):
        m = fan if f else parallel
        sinos.append(m.distance_driven_fp(p, torch.from_numpy(img).cuda(), a))
    log('[bench] fbp batch: plan (%d jobs)' % len(work))
    fb = FBPBatch([DDJob(p, w.angles, fan=f) for p, (f, w, _) in zip(projs, work)], filter_type='hann',
                  autotune=True)
    for t, x in zip(fb.prj, sinos):
        t.copy_(x)
    g = fb.capture()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    best = filt = bp = 1e30
    log('[bench] fbp batch timed x%d' % steps)
    for _ in range(steps + 1):
        torch.cuda.synchronize()
        ev[0].record()
        g.replay()
        ev[1].record()
        torch.cuda.synchronize()
        best = min(best, ev[0].elapsed_time(ev[1]))
        # the two stages on their own (eager launches)
        ev[0].record()
        fb.filter.run()
        ev[1].record()
        fb.dd.backward()
        ev[2].record()
        torch.cuda.synchronize()
        filt = min(filt, ev[0].elapsed_time(ev[1]))
        bp = min(bp, ev[1].elapsed_time(ev[2]))
    batch_img = [t.clone() for t in fb.img]
    fb.close()
    del fb
    log('[bench] fbp batch: sequential per-call entry points')
    seq = 1e30
    worst = 0.0
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        refs = []
        for p, (f, _, _), y, a in zip(projs, work, sinos, dang):
            m = fan if f else parallel
            refs.append(m.distance_driven_bp(p, m.ramp_filter(p, y, 'hann'), a, is_fbp=True))
        torch.cuda.synchronize()
        seq = min(seq, (time.perf_counter() - t0) * 1e3)
    for a, b in zip(batch_img, refs):
        worst = max(worst, float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b).clamp_min(1e-30)))
    return {
        'workload': 'uneven FBP batch: %d fan/parallel jobs (phantom.DD_BATCH_SAMPLES), Hann ramp filter + '
                    'FBP-weighted DD backprojection of each; HIP-graph replay' % len(work),
        'ms_per_step': round(best, 3),
        'filter_ms': round(filt, 3),
        'bp_ms': round(bp, 3),
        'value_gvu_s': round(nvu / (best / 1e3) / 1e9, 2),
        'sequential_calls_ms_per_step': round(seq, 3),
        'speedup_vs_sequential_calls': round(seq / best, 3),
        'max_rel_l2_vs_single_calls': worst,
    }

def serving_extra(count=200, seed=1005, sub_count=32):
    """Served reconstruction stream: Poisson arrivals (200/s) of half fan-beam
    FBP (256^2-512^2 x 1-64 slices, 360-984 views) and half cone-beam SQS-TV
    (128^3-256^3, 10-50 iterations, 360-984 views), continuous batching with
    HBM-budget admission, against sequential plain launches of the same
    requests.  Then the same mix below saturation (arrivals at half the
    served rate) for its latency.  The sequential run also times the
    library's kernels: their summed duration over the run's wall time is how
    busy one request at a time keeps the GPU, so wall / kernel time bounds
    what overlapping requests alone can gain; batching gains come on top only
    where launches of several requests do less work than their sum (the cone
    kernels share traversals and chords across volumes of one geometry)."""
    import torch
    from ctprojector_amd.serving import Server, make_stream, run_sequential
    log('[bench] serving: %d requests (simulating sinograms)' % count)
    reqs = make_stream(count, seed=seed)
    free, _ = torch.cuda.mem_get_info()
    srv = Server(hbm_budget=int(0.5 * free), workers=4)
    srv.warm(reqs)
    log('[bench] serving: served run')
    st = srv.serve(reqs)
    for q in reqs:
        q.result = None
    log('[bench] serving: sequential run')
    kt = {}
    seq_s, seq_out = run_sequential(reqs, kernel_time=kt)
    del seq_out, reqs
    seq_rate = count / seq_s
    busy = kt['library_kernels_ms'] / 1e3 / seq_s
    out = {
        'workload': 'served stream of %d requests (Poisson 200/s, seed %d): fan FBP 256^2-512^2 x 1-64 slices and '
                    'cone SQS-TV 128^3-256^3 x 10-50 iterations, 360-984 views' % (count, seed),
        'recons_per_s': round(st.recons_per_s, 3),
        'latency_p50_s': round(st.latency_p50_s, 3),
        'latency_p99_s': round(st.latency_p99_s, 3),
        'makespan_s': round(st.makespan_s, 3),
        'sequential_recons_per_s': round(seq_rate, 3),
        'speedup_vs_sequential': round(st.recons_per_s / seq_rate, 3),
        'sequential_gpu_busy': round(busy, 3),
        'overlap_bound': round(1.0 / busy, 3) if busy > 0 else None,
        'peak_concurrent_requests': st.peak_concurrent,
        'batching': st.launches_batched,
    }
    if sub_count:
        rate = 0.5 * st.recons_per_s
        log('[bench] serving: %d requests at %.2f/s (half the served rate)' % (sub_count, rate))
        sub = make_stream(sub_count, rate=rate, seed=seed + 100)
        srv.warm(sub)
        s2 = srv.serve(sub)
        out['below_saturation'] = {'requests': sub_count, 'arrivals_per_s': round(rate, 3),
                                   'recons_per_s': round(s2.recons_per_s, 3),
                                   'latency_p50_s': round(s2.latency_p50_s, 3),
                                   'latency_p99_s': round(s2.latency_p99_s, 3)}
    return out


def fft_n(nu):
    """FFT length of the ramp filter: the smallest power of two >= 2nu - 1."""
    n = 8
    while n < 2 * nu - 1:
        n *= 2
    return n


def fbp_extra(steps, w, img, outputs=None):
    import torch
    from ctprojector_amd import _lib
    from ctprojector_amd.projector.cupy import ct_projector, fan_equiangular as fan

    p = ct_projector()
    for k in ('nx', 'ny', 'nz', 'dx', 'dy', 'dz', 'cx', 'cy', 'cz', 'nu', 'nv', 'du', 'dv', 'off_u', 'off_v',
              'dsd', 'dso', 'nview'):
        setattr(p, k, getattr(w, k))
    dev = torch.device('cuda', 0)
    ang_d = torch.from_numpy(w.angles).to(dev)
    prj = fan.distance_driven_fp(p, torch.from_numpy(img).to(dev), ang_d)
    fprj = fan.ramp_filter(p, prj, 'hann')
    _ = fan.fbp_bp(p, fprj, ang_d)
    torch.cuda.synchronize()
    _lib.set_kernel_timing(True)
    _lib.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        fprj = fan.ramp_filter(p, prj, 'hann')
        rec = fan.fbp_bp(p, fprj, ang_d)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if outputs is not None:
        outputs['fl'] = fprj.cpu().numpy()
        outputs['rec'] = rec.cpu().numpy()
    times = _lib.kernel_times()
    _lib.set_kernel_timing(False)
    nvu = w.nx * w.ny * w.nz * w.nview
    nsamp = w.nu * w.nv * w.nview
    rf_ms, _ = kernel_stats(times, 'ramp_filter')
    bp_ms, _ = kernel_stats(times, 'fan_fbp_bp')
    bp_traffic = read_traffic('fan_fbp_bp')
    return {
        'workload': 'fan ramp_filter (Hann) + fbp_bp, same volume/scan',
        'ms_per_step': round((t1 - t0) / steps * 1e3, 3),
        'kernel_ms': {'ramp_filter': round(rf_ms, 3), 'fan_fbp_bp': round(bp_ms, 3)},
        # the ramp filter streams 8 B per detector sample (read + write); its
        # FFT convolution (length N = 2^p >= 2nu - 1, two rows per complex
        # transform, a forward and an inverse) does ~5 N log2 N flops per row
        'ramp_filter_hbm_frac': round(8 * nsamp / (rf_ms / 1e3) / 1e9 / PEAK_HBM_GBS, 4) if rf_ms else None,
        'ramp_filter_fft_tflop_s': round(5.0 * fft_n(w.nu) * np.log2(fft_n(w.nu)) * nsamp / w.nu / (rf_ms / 1e3)
                                         / 1e12, 2) if rf_ms else None,
        'fbp_bp_gvu_s': round(nvu / (bp_ms / 1e3) / 1e9, 2) if bp_ms else None,
        # HBM traffic of the backprojection (counters, L2-miss bytes) over its
        # time; the 4 B per voxel-update ratio beside it is a throughput ratio
        # (caches serve most of those reads), not a bound
        'fbp_bp_hbm_frac': round(bp_traffic / (bp_ms / 1e3) / 1e9 / PEAK_HBM_GBS, 4) if bp_ms and bp_traffic else None,
        'fbp_bp_algorithmic_4B_ratio_to_peak': round(4 * nvu / (bp_ms / 1e3) / 1e9 / PEAK_HBM_GBS, 4)
        if bp_ms else None,
    }


def oracle_parity(workload, outputs):
    """Relative L2 of the GPU outputs of one benchmark workload against the
    FP64 oracle over the whole workload (all views, full volume), computed on
    the host beside the CPU baseline by oracle/cpu_baseline.py --check from
    the float32 outputs saved to a scratch directory."""
    import shutil
    import subprocess
    import tempfile
    d = tempfile.mkdtemp(prefix='ctp_parity_')
    try:
        for k, v in outputs.items():
            np.save(os.path.join(d, k + '.npy'), v)
        log('[bench] oracle parity (%s, FP64, whole workload)' % workload)
        r = subprocess.run([sys.executable, os.path.join(ROOT, 'oracle', 'cpu_baseline.py'), '--workload', workload,
                            '--check', d], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, cwd=ROOT,
                           timeout=900)
        if r.returncode != 0:
            return {'error': 'oracle check failed (%d): %s' % (r.returncode, r.stderr.strip()[-500:])}
        return json.loads(r.stdout.strip().splitlines()[-1])
    finally:
        shutil.rmtree(d, ignore_errors=True)


def calibrate_copy():
    """Three plain device copies of COPY_BYTES (the counter calibration)."""
    import torch
    a = torch.ones(COPY_BYTES // 4, device='cuda')
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log('[bench] calibration copies: 3 x %d bytes' % COPY_BYTES)
    del a, b


def short_name(kname):
    """Kernel symbol without namespace and parameter list, template arguments
    kept (instances are reported separately)."""
    return kname.split('(')[0].replace('void ', '').replace('ctp::', '').strip()


# timing names (HIP-event scopes in the library) -> kernel symbols in rocprofv3 output
# (the instances the benchmark workloads launch)
TIMED_KERNELS = {'dd_fan_fp': 'dd_fp_split_kernel<32, 4, true>', 'dd_fan_bp': 'dd_bp_split_kernel<32, 4, true>',
                 'siddon_cone_fp': 'siddon_cone_fp_col_kernel<1>',
                 'siddon_cone_bp': 'siddon_cone_bp_col_kernel<2, 2, 8, 1>',
                 'fan_fbp_bp': 'fan_fbp_bp_stage_kernel<32, 4, true>', 'ramp_filter': 'ramp_fft16_kernel<2048>'}


def roofline(dom, dom_ms, nvu, compulsory_bytes, times):
    """Roofline entry of the dominant kernel.  'achieved' is the HBM traffic of
    one launch (rocprofv3 FETCH_SIZE scaled by a device copy's calibration,
    plus WRITE_SIZE: bytes that missed the L2, from the committed counter
    summary) over the launch time measured here with HIP events on the
    launching stream.  The survey's 4 B per voxel-update figure (the voxel
    operand streamed once per view) is reported beside it as a throughput
    ratio: caches serve most of those reads, so it exceeds the HBM peak and is
    not a bound."""
    traffic = read_traffic(dom)
    alg_bytes = 4.0 * nvu
    alg_gbs = alg_bytes / (dom_ms / 1e3) / 1e9 if dom_ms > 0 else 0.0
    if traffic and dom_ms > 0:
        achieved, basis = traffic / (dom_ms / 1e3) / 1e9, 'counters'
    else:
        achieved, basis = alg_gbs, 'algorithmic (no counter summary lists this kernel)'
    return {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
            'frac': round(achieved / PEAK_HBM_GBS, 4), 'traffic': traffic, 'basis': basis,
            'traffic_source': '%s: FETCH_SIZE x copy calibration + WRITE_SIZE per launch (L2-miss bytes; the '
                              '256 MB Infinity Cache may serve part of them)' % os.path.relpath(TRAFFIC_FILE, ROOT),
            'kernel': dom, 'kernel_symbol': TIMED_KERNELS.get(dom, dom), 'kernel_ms': round(dom_ms, 3),
            'compulsory_frac': round(compulsory_bytes / (dom_ms / 1e3) / 1e9 / PEAK_HBM_GBS, 4) if dom_ms else None,
            'algorithmic_4B_per_voxel_update': {'bytes_per_launch': int(alg_bytes), 'gb_s': round(alg_gbs, 1),
                                                'ratio_to_peak': round(alg_gbs / PEAK_HBM_GBS, 4)},
            'kernels_ms': {k: round(v[0] / max(v[1], 1), 3) for k, v in times.items()}}


def read_traffic(name):
    """HBM bytes per launch (FETCH + WRITE, calibrated) of the kernel timed as
    ``name``, from the committed counter summary, or None."""
    kernel = TIMED_KERNELS.get(name, name)
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                cells = [c.strip().strip('`') for c in line.strip().strip('|').split('|')]
                if len(cells) >= 4 and cells[0] == kernel:
                    return int(float(cells[3]) * (1 << 20))
    except OSError:
        pass
    return None


def summarize(trace_dir, fetch_dir=None, write_dir=None, hits_dir=None):
    """Markdown tables from rocprofv3 result databases (run_results.db)."""
    import sqlite3

    def db(d):
        return sqlite3.connect(os.path.join(d, 'run_results.db')).cursor()

    out = []
    cur = db(trace_dir)
    rows = list(cur.execute('select name, total_calls, total_duration, average, percentage from top_kernels'))
    out.append('| kernel | calls | total ms | average us | % of GPU time |')
    out.append('|---|---:|---:|---:|---:|')
    for name, calls, tot, avg, pct in rows:
        # top_kernels durations are in microseconds
        out.append('| `%s` | %d | %.3f | %.1f | %.2f |' % (short_name(name), calls, tot / 1e3, avg, pct))
    if fetch_dir and write_dir:
        def counters(d, cname):
            res = {}
            for k, v in db(d).execute('select kernel_name, sum(value) from counters_collection '
                                      'where counter_name = ? group by dispatch_id', (cname,)):
                res.setdefault(short_name(k), []).append(v)
            return res
        fetch = counters(fetch_dir, 'FETCH_SIZE')
        write = counters(write_dir, 'WRITE_SIZE')
        cal_f = max(fetch.get('__amd_rocclr_copyBuffer', [0]))
        cal_w = max(write.get('__amd_rocclr_copyBuffer', [0]))
        kib = COPY_BYTES / 1024.0
        sf = kib / cal_f if cal_f else 1.0
        sw = kib / cal_w if cal_w else 1.0
        hits = counters(hits_dir, 'TCC_HIT_sum') if hits_dir else {}
        miss = counters(hits_dir, 'TCC_MISS_sum') if hits_dir else {}
        out.append('')
        out.append('Copy calibration: a %d MiB device copy reads FETCH_SIZE %.1f KiB and WRITE_SIZE %.1f KiB;'
                   ' counters below are scaled by %.4f (fetch) and %.4f (write).'
                   % (COPY_BYTES >> 20, cal_f, cal_w, sf, sw))
        out.append('')
        out.append('| kernel | fetch MiB/launch | write MiB/launch | HBM MiB/launch | L2 hit % | launches |')
        out.append('|---|---:|---:|---:|---:|---:|')
        for k in sorted(fetch, key=lambda k: -max(fetch[k])):
            if k == '__amd_rocclr_copyBuffer':
                continue
            f = sum(fetch[k]) / len(fetch[k]) * sf / 1024
            w = (sum(write.get(k, [0])) / max(len(write.get(k, [])), 1)) * sw / 1024
            h, m = sum(hits.get(k, [0])), sum(miss.get(k, [0]))
            hr = '%.1f' % (100.0 * h / (h + m)) if h + m else '-'
            out.append('| `%s` | %.1f | %.1f | %.1f | %s | %d |' % (k, f, w, f + w, hr, len(fetch[k])))
    print('\n'.join(out))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-extra', action='store_true')
    ap.add_argument('--no-serving', action='store_true', help='skip the served-stream extra')
    ap.add_argument('--no-parity', action='store_true', help='skip the FP64 oracle check of the outputs')
    ap.add_argument('--extras', default='cone,cone_batch,cone_small_batch,dd_batch,fbp_batch,fbp,unrolled_sqs_tv,tv_prior,serving',
                    help='comma-separated extras to run (default: all)')
    ap.add_argument('--serving-requests', type=int, default=200)
    ap.add_argument('--serving-below-saturation', type=int, default=32,
                    help='requests of the run below saturation (0: skip it)')
    ap.add_argument('--calibrate-copy', action='store_true')
    ap.add_argument('--tune', action='append', default=[], metavar='KEY=VALUE',
                    help='library launch-shape knob (ctpSetTuning), repeatable')
    ap.add_argument('--summarize', metavar='TRACE_DIR')
    ap.add_argument('--fetch', metavar='DIR')
    ap.add_argument('--write', metavar='DIR')
    ap.add_argument('--hits', metavar='DIR')
    args = ap.parse_args(argv)
    if args.summarize:
        return summarize(args.summarize, args.fetch, args.write, args.hits)
    if args.gpus != 1:
        print('bench.py measures one GPU only: the projector path runs on a single MI355X here and multi-GPU '
              'sharding (views for FP, z-slabs for BP) is not part of this build; run with --gpus 1.',
              file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print('bench.py needs a HIP device', file=sys.stderr)
        return 1

    if args.tune:
        from ctprojector_amd import _lib
        for kv in args.tune:
            k, v = kv.split('=')
            _lib.set_tuning(k, int(v))
    if args.calibrate_copy:
        calibrate_copy()
    r = fan_workload_bench(args.steps, args.warmup)
    w, nvu, times = r['w'], r['nvu'], r['times']
    fp_ms, fp_n = kernel_stats(times, 'dd_fan_fp')
    bp_ms, bp_n = kernel_stats(times, 'dd_fan_bp')
    dom, dom_ms = ('dd_fan_fp', fp_ms) if fp_ms >= bp_ms else ('dd_fan_bp', bp_ms)
    line = {
        'metric': 'FP+BP Gray-voxel updates/s & % HBM roofline; served recons/s + p50/p99 latency',
        'value': round(2 * nvu / (r['ms_step'] / 1e3) / 1e9, 2),
        'unit': 'GVU/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(r['ms_step'], 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded ellipsoid phantom, rng 1002)',
        'config': {'workload': 'fan-beam equiangular distance-driven FP+BP, volume 1x32x512x512 @0.7 mm, '
                               '984 views/360 deg, 888x32 detector, dso 595, dsd 1085.6',
                   'batch': 1, 'nx': 512, 'ny': 512, 'nz': 32, 'nview': 984, 'nu': 888, 'nv': 32},
        'roofline': roofline(dom, dom_ms, nvu, 4 * (w.nx * w.ny * w.nz + w.nu * w.nv * w.nview), times),
        'finite': r['finite'],
    }
    outputs = {'fan': dict(r['outputs']), 'cone': {}}
    if not args.no_extra:
        extra = {}
        want = set(args.extras.split(','))
        runs = [('cone', lambda: cone_extra(2, cpu=not args.no_cpu_baseline,
                                            outputs=None if args.no_parity else outputs['cone'])),
                ('cone_batch', lambda: cone_batch_extra(3)),
                ('cone_small_batch', lambda: cone_small_batch_extra(5)),
                ('dd_batch', lambda: dd_batch_extra(5)),
                ('fbp_batch', lambda: fbp_batch_extra(5)),
                ('fbp', lambda: fbp_extra(3, w, r['img'], outputs=None if args.no_parity else outputs['fan'])),
                ('unrolled_sqs_tv', lambda: unrolled_sqs_extra(3)),
                ('tv_prior', lambda: tv_prior_extra(5))]
        if not args.no_serving:
            runs.append(('serving', lambda: serving_extra(args.serving_requests,
                                                          sub_count=args.serving_below_saturation)))
        for name, fn in runs:
            if name not in want:
                continue
            try:
                extra[name] = fn()
            except Exception as e:  # report, do not hide
                extra[name] = {'error': repr(e)}
        line['extra'] = extra
    if not args.no_cpu_baseline:
        line['cpu_baseline'] = run_cpu_baseline('fan')
    if not args.no_parity:
        par = {'tolerance': 1e-4, 'fan': oracle_parity('fan', outputs['fan'])}
        if outputs['cone']:
            par['cone'] = oracle_parity('cone', outputs['cone'])
        par['note'] = ('DD backprojection main axis decided per view in FP64 (the reference: FP32 __cosf/__sinf); '
                       'at the 4 diagonal views of 984 an FP32 decision can swap the axis of whole views, up to '
                       '0.8% rel-L2 against an FP32-decided result')
        line['parity_rel_l2'] = par
    print(json.dumps(line), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
