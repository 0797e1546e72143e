"""Batched projection of uneven jobs: many volumes or slice stacks of different
sizes, each with its own geometry, forward- and back-projected by a few
launches per direction, with launch shapes chosen by the library's autotuner
and the launches capturable in a HIP graph.

* ``ConeBatch``: cone-beam Siddon jobs, each with its own trajectory and
  detector; one persistent launch forward and one backward (each block runs the
  backprojector compiled for its job's candidate window; ``ctpConePlan*`` of
  libprojector.so).
* ``DDBatch``: fan-beam (equiangular) and parallel-beam distance-driven jobs,
  each with its own grid, detector, angles and batch count (``ctpDDBatch*``);
  the jobs whose kernels are the same template instance share a launch.
* ``RampFilterBatch``: the ramp filters of uneven fan / parallel jobs
  (``ctpFilterBatch*``), one launch per transform size; ``FBPBatch`` chains it
  into a DDBatch with FBP weights: filtered backprojection of uneven jobs.

The reference projects a batch by looping over its items on the host with a
launch and a device synchronisation per item (cuda/siddonCone.cu:76-135); its
batches must share one geometry and one size, and its DD entry points take one
geometry per call (cuda/distanceDriven2D.cu:109-634).

    jobs = [ConeJob(projector_a, geo_a), ConeJob(projector_b, geo_b), ...]
    batch = ConeBatch(jobs, autotune=True)
    batch.img[0].copy_(volume_a); ...
    batch.forward()                       # batch.prj[j] = A_j batch.img[j]
    g = batch.capture('forward')          # HIP graph of the same launches
    g.replay()

Buffers ``img[j]`` ([nz, ny, nx]) and ``prj[j]`` ([nview, nv, nu]) are device
tensors owned by the batch and bound to the plan; forward overwrites prj,
backward overwrites img.  Calls are asynchronous on torch's current stream.
"""

import ctypes
from ctypes import c_float, c_int, c_size_t, c_void_p
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np
import torch

from . import _lib


class _CJob(ctypes.Structure):
    _fields_ = [('nx', c_size_t), ('ny', c_size_t), ('nz', c_size_t),
                ('dx', c_float), ('dy', c_float), ('dz', c_float), ('cx', c_float), ('cy', c_float), ('cz', c_float),
                ('nu', c_size_t), ('nv', c_size_t), ('nview', c_size_t),
                ('du', c_float), ('dv', c_float), ('off_u', c_float), ('off_v', c_float),
                ('det_center', c_void_p), ('det_u', c_void_p), ('det_v', c_void_p), ('src', c_void_p),
                ('img', c_void_p), ('prj', c_void_p)]


@dataclass
class ConeJob:
    """One volume: ``projector`` carries nx..nz, dx..cz, nu, nv, du, dv, off_u,
    off_v (the reference's ct_projector attributes); ``geometry`` has
    det_center, det_u, det_v, src as [nview, 3] arrays (x, y, z)."""
    projector: object
    geometry: object


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class ConeBatch:
    def __init__(self, jobs: Sequence[ConeJob], autotune: bool = False, device=None):
        self.device = torch.device('cuda', torch.cuda.current_device()) if device is None else torch.device(device)
        self.img: List[torch.Tensor] = []
        self.prj: List[torch.Tensor] = []
        self._host = []
        cj = (_CJob * len(jobs))()
        for k, job in enumerate(jobs):
            p, geo = job.projector, job.geometry
            arrs = [_f32(geo.det_center), _f32(geo.det_u), _f32(geo.det_v), _f32(geo.src)]
            nview = arrs[0].shape[0]
            self._host.append(arrs)
            img = torch.zeros([p.nz, p.ny, p.nx], dtype=torch.float32, device=self.device)
            prj = torch.zeros([nview, p.nv, p.nu], dtype=torch.float32, device=self.device)
            self.img.append(img)
            self.prj.append(prj)
            cj[k] = _CJob(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, p.cx, p.cy, p.cz, p.nu, p.nv, nview,
                          p.du, p.dv, p.off_u, p.off_v, *[a.ctypes.data for a in arrs],
                          img.data_ptr(), prj.data_ptr())
        self._plan = c_void_p()
        torch.cuda.synchronize(self.device)
        with torch.cuda.device(self.device):
            err = _lib.module().ctpConePlanCreate(ctypes.byref(self._plan), cj, c_int(len(jobs)),
                                                  c_int(1 if autotune else 0))
        if err != 0:
            raise RuntimeError('ctpConePlanCreate failed (%d): %s' % (err, _lib.last_error_message()))

    def _stream(self):
        return c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def forward(self) -> None:
        """prj[j] = A_j img[j] for every job (asynchronous on the current stream)."""
        with torch.cuda.device(self.device):
            err = _lib.module().ctpConePlanForward(self._plan, self._stream())
        if err != 0:
            raise RuntimeError('ctpConePlanForward failed (%d): %s' % (err, _lib.last_error_message()))

    def backward(self) -> None:
        """img[j] = A_j^T prj[j] for every job (asynchronous on the current stream)."""
        with torch.cuda.device(self.device):
            err = _lib.module().ctpConePlanBackward(self._plan, self._stream())
        if err != 0:
            raise RuntimeError('ctpConePlanBackward failed (%d): %s' % (err, _lib.last_error_message()))

    def config(self) -> dict:
        """Launch shapes in use (after autotuning): forward tile width (0: the
        column-driven form), XCD block order, backprojection z-run."""
        a, b, c = c_int(), c_int(), c_int()
        _lib.module().ctpConePlanConfig(self._plan, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {'cone_fp_tile_u': a.value, 'cone_fp_xcd_order': b.value, 'cone_bp_zrun': c.value}

    def capture(self, which: str = 'forward') -> torch.cuda.CUDAGraph:
        """A HIP graph of ``forward``, ``backward`` or ``both`` (forward then
        backward); replay() reruns it on the bound buffers."""
        steps = {'forward': [self.forward], 'backward': [self.backward],
                 'both': [self.forward, self.backward]}[which]
        for f in steps:                      # warm-up outside the capture
            f()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for f in steps:
                f()
        return g

    def close(self) -> None:
        if getattr(self, '_plan', None) is not None and self._plan.value:
            torch.cuda.synchronize(self.device)
            _lib.module().ctpConePlanDestroy(self._plan)
            self._plan = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _DJob(ctypes.Structure):
    _fields_ = [('fan', c_int), ('typeProjector', c_int),
                ('nBatches', c_size_t), ('nx', c_size_t), ('ny', c_size_t), ('nz', c_size_t),
                ('dx', c_float), ('dy', c_float), ('dz', c_float), ('cx', c_float), ('cy', c_float), ('cz', c_float),
                ('nu', c_size_t), ('nv', c_size_t), ('nview', c_size_t),
                ('da', c_float), ('dv', c_float), ('off_a', c_float), ('off_v', c_float),
                ('dsd', c_float), ('dso', c_float),
                ('deg', c_void_p), ('img', c_void_p), ('prj', c_void_p)]


@dataclass
class DDJob:
    """One distance-driven job: ``projector`` carries nx..nz, dx..cz, nu, nv,
    du, dv, off_u, off_v and, for the fan beam, dsd and dso (the reference's
    ct_projector attributes; du is the detector spacing at dsd, as for
    fan_equiangular.distance_driven_fp); ``angles`` are the view angles in
    radians; ``batch`` items share the geometry; ``is_fbp`` (1),
    ``branchless`` (2) and ``force_fbp`` (4) are the typeProjector bits of
    the reference's DD entry points."""
    projector: object
    angles: object
    fan: bool = True
    batch: int = 1
    is_fbp: bool = False
    branchless: bool = False
    force_fbp: bool = False


class DDBatch:
    """Buffers ``img[j]`` ([batch, nz, ny, nx]) and ``prj[j]`` ([batch, nview,
    nv, nu]) are device tensors owned by the batch and bound to it; forward
    writes prj[j] = A_j img[j] (fan_equiangular / parallel
    distance_driven_fp), backward writes img[j] = the reference's
    backprojection of prj[j] (distance_driven_bp with the job's bits).  Every
    job's result is that of its single call.  Calls are asynchronous on
    torch's current stream; calls on one batch must not overlap in time (they
    share its scratch and work counters), so issue them on one stream."""

    def __init__(self, jobs: Sequence[DDJob], autotune: bool = False, device=None):
        self.device = torch.device('cuda', torch.cuda.current_device()) if device is None else torch.device(device)
        self.img: List[torch.Tensor] = []
        self.prj: List[torch.Tensor] = []
        self._host = []
        cj = (_DJob * len(jobs))()
        for k, job in enumerate(jobs):
            p = job.projector
            ang = job.angles
            if isinstance(ang, torch.Tensor):
                ang = ang.detach().cpu().numpy()
            ang = _f32(ang).ravel()
            self._host.append(ang)
            nview, b = int(ang.shape[0]), int(job.batch)
            img = torch.zeros([b, p.nz, p.ny, p.nx], dtype=torch.float32, device=self.device)
            prj = torch.zeros([b, nview, p.nv, p.nu], dtype=torch.float32, device=self.device)
            self.img.append(img)
            self.prj.append(prj)
            fan = bool(job.fan)
            typ = (1 if job.is_fbp else 0) + (2 if job.branchless else 0) + (4 if job.force_fbp else 0)
            cj[k] = _DJob(1 if fan else 0, typ, b, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, p.cx, p.cy, p.cz,
                          p.nu, p.nv, nview, p.du / p.dsd if fan else p.du, p.dv, p.off_u, p.off_v,
                          p.dsd if fan else 0.0, p.dso if fan else 0.0, ang.ctypes.data, img.data_ptr(),
                          prj.data_ptr())
        self._plan = c_void_p()
        torch.cuda.synchronize(self.device)
        with torch.cuda.device(self.device):
            err = _lib.module().ctpDDBatchCreate(ctypes.byref(self._plan), cj, c_int(len(jobs)),
                                                 c_int(1 if autotune else 0))
        if err != 0:
            raise RuntimeError('ctpDDBatchCreate failed (%d): %s' % (err, _lib.last_error_message()))

    def _stream(self):
        return c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def forward(self) -> None:
        """prj[j] = A_j img[j] for every job (asynchronous on the current stream)."""
        with torch.cuda.device(self.device):
            err = _lib.module().ctpDDBatchForward(self._plan, self._stream())
        if err != 0:
            raise RuntimeError('ctpDDBatchForward failed (%d): %s' % (err, _lib.last_error_message()))

    def backward(self) -> None:
        """img[j] = the reference's backprojection of prj[j] for every job."""
        with torch.cuda.device(self.device):
            err = _lib.module().ctpDDBatchBackward(self._plan, self._stream())
        if err != 0:
            raise RuntimeError('ctpDDBatchBackward failed (%d): %s' % (err, _lib.last_error_message()))

    def config(self) -> dict:
        """Launch shapes in use (after autotuning) and kernel launches per call."""
        v = [c_int() for _ in range(8)]
        err = _lib.module().ctpDDBatchConfig(self._plan, *[ctypes.byref(x) for x in v])
        if err != 0:
            raise RuntimeError('ctpDDBatchConfig failed (%d): %s' % (err, _lib.last_error_message()))
        keys = ('dd_fp_lanes_per_column', 'dd_bp_lanes_per_pixel', 'fp_persistent', 'bp_persistent',
                'fp_streams', 'bp_streams', 'fp_launches', 'bp_launches')
        return {k: x.value for k, x in zip(keys, v)}

    def capture(self, which: str = 'forward') -> torch.cuda.CUDAGraph:
        """A HIP graph of ``forward``, ``backward`` or ``both``."""
        steps = {'forward': [self.forward], 'backward': [self.backward],
                 'both': [self.forward, self.backward]}[which]
        for f in steps:
            f()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for f in steps:
                f()
        return g

    def close(self) -> None:
        if getattr(self, '_plan', None) is not None and self._plan.value:
            torch.cuda.synchronize(self.device)
            _lib.module().ctpDDBatchDestroy(self._plan)
            self._plan = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _FJob(ctypes.Structure):
    _fields_ = [('fan', c_int), ('typeFilter', c_int),
                ('nBatches', c_size_t), ('nu', c_size_t), ('nv', c_size_t), ('nview', c_size_t),
                ('da', c_float), ('off_a', c_float), ('dsd', c_float), ('dso', c_float),
                ('prj', c_void_p), ('fprj', c_void_p)]


@dataclass
class FilterJob:
    """One ramp-filter job: ``projector`` carries nu, nv, du, off_u and, for
    the fan beam, dsd and dso (as for fan_equiangular / parallel
    ramp_filter); ``nview`` views; ``batch`` items share the geometry;
    ``filter_type`` as for ramp_filter ('hann', 'hamming', 'cosine', else
    Ram-Lak)."""
    projector: object
    nview: int
    fan: bool = True
    batch: int = 1
    filter_type: str = 'hann'


class RampFilterBatch:
    """Ramp filters of uneven jobs (``ctpFilterBatch*``): ``fprj[j]`` = the
    job's fan_equiangular / parallel ramp_filter of ``prj[j]`` (both
    [batch, nview, nv, nu] device tensors bound to the batch; ``outputs``
    may supply the fprj tensors), in one launch per transform size (rows of
    up to 2048 columns) plus one per job of 2049 to 5000 columns (the single
    call's direct convolution).  ``run`` is asynchronous on torch's current
    stream and can be captured into a HIP graph."""

    def __init__(self, jobs: Sequence[FilterJob], device=None, outputs: Sequence[torch.Tensor] = None):
        from .projector.geometry import filter_code
        self.device = torch.device('cuda', torch.cuda.current_device()) if device is None else torch.device(device)
        self.prj: List[torch.Tensor] = []
        self.fprj: List[torch.Tensor] = []
        cj = (_FJob * len(jobs))()
        for k, job in enumerate(jobs):
            p = job.projector
            shape = [int(job.batch), int(job.nview), int(p.nv), int(p.nu)]
            prj = torch.zeros(shape, dtype=torch.float32, device=self.device)
            if outputs is not None:
                fprj = outputs[k]
                if list(fprj.shape) != shape or fprj.dtype != torch.float32 or not fprj.is_contiguous() \
                        or fprj.device != self.device:
                    raise ValueError('filter output %d: a contiguous float32 %s tensor on %s is needed'
                                     % (k, shape, self.device))
            else:
                fprj = torch.zeros(shape, dtype=torch.float32, device=self.device)
            self.prj.append(prj)
            self.fprj.append(fprj)
            fan = bool(job.fan)
            cj[k] = _FJob(1 if fan else 0, filter_code(job.filter_type), shape[0], shape[3], shape[2], shape[1],
                          p.du / p.dsd if fan else p.du, p.off_u, p.dsd if fan else 0.0, p.dso if fan else 0.0,
                          prj.data_ptr(), fprj.data_ptr())
        self._plan = c_void_p()
        torch.cuda.synchronize(self.device)
        with torch.cuda.device(self.device):
            err = _lib.module().ctpFilterBatchCreate(ctypes.byref(self._plan), cj, c_int(len(jobs)))
        if err != 0:
            raise RuntimeError('ctpFilterBatchCreate failed (%d): %s' % (err, _lib.last_error_message()))

    def run(self) -> None:
        """fprj[j] = the ramp filter of prj[j] for every job (asynchronous)."""
        with torch.cuda.device(self.device):
            err = _lib.module().ctpFilterBatchRun(self._plan,
                                                  c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        if err != 0:
            raise RuntimeError('ctpFilterBatchRun failed (%d): %s' % (err, _lib.last_error_message()))

    def close(self) -> None:
        if getattr(self, '_plan', None) is not None and self._plan.value:
            torch.cuda.synchronize(self.device)
            _lib.module().ctpFilterBatchDestroy(self._plan)
            self._plan = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FBPBatch:
    """Filtered backprojection of uneven fan / parallel jobs: ``prj[j]`` in,
    ``img[j]`` out ([batch, nz, ny, nx]).  ``run`` filters every job
    (RampFilterBatch, one launch per transform size) into the sinograms of a
    DDBatch whose jobs carry the FBP weights (typeProjector bit 0, as
    distance_driven_bp(..., is_fbp=True)) and backprojects them (one launch
    per kernel class): each job's image equals ramp_filter followed by
    distance_driven_bp(is_fbp=True) on its own.  ``jobs`` are DDJob records
    (their is_fbp flag is set here); ``filter_type`` applies to every job.
    The DD batch is autotuned by default (lanes, block distribution and side
    streams; untuned, sixteen mostly single-slice jobs ran 3.02 against 2.86
    ms per step)."""

    def __init__(self, jobs: Sequence[DDJob], filter_type: str = 'hann', autotune: bool = True, device=None):
        jobs = [DDJob(j.projector, j.angles, fan=j.fan, batch=j.batch, is_fbp=True, branchless=j.branchless,
                      force_fbp=j.force_fbp) for j in jobs]
        self.dd = DDBatch(jobs, autotune=autotune, device=device)
        self.device = self.dd.device
        fjobs = [FilterJob(j.projector, int(t.shape[1]), fan=j.fan, batch=j.batch, filter_type=filter_type)
                 for j, t in zip(jobs, self.dd.prj)]
        self.filter = RampFilterBatch(fjobs, device=self.device, outputs=self.dd.prj)
        self.prj = self.filter.prj
        self.img = self.dd.img

    def run(self) -> None:
        """img[j] = FBP of prj[j] for every job (asynchronous on the current stream)."""
        self.filter.run()
        self.dd.backward()

    def capture(self) -> torch.cuda.CUDAGraph:
        """A HIP graph of ``run``."""
        self.run()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        return g

    def close(self) -> None:
        self.filter.close()
        self.dd.close()
