"""Reconstruction serving on one GPU: continuous batching of mixed requests.

Requests are fan-beam FBP reconstructions and cone-beam SQS-TV
reconstructions of different sizes arriving over time.  The server

* groups requests by geometry class (the scanner geometry they share): FBP
  requests of one class are reconstructed together in one filter launch and
  one backprojection launch, whatever their slice counts (every slice is a
  batch item); SQS-TV requests of one class iterate together as one batch, new
  arrivals joining and finished requests leaving at iteration boundaries
  (continuous batching);
* admits a request only when its device-memory footprint fits the HBM budget
  next to the requests already admitted, and keeps sinograms, volumes and SQS
  state resident on the device until the request completes;
* runs classes concurrently on worker threads, each with its own HIP stream
  (library calls follow the thread's stream), so launches of different
  classes overlap on the device;
* records arrival, admission and completion times; latency = completion -
  arrival.

``run_sequential`` processes the same requests one at a time with plain
launches (one request per call, no batching, no overlap): the baseline.

The reconstruction arithmetic is the library's: FBP = ramp filter +
fbp_bp (fan_equiangular), SQS-TV = sqs_tv_one_step with the Siddon cone
projector (recon/sqs_algorithms.py).  A request's result does not depend on
the requests it is batched with.
"""

import math
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from .phantom import circular_cone_geometry, ellipsoid_phantom


# --------------------------------------------------------------- geometry
class _Geo:
    """Attribute holder with the reference's ct_projector attribute names."""

    def __init__(self, **kw):
        self.cx = self.cy = self.cz = 0.0
        self.off_u = self.off_v = 0.0
        self.__dict__.update(kw)


FAN_NU, FAN_DU, FAN_DSO, FAN_DSD, FAN_FOV = 888, 1.2858, 595.0, 1085.6, 358.4
CONE_NU, CONE_NV, CONE_DSO, CONE_DSD, CONE_FOV = 512, 384, 595.0, 1085.6, 128.0


def fan_geometry(n: int, nz: int) -> _Geo:
    d = FAN_FOV / n
    return _Geo(nx=n, ny=n, nz=nz, dx=d, dy=d, dz=1.0, nu=FAN_NU, nv=nz, du=FAN_DU, dv=1.0, dsd=FAN_DSD,
                dso=FAN_DSO)


def cone_geometry(n: int) -> _Geo:
    # at most 1 mm voxels: the voxel-driven backprojector supports voxel shadows
    # of up to 4 detector cells (small volumes cover a smaller field of view)
    d = min(CONE_FOV / n, 1.0)
    return _Geo(nx=n, ny=n, nz=n, dx=d, dy=d, dz=d, nu=CONE_NU, nv=CONE_NV, du=1.0, dv=1.0, dsd=CONE_DSD,
                dso=CONE_DSO)


# --------------------------------------------------------------- requests
@dataclass
class Request:
    rid: int
    kind: str                       # 'fbp' or 'sqs'
    n: int                          # in-plane size (FBP) or cube size (SQS)
    nz: int                         # slices (FBP); = n for SQS
    nview: int
    niter: int                      # SQS iterations (0 for FBP)
    arrival: float                  # seconds after the stream starts
    prj: Optional[np.ndarray] = None    # host sinogram [1, nview, nv, nu]
    result: Optional[torch.Tensor] = None
    t_admit: float = math.nan
    t_done: float = math.nan

    @property
    def key(self) -> Tuple:
        """Geometry class: requests with equal keys share launches."""
        return (self.kind, self.n, self.nview)

    def nbytes(self) -> int:
        """Device-memory footprint while active (sinogram, volume, SQS state)."""
        if self.kind == 'fbp':
            return 4 * (2 * self.nview * self.nz * FAN_NU + self.nz * self.n * self.n)
        return 4 * (3 * self.nview * CONE_NV * CONE_NU + 4 * self.n ** 3)


def make_stream(count: int, rate: float = 200.0, seed: int = 1005, fbp_sizes=(256, 384, 512),
                fbp_views=(360, 492, 720, 984), sqs_sizes=(128, 192, 256), sqs_views=(360, 492, 720, 984),
                sqs_iters=(10, 50), max_slices: int = 64, with_data: bool = True) -> List[Request]:
    """Seeded mixed request stream: Poisson arrivals (mean ``rate`` per second),
    half fan-beam FBP (n^2 x nz slices, nz in 1..max_slices), half cone-beam
    SQS-TV (n^3, iterations in sqs_iters); sinograms are projections of seeded
    ellipsoid phantoms (made on the device, kept on the host)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        if rng.uniform() < 0.5:
            r = Request(i, 'fbp', int(rng.choice(fbp_sizes)), int(rng.integers(1, max_slices + 1)),
                        int(rng.choice(fbp_views)), 0, t)
        else:
            n = int(rng.choice(sqs_sizes))
            r = Request(i, 'sqs', n, n, int(rng.choice(sqs_views)), int(rng.integers(sqs_iters[0], sqs_iters[1] + 1)),
                        t)
        reqs.append(r)
    if with_data:
        for r in reqs:
            r.prj = _simulate(r, seed + 1 + r.rid)
    return reqs


def _simulate(r: Request, seed: int) -> np.ndarray:
    from .projector.cupy import cone, fan_equiangular as fan
    dev = torch.device('cuda', torch.cuda.current_device())
    if r.kind == 'fbp':
        g = fan_geometry(r.n, r.nz)
        img = ellipsoid_phantom(r.nz, r.n, r.n, g.dz, g.dy, g.dx, seed=seed)[None]
        ang = torch.from_numpy(_fan_angles(r.nview)).to(dev)
        return fan.distance_driven_fp(g, torch.from_numpy(img).to(dev), ang).cpu().numpy()
    g = cone_geometry(r.n)
    img = ellipsoid_phantom(r.n, r.n, r.n, g.dz, g.dy, g.dx, seed=seed)[None]
    geo = _cone_geo(r.nview, dev)
    return cone.siddon_fp_arbitrary(g, torch.from_numpy(img).to(dev), *geo).cpu().numpy()


def _fan_angles(nview: int) -> np.ndarray:
    return (np.arange(nview) * 2 * np.pi / nview).astype(np.float32)


def _cone_geo(nview: int, dev) -> List[torch.Tensor]:
    geo = circular_cone_geometry(nview, CONE_DSO, CONE_DSD)
    return [torch.from_numpy(a).to(dev) for a in (geo.det_center, geo.det_u, geo.det_v, geo.src)]


# ------------------------------------------------------------ class engines
class _FBPClass:
    """FBP of every slice of every member: slices are batch items of a
    single-row geometry (slice z of a request is detector row z)."""

    def __init__(self, n: int, nview: int, dev):
        self.g = fan_geometry(n, 1)
        self.angles = torch.from_numpy(_fan_angles(nview)).to(dev)

    def run(self, prjs: List[torch.Tensor]) -> List[torch.Tensor]:
        from .projector.cupy import fan_equiangular as fan
        counts = [p.shape[2] for p in prjs]
        # [1, nview, nz, nu] -> [nz, nview, 1, nu], stacked over requests
        stack = torch.cat([p[0].permute(1, 0, 2).unsqueeze(2) for p in prjs], 0).contiguous()
        rec = fan.fbp_bp(self.g, fan.ramp_filter(self.g, stack, 'hann'), self.angles)   # [sum nz, 1, n, n]
        out, k = [], 0
        for c in counts:
            out.append(rec[k:k + c, 0].unsqueeze(0))
            k += c
        return out


class _SQSClass:
    """SQS-TV of a batch of volumes sharing one cone-beam scan geometry."""
    BETA, WEIGHTS, EPS = 1e-3, (1.0, 1.0, 1.0), 1e-8

    def __init__(self, n: int, nview: int, dev):
        from .projector.cupy import cone, ct_projector
        self.p = ct_projector()
        self.p.__dict__.update(cone_geometry(n).__dict__)
        geo = _cone_geo(nview, dev)
        self.p.set_projector(cone.siddon_fp_arbitrary, det_center=geo[0], det_u=geo[1], det_v=geo[2], src=geo[3])
        self.p.set_backprojector(cone.siddon_bp_arbitrary, det_center=geo[0], det_u=geo[1], det_v=geo[2],
                                 src=geo[3])
        self.pn = self.p.calc_projector_norm(niter=5, verbose=False)
        self.norm_img = self.p.calc_norm_img() / self.pn / self.pn

    def step(self, x: torch.Tensor, prj: torch.Tensor) -> torch.Tensor:
        from .recon.sqs_algorithms import sqs_tv_one_step
        return sqs_tv_one_step(self.p, x, prj, self.norm_img, self.pn, self.BETA, self.WEIGHTS, self.EPS)


# ---------------------------------------------------------------- server
@dataclass
class ServeStats:
    completed: int
    makespan_s: float
    recons_per_s: float
    latency_p50_s: float
    latency_p99_s: float
    latency_mean_s: float
    max_hbm_in_use: int
    peak_concurrent: int
    launches_batched: Dict[str, int] = field(default_factory=dict)


class Server:
    """Continuous-batching reconstruction server (see the module docstring)."""

    def __init__(self, hbm_budget: int, workers: int = 4, max_fbp_slices: int = 256, max_sqs_batch: int = 4,
                 device=None):
        self.dev = torch.device('cuda', torch.cuda.current_device()) if device is None else torch.device(device)
        self.budget = int(hbm_budget)
        self.workers = workers
        self.max_fbp_slices = max_fbp_slices
        self.max_sqs_batch = max_sqs_batch
        self._lock = threading.Condition()
        self._engines: Dict[Tuple, object] = {}
        self._engine_lock = threading.Lock()

    def _engine(self, key):
        with self._engine_lock:
            e = self._engines.get(key)
            if e is None:
                kind, n, nview = key
                e = _FBPClass(n, nview, self.dev) if kind == 'fbp' else _SQSClass(n, nview, self.dev)
                self._engines[key] = e
            return e

    def warm(self, requests: List[Request]) -> None:
        """Build the class engines (geometry tables, SQS norms) before serving."""
        for k in sorted({r.key for r in requests}):
            self._engine(k)
        torch.cuda.synchronize(self.dev)

    def serve(self, requests: List[Request]) -> ServeStats:
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if r.nbytes() > self.budget:
                raise ValueError('request %d needs %d bytes, above the HBM budget' % (r.rid, r.nbytes()))
        arrived: List[Request] = []          # arrived, not admitted
        busy = set()                         # class keys being worked on by a worker
        state = {'in_use': 0, 'max_in_use': 0, 'active': 0, 'peak': 0, 'done': 0, 'next': 0,
                 'batched': {'fbp_launches': 0, 'fbp_requests': 0, 'sqs_iterations': 0, 'sqs_member_iterations': 0}}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        total = len(reqs)

        def release_arrivals():
            while state['next'] < total and reqs[state['next']].arrival <= now():
                arrived.append(reqs[state['next']])
                state['next'] += 1

        def admit(r: Request) -> bool:
            if state['in_use'] + r.nbytes() > self.budget:
                return False
            state['in_use'] += r.nbytes()
            state['max_in_use'] = max(state['max_in_use'], state['in_use'])
            state['active'] += 1
            state['peak'] = max(state['peak'], state['active'])
            r.t_admit = now()
            arrived.remove(r)
            return True

        def finish(r: Request):
            r.t_done = now()
            state['in_use'] -= r.nbytes()
            state['active'] -= 1
            state['done'] += 1

        def pick():
            """Oldest arrived request whose class is free; admit it and the
            class's other arrived requests that fit (locked)."""
            for r in sorted(arrived, key=lambda q: q.arrival):
                if r.key in busy:
                    continue
                if not admit(r):
                    continue
                group = [r]
                if r.kind == 'fbp':
                    slices = r.nz
                    for q in sorted(arrived, key=lambda q: q.arrival):
                        if q.key == r.key and slices + q.nz <= self.max_fbp_slices and admit(q):
                            group.append(q)
                            slices += q.nz
                busy.add(r.key)
                return r.key, group
            return None, None

        def join_sqs(key, members):
            """Arrived requests of an SQS class join at an iteration boundary (locked)."""
            joined = []
            for q in sorted(arrived, key=lambda q: q.arrival):
                if len(members) + len(joined) >= self.max_sqs_batch:
                    break
                if q.key == key and admit(q):
                    joined.append(q)
            return joined

        def worker():
            stream = torch.cuda.Stream(device=self.dev)
            with torch.cuda.device(self.dev), torch.cuda.stream(stream):
                while True:
                    with self._lock:
                        while True:
                            if state['done'] >= total:
                                self._lock.notify_all()
                                return
                            release_arrivals()
                            key, group = pick()
                            if key is not None:
                                break
                            wait = 0.002
                            if state['next'] < total:
                                wait = min(wait, max(reqs[state['next']].arrival - now(), 0.0002))
                            self._lock.wait(timeout=wait)
                    eng = self._engine(key)
                    if key[0] == 'fbp':
                        prjs = [torch.from_numpy(r.prj).to(self.dev, non_blocking=False) for r in group]
                        outs = eng.run(prjs)
                        stream.synchronize()
                        with self._lock:
                            for r, o in zip(group, outs):
                                r.result = o
                                finish(r)
                            state['batched']['fbp_launches'] += 1
                            state['batched']['fbp_requests'] += len(group)
                            busy.discard(key)
                            self._lock.notify_all()
                        continue
                    # SQS class: iterate the batch until every member is done
                    members = list(group)
                    prj = {r.rid: torch.from_numpy(r.prj).to(self.dev) for r in members}
                    xs = {r.rid: torch.zeros([1, r.n, r.n, r.n], dtype=torch.float32, device=self.dev)
                          for r in members}
                    it = {r.rid: 0 for r in members}
                    while members:
                        x = torch.cat([xs[r.rid] for r in members], 0)
                        p = torch.cat([prj[r.rid] for r in members], 0)
                        x = eng.step(x, p)
                        for k, r in enumerate(members):
                            xs[r.rid] = x[k:k + 1]
                            it[r.rid] += 1
                        done = [r for r in members if it[r.rid] >= r.niter]
                        if done:
                            stream.synchronize()
                        with self._lock:
                            state['batched']['sqs_iterations'] += 1
                            state['batched']['sqs_member_iterations'] += len(members)
                            for r in done:
                                r.result = xs.pop(r.rid).clone()
                                prj.pop(r.rid)
                                finish(r)
                            members = [r for r in members if it[r.rid] < r.niter]
                            release_arrivals()
                            for q in join_sqs(key, members):
                                members.append(q)
                                prj[q.rid] = torch.from_numpy(q.prj).to(self.dev)
                                xs[q.rid] = torch.zeros([1, q.n, q.n, q.n], dtype=torch.float32, device=self.dev)
                                it[q.rid] = 0
                            if not members:
                                busy.discard(key)
                            self._lock.notify_all()

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.workers)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        torch.cuda.synchronize(self.dev)
        return _stats(reqs, state['max_in_use'], state['peak'], state['batched'])


def _stats(reqs: List[Request], max_in_use: int, peak: int, batched: dict) -> ServeStats:
    lat = np.array([r.t_done - r.arrival for r in reqs])
    t_end = max(r.t_done for r in reqs)
    t_start = min(r.arrival for r in reqs)
    span = t_end - t_start
    return ServeStats(len(reqs), span, len(reqs) / span, float(np.percentile(lat, 50)), float(np.percentile(lat, 99)),
                      float(lat.mean()), max_in_use, peak, batched)


def run_sequential(requests: List[Request], device=None, kernel_time: Optional[dict] = None
                   ) -> Tuple[float, List[torch.Tensor]]:
    """Baseline: one request at a time in arrival order, plain launches (one
    request per call, the reference's sqs_one_step driver with tv_sqs, no
    batching, no overlap), all requests available from the start.  Scan tables
    shared by a geometry class (projector norm, norm image) are prepared before
    timing, as the server's are.  Returns (seconds, results in request order).
    With ``kernel_time`` (a dict), the library's kernels are timed with HIP
    events during the run and their summed duration (ms) is stored under
    'library_kernels_ms': the device time the projector, filter and prior
    kernels take when nothing overlaps them."""
    from .projector.cupy import fan_equiangular as fan
    from .prior.cupy import tv_sqs
    from .recon.sqs_algorithms import sqs_one_step
    dev = torch.device('cuda', torch.cuda.current_device()) if device is None else torch.device(device)
    classes = {r.key: _SQSClass(r.n, r.nview, dev) for r in requests if r.kind == 'sqs'}
    out = []
    from . import _lib
    if kernel_time is not None:
        _lib.set_kernel_timing(True)
        _lib.reset_kernel_times()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda q: q.arrival):
        prj = torch.from_numpy(r.prj).to(dev)
        if r.kind == 'fbp':
            g = fan_geometry(r.n, r.nz)
            ang = torch.from_numpy(_fan_angles(r.nview)).to(dev)
            out.append(fan.fbp_bp(g, fan.ramp_filter(g, prj, 'hann'), ang))
        else:
            c = classes[r.key]
            x = torch.zeros([1, r.n, r.n, r.n], dtype=torch.float32, device=dev)
            for _ in range(r.niter):
                x = sqs_one_step(c.p, x, prj, c.norm_img, c.pn, _SQSClass.BETA, tv_sqs,
                                 {'weights': _SQSClass.WEIGHTS, 'eps': _SQSClass.EPS})
            out.append(x)
        torch.cuda.synchronize(dev)
    elapsed = time.perf_counter() - t0
    if kernel_time is not None:
        kernel_time['library_kernels_ms'] = sum(v[0] for v in _lib.kernel_times().values())
        _lib.set_kernel_timing(False)
    return elapsed, out
