"""GPU: the reconstruction server (ctprojector_amd.serving).

* every request's result is the same served in a mix (batched with other
  requests of its class, overlapping with other classes) as served alone, and
  agrees with the sequential plain-launch baseline;
* admission keeps the device-memory in use within the HBM budget.
"""

import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def torch():
    import torch as T
    assert T.cuda.is_available(), 'torch sees no HIP device'
    return T


def _stream(seed=3, count=10):
    from ctprojector_amd.serving import make_stream
    return make_stream(count, rate=400.0, seed=seed, fbp_sizes=(64, 96), fbp_views=(60, 90),
                       sqs_sizes=(24, 32), sqs_views=(40, 60), sqs_iters=(2, 4), max_slices=6)


def test_results_same_alone_and_in_mix(torch):
    from ctprojector_amd.serving import Server, run_sequential
    reqs = _stream()
    assert {r.kind for r in reqs} == {'fbp', 'sqs'}
    srv = Server(hbm_budget=8 << 30, workers=3)
    srv.warm(reqs)
    stats = srv.serve(reqs)
    assert stats.completed == len(reqs)
    assert stats.latency_p99_s >= stats.latency_p50_s > 0
    mixed = {r.rid: r.result.clone() for r in reqs}
    alone = _stream()
    for r in alone:
        Server(hbm_budget=8 << 30, workers=1).serve([r])
        assert rel_l2(r.result.cpu().numpy(), mixed[r.rid].cpu().numpy()) < 1e-6, r.rid
    _, seq = run_sequential(_stream())
    for r, s in zip(sorted(reqs, key=lambda q: q.arrival), seq):
        tol = 1e-6 if r.kind == 'fbp' else 1e-5
        assert rel_l2(mixed[r.rid].cpu().numpy(), s.cpu().numpy()) < tol, (r.rid, r.kind)


def test_admission_respects_hbm_budget(torch):
    from ctprojector_amd.serving import Server
    reqs = _stream(seed=4, count=8)
    budget = int(1.2 * max(r.nbytes() for r in reqs))
    srv = Server(hbm_budget=budget, workers=3)
    srv.warm(reqs)
    stats = srv.serve(reqs)
    assert stats.completed == len(reqs)
    assert stats.max_hbm_in_use <= budget
