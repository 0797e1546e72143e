"""GPU: the FFT ramp filter (two detector rows per complex LDS-resident FFT,
N = 2^p >= 2nu - 1) against the FP64 oracle, the direct convolution
(ramp_mode 1) and the generic mixed-radix kernel (ramp_mode 2), for detector
widths that take the three-pass radix-16 kernel (N = 1024, 2048, 4096) and
the generic one (N = 8, 8 + 4 passes, 8 + 2 passes, ...), odd row counts (a
last block with a missing row), every window, both beams, with offsets."""

import numpy as np
import pytest

from ctprojector_amd.projector.numpy import fan_equiangluar as fan, parallel
from ctprojector_amd.projector.numpy.ct_projector import ct_projector

from helpers import Geo, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture
def tuning():
    from ctprojector_amd import _lib
    keys = []

    def set_(k, v):
        keys.append(k)
        _lib.set_tuning(k, v)
    yield set_
    for k in keys:
        _lib.set_tuning(k, 0)


@pytest.mark.parametrize('nu', [1, 3, 64, 200, 367, 888, 2048])
@pytest.mark.parametrize('beam', ['fan', 'parallel'])
def test_fft_ramp_filter(oracle, tuning, nu, beam):
    rng = np.random.default_rng(nu)
    g = Geo(nu=nu, nv=5, du=1.2858 if beam == 'fan' else 0.9, dv=1.0, off_u=0.35, dsd=1085.6, dso=595.0)
    prj = (1.0 + rng.uniform(size=[1, 3, 5, nu])).astype(np.float32)      # 15 rows: an odd count
    p = g.to(ct_projector())
    mod = fan if beam == 'fan' else parallel
    ref_fn = oracle.fan_ramp_filter if beam == 'fan' else oracle.par_ramp_filter
    for filt in ('hann', 'hamming', 'cosine', 'ramp'):
        print('[ramp-fft]', beam, nu, filt, flush=True)
        a = mod.ramp_filter(p, prj, filt)
        e = rel_l2(a, ref_fn(g, prj, filt))
        tuning('ramp_mode', 1)
        b = mod.ramp_filter(p, prj, filt)
        tuning('ramp_mode', 2)
        c = mod.ramp_filter(p, prj, filt)
        tuning('ramp_mode', 0)
        print('[ramp-fft] rel-L2 %.2e vs direct %.2e vs generic FFT %.2e' % (e, rel_l2(a, b), rel_l2(a, c)),
              flush=True)
        assert e < 1e-4
        assert rel_l2(a, b) < 1e-4 and rel_l2(c, b) < 1e-4
