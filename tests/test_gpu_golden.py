"""GPU: every hot-path entry point (through the C ABI, host-array front end)
against the committed golden fixtures (FP64 oracle vectors and the notebook
known answer), rel-L2 <= 1e-4."""

import numpy as np
import pytest

import golden_fixtures
from helpers import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-4


@pytest.fixture(scope='module')
def fx():
    return golden_fixtures.load()


def _p(g):
    from ctprojector_amd.projector.numpy.ct_projector import ct_projector
    return g.to(ct_projector())


def test_known_answer(fx):
    from ctprojector_amd.projector.numpy import parallel
    g, a = fx['kat_parallel']
    assert np.abs(parallel.distance_driven_fp(_p(g), a['img'], a['angles']) - a['fp']).max() < 1e-6


def test_siddon(fx):
    from ctprojector_amd.projector.numpy import cone
    g, a = fx['siddon_cone']
    geo = (a['det_center'], a['det_u'], a['det_v'], a['src'])
    assert rel_l2(cone.siddon_fp_arbitrary(_p(g), a['img'], *geo), a['fp']) < TOL
    assert rel_l2(cone.siddon_bp_arbitrary(_p(g), a['prj'], *geo), a['bp']) < TOL


@pytest.mark.parametrize('branchless', [False, True])
def test_dd(fx, branchless):
    from ctprojector_amd.projector.numpy import fan_equiangluar as fan, parallel
    tag = '_branchless' if branchless else ''
    g, a = fx['dd_fan' + tag]
    assert rel_l2(fan.distance_driven_fp(_p(g), a['img'], a['angles'], branchless=branchless), a['fp']) < TOL
    assert rel_l2(fan.distance_driven_bp(_p(g), a['prj'], a['angles'], branchless=branchless), a['bp']) < TOL
    assert rel_l2(fan.distance_driven_bp(_p(g), a['prj'], a['angles'], is_fbp=True, branchless=branchless),
                  a['bp_fbp']) < TOL
    g, a = fx['dd_par' + tag]
    assert rel_l2(parallel.distance_driven_fp(_p(g), a['img'], a['angles'], branchless=branchless), a['fp']) < TOL
    assert rel_l2(parallel.distance_driven_bp(_p(g), a['prj'], a['angles'], branchless=branchless), a['bp']) < TOL


def test_tomo(fx):
    from ctprojector_amd.projector.numpy import tomo
    g, a = fx['tomo']
    assert rel_l2(tomo.distance_driven_fp(_p(g), a['img'], a['det_center'], a['src']), a['fp']) < TOL
    assert rel_l2(tomo.distance_driven_bp(_p(g), a['prj'], a['det_center'], a['src']), a['bp']) < TOL


def test_filters_and_fbp(fx):
    from ctprojector_amd.projector.numpy import fan_equiangluar as fan, parallel
    g, a = fx['filters']
    assert rel_l2(fan.ramp_filter(_p(g), a['prj'], 'hann'), a['fan_hann']) < TOL
    assert rel_l2(parallel.ramp_filter(_p(g), a['prj'], 'hamming'), a['par_hamming']) < TOL
    assert rel_l2(fan.fbp_bp(_p(g), a['prj'], a['angles']), a['fbp_bp']) < TOL


def test_tv(fx):
    from ctprojector_amd.prior.numpy import tv_sqs
    _, a = fx['tv']
    for got, want in zip(tv_sqs(a['img'], (0.5, 1.0, 2.0), 1e-4), (a['s1'], a['s2'], a['var'])):
        assert rel_l2(got, want) < 1e-5
