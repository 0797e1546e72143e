"""The committed golden fixtures: the notebook known answer holds, and the
oracle reproduces every fixture (guards the oracle against regressions)."""

import numpy as np
import pytest

import golden_fixtures
from helpers import rel_l2


@pytest.fixture(scope='module')
def fx():
    return golden_fixtures.load()


def test_known_answer_fixture(fx, oracle):
    g, a = fx['kat_parallel']
    assert np.abs(oracle.dd_par_fp(g, a['img'], a['angles']) - a['fp']).max() < 1e-12


def test_oracle_reproduces_fixtures(fx, oracle):
    O = oracle
    g, a = fx['siddon_cone']
    geo = (a['det_center'], a['det_u'], a['det_v'], a['src'])
    assert rel_l2(O.siddon_cone_fp(g, a['img'], *geo), a['fp']) < 1e-12
    assert rel_l2(O.siddon_cone_bp(g, a['prj'], *geo), a['bp']) < 1e-12
    for br, tag in ((False, ''), (True, '_branchless')):
        g, a = fx['dd_fan' + tag]
        assert rel_l2(O.dd_fan_fp(g, a['img'], a['angles'], branchless=br), a['fp']) < 1e-12
        assert rel_l2(O.dd_fan_bp(g, a['prj'], a['angles'], branchless=br), a['bp']) < 1e-12
        assert rel_l2(O.dd_fan_bp(g, a['prj'], a['angles'], is_fbp=True, branchless=br), a['bp_fbp']) < 1e-12
        g, a = fx['dd_par' + tag]
        assert rel_l2(O.dd_par_fp(g, a['img'], a['angles'], branchless=br), a['fp']) < 1e-12
        assert rel_l2(O.dd_par_bp(g, a['prj'], a['angles'], branchless=br), a['bp']) < 1e-12
    g, a = fx['tomo']
    assert rel_l2(O.dd_tomo_fp(g, a['img'], a['det_center'], a['src']), a['fp']) < 1e-12
    assert rel_l2(O.dd_tomo_bp(g, a['prj'], a['det_center'], a['src']), a['bp']) < 1e-12
    g, a = fx['filters']
    assert rel_l2(O.fan_ramp_filter(g, a['prj'], 'hann'), a['fan_hann']) < 1e-12
    assert rel_l2(O.par_ramp_filter(g, a['prj'], 'hamming'), a['par_hamming']) < 1e-12
    assert rel_l2(O.fan_fbp_bp(g, a['prj'], a['angles']), a['fbp_bp']) < 1e-12
    _, a = fx['tv']
    for got, want in zip(O.tv_sqs(a['img'], (0.5, 1.0, 2.0), 1e-4), (a['s1'], a['s2'], a['var'])):
        assert rel_l2(got, want) < 1e-12
