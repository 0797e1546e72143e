"""Loader of tests/golden/fixtures.npz (written by tests/golden/make_golden.py)."""

import json
import os

import numpy as np

from helpers import Geo

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'fixtures.npz')


def load():
    """{case: (Geo, {array name: array})}"""
    d = np.load(PATH)          # allow_pickle stays False: plain arrays and strings only
    out = {}
    for k in d.files:
        case, name = k.split('/')
        geo, arrs = out.setdefault(case, [None, {}])
        if name == 'geo':
            out[case][0] = Geo(**json.loads(str(d[k])))
        else:
            arrs[name] = d[k]
    return {k: (v[0], v[1]) for k, v in out.items()}
