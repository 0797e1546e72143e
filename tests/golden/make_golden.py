"""Writes tests/golden/fixtures.npz: small input/output vectors of every
hot-path function, computed by the FP64 oracle (oracle/oracle.cpp), plus the
known answer of the reference notebook example/ex_dd_fbp_forbild.ipynb cells
2-3.  The reference ships no golden vectors of its own; these pin the oracle
against regressions and let the GPU tests check the HIP path without the
oracle.  Regenerate with:  python tests/golden/make_golden.py
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', '..'))
sys.path.insert(0, os.path.join(HERE, '..'))

from helpers import Geo  # noqa: E402
from oracle import oracle as O  # noqa: E402


def cases():
    rng = np.random.default_rng(2024)
    out = {}

    def put(name, geo, **arrays):
        out[name + '/geo'] = np.array(json.dumps(geo.__dict__))
        for k, v in arrays.items():
            out[name + '/' + k] = np.asarray(v)

    # notebook known answer: 2x2 block of ones, dx = du = 0.25, angle 0 -> 0.5 at iu 7, 8
    g = Geo(nx=16, ny=16, nz=1, dx=0.25, dy=0.25, dz=1.0, nu=16, nv=1, du=0.25, dv=1.0)
    img = np.zeros([1, 1, 16, 16], np.float32)
    img[0, 0, 7:9, 7:9] = 1
    kat = np.zeros([1, 1, 1, 16], np.float32)
    kat[0, 0, 0, 7:9] = 0.5
    put('kat_parallel', g, img=img, angles=np.zeros(1, np.float32), fp=kat)

    # Siddon cone
    g = Geo(nx=9, ny=8, nz=7, dx=1.1, dy=0.9, dz=1.2, cx=0.3, cy=-0.2, cz=0.1, nu=12, nv=10, du=1.7, dv=1.6,
            off_u=0.2, off_v=-0.3)
    t = np.array([0.3, 1.9, 4.1])
    src = np.stack([60 * np.cos(t), 60 * np.sin(t), [1.0, -1.0, 0.5]], 1).astype(np.float32)
    dc = np.stack([-40 * np.cos(t), -40 * np.sin(t), [0.5, 0.0, -0.5]], 1).astype(np.float32)
    du = np.stack([-np.sin(t), np.cos(t), np.zeros(3)], 1).astype(np.float32)
    dv = np.tile(np.float32([[0, 0, 1]]), (3, 1))
    img = rng.uniform(size=[1, 7, 8, 9]).astype(np.float32)
    y = rng.uniform(size=[1, 3, 10, 12]).astype(np.float32)
    put('siddon_cone', g, img=img, prj=y, det_center=dc, det_u=du, det_v=dv, src=src,
        fp=O.siddon_cone_fp(g, img, dc, du, dv, src), bp=O.siddon_cone_bp(g, y, dc, du, dv, src))

    # DD fan / parallel, BoxInt and branchless, with 45-degree views
    ang = np.array([0.0, 0.5, np.pi / 4, 2.0, 3.9], np.float32)
    gf = Geo(nx=12, ny=12, nz=2, dx=1.0, dy=1.0, dz=1.0, nu=20, nv=2, du=1.5, dv=1.0, dsd=200.0, dso=120.0,
             off_u=0.3)
    gp = Geo(nx=12, ny=12, nz=2, dx=1.0, dy=1.0, dz=1.0, nu=20, nv=2, du=0.9, dv=1.0, off_u=-0.2)
    img = rng.uniform(size=[1, 2, 12, 12]).astype(np.float32)
    y = rng.uniform(size=[1, 5, 2, 20]).astype(np.float32)
    for br in (False, True):
        tag = '_branchless' if br else ''
        put('dd_fan' + tag, gf, img=img, prj=y, angles=ang, fp=O.dd_fan_fp(gf, img, ang, branchless=br),
            bp=O.dd_fan_bp(gf, y, ang, branchless=br), bp_fbp=O.dd_fan_bp(gf, y, ang, is_fbp=True, branchless=br))
        put('dd_par' + tag, gp, img=img, prj=y, angles=ang, fp=O.dd_par_fp(gp, img, ang, branchless=br),
            bp=O.dd_par_bp(gp, y, ang, branchless=br))

    # tomosynthesis
    gt = Geo(nx=10, ny=9, nz=4, dx=1.0, dy=1.0, dz=1.0, cz=15.0, nu=14, nv=13, du=1.0, dv=1.0)
    tt = np.array([-0.3, 0.0, 0.3])
    src = np.stack([200 * np.sin(tt), 3 + 0 * tt, 200 * np.cos(tt) + 15], 1).astype(np.float32)
    dc = np.stack([0 * tt, 0 * tt, -8 + 0 * tt], 1).astype(np.float32)
    img = rng.uniform(size=[1, 4, 9, 10]).astype(np.float32)
    y = rng.uniform(size=[1, 3, 13, 14]).astype(np.float32)
    put('tomo', gt, img=img, prj=y, det_center=dc, src=src, fp=O.dd_tomo_fp(gt, img, dc, src),
        bp=O.dd_tomo_bp(gt, y, dc, src))

    # ramp filters and fan FBP backprojection
    y = rng.uniform(size=[1, 6, 2, 25]).astype(np.float32)
    gfl = Geo(nx=10, ny=10, nz=2, dx=1.2, dy=1.2, dz=1.0, nu=25, nv=2, du=1.3, dv=1.0, dsd=300.0, dso=180.0,
              off_u=0.4)
    angf = (np.arange(6) * np.pi / 3).astype(np.float32)
    put('filters', gfl, prj=y, angles=angf, fan_hann=O.fan_ramp_filter(gfl, y, 'hann'),
        par_hamming=O.par_ramp_filter(gfl, y, 'hamming'), fbp_bp=O.fan_fbp_bp(gfl, y, angf))

    # TV prior
    x = rng.uniform(size=[2, 3, 5, 6]).astype(np.float32)
    s1, s2, var = O.tv_sqs(x, (0.5, 1.0, 2.0), 1e-4)
    put('tv', Geo(), img=x, s1=s1, s2=s2, var=var)
    return out


if __name__ == '__main__':
    O.build()
    path = os.path.join(HERE, 'fixtures.npz')
    np.savez_compressed(path, **cases())
    print(path, os.path.getsize(path), 'bytes')
