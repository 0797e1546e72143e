"""Pins the FP64 oracle network (tests/oracle_network.py) that the GPU test of
the full-size unrolled SQS-TV gradient compares with: its autograd gradient
equals FP64 central finite differences of the network along random
directions, and the reference convention's gradient is a different one."""

import numpy as np
import torch

from helpers import Geo
from oracle_network import make_network


def test_oracle_network_gradient_vs_central_differences(oracle):
    rng = np.random.default_rng(12)
    g = Geo(nx=48, ny=48, nz=1, dx=1.4, dy=1.4, nu=80, nv=1, du=2.5716 * 444 / 80 / 2, dv=1.0, dsd=1085.6, dso=595.0)
    ang = (np.arange(60) * 2 * np.pi / 60).astype(np.float32)
    x_true = np.zeros([2, 1, 48, 48])
    x_true[:, :, 12:36, 10:30] = 0.02
    x_true[1, :, 20:28, 20:40] += 0.01
    line = oracle.dd_fan_fp(g, x_true.astype(np.float32), ang)
    counts = rng.poisson(1e5 * np.exp(-line))
    p0 = -np.log(np.maximum(counts, 1) / 1e5)
    w = counts / 1e5
    pn = 40.0
    norm_img = oracle.dd_fan_fp_adjoint(g, (w * oracle.dd_fan_fp(g, np.ones([2, 1, 48, 48], np.float32), ang))
                                        .astype(np.float32), ang) / pn / pn
    target = torch.as_tensor(x_true)
    for conv in ('exact', 'reference'):
        run = make_network(oracle, g, ang, norm_img, pn, 2e-3, 4, w, 1e-4, conv)
        p = torch.tensor(p0, requires_grad=True)
        loss = ((run(p) - target) ** 2).sum()
        loss.backward()
        if conv == 'exact':
            grad = p.grad.clone()
            for k in range(2):
                dp = torch.as_tensor(np.random.default_rng(30 + k).normal(size=p0.shape))
                h = 1e-4
                with torch.no_grad():
                    lp = float(((run(torch.as_tensor(p0) + h * dp) - target) ** 2).sum())
                    lm = float(((run(torch.as_tensor(p0) - h * dp) - target) ** 2).sum())
                fd = (lp - lm) / (2 * h)
                ad = float((grad * dp).sum())
                assert abs(ad - fd) < 1e-6 * abs(fd), (ad, fd)
        else:
            # the reference's convention differentiates with its backprojector
            rel = float((p.grad - grad).norm() / grad.norm())
            assert rel > 1e-2, rel
