"""CPU: the gradient of the TV surrogate.

torch autograd through ``tv_sqs_torch`` (the FP64 reference the HIP gradient
kernel is tested against) equals central finite differences of the FP64
oracle's surrogate (oracle.tv_sqs, which follows prior-cuda/TotalVariation.cu)
for sum(g1 s1 + g2 s2 + gv var), on volumes whose edges touch every branch
(lower and upper edges, single-voxel axes).  A planted mistake in the oracle
(the upper-edge variation taken from the neighbour's lower difference) must be
caught by the same comparison.
"""

import numpy as np
import pytest

from oracle import oracle as O


def _fd_grad(f, x, h=1e-5):
    g = np.zeros_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e.flat[i] = h
        g.flat[i] = (f(x + e) - f(x - e)) / (2 * h)
    return g


def _loss(tv, g1, g2, gv, w, eps):
    def f(x):
        s1, s2, var = tv(x, w, eps)
        return float((g1 * s1).sum() + (g2 * s2).sum() + (gv * var).sum())
    return f


def _autograd(x, g1, g2, gv, w, eps):
    import torch
    from ctprojector_amd.autograd import tv_sqs_torch
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    s1, s2, var = tv_sqs_torch(xt, w, eps)
    loss = sum((a * torch.from_numpy(b)).sum() for a, b in ((s1, g1), (s2, g2), (var, gv)))
    loss.backward()
    return xt.grad.numpy()


@pytest.mark.parametrize('shape', [(1, 3, 4, 5), (2, 1, 3, 4), (1, 2, 1, 3)])
def test_torch_surrogate_gradient_vs_oracle_finite_differences(shape):
    rng = np.random.default_rng(11)
    x = rng.uniform(size=shape)
    g1, g2, gv = (rng.normal(size=shape) for _ in range(3))
    w, eps = (0.5, 1.0, 2.0), 2.0 ** -10
    fd = _fd_grad(_loss(O.tv_sqs, g1, g2, gv, w, eps), x)
    ad = _autograd(x, g1, g2, gv, w, eps)
    assert np.linalg.norm(ad - fd) <= 1e-7 * np.linalg.norm(fd)


def test_planted_upper_edge_mistake_is_caught():
    def wrong(x, w, eps):
        s1, s2, var = O.tv_sqs(x, w, eps)
        wz, wy, wx = w
        # the upper x edge should use the voxel's own variation; use the
        # neighbour's lower-edge variation instead
        s2 = s2.copy()
        s2[..., -1] += wx / var[..., -2] - wx / var[..., -1]
        return s1, s2, var

    rng = np.random.default_rng(12)
    shape = (1, 2, 3, 4)
    x = rng.uniform(size=shape)
    g1, g2, gv = (rng.normal(size=shape) for _ in range(3))
    w, eps = (0.5, 1.0, 2.0), 2.0 ** -10
    fd = _fd_grad(_loss(wrong, g1, g2, gv, w, eps), x)
    ad = _autograd(x, g1, g2, gv, w, eps)
    assert np.linalg.norm(ad - fd) > 1e-3 * np.linalg.norm(fd)
