"""Pins the oracle's TV surrogate (oracle.tv_sqs, following
prior-cuda/TotalVariation.cu:21-139):

* s1 is the exact gradient of TV(x) = sum of the variation map (checked by
  central finite differences in FP64, every voxel incl. all edges and corners);
* closed forms: a constant image has var = sqrt(eps), s1 = 0 and
  s2 = 2 (wx + wy + wz) / sqrt(eps); an x ramp of slope a has
  var = sqrt(wx a^2 + eps) away from the lower x edge;
* a plausible mistake (weights applied in x, y, z order instead of the
  wrapper's z, y, x) is caught.
"""

import numpy as np

from oracle import oracle as O


def test_s1_is_gradient_of_total_variation():
    rng = np.random.default_rng(0)
    x = rng.uniform(size=[2, 4, 5, 6])
    w = (0.7, 1.3, 0.4)
    eps = 1e-3
    s1, _, _ = O.tv_sqs(x, w, eps)
    h = 1e-6
    fd = np.zeros_like(x)
    it = np.nditer(x, flags=['multi_index'])
    for _ in it:
        i = it.multi_index
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd[i] = (O.tv_sqs(xp, w, eps)[2].sum() - O.tv_sqs(xm, w, eps)[2].sum()) / (2 * h)
    assert np.abs(fd - s1).max() < 1e-6 * np.abs(s1).max()


def test_constant_image():
    x = np.full([1, 3, 4, 5], 2.5)
    w = (1.0, 2.0, 3.0)
    s1, s2, var = O.tv_sqs(x, w, 1e-4)
    e = float(np.float32(1e-4))
    assert np.allclose(var, np.sqrt(e))
    assert np.abs(s1).max() == 0
    assert np.allclose(s2, 2 * 6.0 / np.sqrt(e))


def test_ramp_and_weight_order():
    a = 0.3
    x = np.broadcast_to(a * np.arange(7.0), [1, 3, 4, 7]).copy()
    w = (5.0, 7.0, 2.0)           # (wz, wy, wx)
    _, _, var = O.tv_sqs(x, w, 1e-8)
    assert np.allclose(var[..., 1:], np.sqrt(2.0 * a * a + float(np.float32(1e-8))))
    _, _, wrong = O.tv_sqs(x, w[::-1], 1e-8)
    assert not np.allclose(wrong[..., 1:], var[..., 1:])
