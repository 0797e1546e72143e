"""FP64 unrolled SQS-TV network built from the oracle (test infrastructure).

The network of ctprojector_amd.autograd.sqs_tv_unrolled,
    x <- x - (A^T (w (A x / n - p / n)) / n + beta s1) / (norm_img + beta s2),
with A the oracle's FP64 fan-beam DD forward projector (dd_fan_fp), A^T its
FP64 exact transpose (dd_fan_fp_adjoint) and the TV surrogate in torch FP64
arithmetic, as a torch autograd graph on the CPU.  The oracle takes float32
operands, so every FP64 operand is split x = hi + lo (two float32 arrays) and
the linear map applied to both: the sum keeps ~2^-48 relative precision.

Gradient conventions: 'exact' (the VJP of A is A^T, of A^T is A) and
'reference' (the VJP of A is the reference's backprojector dd_fan_bp with
pre-weights, as its TensorFlow ops register, cuda/distanceDriven2DBoxInt.cu:319-396
through src/ct_projector/projector/tensorflow/circular_2d.py:140-191)."""

import numpy as np
import torch

from ctprojector_amd.autograd import tv_sqs_torch


def split_apply(f, x):
    hi = x.astype(np.float32)
    lo = (x - hi.astype(np.float64)).astype(np.float32)
    return f(hi) + f(lo)


class _Lin(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, fwd, vjp):
        ctx.fwd, ctx.vjp = fwd, vjp
        return torch.from_numpy(split_apply(fwd, x.detach().numpy()))

    @staticmethod
    def backward(ctx, g):
        return _Lin.apply(g.contiguous(), ctx.vjp, ctx.fwd), None, None


def make_network(oracle, g, angles, norm_img, pn, beta, niter, weight, eps, convention='exact',
                 weights=(1.0, 1.0, 1.0)):
    A = lambda x: oracle.dd_fan_fp(g, x, angles)
    AT = lambda y: oracle.dd_fan_fp_adjoint(g, y, angles)
    BP = lambda y: oracle.dd_fan_bp(g, y, angles)
    vjp_a = AT if convention == 'exact' else BP
    norm_t = torch.as_tensor(np.asarray(norm_img, np.float64))
    w_t = torch.as_tensor(np.asarray(weight, np.float64))

    def run(p):
        x = torch.zeros(norm_t.shape, dtype=torch.float64)
        for _ in range(niter):
            r = _Lin.apply(x, A, vjp_a) / pn - p / pn
            gr = _Lin.apply(r * w_t, AT, A) / pn
            s1, s2, _ = tv_sqs_torch(x, weights, eps)
            x = x - (gr + beta * s1) / (norm_t + beta * s2)
        return x
    return run
