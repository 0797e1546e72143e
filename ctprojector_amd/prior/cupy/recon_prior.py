"""TV prior on device tensors: the reference's
src/ct_projector/prior/cupy/recon_prior.py:85-135 (``tv_sqs``) with torch
tensors, plus ``tv_sqs_update``, the fused SQS-TV step.  The NLM priors of that
module are not on the projector path and are not provided."""

from ctypes import c_float, c_ulong, c_void_p
from typing import Optional, Tuple

import torch

from ... import _lib
from ...projector.cupy._dev import dev_f32, device_of, on_device, ptr


def _stream(dev) -> c_void_p:
    return c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def tv_sqs(img: torch.Tensor, weights: Tuple[float, float, float], eps: float = 1e-8):
    """Anisotropic TV surrogate of img [batch, nz, ny, nx]; weights along (z, y, x).
    Returns (s1, s2, var) as device tensors, queued on torch's current stream
    like torch's own operations (ctpTVSQS3DAsync)."""
    dev = device_of(img)
    img = dev_f32(img, dev)
    s1 = torch.empty_like(img)
    s2 = torch.empty_like(img)
    var = torch.empty_like(img)
    with torch.cuda.device(dev):
        err = _lib.module().ctpTVSQS3DAsync(
            ptr(s1), ptr(s2), ptr(var), ptr(img), c_float(weights[2]), c_float(weights[1]), c_float(weights[0]),
            c_ulong(img.shape[0]), c_ulong(img.shape[3]), c_ulong(img.shape[2]), c_ulong(img.shape[1]), c_float(eps),
            _stream(dev))
    if err != 0:
        raise RuntimeError('ctpTVSQS3DAsync failed (%d): %s' % (err, _lib.last_error_message()))
    return s1, s2, var


def tv_sqs_backward(img: torch.Tensor, g1: torch.Tensor, g2: torch.Tensor, weights: Tuple[float, float, float],
                    eps: float = 1e-8, gvar: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Vector-Jacobian product of ``tv_sqs``: the gradient with respect to img of
    sum(g1 * s1) + sum(g2 * s2) + sum(gvar * var) (no counterpart in the
    reference; used by the differentiable unrolled SQS-TV network), queued on
    torch's current stream (ctpTVSQSBackward3DAsync)."""
    dev = device_of(img)
    img = dev_f32(img, dev)
    g1 = dev_f32(g1, dev).expand_as(img).contiguous()
    g2 = dev_f32(g2, dev).expand_as(img).contiguous()
    gv = dev_f32(gvar, dev).expand_as(img).contiguous() if gvar is not None else None
    out = torch.empty_like(img)
    scratch = torch.empty((2,) + tuple(img.shape), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        err = _lib.module().ctpTVSQSBackward3DAsync(
            ptr(out), ptr(img), ptr(g1), ptr(g2), ptr(gv) if gv is not None else None, ptr(scratch), c_float(weights[2]),
            c_float(weights[1]), c_float(weights[0]), c_ulong(img.shape[0]), c_ulong(img.shape[3]),
            c_ulong(img.shape[2]), c_ulong(img.shape[1]), c_float(eps), _stream(dev))
    if err != 0:
        raise RuntimeError('ctpTVSQSBackward3DAsync failed (%d): %s' % (err, _lib.last_error_message()))
    return out


def tv_sqs_update(img: torch.Tensor, bp: torch.Tensor, norm_img: torch.Tensor, beta: float,
                  weights: Tuple[float, float, float], eps: float = 1e-8, acc_fac: float = 1.0,
                  return_var: bool = False):
    """One fused SQS step with the TV prior (one pass over the volume):
    img - (acc_fac * bp + beta * s1) / (norm_img + beta * s2), with s1, s2 the
    surrogate of ``tv_sqs(img, weights, eps)``.  Returns the new image (and the
    variation map of ``img`` when ``return_var``)."""
    dev = device_of(img)
    img = dev_f32(img, dev)
    bp = dev_f32(bp, dev)
    norm_img = dev_f32(norm_img, dev).expand_as(img).contiguous()
    out = torch.empty_like(img)
    var = torch.empty_like(img) if return_var else None
    with on_device(dev):
        err = _lib.module().ctpTVSQSUpdate3D(
            ptr(out), ptr(var) if var is not None else None, ptr(img), ptr(bp), ptr(norm_img), c_float(beta),
            c_float(acc_fac), c_float(weights[2]), c_float(weights[1]), c_float(weights[0]),
            c_ulong(img.shape[0]), c_ulong(img.shape[3]), c_ulong(img.shape[2]), c_ulong(img.shape[1]), c_float(eps))
    if err != 0:
        raise RuntimeError('ctpTVSQSUpdate3D failed (%d): %s' % (err, _lib.last_error_message()))
    return (out, var) if return_var else out
