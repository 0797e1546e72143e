"""SQS iterations with a generic forward / back projector: the reference's
src/ct_projector/recon/{numpy,cupy}/sqs_algorithms.py:22-57 (``sqs_one_step``)
and :132-179 (``nesterov_acceleration``), written once for numpy arrays and
torch tensors (both provide the arithmetic and ``.sum()`` used here).

``sqs_tv_one_step`` is the same step as ``sqs_one_step`` with the TV prior,
with the prior and the update fused into one pass over the volume on the
device (``prior.cupy.tv_sqs_update``).  ``sqs_gaussian_one_step`` of the
reference needs the NLM prior, which is not part of this package."""

from typing import Any, Callable, Tuple, Union


def sqs_one_step(projector, img, prj, norm_img, projector_norm: float, beta: float,
                 prior_func: Callable[..., Tuple[Any, ...]], prior_kwargs: dict, acc_fac: float = 1,
                 fp_kwargs: dict = {}, bp_kwargs: dict = {}, weight=None, return_loss: bool = False):
    """x <- x - (acc_fac * A^T w (A x - p) + beta * prior[0]) / (norm_img + beta * prior[1]),
    with the projector's fp / bp scaled by 1 / projector_norm
    (recon/cupy/sqs_algorithms.py:22-57)."""
    if weight is None:
        weight = 1
    fp = projector.fp(img, **fp_kwargs) / projector_norm
    fp = fp - prj / projector_norm
    bp = projector.bp(fp * weight, **bp_kwargs) / projector_norm
    prior_kwargs['img'] = img
    prior_res = prior_func(**prior_kwargs)
    img = img - (bp * acc_fac + beta * prior_res[0]) / (norm_img + beta * prior_res[1])
    if return_loss:
        fp = projector.fp(img) / projector_norm
        data_loss = 0.5 * (weight * (fp - prj / projector_norm) ** 2).sum()
        prior_loss = prior_res[2].sum()
        return img, data_loss, prior_loss
    return img


def sqs_tv_one_step(projector, img, prj, norm_img, projector_norm: float, beta: float,
                    weights: Tuple[float, float, float] = (1, 1, 1), eps: float = 1e-8, acc_fac: float = 1,
                    fp_kwargs: dict = {}, bp_kwargs: dict = {}, weight=None, return_loss: bool = False):
    """``sqs_one_step`` with ``prior_func = tv_sqs`` (device tensors), the prior
    and the update fused into one kernel."""
    from ..prior.cupy import tv_sqs_update
    if weight is None:
        weight = 1
    fp = projector.fp(img, **fp_kwargs) / projector_norm
    fp = fp - prj / projector_norm
    bp = projector.bp(fp * weight, **bp_kwargs) / projector_norm
    new, var = tv_sqs_update(img, bp, norm_img, beta, weights, eps, acc_fac, return_var=True)
    if return_loss:
        fp = projector.fp(new) / projector_norm
        data_loss = 0.5 * (weight * (fp - prj / projector_norm) ** 2).sum()
        return new, data_loss, var.sum()
    return new


def nesterov_acceleration(recon_func: Callable[..., Union[Any, Tuple[Any, ...]]], recon_kwargs: dict, img,
                          img_nesterov, nesterov: float = 0.5):
    """Nesterov momentum around one iteration of ``recon_func``
    (recon/cupy/sqs_algorithms.py:132-179)."""
    recon_kwargs['img'] = img
    res = recon_func(**recon_kwargs)
    if type(res) is tuple:
        img = res[0] + nesterov * (res[0] - img_nesterov)
        img_nesterov = res[0]
        return tuple([img, img_nesterov] + list(res[1:]))
    img = res + nesterov * (res - img_nesterov)
    img_nesterov = res
    return img, img_nesterov
