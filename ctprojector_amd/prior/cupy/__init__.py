"""Device-tensor priors (the reference's ct_projector.prior.cupy, torch tensors)."""
from .recon_prior import tv_sqs, tv_sqs_backward, tv_sqs_update  # noqa: F401
