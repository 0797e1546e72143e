"""Host-array priors (the reference's ct_projector.prior.numpy)."""
from .recon_prior import tv_sqs  # noqa: F401
