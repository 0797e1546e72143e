"""Priors for the SQS reconstruction loops (the reference's ct_projector.prior)."""
