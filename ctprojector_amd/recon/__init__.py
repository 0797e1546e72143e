"""SQS reconstruction loops (the reference's ct_projector.recon)."""
