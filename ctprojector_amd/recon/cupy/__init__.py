"""SQS loops for the cupy front end (the reference's ct_projector.recon.cupy)."""
BACKEND = 'cupy'
