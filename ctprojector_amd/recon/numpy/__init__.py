"""SQS loops for the numpy front end (the reference's ct_projector.recon.numpy)."""
BACKEND = 'numpy'
