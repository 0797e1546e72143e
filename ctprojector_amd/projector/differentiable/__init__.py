"""Differentiable projectors with per-sample geometry: the reference's
TensorFlow ops and Keras layers (src/ct_projector/projector/tensorflow/) as
torch autograd functions and modules."""
from . import circular_2d, cone  # noqa: F401
