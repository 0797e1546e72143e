"""Host-array projector front end (the reference's ct_projector.projector.numpy)."""
from .ct_projector import ct_projector, set_device  # noqa: F401
