"""Device-tensor projector front end: the reference's
ct_projector.projector.cupy with PyTorch-ROCm tensors in place of cupy arrays."""
from .ct_projector import ct_projector, set_device  # noqa: F401
