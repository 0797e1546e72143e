"""MI355X-native CT projector hot path (Siddon cone, distance-driven fan /
parallel / tomosynthesis, ramp filter and fan FBP backprojection).

Drop-in counterpart of the ``ct_projector.projector.{numpy,cupy}`` modules of
wudufan/CTProjector: ``ctprojector_amd.projector.numpy`` takes host arrays,
``ctprojector_amd.projector.cupy`` takes device tensors (PyTorch-ROCm in place
of cupy).  The kernels live in ``kernel/bin/libprojector.so`` (HIP, gfx950).
"""

__version__ = '0.1.0'
