"""Loader of the HIP projector library (libprojector.so).

The library is loaded from ``ctprojector_amd/kernel/bin/libprojector.so`` as
the reference loads its CUDA library with ctypes
(src/ct_projector/projector/numpy/ct_projector.py:15-17).  There is no CPU
fallback: if the library is missing or cannot be loaded, every projector call
raises ``ProjectorLibraryError``.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'kernel', 'bin', 'libprojector.so')

# Exported C symbols and their declarations in include/ctprojector.h.
SYMBOLS = [
    'SetDevice',
    'cSiddonConeProjectionArbitrary', 'cupySiddonConeProjectionArbitrary',
    'cSiddonConeBackprojectionArbitrary', 'cupySiddonConeBackprojectionArbitrary',
    'cDistanceDrivenFanProjection', 'cupyDistanceDrivenFanProjection',
    'cDistanceDrivenParallelProjection', 'cupyDistanceDrivenParallelProjection',
    'cDistanceDrivenFanBackprojection', 'cupyDistanceDrivenFanBackprojection',
    'cDistanceDrivenParallelBackprojection', 'cupyDistanceDrivenParallelBackprojection',
    'cDistanceDrivenTomoProjection', 'cupyDistanceDrivenTomoProjection',
    'cDistanceDrivenTomoBackprojection', 'cupyDistanceDrivenTomoBackprojection',
    'cfbpFanFilter', 'cupyfbpFanFilter', 'cfbpParallelFilter', 'cupyfbpParallelFilter',
    'cfbpFanBackprojection', 'cupyfbpFanBackprojection',
    'cSiddonFanProjection', 'cSiddonFanBackprojection', 'cfbpParallelBackprojection', 'cupyFbpTomoFilter',
    'ctpDistanceDrivenFanProjectionAdjoint', 'ctpDistanceDrivenParallelProjectionAdjoint',
    'ctpDistanceDrivenFanProjectionNoAlloc',
    'cupyTVSQS3D', 'cTVSQS3D', 'ctpTVSQSUpdate3D', 'ctpTVSQSBackward3D', 'ctpTVSQS3DAsync', 'ctpTVSQSBackward3DAsync',
    'ctpDDPlanCreate', 'ctpDDPlanForward', 'ctpDDPlanBackward', 'ctpDDPlanAdjoint', 'ctpDDPlanDestroy',
    'ctpConePlanCreate', 'ctpConePlanForward', 'ctpConePlanBackward', 'ctpConePlanConfig', 'ctpConePlanDestroy',
    'ctpDDBatchCreate', 'ctpDDBatchForward', 'ctpDDBatchBackward', 'ctpDDBatchConfig', 'ctpDDBatchDestroy',
    'ctpFilterBatchCreate', 'ctpFilterBatchRun', 'ctpFilterBatchDestroy',
    'ctpGetLastErrorMessage', 'ctpReleaseWorkspace', 'ctpSetStream',
    'ctpSetKernelTiming', 'ctpResetKernelTimes', 'ctpGetKernelTime', 'ctpKernelTimeReport',
    'ctpSetTuning', 'ctpGetTuning',
]


class ProjectorLibraryError(RuntimeError):
    pass


_module = None


def _init_torch_runtime():
    """PyTorch-ROCm bundles its own HIP runtime.  When both runtimes live in one
    process, the one that initialises the device second may find no device if
    it is torch's; so torch's runtime, when torch is installed, initialises
    first.  Device pointers are valid across the two (one HSA address space)."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


def module():
    """The loaded ctypes library (loaded on first use)."""
    global _module
    if _module is None:
        _init_torch_runtime()
        if not os.path.exists(LIB_PATH):
            raise ProjectorLibraryError(
                '%s is not built; run `python -m ctprojector_amd.build` (hipcc, gfx950)' % LIB_PATH)
        try:
            lib = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise ProjectorLibraryError('cannot load %s: %s' % (LIB_PATH, e))
        for name in SYMBOLS:
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
        lib.ctpGetLastErrorMessage.restype = ctypes.c_char_p
        _module = lib
    return _module


def last_error_message() -> str:
    msg = module().ctpGetLastErrorMessage()
    return msg.decode() if msg else ''


def release_workspace() -> int:
    return module().ctpReleaseWorkspace()


def set_tuning(key: str, value: int) -> None:
    """Set a launch-shape knob (0 restores the built-in default)."""
    if module().ctpSetTuning(key.encode(), ctypes.c_int(int(value))) != 0:
        raise KeyError('unknown tuning key %r' % key)


def get_tuning(key: str) -> int:
    return module().ctpGetTuning(key.encode())


def set_kernel_timing(enable: bool) -> None:
    """Bracket every kernel launch with HIP events on its stream (measurement)."""
    module().ctpSetKernelTiming(ctypes.c_int(1 if enable else 0))


def reset_kernel_times() -> None:
    module().ctpResetKernelTimes()


def kernel_times() -> dict:
    """{kernel name: (total ms, launches)} accumulated since the last reset."""
    n = module().ctpKernelTimeReport(None, ctypes.c_size_t(0))
    buf = ctypes.create_string_buffer(n + 1)
    module().ctpKernelTimeReport(buf, ctypes.c_size_t(n + 1))
    out = {}
    for line in buf.value.decode().splitlines():
        name, ms, cnt = line.split()
        out[name] = (float(ms), int(cnt))
    return out


def set_stream(handle) -> None:
    """Run the calling thread's subsequent library calls on this HIP stream
    (a hipStream_t as an int, e.g. torch.cuda.Stream.cuda_stream; 0 = null stream)."""
    module().ctpSetStream(ctypes.c_void_p(int(handle) if handle else None))
