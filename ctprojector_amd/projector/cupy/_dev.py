"""Device-pointer plumbing shared by the tensor front end.

Tensors are passed to the ``cupy*`` entry points of libprojector.so as raw
device pointers.  Each call runs on torch's current stream of the tensors'
device (``ctpSetStream``; torch and the library share one HIP runtime, see
``_lib``), so it is ordered after the work torch has queued there, and returns
when its results are complete."""

from ctypes import c_void_p

import numpy as np
import torch

from ... import _lib


def device_of(*tensors):
    for t in tensors:
        if isinstance(t, torch.Tensor) and t.is_cuda:
            return t.device
    return torch.device('cuda', torch.cuda.current_device())


def dev_f32(x, device):
    """Contiguous float32 tensor on ``device`` (numpy arrays and CPU tensors are copied)."""
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32))
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(x, dtype=torch.float32)
    x = x.to(device=device, dtype=torch.float32)
    return x.contiguous()


def ptr(t: torch.Tensor) -> c_void_p:
    return c_void_p(t.data_ptr())


class on_device:
    """Make ``device`` current and run the library call on torch's current
    stream of that device."""

    def __init__(self, device):
        self.device = device
        self._ctx = torch.cuda.device(device)

    def __enter__(self):
        self._ctx.__enter__()
        _lib.set_stream(torch.cuda.current_stream(self.device).cuda_stream)
        return self

    def __exit__(self, *exc):
        return self._ctx.__exit__(*exc)
