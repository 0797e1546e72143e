"""Device-pointer plumbing shared by the tensor front end.

Tensors are passed to the ``cupy*`` entry points of libprojector.so as raw
device pointers.  The library runs on the current HIP device's null stream and
returns when its results are complete; work queued by torch on another stream
is drained first so that inputs are ready."""

from ctypes import c_void_p

import numpy as np
import torch


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
    """Make ``device`` current for the library call and order it after torch's
    pending work on the current stream."""

    def __init__(self, device):
        self.device = device
        self._ctx = torch.cuda.device(device)

    def __enter__(self):
        self._ctx.__enter__()
        s = torch.cuda.current_stream(self.device)
        if s.cuda_stream != 0:
            s.synchronize()
        return self

    def __exit__(self, *exc):
        return self._ctx.__exit__(*exc)
