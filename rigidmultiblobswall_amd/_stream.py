"""The current HIP stream as the integer handle the C ABI takes.

``torch.cuda.current_stream(dev).cuda_stream`` builds a torch Stream object
on every call (several microseconds of Python); the solvers pass a stream to
each of their many short launches, so the raw handle is read directly when
torch offers it.
"""
from __future__ import annotations

import torch

_raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def current(device=None) -> int:
    """Handle of the current stream of `device` (a torch.device, an index or
    None for the current device)."""
    if device is None:
        idx = torch.cuda.current_device()
    elif isinstance(device, torch.device):
        idx = torch.cuda.current_device() if device.index is None else device.index
    else:
        idx = int(device)
    if _raw is not None:
        return int(_raw(idx))
    return torch.cuda.current_stream(idx).cuda_stream
