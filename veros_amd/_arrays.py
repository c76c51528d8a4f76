"""Array plumbing between callers and the C ABI: every array handed to the
library is a CUDA (HIP) torch tensor; NumPy inputs are copied to the device
and results copied back; other DLPack producers are wrapped zero-copy."""
from __future__ import annotations

import numpy as np
import torch


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("veros_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def to_device(x, dtype=None):
    """(tensor on the GPU, came_from_numpy)."""
    if isinstance(x, torch.Tensor):
        t = x
        from_np = False
    elif isinstance(x, np.ndarray) or np.isscalar(x):
        t = torch.from_numpy(np.ascontiguousarray(x))
        from_np = True
    elif hasattr(x, "__dlpack__"):
        t = torch.from_dlpack(x)
        from_np = False
    else:
        t = torch.as_tensor(x)
        from_np = True
    if t.device.type != "cuda":
        t = t.to(device())
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t, from_np


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def require_contiguous(name, t):
    if not t.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous (got strides {tuple(t.stride())})")


_workspaces: dict = {}
# Every workspace buffer ever handed out stays referenced until the process
# ends: a HIP graph captured against a buffer keeps its address, so freeing it
# when a larger grid needs a larger buffer would let a later replay of that
# graph write into memory torch has handed to another tensor.
_retired: list = []


def workspace(nbytes: int, dev: torch.device) -> torch.Tensor:
    """A cached device buffer of at least ``nbytes`` (grown, never shrunk or
    freed), allocated on first use so the hot calls allocate nothing
    afterwards. Calls that share it must be ordered on one stream."""
    buf = _workspaces.get(dev)
    if buf is None or buf.numel() < nbytes:
        if buf is not None:
            _retired.append(buf)
        buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)
        _workspaces[dev] = buf
    return buf
