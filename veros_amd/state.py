"""Minimal state container on the GPU: ``.variables`` with veros-named torch
tensors and ``.settings``, the two attributes the kernels read
(veros/state.py:219-243 validates shapes and dtypes the same way for its own
variables)."""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import torch

from ._arrays import device
from .layout import LEVELS_INNER, LEVELS_OUTER, set_state_layout

INT_VARS = ("kbot",)


def to_device_state(state, dev=None, levels=LEVELS_INNER):
    """Copy a host state (``.variables`` with NumPy arrays) to the GPU.
    Float arrays become float64, kbot int32, boolean masks stay boolean;
    scalars (tau, taup1, ...) are kept as Python values. ``levels`` selects
    the storage of the time-level variables (veros_amd.layout): the
    reference's ``LEVELS_INNER`` or ``LEVELS_OUTER`` ("inner"/"outer" also
    accepted)."""
    levels = {"inner": LEVELS_INNER, "outer": LEVELS_OUTER}.get(levels, levels)
    dev = dev or device()
    vs = state.variables
    out = {}
    for k, v in vars(vs).items():
        if isinstance(v, np.ndarray):
            if k in INT_VARS:
                out[k] = torch.from_numpy(np.ascontiguousarray(v, dtype=np.int32)).to(dev)
            elif v.dtype == np.bool_:
                out[k] = torch.from_numpy(np.ascontiguousarray(v)).to(dev)
            else:
                out[k] = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(dev)
        else:
            out[k] = v
    ds = SimpleNamespace(variables=SimpleNamespace(**out), settings=state.settings)
    if levels != LEVELS_INNER:
        set_state_layout(ds, levels)
    return ds


def to_host_arrays(tensors: dict):
    return {k: (v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in tensors.items()}
