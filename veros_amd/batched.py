"""Many independent kbot-masked tridiagonal problems in one persistent launch
per depth bucket (C ABI: vx_tdma_plan_* in include/vxmix.h), in float64 or
mixed precision (``precision="mixed"``: float32 sweeps, one float64
refinement step), on the current stream or replayed as one captured hipGraph.

Each system is (a, b, c, d, kbot, x[, b_edge, d_edge]) with the semantics of
``solve_implicit_kbot`` (veros/core/utilities.py:51-59): torch float64 tensors
on the GPU of shape (nx, ny, nz) sharing one (x, y) stride pair per system and
z contiguous, kbot int32 (nx, ny). Shapes and depths may differ between
systems. The plan keeps references to the tensors; their storage must not
move while the plan is alive.
"""
from __future__ import annotations

import ctypes

import torch

from . import _native
from ._arrays import stream_handle

MODES = {"fp64": 0, "mixed": 1}


class TdmaBatch:
    def __init__(self, systems, precision="fp64"):
        if precision not in MODES:
            raise ValueError(f"precision must be one of {tuple(MODES)}")
        self.precision = precision
        self._keep = []
        arr = (_native.VxTdmaSystem * len(systems))()
        for q, sysd in enumerate(systems):
            a, b, c, d, kbot, x = (sysd[k] for k in ("a", "b", "c", "d", "kbot", "x"))
            be, de = sysd.get("b_edge"), sysd.get("d_edge")
            ts = [t for t in (a, b, c, d, x, be, de) if t is not None]
            for t in ts:
                if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
                    raise TypeError(f"system {q}: arrays must be torch tensors on the GPU")
                if t.dtype != torch.float64:
                    raise TypeError(f"system {q}: unsupported dtype {t.dtype} (float64 expected)")
                if t.shape != a.shape or t.stride() != a.stride():
                    raise ValueError(f"system {q}: all arrays must share shape and strides")
            if a.dim() != 3 or a.stride(2) != 1:
                raise ValueError(f"system {q}: expected (nx, ny, nz) arrays with z contiguous")
            if kbot.dtype != torch.int32 or tuple(kbot.shape) != tuple(a.shape[:2]):
                raise ValueError(f"system {q}: kbot must be int32 of shape {tuple(a.shape[:2])}")
            y = arr[q]
            y.a, y.b, y.c, y.d, y.x = (t.data_ptr() for t in (a, b, c, d, x))
            y.b_edge = be.data_ptr() if be is not None else None
            y.d_edge = de.data_ptr() if de is not None else None
            y.kbot = kbot.data_ptr()
            y.kbot_sx, y.kbot_sy = kbot.stride()
            y.nx, y.ny, y.nz = a.shape
            y.sx, y.sy = a.stride(0), a.stride(1)
            self._keep.extend(ts + [kbot])
        self._lib = _native.lib()
        handle = ctypes.c_void_p()
        _native.check(self._lib.vx_tdma_plan_create(arr, len(systems), MODES[precision], ctypes.byref(handle)),
                      "vx_tdma_plan_create")
        self._plan = handle

    def run(self):
        _native.check(self._lib.vx_tdma_plan_run(self._plan, stream_handle()), "vx_tdma_plan_run")

    def tune(self):
        _native.check(self._lib.vx_tdma_plan_tune(self._plan, stream_handle()), "vx_tdma_plan_tune")
        return self.info()

    def graph_launch(self):
        _native.check(self._lib.vx_tdma_plan_graph_launch(self._plan, stream_handle()),
                      "vx_tdma_plan_graph_launch")

    def info(self):
        nb = ctypes.c_int32()
        tot = ctypes.c_int64()
        ch = (ctypes.c_int32 * 24)()
        _native.check(self._lib.vx_tdma_plan_info(self._plan, ctypes.byref(nb), ctypes.byref(tot), ch, 8),
                      "vx_tdma_plan_info")
        buckets = [dict(nz_bucket=ch[3 * q], block=ch[3 * q + 1], blocks_per_cu=ch[3 * q + 2])
                   for q in range(min(nb.value, 8))]
        return dict(total_columns=tot.value, buckets=buckets)

    def close(self):
        if getattr(self, "_plan", None):
            self._lib.vx_tdma_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_systems(n, seed=1003, nxy=(48, 720), nz_range=(15, 100), land_frac=0.3, device="cuda"):
    """``n`` vertical-diffusion systems with nx, ny ~ U{nxy}, nz ~ U{nz_range},
    kbot ~ U{1..nz} with a land fraction set to 0; a = c = -dt*kappa/dz^2,
    kappa ~ LogU(1e-5, 1e-2), b = 1 - a - c, d ~ N(0, 1); per-system seed
    ``seed + g`` (the batched 64-grid workload)."""
    import math

    import numpy as np

    from .workloads import geometric_dz

    out = []
    for gi in range(n):
        rng = np.random.default_rng(seed + gi)
        nx, ny = (int(v) for v in rng.integers(nxy[0], nxy[1] + 1, size=2))
        nz = int(rng.integers(nz_range[0], nz_range[1] + 1))
        g = torch.Generator(device=device)
        g.manual_seed(seed + gi)
        f64 = dict(dtype=torch.float64, device=device)
        dz = torch.tensor(geometric_dz(nz, 10.0, 250.0), **f64)
        lo, hi = math.log(1e-5), math.log(1e-2)
        kappa = torch.exp(lo + (hi - lo) * torch.rand((nx, ny, nz), generator=g, **f64))
        a = -3600.0 * kappa / dz ** 2
        del kappa
        c = a.clone()
        b = 1.0 - a - c
        d = torch.randn((nx, ny, nz), generator=g, **f64)
        kb = rng.integers(1, nz + 1, size=(nx, ny)).astype(np.int32)
        kb[rng.random((nx, ny)) < land_frac] = 0
        out.append(dict(a=a, b=b, c=c, d=d, kbot=torch.from_numpy(kb).to(device), x=torch.empty_like(d)))
    return out
