"""Drop-in for the TDMA half of veros.core.utilities on the GPU.

``solve_implicit`` keeps the reference signature
``solve_implicit(a, b, c, d, water_mask, edge_mask, b_edge=None, d_edge=None)``
(veros/core/utilities.py:51-59) and returns the same array: the solution on
wet cells, 0 elsewhere. The masks must be the ones ``create_water_masks``
builds from ``kbot`` (veros/core/utilities.py:38-48), i.e. each column's wet
cells are a suffix starting at its edge cell; they are reduced to ``kbot`` on
the device. Callers that already hold ``kbot`` pass it as a keyword and skip
the masks entirely (``solve_implicit_kbot``).
"""
from __future__ import annotations

import torch

from . import _native
from ._arrays import ptr, stream_handle, to_device

PRECISIONS = ("fp64", "mixed")


def create_water_masks(ks, nz):
    """(land, water, edge) boolean masks from 1-based bottom indices
    (veros/core/utilities.py:38-48)."""
    ks, _ = to_device(ks)
    ks0 = ks.to(torch.int64) - 1
    land = ks0 >= 0
    levels = torch.arange(nz, device=ks.device)
    water = land[:, :, None] & (levels[None, None, :] >= ks0[:, :, None])
    edge = land[:, :, None] & (levels[None, None, :] == ks0[:, :, None])
    return land, water, edge


def kbot_from_masks(water_mask, edge_mask):
    """kbot = (index of the edge cell) + 1, 0 where a column has no edge."""
    edge, _ = to_device(edge_mask)
    e = edge.to(torch.uint8)
    has = e.amax(dim=-1) > 0
    first = torch.argmax(e, dim=-1)
    return torch.where(has, first + 1, torch.zeros_like(first)).to(torch.int32).contiguous()


def _same_layout(ts):
    s = ts[0].stride()
    return all(t.stride() == s for t in ts) and s[-1] == 1


def solve_implicit_kbot(a, b, c, d, kbot, b_edge=None, d_edge=None, *, precision="fp64",
                        refine_steps=1, out=None):
    """Solve the kbot-masked tridiagonal systems; see include/vxmix.h (vx_tdma_f64)."""
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}")
    arrs = {"a": a, "b": b, "c": c, "d": d, "b_edge": b_edge, "d_edge": d_edge}
    dev = {}
    from_np = False
    for k, v in arrs.items():
        if v is None:
            dev[k] = None
            continue
        t, was_np = to_device(v)
        from_np |= was_np
        dev[k] = t
    shapes = {t.shape for t in dev.values() if t is not None}
    if len(shapes) != 1:
        raise ValueError("all inputs must have identical shape")
    dtypes = {t.dtype for t in dev.values() if t is not None}
    if len(dtypes) != 1:
        raise ValueError("all inputs must have the same dtype")
    if dtypes != {torch.float64}:
        raise TypeError(f"unsupported dtype {dtypes.pop()} (float64 expected)")
    (shape,) = shapes
    if len(shape) != 3:
        raise ValueError(f"expected (nx, ny, nz) arrays, got shape {tuple(shape)}")
    nx, ny, nz = shape
    kb, _ = to_device(kbot, torch.int32)
    if tuple(kb.shape) != (nx, ny):
        raise ValueError(f"kbot must have shape {(nx, ny)}, got {tuple(kb.shape)}")
    ins = [t for t in dev.values() if t is not None]
    if not _same_layout(ins):
        for k in dev:
            if dev[k] is not None:
                dev[k] = dev[k].contiguous()
    if out is None:
        out = torch.empty_strided(shape, dev["a"].stride(), dtype=torch.float64, device=dev["a"].device)
    elif out.stride() != dev["a"].stride() or out.shape != shape:
        raise ValueError("out must match the inputs' shape and strides")
    sx, sy, _ = dev["a"].stride()
    L = _native.lib()
    args = [stream_handle(), ptr(dev["a"]), ptr(dev["b"]), ptr(dev["c"]), ptr(dev["d"]),
            ptr(dev["b_edge"]), ptr(dev["d_edge"]), ptr(kb), kb.stride(0), kb.stride(1), ptr(out),
            nx, ny, nz, sx, sy]
    if precision == "fp64":
        _native.check(L.vx_tdma_f64(*args), "vx_tdma_f64")
    else:
        _native.check(L.vx_tdma_mixed(*args, int(refine_steps)), "vx_tdma_mixed")
    if from_np:
        return out.cpu().numpy()
    return out


def solve_implicit(a, b, c, d, water_mask=None, edge_mask=None, b_edge=None, d_edge=None, *, kbot=None,
                   precision="fp64", refine_steps=1, out=None):
    """veros.core.utilities.solve_implicit on the GPU (veros/core/utilities.py:51-59)."""
    if kbot is None:
        if edge_mask is None:
            raise ValueError("either edge_mask or kbot is required")
        kbot = kbot_from_masks(water_mask, edge_mask)
    return solve_implicit_kbot(a, b, c, d, kbot, b_edge=b_edge, d_edge=d_edge, precision=precision,
                               refine_steps=refine_steps, out=out)


# most launches one enforce_boundaries call may split a view into
_MAX_BOUNDARY_RUNS = 64


def _even_run(t):
    """(elements, stride) when the offsets of one x row of ``t`` (all axes
    but the first) are exactly {e * stride : 0 <= e < elements}, else None."""
    dims = sorted((int(st), int(d)) for d, st in zip(t.shape[1:], t.stride()[1:]) if d > 1)
    if not dims:
        return 1, 1
    es, n = dims[0][0], 1
    for st, d in dims:
        if st != n * es:
            return None
        n *= d
    return n, es


def _boundary_runs(t, out):
    """Split ``t`` along its widest-strided row axis until every piece's rows
    are one evenly strided run; append (view, elements, stride) to ``out``."""
    run = _even_run(t)
    if run is not None:
        out.append((t,) + run)
        return
    ax = max((a for a in range(1, t.dim()) if t.shape[a] > 1), key=lambda a: t.stride(a))
    for i in range(t.shape[ax]):
        _boundary_runs(t.select(ax, i), out)
        if len(out) > _MAX_BOUNDARY_RUNS:
            raise ValueError(f"a view of strides {tuple(t.stride())} splits into more than "
                             f"{_MAX_BOUNDARY_RUNS} evenly strided runs per x row")


def enforce_boundaries(arr, enable_cyclic_x, local=False):
    """The one-process halo update of veros/core/utilities.py:7-21, in place
    on the GPU (vx_enforce_boundaries_f64): with cyclic x, arr[-2:] =
    arr[2:4] and arr[:2] = arr[-4:-2] along the first axis. ``arr`` is a
    float64 torch tensor on the GPU: any contiguous array, a whole
    time-level variable or one level of it in either layout
    (veros_amd.layout), or another view whose x rows split into at most 64
    evenly strided runs (one launch each). Returns ``arr``."""
    if not isinstance(arr, torch.Tensor) or arr.device.type != "cuda":
        raise TypeError("arr must be a torch tensor on the GPU")
    if arr.dtype != torch.float64:
        raise TypeError(f"arr must be torch.float64 (got {arr.dtype})")
    if not enable_cyclic_x:
        return arr
    X = arr.shape[0]
    runs = []
    _boundary_runs(arr, runs)
    for view, n, es in runs:
        rc = _native.lib().vx_enforce_boundaries_f64(stream_handle(), view.data_ptr(), int(X), int(n),
                                                     int(view.stride(0)), int(es), 1)
        _native.check(rc, "enforce_boundaries")
    return arr
