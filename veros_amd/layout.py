"""Device state layout of the time-level variables.

veros stores every prognostic with three time levels as one array with the
level innermost, ``temp (X, Y, nz, 3)`` (veros/variables.py:338, ``time``
dimension last), and steps by rotating the indices ``tau, taup1, taum1``
(veros/veros.py:322), never the data. With the level innermost, a 64-byte
line of ``temp`` holds 2-3 levels of a few cells, and a kernel that reads
one level brings in all three.

``to_levels_outer`` re-stores such a variable with the levels outermost -- a
``(3, X, Y, nz)`` buffer, level ``l`` contiguous -- and hands back the
``(X, Y, nz, 3)`` view of it (strides ``(Y*nz, nz, 1, X*Y*nz)``). Code that
indexes the variable as veros does (``temp[..., tau]``, ``temp[2:-2, 2:-2,
:, taup1]``) sees the same values; the kernels detect the layout from the
strides (``levels_layout``) and read one level per line. The rotation of the
time indices stays index arithmetic (``rotate_time_levels``).
"""
from __future__ import annotations

import torch

from ._native import VX_LEVELS_INNER as LEVELS_INNER  # (X, Y, nz, 3) C-contiguous (the reference's layout)
from ._native import VX_LEVELS_OUTER as LEVELS_OUTER  # the (X, Y, nz, 3) view of a (3, X, Y, nz) buffer

# the time-level variables of the mixing path
TIME_LEVEL_VARS = ("temp", "salt", "int_drhodT", "int_drhodS", "u", "v", "tke", "eke", "E_iw")


def levels_layout(t: torch.Tensor) -> int:
    """LEVELS_INNER or LEVELS_OUTER for an (X, Y, nz, 3) tensor; ValueError
    for any other layout."""
    if t.dim() != 4 or t.shape[3] != 3:
        raise ValueError(f"expected an (X, Y, nz, 3) array, got shape {tuple(t.shape)}")
    X, Y, nz, _ = t.shape
    if t.is_contiguous():
        return LEVELS_INNER
    if t.stride() == (Y * nz, nz, 1, X * Y * nz):
        return LEVELS_OUTER
    raise ValueError(f"time-level array with unsupported strides {t.stride()} (expected levels innermost, "
                     f"{(Y * nz * 3, nz * 3, 3, 1)}, or outermost, {(Y * nz, nz, 1, X * Y * nz)})")


def to_levels_outer(t: torch.Tensor) -> torch.Tensor:
    """The (X, Y, nz, 3) view of a new (3, X, Y, nz) buffer holding ``t``'s values."""
    buf = t.permute(3, 0, 1, 2).contiguous()
    return buf.permute(1, 2, 3, 0)


def to_levels_inner(t: torch.Tensor) -> torch.Tensor:
    """The reference's C-contiguous (X, Y, nz, 3) copy of ``t``."""
    return t.contiguous()


def set_state_layout(state, layout: int, names=TIME_LEVEL_VARS):
    """Re-store the state's time-level variables in ``layout`` (in place on
    ``state.variables``; variables the state lacks are skipped)."""
    vs = state.variables
    conv = to_levels_outer if layout == LEVELS_OUTER else to_levels_inner
    for n in names:
        t = getattr(vs, n, None)
        if isinstance(t, torch.Tensor) and levels_layout(t) != layout:
            setattr(vs, n, conv(t))
    return state


def common_layout(**arrays) -> int:
    """The one layout shared by the given (X, Y, nz, 3) arrays (None entries
    skipped); ValueError if they differ."""
    found = {n: levels_layout(t) for n, t in arrays.items() if t is not None}
    kinds = set(found.values())
    if len(kinds) > 1:
        raise ValueError(f"time-level arrays of one call must share a layout (got {found})")
    return kinds.pop() if kinds else LEVELS_INNER


def rotate_time_levels(vs):
    """tau, taup1, taum1 = taup1, taum1, tau (veros/veros.py:322): the level
    indices move, the data stays."""
    vs.tau, vs.taup1, vs.taum1 = vs.taup1, vs.taum1, vs.tau
    return vs
