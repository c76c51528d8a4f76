"""Oracle: kbot-masked batched tridiagonal solve (TEST INFRASTRUCTURE ONLY).

Restates veros/core/utilities.py:38-59 and veros/core/operators.py:60-77,133-156.

Pins:
* ``solve_tridiagonal_dgtsv`` / ``solve_implicit`` — golden vectors written by
  tests/golden/make_golden.py from the reference's NumPy path; dense
  ``numpy.linalg.solve`` per column on tiny grids; closed forms (``a=c=0``,
  single wet cell, constant-preserving diffusion rows).
* ``solve_tridiagonal_thomas`` — agrees with ``dgtsv`` to O(cond*eps) on
  diagonally dominant systems (tests/test_oracle_tdma.py).
"""
from __future__ import annotations

import numpy as np


def create_water_masks(ks, nz):
    """(land, water, edge) masks from 1-based bottom indices
    (veros/core/utilities.py:38-48). ``ks`` is kbot; 0 means land."""
    ks0 = np.asarray(ks) - 1
    land = ks0 >= 0
    levels = np.arange(nz)[None, None, :]
    water = land[:, :, None] & (levels >= ks0[:, :, None])
    edge = land[:, :, None] & (levels == ks0[:, :, None])
    return land, water, edge


def kbot_from_masks(water_mask, edge_mask):
    """Inverse of create_water_masks for masks of that form: kbot = edge index + 1."""
    has_edge = edge_mask.any(axis=-1)
    return np.where(has_edge, np.argmax(edge_mask, axis=-1) + 1, 0).astype(np.int32)


def solve_tridiagonal_dgtsv(a, b, c, d, water_mask, edge_mask):
    """All wet cells form one long system, decoupled at the edges by zeroing
    the sub-diagonal there and the super-diagonal at the top level; LAPACK
    ``dgtsv`` solves it and the solution is scattered back, zeros elsewhere
    (veros/core/operators.py:60-77)."""
    from scipy.linalg import lapack

    out = np.zeros(a.shape, dtype=a.dtype)
    if not np.any(water_mask):
        return out
    sub = np.array(a, copy=True)
    sup = np.array(c, copy=True)
    sub[edge_mask] = 0
    sup[..., -1] = 0
    dl = sub[water_mask][1:]
    du = sup[water_mask][:-1]
    x = lapack.dgtsv(dl, np.asarray(b)[water_mask], du, np.asarray(d)[water_mask])[3]
    out[water_mask] = x
    return out


def solve_tridiagonal_thomas(a, b, c, d, water_mask, edge_mask):
    """Column-wise Thomas sweep with land rows turned into identity rows
    (veros/core/operators.py:133-156): z runs upward from the edge cell."""
    w = water_mask
    a = np.where(w & ~edge_mask, a, 0.0)
    b = np.where(w, b, 1.0)
    c = np.where(w, c, 0.0)
    d = np.where(w, d, 0.0)
    nz = a.shape[-1]
    cp = np.zeros(a.shape[:-1])
    dp = np.zeros(a.shape[:-1])
    cps = np.empty(a.shape)
    dps = np.empty(a.shape)
    for k in range(nz):
        den = b[..., k] - a[..., k] * cp
        cp = c[..., k] / den
        dp = (d[..., k] - a[..., k] * dp) / den
        cps[..., k] = cp
        dps[..., k] = dp
    x = np.zeros(a.shape)
    nxt = np.zeros(a.shape[:-1])
    for k in range(nz - 1, -1, -1):
        nxt = dps[..., k] - cps[..., k] * nxt
        x[..., k] = nxt
    return np.where(w, x, 0.0)


def solve_tridiagonal_dense(a, b, c, d, water_mask, edge_mask):
    """Brute force: assemble each column's wet block densely and call
    ``numpy.linalg.solve``. Tiny grids only."""
    out = np.zeros(a.shape)
    nx, ny, nz = a.shape
    for i in range(nx):
        for j in range(ny):
            wet = np.nonzero(water_mask[i, j])[0]
            if wet.size == 0:
                continue
            k0 = wet[0]
            n = nz - k0
            A = np.zeros((n, n))
            for r in range(n):
                k = k0 + r
                A[r, r] = b[i, j, k]
                if r > 0:
                    A[r, r - 1] = a[i, j, k]
                if r < n - 1:
                    A[r, r + 1] = c[i, j, k]
            out[i, j, k0:] = np.linalg.solve(A, d[i, j, k0:])
    return out


def solve_implicit(a, b, c, d, water_mask, edge_mask, b_edge=None, d_edge=None, solver="dgtsv"):
    """Edge substitution, then the masked tridiagonal solve
    (veros/core/utilities.py:51-59)."""
    if b_edge is not None:
        b = np.where(edge_mask, b_edge, b)
    if d_edge is not None:
        d = np.where(edge_mask, d_edge, d)
    fn = {"dgtsv": solve_tridiagonal_dgtsv, "thomas": solve_tridiagonal_thomas,
          "dense": solve_tridiagonal_dense}[solver]
    return fn(a, b, c, d, water_mask, edge_mask)


def solve_implicit_kbot(a, b, c, d, kbot, b_edge=None, d_edge=None, solver="dgtsv"):
    """Convenience: masks from kbot (nx, ny), then solve_implicit."""
    _, water, edge = create_water_masks(kbot, a.shape[-1])
    return solve_implicit(a, b, c, d, water, edge, b_edge=b_edge, d_edge=d_edge, solver=solver)
