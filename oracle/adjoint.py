"""Oracle: reverse-mode derivatives of the kbot-masked TDMA solve and of one
isoneutral tracer step (TEST INFRASTRUCTURE ONLY).

The reference differentiates these paths with JAX's autodiff of its NumPy/JAX
code (veros/core/utilities.py:51-59, veros/core/isoneutral/diffusion.py:172-207);
JAX is not available here, so the derivatives are restated from the
mathematics, independently of the GPU adjoint kernels:

* ``tdma_vjp`` — per wet column, the dense matrix A of the masked system
  (veros/core/operators.py:60-77 with the edge substitution of
  veros/core/utilities.py:51-59) is built and ``numpy.linalg.solve(A.T, gx)``
  gives lam; then dL/dd = lam and dL/dA[r, s] = -lam[r] x[s] on the three
  diagonals.
* ``iso_step_vjp`` — for a fixed isoneutral tensor the step
  (T[tau], T[taup1], K_iso) -> T_new is affine in (T[tau], T[taup1]) and
  affine in K_iso, so its Jacobian is read off exactly (to rounding) by
  unit perturbations through the oracle forward step, one input entry at a
  time (brute force; tiny grids only).

Pins: tests/test_oracle_adjoint.py checks both against central finite
differences (eps = 1e-6) of the oracle forward and checks the affine
property the brute force relies on.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from .isoneutral import isoneutral_diffusion_tracer
from .tdma import solve_implicit_kbot


def tdma_vjp(a, b, c, kbot, x, gx, b_edge=None, d_edge=None):
    """(ga, gb, gc, gd, gb_edge, gd_edge) for x = solve_implicit_kbot(a, b, c, d, kbot, b_edge, d_edge).

    gb_edge / gd_edge are None when b_edge / d_edge are; otherwise the edge
    row's derivative is reported there (and is zero in gb / gd)."""
    nx, ny, nz = a.shape
    ga, gb, gc, gd = (np.zeros((nx, ny, nz)) for _ in range(4))
    gbe = None if b_edge is None else np.zeros((nx, ny, nz))
    gde = None if d_edge is None else np.zeros((nx, ny, nz))
    for i in range(nx):
        for j in range(ny):
            kb = int(kbot[i, j])
            if kb < 1 or kb > nz:
                continue
            k0 = kb - 1
            n = nz - k0
            A = np.zeros((n, n))
            for r in range(n):
                k = k0 + r
                A[r, r] = b_edge[i, j, k] if (r == 0 and b_edge is not None) else b[i, j, k]
                if r > 0:
                    A[r, r - 1] = a[i, j, k]
                if r < n - 1:
                    A[r, r + 1] = c[i, j, k]
            lam = np.linalg.solve(A.T, gx[i, j, k0:])
            xs = x[i, j, k0:]
            gd[i, j, k0:] = lam
            gb[i, j, k0:] = -lam * xs
            ga[i, j, k0 + 1:] = -lam[1:] * xs[:-1]
            gc[i, j, k0:nz - 1] = -lam[:-1] * xs[1:]
            if gde is not None:
                gde[i, j, k0] = gd[i, j, k0]
                gd[i, j, k0] = 0.0
            if gbe is not None:
                gbe[i, j, k0] = gb[i, j, k0]
                gb[i, j, k0] = 0.0
    return ga, gb, gc, gd, gbe, gde


def _with_K(state, K_iso):
    vs = SimpleNamespace(**vars(state.variables))
    vs.K_iso = K_iso
    return SimpleNamespace(variables=vs, settings=state.settings)


def iso_step(state, T_tau, T_taup1, K_iso):
    """T_new = tracer[..., taup1] after one isoneutral step (iso=True) of the
    tracer whose levels tau / taup1 are T_tau / T_taup1, with K_iso replaced
    and the state's isoneutral tensor held fixed
    (veros/core/isoneutral/diffusion.py:172-207)."""
    vs = state.variables
    X, Y, nz = vs.K_iso.shape
    tr = np.zeros((X, Y, nz, 3))
    tr[..., vs.tau] = T_tau
    tr[..., vs.taup1] = T_taup1
    st = _with_K(state, K_iso)
    new, _, _, _, _ = isoneutral_diffusion_tracer(st, tr, np.zeros((X, Y, nz)), iso=True)
    return new[..., vs.taup1]


def iso_step_vjp(state, T_tau, T_taup1, K_iso, g_new):
    """(dL/dT_tau, dL/dT_taup1, dL/dK_iso) for L = sum(g_new * iso_step(...)),
    by unit perturbations (exact for the affine dependence on each input)."""
    base = iso_step(state, T_tau, T_taup1, K_iso)
    L0 = float(np.sum(g_new * base))
    out = []
    inputs = [T_tau, T_taup1, K_iso]
    for which in range(3):
        x = inputs[which]
        h = max(float(np.max(np.abs(x))), 1.0)
        g = np.zeros(x.shape)
        for idx in np.ndindex(x.shape):
            xp = np.array(x, copy=True)
            xp[idx] += h
            args = list(inputs)
            args[which] = xp
            g[idx] = (float(np.sum(g_new * iso_step(state, *args))) - L0) / h
        out.append(g)
    return tuple(out)


def mixing_step(state, a, b, c, d, T_tau, K_iso, b_edge=None, d_edge=None):
    """The differentiated step of the adjoint workload: a solve_implicit
    vertical diffusion of d on the interior columns (kbot of the state), whose
    result becomes the tracer's taup1 level (halo from the state's tracer),
    followed by one isoneutral step with T[tau] = T_tau and K_iso."""
    vs = state.variables
    kb = vs.kbot[2:-2, 2:-2]
    x = solve_implicit_kbot(a, b, c, d, kb, b_edge=b_edge, d_edge=d_edge)
    T_taup1 = np.array(state.variables.temp[..., vs.taup1], copy=True)
    T_taup1[2:-2, 2:-2, :] = x
    return iso_step(state, T_tau, T_taup1, K_iso)
