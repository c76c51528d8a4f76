"""Oracle: implicit vertical friction of the horizontal velocity (TEST
INFRASTRUCTURE ONLY).

Restates, in NumPy float64,
* ``implicit_vert_friction``   veros/core/friction.py:90-205
* ``ugrid_to_tgrid``           veros/core/numerics.py:313-324
* ``vgrid_to_tgrid``           veros/core/numerics.py:328-336
on top of this package's ``solve_implicit`` (veros/core/utilities.py:51-59).

Pins (tests/test_oracle_friction.py): with kappaM = 0 the solve is the
identity, so u and v at taup1 take the tau values on the wet faces and the
dissipation is zero; a vertically uniform velocity column is a fixed point;
the solve conserves sum_k dzt * u per face column; every wet face column
agrees with a dense solve of the matrix built from the formulas of the
reference; and on a two-level grid with one moving face the dissipation and
its interpolation onto the two adjacent T cells have a closed form.

Each function works on a ``state`` with ``.variables`` (arrays shaped as veros
variables, 2-cell halos, z contiguous, k = 0 deepest) and ``.settings``.
Inputs are never modified.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from .tdma import create_water_masks, solve_implicit


def ugrid_to_tgrid(dxt, dxu, a):
    """veros/core/numerics.py:313-324."""
    b = np.array(a, copy=True)
    b[2:-2] = (dxu[2:-2, None, None] * a[2:-2] + dxu[1:-3, None, None] * a[1:-3]) / (2 * dxt[2:-2, None, None])
    return b


def vgrid_to_tgrid(area_t, area_v, a):
    """veros/core/numerics.py:328-336."""
    b = np.array(a, copy=True)
    b[:, 2:-2] = (area_v[:, 2:-2, None] * a[:, 2:-2] + area_v[:, 1:-3, None] * a[:, 1:-3]) / (2 * area_t[:, 2:-2, None])
    return b


def _face_system(vs, st, kap_face, mask, kss):
    """(a, b, c, b_edge, water, edge) of one velocity component on the
    [1:-2, 1:-2] face columns (veros/core/friction.py:111-126, 162-178)."""
    nz = kap_face.shape[2]
    dzt, dzw = np.asarray(vs.dzt, np.float64), np.asarray(vs.dzw, np.float64)
    _, water, edge = create_water_masks(kss, nz)
    m = np.asarray(mask, np.float64)[1:-2, 1:-2]
    delta = np.zeros(kap_face.shape)
    delta[:, :, :-1] = st.dt_mom / dzw[None, None, :-1] * kap_face[:, :, :-1] * m[:, :, 1:] * m[:, :, :-1]
    a = np.zeros(kap_face.shape)
    b = np.zeros(kap_face.shape)
    c = np.zeros(kap_face.shape)
    a[:, :, 1:] = -delta[:, :, :-1] / dzt[None, None, 1:]
    b[:, :, 1:] = 1 + delta[:, :, :-1] / dzt[None, None, 1:]
    b[:, :, 1:-1] += delta[:, :, 1:-1] / dzt[None, None, 1:-1]
    b_edge = 1 + delta / dzt[None, None, :]
    c[:, :, :-1] = -delta[:, :, :-1] / dzt[None, None, :-1]
    return a, b, c, b_edge, water, edge


def _face_dissipation(vs, kap_face, mask, vel, tau, taup1):
    """The vertical-friction dissipation on the faces before interpolation
    (veros/core/friction.py:137-153, 189-201)."""
    X, Y, nz = vel.shape[:3]
    dzw = np.asarray(vs.dzw, np.float64)
    m = np.asarray(mask, np.float64)[1:-2, 1:-2]
    new = vel[1:-2, 1:-2, :, taup1]
    old = vel[1:-2, 1:-2, :, tau]
    flux = kap_face[:, :, :-1] * (new[:, :, 1:] - new[:, :, :-1]) / dzw[:-1] * m[:, :, 1:] * m[:, :, :-1]
    diss = np.zeros((X, Y, nz))
    diss[1:-2, 1:-2, :-1] = (old[:, :, 1:] - old[:, :, :-1]) * flux / dzw[:-1]
    diss[:, :, -1] = 0.0
    return diss


def implicit_vert_friction(state, solver="dgtsv"):
    """The kernel (veros/core/friction.py:90-205): returns
    SimpleNamespace(u, v, du_mix, dv_mix, K_diss_v) as new arrays."""
    vs, st = state.variables, state.settings
    tau, taup1 = int(vs.tau), int(vs.taup1)
    kbot = np.asarray(vs.kbot)
    kap = np.asarray(vs.kappaM, np.float64)
    dt = st.dt_mom
    u = np.array(vs.u, dtype=np.float64, copy=True)
    v = np.array(vs.v, dtype=np.float64, copy=True)
    du_mix = np.array(vs.du_mix, dtype=np.float64, copy=True)
    dv_mix = np.array(vs.dv_mix, dtype=np.float64, copy=True)
    K = np.array(vs.K_diss_v, dtype=np.float64, copy=True)

    # zonal momentum (:108-136): faces between T columns i and i+1
    kss = np.maximum(kbot[1:-2, 1:-2], kbot[2:-1, 1:-2])
    fxa = 0.5 * (kap[1:-2, 1:-2] + kap[2:-1, 1:-2])
    a, b, c, be, water, edge = _face_system(vs, st, fxa, vs.maskU, kss)
    sol = solve_implicit(a, b, c, u[1:-2, 1:-2, :, tau], water, edge, b_edge=be, solver=solver)
    u[1:-2, 1:-2, :, taup1] = np.where(water, sol, u[1:-2, 1:-2, :, taup1])
    du_mix[1:-2, 1:-2] = (u[1:-2, 1:-2, :, taup1] - u[1:-2, 1:-2, :, tau]) / dt
    diss = _face_dissipation(vs, fxa, vs.maskU, u, tau, taup1)
    K = K + ugrid_to_tgrid(np.asarray(vs.dxt), np.asarray(vs.dxu), diss)

    # meridional momentum (:158-203): faces between T rows j and j+1
    kss = np.maximum(kbot[1:-2, 1:-2], kbot[1:-2, 2:-1])
    fxa = 0.5 * (kap[1:-2, 1:-2] + kap[1:-2, 2:-1])
    a, b, c, be, water, edge = _face_system(vs, st, fxa, vs.maskV, kss)
    sol = solve_implicit(a, b, c, v[1:-2, 1:-2, :, tau], water, edge, b_edge=be, solver=solver)
    v[1:-2, 1:-2, :, taup1] = np.where(water, sol, v[1:-2, 1:-2, :, taup1])
    dv_mix[1:-2, 1:-2] = (v[1:-2, 1:-2, :, taup1] - v[1:-2, 1:-2, :, tau]) / dt
    diss = _face_dissipation(vs, fxa, vs.maskV, v, tau, taup1)
    K = K + vgrid_to_tgrid(np.asarray(vs.area_t), np.asarray(vs.area_v), diss)
    return SimpleNamespace(u=u, v=v, du_mix=du_mix, dv_mix=dv_mix, K_diss_v=K)
