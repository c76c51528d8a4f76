"""Oracle: implicit vertical mixing of temperature and salinity (TEST
INFRASTRUCTURE ONLY).

Restates, in NumPy float64,
* ``vertmix_tempsalt``     veros/core/thermodynamics.py:248-300
* ``enforce_boundaries``   veros/core/utilities.py:7-21 (one process: the
                           cyclic x halo copy)
on top of this package's ``solve_implicit`` (veros/core/utilities.py:51-59).

Pins (tests/test_oracle_vertmix.py): with kappaH = 0 the solve is the
identity, so the step adds dt * forcing / dzt[nz-1] to the top level of the
wet columns only and leaves every other cell unchanged; the diffusion rows sum
to one, so a vertically uniform tracer without forcing is a fixed point; the
solve conserves sum_k dzt * tracer per column up to the forcing,
dt * forcing; every wet column agrees with a dense solve of the matrix built
from the formulas of the reference; the tendencies are (new - old) / dt and
exactly zero off the wet cells; the cyclic halos copy rows 2, 3 and X-4, X-3.

Each function works on a ``state`` with ``.variables`` (arrays shaped as veros
variables, 2-cell halos, z contiguous, k = 0 deepest) and ``.settings``.
Inputs are never modified.
"""
from __future__ import annotations

import numpy as np

from .isoneutral import KernelOutput
from .tdma import create_water_masks, solve_implicit


def enforce_boundaries(arr, enable_cyclic_x):
    """One-process halo update (veros/core/utilities.py:7-21)."""
    arr = np.array(arr, copy=True)
    if enable_cyclic_x:
        arr[-2:, ...] = arr[2:4, ...]
        arr[:2, ...] = arr[-4:-2, ...]
    return arr


def vertmix_coefficients(state):
    """(a, b, c, b_edge) of the implicit vertical mixing system on the
    interior columns (veros/core/thermodynamics.py:260-276)."""
    vs, st = state.variables, state.settings
    nz = vs.kappaH.shape[2]
    dzt = np.asarray(vs.dzt, dtype=np.float64)
    dzw = np.asarray(vs.dzw, dtype=np.float64)
    kap = np.asarray(vs.kappaH, dtype=np.float64)[2:-2, 2:-2]
    shape = kap.shape
    delta = np.zeros(shape)
    delta[:, :, :-1] = st.dt_tracer / dzw[None, None, :-1] * kap[:, :, :-1]
    delta[:, :, -1] = 0.0
    a = np.zeros(shape)
    b = np.zeros(shape)
    c = np.zeros(shape)
    a[:, :, 1:] = -delta[:, :, :-1] / dzt[None, None, 1:]
    b[:, :, 1:] = 1 + (delta[:, :, 1:] + delta[:, :, :-1]) / dzt[None, None, 1:]
    b_edge = 1 + delta / dzt[None, None, :]
    c[:, :, :-1] = -delta[:, :, :-1] / dzt[None, None, :-1]
    return a, b, c, b_edge


def vertmix_tempsalt(state, solver="dgtsv"):
    """The kernel (veros/core/thermodynamics.py:248-300): returns
    KernelOutput(dtemp_vmix, temp, dsalt_vmix, salt) as new arrays."""
    vs, st = state.variables, state.settings
    nz = vs.kappaH.shape[2]
    taup1 = int(vs.taup1)
    dt = st.dt_tracer
    dzt = np.asarray(vs.dzt, dtype=np.float64)
    _, water, edge = create_water_masks(np.asarray(vs.kbot)[2:-2, 2:-2], nz)
    a, b, c, b_edge = vertmix_coefficients(state)
    out = {}
    for name, forc, dname in (("temp", vs.forc_temp_surface, "dtemp_vmix"),
                              ("salt", vs.forc_salt_surface, "dsalt_vmix")):
        tr = np.array(getattr(vs, name), dtype=np.float64, copy=True)
        old = tr[:, :, :, taup1].copy()
        d = tr[2:-2, 2:-2, :, taup1].copy()
        d[:, :, -1] += dt * np.asarray(forc, dtype=np.float64)[2:-2, 2:-2] / dzt[-1]
        sol = solve_implicit(a, b, c, d, water, edge, b_edge=b_edge, solver=solver)
        tr[2:-2, 2:-2, :, taup1] = np.where(water, sol, tr[2:-2, 2:-2, :, taup1])
        out[dname] = (tr[:, :, :, taup1] - old) / dt
        tr[..., taup1] = enforce_boundaries(tr[..., taup1], bool(st.enable_cyclic_x))
        out[name] = tr
    return KernelOutput(dtemp_vmix=out["dtemp_vmix"], temp=out["temp"], dsalt_vmix=out["dsalt_vmix"],
                        salt=out["salt"])
