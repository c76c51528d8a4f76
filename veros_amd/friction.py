"""Drop-in for veros.core.friction.implicit_vert_friction on the GPU.

* ``implicit_vert_friction(state) -> KernelOutput(u, v, du_mix, dv_mix,
  K_diss_v)`` (veros/core/friction.py:90-205): implicit vertical friction of
  the horizontal velocity with the viscosity ``kappaM``, the mixing tendencies
  ``du_mix`` / ``dv_mix``, and the dissipation added to ``K_diss_v``.

``state`` is any object with ``.variables`` and ``.settings`` holding the veros
variables and settings read here, as torch float64 tensors on the GPU
(``kbot`` int32). The outputs are written IN PLACE into the state's arrays and
returned in the KernelOutput; as in the reference, ``du_mix`` / ``dv_mix`` are
overwritten on the face columns ``[1:-2, 1:-2]`` only.
"""
from __future__ import annotations

import ctypes

import torch

from . import _native
from ._arrays import ptr, stream_handle
from .isoneutral import KernelOutput, _check_array, _check_levels
from .layout import common_layout


def implicit_vert_friction(state):
    vs, st = state.variables, state.settings
    X, Y, nz = vs.kappaM.shape
    kbot = _check_array("kbot", vs.kbot, (X, Y), dtype=torch.int32)
    dzt = _check_array("dzt", vs.dzt, (nz,))
    dzw = _check_array("dzw", vs.dzw, (nz,))
    dxt = _check_array("dxt", vs.dxt, (X,))
    dxu = _check_array("dxu", vs.dxu, (X,))
    u = _check_levels("u", vs.u, (X, Y, nz, 3))
    v = _check_levels("v", vs.v, (X, Y, nz, 3))
    K = _check_array("kappaM", vs.kappaM, (X, Y, nz))
    at = _check_array("area_t", vs.area_t, (X, Y))
    av = _check_array("area_v", vs.area_v, (X, Y))
    du = _check_array("du_mix", vs.du_mix, (X, Y, nz))
    dv = _check_array("dv_mix", vs.dv_mix, (X, Y, nz))
    kd = _check_array("K_diss_v", vs.K_diss_v, (X, Y, nz))
    g = _native.VxGrid()
    g.kbot, g.dzt, g.dzw, g.dxt, g.dxu = ptr(kbot), ptr(dzt), ptr(dzw), ptr(dxt), ptr(dxu)
    g.nx, g.ny, g.nz = X - 4, Y - 4, nz
    g.time_layout = common_layout(u=u, v=v)
    L = _native.lib()
    rc = L.vx_implicit_vert_friction_f64(
        stream_handle(), ctypes.byref(g), float(st.dt_mom), ptr(u), ptr(v), int(vs.tau), int(vs.taup1), ptr(K),
        ptr(at), ptr(av), ptr(du), ptr(dv), ptr(kd))
    _native.check(rc, "implicit_vert_friction")
    return KernelOutput(u=u, v=v, du_mix=du, dv_mix=dv, K_diss_v=kd)
