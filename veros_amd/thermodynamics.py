"""Drop-in for veros.core.thermodynamics.vertmix_tempsalt on the GPU.

* ``vertmix_tempsalt(state) -> KernelOutput(dtemp_vmix, temp, dsalt_vmix, salt)``
  (veros/core/thermodynamics.py:248-300): implicit vertical mixing of
  temperature and salinity with ``kappaH`` and the surface forcing
  ``forc_temp_surface`` / ``forc_salt_surface``, the tendencies
  ``dtemp_vmix`` / ``dsalt_vmix``, and the one-process halo update of level
  ``taup1`` (``enforce_boundaries``, veros/core/utilities.py:7-21).

``state`` is any object with ``.variables`` and ``.settings`` holding the veros
variables and settings read here, as torch float64 tensors on the GPU
(``kbot`` int32). The outputs are written IN PLACE into the state's arrays and
returned in the KernelOutput; the tendencies are overwritten everywhere, as in
the reference, with zeros off the wet cells.
"""
from __future__ import annotations

import ctypes

import torch

from . import _native
from ._arrays import ptr, stream_handle
from .isoneutral import KernelOutput, _check_array, _check_levels
from .layout import common_layout


def vertmix_tempsalt(state):
    vs, st = state.variables, state.settings
    X, Y, nz = vs.kappaH.shape
    kbot = _check_array("kbot", vs.kbot, (X, Y), dtype=torch.int32)
    dzt = _check_array("dzt", vs.dzt, (nz,))
    dzw = _check_array("dzw", vs.dzw, (nz,))
    T = _check_levels("temp", vs.temp, (X, Y, nz, 3))
    S = _check_levels("salt", vs.salt, (X, Y, nz, 3))
    K = _check_array("kappaH", vs.kappaH, (X, Y, nz))
    fT = _check_array("forc_temp_surface", vs.forc_temp_surface, (X, Y))
    fS = _check_array("forc_salt_surface", vs.forc_salt_surface, (X, Y))
    dT = _check_array("dtemp_vmix", vs.dtemp_vmix, (X, Y, nz))
    dS = _check_array("dsalt_vmix", vs.dsalt_vmix, (X, Y, nz))
    g = _native.VxGrid()
    g.kbot, g.dzt, g.dzw = ptr(kbot), ptr(dzt), ptr(dzw)
    g.nx, g.ny, g.nz = X - 4, Y - 4, nz
    g.time_layout = common_layout(temp=T, salt=S)
    rc = _native.lib().vx_vertmix_tempsalt_f64(
        stream_handle(), ctypes.byref(g), float(st.dt_tracer), ptr(T), ptr(S), int(vs.taup1), ptr(K), ptr(fT),
        ptr(fS), ptr(dT), ptr(dS), int(bool(st.enable_cyclic_x)))
    _native.check(rc, "vertmix_tempsalt")
    return KernelOutput(dtemp_vmix=dT, temp=T, dsalt_vmix=dS, salt=S)
