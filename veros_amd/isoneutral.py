"""Drop-in for veros.core.isoneutral's hot pair on the GPU.

* ``isoneutral_diffusion_pre(state) -> KernelOutput(Ai_ez, Ai_nz, Ai_bx, Ai_by,
  K_11, K_22, K_33)``  (veros/core/isoneutral/isoneutral.py:18-229)
* ``isoneutral_diffusion_kernel(state, tr, istemp, iso=True) -> KernelOutput(temp|salt,
  dtemp_iso|dsalt_iso[, P_diss_iso])``  (veros/core/isoneutral/diffusion.py:210-283)
* ``isoneutral_diffusion(state, tr, istemp)``, the routine that stores the
  kernel's outputs in ``state.variables`` (veros/core/isoneutral/diffusion.py:286-295)
* ``isoneutral_skew_diffusion(state, tr, istemp)`` (veros/core/isoneutral/diffusion.py:298-307)
* ``isoneutral_diffusion_temp_salt(state)``: the temperature then salinity
  calls of veros/core/thermodynamics.py:431-432 in one pass
* ``isoneutral_pre_and_diffusion_temp_salt(state)``: pre followed by both
  calls (veros/core/thermodynamics.py:430-432) in one pass

``state`` is any object with ``.variables`` and ``.settings`` holding the veros
variables and settings these functions read (a ``VerosState`` whose arrays
are on the GPU, or ``veros_amd.state.to_device_state(...)``). Arrays are the
full veros arrays with 2-cell halos, C-contiguous, float64, on the GPU.

The outputs are written IN PLACE into the state's arrays over exactly the
index ranges the reference writes; entries outside those ranges keep their
values, which is what the reference's copy-and-update yields. The returned
KernelOutput holds the same tensors.

The equation of state must be the linear one (``eq_of_state_type == 1``,
whose drho/dT and drho/dS are constants, veros/core/density/get_rho.py:93-132);
the constants are taken from ``settings.drhodT`` / ``settings.drhodS`` when
present, else from veros's linear EOS module values
(veros/core/density/linear_eq.py:10-15,36-43).
"""
from __future__ import annotations

from collections import namedtuple

import torch

from . import _native
from ._arrays import ptr, require_contiguous, stream_handle, workspace
from .layout import LEVELS_INNER, common_layout, levels_layout

# veros/core/density/linear_eq.py:10-15
LINEAR_EOS_RHO0 = 1024.0
LINEAR_EOS_BETA_T = 1.67e-4
LINEAR_EOS_BETA_S = 0.78e-3


def KernelOutput(**kw):
    """namedtuple built from keyword arguments (veros/state.py:16-20)."""
    return namedtuple("KernelOutput", list(kw.keys()))(*kw.values())


def eos_derivatives(settings):
    if getattr(settings, "eq_of_state_type", 1) != 1:
        raise NotImplementedError("only the linear equation of state (eq_of_state_type=1) is supported")
    drhodT = getattr(settings, "drhodT", -LINEAR_EOS_BETA_T * LINEAR_EOS_RHO0)
    drhodS = getattr(settings, "drhodS", LINEAR_EOS_BETA_S * LINEAR_EOS_RHO0)
    return float(drhodT), float(drhodS)


def _check_array(name, t, shape, dtype=torch.float64):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"{name} must be a torch tensor on the GPU")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype} (got {t.dtype})")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"unexpected shape for {name} (expected {tuple(shape)}, got {tuple(t.shape)})")
    require_contiguous(name, t)
    return t


def _check_levels(name, t, shape):
    """A time-level array: (X, Y, nz, 3) float64 on the GPU with the levels
    innermost or outermost (veros_amd.layout)."""
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"{name} must be a torch tensor on the GPU")
    if t.dtype != torch.float64:
        raise TypeError(f"{name} must be {torch.float64} (got {t.dtype})")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"unexpected shape for {name} (expected {tuple(shape)}, got {tuple(t.shape)})")
    levels_layout(t)
    return t


def grid_struct(vs, keep, layout=LEVELS_INNER):
    """vx_grid for the state's grid (time-level arrays in ``layout``); ``keep``
    collects the tensors it points into."""
    kbot = vs.kbot
    X, Y = kbot.shape
    nz = vs.dzt.shape[0]
    kb = kbot if (kbot.dtype == torch.int32 and kbot.is_contiguous()) else kbot.to(torch.int32).contiguous()
    _check_array("kbot", kb, (X, Y), torch.int32)
    g = _native.VxGrid()
    g.kbot = kb.data_ptr()
    for name, n in (("dxt", X), ("dxu", X), ("dyt", Y), ("dyu", Y), ("cost", Y), ("cosu", Y),
                    ("dzt", nz), ("dzw", nz)):
        t = _check_array(name, getattr(vs, name), (n,))
        setattr(g, name, t.data_ptr())
        keep.append(t)
    keep.append(kb)
    g.nx, g.ny, g.nz = X - 4, Y - 4, nz
    g.time_layout = int(layout)
    return g


def params_struct(settings):
    p = _native.VxIsoParams()
    p.drhodT, p.drhodS = eos_derivatives(settings)
    p.K_iso_steep = float(settings.K_iso_steep)
    p.iso_slopec = float(settings.iso_slopec)
    p.iso_dslope = float(settings.iso_dslope)
    p.dt_tracer = float(settings.dt_tracer)
    p.grav = float(settings.grav)
    p.rho_0 = float(settings.rho_0)
    return p


def isoneutral_diffusion_pre(state):
    """Triad slopes, tapers and the isoneutral tensor, in place."""
    vs, st = state.variables, state.settings
    X, Y, nz = vs.K_iso.shape
    keep = []
    T = _check_levels("temp", vs.temp, (X, Y, nz, 3))
    S = _check_levels("salt", vs.salt, (X, Y, nz, 3))
    g = grid_struct(vs, keep, common_layout(temp=T, salt=S))
    p = params_struct(st)
    K = _check_array("K_iso", vs.K_iso, (X, Y, nz))
    outs = {}
    for name in ("Ai_ez", "Ai_nz", "Ai_bx", "Ai_by"):
        outs[name] = _check_array(name, getattr(vs, name), (X, Y, nz, 2, 2))
    for name in ("K_11", "K_22", "K_33"):
        outs[name] = _check_array(name, getattr(vs, name), (X, Y, nz))
    rc = _native.lib().vx_iso_pre_f64(stream_handle(), g, p, ptr(T), ptr(S), int(vs.tau), ptr(K),
                                      *(ptr(outs[n]) for n in ("Ai_ez", "Ai_nz", "Ai_bx", "Ai_by",
                                                               "K_11", "K_22", "K_33")))
    _native.check(rc, "isoneutral_diffusion_pre")
    return KernelOutput(**outs)


def isoneutral_diffusion_kernel(state, tr, istemp, iso=True):
    """One tracer's isoneutral (``iso=True``) or skew (``iso=False``, K_skew =
    K_gm) step, in place on ``tr`` and the state's dtemp_iso/dsalt_iso and
    P_diss_iso / P_diss_skew (veros/core/isoneutral/diffusion.py:210-283)."""
    vs, st = state.variables, state.settings
    X, Y, nz = vs.K_iso.shape if iso else vs.K_gm.shape
    keep = []
    p = params_struct(st)
    tr = _check_levels("tr", tr, (X, Y, nz, 3))
    dname = "dtemp_iso" if istemp else "dsalt_iso"
    dtr = _check_array(dname, getattr(vs, dname), (X, Y, nz))
    Kname = "K_iso" if iso else "K_gm"
    arrays = [_check_array(Kname, getattr(vs, Kname), (X, Y, nz))]
    arrays += [_check_array(n, getattr(vs, n), (X, Y, nz, 2, 2)) for n in ("Ai_ez", "Ai_nz", "Ai_bx", "Ai_by")]
    arrays += [_check_array(n, getattr(vs, n), (X, Y, nz)) for n in ("K_11", "K_22")]
    if iso:
        arrays.append(_check_array("K_33", vs.K_33, (X, Y, nz)))
    energy = bool(st.enable_conserve_energy)
    pname = "P_diss_iso" if iso else "P_diss_skew"
    if energy:
        I = _check_levels("int_drhodT" if istemp else "int_drhodS",
                          vs.int_drhodT if istemp else vs.int_drhodS, (X, Y, nz, 3))
        P = _check_array(pname, getattr(vs, pname), (X, Y, nz))
    else:
        I = P = None
    g = grid_struct(vs, keep, common_layout(tr=tr, int_drhodX=I))
    L = _native.lib()
    ws = workspace(L.vx_iso_workspace_bytes(g, 1), tr.device)
    fn = L.vx_iso_diffusion_f64 if iso else L.vx_iso_skew_diffusion_f64
    rc = fn(stream_handle(), g, p, ptr(tr), int(vs.tau), int(vs.taup1),
            *(ptr(a) for a in arrays), ptr(dtr), ptr(I), ptr(P), ptr(ws))
    _native.check(rc, "isoneutral_diffusion" if iso else "isoneutral_skew_diffusion")
    out = {("temp" if istemp else "salt"): tr, dname: dtr}
    if energy:
        out[pname] = P
    return KernelOutput(**out)


def _store(state, out):
    for k, v in out._asdict().items():
        setattr(state.variables, k, v)


def isoneutral_diffusion(state, tr, istemp):
    """The routine: run the kernel and store its outputs in state.variables
    (veros/core/isoneutral/diffusion.py:286-295)."""
    _store(state, isoneutral_diffusion_kernel(state, tr, istemp, iso=True))


def isoneutral_diffusion_temp_salt(state):
    """``isoneutral_diffusion(state, temp, True)`` then
    ``isoneutral_diffusion(state, salt, False)`` (the pair thermodynamics runs,
    veros/core/thermodynamics.py:431-432) in one pass over the isoneutral
    tensor; same results as the two calls."""
    vs, st = state.variables, state.settings
    X, Y, nz = vs.K_iso.shape
    keep = []
    p = params_struct(st)
    T = _check_levels("temp", vs.temp, (X, Y, nz, 3))
    S = _check_levels("salt", vs.salt, (X, Y, nz, 3))
    arrays = [_check_array("K_iso", vs.K_iso, (X, Y, nz))]
    arrays += [_check_array(n, getattr(vs, n), (X, Y, nz, 2, 2)) for n in ("Ai_ez", "Ai_nz", "Ai_bx", "Ai_by")]
    arrays += [_check_array(n, getattr(vs, n), (X, Y, nz)) for n in ("K_11", "K_22", "K_33")]
    dT = _check_array("dtemp_iso", vs.dtemp_iso, (X, Y, nz))
    dS = _check_array("dsalt_iso", vs.dsalt_iso, (X, Y, nz))
    if st.enable_conserve_energy:
        IT = _check_levels("int_drhodT", vs.int_drhodT, (X, Y, nz, 3))
        IS = _check_levels("int_drhodS", vs.int_drhodS, (X, Y, nz, 3))
        P = _check_array("P_diss_iso", vs.P_diss_iso, (X, Y, nz))
    else:
        IT = IS = P = None
    g = grid_struct(vs, keep, common_layout(temp=T, salt=S, int_drhodT=IT, int_drhodS=IS))
    L = _native.lib()
    ws = workspace(L.vx_iso_workspace_bytes(g, 2), T.device)
    rc = L.vx_iso_diffusion_ts_f64(stream_handle(), g, p, ptr(T), ptr(S), int(vs.tau), int(vs.taup1),
                                   *(ptr(a) for a in arrays), ptr(dT), ptr(dS), ptr(IT), ptr(IS), ptr(P), ptr(ws))
    _native.check(rc, "isoneutral_diffusion (temp, salt)")


def isoneutral_skew_diffusion(state, tr, istemp):
    """Skew (GM) variant of the routine (veros/core/isoneutral/diffusion.py:298-307)."""
    _store(state, isoneutral_diffusion_kernel(state, tr, istemp, iso=False))


def isoneutral_pre_and_diffusion_temp_salt(state):
    """``isoneutral_diffusion_pre(state)``, then
    ``isoneutral_diffusion(state, temp, True)`` and
    ``isoneutral_diffusion(state, salt, False)`` (the sequence of
    veros/core/thermodynamics.py:430-432), in place, in one pass: the face
    fluxes are formed while the isoneutral tensor is computed. Same results
    as the three calls."""
    vs, st = state.variables, state.settings
    X, Y, nz = vs.K_iso.shape
    keep = []
    p = params_struct(st)
    T = _check_levels("temp", vs.temp, (X, Y, nz, 3))
    S = _check_levels("salt", vs.salt, (X, Y, nz, 3))
    K = _check_array("K_iso", vs.K_iso, (X, Y, nz))
    tens = [_check_array(n, getattr(vs, n), (X, Y, nz, 2, 2)) for n in ("Ai_ez", "Ai_nz", "Ai_bx", "Ai_by")]
    tens += [_check_array(n, getattr(vs, n), (X, Y, nz)) for n in ("K_11", "K_22", "K_33")]
    dT = _check_array("dtemp_iso", vs.dtemp_iso, (X, Y, nz))
    dS = _check_array("dsalt_iso", vs.dsalt_iso, (X, Y, nz))
    if st.enable_conserve_energy:
        IT = _check_levels("int_drhodT", vs.int_drhodT, (X, Y, nz, 3))
        IS = _check_levels("int_drhodS", vs.int_drhodS, (X, Y, nz, 3))
        P = _check_array("P_diss_iso", vs.P_diss_iso, (X, Y, nz))
    else:
        IT = IS = P = None
    g = grid_struct(vs, keep, common_layout(temp=T, salt=S, int_drhodT=IT, int_drhodS=IS))
    L = _native.lib()
    ws = workspace(L.vx_iso_pre_diffusion_ts_workspace_bytes(g), T.device)
    rc = L.vx_iso_pre_diffusion_ts_f64(stream_handle(), g, p, ptr(T), ptr(S), int(vs.tau), int(vs.taup1), ptr(K),
                                       *(ptr(t) for t in tens), ptr(dT), ptr(dS), ptr(IT), ptr(IS), ptr(P),
                                       ptr(ws))
    _native.check(rc, "isoneutral_diffusion_pre + isoneutral_diffusion (temp, salt)")


def isoneutral_diag_streamfunction_kernel(state):
    """B1_gm, B2_gm from K_gm and the slopes (veros/core/isoneutral/isoneutral.py:232-255),
    in place (vx_iso_diag_streamfunction_f64)."""
    vs = state.variables
    X, Y, nz = vs.K_gm.shape
    keep = []
    g = grid_struct(vs, keep)
    K = _check_array("K_gm", vs.K_gm, (X, Y, nz))
    Aez = _check_array("Ai_ez", vs.Ai_ez, (X, Y, nz, 2, 2))
    Anz = _check_array("Ai_nz", vs.Ai_nz, (X, Y, nz, 2, 2))
    B1 = _check_array("B1_gm", vs.B1_gm, (X, Y, nz))
    B2 = _check_array("B2_gm", vs.B2_gm, (X, Y, nz))
    rc = _native.lib().vx_iso_diag_streamfunction_f64(stream_handle(), g, ptr(K), ptr(Aez), ptr(Anz), ptr(B1),
                                                      ptr(B2))
    _native.check(rc, "isoneutral_diag_streamfunction")
    return KernelOutput(B1_gm=B1, B2_gm=B2)


def isoneutral_diag_streamfunction(state):
    """The routine (veros/core/isoneutral/isoneutral.py:257-269): only with
    neutral and skew diffusion enabled."""
    st = state.settings
    if not (getattr(st, "enable_neutral_diffusion", True) and getattr(st, "enable_skew_diffusion", True)):
        return
    _store(state, isoneutral_diag_streamfunction_kernel(state))
