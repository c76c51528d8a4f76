"""Drop-in GPU versions of the TKE, EKE and IDEMIX integrations, the
remaining ``solve_implicit`` callers of the mixing path:

* ``integrate_tke_kernel(state)``     veros/core/tke.py:123-321
* ``integrate_eke_kernel(state)``     veros/core/eke.py:93-255
* ``integrate_idemix_kernel(state)``  veros/core/idemix.py:47-265

and the routines ``integrate_tke`` / ``integrate_eke`` / ``integrate_idemix``
that store the outputs (veros/core/tke.py:117-120 and the like). Each
kernel is one C-ABI call (vx_integrate_{tke,eke,idemix}_f64): the forcing is
assembled in the kernel from the dissipation terms the settings branches
select (passed as a list of signed terms, in the reference's order), the
implicit vertical system is formed in registers and solved, the dissipation
diagnostic and the lateral diffusion are written in the same pass. The
superbee / upwind advection branches (another veros component) raise
NotImplementedError. Arrays are updated in place; the returned KernelOutput
holds the same tensors.
"""
from __future__ import annotations

import ctypes

import torch

from . import _native
from ._arrays import ptr, stream_handle
from .isoneutral import KernelOutput, _check_array, _check_levels, grid_struct
from .layout import common_layout


def _terms(vs, names_signs, X, Y, nz):
    t = _native.VxTerms()
    if len(names_signs) > 12:
        raise ValueError("at most 12 forcing terms")
    keep = []
    for q, (name, sgn) in enumerate(names_signs):
        a = _check_array(name, getattr(vs, name), (X, Y, nz))
        t.arr[q] = a.data_ptr()
        t.sign[q] = float(sgn)
        keep.append(a)
    t.n = len(names_signs)
    return t, keep


def tke_forcing_terms(st):
    """The terms of the TKE forcing (veros/core/tke.py:139-177), in order."""
    terms = [("K_diss_v", 1), ("P_diss_v", -1), ("P_diss_adv", -1)]
    if not st.enable_store_cabbeling_heat:
        terms.append(("P_diss_nonlin", -1))
    if st.enable_eke:
        terms.append(("eke_diss_tke", 1))
    if st.enable_idemix:
        terms.append(("iw_diss", 1))
        if st.enable_store_bottom_friction_tke:
            terms.append(("K_diss_bot", 1))
    else:
        if st.enable_eke:
            terms.append(("eke_diss_iw", 1))
        else:
            terms += [("K_diss_gm", 1), ("K_diss_h", 1), ("P_diss_skew", -1)]
            if st.enable_store_cabbeling_heat:
                terms += [("P_diss_hmix", -1), ("P_diss_iso", -1)]
        terms.append(("K_diss_bot", 1))
    return terms


def eke_forcing_terms(st):
    """veros/core/eke.py:109-118."""
    terms = [("K_diss_gm", 1), ("K_diss_h", 1), ("P_diss_skew", -1)]
    if not st.enable_store_cabbeling_heat:
        terms += [("P_diss_hmix", -1), ("P_diss_iso", -1)]
    return terms


def idemix_forcing_terms(st):
    """(terms before the optional move to the bottom/surface cells, terms
    added after it) (veros/core/idemix.py:59-115)."""
    if st.enable_eke:
        base = [("eke_diss_iw", 1)]
    else:
        base = [("K_diss_gm", 1), ("K_diss_h", 1), ("P_diss_skew", -1)]
        if st.enable_store_cabbeling_heat:
            base += [("P_diss_hmix", -1), ("P_diss_iso", -1)]
    post = [] if st.enable_store_bottom_friction_tke else [("K_diss_bot", 1)]
    return base, post


def _grid(vs, keep, layout):
    return grid_struct(vs, keep, layout)


def integrate_tke_kernel(state):
    vs, st = state.variables, state.settings
    if st.enable_tke_superbee_advection or st.enable_tke_upwind_advection:
        raise NotImplementedError("TKE advection (superbee / upwind) is not part of this path")
    X, Y, nz = vs.kappaM.shape
    tke = _check_levels("tke", vs.tke, (X, Y, nz, 3))
    keep = []
    g = _grid(vs, keep, common_layout(tke=tke))
    f3 = [_check_array(n, getattr(vs, n), (X, Y, nz)) for n in ("kappaM", "sqrttke", "mxl")]
    fs = _check_array("forc_tke_surface", vs.forc_tke_surface, (X, Y))
    sc = _check_array("tke_surf_corr", vs.tke_surf_corr, (X, Y))
    diss = _check_array("tke_diss", vs.tke_diss, (X, Y, nz))
    terms, tk = _terms(vs, tke_forcing_terms(st), X, Y, nz)
    p = _native.VxTkeParams(float(st.dt_mom), float(st.alpha_tke), float(st.c_eps), float(st.K_h_tke),
                            int(bool(st.enable_tke_hor_diffusion)), 0)
    rc = _native.lib().vx_integrate_tke_f64(stream_handle(), g, ctypes.byref(p), ptr(tke), int(vs.tau),
                                            int(vs.taup1), *(ptr(a) for a in f3), ptr(fs), ctypes.byref(terms),
                                            ptr(sc), ptr(diss))
    _native.check(rc, "integrate_tke_kernel")
    return KernelOutput(tke=tke, tke_surf_corr=sc, tke_diss=diss)


def integrate_eke_kernel(state):
    vs, st = state.variables, state.settings
    if st.enable_eke_superbee_advection or st.enable_eke_upwind_advection:
        raise NotImplementedError("EKE advection (superbee / upwind) is not part of this path")
    X, Y, nz = vs.kappaM.shape
    eke = _check_levels("eke", vs.eke, (X, Y, nz, 3))
    keep = []
    g = _grid(vs, keep, common_layout(eke=eke))
    f3 = [_check_array(n, getattr(vs, n), (X, Y, nz)) for n in ("kappaM", "sqrteke", "eke_len", "K_gm")]
    diw = _check_array("eke_diss_iw", vs.eke_diss_iw, (X, Y, nz))
    dtk = _check_array("eke_diss_tke", vs.eke_diss_tke, (X, Y, nz))
    terms, tk = _terms(vs, eke_forcing_terms(st), X, Y, nz)
    p = _native.VxEkeParams(float(st.dt_tracer), float(st.alpha_eke), float(st.eke_c_eps))
    rc = _native.lib().vx_integrate_eke_f64(stream_handle(), g, ctypes.byref(p), ptr(eke), int(vs.tau),
                                            int(vs.taup1), *(ptr(a) for a in f3), ctypes.byref(terms), ptr(diw),
                                            ptr(dtk))
    _native.check(rc, "integrate_eke_kernel")
    return KernelOutput(eke=eke, eke_diss_iw=diw, eke_diss_tke=dtk)


def integrate_idemix_kernel(state):
    vs, st = state.variables, state.settings
    if st.enable_idemix_superbee_advection or st.enable_idemix_upwind_advection:
        raise NotImplementedError("IDEMIX advection (superbee / upwind) is not part of this path")
    X, Y, nz = vs.c0.shape
    E = _check_levels("E_iw", vs.E_iw, (X, Y, nz, 3))
    keep = []
    g = _grid(vs, keep, common_layout(E_iw=E))
    f3 = [_check_array(n, getattr(vs, n), (X, Y, nz)) for n in ("c0", "v0", "alpha_c")]
    fb = _check_array("forc_iw_bottom", vs.forc_iw_bottom, (X, Y))
    fs = _check_array("forc_iw_surface", vs.forc_iw_surface, (X, Y))
    iwd = _check_array("iw_diss", vs.iw_diss, (X, Y, nz))
    base, post = idemix_forcing_terms(st)
    terms, tk = _terms(vs, base, X, Y, nz)
    terms_post, tk2 = _terms(vs, post, X, Y, nz)
    mode = 0
    if st.enable_eke and (st.enable_eke_diss_bottom or st.enable_eke_diss_surfbot):
        mode = 1 if st.enable_eke_diss_bottom else 2
    p = _native.VxIdemixParams(float(st.dt_tracer), float(st.tau_v), float(st.tau_h),
                               float(st.eke_diss_surfbot_frac), int(bool(st.enable_idemix_hor_diffusion)), mode)
    rc = _native.lib().vx_integrate_idemix_f64(stream_handle(), g, ctypes.byref(p), ptr(E), int(vs.tau),
                                               int(vs.taup1), *(ptr(a) for a in f3), ptr(fb), ptr(fs),
                                               ctypes.byref(terms), ctypes.byref(terms_post), ptr(iwd))
    _native.check(rc, "integrate_idemix_kernel")
    return KernelOutput(E_iw=E, dE_iw=vs.dE_iw, iw_diss=iwd)


def _store(state, out):
    for k, v in out._asdict().items():
        setattr(state.variables, k, v)


def integrate_tke(state):
    """veros/core/tke.py:117-120."""
    _store(state, integrate_tke_kernel(state))


def integrate_eke(state):
    _store(state, integrate_eke_kernel(state))


def integrate_idemix(state):
    _store(state, integrate_idemix_kernel(state))
