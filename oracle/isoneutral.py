"""Oracle: isoneutral triad-tensor stencils (TEST INFRASTRUCTURE ONLY).

Restates, in NumPy float64, the reference's
* ``dm_taper`` and ``isoneutral_diffusion_pre``   veros/core/isoneutral/isoneutral.py:10-229
* ``_calc_tracer_fluxes``                         veros/core/isoneutral/diffusion.py:9-113
* ``_calc_explicit_part``                         veros/core/isoneutral/diffusion.py:116-139
* ``_calc_implicit_part``                         veros/core/isoneutral/diffusion.py:142-169
* ``isoneutral_diffusion_tracer`` / ``_kernel``   veros/core/isoneutral/diffusion.py:172-283
* ``compute_dissipation``, ``dissipation_on_wgrid`` veros/core/diffusion.py:9-62
* ``pad_z_edges``                                 veros/core/utilities.py:24-35

The equation of state is the linear one (eq_of_state_type 1), whose
derivatives are constants (veros/core/density/get_rho.py:93-132,
veros/core/density/linear_eq.py:36-43); they are read from
``settings.drhodT`` / ``settings.drhodS``.

Pins: golden vectors from the reference's NumPy path
(tests/golden/make_golden.py), plus closed forms and invariants in
tests/test_oracle_isoneutral.py (uniform tracers give zero slopes and the
K_11 closed form; a linear stratification on a Cartesian grid gives the
constant-slope closed form; the explicit step conserves tracer content; the
implicit step conserves column content; the step is linear in the tracer).

Each function works on a ``state`` with ``.variables`` (arrays shaped as veros
variables, 2-cell halos, z contiguous, k = 0 deepest) and ``.settings``. Inputs
are never modified; outputs are new arrays carrying the inputs' values outside
the written ranges (veros's update semantics).
"""
from __future__ import annotations

from collections import namedtuple

import numpy as np

from .tdma import create_water_masks, solve_implicit

EPSLN = 1e-20  # veros/core/isoneutral/isoneutral.py:28


def KernelOutput(**kw):
    """Namedtuple built from keyword arguments (veros/state.py:16-20)."""
    return namedtuple("KernelOutput", list(kw.keys()))(*kw.values())


def dm_taper(sx, slopec, dslope):
    """0.5*(1 + tanh((slopec - |s|)/dslope)) (veros/core/isoneutral/isoneutral.py:10-15)."""
    return 0.5 * (1.0 + np.tanh((-np.abs(sx) + slopec) / dslope))


def pad_z_edges(arr):
    """Repeat the first and last level once along z (veros/core/utilities.py:24-35)."""
    if arr.ndim == 1:
        return np.pad(arr, 1, mode="edge")
    return np.pad(arr, ((0, 0), (0, 0), (1, 1)), mode="edge")


def _f(mask):
    return np.asarray(mask, dtype=np.float64)


def _gradients(vs, st, fld):
    """Masked face gradients of one field (veros/core/isoneutral/isoneutral.py:46-95):
    top faces (W points), east faces (U points, metric dxu*cost) and north faces
    (V points, metric dyu). The last index along each axis stays 0."""
    gz = np.zeros(fld.shape)
    gx = np.zeros(fld.shape)
    gy = np.zeros(fld.shape)
    gz[:, :, :-1] = _f(vs.maskW)[:, :, :-1] * (fld[:, :, 1:] - fld[:, :, :-1]) / vs.dzw[None, None, :-1]
    gx[:-1] = _f(vs.maskU)[:-1] * (fld[1:] - fld[:-1]) / (vs.dxu[:-1, None, None] * vs.cost[None, :, None])
    gy[:, :-1] = _f(vs.maskV)[:, :-1] * (fld[:, 1:] - fld[:, :-1]) / vs.dyu[None, :-1, None]
    return gx, gy, gz


def isoneutral_diffusion_pre(state):
    """Triad slopes, tapers and the isoneutral tensor components
    (veros/core/isoneutral/isoneutral.py:18-229)."""
    vs, st = state.variables, state.settings
    X, Y, nz = vs.K_iso.shape
    T = vs.temp[..., vs.tau]
    S = vs.salt[..., vs.tau]
    K = vs.K_iso
    mT, mU, mV, mW = (_f(m) for m in (vs.maskT, vs.maskU, vs.maskV, vs.maskW))

    # drho/dT and drho/dS at T points (:40-41), constants for the linear EOS
    drdT = mT * st.drhodT
    drdS = mT * st.drhodS
    dTdx, dTdy, dTdz = _gradients(vs, st, T)
    dSdx, dSdy, dSdz = _gradients(vs, st, S)

    def taper(s):
        return dm_taper(s, st.iso_slopec, st.iso_dslope)

    Ai_ez = np.array(vs.Ai_ez, copy=True)
    Ai_nz = np.array(vs.Ai_nz, copy=True)
    Ai_bx = np.array(vs.Ai_bx, copy=True)
    Ai_by = np.array(vs.Ai_by, copy=True)
    K_11 = np.array(vs.K_11, copy=True)
    K_22 = np.array(vs.K_22, copy=True)
    K_33 = np.array(vs.K_33, copy=True)

    # --- east faces: i in [1, X-2), j in [2, Y-2), all k (:97-132)
    I0, J0 = slice(1, X - 2), slice(2, Y - 2)
    diff_e = np.zeros((X - 3, Y - 4, nz))
    diff_e[:, :, 1:] = 0.25 * (K[1:-2, 2:-2, 1:] + K[1:-2, 2:-2, :-1] + K[2:-1, 2:-2, 1:] + K[2:-1, 2:-2, :-1])
    diff_e[:, :, 0] = 0.5 * (K[1:-2, 2:-2, 0] + K[2:-1, 2:-2, 0])
    acc = np.zeros((X - 3, Y - 4, nz))
    for kr in (0, 1):
        k_lo = 1 - kr                      # levels k >= k_lo
        kz = slice(0, nz - 1 + kr)         # vertical gradient at k-1+kr
        for ip in (0, 1):
            Ii = slice(1 + ip, X - 2 + ip)  # i + ip
            num = drdT[Ii, J0, k_lo:] * dTdx[I0, J0, k_lo:] + drdS[Ii, J0, k_lo:] * dSdx[I0, J0, k_lo:]
            den = drdT[Ii, J0, k_lo:] * dTdz[Ii, J0, kz] + drdS[Ii, J0, k_lo:] * dSdz[Ii, J0, kz]
            s = -num / (np.minimum(0.0, den) - EPSLN)
            t = taper(s)
            acc[:, :, k_lo:] += vs.dzw[None, None, kz] * mU[I0, J0, k_lo:] * np.maximum(st.K_iso_steep, diff_e[:, :, k_lo:] * t)
            Ai_ez[I0, J0, k_lo:, ip, kr] = t * s * mU[I0, J0, k_lo:]
    K_11[I0, J0, :] = acc / (4.0 * vs.dzt[None, None, :])

    # --- north faces: i in [2, X-2), j in [1, Y-2) (:134-168)
    I0, J0 = slice(2, X - 2), slice(1, Y - 2)
    diff_n = np.zeros((X - 4, Y - 3, nz))
    diff_n[:, :, 1:] = 0.25 * (K[2:-2, 1:-2, 1:] + K[2:-2, 1:-2, :-1] + K[2:-2, 2:-1, 1:] + K[2:-2, 2:-1, :-1])
    diff_n[:, :, 0] = 0.5 * (K[2:-2, 1:-2, 0] + K[2:-2, 2:-1, 0])
    acc = np.zeros((X - 4, Y - 3, nz))
    for kr in (0, 1):
        k_lo = 1 - kr
        kz = slice(0, nz - 1 + kr)
        for jp in (0, 1):
            Jj = slice(1 + jp, Y - 2 + jp)
            num = drdT[I0, Jj, k_lo:] * dTdy[I0, J0, k_lo:] + drdS[I0, Jj, k_lo:] * dSdy[I0, J0, k_lo:]
            den = drdT[I0, Jj, k_lo:] * dTdz[I0, Jj, kz] + drdS[I0, Jj, k_lo:] * dSdz[I0, Jj, kz]
            s = -num / (np.minimum(0.0, den) - EPSLN)
            t = taper(s)
            acc[:, :, k_lo:] += vs.dzw[None, None, kz] * mV[I0, J0, k_lo:] * np.maximum(st.K_iso_steep, diff_n[:, :, k_lo:] * t)
            Ai_nz[I0, J0, k_lo:, jp, kr] = t * s * mV[I0, J0, k_lo:]
    K_22[I0, J0, :] = acc / (4.0 * vs.dzt[None, None, :])

    # --- top faces: i in [2, X-2), j in [2, Y-2), k in [0, nz-1) (:170-225)
    I0, J0 = slice(2, X - 2), slice(2, Y - 2)
    accx = np.zeros((X - 4, Y - 4, nz - 1))
    accy = np.zeros((X - 4, Y - 4, nz - 1))
    Kw = K[I0, J0, :-1]
    mWi = mW[I0, J0, :-1]
    for kr in (0, 1):
        kk = slice(kr, nz - 1 + kr)        # k + kr
        den = drdT[I0, J0, kk] * dTdz[I0, J0, :-1] + drdS[I0, J0, kk] * dSdz[I0, J0, :-1]
        for ip in (0, 1):
            Ii = slice(1 + ip, X - 3 + ip)  # i - 1 + ip
            num = drdT[I0, J0, kk] * dTdx[Ii, J0, kk] + drdS[I0, J0, kk] * dSdx[Ii, J0, kk]
            s = -num / (np.minimum(0.0, den) - EPSLN)
            t = taper(s)
            accx = accx + vs.dxu[Ii, None, None] * Kw * t * s ** 2 * mWi
            Ai_bx[I0, J0, :-1, ip, kr] = t * s * mWi
        for jp in (0, 1):
            Jj = slice(1 + jp, Y - 3 + jp)  # j - 1 + jp
            fy = vs.cosu[Jj] * vs.dyu[Jj]
            num = drdT[I0, J0, kk] * dTdy[I0, Jj, kk] + drdS[I0, J0, kk] * dSdy[I0, Jj, kk]
            s = -num / (np.minimum(0.0, den) - EPSLN)
            t = taper(s)
            accy = accy + fy[None, :, None] * Kw * t * s ** 2 * mWi
            Ai_by[I0, J0, :-1, jp, kr] = t * s * mWi
    K_33[I0, J0, :-1] = (accx / (4 * vs.dxt[I0, None, None])
                         + accy / (4 * vs.dyt[None, J0, None] * vs.cost[None, J0, None]))
    K_33[..., -1] = 0.0

    return KernelOutput(Ai_ez=Ai_ez, Ai_nz=Ai_nz, Ai_bx=Ai_bx, Ai_by=Ai_by, K_11=K_11, K_22=K_22, K_33=K_33)


def calc_tracer_fluxes(state, tr, K_iso, K_skew):
    """East, north and top isoneutral fluxes of one tracer
    (veros/core/isoneutral/diffusion.py:9-113). ``K_iso``/``K_skew`` may be
    scalars (0 for the skew or the isoneutral variant)."""
    vs = state.variables
    X, Y, nz = vs.K_iso.shape
    trc = tr[..., vs.tau]
    trp = pad_z_edges(trc)
    K1 = np.broadcast_to(np.asarray(K_iso - K_skew, dtype=np.float64), (X, Y, nz))
    K2 = np.broadcast_to(np.asarray(K_iso + K_skew, dtype=np.float64), (X, Y, nz))
    fe = np.zeros((X, Y, nz))
    fn = np.zeros((X, Y, nz))
    ft = np.zeros((X, Y, nz))

    # east (:22-47): clamped vertical differences tr[min(k+kr,nz-1)] - tr[max(k+kr-1,0)]
    dl = np.zeros((X - 3, Y - 4, nz))
    dl[:, :, 1:] = 0.25 * (K1[1:-2, 2:-2, 1:] + K1[1:-2, 2:-2, :-1] + K1[2:-1, 2:-2, 1:] + K1[2:-1, 2:-2, :-1])
    dl[:, :, 0] = 0.5 * (K1[1:-2, 2:-2, 0] + K1[2:-1, 2:-2, 0])
    acc = 0.0
    for kr in (0, 1):
        for ip in (0, 1):
            Ii = slice(1 + ip, X - 2 + ip)
            upper = trp[Ii, 2:-2, 1 + kr:nz + 1 + kr]
            lower = trp[Ii, 2:-2, kr:nz + kr]
            acc = acc + dl * vs.Ai_ez[1:-2, 2:-2, :, ip, kr] * (upper - lower)
    fe[1:-2, 2:-2, :] = (acc / (4.0 * vs.dzt[None, None, :])
                         + (trc[2:-1, 2:-2] - trc[1:-2, 2:-2])
                         / (vs.cost[None, 2:-2, None] * vs.dxu[1:-2, None, None]) * vs.K_11[1:-2, 2:-2, :])

    # north (:49-77)
    dl = np.zeros((X - 4, Y - 3, nz))
    dl[:, :, 1:] = 0.25 * (K1[2:-2, 1:-2, 1:] + K1[2:-2, 1:-2, :-1] + K1[2:-2, 2:-1, 1:] + K1[2:-2, 2:-1, :-1])
    dl[:, :, 0] = 0.5 * (K1[2:-2, 1:-2, 0] + K1[2:-2, 2:-1, 0])
    acc = 0.0
    for kr in (0, 1):
        for jp in (0, 1):
            Jj = slice(1 + jp, Y - 2 + jp)
            upper = trp[2:-2, Jj, 1 + kr:nz + 1 + kr]
            lower = trp[2:-2, Jj, kr:nz + kr]
            acc = acc + dl * vs.Ai_nz[2:-2, 1:-2, :, jp, kr] * (upper - lower)
    fn[2:-2, 1:-2, :] = vs.cosu[None, 1:-2, None] * (
        acc / (4.0 * vs.dzt[None, None, :])
        + (trc[2:-2, 2:-1] - trc[2:-2, 1:-2]) / vs.dyu[None, 1:-2, None] * vs.K_22[2:-2, 1:-2, :])

    # top (:79-111): K31/K32 parts only
    dl = K2[2:-2, 2:-2, :-1]
    accx = 0.0
    for ip in (0, 1):
        for kr in (0, 1):
            kk = slice(kr, nz - 1 + kr)
            accx = accx + dl * vs.Ai_bx[2:-2, 2:-2, :-1, ip, kr] / vs.cost[None, 2:-2, None] * (
                trc[2 + ip:X - 2 + ip, 2:-2, kk] - trc[1 + ip:X - 3 + ip, 2:-2, kk])
    accy = 0.0
    for jp in (0, 1):
        for kr in (0, 1):
            kk = slice(kr, nz - 1 + kr)
            accy = accy + dl * vs.Ai_by[2:-2, 2:-2, :-1, jp, kr] * vs.cosu[None, 1 + jp:Y - 3 + jp, None] * (
                trc[2:-2, 2 + jp:Y - 2 + jp, kk] - trc[2:-2, 1 + jp:Y - 3 + jp, kk])
    ft[2:-2, 2:-2, :-1] = (accx / (4 * vs.dxt[2:-2, None, None])
                           + accy / (4 * vs.dyt[None, 2:-2, None] * vs.cost[None, 2:-2, None]))
    ft[:, :, -1] = 0.0
    return fe, fn, ft


def calc_explicit_part(state, fe, fn, ft):
    """Flux divergence (veros/core/isoneutral/diffusion.py:116-139)."""
    vs = state.variables
    mT = _f(vs.maskT)
    out = np.zeros(fe.shape)
    out[2:-2, 2:-2, :] = mT[2:-2, 2:-2, :] * (
        (fe[2:-2, 2:-2, :] - fe[1:-3, 2:-2, :]) / (vs.cost[None, 2:-2, None] * vs.dxt[2:-2, None, None])
        + (fn[2:-2, 2:-2, :] - fn[2:-2, 1:-3, :]) / (vs.cost[None, 2:-2, None] * vs.dyt[None, 2:-2, None]))
    out[:, :, 0] += mT[:, :, 0] * ft[:, :, 0] / vs.dzt[0]
    out[:, :, 1:] += mT[:, :, 1:] * (ft[:, :, 1:] - ft[:, :, :-1]) / vs.dzt[None, None, 1:]
    return out


def implicit_coefficients(state):
    """K33 vertical diffusion rows (veros/core/isoneutral/diffusion.py:147-164):
    (a, b, c, b_edge) on the interior (nx, ny, nz)."""
    vs, st = state.variables, state.settings
    X, Y, nz = vs.K_iso.shape
    shp = (X - 4, Y - 4, nz)
    delta = np.zeros(shp)
    delta[:, :, :-1] = st.dt_tracer / vs.dzw[None, None, :-1] * vs.K_33[2:-2, 2:-2, :-1]
    a = np.zeros(shp)
    b = np.zeros(shp)
    c = np.zeros(shp)
    a[:, :, 1:] = -delta[:, :, :-1] / vs.dzt[None, None, 1:]
    b[:, :, 1:-1] = 1 + (delta[:, :, 1:-1] + delta[:, :, :-2]) / vs.dzt[None, None, 1:-1]
    b[:, :, -1] = 1 + delta[:, :, -2] / vs.dzt[None, None, -1]
    b_edge = 1 + delta / vs.dzt[None, None, :]
    c[:, :, :-1] = -delta[:, :, :-1] / vs.dzt[None, None, :-1]
    return a, b, c, b_edge


def calc_implicit_part(state, tr, solver="dgtsv"):
    """Implicit K33 step on the interior (veros/core/isoneutral/diffusion.py:142-169)."""
    vs, st = state.variables, state.settings
    nz = vs.K_iso.shape[2]
    _, water, edge = create_water_masks(vs.kbot[2:-2, 2:-2], nz)
    a, b, c, b_edge = implicit_coefficients(state)
    rhs = tr[2:-2, 2:-2, :, vs.taup1]
    sol = solve_implicit(a, b, c, rhs, water, edge, b_edge=b_edge, solver=solver)
    return np.where(water, sol, rhs)


def isoneutral_diffusion_tracer(state, tr, dtracer_iso, iso=True, skew=False, solver="dgtsv"):
    """Explicit flux divergence, then (iso only) the implicit K33 solve
    (veros/core/isoneutral/diffusion.py:172-207)."""
    vs, st = state.variables, state.settings
    K_iso = vs.K_iso if iso else 0.0
    K_skew = vs.K_gm if skew else 0.0
    fe, fn, ft = calc_tracer_fluxes(state, tr, K_iso, K_skew)
    dtr = calc_explicit_part(state, fe, fn, ft)
    dtracer_iso = dtracer_iso + dtr
    tr = np.array(tr, copy=True)
    tr[2:-2, 2:-2, :, vs.taup1] += st.dt_tracer * dtr[2:-2, 2:-2, :]
    if iso:
        new_tr = np.array(tr, copy=True)
        new_tr[2:-2, 2:-2, :, vs.taup1] = calc_implicit_part(state, tr, solver=solver)
        dtracer_iso = dtracer_iso + (new_tr[..., vs.taup1] - tr[..., vs.taup1]) / st.dt_tracer
        tr = new_tr
    return tr, dtracer_iso, fe, fn, ft


def compute_dissipation(state, int_drhodX, fe, fn):
    """Dissipation of dynamic enthalpy by lateral fluxes on T points
    (veros/core/diffusion.py:9-36); written on [1:-1, 1:-1]."""
    vs, st = state.variables, state.settings
    I = int_drhodX
    g = 0.5 * st.grav / st.rho_0
    diss = np.zeros(I.shape)
    diss[1:-1, 1:-1, :] = (
        g * ((I[2:, 1:-1] - I[1:-1, 1:-1]) * fe[1:-1, 1:-1] + (I[1:-1, 1:-1] - I[:-2, 1:-1]) * fe[:-2, 1:-1])
        / (vs.dxt[1:-1, None, None] * vs.cost[None, 1:-1, None])
        + g * ((I[1:-1, 2:] - I[1:-1, 1:-1]) * fn[1:-1, 1:-1] + (I[1:-1, 1:-1] - I[1:-1, :-2]) * fn[1:-1, :-2])
        / (vs.dyt[None, 1:-1, None] * vs.cost[None, 1:-1, None]))
    return diss


def dissipation_on_wgrid(state, diss, ks):
    """Average T-point dissipation onto W points with the bottom-cell
    weighting (veros/core/diffusion.py:39-62)."""
    vs = state.variables
    nz = diss.shape[2]
    land, water, edge = create_water_masks(ks, nz)
    water = water & ~edge
    dzw_pad = pad_z_edges(vs.dzw)
    out = np.zeros(diss.shape)
    avg = 0.5 * (diss[:, :, :-1] + diss[:, :, 1:])
    out[:, :, :-1] = ((avg + 0.5 * (diss[:, :, :-1] * dzw_pad[None, None, :-3] / vs.dzw[None, None, :-1]))
                      * edge[:, :, :-1] + avg * water[:, :, :-1])
    out[:, :, -1] = diss[:, :, -1] * land
    return out


def isoneutral_diffusion_kernel(state, tr, istemp, iso=True, solver="dgtsv"):
    """One tracer's isoneutral (``iso=True``) or skew (``iso=False``) step,
    with the energy-conserving dissipation diagnostics
    (veros/core/isoneutral/diffusion.py:210-283)."""
    vs, st = state.variables, state.settings
    dtracer_iso = vs.dtemp_iso if istemp else vs.dsalt_iso
    tr, dtracer_iso, fe, fn, ft = isoneutral_diffusion_tracer(state, tr, dtracer_iso, iso=iso, skew=not iso,
                                                              solver=solver)
    out = {}
    if istemp:
        out.update(temp=tr, dtemp_iso=dtracer_iso)
    else:
        out.update(salt=tr, dsalt_iso=dtracer_iso)

    if st.enable_conserve_energy:
        I = (vs.int_drhodT if istemp else vs.int_drhodS)[..., vs.tau]
        diss = compute_dissipation(state, I, fe, fn)
        diss_w = dissipation_on_wgrid(state, diss, vs.kbot)
        P = (vs.P_diss_iso if iso else vs.P_diss_skew) + diss_w
        fxa = (-I[2:-2, 2:-2, 1:] + I[2:-2, 2:-2, :-1]) / vs.dzw[None, None, :-1]
        mWi = _f(vs.maskW)[2:-2, 2:-2, :-1]
        if iso:
            P[2:-2, 2:-2, :-1] += -st.grav / st.rho_0 * fxa * (
                ft[2:-2, 2:-2, :-1] * mWi
                + vs.K_33[2:-2, 2:-2, :-1]
                * (tr[2:-2, 2:-2, 1:, vs.taup1] - tr[2:-2, 2:-2, :-1, vs.taup1])
                / vs.dzw[None, None, :-1] * mWi)
            out["P_diss_iso"] = P
        else:
            P[2:-2, 2:-2, :-1] += -st.grav / st.rho_0 * fxa * ft[2:-2, 2:-2, :-1] * mWi
            out["P_diss_skew"] = P
    return KernelOutput(**out)


def isoneutral_diffusion(state, tr_name, istemp):
    """The routine: runs the kernel and writes its outputs back into the state
    (veros/core/isoneutral/diffusion.py:286-295). ``tr_name`` is "temp" or "salt"."""
    out = isoneutral_diffusion_kernel(state, getattr(state.variables, tr_name), istemp, iso=True)
    for k, v in out._asdict().items():
        setattr(state.variables, k, v)
    return out


def isoneutral_diag_streamfunction_kernel(state):
    """B1_gm, B2_gm (veros/core/isoneutral/isoneutral.py:232-255) as new
    arrays: the 4-point average of the z edge-padded K_gm around each face
    times the sum of the face's four slopes, on [1:-2, 2:-2] (B2) and
    [2:-2, 1:-2] (B1)."""
    vs = state.variables
    K = np.asarray(vs.K_gm, dtype=np.float64)
    Kp = np.concatenate([K[..., :1], K, K[..., -1:]], axis=-1)  # pad_z_edges (utilities.py:24-35)
    B1 = np.array(vs.B1_gm, dtype=np.float64, copy=True)
    B2 = np.array(vs.B2_gm, dtype=np.float64, copy=True)
    d = 0.25 * (Kp[1:-2, 2:-2, 1:-1] + Kp[1:-2, 2:-2, :-2] + Kp[2:-1, 2:-2, 1:-1] + Kp[2:-1, 2:-2, :-2])
    B2[1:-2, 2:-2, :] = 0.25 * d * np.sum(np.asarray(vs.Ai_ez)[1:-2, 2:-2, ...], axis=(3, 4))
    d = 0.25 * (Kp[2:-2, 1:-2, 1:-1] + Kp[2:-2, 1:-2, :-2] + Kp[2:-2, 2:-1, 1:-1] + Kp[2:-2, 2:-1, :-2])
    B1[2:-2, 1:-2, :] = -0.25 * d * np.sum(np.asarray(vs.Ai_nz)[2:-2, 1:-2, ...], axis=(3, 4))
    return KernelOutput(B1_gm=B1, B2_gm=B2)
