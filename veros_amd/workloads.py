"""Seeded synthetic inputs for the mixing hot path (host side, NumPy float64).

Everything here builds *inputs* only: grids, topography, masks, tracer fields
and tridiagonal coefficients, laid out exactly like veros state variables:

* 3-D fields ``(X, Y, nz)`` with ``X = nx + 4`` and ``Y = ny + 4`` (2-cell halos),
  z contiguous, ``k = 0`` the deepest level (veros/variables.py:146-149);
* ``temp``/``salt``/``int_drhodT``/``int_drhodS`` are ``(X, Y, nz, 3)`` with the
  time-level axis innermost (veros/variables.py:75,338,346);
* ``Ai_ez/Ai_nz/Ai_bx/Ai_by`` are ``(X, Y, nz, 2, 2)`` (veros/variables.py:672-699).

The grid follows ``calc_grid_spacings_kernel`` (veros/core/numerics.py:26-97,
including the ``u_centered_grid`` staggering of :8-22) and the masks follow
``calc_topo_kernel`` (veros/core/numerics.py:180-221).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, fields
from types import SimpleNamespace

import numpy as np

# veros/settings.py:32-37
PI = math.pi
RADIUS = 6370e3
DEGTOM = RADIUS / 180.0 * PI
GRAV = 9.81
RHO_0 = 1024.0


@dataclass
class MixSettings:
    """The settings the mixing path reads (veros/settings.py:26-91 defaults,
    overridden per workload)."""

    nx: int
    ny: int
    nz: int
    dt_tracer: float = 3600.0
    dt_mom: float = 3600.0
    K_iso_steep: float = 500.0
    iso_slopec: float = 1e-2
    iso_dslope: float = 5e-3
    enable_conserve_energy: bool = True
    enable_cyclic_x: bool = True
    coord_degree: bool = True
    grav: float = GRAV
    rho_0: float = RHO_0
    eq_of_state_type: int = 1
    # derivatives of the linear equation of state (veros/core/density/linear_eq.py:36-43)
    drhodT: float = -1.67e-4 * 1024.0
    drhodS: float = 0.78e-3 * 1024.0

    def copy(self, **kw):
        d = {f.name: getattr(self, f.name) for f in fields(self)}
        d.update(kw)
        return MixSettings(**d)


def u_centered_grid(dt_, du, t, u):
    """Staggered grid from cell widths (veros/core/numerics.py:8-22)."""
    dt_ = np.asarray(dt_, dtype=np.float64)
    u = np.zeros_like(dt_)
    u[1:] = np.cumsum(dt_[1:])
    t = np.zeros_like(dt_)
    t[0] = u[0] - dt_[0] * 0.5
    t[1:] = 2 * u[:-1]
    alt = np.ones_like(t)
    alt[::2] = -1
    t = alt * np.cumsum(alt * t)
    du = np.zeros_like(dt_)
    du[:-1] = t[1:] - t[:-1]
    du[-1] = 2 * dt_[-1] - du[-2]
    return du, t, u


def geometric_dz(nz: int, dz_top: float, dz_bottom: float) -> np.ndarray:
    """Layer thicknesses growing geometrically with depth; index 0 is the deepest."""
    if nz == 1:
        return np.array([dz_top])
    r = (dz_bottom / dz_top) ** (1.0 / (nz - 1))
    from_top = dz_top * r ** np.arange(nz)
    return from_top[::-1].copy()


def make_grid(nx, ny, nz, *, dx_deg, dy_deg, y_origin, dz_top, dz_bottom,
              cyclic=True, coord_degree=True, rng=None, jitter=0.0):
    """1-D metric arrays as calc_grid_spacings + calc_grid_metrics produce them
    (veros/core/numerics.py:26-97, :116-147). ``jitter`` perturbs the widths
    multiplicatively (as the random-state recipe does, veros/pyom_compat.py:463-472)
    so that index mistakes in metric terms do not cancel in tests."""
    X, Y = nx + 4, ny + 4
    dxt = np.full(X, float(dx_deg))
    dyt = np.full(Y, float(dy_deg))
    dzt = geometric_dz(nz, dz_top, dz_bottom)
    if jitter and rng is not None:
        dxt = dxt * (1 + jitter * rng.standard_normal(X))
        dyt = dyt * (1 + jitter * rng.standard_normal(Y))
        dzt = dzt * (1 + jitter * rng.standard_normal(nz))
    if cyclic:
        dxt[-2:] = dxt[2:4]
        dxt[:2] = dxt[-4:-2]
    else:
        dxt[-2:] = dxt[-3]
        dxt[:2] = dxt[2]
    dyt[-2:] = dyt[-3]
    dyt[:2] = dyt[2]
    dxu, xt, xu = u_centered_grid(dxt, None, None, None)
    xt = xt - xu[2]
    if cyclic:
        dxu[-2:] = dxu[2:4]
        dxu[:2] = dxu[-4:-2]
    dyu, yt, yu = u_centered_grid(dyt, None, None, None)
    yt = yt + y_origin - yu[2]
    yu = yu + y_origin - yu[2]
    if coord_degree:
        dxt, dxu, dyt, dyu = (v * DEGTOM for v in (dxt, dxu, dyt, dyu))
        cost = np.cos(yt * PI / 180.0)
        cosu = np.cos(yu * PI / 180.0)
    else:
        cost = np.ones(Y)
        cosu = np.ones(Y)
    dzw, zt, zw = u_centered_grid(dzt, None, None, None)
    zt = zt - zw[-1]
    zw = zw - zw[-1]
    return dict(dxt=dxt, dxu=dxu, dyt=dyt, dyu=dyu, cost=cost, cosu=cosu,
                dzt=dzt, dzw=dzw, zt=zt, zw=zw)


def enforce_cyclic_x(arr):
    arr = arr.copy()
    arr[-2:] = arr[2:4]
    arr[:2] = arr[-4:-2]
    return arr


def make_kbot(rng, nx, ny, nz, *, kmin=1, kmax=None, land_frac=0.0, cyclic=True):
    """Random bottom index per column: ``U{kmin..kmax}``, then a fraction set to
    land (0); the domain is closed in y and in x unless cyclic
    (veros/core/numerics.py:185-193)."""
    kmax = nz if kmax is None else kmax
    X, Y = nx + 4, ny + 4
    kbot = np.zeros((X, Y), dtype=np.int32)
    kbot[2:-2, 2:-2] = rng.integers(kmin, kmax + 1, size=(nx, ny))
    if land_frac > 0:
        land = rng.random((nx, ny)) < land_frac
        kbot[2:-2, 2:-2][land] = 0
    kbot[:, :2] = 0
    kbot[:, -2:] = 0
    if cyclic:
        kbot = enforce_cyclic_x(kbot)
    else:
        kbot[:2, :] = 0
        kbot[-2:, :] = 0
    return kbot


def masks_from_kbot(kbot, nz, cyclic=True):
    """maskT/U/V/W exactly as calc_topo_kernel builds them
    (veros/core/numerics.py:195-221)."""
    ks = np.arange(nz)[None, None, :]
    land = kbot > 0
    maskT = land[..., None] & (kbot[..., None] - 1 <= ks)
    if cyclic:
        maskT = enforce_cyclic_x(maskT)
    maskU = maskT.copy()
    maskU[:-1] = np.minimum(maskT[:-1], maskT[1:])
    if cyclic:
        maskU = enforce_cyclic_x(maskU)
    maskV = maskT.copy()
    maskV[:, :-1] = np.minimum(maskT[:, :-1], maskT[:, 1:])
    if cyclic:
        maskV = enforce_cyclic_x(maskV)
    maskW = maskT.copy()
    maskW[:, :, :-1] = np.minimum(maskT[:, :, :-1], maskT[:, :, 1:])
    return dict(maskT=maskT, maskU=maskU, maskV=maskV, maskW=maskW)


# ----------------------------------------------------------------------------
# tridiagonal systems (solve_implicit inputs)
# ----------------------------------------------------------------------------

@dataclass
class TdmaProblem:
    a: np.ndarray
    b: np.ndarray
    c: np.ndarray
    d: np.ndarray
    kbot: np.ndarray          # (nx, ny) int32, 0 = land
    b_edge: np.ndarray | None = None
    d_edge: np.ndarray | None = None
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.a.shape


def diffusion_tdma(rng, nx, ny, nz, *, dz, dt=3600.0, kmin=1, kmax=None, land_frac=0.3,
                   independent_ac=False, with_b_edge=False, with_d_edge=False):
    """Vertical-diffusion systems: ``a = c = -dt*kappa/dz^2`` with
    ``kappa ~ LogU(1e-5, 1e-2)``, ``b = 1 - a - c``, ``d ~ N(0, 1)``;
    per-column ``kbot ~ U{kmin..kmax}`` with a land fraction set to 0."""
    kmax = nz if kmax is None else kmax
    shape = (nx, ny, nz)
    dz = np.asarray(dz, dtype=np.float64)[None, None, :]
    lo, hi = math.log(1e-5), math.log(1e-2)
    kappa = np.exp(rng.uniform(lo, hi, size=shape))
    a = -dt * kappa / dz ** 2
    if independent_ac:
        kappa2 = np.exp(rng.uniform(lo, hi, size=shape))
        c = -dt * kappa2 / dz ** 2
    else:
        c = a.copy()
    b = 1.0 - a - c
    d = rng.standard_normal(shape)
    kbot = rng.integers(kmin, kmax + 1, size=(nx, ny)).astype(np.int32)
    if land_frac > 0:
        kbot[rng.random((nx, ny)) < land_frac] = 0
    b_edge = (1.0 - c + 0.1 * rng.random(shape)) if with_b_edge else None
    d_edge = rng.standard_normal(shape) if with_d_edge else None
    return TdmaProblem(a, b, c, d, kbot, b_edge, d_edge,
                       meta=dict(nx=nx, ny=ny, nz=nz, land_frac=land_frac))


def tdma_config0(seed=1000):
    """64x64x40 FP64 TDMA, dz stretched 10->500 m, 30% land."""
    rng = np.random.default_rng(seed)
    dz = geometric_dz(40, 10.0, 500.0)
    return diffusion_tdma(rng, 64, 64, 40, dz=dz, land_frac=0.3)


def tdma_config2(seed=1002, nx=1440, ny=720, nz=80):
    """1440x720x80 TDMA, dz 10->250 m, kbot ~ U{5..80}, 35% land."""
    rng = np.random.default_rng(seed)
    dz = geometric_dz(nz, 10.0, 250.0)
    return diffusion_tdma(rng, nx, ny, nz, dz=dz, kmin=min(5, nz), land_frac=0.35)


# ----------------------------------------------------------------------------
# isoneutral states
# ----------------------------------------------------------------------------

def iso_state(rng, nx, ny, nz, *, dx_deg, dy_deg, y_origin, dz_top, dz_bottom,
              kmin=1, land_frac=0.25, cyclic=True, settings=None, tau=0, taup1=1,
              k_iso=1000.0, random_k_iso=False, stratified=True, jitter=0.0,
              random_aux=False):
    """A state holding every variable the isoneutral pair reads.

    ``stratified``: ``temp = 20*exp(zt/800) + 0.1*N(0,1)``,
    ``salt = 35 + 0.3*N(0,1)`` (the ACC-like channel); otherwise both are
    ``N(0,1)`` noise around 10 / 35 so that slopes of every sign occur.
    ``random_aux`` fills the in/out fields (``Ai_*``, ``K_*``, ``dtemp_iso``,
    ``P_diss_iso``, other time levels) with noise so tests see whether untouched
    index ranges are preserved."""
    X, Y = nx + 4, ny + 4
    if settings is None:
        settings = MixSettings(nx=nx, ny=ny, nz=nz, enable_cyclic_x=cyclic)
    grid = make_grid(nx, ny, nz, dx_deg=dx_deg, dy_deg=dy_deg, y_origin=y_origin,
                     dz_top=dz_top, dz_bottom=dz_bottom, cyclic=cyclic,
                     coord_degree=settings.coord_degree, rng=rng, jitter=jitter)
    kbot = make_kbot(rng, nx, ny, nz, kmin=kmin, land_frac=land_frac, cyclic=cyclic)
    masks = masks_from_kbot(kbot, nz, cyclic=cyclic)
    zt = grid["zt"]
    if stratified:
        base_t = 20.0 * np.exp(zt / 800.0)[None, None, :]
        temp1 = base_t + 0.1 * rng.standard_normal((X, Y, nz))
        salt1 = 35.0 + 0.3 * rng.standard_normal((X, Y, nz))
    else:
        temp1 = 10.0 + rng.standard_normal((X, Y, nz))
        salt1 = 35.0 + rng.standard_normal((X, Y, nz))
    if cyclic:
        temp1 = enforce_cyclic_x(temp1)
        salt1 = enforce_cyclic_x(salt1)
    temp = np.repeat(temp1[..., None], 3, axis=-1)
    salt = np.repeat(salt1[..., None], 3, axis=-1)
    if random_aux:
        temp += 0.01 * rng.standard_normal(temp.shape)
        salt += 0.01 * rng.standard_normal(salt.shape)
    if random_k_iso:
        K_iso = k_iso * (0.5 + rng.random((X, Y, nz)))
    else:
        K_iso = np.full((X, Y, nz), float(k_iso))
    absz = np.abs(zt)[None, None, :]
    int_drhodT = np.repeat((absz * settings.drhodT * np.ones((X, Y, 1)))[..., None], 3, axis=-1)
    int_drhodS = np.repeat((absz * settings.drhodS * np.ones((X, Y, 1)))[..., None], 3, axis=-1)
    if random_aux:
        int_drhodT = int_drhodT * (1 + 0.1 * rng.standard_normal(int_drhodT.shape))
        int_drhodS = int_drhodS * (1 + 0.1 * rng.standard_normal(int_drhodS.shape))

    def aux(shape):
        return rng.standard_normal(shape) if random_aux else np.zeros(shape)

    vs = SimpleNamespace(
        nx=nx, ny=ny, nz=nz, tau=tau, taup1=taup1,
        temp=temp, salt=salt, K_iso=K_iso, kbot=kbot,
        int_drhodT=int_drhodT, int_drhodS=int_drhodS,
        Ai_ez=aux((X, Y, nz, 2, 2)), Ai_nz=aux((X, Y, nz, 2, 2)),
        Ai_bx=aux((X, Y, nz, 2, 2)), Ai_by=aux((X, Y, nz, 2, 2)),
        K_11=aux((X, Y, nz)), K_22=aux((X, Y, nz)), K_33=aux((X, Y, nz)),
        dtemp_iso=aux((X, Y, nz)), dsalt_iso=aux((X, Y, nz)), P_diss_iso=aux((X, Y, nz)),
        **grid, **masks,
    )
    return SimpleNamespace(variables=vs, settings=settings)


def iso_config1(seed=1001, nx=360, ny=160, nz=60, **kw):
    """360x160x60 ACC-like channel: 1 degree, dz 10->250 m, 25% land, cyclic x."""
    rng = np.random.default_rng(seed)
    return iso_state(rng, nx, ny, nz, dx_deg=1.0, dy_deg=1.0, y_origin=-79.0,
                     dz_top=10.0, dz_bottom=250.0, land_frac=0.25, **kw)


def iso_config2(seed=1002, nx=1440, ny=720, nz=80, **kw):
    """1440x720x80 quarter-degree global-like grid, kbot ~ U{5..80}, 35% land."""
    rng = np.random.default_rng(seed + 100)
    return iso_state(rng, nx, ny, nz, dx_deg=0.25, dy_deg=0.25, y_origin=-89.75,
                     dz_top=10.0, dz_bottom=250.0, kmin=min(5, nz), land_frac=0.35, **kw)


def mixing_config4(seed=1004, nx=180, ny=90, nz=40):
    """The differentiated step's inputs on a 180x90x40 grid: a config-1 style
    state (2 degree channel, dz 10->250 m, 25% land), config-0 style
    vertical-diffusion coefficients on its interior columns (same kbot),
    and loss weights w ~ N(0, 1) on T_new. Returns (state, TdmaProblem, w)."""
    rng = np.random.default_rng(seed)
    st = iso_state(rng, nx, ny, nz, dx_deg=2.0, dy_deg=1.0, y_origin=-79.0,
                   dz_top=10.0, dz_bottom=250.0, land_frac=0.25)
    st.settings.enable_conserve_energy = False
    dz = geometric_dz(nz, 10.0, 250.0)
    p = diffusion_tdma(rng, nx, ny, nz, dz=dz, land_frac=0.0, with_b_edge=True)
    p.kbot = np.ascontiguousarray(st.variables.kbot[2:-2, 2:-2]).astype(np.int32)
    w = rng.standard_normal(st.variables.K_iso.shape)
    return st, p, w


def add_vertmix_fields(rng, state, *, kappa_lo=1e-5, kappa_hi=1e-2, forc_scale=1e-5, random_aux=False):
    """Add the fields vertmix_tempsalt reads and writes to an isoneutral state:
    ``kappaH ~ LogU(kappa_lo, kappa_hi)`` m^2/s per cell (the vertical
    diffusivity of the standalone TDMA configs), surface forcing
    ``forc_temp_surface``, ``forc_salt_surface ~ forc_scale * N(0, 1)`` per
    column (K m/s, g/kg m/s), and the outputs ``dtemp_vmix``, ``dsalt_vmix``
    (noise with ``random_aux``, so tests see that every cell is written)."""
    vs = state.variables
    X, Y, nz = vs.K_iso.shape
    vs.kappaH = np.exp(rng.uniform(np.log(kappa_lo), np.log(kappa_hi), size=(X, Y, nz)))
    vs.forc_temp_surface = forc_scale * rng.standard_normal((X, Y))
    vs.forc_salt_surface = 0.1 * forc_scale * rng.standard_normal((X, Y))
    vs.dtemp_vmix = rng.standard_normal((X, Y, nz)) if random_aux else np.zeros((X, Y, nz))
    vs.dsalt_vmix = rng.standard_normal((X, Y, nz)) if random_aux else np.zeros((X, Y, nz))
    return state


def add_friction_fields(rng, state, *, kappa_lo=1e-5, kappa_hi=1e-2, u_scale=0.1, random_aux=False):
    """Add the fields implicit_vert_friction reads and writes to an isoneutral
    state: velocities ``u``, ``v ~ u_scale * N(0, 1)`` m/s at every time level,
    ``kappaM ~ LogU(kappa_lo, kappa_hi)`` m^2/s per cell, the cell and
    v-face areas ``area_t = dxt cost dyt``, ``area_v = dxt cosu dyu``
    (veros/core/numerics.py:136-138), and the in/out fields ``du_mix``,
    ``dv_mix``, ``K_diss_v`` (noise with ``random_aux``, so tests see which
    cells are written)."""
    vs = state.variables
    X, Y, nz = vs.K_iso.shape
    vs.u = u_scale * rng.standard_normal((X, Y, nz, 3))
    vs.v = u_scale * rng.standard_normal((X, Y, nz, 3))
    vs.kappaM = np.exp(rng.uniform(np.log(kappa_lo), np.log(kappa_hi), size=(X, Y, nz)))
    vs.area_t = vs.cost[None, :] * vs.dyt[None, :] * vs.dxt[:, None]
    vs.area_v = vs.cosu[None, :] * vs.dyu[None, :] * vs.dxt[:, None]
    vs.du_mix = rng.standard_normal((X, Y, nz)) if random_aux else np.zeros((X, Y, nz))
    vs.dv_mix = rng.standard_normal((X, Y, nz)) if random_aux else np.zeros((X, Y, nz))
    vs.K_diss_v = np.abs(rng.standard_normal((X, Y, nz))) * 1e-9 if random_aux else np.zeros((X, Y, nz))
    return state


def add_turbulence_fields(rng, state, *, random_aux=False):
    """Add the fields the TKE, EKE and IDEMIX integrations read and write
    (veros/core/tke.py:123-321, eke.py:93-255, idemix.py:47-265) to an
    isoneutral state: energies tke, eke, E_iw ``(X, Y, nz, 3)`` (E_iw with a
    few negative values, so that max(0, E_iw) matters; the surface TKE
    forcing with both signs, so that the surface correction acts), their
    diffusivities and scales (kappaM ~ LogU(1e-5, 1e-2), sqrttke =
    sqrt(tke[tau]), mxl ~ U(1, 100) m, sqrteke, eke_len ~ U(1e3, 1e5) m,
    K_gm ~ U(150, 450)), the IDEMIX parameters c0, v0, alpha_c (times maskW,
    as set_idemix_parameter forms them), and the dissipation terms that feed
    the three (~1e-9 W/kg). ``random_aux`` fills the outputs with noise so
    tests see which cells are written."""
    vs = state.variables
    X, Y, nz = vs.K_iso.shape
    tau = int(vs.tau)
    maskW = vs.maskW.astype(np.float64)
    f3 = (X, Y, nz)
    if not hasattr(vs, "kappaM"):
        vs.kappaM = np.exp(rng.uniform(np.log(1e-5), np.log(1e-2), size=f3))
    vs.tke = rng.uniform(1e-5, 1e-2, size=(X, Y, nz, 3))
    vs.sqrttke = np.sqrt(np.maximum(0.0, vs.tke[..., tau]))
    vs.mxl = rng.uniform(1.0, 100.0, size=f3)
    vs.forc_tke_surface = 2e-5 * rng.standard_normal((X, Y))
    vs.eke = rng.uniform(1e-4, 1e-2, size=(X, Y, nz, 3))
    vs.sqrteke = np.sqrt(vs.eke[..., tau])
    vs.eke_len = rng.uniform(1e3, 1e5, size=f3)
    vs.K_gm = rng.uniform(150.0, 450.0, size=f3)
    vs.E_iw = rng.uniform(-1e-5, 1e-3, size=(X, Y, nz, 3))
    vs.dE_iw = rng.standard_normal((X, Y, nz, 3)) if random_aux else np.zeros((X, Y, nz, 3))
    vs.c0 = rng.uniform(0.0, 1.0, size=f3) * maskW
    vs.v0 = rng.uniform(0.0, 0.5, size=f3) * maskW
    vs.alpha_c = rng.uniform(0.0, 1e-3, size=f3) * maskW
    vs.forc_iw_bottom = 1e-6 * rng.uniform(0.0, 1.0, size=(X, Y))
    vs.forc_iw_surface = 1e-6 * rng.uniform(0.0, 1.0, size=(X, Y))
    for k in ("K_diss_v", "P_diss_v", "P_diss_adv", "P_diss_nonlin", "eke_diss_tke", "iw_diss", "K_diss_bot",
              "eke_diss_iw", "K_diss_gm", "K_diss_h", "P_diss_skew", "P_diss_hmix", "P_diss_iso"):
        setattr(vs, k, 1e-9 * rng.standard_normal(f3) * maskW)
    vs.tke_surf_corr = rng.standard_normal((X, Y)) if random_aux else np.zeros((X, Y))
    vs.tke_diss = rng.standard_normal(f3) if random_aux else np.zeros(f3)
    return state


def turbulence_settings(**kw):
    """The settings the TKE / EKE / IDEMIX integrations read, with veros's
    defaults (veros/settings.py:28-128), overridden by ``kw``."""
    d = dict(nz=None, dt_mom=3600.0, dt_tracer=3600.0, alpha_tke=1.0, c_eps=0.7, K_h_tke=2000.0,
             enable_tke_hor_diffusion=False, enable_tke_superbee_advection=False, enable_tke_upwind_advection=False,
             enable_store_cabbeling_heat=False, enable_store_bottom_friction_tke=False, enable_eke=False,
             enable_idemix=False, eke_c_eps=1.0, alpha_eke=1.0, enable_eke_superbee_advection=False,
             enable_eke_upwind_advection=False, tau_v=2.0 * 86400.0, tau_h=15.0 * 86400.0,
             enable_idemix_hor_diffusion=False, enable_eke_diss_bottom=False, enable_eke_diss_surfbot=False,
             eke_diss_surfbot_frac=1.0, enable_idemix_superbee_advection=False,
             enable_idemix_upwind_advection=False, enable_cyclic_x=True)
    d.update(kw)
    return SimpleNamespace(**d)


def friction_config2(seed=1002, nx=1440, ny=720, nz=80):
    """The 1440x720x80 grid of the mixing step with the vertical-friction fields."""
    st = iso_config2(seed, nx, ny, nz)
    return add_friction_fields(np.random.default_rng(seed + 300), st)


def vertmix_config2(seed=1002, nx=1440, ny=720, nz=80):
    """The 1440x720x80 grid of the mixing step with the vertical-mixing fields."""
    st = iso_config2(seed, nx, ny, nz)
    return add_vertmix_fields(np.random.default_rng(seed + 200), st)


def config2_settings(nx=1440, ny=720, nz=80):
    """Linear EOS rho = 1025*(1 - 2e-4*T + 7.6e-4*(S-35)) and ACC mixing constants."""
    return MixSettings(nx=nx, ny=ny, nz=nz, drhodT=-1025.0 * 2e-4, drhodS=1025.0 * 7.6e-4)


def state_to_numpy_copy(state):
    """Deep copy of a SimpleNamespace state (arrays copied)."""
    vs = state.variables
    new = SimpleNamespace(**{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in vars(vs).items()})
    return SimpleNamespace(variables=new, settings=state.settings.copy())


# ----------------------------------------------------------------------------
# sub-windows of a state (for checking a large grid on a block of columns)
# ----------------------------------------------------------------------------

X_ARRAYS = ("dxt", "dxu", "xt", "xu")
Y_ARRAYS = ("dyt", "dyu", "cost", "cosu", "yt", "yu")


def _host(v):
    if hasattr(v, "detach"):
        return v.detach().cpu().numpy()
    return np.asarray(v)


def whole_grid_masks(state):
    """maskT/U/V/W of the whole grid (NumPy), as window_state builds them."""
    st = state.settings
    return masks_from_kbot(_host(state.variables.kbot), st.nz, cyclic=bool(getattr(st, "enable_cyclic_x", True)))


def window_state(state, i0, i1, j0, j1, *, masks=True, names=None, full_masks=None):
    """A NumPy copy of the columns ``[i0, i1) x [j0, j1)`` of ``state`` (indices
    of the halo-inclusive arrays; the window's outer two rows and columns act
    as its halos). Arrays with leading ``(X, Y)`` axes are sliced on both,
    the 1-D x metrics on ``[i0, i1)``, the y metrics on ``[j0, j1)``; other
    arrays and scalars are copied. With ``masks``, maskT/U/V/W are built from
    the whole grid's ``kbot`` (cyclic in x as ``settings.enable_cyclic_x``;
    or taken from ``full_masks``) and then sliced, so they equal the whole
    grid's masks on the window.

    On the window's cells ``[2:-2, 2:-2]`` the isoneutral and vertical-mixing
    operators read only cells of the window, so their results there equal the
    whole grid's."""
    vs = state.variables
    X, Y = tuple(vs.kbot.shape)
    out = {}
    for k, v in vars(vs).items():
        if names is not None and k not in names and k not in ("tau", "taup1", "taum1", "kbot"):
            continue
        if not hasattr(v, "shape") or np.ndim(v) == 0:
            out[k] = v
            continue
        shp = tuple(v.shape)
        if len(shp) >= 2 and shp[:2] == (X, Y):
            out[k] = _host(v[i0:i1, j0:j1]).copy()
        elif len(shp) == 1 and shp[0] == X and k in X_ARRAYS:
            out[k] = _host(v[i0:i1]).copy()
        elif len(shp) == 1 and shp[0] == Y and k in Y_ARRAYS:
            out[k] = _host(v[j0:j1]).copy()
        else:
            out[k] = _host(v).copy()
    st = state.settings
    if masks:
        full = full_masks if full_masks is not None else whole_grid_masks(state)
        for k, m in full.items():
            out[k] = m[i0:i1, j0:j1].copy()
    settings = st.copy(nx=i1 - i0 - 4, ny=j1 - j0 - 4) if hasattr(st, "copy") else st
    return SimpleNamespace(variables=SimpleNamespace(**out), settings=settings)
