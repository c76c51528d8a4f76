"""Pins of the TKE / EKE / IDEMIX oracle (oracle/turbulence.py).

* The reference's own outputs (tests/golden/{tke,eke,idemix}_*.npz, written by
  tests/golden/make_golden.py from veros.core.tke.integrate_tke_kernel,
  veros.core.eke.integrate_eke_kernel and veros.core.idemix.integrate_idemix_kernel
  on seeded states and several settings combinations): 1e-13 relative per
  output (both sides solve with LAPACK dgtsv; the forcing sums may associate
  differently, and IDEMIX's column integral is summed in another order).
* Independent of the reference: with the lateral terms off, every wet column
  equals a dense solve of the matrix built from the formulas, and a field
  that is uniform in the vertical stays uniform without forcing or dissipation.
"""
import numpy as np
import pytest

from conftest import golden_files, load_golden, rel_err, turbulence_state_from_golden
from oracle import turbulence as OT
from oracle.tdma import create_water_masks
from veros_amd import workloads as W

FNS = {"tke": OT.integrate_tke_kernel, "eke": OT.integrate_eke_kernel, "idemix": OT.integrate_idemix_kernel}
FILES = golden_files("tke_") + golden_files("eke_") + golden_files("idemix_")


@pytest.mark.parametrize("path", FILES, ids=lambda p: p.stem)
def test_oracle_matches_reference(path):
    rec = load_golden(path)
    st = turbulence_state_from_golden(rec)
    out = FNS[str(rec["model"])](st)
    outs = [k[4:] for k in rec if k.startswith("out_")]
    assert len(outs) == 3
    for k in outs:
        assert rel_err(getattr(out, k), rec["out_" + k]) <= 1e-13, k


def test_fixtures_cover_written_ranges():
    for path in FILES:
        rec = load_golden(path)
        name = {"tke": "tke", "eke": "eke", "idemix": "E_iw"}[str(rec["model"])]
        a, b = rec["in_" + name], rec["out_" + name]
        assert np.any(a != b) and np.any(a == b), path.stem


def _state(seed=9, nx=9, ny=7, nz=7, **kw):
    rng = np.random.default_rng(seed)
    st = W.iso_state(rng, nx, ny, nz, dx_deg=1.0, dy_deg=1.0, y_origin=-40.0, dz_top=10.0, dz_bottom=80.0,
                     land_frac=0.3, tau=0, taup1=1, random_aux=True)
    W.add_turbulence_fields(rng, st, random_aux=True)
    st.settings = OT.turbulence_settings(nz=nz, **kw)
    return st


def _dense_column(a, b, c, d, k0):
    n = len(b) - k0
    M = np.diag(b[k0:]) + np.diag(a[k0 + 1:], -1) + np.diag(c[k0:-1], 1)
    return np.linalg.solve(M, d[k0:])


def test_tke_columns_equal_dense_solves():
    """Rows written out from veros/core/tke.py:183-225 level by level."""
    st = _state()
    vs, s = st.variables, st.settings
    s.K_h_tke = 0.0
    out = OT.integrate_tke_kernel(st)
    nz, dt = s.nz, s.dt_mom
    forc = OT.tke_forcing(vs, s)
    for i in range(2, vs.kbot.shape[0] - 2):
        for j in range(2, vs.kbot.shape[1] - 2):
            kb = vs.kbot[i, j]
            if kb < 1:
                continue
            k0 = kb - 1
            km, sq, ml = vs.kappaM[i, j], vs.sqrttke[i, j], vs.mxl[i, j]
            dl = [dt / vs.dzt[k + 1] * s.alpha_tke * 0.5 * (km[k] + km[k + 1]) if k < nz - 1 else 0.0
                  for k in range(nz)]
            a, b, c = np.zeros(nz), np.zeros(nz), np.zeros(nz)
            for k in range(nz):
                h = 0.5 * vs.dzw[k] if k == nz - 1 else vs.dzw[k]
                a[k] = -dl[k - 1] / h if k >= 1 else 0.0
                b[k] = 1 + (dl[k] + (dl[k - 1] if k >= 1 else 0.0)) / h + dt * s.c_eps * sq[k] / ml[k]
                c[k] = -dl[k] / vs.dzw[k] if k < nz - 1 else 0.0
            b[k0] = 1 + dl[k0] / vs.dzw[k0] + dt * s.c_eps * sq[k0] / ml[k0]
            d = vs.tke[i, j, :, 0] + dt * forc[i, j]
            d[-1] += dt * vs.forc_tke_surface[i, j] / (0.5 * vs.dzw[-1])
            x = _dense_column(a, b, c, d, k0)
            x[-1] = max(0.0, x[-1])
            np.testing.assert_allclose(out.tke[i, j, k0:, 1], x, rtol=1e-10, atol=1e-18)


def test_idemix_uniform_fixed_point():
    """Without forcing or dissipation, a vertically uniform E_iw is a fixed
    point: the diffusion rows sum to one (c0 enters a and c as the flux form
    delta * (c0[k+1] E[k+1] - c0[k] E[k]) does only when c0 is uniform)."""
    st = _state(enable_eke=False)
    vs = st.variables
    for k in ("K_diss_gm", "K_diss_h", "P_diss_skew", "K_diss_bot", "eke_diss_iw"):
        setattr(vs, k, np.zeros_like(getattr(vs, k)))
    vs.alpha_c[:] = 0.0
    vs.forc_iw_bottom[:] = 0.0
    vs.forc_iw_surface[:] = 0.0
    vs.c0[:] = 0.7
    vs.E_iw[..., 0] = 2.5e-4
    out = OT.integrate_idemix_kernel(st)
    _, water, _ = create_water_masks(vs.kbot[2:-2, 2:-2], st.settings.nz)
    got = out.E_iw[2:-2, 2:-2, :, 1][water]
    # the rows' entries are dt * tau_v * c0^2 / dz^2 ~ 1e7 (tau_v = 2 days), so
    # the solve keeps about 1e7 * eps of the uniform value
    np.testing.assert_allclose(got, 2.5e-4, rtol=1e-8)


def test_advection_branches_raise():
    st = _state(enable_tke_upwind_advection=True)
    with pytest.raises(NotImplementedError):
        OT.integrate_tke_kernel(st)
