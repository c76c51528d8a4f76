"""Pins of the vertical-friction oracle (oracle/friction.py) against what the
mathematics fixes: identity solve, fixed points, conservation, dense solves,
and a two-level closed form of the dissipation and its interpolation."""
import numpy as np

from oracle import friction as F
from veros_amd import workloads as W


def _state(seed=5, nx=9, ny=7, nz=6, cyclic=True, land=0.3, **kw):
    rng = np.random.default_rng(seed)
    st = W.iso_state(rng, nx, ny, nz, dx_deg=1.0, dy_deg=1.0, y_origin=-40.0, dz_top=10.0, dz_bottom=80.0,
                     land_frac=land, cyclic=cyclic, tau=0, taup1=1, random_aux=True, jitter=0.1)
    return W.add_friction_fields(rng, st, random_aux=True, **kw)


def _face_wet(vs, axis):
    """Water mask of the [1:-2, 1:-2] face columns: k >= max(kbot, kbot_next) - 1."""
    kb = vs.kbot
    nb = kb[2:-1, 1:-2] if axis == 0 else kb[1:-2, 2:-1]
    kss = np.maximum(kb[1:-2, 1:-2], nb)
    k = np.arange(vs.kappaM.shape[2])[None, None, :]
    return (kss[..., None] >= 1) & (k >= kss[..., None] - 1)


def test_zero_viscosity_is_identity_and_dissipates_nothing():
    st = _state()
    vs = st.variables
    vs.kappaM[:] = 0.0
    out = F.implicit_vert_friction(st)
    for ax, name in ((0, "u"), (1, "v")):
        wet = _face_wet(vs, ax)
        old, new = getattr(vs, name), getattr(out, name)
        want = np.where(wet, old[1:-2, 1:-2, :, 0], old[1:-2, 1:-2, :, 1])
        assert np.array_equal(new[1:-2, 1:-2, :, 1], want)
        d = getattr(out, "d" + name + "_mix")
        assert np.all(d[1:-2, 1:-2][wet] == 0.0)
        # outside the face range the tendencies keep their values
        old_d = getattr(vs, "d" + name + "_mix")
        assert np.array_equal(d[0], old_d[0]) and np.array_equal(d[-2:], old_d[-2:])
    np.testing.assert_array_equal(out.K_diss_v, vs.K_diss_v)


def test_uniform_column_is_fixed_point():
    st = _state(seed=6)
    vs = st.variables
    vs.u[..., 0] = 0.3
    vs.v[..., 0] = -0.2
    out = F.implicit_vert_friction(st)
    for ax, name, val in ((0, "u", 0.3), (1, "v", -0.2)):
        wet = _face_wet(vs, ax)
        np.testing.assert_allclose(getattr(out, name)[1:-2, 1:-2, :, 1][wet], val, rtol=1e-14)
    # no vertical shear at tau: no dissipation
    np.testing.assert_array_equal(out.K_diss_v, vs.K_diss_v)


def test_solve_conserves_column_momentum():
    st = _state(seed=7, land=0.2)
    vs = st.variables
    out = F.implicit_vert_friction(st)
    dzt = vs.dzt[None, None, :]
    for ax, name in ((0, "u"), (1, "v")):
        wet = _face_wet(vs, ax)
        old = np.where(wet, getattr(vs, name)[1:-2, 1:-2, :, 0], 0.0)
        new = np.where(wet, getattr(out, name)[1:-2, 1:-2, :, 1], 0.0)
        np.testing.assert_allclose((new * dzt).sum(-1), (old * dzt).sum(-1), rtol=1e-12, atol=1e-14)


def test_wet_columns_match_dense_solve():
    st = _state(seed=8, nx=6, ny=5, nz=7)
    vs, s = st.variables, st.settings
    out = F.implicit_vert_friction(st)
    dt, dzt, dzw = s.dt_mom, vs.dzt, vs.dzw
    X, Y, nz = vs.kappaM.shape
    nchecked = 0
    for ax, name, mask in ((0, "u", vs.maskU), (1, "v", vs.maskV)):
        for i in range(1, X - 2):
            for j in range(1, Y - 2):
                ii, jj = (i + 1, j) if ax == 0 else (i, j + 1)
                kss = max(vs.kbot[i, j], vs.kbot[ii, jj])
                if kss < 1:
                    continue
                k0 = kss - 1
                n = nz - k0
                # the diffusion matrix of the face column from its definition:
                # flux between levels k and k+1 is delta_k (x_{k+1} - x_k)
                kap = 0.5 * (vs.kappaM[i, j] + vs.kappaM[ii, jj])
                delta = np.array([dt / dzw[k] * kap[k] * mask[i, j, k] * mask[i, j, k + 1] for k in range(nz - 1)] + [0.0])
                A = np.eye(n)
                for r in range(n):
                    k = k0 + r
                    if k < nz - 1:
                        A[r, r] += delta[k] / dzt[k]
                        A[r, r + 1] -= delta[k] / dzt[k]
                    if r > 0:
                        A[r, r] += delta[k - 1] / dzt[k]
                        A[r, r - 1] -= delta[k - 1] / dzt[k]
                want = np.linalg.solve(A, getattr(vs, name)[i, j, k0:, 0])
                np.testing.assert_allclose(getattr(out, name)[i, j, k0:, 1], want, rtol=1e-12, atol=1e-15)
                nchecked += 1
    assert nchecked > 20


def _closed_form_state(ax):
    """All-wet two-level grid, uniform kappaM, one moving face at (4, 3)."""
    st = _state(seed=9, nx=6, ny=5, nz=2, land=0.0)
    vs = st.variables
    vs.kbot[:] = 1
    for m in ("maskT", "maskU", "maskV", "maskW"):
        getattr(vs, m)[:] = True
    vs.kappaM[:] = 2e-3
    vs.u[:] = 0.0
    vs.v[:] = 0.0
    vel = vs.u if ax == 0 else vs.v
    vel[4, 3, 0, 0], vel[4, 3, 1, 0] = 0.05, 0.25
    return st


def test_dissipation_closed_form_and_interpolation():
    for ax in (0, 1):
        st = _closed_form_state(ax)
        vs, s = st.variables, st.settings
        out = F.implicit_vert_friction(st)
        d0, d1 = 0.05, 0.25
        kap, dt, dzw, dzt = 2e-3, s.dt_mom, vs.dzw[0], vs.dzt
        delta = dt / dzw * kap
        # x1 - x0 = (d1 - d0) / (1 + delta (1/dzt0 + 1/dzt1)) from the two rows
        dx_new = (d1 - d0) / (1 + delta * (1 / dzt[0] + 1 / dzt[1]))
        vel = out.u if ax == 0 else out.v
        np.testing.assert_allclose(vel[4, 3, 1, 1] - vel[4, 3, 0, 1], dx_new, rtol=1e-13)
        diss = (d1 - d0) * kap * dx_new / dzw / dzw
        dK = out.K_diss_v - vs.K_diss_v
        want = np.zeros_like(dK)
        if ax == 0:
            # a u face between T columns 4 and 5 feeds both, weighted dxu / (2 dxt)
            want[4, 3, 0] = vs.dxu[4] * diss / (2 * vs.dxt[4])
            want[5, 3, 0] = vs.dxu[4] * diss / (2 * vs.dxt[5])
        else:
            want[4, 3, 0] = vs.area_v[4, 3] * diss / (2 * vs.area_t[4, 3])
            want[4, 4, 0] = vs.area_v[4, 3] * diss / (2 * vs.area_t[4, 4])
        np.testing.assert_allclose(dK, want, rtol=1e-12, atol=1e-12 * abs(diss))
