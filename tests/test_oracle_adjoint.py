"""Pins of the oracle's derivatives (oracle/adjoint.py) on the CPU: central
finite differences (eps = 1e-6) of the oracle forward along random
directions, the affine dependence the brute-force step VJP relies on, and
the closed form of a single wet cell."""
import numpy as np
import pytest

from oracle import adjoint as A
from oracle import isoneutral as O
from oracle import tdma as T
from veros_amd import workloads as W

EPS = 1e-6


def _tdma(seed, nx=4, ny=3, nz=7, edges=True):
    rng = np.random.default_rng(seed)
    dz = W.geometric_dz(nz, 10.0, 500.0)
    p = W.diffusion_tdma(rng, nx, ny, nz, dz=dz, land_frac=0.2, independent_ac=True, with_b_edge=edges,
                         with_d_edge=edges)
    p.kbot[0, 0] = 1          # a full column
    p.kbot[-1, -1] = nz       # a single wet cell
    return p, rng


@pytest.mark.parametrize("edges", [True, False])
def test_tdma_vjp_matches_central_differences(edges):
    p, rng = _tdma(11 + edges, edges=edges)
    gx = rng.standard_normal(p.a.shape)
    names = ["a", "b", "c", "d"] + (["b_edge", "d_edge"] if edges else [])
    args = {n: getattr(p, n) for n in ("a", "b", "c", "d", "b_edge", "d_edge")}

    def loss(**kw):
        v = {**args, **kw}
        x = T.solve_implicit_kbot(v["a"], v["b"], v["c"], v["d"], p.kbot, b_edge=v["b_edge"], d_edge=v["d_edge"])
        return float(np.sum(gx * x))

    x = T.solve_implicit_kbot(p.a, p.b, p.c, p.d, p.kbot, b_edge=p.b_edge, d_edge=p.d_edge)
    grads = dict(zip(("a", "b", "c", "d", "b_edge", "d_edge"),
                     A.tdma_vjp(p.a, p.b, p.c, p.kbot, x, gx, b_edge=p.b_edge, d_edge=p.d_edge)))
    for n in names:
        v = np.max(np.abs(args[n])) * rng.standard_normal(p.a.shape)
        fd = (loss(**{n: args[n] + EPS * v}) - loss(**{n: args[n] - EPS * v})) / (2 * EPS)
        ad = float(np.sum(grads[n] * v))
        assert abs(fd - ad) <= 1e-6 * abs(ad), (n, fd, ad)
    if not edges:
        assert grads["b_edge"] is None and grads["d_edge"] is None


def test_tdma_vjp_single_cell_closed_form():
    """One wet cell: x = d / b, so dL/dd = g / b and dL/db = -g d / b^2."""
    a = np.zeros((1, 1, 3))
    b = np.array([[[1.0, 2.0, 4.0]]])
    c = np.zeros((1, 1, 3))
    d = np.array([[[7.0, 3.0, 5.0]]])
    kb = np.array([[3]], dtype=np.int32)
    x = T.solve_implicit_kbot(a, b, c, d, kb, solver="dense")  # LAPACK dgtsv rejects a 1-cell system
    g = np.array([[[1.0, 1.0, 2.0]]])
    ga, gb, gc, gd, _, _ = A.tdma_vjp(a, b, c, kb, x, g)
    assert np.allclose(gd, [[[0, 0, 0.5]]]) and np.allclose(gb, [[[0, 0, -2.0 * 5 / 16]]])
    assert not ga.any() and not gc.any()


def _iso_small(seed, nx=5, ny=4, nz=4):
    rng = np.random.default_rng(seed)
    st = W.iso_state(rng, nx, ny, nz, dx_deg=1.0, dy_deg=1.0, y_origin=-50.0, dz_top=10.0, dz_bottom=80.0,
                     land_frac=0.25, stratified=False, random_k_iso=True, jitter=0.02, tau=1, taup1=0)
    st.settings.K_iso_steep = 20.0
    pre = O.isoneutral_diffusion_pre(st)
    for k, v in pre._asdict().items():
        setattr(st.variables, k, v)
    vs = st.variables
    return st, rng, vs.temp[..., vs.tau].copy(), vs.temp[..., vs.taup1].copy() + rng.standard_normal(vs.K_iso.shape)


def test_iso_step_is_affine():
    st, rng, T0, T1, = _iso_small(3)
    K = st.variables.K_iso
    f = lambda a, b, k: A.iso_step(st, a, b, k)
    dT = rng.standard_normal(T0.shape)
    dK = K * rng.random(K.shape)
    lhs = f(T0 + dT, T1 + dT, K) - f(T0, T1, K)
    rhs = f(dT, dT, K) - f(0 * T0, 0 * T1, K)
    assert np.max(np.abs(lhs - rhs)) <= 1e-12 * np.max(np.abs(lhs))
    lhs = f(T0, T1, K + dK) - f(T0, T1, K)
    rhs = f(T0, T1, K + 2 * dK) - f(T0, T1, K + dK)
    assert np.max(np.abs(lhs - rhs)) <= 1e-12 * np.max(np.abs(f(T0, T1, K)))


def test_iso_step_vjp_matches_central_differences():
    st, rng, T0, T1 = _iso_small(5)
    K = st.variables.K_iso
    g = rng.standard_normal(K.shape)
    gT0, gT1, gK = A.iso_step_vjp(st, T0, T1, K, g)
    loss = lambda a, b, k: float(np.sum(g * A.iso_step(st, a, b, k)))
    # The gradients do not depend on T_taup1 (the step is affine in it), so the
    # differences are taken at T_taup1 = 0: T_new then holds only the diffusive
    # increment and rounding in the loss stays far below the signal.
    Z = np.zeros(K.shape)
    for grad, which, scale in ((gT0, 0, 1.0), (gT1, 1, 1.0), (gK, 2, K)):
        v = scale * rng.standard_normal(K.shape)
        args_p = [T0, Z, K]
        args_m = [T0, Z, K]
        args_p[which] = args_p[which] + EPS * v
        args_m[which] = args_m[which] - EPS * v
        fd = (loss(*args_p) - loss(*args_m)) / (2 * EPS)
        ad = float(np.sum(grad * v))
        assert abs(fd - ad) <= 1e-6 * abs(ad), (which, fd, ad)
    # the halo's outer ring never reaches T_new
    assert not gT0[0].any() and not gT0[-1].any() and not gK[0].any()
    # T_taup1 passes through unchanged outside the interior
    assert np.allclose(gT1[:2], g[:2], rtol=0, atol=1e-10)
