"""GPU adjoint kernels behind torch.autograd against the oracle's derivatives
(oracle/adjoint.py), central finite differences (eps = 1e-6, bar 1e-6
relative) and dot-product identities, on small grids and on the 180x90x40
differentiated mixing step."""
import numpy as np
import pytest
import torch

from conftest import rel_err
from oracle import adjoint as OA
from oracle import isoneutral as O
from oracle import tdma as OT
from veros_amd import workloads as W

pytestmark = pytest.mark.gpu
EPS = 1e-6
FD_TOL = 1e-6


def _t(a, grad=False):
    if a is None:
        return None
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t.requires_grad_(grad)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("shape", [(13, 11, 40), (7, 5, 1), (9, 6, 17), (5, 4, 100)])
@pytest.mark.parametrize("edges", [True, False])
def test_tdma_adjoint_matches_oracle(shape, edges):
    from veros_amd.autodiff import solve_implicit_ad
    rng = np.random.default_rng(sum(shape) + edges)
    nx, ny, nz = shape
    p = W.diffusion_tdma(rng, nx, ny, nz, dz=W.geometric_dz(nz, 10.0, 500.0), land_frac=0.2,
                         independent_ac=True, with_b_edge=edges, with_d_edge=edges)
    p.kbot[0, 0] = 1
    p.kbot[-1, -1] = nz
    gx = rng.standard_normal(shape)
    ts = {n: _t(getattr(p, n), True) for n in ("a", "b", "c", "d", "b_edge", "d_edge")}
    x = solve_implicit_ad(ts["a"], ts["b"], ts["c"], ts["d"], _t(p.kbot), ts["b_edge"], ts["d_edge"])
    x.backward(_t(gx))
    xr = OT.solve_implicit_kbot(p.a, p.b, p.c, p.d, p.kbot, b_edge=p.b_edge, d_edge=p.d_edge, solver="dense")
    assert rel_err(_np(x), xr) <= 1e-12
    ref = dict(zip(("a", "b", "c", "d", "b_edge", "d_edge"),
                   OA.tdma_vjp(p.a, p.b, p.c, p.kbot, xr, gx, b_edge=p.b_edge, d_edge=p.d_edge)))
    for n, t in ts.items():
        if t is None:
            continue
        assert rel_err(_np(t.grad), ref[n]) <= 1e-12, n


def test_tdma_adjoint_dot_product_config0():
    """<gx, A^-1 v> == <A^-T gx, v> on the 64x64x40 config (x linear in d)."""
    from veros_amd.autodiff import solve_implicit_ad
    p = W.tdma_config0()
    rng = np.random.default_rng(5)
    gx = _t(rng.standard_normal(p.a.shape))
    v = _t(rng.standard_normal(p.a.shape))
    d = _t(p.d, True)
    x = solve_implicit_ad(_t(p.a), _t(p.b), _t(p.c), d, _t(p.kbot))
    x.backward(gx)
    xv = solve_implicit_ad(_t(p.a), _t(p.b), _t(p.c), v, _t(p.kbot))
    lhs = float(torch.sum(gx * xv))
    rhs = float(torch.sum(d.grad * v))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def _iso_state(seed, nx, ny, nz, cyclic=True):
    rng = np.random.default_rng(seed)
    st = W.iso_state(rng, nx, ny, nz, dx_deg=1.0, dy_deg=1.0, y_origin=-50.0, dz_top=10.0, dz_bottom=80.0,
                     land_frac=0.25, stratified=False, random_k_iso=True, jitter=0.02, tau=1, taup1=0,
                     cyclic=cyclic)
    st.settings.K_iso_steep = 20.0
    st.settings.enable_conserve_energy = False
    pre = O.isoneutral_diffusion_pre(st)
    for k, v in pre._asdict().items():
        setattr(st.variables, k, v)
    return st, rng


@pytest.mark.parametrize("shape,cyclic", [((5, 4, 4), True), ((6, 5, 3), False)])
def test_iso_adjoint_matches_oracle(shape, cyclic):
    from veros_amd.autodiff import isoneutral_step
    from veros_amd.state import to_device_state
    st, rng = _iso_state(31 + sum(shape), *shape, cyclic=cyclic)
    vs = st.variables
    X, Y, nz = vs.K_iso.shape
    T0 = vs.temp[..., vs.tau].copy()
    T1 = rng.standard_normal((X, Y, nz))
    K = vs.K_iso.copy()
    g = rng.standard_normal((X, Y, nz))
    ds = to_device_state(st)
    t0, t1, k = _t(T0, True), _t(T1, True), _t(K, True)
    out = isoneutral_step(ds, t0, t1, k)
    assert rel_err(_np(out), OA.iso_step(st, T0, T1, K)) <= 1e-12
    out.backward(_t(g))
    # the gradients do not depend on T_taup1: the brute force runs at T_taup1 = 0
    ref = OA.iso_step_vjp(st, T0, np.zeros_like(T1), K, g)
    for name, t, r in (("T_tau", t0, ref[0]), ("T_taup1", t1, ref[1]), ("K_iso", k, ref[2])):
        assert rel_err(_np(t.grad), r) <= 1e-10, name


def _config4():
    from veros_amd import isoneutral
    from veros_amd.state import to_device_state
    st, p, w = W.mixing_config4()
    ds = to_device_state(st)
    isoneutral.isoneutral_diffusion_pre(ds)   # the tensor the step holds fixed
    return ds, p, w


def _directional_fd(f, x0, v):
    with torch.no_grad():
        return (float(f(x0 + EPS * v)) - float(f(x0 - EPS * v))) / (2 * EPS)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config4_mixing_step_gradients(precision):
    """loss = sum(w * T_new) of solve_implicit then isoneutral_diffusion on
    180x90x40; gradients for d, a, b, c, b_edge, T and K_iso against central
    differences along random directions. The tracer and K_iso enter only the
    isoneutral step's diffusive increment, whose gradients do not depend on the
    taup1 level; their differences are taken through isoneutral_step at
    T_taup1 = 0 so the loss's rounding stays far below the signal."""
    from veros_amd.autodiff import isoneutral_step, mixing_step
    ds, p, w = _config4()
    vs = ds.variables
    wt = _t(w)
    kw = dict(precision=precision, refine_steps=2)
    leaves = {n: _t(getattr(p, n), True) for n in ("a", "b", "c", "d", "b_edge")}
    T_tau = vs.temp[..., vs.tau].contiguous().requires_grad_(True)
    K = vs.K_iso.clone().requires_grad_(True)
    L = torch.sum(wt * mixing_step(ds, leaves["a"], leaves["b"], leaves["c"], leaves["d"], T_tau, K,
                                   b_edge=leaves["b_edge"], **kw))
    L.backward()
    rng = np.random.default_rng(44)
    base = {n: t.detach() for n, t in leaves.items()}
    for n in ("a", "b", "c", "d", "b_edge"):
        v = float(base[n].abs().max()) * _t(rng.standard_normal(p.a.shape))

        def f(x, n=n):
            args = dict(base)
            args[n] = x
            T_new = mixing_step(ds, args["a"], args["b"], args["c"], args["d"], T_tau.detach(), K.detach(),
                                b_edge=args["b_edge"], **kw)
            return torch.sum(wt * T_new)

        fd = _directional_fd(f, base[n], v)
        ad = float(torch.sum(leaves[n].grad * v))
        print(precision, n, fd, ad)
        assert abs(fd - ad) <= FD_TOL * abs(ad), (n, fd, ad)
    Z = torch.zeros_like(K)
    for n, leaf, scale in (("T", T_tau, 1.0), ("K_iso", K, K.detach())):
        v = scale * _t(rng.standard_normal(w.shape))
        if n == "T":
            f = lambda x: torch.sum(wt * isoneutral_step(ds, x, Z, K.detach()))
        else:
            f = lambda x: torch.sum(wt * isoneutral_step(ds, T_tau.detach(), Z, x))
        fd = _directional_fd(f, leaf.detach(), v)
        ad = float(torch.sum(leaf.grad * v))
        print(precision, n, fd, ad)
        assert abs(fd - ad) <= FD_TOL * abs(ad), (n, fd, ad)


def test_config4_mixed_gradients_match_fp64():
    from veros_amd.autodiff import mixing_step
    ds, p, w = _config4()
    vs = ds.variables
    grads = {}
    for prec in ("fp64", "mixed"):
        leaves = {n: _t(getattr(p, n), True) for n in ("a", "b", "c", "d", "b_edge")}
        T_tau = vs.temp[..., vs.tau].contiguous().requires_grad_(True)
        K = vs.K_iso.clone().requires_grad_(True)
        L = torch.sum(_t(w) * mixing_step(ds, leaves["a"], leaves["b"], leaves["c"], leaves["d"], T_tau, K,
                                          b_edge=leaves["b_edge"], precision=prec, refine_steps=2))
        L.backward()
        grads[prec] = {n: _np(t.grad) for n, t in leaves.items()}
        grads[prec].update(T=_np(T_tau.grad), K_iso=_np(K.grad))
    for n, g in grads["fp64"].items():
        assert rel_err(grads["mixed"][n], g) <= 1e-10, n


def test_iso_adjoint_repeatable():
    """Float64 atomics in the scatter: repeated runs agree to rounding."""
    from veros_amd.autodiff import isoneutral_step
    ds, p, w = _config4()
    vs = ds.variables
    res = []
    for _ in range(2):
        T_tau = vs.temp[..., vs.tau].contiguous().requires_grad_(True)
        K = vs.K_iso.clone().requires_grad_(True)
        out = isoneutral_step(ds, T_tau, vs.temp[..., vs.taup1].contiguous(), K)
        out.backward(_t(w))
        res.append((_np(T_tau.grad), _np(K.grad)))
    for a, b in zip(res[0], res[1]):
        assert rel_err(a, b) <= 1e-13
