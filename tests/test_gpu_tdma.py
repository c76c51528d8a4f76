"""GPU parity of the TDMA entry points against the oracle (bar: 1e-12 relative,
max|x - x_ref| / max|x_ref|)."""
import numpy as np
import pytest
import torch

from conftest import golden_files, load_golden, rel_err
from oracle import tdma as O
from veros_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _dev(*arrs):
    return [None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


def _solve(p_or_arrays, kbot, precision="fp64", refine=1, b_edge=None, d_edge=None):
    from veros_amd.utilities import solve_implicit_kbot
    a, b, c, d = _dev(*p_or_arrays)
    be, de = _dev(b_edge, d_edge)
    (kb,) = _dev(kbot.astype(np.int32))
    x = solve_implicit_kbot(a, b, c, d, kb, b_edge=be, d_edge=de, precision=precision, refine_steps=refine)
    return x.cpu().numpy()


def _max_cond(g):
    """Largest 2-norm condition number of the fixture's wet systems."""
    a, b, c, kb = g["a"], g.get("b_edge", g["b"]), g["c"], g["kbot"]
    nz = a.shape[-1]
    worst = 1.0
    for i, j in zip(*np.nonzero((kb >= 1) & (kb <= nz))):
        k0 = kb[i, j] - 1
        diag = g["b"][i, j, k0:].copy()
        if "b_edge" in g:
            diag[0] = b[i, j, k0]
        m = np.diag(diag) + np.diag(a[i, j, k0 + 1:], -1) + np.diag(c[i, j, k0:-1], 1)
        worst = max(worst, np.linalg.cond(m))
    return worst


@pytest.mark.parametrize("path", golden_files("tdma_"), ids=lambda p: p.stem)
@pytest.mark.parametrize("mode", ["fp64", "mixed1", "mixed2"])
def test_golden(path, mode):
    """Every reference fixture, including the standard-normal coefficients of
    the reference's own tridiagonal test (dgtsv pivots there; these systems
    need no pivot, so an elimination without pivoting matches it to rounding).

    The mixed form's error after R refinement steps is about
    (cond * u32)^(R+1) of the solution (u32 = 2^-24, each float32 solve
    leaves a relative error of about cond * u32 and the float64 residual
    step multiplies it by the same factor). On the diffusion fixtures
    cond < 10, so every mode meets 1e-12; the standard-normal systems reach
    cond = 126, where one step guarantees only about 6e-11 and two steps
    1e-15. The bar is therefore max(1e-12, 8 * (cond * u32)^(R+1))."""
    g = load_golden(path)
    prec, refine = ("fp64", 0) if mode == "fp64" else ("mixed", int(mode[-1]))
    x = _solve((g["a"], g["b"], g["c"], g["d"]), g["kbot"], prec, refine, g.get("b_edge"), g.get("d_edge"))
    tol = TOL if prec == "fp64" else max(TOL, 8.0 * (_max_cond(g) * 2.0 ** -24) ** (refine + 1))
    assert rel_err(x, g["x_implicit"]) <= tol
    assert np.all(x[~g["water"]] == 0)


def test_strongly_coupled_friction_systems(monkeypatch):
    """The u and v systems of the strong-viscosity friction fixture
    (tests/golden/friction_16x12x20_strong.npz), captured from the oracle and
    solved through vx_tdma_f64: a few columns have interface rows that are
    not diagonally dominant at all (|a| + |c| up to 5 b), where parallel
    cyclic reduction alone leaves 4e-12 and the interface refinement step
    brings the solve back to the accuracy of a sequential sweep (dgtsv
    against a column Thomas sweep differ by 4e-15 here)."""
    from conftest import column_state_from_golden
    from oracle import friction as F
    rec = load_golden([p for p in golden_files("friction_") if "strong" in p.stem][0])
    captured = []
    dgtsv = O.solve_tridiagonal_dgtsv

    def spy(a, b, c, d, water, edge):
        x = dgtsv(a, b, c, d, water, edge)
        captured.append((a, b, c, d, water, edge, x))
        return x

    monkeypatch.setattr(O, "solve_tridiagonal_dgtsv", spy)
    F.implicit_vert_friction(column_state_from_golden(rec))
    assert len(captured) == 2
    for a, b, c, d, water, edge, ref in captured:
        kbot = np.where(edge.any(-1), np.argmax(edge, axis=-1) + 1, 0).astype(np.int32)
        thomas = O.solve_tridiagonal_thomas(a, b, c, d, water, edge)
        assert rel_err(thomas, ref) <= 1e-14
        x = _solve((a, b, c, d), kbot)
        assert rel_err(np.where(water, x, 0.0), np.where(water, ref, 0.0)) <= TOL


def test_reference_mask_signature():
    from veros_amd.utilities import create_water_masks, solve_implicit
    g = load_golden(golden_files("tdma_diffusion")[0])
    a, b, c, d, be, de = _dev(g["a"], g["b"], g["c"], g["d"], g["b_edge"], g["d_edge"])
    (kb,) = _dev(g["kbot"])
    _, water, edge = create_water_masks(kb, g["a"].shape[-1])
    assert np.array_equal(water.cpu().numpy(), g["water"])
    assert np.array_equal(edge.cpu().numpy(), g["edge"])
    x = solve_implicit(a, b, c, d, water, edge, b_edge=be, d_edge=de)
    assert rel_err(x.cpu().numpy(), g["x_implicit"]) <= TOL
    # NumPy inputs go through the device and come back as NumPy
    xn = solve_implicit(g["a"], g["b"], g["c"], g["d"], g["water"], g["edge"], b_edge=g["b_edge"], d_edge=g["d_edge"])
    assert isinstance(xn, np.ndarray) and rel_err(xn, g["x_implicit"]) <= TOL


@pytest.mark.parametrize("nz", [1, 2, 5, 16, 17, 33, 40, 64, 79, 80, 96, 100, 113, 128])
def test_depth_sweep(nz):
    rng = np.random.default_rng(100 + nz)
    p = W.diffusion_tdma(rng, 23, 19, nz, dz=W.geometric_dz(nz, 10.0, 300.0), land_frac=0.2,
                         independent_ac=True, with_b_edge=True)
    p.kbot[0, :] = nz
    p.kbot[1, :] = 1
    ref = O.solve_implicit_kbot(p.a, p.b, p.c, p.d, p.kbot, b_edge=p.b_edge)
    for prec, refine in (("fp64", 0), ("mixed", 1), ("mixed", 2)):
        x = _solve((p.a, p.b, p.c, p.d), p.kbot, prec, refine, p.b_edge)
        assert rel_err(x, ref) <= TOL, (prec, refine)


def test_mixed_without_refinement_is_float32_accurate():
    p = W.tdma_config0()
    ref = O.solve_implicit_kbot(p.a, p.b, p.c, p.d, p.kbot)
    x = _solve((p.a, p.b, p.c, p.d), p.kbot, "mixed", 0)
    e = rel_err(x, ref)
    assert 1e-12 < e < 1e-5


def test_config0_against_dgtsv():
    p = W.tdma_config0()
    ref = O.solve_implicit_kbot(p.a, p.b, p.c, p.d, p.kbot)
    for prec, refine in (("fp64", 0), ("mixed", 1), ("mixed", 2)):
        assert rel_err(_solve((p.a, p.b, p.c, p.d), p.kbot, prec, refine), ref) <= TOL


def test_odd_kbot_values_give_zero_columns():
    rng = np.random.default_rng(4)
    p = W.diffusion_tdma(rng, 8, 8, 12, dz=np.ones(12), land_frac=0.0)
    p.kbot[0, :4] = [-3, 0, 13, 1000]
    ref = O.solve_implicit_kbot(p.a, p.b, p.c, p.d, p.kbot)
    x = _solve((p.a, p.b, p.c, p.d), p.kbot)
    assert np.all(x[0, :4] == 0)
    assert rel_err(x, ref) <= TOL


def test_strided_interior_views_and_inplace():
    """Interior slices of halo arrays are solved without copies; x may alias d."""
    from veros_amd.utilities import solve_implicit_kbot
    rng = np.random.default_rng(8)
    nx, ny, nz = 30, 21, 24
    full = [rng.standard_normal((nx + 4, ny + 4, nz)) for _ in range(4)]
    full[0] *= 0.1
    full[2] *= 0.1
    full[1] = 2.0 + np.abs(full[1])
    kbot = rng.integers(0, nz + 1, size=(nx, ny)).astype(np.int32)
    inner = (slice(2, -2), slice(2, -2))
    ref = O.solve_implicit_kbot(*(f[inner] for f in full), kbot)
    t = [torch.from_numpy(f).cuda() for f in full]
    views = [f[inner] for f in t]
    assert not views[0].is_contiguous()
    kb = torch.from_numpy(kbot).cuda()
    x = solve_implicit_kbot(*views, kb)
    assert x.stride() == views[0].stride()
    assert rel_err(x.cpu().numpy(), ref) <= TOL
    d_before = t[3].clone()
    solve_implicit_kbot(*views, kb, out=views[3])
    assert rel_err(views[3].cpu().numpy(), ref) <= TOL
    halo = torch.ones_like(t[3], dtype=torch.bool)
    halo[2:-2, 2:-2] = False
    assert torch.equal(t[3][halo], d_before[halo])


def test_large_config2_residual_and_sampled_columns():
    """1440x720x80 (the mixed-precision workload): residual ~eps on the GPU and
    a random sample of columns against dgtsv."""
    from veros_amd.utilities import solve_implicit_kbot
    p = W.tdma_config2()
    a, b, c, d = _dev(p.a, p.b, p.c, p.d)
    (kb,) = _dev(p.kbot)
    water = torch.from_numpy(O.create_water_masks(p.kbot, 80)[1]).cuda()
    edge = torch.from_numpy(O.create_water_masks(p.kbot, 80)[2]).cuda()
    rng = np.random.default_rng(0)
    ii = rng.integers(0, 1440, 400)
    jj = rng.integers(0, 720, 400)
    sub = [v[ii, jj][:, None, :] for v in (p.a, p.b, p.c, p.d)]
    ref = O.solve_implicit_kbot(*sub, p.kbot[ii, jj][:, None])
    for prec, refine in (("fp64", 0), ("mixed", 1)):
        x = solve_implicit_kbot(a, b, c, d, kb, precision=prec, refine_steps=refine)
        aa = torch.where(edge, torch.zeros_like(a), a)
        r = b * x - d
        r[..., 1:] += aa[..., 1:] * x[..., :-1]
        r[..., :-1] += c[..., :-1] * x[..., 1:]
        res = (r[water].abs().max() / d.abs().max()).item()
        assert res <= 1e-13, (prec, res)
        assert torch.all(x[~water] == 0)
        xs = x.cpu().numpy()[ii, jj][:, None, :]
        assert rel_err(xs, ref) <= TOL, prec


def test_deterministic():
    p = W.tdma_config0()
    x1 = _solve((p.a, p.b, p.c, p.d), p.kbot)
    x2 = _solve((p.a, p.b, p.c, p.d), p.kbot)
    assert np.array_equal(x1, x2)


def test_errors():
    from veros_amd.utilities import solve_implicit_kbot
    a = torch.ones(4, 4, 130, dtype=torch.float64, device="cuda")
    kb = torch.ones(4, 4, dtype=torch.int32, device="cuda")
    with pytest.raises(NotImplementedError):
        solve_implicit_kbot(a, a, a, a, kb)
    with pytest.raises(ValueError):
        solve_implicit_kbot(a, a, a, a[:, :, :5], kb)
    f = torch.ones(4, 4, 8, dtype=torch.float32, device="cuda")
    with pytest.raises(TypeError):
        solve_implicit_kbot(f, f, f, f, kb)
