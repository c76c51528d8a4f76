"""GPU parity of vertmix_tempsalt (vx_vertmix_tempsalt_f64) against the oracle
(bar: 1e-12 relative per output, max|x - x_ref| / max|x_ref| over the whole
array, halos included), and size-independent properties on the 1440x720x80
grid.

The tendencies dtemp_vmix = (new - old) / dt difference two tracer values of
order 10-35 whose change over a step can be 1e-4 of that: two tridiagonal
solvers that agree to a few ulp of the tracer (the oracle's LAPACK dgtsv
eliminates in another order than the GPU's Thomas sweep) differ there by a few
ulp of the tracer / dt. Their error is therefore measured in units of the
tracer: max|dv - dv_ref| * dt / max|tracer_ref[..., taup1]| <= 1e-12."""
import numpy as np
import pytest
import torch

from types import SimpleNamespace

from conftest import column_state_from_golden, golden_files, load_golden, rel_err
from oracle import vertmix as V
from veros_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-12
KEYS = ("temp", "salt", "dtemp_vmix", "dsalt_vmix")


def _check(out, ref, st):
    t1, dt = st.variables.taup1, st.settings.dt_tracer
    for k in ("temp", "salt"):
        assert rel_err(_np(getattr(out, k)), getattr(ref, k)) <= TOL, k
    for k, tr in (("dtemp_vmix", "temp"), ("dsalt_vmix", "salt")):
        scale = np.abs(getattr(ref, tr)[..., t1]).max()
        err = np.abs(_np(getattr(out, k)) - getattr(ref, k)).max() * dt / scale
        assert err <= TOL, (k, err)


def _state(seed, nx, ny, nz, cyclic=True, land=0.3, kmin=1):
    rng = np.random.default_rng(seed)
    st = W.iso_state(rng, nx, ny, nz, dx_deg=1.0, dy_deg=1.0, y_origin=-40.0, dz_top=10.0, dz_bottom=250.0,
                     land_frac=land, cyclic=cyclic, kmin=kmin, tau=0, taup1=1, random_aux=True)
    return W.add_vertmix_fields(rng, st, random_aux=True)


def _dev(st, levels="inner"):
    from veros_amd.state import to_device_state
    return to_device_state(st, levels=levels)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("shape,cyclic,land", [((9, 7, 6), True, 0.3), ((14, 11, 33), False, 0.3),
                                               ((12, 9, 2), True, 0.0), ((8, 10, 80), True, 0.35),
                                               ((6, 5, 128), False, 0.2), ((7, 6, 17), True, 1.0),
                                               ((10, 12, 100), True, 0.3), ((11, 9, 15), False, 0.25)])
@pytest.mark.parametrize("levels", ["inner", "outer"])
def test_vertmix_random_against_oracle(shape, cyclic, land, levels):
    """levels outermost with nz even runs the LDS-staged kernel; the others
    the kernel whose lanes load their own levels."""
    st = _state(sum(shape), *shape, cyclic=cyclic, land=land)
    ref = V.vertmix_tempsalt(st)
    ds = _dev(st, levels)
    from veros_amd import thermodynamics
    out = thermodynamics.vertmix_tempsalt(ds)
    _check(out, ref, st)
    for k in KEYS:
        assert getattr(out, k) is getattr(ds.variables, k)  # in place


@pytest.mark.parametrize("path", golden_files("vertmix_"), ids=lambda p: p.stem)
@pytest.mark.parametrize("levels", ["inner", "outer"])
def test_vertmix_golden(path, levels):
    """Against the reference's own outputs (veros.core.thermodynamics.vertmix_tempsalt,
    written by tests/golden/make_golden.py), with the bar of _check."""
    rec = load_golden(path)
    st = column_state_from_golden(rec)
    ref = SimpleNamespace(**{k[4:]: v for k, v in rec.items() if k.startswith("out_")})
    from veros_amd import thermodynamics
    out = thermodynamics.vertmix_tempsalt(_dev(st, levels))
    _check(out, ref, st)


def test_vertmix_channel_360x160x60_against_oracle():
    st = W.add_vertmix_fields(np.random.default_rng(7), W.iso_config1(), random_aux=True)
    ref = V.vertmix_tempsalt(st)
    from veros_amd import thermodynamics
    out = thermodynamics.vertmix_tempsalt(_dev(st))
    _check(out, ref, st)


@pytest.mark.parametrize("levels", ["inner", "outer"])
def test_vertmix_full_grid_properties(levels):
    """1440x720x80: the column content sum_k dzt * tracer over the wet cells
    changes by exactly dt * forcing; a vertically uniform tracer without
    forcing is a fixed point; the tendencies vanish off the wet cells."""
    from veros_amd import thermodynamics
    st = W.vertmix_config2()
    ds = _dev(st, levels)
    vs = ds.variables
    t1, dt = vs.taup1, st.settings.dt_tracer
    nz = vs.kappaH.shape[2]
    kb = vs.kbot.long()
    k = torch.arange(nz, device=kb.device)
    wet = (kb[..., None] >= 1) & (k[None, None, :] >= kb[..., None] - 1)
    wet[:2] = False
    wet[-2:] = False
    wet[:, :2] = False
    wet[:, -2:] = False
    dzt = vs.dzt[None, None, :]
    before = {n: (torch.where(wet, getattr(vs, n)[..., t1], 0.0) * dzt).sum(dim=2) for n in ("temp", "salt")}
    thermodynamics.vertmix_tempsalt(ds)
    for n, f in (("temp", vs.forc_temp_surface), ("salt", vs.forc_salt_surface)):
        after = (torch.where(wet, getattr(vs, n)[..., t1], 0.0) * dzt).sum(dim=2)
        want = before[n] + torch.where(wet[..., -1], dt * f, 0.0)
        err = (after - want).abs().max().item() / before[n].abs().max().item()
        assert err <= 1e-13, (n, err)
    for d in (vs.dtemp_vmix, vs.dsalt_vmix):
        assert torch.count_nonzero(d[~wet]).item() == 0
    vs.forc_temp_surface.zero_()
    vs.temp[..., t1] = 12.5
    thermodynamics.vertmix_tempsalt(ds)
    got = vs.temp[2:-2, 2:-2, :, t1]
    assert (got - 12.5).abs().max().item() <= 12.5 * 4e-15


def test_vertmix_argument_errors():
    from veros_amd import thermodynamics
    st = _state(3, 3, 5, 4, cyclic=True)
    ds = _dev(st)
    with pytest.raises(NotImplementedError):
        thermodynamics.vertmix_tempsalt(ds)  # cyclic x needs nx >= 4
    ds.settings.enable_cyclic_x = False
    thermodynamics.vertmix_tempsalt(ds)
    ds.variables.kappaH = ds.variables.kappaH.float()
    with pytest.raises(TypeError):
        thermodynamics.vertmix_tempsalt(ds)
