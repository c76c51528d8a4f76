"""The time-level layout adapter (veros_amd.layout): the (X, Y, nz, 3) view
of a levels-outermost buffer holds the same values as the reference's
levels-innermost array; layouts are detected from strides; the rotation of
the time indices follows veros/veros.py:322. With the GPU: every kernel that
reads or writes time levels gives the same outputs, value for value, in both
layouts."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from veros_amd import layout as LY


def test_views_and_detection():
    t = torch.randn(7, 6, 5, 3, dtype=torch.float64)
    o = LY.to_levels_outer(t)
    assert LY.levels_layout(t) == LY.LEVELS_INNER
    assert LY.levels_layout(o) == LY.LEVELS_OUTER
    assert torch.equal(o, t)
    assert o.stride() == (6 * 5, 5, 1, 7 * 6 * 5)
    # level l is one contiguous block of the buffer
    buf = o.permute(3, 0, 1, 2)
    assert buf.is_contiguous()
    for lev in range(3):
        assert torch.equal(buf[lev], t[..., lev])
    back = LY.to_levels_inner(o)
    assert back.is_contiguous() and torch.equal(back, t)
    # writes through veros-style indexing land in the buffer
    o[2:-2, 2:-2, :, 1] = 7.0
    assert torch.all(buf[1, 2:-2, 2:-2] == 7.0)


def test_rejects_other_layouts_and_mixed_calls():
    t = torch.randn(7, 6, 5, 3, dtype=torch.float64)
    with pytest.raises(ValueError):
        LY.levels_layout(t.transpose(0, 1))
    with pytest.raises(ValueError):
        LY.levels_layout(torch.zeros(7, 6, 5, dtype=torch.float64))
    with pytest.raises(ValueError):
        LY.common_layout(temp=t, salt=LY.to_levels_outer(t))
    assert LY.common_layout(temp=t, salt=None) == LY.LEVELS_INNER
    assert LY.common_layout() == LY.LEVELS_INNER


def test_set_state_layout_and_rotation():
    vs = SimpleNamespace(temp=torch.randn(6, 5, 4, 3, dtype=torch.float64), K_iso=torch.zeros(6, 5, 4),
                         tau=0, taup1=1, taum1=2)
    ref = vs.temp.clone()
    st = SimpleNamespace(variables=vs)
    LY.set_state_layout(st, LY.LEVELS_OUTER)
    assert LY.levels_layout(vs.temp) == LY.LEVELS_OUTER and torch.equal(vs.temp, ref)
    LY.set_state_layout(st, LY.LEVELS_INNER)
    assert vs.temp.is_contiguous() and torch.equal(vs.temp, ref)
    LY.rotate_time_levels(vs)
    assert (vs.tau, vs.taup1, vs.taum1) == (1, 2, 0)
    for _ in range(2):
        LY.rotate_time_levels(vs)
    assert (vs.tau, vs.taup1, vs.taum1) == (0, 1, 2)


def _full_state(seed, nx, ny, nz):
    from veros_amd import workloads as W
    rng = np.random.default_rng(seed)
    st = W.iso_state(rng, nx, ny, nz, dx_deg=1.0, dy_deg=1.0, y_origin=-40.0, dz_top=10.0, dz_bottom=200.0,
                     land_frac=0.3, random_aux=True, random_k_iso=True, stratified=False, tau=2, taup1=0)
    W.add_vertmix_fields(rng, st, random_aux=True)
    W.add_friction_fields(rng, st, random_aux=True)
    vs = st.variables
    vs.K_gm = 300.0 * (0.5 + rng.random(vs.K_iso.shape))
    vs.P_diss_skew = rng.standard_normal(vs.K_iso.shape)
    return st


TL = ("temp", "salt", "u", "v")
OUTS = ("temp", "salt", "Ai_ez", "Ai_nz", "Ai_bx", "Ai_by", "K_11", "K_22", "K_33", "dtemp_iso", "dsalt_iso",
        "P_diss_iso", "dtemp_vmix", "dsalt_vmix", "u", "v", "du_mix", "dv_mix", "K_diss_v")


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(40, 30, 17), (130, 90, 48)])
def test_every_call_equal_in_both_layouts(shape):
    """The one call (pre + T + S), the separate pre and T+S calls, vertical
    mixing and vertical friction: outputs identical value for value."""
    from veros_amd import friction, isoneutral, thermodynamics
    from veros_amd.state import to_device_state
    st = _full_state(sum(shape), *shape)
    res = {}
    for lv in ("inner", "outer"):
        a = to_device_state(st, levels=lv)
        b = to_device_state(st, levels=lv)
        for n in TL:
            assert LY.levels_layout(getattr(a.variables, n)) == (LY.LEVELS_INNER if lv == "inner"
                                                                 else LY.LEVELS_OUTER)
        isoneutral.isoneutral_pre_and_diffusion_temp_salt(a)
        thermodynamics.vertmix_tempsalt(a)
        friction.implicit_vert_friction(a)
        isoneutral.isoneutral_diffusion_pre(b)
        isoneutral.isoneutral_diffusion_temp_salt(b)
        isoneutral.isoneutral_skew_diffusion(b, b.variables.temp, True)
        torch.cuda.synchronize()
        res[lv] = ({k: getattr(a.variables, k).clone() for k in OUTS},
                   {k: getattr(b.variables, k).clone() for k in ("temp", "salt", "dtemp_iso", "P_diss_skew")})
    for r in range(2):
        for k, v in res["inner"][r].items():
            w = res["outer"][r][k]
            assert torch.equal(v, w), (r, k, (v - w).abs().max().item())
