"""CPU checks of the C ABI: the library loads, exports every symbol the header
declares, and rejects bad arguments before touching the GPU."""
import re
from pathlib import Path

import pytest

from veros_amd import _native

HEADER = Path(__file__).resolve().parent.parent / "include" / "vxmix.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vx_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_reports():
    L = _native.lib()
    assert _native.version().startswith("vxmix")
    assert _native.max_nz() >= 100
    assert L is _native.lib()


def test_every_declared_symbol_is_exported_and_bound():
    L = _native.lib()
    names = declared_symbols()
    assert len(names) >= 8
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_native.exported_symbols())


def test_tdma_argument_errors_without_gpu():
    L = _native.lib()
    args = dict(a=1, b=1, c=1, d=1, be=None, de=None, kb=1, kx=1, ky=1, x=1, nx=2, ny=2, nz=4, sx=8, sy=4)

    def call(**kw):
        v = {**args, **kw}
        return L.vx_tdma_f64(None, v["a"], v["b"], v["c"], v["d"], v["be"], v["de"], v["kb"], v["kx"], v["ky"],
                             v["x"], v["nx"], v["ny"], v["nz"], v["sx"], v["sy"])

    assert call(nz=0) == _native.VX_ERR_INVALID_ARGUMENT
    assert call(nz=_native.max_nz() + 1) == _native.VX_ERR_UNSUPPORTED
    assert call(nx=-1) == _native.VX_ERR_INVALID_ARGUMENT
    assert call(a=None) == _native.VX_ERR_INVALID_ARGUMENT
    assert call(sy=2) == _native.VX_ERR_INVALID_ARGUMENT      # columns would overlap
    assert "vx_tdma_f64" in L.vx_last_error().decode()
    assert call(nx=0) == _native.VX_OK                        # nothing to do: no launch
    with pytest.raises(ValueError):
        _native.check(call(nz=0), "t")
    with pytest.raises(NotImplementedError):
        _native.check(call(nz=999), "t")
    v = {**args}
    rc = L.vx_tdma_mixed(None, v["a"], v["b"], v["c"], v["d"], None, None, v["kb"], 1, 1, v["x"], 2, 2, 4, 8, 4, 3)
    assert rc == _native.VX_ERR_INVALID_ARGUMENT


def test_tdma_adjoint_argument_errors_without_gpu():
    L = _native.lib()

    def call(be=None, de=None, gbe=None, gde=None, gd=1, nz=4, nx=2):
        return L.vx_tdma_adjoint_f64(None, 1, 1, 1, be, de, 1, 1, 1, 1, 1, 1, 1, 1, gd, gbe, gde,
                                     nx, 2, nz, 8, 4)

    assert call(be=1) == _native.VX_ERR_INVALID_ARGUMENT            # b_edge without gb_edge
    assert call(gde=1) == _native.VX_ERR_INVALID_ARGUMENT           # gd_edge without d_edge
    assert call(gd=None) == _native.VX_ERR_INVALID_ARGUMENT
    assert call(nz=0) == _native.VX_ERR_INVALID_ARGUMENT
    assert call(nz=_native.max_nz() + 1) == _native.VX_ERR_UNSUPPORTED
    assert call(nx=0) == _native.VX_OK
    assert "vx_tdma_adjoint_f64" in L.vx_last_error().decode()


def test_iso_argument_errors_without_gpu():
    L = _native.lib()
    g = _native.VxGrid()
    p = _native.VxIsoParams()
    g.nx, g.ny, g.nz = 4, 4, 1
    assert L.vx_iso_pre_f64(None, g, p, 1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1) == _native.VX_ERR_INVALID_ARGUMENT
    g.nz = 8
    assert L.vx_iso_pre_f64(None, g, p, 1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1) == _native.VX_ERR_INVALID_ARGUMENT  # null grid arrays
    for name in ("kbot", "dxt", "dxu", "dyt", "dyu", "cost", "cosu", "dzt", "dzw"):
        setattr(g, name, 16)
    assert L.vx_iso_pre_f64(None, g, p, 1, 1, 5, 1, 1, 1, 1, 1, 1, 1, 1) == _native.VX_ERR_INVALID_ARGUMENT  # tau
    # three face-flux arrays per tracer + the wet-cell plan (a few ints per
    # column of the 6 x 6 step range, padded to whole strips)
    plan = L.vx_iso_workspace_bytes(g, 2) - 6 * 8 * 8 * 8 * 8
    assert plan >= (3 * 36 + 14) * 4 + 8 and plan % 8 == 0
    assert L.vx_iso_workspace_bytes(g, 1) == 3 * 8 * 8 * 8 * 8 + plan
    assert L.vx_iso_workspace_bytes(g, 3) == -1
    assert L.vx_iso_pre_diffusion_ts_workspace_bytes(g) == L.vx_iso_workspace_bytes(g, 2)
    assert L.vx_iso_pre_diffusion_ts_workspace_bytes(None) == -1
    rc = L.vx_iso_diffusion_f64(None, g, p, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, None, 1, 1)
    assert rc == _native.VX_ERR_INVALID_ARGUMENT  # P_diss without int_drhodX
    rc = L.vx_iso_diffusion_f64(None, g, p, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, None, None, None)
    assert rc == _native.VX_ERR_INVALID_ARGUMENT  # no workspace
    rc = L.vx_iso_diffusion_ts_f64(None, g, p, 1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1)
    assert rc == _native.VX_ERR_INVALID_ARGUMENT  # temp and salt alias
    rc = L.vx_iso_pre_diffusion_ts_f64(None, g, p, 1, 2, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, None, None, None, None)
    assert rc == _native.VX_ERR_INVALID_ARGUMENT  # no workspace
    rc = L.vx_iso_pre_diffusion_ts_f64(None, g, p, 1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, None, None, None, 1)
    assert rc == _native.VX_ERR_INVALID_ARGUMENT  # temp and salt alias
    rc = L.vx_iso_skew_diffusion_f64(None, g, p, 1, 0, 1, None, 1, 1, 1, 1, 1, 1, 1, None, None, 1)
    assert rc == _native.VX_ERR_INVALID_ARGUMENT  # null K_gm
    assert L.vx_iso_adjoint_workspace_bytes(g) == 4 * 8 * 8 * 8 * 8
    assert L.vx_iso_adjoint_workspace_bytes(None) == -1
    rc = L.vx_iso_diffusion_adjoint_f64(None, g, p, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, None)
    assert rc == _native.VX_ERR_INVALID_ARGUMENT  # no workspace
    rc = L.vx_iso_diffusion_adjoint_f64(None, g, p, 1, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1)
    assert rc == _native.VX_ERR_INVALID_ARGUMENT  # tau out of range


def test_vertmix_argument_errors_without_gpu():
    L = _native.lib()
    g = _native.VxGrid()
    g.nx, g.ny, g.nz = 4, 4, 8

    def call(grid=g, dt=3600.0, temp=1, salt=2, taup1=1, kap=1, cyc=0):
        return L.vx_vertmix_tempsalt_f64(None, grid, dt, temp, salt, taup1, kap, 1, 1, 3, 4, cyc)

    assert call() == _native.VX_ERR_INVALID_ARGUMENT                # null kbot/dzt/dzw
    g.kbot, g.dzt, g.dzw = 16, 16, 16
    assert call(grid=None) == _native.VX_ERR_INVALID_ARGUMENT
    assert call(taup1=3) == _native.VX_ERR_INVALID_ARGUMENT
    assert call(temp=2) == _native.VX_ERR_INVALID_ARGUMENT           # temp and salt alias
    assert call(kap=None) == _native.VX_ERR_INVALID_ARGUMENT
    assert call(dt=0.0) == _native.VX_ERR_INVALID_ARGUMENT
    g.nz = _native.max_nz() + 1
    assert call() == _native.VX_ERR_UNSUPPORTED
    g.nz, g.nx = 8, 3
    assert call(cyc=1) == _native.VX_ERR_UNSUPPORTED                 # cyclic halo copy needs nx >= 4
    assert "vx_vertmix_tempsalt_f64" in L.vx_last_error().decode()


def test_friction_argument_errors_without_gpu():
    L = _native.lib()
    g = _native.VxGrid()
    g.nx, g.ny, g.nz = 4, 4, 8

    def call(grid=g, dt=3600.0, u=1, v=2, tau=0, taup1=1, kap=3, du=4, dv=5, kd=6):
        return L.vx_implicit_vert_friction_f64(None, grid, dt, u, v, tau, taup1, kap, 7, 8, du, dv, kd)

    assert call() == _native.VX_ERR_INVALID_ARGUMENT                # null kbot/dzt/dzw/dxt/dxu
    g.kbot, g.dzt, g.dzw, g.dxt, g.dxu = 16, 16, 16, 16, 16
    assert call(grid=None) == _native.VX_ERR_INVALID_ARGUMENT
    assert call(tau=1) == _native.VX_ERR_INVALID_ARGUMENT            # tau == taup1
    assert call(taup1=3) == _native.VX_ERR_INVALID_ARGUMENT
    assert call(v=1) == _native.VX_ERR_INVALID_ARGUMENT              # u and v alias
    assert call(kd=4) == _native.VX_ERR_INVALID_ARGUMENT             # K_diss_v aliases du_mix
    assert call(kap=None) == _native.VX_ERR_INVALID_ARGUMENT
    assert call(dt=-1.0) == _native.VX_ERR_INVALID_ARGUMENT
    g.nz = _native.max_nz() + 1
    assert call() == _native.VX_ERR_UNSUPPORTED
    assert "vx_implicit_vert_friction_f64" in L.vx_last_error().decode()


def test_batch_plan_argument_errors_without_gpu():
    import ctypes
    L = _native.lib()
    h = ctypes.c_void_p()
    arr = (_native.VxTdmaSystem * 1)()
    assert L.vx_tdma_plan_create(arr, 0, 0, ctypes.byref(h)) == _native.VX_ERR_INVALID_ARGUMENT
    assert L.vx_tdma_plan_create(arr, 1, 5, ctypes.byref(h)) == _native.VX_ERR_INVALID_ARGUMENT
    assert L.vx_tdma_plan_create(arr, 1, 1, ctypes.byref(h)) == _native.VX_ERR_INVALID_ARGUMENT  # nz = 0
    arr[0].nz = 200
    arr[0].nx = arr[0].ny = 2
    assert L.vx_tdma_plan_create(arr, 1, 0, ctypes.byref(h)) == _native.VX_ERR_UNSUPPORTED
    arr[0].nz = 8
    assert L.vx_tdma_plan_create(arr, 1, 0, ctypes.byref(h)) == _native.VX_ERR_INVALID_ARGUMENT  # null arrays
    assert not h.value
    assert L.vx_tdma_plan_run(None, None) == _native.VX_ERR_INVALID_ARGUMENT
    L.vx_tdma_plan_destroy(None)
