"""The 1440x720x80 quarter-degree step (35% land, kbot U{5..80}) against the
oracle, output by output.

The whole grid runs through the GPU; the oracle runs on windows of the same
inputs (veros_amd.workloads.window_state: masks from the whole grid's kbot,
the window's outer two rows and columns as its halos), whose cells
[2:-2, 2:-2] depend on window cells only, so the oracle's values there are the
whole grid's. Bar: max|x - x_ref| / max|x_ref| <= 1e-12 over those cells.

P_diss_iso sums the temperature and salinity buoyancy-flux terms, which
isoneutral mixing makes cancel; its error is measured against the larger of
max|P_ref| and the largest single-tracer increment (the same scale as the
360x160x60 channel test)."""
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from conftest import rel_err
from oracle import isoneutral as OI
from oracle import tdma as OT
from veros_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-12
REPO = Path(__file__).resolve().parent.parent
PRE_KEYS = ("Ai_ez", "Ai_nz", "Ai_bx", "Ai_by", "K_11", "K_22", "K_33")
STEP_KEYS = ("temp", "salt", "dtemp_iso", "dsalt_iso")
# west edge with the cyclic halo, the interior, the north-east corner
WINDOWS = ((0, 100, 0, 84), (700, 800, 300, 384), (1344, 1444, 640, 724))


@pytest.fixture(scope="module")
def workload():
    if str(REPO) not in sys.path:
        sys.path.insert(0, str(REPO))
    import bench
    st, td = bench.build_workload(1440, 720, 80)
    torch.cuda.synchronize()
    yield st, td
    del st, td
    torch.cuda.empty_cache()


def _oracle_step(w):
    P0 = np.array(w.variables.P_diss_iso, copy=True)
    pre = OI.isoneutral_diffusion_pre(w)
    for k, v in pre._asdict().items():
        setattr(w.variables, k, v)
    OI.isoneutral_diffusion(w, "temp", True)
    Pt = np.array(w.variables.P_diss_iso, copy=True)
    OI.isoneutral_diffusion(w, "salt", False)
    return P0, Pt - P0, w.variables.P_diss_iso - Pt


def test_one_call_windows_against_oracle(workload):
    from veros_amd import isoneutral
    st, _ = workload
    names = set(PRE_KEYS + STEP_KEYS) | {"P_diss_iso", "K_iso", "int_drhodT", "int_drhodS", "dxt", "dxu", "dyt",
                                         "dyu", "cost", "cosu", "dzt", "dzw", "zt"}
    before = [W.window_state(st, *w, names=names) for w in WINDOWS]
    isoneutral.isoneutral_pre_and_diffusion_temp_salt(st)
    torch.cuda.synchronize()
    worst = {}
    for win, hs in zip(WINDOWS, before):
        got = W.window_state(st, *win, masks=False, names=names).variables
        _, dPt, dPs = _oracle_step(hs)
        ref = hs.variables
        tau1 = ref.taup1
        for k in PRE_KEYS + STEP_KEYS:
            g, r = getattr(got, k), getattr(ref, k)
            if k in ("temp", "salt"):
                g, r = g[..., tau1], r[..., tau1]
            e = rel_err(g[2:-2, 2:-2], r[2:-2, 2:-2])
            worst[k] = max(worst.get(k, 0.0), e)
            assert e <= TOL, (win, k, e)
        sl = np.s_[2:-2, 2:-2]
        Pr = ref.P_diss_iso[sl]
        scale = max(np.abs(Pr).max(), np.abs(dPt[sl]).max(), np.abs(dPs[sl]).max())
        e = np.abs(got.P_diss_iso[sl] - Pr).max() / scale
        worst["P_diss_iso"] = max(worst.get("P_diss_iso", 0.0), e)
        assert e <= TOL, (win, "P_diss_iso", e)
    print("config-2 window errors:", {k: f"{v:.2e}" for k, v in worst.items()})


@pytest.mark.parametrize("precision,refine", [("fp64", 0), ("mixed", 1), ("mixed", 2)])
def test_tdma_sampled_rows_against_oracle(workload, precision, refine):
    """The step's 1440x720x80 TDMA (kbot U{5..80}, 35% land) on every column,
    checked on 12 whole rows of columns (8640 columns) against dgtsv."""
    from veros_amd.utilities import solve_implicit_kbot
    _, td = workload
    x = solve_implicit_kbot(td["a"], td["b"], td["c"], td["d"], td["kbot"], precision=precision,
                            refine_steps=refine)
    torch.cuda.synchronize()
    rows = np.random.default_rng(5).choice(1440, size=12, replace=False)
    rows[0], rows[1] = 0, 1439
    idx = torch.from_numpy(rows).cuda()
    sub = {k: td[k].index_select(0, idx).cpu().numpy() for k in ("a", "b", "c", "d", "kbot")}
    ref = OT.solve_implicit_kbot(sub["a"], sub["b"], sub["c"], sub["d"], sub["kbot"])
    got = x.index_select(0, idx).cpu().numpy()
    assert rel_err(got, ref) <= TOL
    assert np.all(got[sub["kbot"] == 0] == 0)
