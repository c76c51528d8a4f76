"""GPU parity of the batched TDMA plan: every system equals its single-system
solve bit for bit and the oracle to 1e-12; tuning and hipGraph replay keep it."""
import numpy as np
import pytest
import torch

from conftest import rel_err
from oracle import tdma as O

pytestmark = pytest.mark.gpu


def _systems(seed=5):
    from veros_amd.batched import random_systems
    # small grids spanning every depth bucket, one with edge arrays and one strided view
    sys = random_systems(7, seed=seed, nxy=(5, 41), nz_range=(2, 128))
    rng = np.random.default_rng(seed)
    s0 = sys[0]
    s0["b_edge"] = s0["b"] + 0.1 * torch.rand_like(s0["b"])
    s0["d_edge"] = torch.randn_like(s0["d"])
    full = [torch.randn(30, 20, 33, dtype=torch.float64, device="cuda") for _ in range(5)]
    full[0] *= 0.1
    full[2] *= 0.1
    full[1] = 2.0 + full[1].abs()
    view = [f[2:-2, 3:-1] for f in full]
    kb = torch.from_numpy(rng.integers(0, 34, size=(26, 16)).astype(np.int32)).cuda()
    sys.append(dict(a=view[0], b=view[1], c=view[2], d=view[3], kbot=kb, x=view[4]))
    return sys


def _single(s, precision):
    from veros_amd.utilities import solve_implicit_kbot
    return solve_implicit_kbot(s["a"], s["b"], s["c"], s["d"], s["kbot"], b_edge=s.get("b_edge"),
                               d_edge=s.get("d_edge"), precision=precision, refine_steps=1)


def _oracle(s):
    f = lambda t: None if t is None else t.cpu().numpy()
    return O.solve_implicit_kbot(f(s["a"]), f(s["b"]), f(s["c"]), f(s["d"]), f(s["kbot"]),
                                 b_edge=f(s.get("b_edge")), d_edge=f(s.get("d_edge")))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_plan_matches_single_and_oracle(precision):
    from veros_amd.batched import TdmaBatch
    sys = _systems()
    plan = TdmaBatch(sys, precision)
    info = plan.info()
    assert info["total_columns"] == sum(s["a"].shape[0] * s["a"].shape[1] for s in sys)
    assert len(info["buckets"]) >= 5
    plan.run()
    torch.cuda.synchronize()
    for q, s in enumerate(sys):
        single = _single(s, precision)
        assert torch.equal(s["x"], single), q
        assert rel_err(s["x"].cpu().numpy(), _oracle(s)) <= 1e-12, q
    plan.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tune_and_graph_replays(precision):
    from veros_amd.batched import TdmaBatch
    sys = _systems(seed=9)
    plan = TdmaBatch(sys, precision)
    info = plan.tune()
    for b in info["buckets"]:
        assert b["block"] in (64, 128, 256) and b["blocks_per_cu"] in (0, 1, 2, 4, 8)
    for rep in range(3):
        for s in sys:                       # eager work between replays
            s["d"].mul_(1.5).add_(0.25)
            s["x"].zero_()
        torch.cuda.synchronize()
        plan.graph_launch()
        torch.cuda.synchronize()
        for q, s in enumerate(sys):
            assert torch.equal(s["x"], _single(s, precision)), (rep, q)
    plan.close()


def test_plan_rejects_bad_systems():
    from veros_amd.batched import TdmaBatch
    sys = _systems()
    bad = dict(sys[1])
    bad["d"] = bad["d"].float()
    with pytest.raises(TypeError):
        TdmaBatch([sys[0], bad])
    with pytest.raises(ValueError):
        TdmaBatch(sys, "bf16")
    with pytest.raises(TypeError):
        TdmaBatch([sys[0], bad], "mixed")


# ---------------------------------------------------------------------------
# The 64-grid batched workload at its real sizes: nx, ny ~ U{48..720},
# nz ~ U{15..100}, kbot ~ U{1..nz} with 30% land (about 9.4 M columns and
# 540 M cells). Every system is checked on 48 sampled columns per system
# (dry columns included) against the oracle's dgtsv, and equals its own
# single-grid solve value for value.
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def batch64():
    from veros_amd.batched import random_systems
    sys64 = random_systems(64)
    torch.cuda.synchronize()
    rng = np.random.default_rng(64)
    picks = []
    for s in sys64:
        nx, ny, _ = s["a"].shape
        ii = torch.from_numpy(rng.integers(0, nx, size=48)).cuda()
        jj = torch.from_numpy(rng.integers(0, ny, size=48)).cuda()
        cols = {k: s[k][ii, jj][None].cpu().numpy() for k in ("a", "b", "c", "d")}
        kb = s["kbot"][ii, jj][None].cpu().numpy()
        picks.append((ii, jj, O.solve_implicit_kbot(cols["a"], cols["b"], cols["c"], cols["d"], kb)))
    yield sys64, picks
    del sys64
    torch.cuda.empty_cache()


def _check_batch64(sys64, picks, precision, what):
    torch.cuda.synchronize()
    for q, (s, (ii, jj, ref)) in enumerate(zip(sys64, picks)):
        got = s["x"][ii, jj][None].cpu().numpy()
        assert rel_err(got, ref) <= 1e-12, (what, q, s["a"].shape)
    for q in (0, 17, 63):                 # value for value against the single-grid solve
        assert torch.equal(sys64[q]["x"], _single(sys64[q], precision)), (what, q)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch64_run_tune_and_graph(batch64, precision):
    from veros_amd.batched import TdmaBatch
    sys64, picks = batch64
    cols = sum(s["a"].shape[0] * s["a"].shape[1] for s in sys64)
    assert cols > 5_000_000
    plan = TdmaBatch(sys64, precision)
    assert plan.info()["total_columns"] == cols
    for s in sys64:
        s["x"].fill_(float("nan"))
    plan.run()
    _check_batch64(sys64, picks, precision, "run")
    plan.tune()
    for s in sys64:
        s["x"].fill_(float("nan"))
    plan.run()
    _check_batch64(sys64, picks, precision, "tuned run")
    for rep in range(2):
        for s in sys64:
            s["x"].fill_(float("nan"))
        torch.cuda.synchronize()
        plan.graph_launch()
        _check_batch64(sys64, picks, precision, f"graph replay {rep}")
    plan.close()


def test_tune_rejects_in_place_plan():
    """A plan whose x aliases d solves correctly, but tuning (many repeated
    solves) would overwrite d, so tune refuses it."""
    from veros_amd.batched import TdmaBatch
    sys = _systems(seed=13)[:3]
    ref = [_single(s, "fp64") for s in sys]
    for s in sys:
        s["x"] = s["d"]
    plan = TdmaBatch(sys, "fp64")
    plan.run()
    torch.cuda.synchronize()
    for s, r in zip(sys, ref):
        assert torch.equal(s["x"], r)
    with pytest.raises(ValueError):
        plan.tune()
    plan.close()
