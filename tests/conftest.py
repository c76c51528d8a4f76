import os
import sys
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
GOLDEN = REPO / "tests" / "golden"
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def pytest_collection_modifyitems(config, items):
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


def golden_files(prefix):
    return sorted(GOLDEN.glob(prefix + "*.npz"))


def load_golden(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def state_from_golden(rec):
    """Rebuild a (variables, settings) state from an isoneutral fixture."""
    from veros_amd.workloads import MixSettings

    vs = SimpleNamespace(**{k[3:]: np.array(v) for k, v in rec.items() if k.startswith("in_")})
    kw = {k[4:]: v.item() for k, v in rec.items() if k.startswith("set_") and k not in ("set_tau", "set_taup1")}
    vs.tau = int(rec["set_tau"])
    vs.taup1 = int(rec["set_taup1"])
    st = MixSettings(**kw)
    return SimpleNamespace(variables=vs, settings=st)


def rel_err(x, ref, mask=None):
    """max|x - ref| / max|ref| (over ``mask`` when given)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if mask is not None:
        x = x[mask]
        ref = ref[mask]
    scale = np.max(np.abs(ref)) if ref.size else 0.0
    diff = np.max(np.abs(x - ref)) if ref.size else 0.0
    if scale == 0.0:
        return float(diff)
    return float(diff / scale)


def column_state_from_golden(rec):
    """Rebuild the state of a vertmix_* / friction_* fixture (inputs in_*,
    settings set_*) with veros_amd.workloads.MixSettings."""
    from veros_amd.workloads import MixSettings

    vs = SimpleNamespace(**{k[3:]: np.array(v) for k, v in rec.items() if k.startswith("in_")})
    vs.tau = int(rec["set_tau"])
    vs.taup1 = int(rec["set_taup1"])
    kw = {k[4:]: v.item() for k, v in rec.items() if k.startswith("set_") and k not in ("set_tau", "set_taup1")}
    return SimpleNamespace(variables=vs, settings=MixSettings(**kw))


def turbulence_state_from_golden(rec):
    """The state of a tke_* / eke_* / idemix_* fixture: inputs in_*, the
    settings set_* and the time levels."""
    vs = SimpleNamespace(**{k[3:]: np.array(v) for k, v in rec.items() if k.startswith("in_")})
    vs.tau = int(rec["set_tau"])
    vs.taup1 = int(rec["set_taup1"])
    vs.taum1 = int(rec["set_taum1"])
    st = SimpleNamespace(**{k[4:]: v.item() for k, v in rec.items()
                            if k.startswith("set_") and k not in ("set_tau", "set_taup1", "set_taum1")})
    return SimpleNamespace(variables=vs, settings=st)
