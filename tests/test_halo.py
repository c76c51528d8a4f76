"""enforce_boundaries (veros/core/utilities.py:7-21) and
isoneutral_diag_streamfunction_kernel (veros/core/isoneutral/isoneutral.py:232-255):
oracle and GPU against the reference's outputs (tests/golden/boundaries_*.npz,
streamfunction_*.npz), bit for bit for the copy and 1e-14 relative for the
diagnostic; the GPU copy also on time-level slices of both layouts."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from conftest import golden_files, load_golden, rel_err
from oracle import isoneutral as OI
from oracle import vertmix as OV


@pytest.mark.parametrize("path", golden_files("boundaries_"), ids=lambda p: p.stem)
def test_oracle_boundaries(path):
    g = load_golden(path)
    assert np.array_equal(OV.enforce_boundaries(g["in_arr"], bool(g["set_cyclic"])), g["out_arr"])


@pytest.mark.parametrize("path", golden_files("streamfunction_"), ids=lambda p: p.stem)
def test_oracle_streamfunction(path):
    g = load_golden(path)
    vs = SimpleNamespace(**{k[3:]: v for k, v in g.items() if k.startswith("in_")})
    out = OI.isoneutral_diag_streamfunction_kernel(SimpleNamespace(variables=vs))
    for k in ("B1_gm", "B2_gm"):
        assert rel_err(getattr(out, k), g["out_" + k]) <= 1e-14, k


@pytest.mark.gpu
@pytest.mark.parametrize("path", golden_files("boundaries_"), ids=lambda p: p.stem)
def test_gpu_boundaries_golden(path):
    from veros_amd.utilities import enforce_boundaries
    g = load_golden(path)
    a = torch.from_numpy(g["in_arr"]).cuda()
    enforce_boundaries(a, bool(g["set_cyclic"]))
    assert np.array_equal(a.cpu().numpy(), g["out_arr"])


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(8, 5), (14, 11, 7), (37, 20, 33), (12, 6, 4, 3)])
@pytest.mark.parametrize("levels", ["inner", "outer"])
def test_gpu_boundaries_levels(shape, levels):
    from veros_amd.layout import to_levels_outer
    from veros_amd.utilities import enforce_boundaries
    rng = np.random.default_rng(sum(shape))
    host = rng.standard_normal(shape + (3,))
    t = torch.from_numpy(host).cuda()
    if levels == "outer" and len(shape) == 3:
        t = to_levels_outer(t)
    for lev in range(3):
        enforce_boundaries(t[..., lev], True)
        host[..., lev] = OV.enforce_boundaries(host[..., lev], True)
    assert np.array_equal(t.cpu().numpy(), host)
    if t.dim() >= 4:  # trailing axes that are not one evenly strided run
        with pytest.raises(ValueError):
            enforce_boundaries(t[:, :, :-1, ..., 0], True)


@pytest.mark.gpu
@pytest.mark.parametrize("path", golden_files("streamfunction_"), ids=lambda p: p.stem)
def test_gpu_streamfunction_golden(path):
    from veros_amd import isoneutral
    g = load_golden(path)
    K = g["in_K_gm"]
    X, Y, nz = K.shape
    vs = SimpleNamespace(**{k[3:]: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in g.items()
                            if k.startswith("in_")})
    vs.kbot = torch.zeros((X, Y), dtype=torch.int32, device="cuda")
    for n, m in (("dxt", X), ("dxu", X), ("dyt", Y), ("dyu", Y), ("cost", Y), ("cosu", Y), ("dzt", nz), ("dzw", nz)):
        setattr(vs, n, torch.ones(m, dtype=torch.float64, device="cuda"))
    out = isoneutral.isoneutral_diag_streamfunction_kernel(SimpleNamespace(variables=vs))
    for k in ("B1_gm", "B2_gm"):
        assert rel_err(getattr(out, k).cpu().numpy(), g["out_" + k]) <= 1e-14, k
