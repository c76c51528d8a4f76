"""The vertical-mixing and vertical-friction oracles against the reference's
own outputs (tests/golden/vertmix_*.npz, friction_*.npz, written by
tests/golden/make_golden.py from veros.core.thermodynamics.vertmix_tempsalt,
veros/core/thermodynamics.py:248-300, and
veros.core.friction.implicit_vert_friction, veros/core/friction.py:90-205).
Both sides solve with LAPACK dgtsv in the same order, so the bar is 1e-14
relative per output (max|x - x_ref| / max|x_ref| over the whole array)."""
import numpy as np
import pytest

from conftest import column_state_from_golden, golden_files, load_golden, rel_err
from oracle import friction as F
from oracle import vertmix as V

TOL = 1e-14


@pytest.mark.parametrize("path", golden_files("vertmix_"), ids=lambda p: p.stem)
def test_vertmix_oracle_matches_reference(path):
    rec = load_golden(path)
    st = column_state_from_golden(rec)
    out = V.vertmix_tempsalt(st)
    for k in ("temp", "salt", "dtemp_vmix", "dsalt_vmix"):
        assert rel_err(getattr(out, k), rec["out_" + k]) <= TOL, k


@pytest.mark.parametrize("path", golden_files("friction_"), ids=lambda p: p.stem)
def test_friction_oracle_matches_reference(path):
    rec = load_golden(path)
    st = column_state_from_golden(rec)
    out = F.implicit_vert_friction(st)
    for k in ("u", "v", "du_mix", "dv_mix", "K_diss_v"):
        assert rel_err(getattr(out, k), rec["out_" + k]) <= TOL, k


def test_fixtures_cover_written_ranges():
    """The fixtures' in/out fields are noisy, so untouched ranges are visible:
    at least one cell of every output changes, and every output except the
    vertical-mixing tendencies (written over the whole array,
    veros/core/thermodynamics.py:286-287) keeps at least one input value."""
    for path in golden_files("vertmix_") + golden_files("friction_"):
        rec = load_golden(path)
        for k in [k[4:] for k in rec if k.startswith("out_")]:
            a, b = rec["in_" + k], rec["out_" + k]
            assert np.any(a != b), (path.stem, k)
            if k not in ("dtemp_vmix", "dsalt_vmix"):
                assert np.any(a == b), (path.stem, k)
