"""Land-use table, climate and socioeconomic pathways and agreement criteria against the values
the reference's own functions returned (tests/golden/make_tables_golden.py): the eight classes'
parameters, the rain factor and sea-level rise of every RCP at every horizon year from 2010 to
2100, the SSP modifiers, and every metric, limit and pass/fail result of ``agreement`` on five seeded pairs
of runs (passing, failing, no flooded area in the first run, no deep cell).  Equal to the last bit
except the agreement metrics, which are sums and percentiles over arrays and are held to 1e-12
relative."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from his_amd import criteria as C
from his_amd import scenarios as S
from his_amd.domain import landuse as LU

GOLDEN = Path(__file__).resolve().parent / "golden"
REF = json.loads((GOLDEN / "tables_ref.json").read_text())


def test_landuse_table_matches():
    for attr, col in REF["landuse"].items():
        got = LU.lookup(np.arange(LU.N_CLASSES), attr)
        assert got.tolist() == col, attr


def test_pathways_match():
    for rcp, v in REF["rain_factor"].items():
        assert S.rain_factor(rcp) == v, rcp
    for rcp, years in REF["slr_m"].items():
        for y, v in years.items():
            assert S.sea_level_rise(rcp, int(y)) == v and S.slr_m(rcp, int(y)) == v, (rcp, y)
    for ssp, (urban, pop, drain) in REF["ssp"].items():
        assert S.SSP_TABLE[ssp][0] == urban and S.population_factor(ssp) == pop and S.drainage_factor(ssp) == drain
        st = S.ssp_state(ssp)
        assert (st.ssp, st.urban_expansion_rate, st.population_factor, st.drainage_investment_factor) == \
            (ssp, urban, pop, drain)


def _close(got, want, where):
    if isinstance(want, dict):
        assert isinstance(got, dict), where
        assert {str(k) for k in got} == set(want), f"{where}: {sorted(map(str, got))} vs {sorted(want)}"
        g = {str(k): v for k, v in got.items()}
        for k in want:
            _close(g[k], want[k], f"{where}.{k}")
    elif isinstance(want, bool):
        assert bool(got) is want, where
    elif isinstance(want, (int, float)):
        w, x = float(want), float(got)
        if math.isinf(w):
            assert math.isinf(x) and (x > 0) == (w > 0), where
        else:
            assert abs(x - w) <= 1e-12 * max(abs(w), 1e-30), f"{where}: {x!r} vs {w!r}"
    else:
        assert got == want, where


@pytest.mark.parametrize("case", range(5))
def test_agreement_matches(case):
    d = np.load(GOLDEN / "tables_inputs.npz")
    a = {k[len(f"c{case}_"):]: d[k] for k in d.files if k.startswith(f"c{case}_")}
    dx, h_dry = a["scalars"]
    got = C.agreement(a["h_a"], a["h_b"], a["land"], float(dx), float(h_dry), a["q_a"], a["q_b"], a["reach_of"])
    _close(got, REF["agreement"][case], f"case {case}")
