"""Write golden values of hydrointel's land-use table, climate and socioeconomic pathways and
agreement criteria.

Run once, in a container that holds the reference checkout, with
    PYTHONPATH=/root/reference python tests/golden/make_tables_golden.py
It calls the reference's hydrointel.domain.landuse.table_array (landuse.py:50-52) for every class
parameter, hydrointel.forcing.scenarios.rain_factor / slr_m / ssp_state (scenarios.py:48-78) for
every pathway and horizon year, and hydrointel.criteria.agreement (criteria.py:52-100) on seeded
pairs of synthetic runs, and stores the inputs and what the reference returned:

  tables_inputs.npz   the agreement cases' depth series, land masks, discharges and reach maps
  tables_ref.json     the table columns, pathway values and agreement results

Nothing from the reference is stored.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

import hydrointel.criteria as RCRIT                       # noqa: E402
import hydrointel.domain.landuse as RLU                   # noqa: E402
import hydrointel.forcing.scenarios as RSC                # noqa: E402

ATTRS = ("manning", "f0_mmh", "fc_mmh", "k_per_h", "storage_m", "exposure")
YEARS = list(range(2010, 2101, 5)) + [2021, 2049, 2051, 2099]


def agreement_cases():
    """Seeded pairs of depth series (B a perturbed A), with the land mask, spacing, dry depth and
    channel discharge series of each case."""
    rng = np.random.default_rng(4242)
    cases = []
    for k, (scale, dq, n_reach) in enumerate([(0.0005, 0.005, 3), (0.004, 0.03, 4), (0.0002, 0.0, 2),
                                              (0.02, 0.1, 1), (0.0, 0.0, 3)]):
        nt, ny, nx = 5, 12, 15
        base = np.clip(rng.random((ny, nx)) * 0.9 - 0.2, 0.0, None)
        growth = np.linspace(0.2, 1.0, nt)[:, None, None]
        h_a = base[None] * growth
        h_b = np.clip(h_a + scale * rng.standard_normal(h_a.shape), 0.0, None)
        if k == 2:
            h_a[:, :, :] = np.minimum(h_a, 0.05)          # no deep cell and no flooded area in A
        land = np.ones((ny, nx), bool)
        land[:, -3:] = False
        n_nodes = 4 * n_reach
        q_a = 5.0 + 20.0 * rng.random((nt, n_nodes))
        q_b = q_a * (1.0 + dq * rng.standard_normal(q_a.shape))
        reach_of = np.repeat(np.arange(n_reach), 4)
        cases.append(dict(h_a=h_a, h_b=h_b, land=land, q_a=q_a, q_b=q_b, reach_of=reach_of, dx=20.0 + 10 * k,
                          h_dry=1e-3))
    return cases


def main() -> None:
    out = {"landuse": {a: RLU.table_array(a).astype(float).tolist() for a in ATTRS},
           "rain_factor": {r: RSC.rain_factor(r) for r in sorted(RSC.RCP_TABLE)},
           "slr_m": {r: {str(y): RSC.slr_m(r, y) for y in YEARS} for r in sorted(RSC.RCP_TABLE)},
           "ssp": {s: [RSC.ssp_state(s).urban_expansion_rate, RSC.ssp_state(s).population_factor,
                       RSC.ssp_state(s).drainage_investment_factor] for s in sorted(RSC.SSP_TABLE)},
           "agreement": []}
    arrays = {}
    for i, c in enumerate(agreement_cases()):
        for k, v in c.items():
            if isinstance(v, np.ndarray):
                arrays[f"c{i}_{k}"] = v
        arrays[f"c{i}_scalars"] = np.array([c["dx"], c["h_dry"]])
        r = RCRIT.agreement(c["h_a"], c["h_b"], c["land"], c["dx"], c["h_dry"], c["q_a"], c["q_b"], c["reach_of"])
        out["agreement"].append(r)
    np.savez_compressed(HERE / "tables_inputs.npz", **arrays)
    # json keys must be strings (reach indices); infinity is written as the token Infinity
    (HERE / "tables_ref.json").write_text(json.dumps(out, indent=1, sort_keys=True, default=float))
    assert not any(isinstance(v, float) and math.isnan(v) for v in out["slr_m"]["RCP2.6"].values())
    print("wrote tables_inputs.npz, tables_ref.json")


if __name__ == "__main__":
    main()
