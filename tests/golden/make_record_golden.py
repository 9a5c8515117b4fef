"""Write a golden simulation record with the reference's own record writer.

Run once, in a container that holds the reference checkout, with
    PYTHONPATH=/root/reference python tests/golden/make_record_golden.py
It feeds hydrointel.data.generate._write_sim (generate.py:355-397) small synthetic inputs -- a
flood result of 4 output times on a 10 x 14 grid with 9 channel cells, a land-use raster, the
storm's fields, a hyetograph starting after a half-hour spin-up -- and stores what it wrote:

  record_inputs.npz     the synthetic inputs (arrays and the scalars that define the case)
  record_ref.pt         the record the reference wrote (tensors, numbers, strings and
                        containers of them; loads with torch.load(weights_only=True))
  record_ref_index.json the index entry the reference wrote for it

The tide and the inflow are this repository's series objects, called through plain functions, so
both writers sample the same functions; the rain is the reference's Hyetograph, so the rain
resampling is the reference's.  Nothing from the reference is stored.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from his_amd.forcing import PiecewiseLinear, sine_tide          # noqa: E402

import hydrointel.data.generate as RG                            # noqa: E402
from hydrointel.forcing.hyetograph import Hyetograph             # noqa: E402
from hydrointel.solver.coupling import RunLog as RefRunLog       # noqa: E402

NT, NY, NX, NC, NR, F = 4, 10, 14, 9, 3, 2
SIM_ID = 7


def inputs() -> dict:
    rng = np.random.default_rng(20261204)
    h = np.clip(rng.random((NT, NY, NX)) - 0.35, 0.0, None).astype(np.float32)
    h[0] = 0.0
    return dict(
        times=np.arange(NT) * 1200.0,
        h=h, u=rng.standard_normal((NT, NY, NX)).astype(np.float32), v=rng.standard_normal((NT, NY, NX)).astype(np.float32),
        eta1=(1.0 + rng.random((NT, NC))).astype(np.float64), q1=(10.0 * rng.random((NT, NC))).astype(np.float64),
        landuse=rng.integers(0, 8, (NY, NX)).astype(np.int64),
        storage=0.002 + 0.008 * rng.random((NY, NX)), kappa=0.7 + 0.6 * rng.random((NY, NX)),
        manning=0.02 + 0.06 * rng.random((NY, NX)), gamma=0.8 + 0.4 * rng.random(NR),
        sea=(np.arange(NX)[None, :] >= NX - 3).repeat(NY, axis=0),
        rain_mmh=np.array([5.0, 20.0, 60.0, 35.0, 10.0, 0.0]), rain_step_s=np.array(600.0),
        storm_start=np.array(1800.0), t_end=np.array(7200.0),
        tide=np.array([0.4, 44712.0, 0.1]), inflow=np.array([[5.0, 40.0, 12.0, 6.0]]), inflow_step_s=np.array(1800.0),
        ledger=np.array([[10.0 * t + k + 0.25 for k in range(9)] for t in range(NT)]),
        series=np.array([[1200.0 * t, 100.0 + t, 1e-7 * (t + 1)] for t in range(NT)]),
        scalars=np.array([1.5, 3.0, 2.0, 20.0, 1.1, 0.22, 0.93, 4.2e-7, 123.4, 45.6]),
        run_log=np.array([812.0, 0.75, 0.21, 1.9, 2.0, 3.0, 55.5, 44.4]),
    )


META = {"return_period_yr": 25, "rcp": "RCP4.5", "ssp": "SSP2", "horizon_year": 2060, "tide_on": True,
        "storm_tide_offset_h": 1.5, "site": "synthetic"}
LEDGER_KEYS = ["v2d", "v1d", "storage", "rain", "infiltration", "bnd2d_in", "bnd2d_out", "bnd1d_in", "bnd1d_out"]


def main() -> None:
    d = inputs()
    np.savez_compressed(HERE / "record_inputs.npz", **d)
    wall, batch_wall, batch_size, dx, rain_factor, slr, arf, mass_err, infil, stored = d["scalars"]
    tide = sine_tide(*d["tide"][:2], [d["tide"][2]])
    inflow = PiecewiseLinear(0.0, float(d["inflow_step_s"]), d["inflow"])
    fo = SimpleNamespace(rain=Hyetograph(float(d["rain_step_s"]), d["rain_mmh"].copy(), float(d["storm_start"])),
                         tide=lambda t: float(np.asarray(tide(t)).reshape(-1)[0]),
                         inflow=lambda t: float(np.asarray(inflow(t)).reshape(-1)[0]),
                         t_end=float(d["t_end"]), storm_start=float(d["storm_start"]), rain_factor=float(rain_factor),
                         slr_m=float(slr), arf_domain=float(arf))
    rl = d["run_log"]
    log = RefRunLog(steps=int(rl[0]), wall_s=float(rl[1]), dt_min=float(rl[2]), dt_max=float(rl[3]),
                    n_sub_max=int(rl[4]), n_dt_clamped=int(rl[5]), exchanged_in_m3=float(rl[6]),
                    exchanged_out_m3=float(rl[7]))
    extras = {"forcing": fo, "run_log": log, "sources": {"rain": 3.5, "infiltration": 1.25, "bnd2d_in": 0.5},
              "ledger": [{k: float(row[j]) for j, k in enumerate(LEDGER_KEYS)} for row in d["ledger"]],
              "mass_series": [{"t": float(r[0]), "volume": float(r[1]), "rel_error": float(r[2])} for r in d["series"]]}
    res = SimpleNamespace(times=torch.as_tensor(d["times"]), depth_series=torch.as_tensor(d["h"]),
                          u=torch.as_tensor(d["u"]), v=torch.as_tensor(d["v"]), channel_eta=torch.as_tensor(d["eta1"]),
                          channel_q=torch.as_tensor(d["q1"]), mass_balance_error=float(mass_err),
                          infiltrated_volume_m3=float(infil), stored_volume_m3=float(stored), in_distribution=True,
                          extras=extras)
    cfg = SimpleNamespace(model=SimpleNamespace(forcing_steps=12))
    dom = SimpleNamespace(dx=float(dx), sea=d["sea"])
    smp = SimpleNamespace(meta=lambda: dict(META))
    out = HERE / "_record_tmp"
    out.mkdir(exist_ok=True)
    path = out / f"sim_{SIM_ID:05d}.pt"
    index: dict = {}
    RG._write_sim(cfg, dom, F, SIM_ID, smp, d["landuse"], d["storage"], d["kappa"], d["manning"], d["gamma"], res,
                  path, index, float(wall), float(batch_wall), int(batch_size))
    rec = torch.load(path, weights_only=True)
    torch.save(rec, HERE / "record_ref.pt")
    (HERE / "record_ref_index.json").write_text(json.dumps(index[str(SIM_ID)], indent=1, sort_keys=True))
    path.unlink()
    out.rmdir()
    print("wrote record_inputs.npz, record_ref.pt, record_ref_index.json")


if __name__ == "__main__":
    main()
