"""The simulation record against one written by the reference's own writer.

``tests/golden/make_record_golden.py`` fed hydrointel's ``_write_sim`` (generate.py:355-397) a
small synthetic flood result and stored the record and index entry it wrote.  Here the same inputs
go through ``records.simulation_record`` and ``records.index_entry``: the record must hold the same
keys at every level, the same tensor dtypes and shapes, the same numbers and strings, and the same
values: bit for bit, except the Horton rates, which this package forms in m/s and the reference in
mm/h (held to float32 round-off; on this case they too agree bit for bit).  Among what this pins:
the block means, the depth-weighted velocities, the rain resampled from a hyetograph that starts
after a spin-up,
the Horton rates measured from the storm's start, the forcing block's rain factor, sea-level rise
and areal reduction factor, the ledger with its coarse-extent surface volume, and the run log.
"""
import json
from pathlib import Path

import numpy as np
import torch

from his_amd import records as R
from his_amd.api import FloodResult, SiteState
from his_amd.engine import Forcing
from his_amd.forcing import PiecewiseConstant, PiecewiseLinear, sine_tide
from his_amd.solver.coupling import RunLog

GOLDEN = Path(__file__).resolve().parent / "golden"
MMH = 1.0 / 3.6e6
META = {"return_period_yr": 25, "rcp": "RCP4.5", "ssp": "SSP2", "horizon_year": 2060, "tide_on": True,
        "storm_tide_offset_h": 1.5, "site": "synthetic"}
F, SIM_ID = 2, 7


def _inputs():
    return dict(np.load(GOLDEN / "record_inputs.npz"))


def _result(d):
    wall, batch_wall, batch_size, dx, rain_factor, slr, arf, mass_err, infil, stored = d["scalars"]
    tide = sine_tide(*d["tide"][:2], [d["tide"][2]])
    inflow = PiecewiseLinear(0.0, float(d["inflow_step_s"]), d["inflow"])
    rain = PiecewiseConstant(float(d["storm_start"]), float(d["rain_step_s"]), [d["rain_mmh"] * MMH])
    fo = Forcing(rain, inflow, tide, float(d["t_end"]), storm_start=float(d["storm_start"]))
    rl = d["run_log"]
    log = RunLog(steps=int(rl[0]), wall_s=float(rl[1]), dt_min=float(rl[2]), dt_max=float(rl[3]), n_sub_max=int(rl[4]),
                 n_dt_clamped=int(rl[5]), exchanged_in_m3=float(rl[6]), exchanged_out_m3=float(rl[7]))
    extras = {"forcing": fo, "run_log": log, "sources": {"rain": 3.5, "infiltration": 1.25, "bnd2d_in": 0.5},
              "ledger": [{k: float(row[j]) for j, k in enumerate(R.LEDGER_KEYS)} for row in d["ledger"]],
              "mass_series": [{"t": float(r[0]), "volume": float(r[1]), "rel_error": float(r[2])} for r in d["series"]]}
    h = torch.as_tensor(d["h"])
    res = FloodResult(depth_max=h.max(0).values, depth_series=h, u=torch.as_tensor(d["u"]), v=torch.as_tensor(d["v"]),
                      channel_eta=torch.as_tensor(d["eta1"]), channel_q=torch.as_tensor(d["q1"]),
                      times=torch.as_tensor(d["times"]), infiltrated_volume_m3=float(infil),
                      stored_volume_m3=float(stored), mass_balance_error=float(mass_err), provenance={}, extras=extras)
    site = SiteState(torch.as_tensor(d["storage"]), torch.as_tensor(d["kappa"]), torch.as_tensor(d["manning"]),
                     torch.as_tensor(d["gamma"]))
    factors = {"rain_factor": float(rain_factor), "slr_m": float(slr), "arf": float(arf)}
    return res, site, factors, dict(wall=float(wall), batch_wall=float(batch_wall), batch_size=int(batch_size),
                                    dx=float(dx))


def _record(d):
    res, site, factors, s = _result(d)
    lu = d["landuse"]
    rec = R.simulation_record(res, F, s["dx"], site=site, meta=dict(META), landuse_frac=R.landuse_fractions(lu),
                              horton_rate_mmh=R.horton_rate_at(lu, d["kappa"], d["times"], F, float(d["storm_start"])),
                              forcing_steps=12, forcing_factors=factors, wall_s=s["wall"])
    return rec, res, s


def _compare(got, want, where=""):
    if isinstance(want, dict):
        assert isinstance(got, dict) and set(got) == set(want), f"{where}: keys {sorted(got)} vs {sorted(want)}"
        for k in want:
            _compare(got[k], want[k], f"{where}.{k}")
    elif isinstance(want, list):
        assert isinstance(got, list) and len(got) == len(want), where
        for i, (g, w) in enumerate(zip(got, want)):
            _compare(g, w, f"{where}[{i}]")
    elif torch.is_tensor(want):
        assert torch.is_tensor(got) and got.dtype == want.dtype and got.shape == want.shape, \
            f"{where}: {getattr(got, 'dtype', type(got))} {tuple(getattr(got, 'shape', ()))} vs {want.dtype} {tuple(want.shape)}"
        if where.endswith("horton_rate_mmh"):
            scale = float(want.abs().max()) if want.numel() else 0.0
            err = float((got - want).abs().max()) if want.numel() else 0.0
            assert err <= 2e-7 * max(scale, 1e-30) + 1e-30, f"{where}: max difference {err:.3e} at scale {scale:.3e}"
        else:
            assert torch.equal(got, want), f"{where}: values differ"
    else:
        assert type(got) is type(want) and got == want, f"{where}: {got!r} vs {want!r}"


def test_record_matches_the_reference_writer():
    d = _inputs()
    rec, _, _ = _record(d)
    want = torch.load(GOLDEN / "record_ref.pt", weights_only=True)
    _compare(rec, want, "record")


def test_record_round_trips_and_index_entry_matches(tmp_path):
    d = _inputs()
    rec, res, s = _record(d)
    path = R.record_path(tmp_path, SIM_ID, False)
    R.save_record(rec, path)
    _compare(R.load_record(path), torch.load(GOLDEN / "record_ref.pt", weights_only=True), "saved record")
    entry = R.index_entry(path, res, d["sea"], dict(META), F, s["wall"], s["batch_wall"], s["batch_size"])
    want = json.loads((GOLDEN / "record_ref_index.json").read_text())
    assert set(entry) == set(want)
    for k, v in want.items():
        if k == "bytes":                       # the file size depends on the pickler, not the format
            assert entry[k] > 0
        elif isinstance(v, float):
            assert abs(entry[k] - v) <= 1e-12 * max(abs(v), 1.0), k
        else:
            assert entry[k] == v, k


def test_horton_series_is_measured_from_the_storm_start():
    """Before the storm the rate is the land use's initial capacity; a series measured from t = 0
    instead of the storm's start differs from the reference's at every output time."""
    d = _inputs()
    lu, kap, times = d["landuse"], d["kappa"], d["times"]
    start = float(d["storm_start"])
    rates = R.horton_rate_at(lu, kap, times, F, start)
    pre = R.horton_rate_at(lu, kap, np.zeros(1), F, 0.0)[0]
    for t, r in zip(times, rates):
        if t <= start:
            assert np.array_equal(r, pre)
    assert not np.allclose(R.horton_rate_at(lu, kap, times, F, 0.0)[1:], rates[1:])
