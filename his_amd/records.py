"""Simulation records in the reference's dataset format, and the loop that writes them.

hydrointel's ``generate`` runs sampled storms through ``simulate_batch`` and writes one record
per storm, ``sim_XXXXX.pt`` or gzip-compressed ``sim_XXXXX.pt.gz`` (hydrointel/data/generate.py:
37-63), plus ``index.json`` with one entry per storm. A record holds (generate.py:355-425):

* the output-time snapshots block-averaged onto the storage grid: depth, and velocities as
  depth-weighted means (``coarsen_mean``, generate.py:104-110);
* the peak depth at full resolution (``depth_max_full``) and the channel level and discharge;
* the storm's fields and its forcing resampled to ``forcing_steps`` values;
* its volumes and mass-balance checks;
* the volume ledger at the output times, with the surface volume inside the coarse extent;
* the run log.

``simulation_record`` builds the same record from a FloodResult of this engine, and
``write_storms`` is the writing half of ``generate``: run a batch, write each record, keep the
index. Several inputs come from the reference's scenario sampler, land-use tables and
design-storm construction, which sit upstream of the engine: the storm's metadata, land-use
fractions, Horton rates per land-use class, and the forcing's design factors. The caller
passes those in.
"""
from __future__ import annotations

import dataclasses
import gzip
import json
import time
from pathlib import Path

import numpy as np
import torch

COMPRESSED_SUFFIX = ".pt.gz"
MMH = 1.0 / 3.6e6                      # m/s per mm/h
FAILED = "failed_numerical"            # index status of a storm the engine could not simulate
LEDGER_KEYS = ("v2d", "v1d", "storage", "rain", "infiltration", "bnd2d_in", "bnd2d_out", "bnd1d_in",
               "bnd1d_out")


def record_path(root, sim_id: int, compress: bool | None = None) -> Path:
    """Where storm ``sim_id`` lives (generate.py:37-43). With ``compress`` None, an existing
    compressed file wins when there is no plain one, so a directory holding either format reads
    back."""
    root = Path(root)
    plain, gz = root / f"sim_{sim_id:05d}.pt", root / f"sim_{sim_id:05d}{COMPRESSED_SUFFIX}"
    if compress is None:
        return gz if gz.exists() and not plain.exists() else plain
    return gz if compress else plain


def save_record(rec: dict, path) -> None:
    """torch.save, through gzip level 1 for a ``.pt.gz`` path (generate.py:46-52)."""
    path = Path(path)
    if path.name.endswith(COMPRESSED_SUFFIX):
        with gzip.open(path, "wb", compresslevel=1) as fh:
            torch.save(rec, fh)
    else:
        torch.save(rec, path)


def load_record(path) -> dict:
    """Read a record back. Records hold only tensors, numbers, strings and containers of
    them, so they load with torch's restricted unpickler; plain files are memory-mapped."""
    if str(path).endswith(COMPRESSED_SUFFIX):
        with gzip.open(path, "rb") as fh:
            return torch.load(fh, weights_only=True)
    return torch.load(path, weights_only=True, mmap=True)


def coarsen_mean(a: np.ndarray, f: int) -> np.ndarray:
    """Block mean over f x f cells of the last two axes, dropping the ragged edge
    (generate.py:104-110)."""
    if f == 1:
        return a
    *lead, ny, nx = a.shape
    cy, cx = ny // f, nx // f
    return a[..., :cy * f, :cx * f].reshape(*lead, cy, f, cx, f).mean(axis=(-1, -3))


def _member_value(series, t: float, b: int) -> float:
    v = np.asarray(series(t), dtype=np.float64).reshape(-1)
    return float(v[b if v.size > 1 else 0])


def resampled_mmh(rain, t_end: float, n: int, b: int = 0) -> np.ndarray:
    """Member b's rain intensity (mm/h) as block means over n equal steps of [0, t_end], from 20
    samples per step (hydrointel/forcing/hyetograph.py:51-55). ``rain`` gives m/s."""
    fine = (np.arange(20 * n) + 0.5) * t_end / (20 * n)
    vals = np.array([_member_value(rain, t, b) for t in fine]) / MMH
    return vals.reshape(n, 20).mean(axis=1)


def _plain(v, b: int):
    """A run-log or ledger value as a plain Python number (member b's, for per-member arrays)."""
    if isinstance(v, np.ndarray) or torch.is_tensor(v):
        a = np.asarray(v.detach().cpu() if torch.is_tensor(v) else v, dtype=np.float64).reshape(-1)
        return float(a[b if a.size > 1 else 0])
    if isinstance(v, (bool, np.bool_)):
        return bool(v)
    if isinstance(v, (int, np.integer)):
        return int(v)
    if isinstance(v, (float, np.floating)):
        return float(v)
    return v


def _to_tensors(obj):
    """numpy arrays to tensors, numpy scalars to Python numbers, recursively (generate.py:420-425)."""
    if isinstance(obj, dict):
        return {k: _to_tensors(v) for k, v in obj.items()}
    if isinstance(obj, (list, tuple)):
        return [_to_tensors(v) for v in obj]
    if isinstance(obj, np.ndarray) and obj.dtype != object and obj.dtype.kind in "fiub":
        return torch.from_numpy(np.ascontiguousarray(obj))
    if isinstance(obj, np.generic):
        return obj.item()
    return obj


def ledger_at_outputs(res, h: np.ndarray, f: int, dx: float) -> dict:
    """The engine's volume ledger at the output times, plus ``v2d_coarse``: the surface volume
    inside the coarse-grid extent (generate.py:409-417)."""
    L = res.extras["ledger"]
    out = {k: np.array([x[k] for x in L], np.float64) for k in LEDGER_KEYS}
    ny, nx = h.shape[-2:]
    out["v2d_coarse"] = h[:, :(ny // f) * f, :(nx // f) * f].astype(np.float64).sum(axis=(1, 2)) * dx * dx
    return out


def simulation_record(res, f: int, dx: float, *, site, meta: dict | None = None, landuse_frac=None,
                      horton_rate_mmh=None, forcing_steps: int = 96, forcing_factors: dict | None = None,
                      wall_s: float = 0.0) -> dict:
    """One storm's record (generate.py:355-397) from its FloodResult and SiteState.

    ``landuse_frac`` (n_classes, ny, nx) and ``horton_rate_mmh`` (nt, ny/f, nx/f) are stored
    when given; ``forcing_factors`` (for instance rain_factor, slr_m, arf) join the forcing
    block."""
    fo = res.extras["forcing"]
    # a result's forcing holds its own storm (one member) or the whole batch's
    b = 0 if getattr(fo.rain, "n", 1) == 1 else int(res.extras.get("batch_index", 0))
    times = res.times.numpy()
    h, u, v = res.depth_series.numpy(), res.u.numpy(), res.v.numpy()
    hc = coarsen_mean(h, f)
    hc_safe = np.where(hc > 0, hc, 1.0)
    uc = np.where(hc > 0, coarsen_mean(h * u, f) / hc_safe, 0.0)
    vc = np.where(hc > 0, coarsen_mean(h * v, f) / hc_safe, 0.0)
    arr = lambda t: np.asarray(t.detach().cpu() if torch.is_tensor(t) else t, dtype=np.float64)
    fields = {"storage": coarsen_mean(arr(site.storage_depth), f).astype(np.float32),
              "kappa": coarsen_mean(arr(site.infil_multiplier), f).astype(np.float32),
              "manning": coarsen_mean(arr(site.manning), f).astype(np.float32),
              "gamma": arr(site.channel_gamma).astype(np.float32)}
    if landuse_frac is not None:
        fields["landuse_frac"] = coarsen_mean(np.asarray(landuse_frac, dtype=np.float64), f).astype(np.float32)
    grid_t = np.linspace(0.0, fo.t_end, forcing_steps)
    forcing = {"rain_mmh": resampled_mmh(fo.rain, fo.t_end, forcing_steps, b).astype(np.float32),
               "tide_m": np.array([_member_value(fo.tide, t, b) for t in grid_t], np.float32),
               "inflow_m3s": np.array([_member_value(fo.inflow, t, b) for t in grid_t], np.float32),
               "t_end": float(fo.t_end), "storm_start": float(fo.storm_start)}
    forcing.update({k: _plain(x, b) for k, x in (forcing_factors or {}).items()})
    if horton_rate_mmh is not None:
        forcing["horton_rate_mmh"] = np.asarray(horton_rate_mmh, dtype=np.float32)
    log = res.extras["run_log"]
    rec = {
        "meta": dict(meta or {}), "wall_s": float(wall_s), "mass_balance_error": float(res.mass_balance_error),
        "times": times.astype(np.float32),
        "h": hc.astype(np.float32), "u": uc.astype(np.float32), "v": vc.astype(np.float32),
        "depth_max_full": h.max(0).astype(np.float32),
        "eta1": res.channel_eta.numpy(), "q1": res.channel_q.numpy(),
        "fields": fields, "forcing": forcing,
        "volumes": {"infiltrated_m3": float(res.infiltrated_volume_m3), "stored_m3": float(res.stored_volume_m3),
                    "sources": {k: _plain(x, b) for k, x in res.extras["sources"].items()},
                    "series": [{k: _plain(x, b) for k, x in r.items()} for r in res.extras["mass_series"]]},
        "ledger": ledger_at_outputs(res, h, f, dx),
        "run_log": {k: _plain(x, b) for k, x in dataclasses.asdict(log).items() if k != "series"},
        "store_coarsen": int(f),
    }
    return _to_tensors(rec)


def index_entry(path: Path, res, sea: np.ndarray, meta: dict, f: int, wall_s: float, batch_wall_s: float,
                batch_size: int) -> dict:
    """The storm's ``index.json`` entry (generate.py:392-397)."""
    h = res.depth_series.numpy()
    land = ~np.asarray(sea, dtype=bool)
    return {**meta, "wall_s": wall_s, "batch_wall_s": batch_wall_s, "batch_size": batch_size,
            "store_coarsen": int(f), "record": path.name, "bytes": path.stat().st_size,
            "mass_err": float(res.mass_balance_error),
            "peak_depth_land": float(h.max(0)[land].max()) if land.any() else 0.0,
            "in_distribution": bool(res.in_distribution)}


def horton_rate_at(landuse, kappa, times, f: int, storm_start: float = 0.0) -> np.ndarray:
    """Horton capacity f(t) = fc + (f0 - fc) exp(-k t) [mm/h] of the land use scaled by kappa, with t
    the time since the storm started (0 before it), at each output time, block-averaged onto the
    storage grid: (nt, ny/f, nx/f) (generate.py:400-406)."""
    from .domain import landuse as LU
    f0, fc, k = LU.horton_fields(landuse, kappa)
    out = [coarsen_mean((fc + (f0 - fc) * np.exp(-k * max(float(t) - storm_start, 0.0))) / MMH, f)
           for t in np.asarray(times)]
    return np.stack(out).astype(np.float32)


def landuse_fractions(landuse) -> np.ndarray:
    """One-hot class map (n_classes, ny, nx); ``simulation_record`` block-averages it into each class's
    share of every storage-grid cell."""
    from .domain import landuse as LU
    lu = np.asarray(landuse)
    return np.stack([(lu == c).astype(np.float64) for c in range(LU.N_CLASSES)])


def write_storms(out, storms: list, f: int = 1, compress: bool = False, forcing_steps: int = 96,
                 output_interval_s: float | None = None) -> dict:
    """Run ``storms`` -- a list of (sim_id, SiteState, Scenario, meta dict) -- as one batch through
    ``api.simulate_batch(sites, scenarios)``, as generate.py:288 calls it, and write each storm's record
    and index entry (the loop body of generate.py:229-321) over the configured domain.  Resumable: a
    storm whose record exists and which the index holds is skipped, and so is one recorded as failed.
    A batch of one that the engine cannot simulate is recorded as failed; in a larger batch the
    failing member cannot be told apart and the error propagates."""
    from . import api
    from .scenarios import rain_factor, sea_level_rise
    from .solver.coupling import NumericalInstabilityError
    ctx = api.context()
    dom = ctx.domain
    out = Path(out)
    out.mkdir(parents=True, exist_ok=True)
    index_path = out / "index.json"
    index = json.loads(index_path.read_text(encoding="utf-8")) if index_path.exists() else {}
    run = [st for st in storms
           if not (record_path(out, st[0]).exists() and str(st[0]) in index)
           and index.get(str(st[0]), {}).get("status") != FAILED]
    if not run:
        return index
    sites, scens = [st[1] for st in run], [st[2] for st in run]
    t0 = time.perf_counter()
    try:
        results = api.simulate_batch(sites, scens, output_interval_s)
    except NumericalInstabilityError as e:
        if len(run) > 1:
            raise
        i, _, _, meta = run[0]
        index[str(i)] = {**meta, "status": FAILED, "error": str(e), "wall_s": time.perf_counter() - t0}
        index_path.write_text(json.dumps(index, indent=1), encoding="utf-8")
        return index
    batch_wall = time.perf_counter() - t0
    wall = batch_wall / len(run)
    for (i, site, scen, meta), res in zip(run, results):
        path = record_path(out, i, compress)
        lu = ctx.landuse(scen.ssp, scen.horizon_year)
        fo = res.extras["forcing"]
        # the design factors of this storm's forcing (generate.py:383-384) when its resolver records
        # them, else the scenario's pathway values and no areal reduction (the default design storm)
        factors = {"rain_factor": getattr(fo, "rain_factor", rain_factor(scen.rcp)),
                   "slr_m": getattr(fo, "slr_m", sea_level_rise(scen.rcp, scen.horizon_year)),
                   "arf": getattr(fo, "arf_domain", 1.0)}
        meta = {**meta, "return_period_yr": scen.return_period_yr, "rcp": scen.rcp, "ssp": scen.ssp,
                "horizon_year": scen.horizon_year, "tide_on": scen.tide_on,
                "storm_tide_offset_h": scen.storm_tide_offset_h}
        rec = simulation_record(res, f, dom.dx, site=site, meta=meta, landuse_frac=landuse_fractions(lu),
                                horton_rate_mmh=horton_rate_at(lu, site.infil_multiplier.double().numpy(),
                                                               res.times.numpy(), f, float(fo.storm_start)),
                                forcing_steps=forcing_steps, wall_s=wall, forcing_factors=factors)
        save_record(rec, path)
        index[str(i)] = index_entry(path, res, dom.sea, meta, f, wall, batch_wall, len(run))
        index_path.write_text(json.dumps(index, indent=1), encoding="utf-8")
    return index
