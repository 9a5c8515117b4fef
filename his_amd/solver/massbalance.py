"""Float64 volume ledger of a coupled run, one column per member.

The terms and the error measure are hydrointel's (hydrointel/solver/massbalance.py:18-83), so a
mass check means the same thing in both:

    error = |V_now - V0 - net| / max(gross, 1e-9 |V0| + 1)
    net   = rain - infiltration + bnd2d_in - bnd2d_out + bnd1d_in - bnd1d_out
    gross = rain + infiltration + bnd2d_in + bnd2d_out + bnd1d_in + bnd1d_out

with ``clip`` (depth added by the h >= 0 clips) kept in the ledger but outside both sums, and
checked per member so that one bad storm cannot hide in a batch average (AssertionError).

The ledger is one (7, B) float64 tensor on the engine's device.  Host-driven steps add their
step volumes to it without a host round trip; the device-resident engine keeps its own
(10, B) ledger in its control block and ``fold_engine_rows`` adds that in one operation.
"""
from __future__ import annotations

import numpy as np
import torch

TERMS = ("rain", "infiltration", "bnd2d_in", "bnd2d_out", "bnd1d_in", "bnd1d_out", "clip")
_ROW = {name: k for k, name in enumerate(TERMS)}
_SIGN = np.array([1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 0.0])      # net = SIGN . ledger
_GROSS = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0])        # gross = GROSS . ledger


def _host(v) -> np.ndarray:
    if torch.is_tensor(v):
        return v.detach().to(torch.float64).cpu().numpy()
    return np.asarray(v, dtype=np.float64)


class MassBalance:
    def __init__(self, device, v0):
        self.device = torch.device(device) if device is not None else torch.device("cpu")
        self.v0 = _host(v0)
        self.batch = self.v0.shape[0] if self.v0.ndim else None
        self.ledger = torch.zeros((len(TERMS), 1 if self.batch is None else self.batch), dtype=torch.float64,
                                  device=self.device)
        self.series: list = []

    @property
    def members(self) -> int:
        return self.ledger.shape[1]

    def _column(self, name: str, value) -> torch.Tensor:
        """A term's value as a float64 (B,) tensor on the ledger's device."""
        v = torch.as_tensor(value, dtype=torch.float64)
        v = v.to(self.device).reshape(-1)
        if v.numel() == 1:
            return v.expand(self.members)
        if v.numel() != self.members:
            raise ValueError(f"mass term '{name}' carries {v.numel()} values for {self.members} members")
        return v

    def add(self, name: str, value) -> None:
        """Add one step's volume of a term (a scalar, or one value per member)."""
        if name not in _ROW:
            raise KeyError(name)
        self.ledger[_ROW[name]] += self._column(name, value)

    def fold_engine_rows(self, acc: torch.Tensor) -> None:
        """Add the device engine's ledger rows for these members, acc = (10, B) float64: rain,
        infiltration, bnd2d_in, bnd2d_out, bnd1d_in, bnd1d_out, clip2d, clip1d, exchanged in/out."""
        acc = acc.to(device=self.device, dtype=torch.float64)
        self.ledger[:6] += acc[:6]
        self.ledger[6] += acc[6] + acc[7]

    def sources(self) -> dict:
        """Accumulated volume of each term: arrays over members, floats when unbatched."""
        rows = _host(self.ledger)
        if self.batch is None:
            return {name: float(rows[k, 0]) for name, k in _ROW.items()}
        return {name: rows[k].copy() for name, k in _ROW.items()}

    def _balance(self, volume_now):
        rows = _host(self.ledger)
        net = _SIGN @ rows
        gross = _GROSS @ rows
        now = _host(volume_now).reshape(-1)
        v0 = self.v0.reshape(-1)
        err = np.abs(now - v0 - net) / np.maximum(gross, 1e-9 * np.abs(v0) + 1.0)
        return now, net, gross, rows[_ROW["clip"]], err

    def check(self, volume_now, t: float, record: bool = True):
        """Relative mass-balance error per member (a float when unbatched); optionally logged."""
        now, net, gross, clip, err = self._balance(volume_now)
        one = (lambda a: a) if self.batch is not None else (lambda a: float(a[0]))
        if record:
            self.series.append({"t": t, "volume": one(now), "net_sources": one(net), "gross": one(gross),
                                "clip": one(clip), "rel_error": one(err)})
        return one(err)

    def assert_ok(self, volume_now, t: float, tol: float, label: str = ""):
        """``check``, raising AssertionError that names every member at or above ``tol``."""
        err = self.check(volume_now, t)
        errs = np.atleast_1d(np.asarray(err, dtype=np.float64))
        bad = np.flatnonzero(~(errs < tol))
        if bad.size:
            where = f" ({label})" if label else ""
            if self.batch is None:
                raise AssertionError(f"mass not conserved{where} at t = {t:.1f} s: relative error "
                                     f"{errs[0]:.3e}, tolerance {tol:.1e}; ledger {self.sources()}")
            shown = "; ".join(f"member {int(b)} error {errs[b]:.3e}" for b in bad[:8])
            raise AssertionError(f"mass not conserved{where} at t = {t:.1f} s in {bad.size} of {self.batch} "
                                 f"members (tolerance {tol:.1e}): {shown}")
        return err
