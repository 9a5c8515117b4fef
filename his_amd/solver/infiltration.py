"""Horton infiltration fields.

Same data model as hydrointel/solver/infiltration.py:15-37: per-cell f0, fc [m/s], k
[1/s] and the decay-clock start t0; f(t) = fc + (f0 - fc) exp(-k max(t - t0, 0)).  The
rate is evaluated inside the 2-D kernel; ``rate`` here is the host-side form for callers.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

MMH_TO_MS = 1.0 / 3.6e6


@dataclass
class Horton:
    f0: torch.Tensor
    fc: torch.Tensor
    k: torch.Tensor
    t0: float = 0.0

    @classmethod
    def constant(cls, rate_ms, shape, device, dtype) -> "Horton":
        r = torch.full(shape, float(rate_ms), device=device, dtype=dtype)
        return cls(r, r.clone(), torch.zeros(shape, device=device, dtype=dtype))

    @classmethod
    def from_landuse(cls, landuse, kappa, device, dtype, t0: float = 0.0) -> "Horton":
        """The land-use classes' capacities times kappa and their decay rates
        (hydrointel/solver/infiltration.py:23-28), formed in float64 and stored in ``dtype``."""
        from ..domain.landuse import horton_fields
        kap = kappa.detach().cpu().numpy() if torch.is_tensor(kappa) else kappa
        t = lambda a: torch.as_tensor(a, dtype=dtype, device=device)
        f0, fc, k = horton_fields(landuse, kap)
        return cls(t(f0), t(fc), t(k), t0)

    def rate(self, t: float) -> torch.Tensor:
        tt = torch.clamp(torch.as_tensor(t) - self.t0, min=0.0)
        return self.fc + (self.f0 - self.fc) * torch.exp(-self.k * tt)
