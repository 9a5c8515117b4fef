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
    def from_mmh(cls, f0_mmh, fc_mmh, k_per_h, device, dtype, t0: float = 0.0) -> "Horton":
        t = lambda a: torch.as_tensor(a, dtype=dtype, device=device)
        return cls(t(f0_mmh) * MMH_TO_MS, t(fc_mmh) * MMH_TO_MS, t(k_per_h) / 3600.0, t0)

    def rate(self, t: float) -> torch.Tensor:
        tt = torch.clamp(torch.as_tensor(t) - self.t0, min=0.0)
        return self.fc + (self.f0 - self.fc) * torch.exp(-self.k * tt)
