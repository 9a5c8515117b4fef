"""Solver configuration for the coupled flood-engine time step.

Mirrors ``SolverConfig`` (hydrointel/config.py:95-131) field for field, with the
same defaults, and ``RunConfig.dtype`` (hydrointel/config.py:272-275).  One
precision mode is added: ``"fp32bulk"`` keeps the 2-D state, reconstruction and
fluxes in float32 while every volume and mass accumulator, the 1-D network,
the weir exchange and simulated time stay float64.  The reference's ``"fp32"``
is accepted as an alias of ``"fp32bulk"``.
"""
from __future__ import annotations

from dataclasses import dataclass, fields

import torch

PRECISIONS = ("fp64", "fp32bulk", "fp32")


@dataclass
class SolverConfig:
    g: float = 9.81
    h_dry: float = 1e-3
    kp_eps: float = 1e-12
    cfl: float = 0.45
    dt_min: float = 1e-4
    dt_max: float = 5.0
    spatial_order: int = 2
    flux: str = "hllc"                  # "hllc" | "hll"
    precision: str = "fp64"             # "fp64" | "fp32bulk" ("fp32" is an alias)
    compile: bool = True                # kept for call-surface compatibility; the HIP path is always used
    mass_tol: float = 1e-3
    recession_h: float = 2.0
    output_interval_s: float = 900.0
    slot_width_frac: float = 1.0
    cfl_1d: float = 0.45
    weir_cw: float = 1.7
    exchange_limit: float = 0.5
    exchange_relax: float = 0.5

    def validate(self) -> "SolverConfig":
        if self.precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}, got {self.precision!r}")
        if self.flux not in ("hllc", "hll"):
            raise ValueError(f"flux must be 'hllc' or 'hll', got {self.flux!r}")
        if self.spatial_order not in (1, 2):
            raise ValueError(f"spatial_order must be 1 or 2, got {self.spatial_order}")
        return self

    @property
    def dtype(self) -> torch.dtype:
        """Working dtype of the 2-D state (hydrointel/config.py:272-275)."""
        return torch.float64 if self.precision == "fp64" else torch.float32

    def replace(self, **kw) -> "SolverConfig":
        d = {f.name: getattr(self, f.name) for f in fields(self)}
        d.update(kw)
        return SolverConfig(**d)
