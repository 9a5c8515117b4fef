"""Storm scenarios and how they become the forcing of a run.

``Scenario`` keeps hydrointel's fields and validation (hydrointel/api.py:55-63, 445-454): a return
period of 10, 25, 50 or 100 years, an RCP and an SSP from the pathway tables, a horizon year within
2020-2100, whether the tide runs, and the offset of the storm from high tide.  The pathway tables are
hydrointel's (hydrointel/forcing/scenarios.py:17-35: RCP rain factor and sea-level rise at 2050 and
2100; SSP urban expansion, population factor and drainage-investment factor), so a scenario means the
same thing in both.

Turning a scenario into rain, tide and inflow is the work of a *resolver*, ``resolver(scenario, domain)
-> Forcing``.  hydrointel's own (IDF alternating-block rainfall, a Clark unit hydrograph, a
four-constituent tide; hydrointel/engine.py:53-70) sits upstream of the engine and is not part of this
package; ``register_resolver`` installs it, or any other.  The default, ``DesignStorm``, is this
package's seeded storm generator, the one its benchmark workloads use: a triangular hyetograph over 6 h
with 2.5 h of recession, peaking at the return period's intensity times the RCP's rain factor; a
sinusoidal tide of 0.8 m amplitude and 12.42 h period about the sea level of the RCP and horizon
year, at high water ``storm_tide_offset_h`` after the rain peak (mean sea level held when the tide
is off); and a triangular upstream inflow hydrograph whose peak follows the return period.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .forcing import Harmonic, triangle

RETURN_PERIODS = (10, 25, 50, 100)
# rain factor, sea-level rise at 2050 (m), at 2100 (m)
RCP_TABLE = {"RCP2.6": (1.05, 0.18, 0.40), "RCP4.5": (1.10, 0.22, 0.55), "RCP6.0": (1.14, 0.24, 0.62),
             "RCP8.5": (1.22, 0.28, 0.85)}
# urban expansion (fraction of convertible land per decade), population factor, drainage investment
SSP_TABLE = {"SSP1": (0.010, 1.10, 1.15), "SSP2": (0.020, 1.25, 1.05), "SSP3": (0.035, 1.45, 0.90),
             "SSP4": (0.025, 1.20, 0.95), "SSP5": (0.040, 1.35, 1.20)}
SLR_BASE_YEAR = 2020


@dataclass(frozen=True)
class Scenario:
    return_period_yr: int
    rcp: str
    ssp: str
    horizon_year: int
    tide_on: bool = True
    storm_tide_offset_h: float = 0.0


def validate(s: Scenario) -> None:
    if s.return_period_yr not in RETURN_PERIODS:
        raise ValueError(f"return_period_yr must be one of {RETURN_PERIODS}, not {s.return_period_yr}")
    if s.rcp not in RCP_TABLE:
        raise ValueError(f"rcp {s.rcp!r} is not one of {sorted(RCP_TABLE)}")
    if s.ssp not in SSP_TABLE:
        raise ValueError(f"ssp {s.ssp!r} is not one of {sorted(SSP_TABLE)}")
    if not SLR_BASE_YEAR <= s.horizon_year <= 2100:
        raise ValueError(f"horizon_year must lie in {SLR_BASE_YEAR}-2100, not {s.horizon_year}")


def rain_factor(rcp: str) -> float:
    return RCP_TABLE[rcp][0]


def sea_level_rise(rcp: str, year: int) -> float:
    """Metres above today's sea level: linear from 0 in 2020 to the 2050 value, then to the 2100 value."""
    _, at2050, at2100 = RCP_TABLE[rcp]
    if year <= SLR_BASE_YEAR:
        return 0.0
    if year <= 2050:
        return at2050 * (year - SLR_BASE_YEAR) / (2050 - SLR_BASE_YEAR)
    return at2050 + (at2100 - at2050) * (min(year, 2100) - 2050) / 50.0


def slr_m(rcp: str, year: int) -> float:
    """hydrointel's name for ``sea_level_rise`` (forcing/scenarios.py:48-57), which, like it, refuses a
    year beyond the projection table."""
    if rcp not in RCP_TABLE:
        raise ValueError(f"unknown RCP {rcp!r} (expected one of {sorted(RCP_TABLE)})")
    if year > 2100:
        raise ValueError(f"horizon_year {year} beyond 2100 is outside the projection table")
    return sea_level_rise(rcp, year)


@dataclass(frozen=True)
class SSPState:
    """A pathway's socioeconomic modifiers, with hydrointel's field names (forcing/scenarios.py:69-82)."""
    ssp: str
    urban_expansion_rate: float
    population_factor: float
    drainage_investment_factor: float


def ssp_state(ssp: str) -> SSPState:
    if ssp not in SSP_TABLE:
        raise ValueError(f"unknown SSP {ssp!r} (expected one of {sorted(SSP_TABLE)})")
    return SSPState(ssp, *SSP_TABLE[ssp])


def drainage_factor(ssp: str) -> float:
    """Multiplier on channel conveyance from the pathway's drainage investment."""
    return SSP_TABLE[ssp][2]


def population_factor(ssp: str) -> float:
    return SSP_TABLE[ssp][1]


class DesignStorm:
    """This package's storm for a scenario (see the module docstring); one member per call."""
    storm_s = 6 * 3600.0
    recession_s = 2.5 * 3600.0
    peak_mmh = {10: 50.0, 25: 65.0, 50: 80.0, 100: 95.0}
    inflow_peak = {10: 150.0, 25: 220.0, 50: 300.0, 100: 380.0}
    inflow_base = 10.0
    tide_amp_m = 0.8
    tide_period_s = 12.42 * 3600.0

    def __call__(self, s: Scenario, domain=None):
        from .engine import Forcing
        validate(s)
        mmh = 1.0 / 3.6e6
        rain = triangle(0.0, self.storm_s, [self.peak_mmh[s.return_period_yr] * rain_factor(s.rcp) * mmh])
        inflow = triangle(0.0, self.storm_s, [self.inflow_peak[s.return_period_yr]], base=self.inflow_base)
        w = 2 * math.pi / self.tide_period_s
        t_high = 0.5 * self.storm_s + 3600.0 * s.storm_tide_offset_h
        # amp cos(w (t - t_high)): high water at t_high
        tide = Harmonic([self.tide_amp_m], [self.tide_period_s], [0.0], [sea_level_rise(s.rcp, s.horizon_year)],
                        [-t_high], on=[bool(s.tide_on)])
        return Forcing(rain, inflow, tide, self.storm_s + self.recession_s, storm_start=0.0, ssp=s.ssp)


_RESOLVER = [DesignStorm()]


def register_resolver(fn) -> None:
    """Install ``fn(scenario, domain) -> Forcing`` (one member, shared time base) for every later run."""
    _RESOLVER[0] = fn


def resolve(s: Scenario, domain=None):
    validate(s)
    return _RESOLVER[0](s, domain)


def label(s: Scenario) -> str:
    return f"RP{s.return_period_yr} {s.rcp} {s.ssp} {s.horizon_year} tide={'on' if s.tide_on else 'off'}"


def scenario_grid(n: int, seed: int = 0) -> list:
    """n scenarios drawn from the tables (seeded), for batch runs and tests."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append(Scenario(int(rng.choice(RETURN_PERIODS)), str(rng.choice(sorted(RCP_TABLE))),
                            str(rng.choice(sorted(SSP_TABLE))), int(rng.integers(2020, 2101)),
                            bool(rng.random() < 0.8), float(rng.uniform(-3.0, 3.0))))
    return out
