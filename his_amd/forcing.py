"""Per-member forcing series that the device evaluates inside the coupled step.

The reference evaluates rainfall, tide and upstream inflow in host numpy every
step (hydrointel/engine.py:104-166, ``BatchForcing``), which forces a host round
trip per step.  Here the same three functional forms are kept as small tables
that a kernel evaluates with the same float64 arithmetic, in the same order:

* ``PiecewiseLinear`` -- linear interpolation on a uniform time base with the end
  values held outside it: the arithmetic of ``BatchForcing.inflow``
  (hydrointel/engine.py:158-166).
* ``PiecewiseConstant`` -- blocks on a uniform base, zero outside:
  ``BatchForcing.rain_rate`` (hydrointel/engine.py:145-149).
* ``Harmonic`` -- base + sum_c amp_c cos(w_c (t + shift_b) - phase_c), switched per
  member: ``BatchForcing.tide`` (hydrointel/engine.py:151-156).

Every series is callable on the host, returning one float64 value per member (a
float when unbatched), so it also serves as the ``EdgeBC.value`` / ``inflow`` /
``stage`` / ``rain`` callable the reference's solver classes take.
"""
from __future__ import annotations

import math

import numpy as np

KIND_CONST, KIND_PL, KIND_PC, KIND_HARM = 0, 1, 2, 3


class Series:
    kind = KIND_CONST
    batched = True

    def __call__(self, t: float):
        v = self.eval(float(t))
        return v if self.batched else float(v[0])

    def eval(self, t: float) -> np.ndarray:          # pragma: no cover - abstract
        raise NotImplementedError

    # device table: (kind, K, t0, step, values(B*K), base(B), shift(B), on(B), amp(K), w(K), phase(K))
    def table(self):
        raise NotImplementedError


class Constant(Series):
    kind = KIND_CONST

    def __init__(self, values, batched: bool = True):
        self.values = np.atleast_1d(np.asarray(values, dtype=np.float64)).copy()
        self.batched = batched
        self.n = self.values.size

    def eval(self, t):
        return self.values.copy()


class PiecewiseLinear(Series):
    kind = KIND_PL

    def __init__(self, t0: float, step: float, values, batched: bool = True):
        v = np.asarray(values, dtype=np.float64)
        self.values = np.atleast_2d(v).copy()
        if step <= 0 or self.values.shape[1] < 1:
            raise ValueError("piecewise-linear series needs step > 0 and at least one knot")
        self.t0, self.step = float(t0), float(step)
        self.batched = batched
        self.n = self.values.shape[0]

    def eval(self, t):
        x = (t - self.t0) / self.step
        q = self.values
        if x <= 0:
            return q[:, 0].copy()
        if x >= q.shape[1] - 1:
            return q[:, -1].copy()
        k = int(x)
        w = x - k
        return (1 - w) * q[:, k] + w * q[:, k + 1]


class PiecewiseConstant(Series):
    kind = KIND_PC

    def __init__(self, t0: float, step: float, values, batched: bool = True):
        self.values = np.atleast_2d(np.asarray(values, dtype=np.float64)).copy()
        if step <= 0:
            raise ValueError("piecewise-constant series needs step > 0")
        self.t0, self.step = float(t0), float(step)
        self.batched = batched
        self.n = self.values.shape[0]

    def eval(self, t):
        k = int(np.floor((t - self.t0) / self.step + 1e-9))
        if k < 0 or k >= self.values.shape[1]:
            return np.zeros(self.n)
        return self.values[:, k].copy()


class Harmonic(Series):
    kind = KIND_HARM

    def __init__(self, amp, period_s, phase, base, shift, on=None, batched: bool = True):
        self.amp = np.atleast_1d(np.asarray(amp, dtype=np.float64)).copy()
        self.w = np.array([2 * np.pi / float(p) for p in np.atleast_1d(period_s)])
        self.phase = np.atleast_1d(np.asarray(phase, dtype=np.float64)).copy()
        self.base = np.atleast_1d(np.asarray(base, dtype=np.float64)).copy()
        self.shift = np.atleast_1d(np.asarray(shift, dtype=np.float64)).copy()
        self.n = max(self.base.size, self.shift.size)
        self.base = np.broadcast_to(self.base, (self.n,)).copy()
        self.shift = np.broadcast_to(self.shift, (self.n,)).copy()
        self.on = np.ones(self.n, bool) if on is None else np.broadcast_to(np.asarray(on, bool), (self.n,)).copy()
        self.batched = batched

    def eval(self, t):
        tt = t + self.shift
        eta = np.zeros(self.n)
        for a, w, p in zip(self.amp, self.w, self.phase):
            eta = eta + a * np.cos(w * tt - p)
        return self.base + np.where(self.on, eta, 0.0)


def sine_tide(amplitude: float, period_s: float, phase, batched: bool = True) -> Harmonic:
    """amplitude * sin(2 pi t / period + phase), one phase per member."""
    ph = np.atleast_1d(np.asarray(phase, dtype=np.float64))
    w = 2 * np.pi / period_s
    return Harmonic([amplitude], [period_s], [0.5 * math.pi], np.zeros(ph.size), ph / w, batched=batched)


def triangle(t_start: float, duration: float, peak, base=0.0, batched: bool = True) -> PiecewiseLinear:
    """Triangle from ``base`` at t_start up to ``peak`` at mid-duration and back, ``base`` outside."""
    pk = np.atleast_1d(np.asarray(peak, dtype=np.float64))
    bs = np.broadcast_to(np.asarray(base, dtype=np.float64), pk.shape)
    return PiecewiseLinear(t_start, 0.5 * duration, np.stack([bs, pk, bs], axis=1), batched=batched)


def stack(members: list) -> Series:
    """One batched series from one-member series of the same form on a shared time base (the same
    knots, or the same tidal constituents): hydrointel's BatchForcing enforces a shared time base too
    (hydrointel/engine.py:104-143).  ValueError when the members do not share one."""
    first = members[0]
    kind = type(first)
    if any(type(m) is not kind for m in members):
        raise ValueError("batched forcing members mix series forms")
    if isinstance(first, (PiecewiseLinear, PiecewiseConstant)):
        if any(m.t0 != first.t0 or m.step != first.step or m.values.shape[1] != first.values.shape[1]
               for m in members):
            raise ValueError("batched forcing members do not share a time base (t0, step, knots)")
        return kind(first.t0, first.step, np.concatenate([m.values for m in members], axis=0))
    if isinstance(first, Harmonic):
        if any(not (np.array_equal(m.amp, first.amp) and np.array_equal(m.w, first.w)
                    and np.array_equal(m.phase, first.phase)) for m in members):
            raise ValueError("batched tide members do not share their constituents")
        out = Harmonic(first.amp, [2 * np.pi / w for w in first.w], first.phase,
                       np.concatenate([m.base for m in members]), np.concatenate([m.shift for m in members]),
                       np.concatenate([m.on for m in members]))
        out.w = first.w.copy()
        return out
    return Constant(np.concatenate([np.atleast_1d(m.values) for m in members]))
