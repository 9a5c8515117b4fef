"""1-D channel network data model: reaches, level-pool junctions, conveyance scaling.

The public types keep hydrointel's names and fields (hydrointel/domain/channels.py:19-72:
``Reach``, ``Junction``, ``ChannelNetwork``) so that networks built for the reference load
here unchanged, and ``validate`` enforces the same rules with the same exception class
(ValueError): per-node fields of the node count and finite, positive width and roughness,
non-negative side slope, a bank crest at least 0.2 m above the bed, known end kinds and
junction indices in range.

A reach end is ("inflow",), ("stage",), ("closed",) or ("junction", j).

``apply_gamma`` scales conveyance per reach (hydrointel/domain/channels.py:172-193): bottom width
and bankfull depth both by sqrt(gamma), the bed moving about the fixed bank crest; every
junction is rescaled once for each reach whose downstream end it is, in reach order.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np

END_KINDS = ("inflow", "stage", "closed", "junction")
NODE_FIELDS = ("bed", "width", "side_slope", "manning", "z_bank")
MIN_BANK_HEIGHT_M = 0.2


@dataclass
class Reach:
    reach_id: str
    x: np.ndarray
    y: np.ndarray
    bed: np.ndarray
    width: np.ndarray
    side_slope: np.ndarray
    manning: np.ndarray
    z_bank: np.ndarray
    up: tuple = ("inflow",)
    down: tuple = ("stage",)

    @property
    def n(self) -> int:
        return len(self.x)

    def _segments(self) -> np.ndarray:
        """Plan-view length of each of the n-1 segments between consecutive nodes."""
        return np.hypot(np.diff(self.x), np.diff(self.y))

    @property
    def spacing(self) -> np.ndarray:
        """Length of each node's cell: the mean of the segments on either side of it, the end
        nodes counting their one segment twice (a single-node reach is 1 m long)."""
        if self.n == 1:
            return np.ones(1)
        seg = self._segments()
        before = np.concatenate((seg[:1], seg))
        after = np.concatenate((seg, seg[-1:]))
        return 0.5 * (before + after)

    @property
    def chainage(self) -> np.ndarray:
        """Distance along the reach from its first node."""
        return np.cumsum(np.concatenate(([0.0], self._segments())))


@dataclass
class Junction:
    x: float
    y: float
    bed: float
    width: float
    side_slope: float
    manning: float
    z_bank: float
    length: float


@dataclass
class ChannelNetwork:
    reaches: list = field(default_factory=list)
    junctions: list = field(default_factory=list)
    inflow_reach: str | None = None

    @property
    def n_reaches(self) -> int:
        return len(self.reaches)

    def reach_index(self, rid: str) -> int:
        for k, r in enumerate(self.reaches):
            if r.reach_id == rid:
                return k
        raise ValueError(f"no reach named {rid!r}")

    def validate(self) -> None:
        for r in self.reaches:
            _check_reach(r, len(self.junctions))


def _check_reach(r: Reach, n_junctions: int) -> None:
    """Raise ValueError naming the reach and the first rule it breaks."""
    def fail(why: str):
        raise ValueError(f"reach {r.reach_id}: {why}")

    for name in NODE_FIELDS:
        a = np.asarray(getattr(r, name))
        if a.shape != np.shape(r.x):
            fail(f"'{name}' holds {a.shape} values for {np.shape(r.x)} nodes")
        if not np.isfinite(a).all():
            fail(f"'{name}' is not finite everywhere")
    if (r.width <= 0).any() or (r.manning <= 0).any() or (r.side_slope < 0).any():
        fail("bottom width and Manning n must be positive and the side slope non-negative")
    if (r.z_bank - r.bed < MIN_BANK_HEIGHT_M).any():
        fail(f"the bank crest must stand at least {MIN_BANK_HEIGHT_M} m above the bed")
    for end in (r.up, r.down):
        if end[0] not in END_KINDS:
            fail(f"end kind {end[0]!r} is not one of {END_KINDS}")
        if end[0] == "junction" and not 0 <= end[1] < n_junctions:
            fail(f"junction {end[1]} does not exist ({n_junctions} junctions)")


def _rescale(width, bed, z_bank, s):
    """Width times s and bankfull depth times s about the bank crest."""
    return width * s, z_bank - (z_bank - bed) * s


def apply_gamma(net: ChannelNetwork, gamma) -> ChannelNetwork:
    """A copy of ``net`` with each reach's conveyance scaled by its gamma (one value per reach, or
    one for all); each junction is rescaled by the factor of every reach ending at it."""
    g = np.asarray(gamma, dtype=float).reshape(-1)
    if g.size == 1 and net.n_reaches != 1:
        g = np.full(net.n_reaches, float(g[0]))
    if g.size != net.n_reaches:
        raise ValueError(f"channel_gamma has {g.size} entries, network has {net.n_reaches} reaches")
    if (g <= 0).any():
        raise ValueError("channel_gamma must be positive")
    s = np.sqrt(g)
    reaches = []
    for r, sr in zip(net.reaches, s):
        width, bed = _rescale(r.width, r.bed, r.z_bank, sr)
        reaches.append(dataclasses.replace(r, x=r.x.copy(), y=r.y.copy(), width=width, bed=bed,
                                           side_slope=r.side_slope.copy(), manning=r.manning.copy(),
                                           z_bank=r.z_bank.copy()))
    junctions = []
    for jk, j in enumerate(net.junctions):
        width, bed = j.width, j.bed
        for r, sr in zip(net.reaches, s):
            if r.down[0] == "junction" and r.down[1] == jk:
                width, bed = _rescale(width, bed, j.z_bank, sr)
        junctions.append(dataclasses.replace(j, width=width, bed=bed))
    out = ChannelNetwork(reaches, junctions, net.inflow_reach)
    out.validate()
    return out
