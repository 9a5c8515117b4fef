"""The site a run simulates: terrain, channel network and land use on one grid.

hydrointel's runs take a ``Domain`` (hydrointel/domain/synthetic.py) built once per configuration;
its engine and API read the bed ``dem2d``, the ``sea`` and ``channel`` masks, the grid (``dx``,
``nx``, ``ny``, ``shape``), the channel ``network`` and the ``landuse`` class raster.  ``Domain``
here carries the same attributes over this package's terrain, plus the edge the sea lies on (the
edge that takes the tidal stage; the synthetic coastal plain has its sea along the east edge).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import landuse as LU


@dataclass
class Domain:
    terrain: object                   # his_amd.workloads.Terrain (z, sea, dx, nx, ny, centerline)
    network: object                   # ChannelNetwork
    landuse: np.ndarray               # (ny, nx) class codes
    sea_edge: str = "east"
    name: str = ""
    _channel: np.ndarray | None = field(default=None, repr=False)

    @property
    def dem2d(self) -> np.ndarray:
        return self.terrain.z

    @property
    def sea(self) -> np.ndarray:
        return self.terrain.sea

    @property
    def dx(self) -> float:
        return self.terrain.dx

    @property
    def nx(self) -> int:
        return self.terrain.nx

    @property
    def ny(self) -> int:
        return self.terrain.ny

    @property
    def shape(self):
        return self.terrain.shape

    @property
    def channel(self) -> np.ndarray:
        """2-D cells that hold an interior 1-D node (the channel footprint)."""
        if self._channel is None:
            self._channel = channel_mask(self.terrain, self.network)
        return self._channel

    @property
    def land(self) -> np.ndarray:
        return ~(self.sea | self.channel)


def channel_mask(terrain, network) -> np.ndarray:
    from ..solver.swe1d import KIND_JUNCTION, build_topology
    tp = build_topology(network)
    inner = tp.kind <= KIND_JUNCTION
    i = np.clip((tp.x[inner] / terrain.dx).astype(int), 0, terrain.nx - 1)
    j = np.clip((tp.y[inner] / terrain.dx).astype(int), 0, terrain.ny - 1)
    m = np.zeros(terrain.shape, bool)
    m[j, i] = True
    return m


def domain_from(terrain, network, seed: int = 7, name: str = "") -> Domain:
    """A domain over a synthetic terrain, with the seeded land-use map painted onto it."""
    ch = channel_mask(terrain, network)
    return Domain(terrain, network, LU.synthetic_landuse(terrain, ch, seed), "east", name, ch)
