"""Land-use classes, their physical parameters, and a seeded land-use map for the synthetic terrain.

The eight classes and their parameters are hydrointel's (hydrointel/domain/landuse.py:13-51), so that a
site described by land use gets the same roughness, retention and Horton infiltration here as there
(roads and open water do not infiltrate; dense urban 2 -> 0.5 mm/h, residential 10 -> 2 mm/h, bare
ground 40 -> 8, grass 60 -> 12, cropland 30 -> 6, mangrove 20 -> 5).  A raster of class codes maps to
per-cell fields with ``lookup``; ``horton_fields`` gives the (f0, fc, k) fields in SI units with the
capacity scaled by an infiltration multiplier kappa (hydrointel/solver/infiltration.py:24-29).

``synthetic_landuse`` paints a map for this package's coastal-plain terrain: water on the sea and under
the channels, a town of dense urban core, residential rings and a road grid beside the river mouth,
mangrove along the shore, and cropland, grass and bare ground inland from a seeded noise field.
"""
from __future__ import annotations

import numpy as np

WATER, ROAD, DENSE_URBAN, RESIDENTIAL, BARE, GRASS, CROPLAND, MANGROVE = range(8)
N_CLASSES = 8
NAMES = ("open water / channel", "paved road", "dense urban", "residential", "bare ground", "grass / park",
         "cropland / ricefield", "mangrove / coastal vegetation")
ATTRS = ("manning", "f0_mmh", "fc_mmh", "k_per_h", "storage_m", "exposure")
# one row per class code, columns ATTRS
TABLE = np.array([
    # manning  f0     fc     k/h   storage  exposure
    [0.030,   0.0,   0.0,   0.0,  0.00,    0.00],   # WATER
    [0.014,   0.0,   0.0,   0.0,  0.00,    0.10],   # ROAD
    [0.050,   2.0,   0.5,   4.0,  0.00,    1.00],   # DENSE_URBAN
    [0.045,  10.0,   2.0,   4.0,  0.00,    0.60],   # RESIDENTIAL
    [0.025,  40.0,   8.0,   3.0,  0.00,    0.02],   # BARE
    [0.035,  60.0,  12.0,   3.0,  0.00,    0.02],   # GRASS
    [0.040,  30.0,   6.0,   2.0,  0.05,    0.05],   # CROPLAND
    [0.120,  20.0,   5.0,   2.0,  0.00,    0.01],   # MANGROVE
], dtype=np.float64)
MMH = 1.0 / 3.6e6


def lookup(raster, attr: str) -> np.ndarray:
    """The per-cell value of one class parameter; ValueError for a code outside the table."""
    codes = np.asarray(raster)
    if codes.size and (codes.min() < 0 or codes.max() >= N_CLASSES):
        raise ValueError(f"land-use codes must lie in 0..{N_CLASSES - 1}; found {codes.min()}..{codes.max()}")
    if attr not in ATTRS:
        raise KeyError(attr)
    return TABLE[:, ATTRS.index(attr)][codes.astype(np.int64)]


def horton_fields(raster, kappa=1.0):
    """(f0 [m/s], fc [m/s], k [1/s]) per cell; kappa scales the two capacities, not the decay."""
    kap = np.asarray(kappa, dtype=np.float64)
    return (lookup(raster, "f0_mmh") * MMH * kap, lookup(raster, "fc_mmh") * MMH * kap,
            lookup(raster, "k_per_h") / 3600.0)


def synthetic_landuse(terrain, channel_mask, seed: int = 7) -> np.ndarray:
    """A class map for a coastal-plain terrain whose sea lies along its last columns (int8, (ny, nx))."""
    ny, nx = terrain.shape
    dx = terrain.dx
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dx
    X, Y = np.meshgrid(x, y)
    shore_x = x[~terrain.sea.all(axis=0)].max() if (~terrain.sea).any() else x[-1]
    d_shore = shore_x - X
    # the town: centred on the river mouth, 1.2 km inland
    cy = float(np.asarray(terrain.centerline(shore_x)))
    r_town = np.hypot(X - (shore_x - 1200.0), Y - cy)
    # smooth noise for the rural mosaic: a few random Fourier modes
    noise = np.zeros_like(X)
    for _ in range(6):
        kx, ky = rng.uniform(-1.0, 1.0, 2) * 2 * np.pi / 2500.0
        noise += np.cos(kx * X + ky * Y + rng.uniform(0, 2 * np.pi))
    lu = np.full((ny, nx), CROPLAND, np.int8)
    lu[noise > 1.0] = GRASS
    lu[noise < -1.6] = BARE
    lu[r_town < 1400.0] = RESIDENTIAL
    lu[r_town < 600.0] = DENSE_URBAN
    grid = (np.mod(X - shore_x, 400.0) < dx) | (np.mod(Y - cy, 400.0) < dx)
    lu[(r_town < 1400.0) & grid] = ROAD
    lu[(d_shore < 250.0) & (np.abs(Y - cy) > 300.0)] = MANGROVE
    lu[np.asarray(channel_mask, bool)] = WATER
    lu[terrain.sea] = WATER
    return lu
