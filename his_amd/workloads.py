"""Seeded synthetic workloads for the coupled flood-engine step.

Five workloads are defined, all on an 8 km x 6.5 km coastal plain:

  plain_cpu     100 x 81 cells at 80 m, one storm, float64
  plain_gpu     400 x 325 cells at 20 m, one storm, float64
  precision     400 x 325 at 20 m, 32 members, float32 bulk
  batching      16 x (100 x 81 @ 80 m) + 8 x (200 x 163 @ 40 m) + 8 x (400 x 325 @ 20 m)
  serving       Poisson(4/s) requests over the three sizes

Terrain: a plane rising 0.002 westward from the shore, the sea on the last five
columns (bed held below -1 m there), 12 Gaussian hills (amplitude U[2,15] m,
sigma U[200,800] m) and a sine-meander trench 3 m deep and 40 m wide running
west to east into the sea.  The hills and the meander are drawn once from the
seed in physical coordinates, so every resolution samples the same landscape.
On grids coarser than 53 m the trench cross-section is widened to 0.75 cell so
that it stays a connected channel.

Channel networks follow the meander.  ``chain`` joins n reaches end to end
through level-pool junctions (upstream inflow, tidal stage at the mouth).
``network60`` is a 3-reach main stem split by 2 junctions, one closed-headed
tributary at each junction and 55 coastal drains (closed head, tidal outlet),
60 reaches in all.  Every interior 1-D node is weir-coupled to the 2-D cell that
contains it, whose bed is the bank crest.

Forcing: a triangular hyetograph, a sinusoidal tide (period 12.42 h) on the
sea edge and at every stage end, and a triangular upstream inflow hydrograph.
Per-member draws follow one seeded ``numpy.random.default_rng``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .domain.channels import ChannelNetwork, Junction, Reach
from .forcing import sine_tide, triangle

LX, LY = 8000.0, 6500.0
SEA_COLS = 5
TIDE_PERIOD_S = 12.42 * 3600.0
MMH = 1.0 / 3.6e6


@dataclass
class Terrain:
    nx: int
    ny: int
    dx: float
    z: np.ndarray              # (ny, nx)
    sea: np.ndarray            # (ny, nx) bool
    meander: tuple             # (amp, wavelength, phase, y0)

    @property
    def shape(self):
        return (self.ny, self.nx)

    def centerline(self, x):
        a, lam, ph, y0 = self.meander
        return y0 + a * np.sin(2 * np.pi * np.asarray(x) / lam + ph)


def _landscape(seed: int):
    rng = np.random.default_rng(seed)
    hills = [(rng.uniform(2.0, 15.0), rng.uniform(200.0, 800.0), rng.uniform(0.0, LX - 1500.0),
              rng.uniform(0.0, LY)) for _ in range(12)]
    meander = (rng.uniform(300.0, 500.0), rng.uniform(2000.0, 3000.0), rng.uniform(0, 2 * np.pi),
               LY * rng.uniform(0.4, 0.6))
    return hills, meander


def make_terrain(nx: int, ny: int, dx: float, seed: int = 20260917) -> Terrain:
    hills, meander = _landscape(seed)
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dx
    X, Y = np.meshgrid(x, y)
    x_shore = (nx - SEA_COLS) * dx
    z = 0.002 * (x_shore - X)
    for a, s, cx, cy in hills:
        z = z + a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * s * s))
    a, lam, ph, y0 = meander
    yc = y0 + a * np.sin(2 * np.pi * X / lam + ph)
    half = max(20.0, 0.75 * dx)
    r = np.abs(Y - yc) / half
    z = z - 3.0 * np.clip(1.0 - r * r, 0.0, None)
    sea = np.zeros((ny, nx), bool)
    sea[:, nx - SEA_COLS:] = True
    z[sea] = np.minimum(z[sea], -1.0 - 0.002 * (X[sea] - x_shore))
    return Terrain(nx, ny, dx, z, sea, meander)


def _cell_z(ter: Terrain, x, y):
    i = np.clip((np.asarray(x) / ter.dx).astype(int), 0, ter.nx - 1)
    j = np.clip((np.asarray(y) / ter.dx).astype(int), 0, ter.ny - 1)
    return ter.z[j, i]


def _reach(rid, xs, ys, ter, depth, width, side, n, up, down):
    zb = _cell_z(ter, xs, ys)
    k = len(xs)
    return Reach(rid, np.asarray(xs, float), np.asarray(ys, float), zb - depth, np.full(k, width),
                 np.full(k, side), np.full(k, n), zb.astype(float), up, down)


def _junction(ter, x, y, depth, width, side, n):
    zb = float(_cell_z(ter, np.array([x]), np.array([y]))[0])
    return Junction(float(x), float(y), zb - depth, width, side, n, zb, 2.0 * width)


def chain_network(ter: Terrain, n_reaches: int = 20, width: float = 20.0, side: float = 2.0,
                  n: float = 0.03, depth: float = 2.5, spacing: float | None = None) -> ChannelNetwork:
    """n reaches end to end along the meander, joined by n-1 level-pool junctions."""
    ds = spacing or ter.dx
    x_end = (ter.nx - SEA_COLS) * ter.dx - 0.5 * ds
    xs = np.arange(ds, x_end, ds)
    ys = ter.centerline(xs)
    npts = len(xs)
    nj = n_reaches - 1
    per = (npts - nj) // n_reaches
    if per < 2:
        raise ValueError(f"{npts} channel points cannot hold {n_reaches} reaches of >= 2 nodes")
    net = ChannelNetwork()
    k = 0
    for r in range(n_reaches):
        cnt = per if r < n_reaches - 1 else npts - k
        up = ("inflow",) if r == 0 else ("junction", r - 1)
        down = ("stage",) if r == n_reaches - 1 else ("junction", r)
        net.reaches.append(_reach(f"r{r:02d}", xs[k:k + cnt], ys[k:k + cnt], ter, depth, width, side, n, up, down))
        k += cnt
        if r < n_reaches - 1:
            net.junctions.append(_junction(ter, xs[k], ys[k], depth, width, side, n))
            k += 1
    net.inflow_reach = "r00"
    net.validate()
    return net


def network60(ter: Terrain, width: float = 20.0, side: float = 2.0, n: float = 0.03, depth: float = 2.5,
              n_drains: int = 55) -> ChannelNetwork:
    """Main stem in 3 reaches split by 2 junctions, a tributary into each junction, and
    ``n_drains`` coastal drains with closed heads and tidal outlets."""
    ds = ter.dx
    x_shore = (ter.nx - SEA_COLS) * ter.dx
    xs = np.arange(ds, x_shore - 0.5 * ds, ds)
    ys = ter.centerline(xs)
    npts = len(xs)
    j0, j1 = npts // 3, (2 * npts) // 3
    net = ChannelNetwork()
    net.junctions.append(_junction(ter, xs[j0], ys[j0], depth, width, side, n))
    net.junctions.append(_junction(ter, xs[j1], ys[j1], depth, width, side, n))
    net.reaches.append(_reach("main0", xs[:j0], ys[:j0], ter, depth, width, side, n, ("inflow",), ("junction", 0)))
    net.reaches.append(_reach("main1", xs[j0 + 1:j1], ys[j0 + 1:j1], ter, depth, width, side, n,
                              ("junction", 0), ("junction", 1)))
    net.reaches.append(_reach("main2", xs[j1 + 1:], ys[j1 + 1:], ter, depth, width, side, n,
                              ("junction", 1), ("stage",)))
    for k, jj in enumerate((j0, j1)):
        sgn = 1.0 if k == 0 else -1.0
        ty = ys[jj] + sgn * (np.arange(30, 0, -1) * ds)
        tx = np.full(ty.shape, xs[jj])
        keep = (ty > ds) & (ty < ter.ny * ter.dx - ds)
        net.reaches.append(_reach(f"trib{k}", tx[keep], ty[keep], ter, 1.5, 6.0, 1.5, 0.035, ("closed",),
                                  ("junction", k)))
    # coastal drains: straight west -> east lines across the last 1.5 km, clear of the meander band
    dxs = np.arange(x_shore - 1500.0, x_shore - 0.5 * ds, ds)
    band = ter.centerline(dxs)
    ly = ter.ny * ter.dx
    cand = np.linspace(2 * ds, ly - 2 * ds, 4 * n_drains)
    ok = [y for y in cand if np.all(np.abs(y - band) > 3 * ds + 40.0)]
    if len(ok) < n_drains:
        raise ValueError(f"only {len(ok)} drain positions clear the channel band")
    pick = np.asarray(ok)[np.linspace(0, len(ok) - 1, n_drains).round().astype(int)]
    for k, yd in enumerate(pick):
        net.reaches.append(_reach(f"drain{k:02d}", dxs, np.full(dxs.shape, yd), ter, 1.5, 4.0, 1.5, 0.035,
                                  ("closed",), ("stage",)))
    net.inflow_reach = "main0"
    net.validate()
    return net


@dataclass
class Member:
    """One storm's per-member parameters (the precision-form draws)."""
    manning: float = 0.035
    retention_m: float = 0.005
    f0_mmh: float = 60.0
    fc_mmh: float = 10.0
    k_per_h: float = 2.0
    gamma: float = 1.0
    rain_peak_mmh: float = 80.0
    tide_phase: float = 0.0
    inflow_peak: float = 300.0


def draw_members(B: int, seed: int) -> list:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(B):
        out.append(Member(manning=rng.uniform(0.02, 0.08), retention_m=rng.uniform(2.0, 10.0) * 1e-3,
                          f0_mmh=rng.uniform(30.0, 90.0), gamma=rng.uniform(0.7, 1.3),
                          rain_peak_mmh=rng.uniform(40.0, 120.0), tide_phase=rng.uniform(0.0, 2 * np.pi),
                          inflow_peak=rng.uniform(100.0, 500.0)))
    return out


@dataclass
class Storm:
    storm_s: float          # hyetograph duration
    recession_s: float
    inflow_base: float = 10.0
    tide_amp: float = 0.8

    @property
    def t_end(self):
        return self.storm_s + self.recession_s


@dataclass
class Workload:
    name: str
    terrain: Terrain
    network: ChannelNetwork
    members: list
    storm: Storm
    precision: str = "fp64"
    slot_width_frac: float = 1.0
    extra: dict = field(default_factory=dict)

    @property
    def batch(self):
        return len(self.members)

    @property
    def cells(self):
        return self.terrain.nx * self.terrain.ny

    def fields(self):
        """Per-member (B, ny, nx) float64 fields: manning, storage capacity, Horton f0, fc, k (SI)."""
        ter, B = self.terrain, self.batch
        land = ~ter.sea
        shp = (B,) + ter.shape
        n = np.empty(shp); S = np.empty(shp); f0 = np.empty(shp); fc = np.empty(shp); k = np.empty(shp)
        for b, m in enumerate(self.members):
            n[b] = m.manning
            S[b] = np.where(land, m.retention_m, 0.0)
            f0[b] = np.where(land, m.f0_mmh * MMH, 0.0)
            fc[b] = np.where(land, m.fc_mmh * MMH, 0.0)
            k[b] = np.where(land, m.k_per_h / 3600.0, 0.0)
        return n, S, f0, fc, k

    def forcing(self, batched: bool = True):
        """(rain [m/s], tide [m], inflow [m^3/s]) series for every member."""
        st = self.storm
        peaks = np.array([m.rain_peak_mmh * MMH for m in self.members])
        rain = triangle(0.0, st.storm_s, peaks, batched=batched)
        tide = sine_tide(st.tide_amp, TIDE_PERIOD_S, [m.tide_phase for m in self.members], batched=batched)
        qpk = np.array([m.inflow_peak for m in self.members])
        inflow = triangle(0.0, st.storm_s, qpk, base=st.inflow_base, batched=batched)
        return rain, tide, inflow


def plain_cpu(seed: int = 20260917) -> Workload:
    ter = make_terrain(100, 81, 80.0, seed)
    m = Member(manning=0.035, retention_m=0.005, f0_mmh=60.0, fc_mmh=10.0, k_per_h=2.0, gamma=1.0,
               rain_peak_mmh=80.0, tide_phase=0.0, inflow_peak=100.0)
    return Workload("plain_cpu 100x81@80m B=1 fp64", ter, chain_network(ter, 20), [m],
                    Storm(1800.0, 5400.0, inflow_base=5.0))


def plain_gpu(seed: int = 20260917, rain_seed: int = 1) -> Workload:
    ter = make_terrain(400, 325, 20.0, seed)
    rng = np.random.default_rng(rain_seed)
    m = Member(rain_peak_mmh=float(rng.uniform(40.0, 120.0)), inflow_peak=300.0)
    return Workload("plain_gpu 400x325@20m B=1 fp64", ter, network60(ter), [m], Storm(6 * 3600.0, 2.5 * 3600.0),
                    slot_width_frac=0.25)


def precision_form(seed: int = 20260917, B: int = 32, member_seed: int = 3) -> Workload:
    ter = make_terrain(400, 325, 20.0, seed)
    return Workload(f"precision 400x325@20m B={B} fp32bulk", ter, network60(ter), draw_members(B, member_seed),
                    Storm(6 * 3600.0, 2.5 * 3600.0), precision="fp32bulk", slot_width_frac=0.25)


def batching_form(seed: int = 20260917, member_seed: int = 4, precision: str = "fp32bulk") -> list:
    """Three groups at their own resolutions; each shares dt within itself."""
    out = []
    for (nx, ny, dx, B, net) in ((100, 81, 80.0, 16, "chain"), (200, 163, 40.0, 8, "net60"),
                                 (400, 325, 20.0, 8, "net60")):
        ter = make_terrain(nx, ny, dx, seed)
        netw = chain_network(ter, 20) if net == "chain" else network60(ter)
        out.append(Workload(f"group {nx}x{ny}@{dx:g}m B={B}", ter, netw, draw_members(B, member_seed + nx),
                            Storm(6 * 3600.0, 2.5 * 3600.0), precision=precision,
                            slot_width_frac=1.0 if net == "chain" else 0.25))
    return out


SERVING_SIZES = ((100, 81, 80.0), (200, 163, 40.0), (400, 325, 20.0))


def serving_requests(n: int = 2000, rate_per_s: float = 4.0, seed: int = 5):
    """Poisson arrivals; each request: grid size index, storm duration U[0.5,6] h, member draws."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, n))
    sizes = rng.integers(0, 3, n)
    dur = rng.uniform(0.5, 6.0, n) * 3600.0
    members = draw_members(n, seed + 1)
    return [dict(id=i, arrival_s=float(t[i]), size=int(sizes[i]), storm_s=float(dur[i]), member=members[i])
            for i in range(n)]

