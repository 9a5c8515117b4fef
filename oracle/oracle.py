"""Python driver of the C oracle -- TEST INFRASTRUCTURE ONLY.

Restates, in float64 on the CPU, everything the coupled flood-engine step needs:

* ``Oracle2D``     the 2-D step and CFL rate (hydrointel/solver/swe2d.py:113-371)
* ``OracleNet``    network topology and sections (hydrointel/solver/swe1d.py:133-231),
                   the 1-D step and CFL rate (swe1d.py:253-368)
* ``oracle_links`` / ``OracleNet.exchange``  links and weir exchange (hydrointel/solver/coupling.py:60-124)
* ``normal_depth`` Manning normal depth by bisection (hydrointel/engine.py:78-101)
* ``OracleCoupled`` the coupled step order and mass ledger (hydrointel/solver/coupling.py:194-227,
                   hydrointel/solver/massbalance.py:30-64)

The per-cell arithmetic lives in swe_oracle.c.  Only tests, the smoke check and the
benchmark's CPU-baseline leg import this module, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBPATH = HERE / "_build" / "liboracle.so"
BC_CODES = {"closed": 0, "transmissive": 1, "stage": 2, "discharge": 3}
EDGES = ("west", "east", "south", "north")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise RuntimeError(f"the oracle library is not built ({LIBPATH}); run python oracle/build.py")
        _lib = C.CDLL(str(LIBPATH))
        _lib.orc_swe2d_step.restype = None
        _lib.orc_swe2d_max_rate.restype = None
        _lib.orc_swe1d_step.restype = None
        _lib.orc_swe1d_max_rate.restype = None
        _lib.orc_exchange.restype = None
    return _lib


P = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)
PL = C.POINTER(C.c_int64)


def _p(a):
    return a.ctypes.data_as(P) if a is not None else None


class _Grid(C.Structure):
    _fields_ = [("B", C.c_int32), ("ny", C.c_int32), ("nx", C.c_int32), ("order", C.c_int32),
                ("flux", C.c_int32), ("bc", C.c_int32 * 4), ("dx", C.c_double), ("dy", C.c_double),
                ("g", C.c_double), ("h_dry", C.c_double), ("kp_eps", C.c_double),
                ("z", P), ("n", P), ("n_bs", C.c_int64), ("S", P), ("S_bs", C.c_int64),
                ("f0", P), ("fc", P), ("k", P), ("hz_bs", C.c_int64), ("t0", C.c_double),
                ("rw", P), ("rw_bs", C.c_int64)]


class _Net(C.Structure):
    _fields_ = [("B", C.c_int32), ("nc", C.c_int32), ("nf", C.c_int32), ("ngh", C.c_int32),
                ("g", C.c_double), ("h_dry", C.c_double),
                ("z", P), ("b", P), ("yb", P), ("ts", P), ("bs", C.c_int64),
                ("m", P), ("nman", P), ("dx", P),
                ("fb", P), ("fyb", P), ("fts", P), ("fbs", C.c_int64), ("fm", P),
                ("fl", PI), ("fr", PI), ("kind", PI), ("g_idx", PI), ("g_src", PI), ("g_sign", P)]


class _Links(C.Structure):
    _fields_ = [("nl", C.c_int32), ("nu", C.c_int32), ("cell1d", PI), ("inv", PI), ("uniq", PL),
                ("zbank", P), ("length", P), ("z2", P), ("area2", C.c_double), ("weir_cw", C.c_double),
                ("limit", C.c_double), ("relax", C.c_double)]


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _field(a, B, shape):
    """(array, member stride) for a field given as (ny,nx) or (B,ny,nx)."""
    if a is None:
        return None, 0
    if np.ndim(a) == 0 or np.size(a) == 1:
        a = np.full(shape, float(np.asarray(a).reshape(-1)[0]))
    a = _f64(a)
    if a.shape == shape:
        return a, 0
    if a.shape == (B,) + shape:
        return a, shape[0] * shape[1]
    raise ValueError(f"field shape {a.shape} does not fit grid {shape} with batch {B}")


class Oracle2D:
    """Float64 CPU 2-D step with the reference's call shape.  ``bcs`` maps edge -> kind name."""

    def __init__(self, z, manning, dx, bcs, B=1, order=2, flux="hllc", g=9.81, h_dry=1e-3, kp_eps=1e-12,
                 storage=None, horton=None, rain_weight=None, dy=None):
        self.z = _f64(z)
        self.ny, self.nx = self.z.shape
        self.B = B
        shp = self.z.shape
        self.n, n_bs = _field(manning, B, shp)
        self.S, S_bs = _field(storage, B, shp)
        if horton is not None:
            f0, fc, k, t0 = horton
            self.f0, hz_bs = _field(f0, B, shp)
            self.fc, _ = _field(fc, B, shp)
            self.k, _ = _field(k, B, shp)
        else:
            self.f0 = self.fc = self.k = None
            hz_bs, t0 = 0, 0.0
        self.rw, rw_bs = _field(rain_weight, B, shp)
        self.dx, self.dy = float(dx), float(dy if dy is not None else dx)
        self.g = g
        kinds = [BC_CODES[bcs[e]] for e in EDGES]
        self.G = _Grid(B, self.ny, self.nx, order, 1 if flux == "hll" else 0, (C.c_int32 * 4)(*kinds),
                       self.dx, self.dy, g, h_dry, kp_eps, _p(self.z), _p(self.n), n_bs, _p(self.S), S_bs,
                       _p(self.f0), _p(self.fc), _p(self.k), hz_bs, float(t0), _p(self.rw), rw_bs)

    def _shape(self, a):
        return _f64(a).reshape(self.B, self.ny, self.nx)

    def step(self, h, hu, hv, s, ic, dt, t, pm):
        h, hu, hv, s, ic = (self._shape(a) for a in (h, hu, hv, s, ic))
        pm = _f64(pm).reshape(9, self.B)
        out = [np.empty_like(h) for _ in range(5)]
        vols = np.zeros((5, self.B))
        lib().orc_swe2d_step(C.byref(self.G), _p(h), _p(hu), _p(hv), _p(s), _p(ic), C.c_double(dt),
                             C.c_double(t), _p(pm), *(_p(o) for o in out), _p(vols))
        return out, vols

    def max_rate(self, h, hu, hv):
        h, hu, hv = (self._shape(a) for a in (h, hu, hv))
        r = np.zeros(self.B)
        lib().orc_swe2d_max_rate(C.byref(self.G), _p(h), _p(hu), _p(hv), _p(r))
        return r


# ------------------------------------------------------------------ 1-D network

def topology(net):
    """swe1d.py:133-166: flatten reaches and junctions into cells and faces."""
    cols = {k: [] for k in ("kind", "reach", "gsrc", "gsign", "z", "b", "m", "zbank", "n", "dx", "x", "y", "chain")}
    fl, fr = [], []

    def add(kind, reach, z, b, m, zb, n, dx, x, y, ch=0.0, gsrc=-1, gsign=0):
        for k, v in zip(cols, (kind, reach, gsrc, gsign, z, b, m, zb, n, dx, x, y, ch)):
            cols[k].append(v)
        return len(cols["kind"]) - 1

    jcell = [add(1, -1, j.bed, j.width, j.side_slope, j.z_bank, j.manning, j.length, j.x, j.y) for j in net.junctions]
    kinds = {"inflow": 2, "stage": 3, "closed": 4}
    for ri, r in enumerate(net.reaches):
        ds, ch = r.spacing, r.chainage
        ids = [add(0, ri, r.bed[k], r.width[k], r.side_slope[k], r.z_bank[k], r.manning[k], ds[k], r.x[k], r.y[k],
                   ch[k]) for k in range(r.n)]
        for a, b in zip(ids[:-1], ids[1:]):
            fl.append(a); fr.append(b)
        for end, first, sign in ((r.up, ids[0], -1), (r.down, ids[-1], +1)):
            if end[0] == "junction":
                other = jcell[end[1]]
            else:
                k = 0 if sign < 0 else r.n - 1
                other = add(kinds[end[0]], ri, r.bed[k], r.width[k], r.side_slope[k], r.z_bank[k], r.manning[k],
                            ds[k], r.x[k], r.y[k], ch[k], gsrc=first, gsign=sign)
            if sign < 0:
                fl.append(other); fr.append(first)
            else:
                fl.append(first); fr.append(other)
    out = {k: np.array(v) for k, v in cols.items()}
    out["fl"], out["fr"] = np.array(fl), np.array(fr)
    return out


class OracleNet:
    """Float64 CPU network (batched over a list of networks that share topology)."""

    def __init__(self, nets, slot_width_frac=1.0, g=9.81, h_dry=1e-3, cfl_1d=0.45, dt_max=5.0):
        nets = list(nets) if isinstance(nets, (list, tuple)) else [nets]
        self.B = len(nets)
        tps = [topology(n) for n in nets]
        tp = self.tp = tps[0]
        self.g, self.h_dry, self.cfl_1d, self.dt_max = g, h_dry, cfl_1d, dt_max
        z = np.stack([t["z"] for t in tps]).astype(float)
        b = np.stack([t["b"] for t in tps]).astype(float)
        m = tp["m"].astype(float)
        zbank = tp["zbank"].astype(float)
        fl, fr = tp["fl"], tp["fr"]
        yb = np.maximum(zbank - z, 0.2)
        ts = slot_width_frac * (b + 2 * m * yb)
        zf = np.maximum(z[..., fl], z[..., fr])
        bank_f = 0.5 * (zbank[fl] + zbank[fr])
        ybf = np.maximum(bank_f - zf, 0.1)
        bf, mf = 0.5 * (b[..., fl] + b[..., fr]), 0.5 * (m[fl] + m[fr])
        fts = slot_width_frac * (bf + 2 * mf * ybf)
        self.arr = dict(z=_f64(z), b=_f64(b), yb=_f64(yb), ts=_f64(ts), m=_f64(m), nman=_f64(tp["n"]),
                        dx=_f64(tp["dx"]), fb=_f64(bf), fyb=_f64(ybf), fts=_f64(fts), fm=_f64(mf))
        kind = tp["kind"].astype(np.int32)
        gmask = kind >= 2
        self.iarr = dict(fl=np.ascontiguousarray(fl, np.int32), fr=np.ascontiguousarray(fr, np.int32),
                         kind=np.ascontiguousarray(kind), g_idx=np.ascontiguousarray(np.nonzero(gmask)[0], np.int32),
                         g_src=np.ascontiguousarray(tp["gsrc"][gmask], np.int32))
        self.g_sign = _f64(tp["gsign"][gmask].astype(float))
        self.nc, self.nf = len(kind), len(fl)
        a, ia = self.arr, self.iarr
        pi = lambda x: x.ctypes.data_as(PI)
        self.N = _Net(self.B, self.nc, self.nf, len(ia["g_idx"]), g, h_dry,
                      _p(a["z"]), _p(a["b"]), _p(a["yb"]), _p(a["ts"]), self.nc, _p(a["m"]), _p(a["nman"]),
                      _p(a["dx"]), _p(a["fb"]), _p(a["fyb"]), _p(a["fts"]), self.nf, _p(a["fm"]),
                      pi(ia["fl"]), pi(ia["fr"]), pi(ia["kind"]), pi(ia["g_idx"]), pi(ia["g_src"]), _p(self.g_sign))
        self.interior = kind <= 1

    def _st(self, a):
        return _f64(a).reshape(self.B, self.nc)

    def area(self, y):
        """Sections.area (swe1d.py:60-63) for every cell."""
        a = self.arr
        yt = np.minimum(y, a["yb"])
        return yt * (a["b"] + a["m"] * yt) + a["ts"] * np.clip(y - a["yb"], 0.0, None)

    def depth(self, A):
        a = self.arr
        A = np.clip(A, 0.0, None)
        ab = a["yb"] * (a["b"] + a["m"] * a["yb"])
        yt = 2.0 * A / (a["b"] + np.sqrt(a["b"] * a["b"] + 4.0 * a["m"] * A))
        return np.where(A <= ab, yt, a["yb"] + (A - ab) / a["ts"])

    def step(self, A, Q, dt, q_in, eta):
        A, Q = self._st(A), self._st(Q)
        q_in = _f64(np.broadcast_to(np.asarray(q_in, float), (self.B,)))
        eta = _f64(np.broadcast_to(np.asarray(eta, float), (self.B,)))
        A1, Q1 = np.empty_like(A), np.empty_like(Q)
        vols = np.zeros((3, self.B))
        lib().orc_swe1d_step(C.byref(self.N), _p(A), _p(Q), C.c_double(dt), _p(q_in), _p(eta), _p(A1), _p(Q1),
                             _p(vols))
        return A1, Q1, vols

    def max_rate(self, A, Q):
        A, Q = self._st(A), self._st(Q)
        r = np.zeros(self.B)
        w = np.zeros(self.B, np.int32)
        lib().orc_swe1d_max_rate(C.byref(self.N), _p(A), _p(Q), _p(r), w.ctypes.data_as(PI))
        return r, w

    def max_dt(self, A, Q):
        """swe1d.py:253-262 over the whole batch."""
        r, w = self.max_rate(A, Q)
        if not w.any():
            return self.dt_max
        return min(self.cfl_1d / float(r.max()), self.dt_max)

    def volume(self, A):
        return (self._st(A) * self.arr["dx"] * self.interior).sum(-1)


def normal_depth(net: OracleNet, Q, slope_min=1e-4):
    """engine.py:78-101, numpy float64."""
    tp = net.tp
    z = net.arr["z"]
    s = np.full(z.shape, slope_min)
    reach = tp["reach"]
    for ri in range(int(reach.max()) + 1 if len(reach) else 0):
        idx = np.nonzero(reach == ri)[0]
        idx = idx[tp["kind"][idx] == 0]
        if len(idx) > 1:
            dz = -np.gradient(z[..., idx], tp["chain"][idx], axis=-1)
            s[..., idx] = np.maximum(dz, slope_min)
    Qa = np.asarray(Q, float).reshape(-1, 1) if np.ndim(Q) else float(Q)
    a = net.arr
    lo, hi = np.zeros_like(s), np.broadcast_to(a["yb"], s.shape).copy()
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        yt = np.minimum(mid, a["yb"])
        A = yt * (a["b"] + a["m"] * yt) + a["ts"] * np.clip(mid - a["yb"], 0.0, None)
        P = a["b"] + 2.0 * yt * np.sqrt(1.0 + a["m"] * a["m"])
        R = A / P
        q = A * R ** (2 / 3) * np.sqrt(s) / a["nman"]
        lo, hi = np.where(q < Qa, mid, lo), np.where(q < Qa, hi, mid)
    return 0.5 * (lo + hi)


def oracle_links(net: OracleNet, z2d, dx2, dy2=None, weir_cw=1.7, limit=0.5, relax=0.5):
    """coupling.py:60-70 plus the exchange constants."""
    dy2 = dy2 or dx2
    tp = net.tp
    ny, nx = np.asarray(z2d).shape
    idx = np.nonzero(tp["kind"] <= 1)[0]
    i = np.clip((tp["x"][idx] / dx2).astype(int), 0, nx - 1)
    j = np.clip((tp["y"][idx] / dy2).astype(int), 0, ny - 1)
    flat = j * nx + i
    uniq, inv = np.unique(flat, return_inverse=True)
    d = dict(cell1d=np.ascontiguousarray(idx, np.int32), inv=np.ascontiguousarray(inv, np.int32),
             uniq=np.ascontiguousarray(uniq, np.int64), zbank=_f64(tp["zbank"][idx]), length=_f64(2.0 * tp["dx"][idx]),
             z2=_f64(np.asarray(z2d).reshape(-1)[flat]))
    L = _Links(len(idx), len(uniq), d["cell1d"].ctypes.data_as(PI), d["inv"].ctypes.data_as(PI),
               d["uniq"].ctypes.data_as(PL), _p(d["zbank"]), _p(d["length"]), _p(d["z2"]), dx2 * dy2, weir_cw,
               limit, relax)
    d["struct"] = L
    return d


def exchange(net: OracleNet, links, h, hu, hv, A, dt):
    """coupling.py:73-124; h/hu/hv (B,ny,nx) and A (B,nc) updated in place.  Returns V (B, links)."""
    B = net.B
    plane = h[0].size
    for a in (h, hu, hv, A):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    V = np.zeros((B, len(links["cell1d"])))
    lib().orc_exchange(C.byref(links["struct"]), C.byref(net.N), C.c_int64(plane), _p(h), _p(hu), _p(hv), _p(A),
                       C.c_double(dt), _p(V))
    return V


TERMS = ("rain", "infiltration", "bnd2d_in", "bnd2d_out", "bnd1d_in", "bnd1d_out", "clip")


class OracleCoupled:
    """coupling.py:194-227: one coupled step = CFL dt (pre-exchange), n_sub, exchange, 2-D step, n_sub 1-D steps."""

    def __init__(self, two: Oracle2D, one: OracleNet | None, links, bcs_values, rain, inflow=None, stage=None,
                 cfl=0.45, dt_min=1e-4, dt_max=5.0, area=None):
        self.two, self.one, self.links = two, one, links
        self.bcv = bcs_values          # edge -> callable(t) -> (B,) or None
        self.rain, self.inflow, self.stage = rain, inflow, stage
        self.cfl, self.dt_min, self.dt_max = cfl, dt_min, dt_max
        B = two.B
        shp = (B, two.ny, two.nx)
        self.h, self.hu, self.hv = np.zeros(shp), np.zeros(shp), np.zeros(shp)
        self.s, self.ic = np.zeros(shp), np.zeros(shp)
        if one is not None:
            self.A, self.Q = np.zeros((B, one.nc)), np.zeros((B, one.nc))
        self.acc = {k: np.zeros(B) for k in TERMS}
        self.ex_in, self.ex_out = np.zeros(B), np.zeros(B)
        self.area2 = two.dx * two.dy
        self.steps = 0
        self.n_dt_clamped = 0

    def _vals(self, t):
        B = self.two.B
        out = []
        for e in EDGES:
            f = self.bcv.get(e)
            out.append(np.broadcast_to(np.asarray(f(t) if f is not None else 0.0, float), (B,)))
        return out

    def max_dt_2d(self):
        m = float(self.two.max_rate(self.h, self.hu, self.hv).max())
        dt = self.cfl / m if m > 0 else self.dt_max
        if dt < self.dt_min or dt > self.dt_max:
            if dt < self.dt_min:
                self.n_dt_clamped += 1
            dt = min(max(dt, self.dt_min), self.dt_max)
        return dt

    def total_volume(self):
        v = self.h.sum(axis=(-2, -1)) * self.area2 + self.s.sum(axis=(-2, -1)) * self.area2
        if self.one is not None:
            v = v + self.one.volume(self.A)
        return v

    def step(self, t, t_limit, dt=None, n_sub=None):
        B = self.two.B
        dt = min(self.max_dt_2d() if dt is None else dt, t_limit - t)
        if self.one is not None:
            if n_sub is None:
                n_sub = max(1, math.ceil(dt / self.one.max_dt(self.A, self.Q) - 1e-9))
            V = exchange(self.one, self.links, self.h, self.hu, self.hv, self.A, dt)
            self.ex_in += np.clip(V, 0, None).sum(-1)
            self.ex_out += np.clip(-V, 0, None).sum(-1)
        pm = np.stack([np.broadcast_to(np.asarray(self.rain(t), float), (B,))] + self._vals(t) + self._vals(t + dt))
        (self.h, self.hu, self.hv, self.s, self.ic), vols = self.two.step(self.h, self.hu, self.hv, self.s, self.ic,
                                                                          dt, t, pm)
        for k, name in enumerate(("bnd2d_in", "bnd2d_out", "rain", "infiltration", "clip")):
            self.acc[name] += vols[k]
        if self.one is not None:
            sub = dt / n_sub
            for k in range(n_sub):
                tk = t + k * sub
                q = self.inflow(tk) if self.inflow is not None else 0.0
                eta = self.stage(tk) if self.stage is not None else 0.0
                self.A, self.Q, v1 = self.one.step(self.A, self.Q, sub, q, eta)
                self.acc["bnd1d_in"] += v1[0]
                self.acc["bnd1d_out"] += v1[1]
                self.acc["clip"] += v1[2]
        self.steps += 1
        self.last = (dt, n_sub)
        return t + dt

    def mass_error(self, v0):
        s = self.acc
        net = s["rain"] - s["infiltration"] + s["bnd2d_in"] - s["bnd2d_out"] + s["bnd1d_in"] - s["bnd1d_out"]
        gross = s["rain"] + s["infiltration"] + s["bnd2d_in"] + s["bnd2d_out"] + s["bnd1d_in"] + s["bnd1d_out"]
        dv = self.total_volume() - v0
        return np.abs(dv - net) / np.maximum(gross, 1e-9 * np.abs(v0) + 1.0)
