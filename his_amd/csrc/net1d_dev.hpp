// Device helpers of the 1-D network and the weir exchange, shared by the network kernels and the
// engine's fused step head: trapezoid-and-slot sections (hydrointel/solver/swe1d.py:50-84) and
// the per-item arithmetic of the exchange (hydrointel/solver/coupling.py:73-124).
#pragma once
#include "net1d.hpp"

namespace his {

struct Sec { double b, m, yb, ts; };

__device__ __forceinline__ double bank_area(const Sec& s) { return s.yb * (s.b + s.m * s.yb); }
__device__ __forceinline__ double sec_area(const Sec& s, double y) {
    double yt = tmin(y, s.yb);
    double at = yt * (s.b + s.m * yt);
    return at + s.ts * clamp0(y - s.yb);
}
__device__ __forceinline__ double sec_top(const Sec& s, double y) { return y <= s.yb ? s.b + 2.0 * s.m * y : s.ts; }
__device__ __forceinline__ double sec_i1(const Sec& s, double y) {
    double yt = tmin(y, s.yb);
    double above = clamp0(y - s.yb);
    double i_t = s.b * yt * yt / 2.0 + s.m * (yt * yt * yt) * (1.0 / 3.0);   // a product, not a division
    double at = yt * (s.b + s.m * yt);
    return i_t + at * above + s.ts * above * above / 2.0;
}
__device__ __forceinline__ double sec_perim(const Sec& s, double y) {
    double yt = tmin(y, s.yb);
    return s.b + 2.0 * yt * sqrt(1.0 + s.m * s.m);
}
// one division for both branches: below the bank y = 0 + 2A / (b + sqrt(b^2 + 4mA)), above it
// y = yb + (A - ab) / ts; adding 0.0 is exact, so the result is that of the two-branch form
__device__ __forceinline__ double sec_depth(const Sec& s, double A) {
    A = clamp0(A);
    const double ab = bank_area(s);
    const bool below = A <= ab;
    const double num = below ? 2.0 * A : A - ab;
    const double den = below ? s.b + sqrt(s.b * s.b + 4.0 * s.m * A) : s.ts;
    return (below ? 0.0 : s.yb) + num / den;
}
__device__ __forceinline__ Sec cell_sec(const his_net1d& N, int b, int c) {
    const int64_t o = (int64_t)b * N.bs + c;
    return Sec{N.b[o], N.m[c], N.yb[o], N.ts[o]};
}
__device__ __forceinline__ Sec face_sec(const his_net1d& N, int b, int f) {
    const int64_t o = (int64_t)b * N.fbs + f;
    return Sec{N.fb[o], N.fm[f], N.fyb[o], N.fts[o]};
}
__device__ __forceinline__ double vel1(const Sec& s, double A, double Q, double h_dry) {
    double y = sec_depth(s, A);
    return y >= h_dry ? Q / (A < 1e-12 ? 1e-12 : A) : 0.0;
}

struct ExWs {
    double *V, *sg, *Vf, *hu_, *av1;
};

__device__ __forceinline__ ExWs ex_ws(const ExchangeArgs& a, int b) {
    const int nl = a.L.nl, nu = a.L.nu;
    double* w = a.ws + (int64_t)b * exchange_ws_stride(a.L, a.N);
    return ExWs{w, w + nl, w + 2 * nl, w + 3 * nl, w + 3 * nl + nu};
}

// donor limit factor of one target: the links' gross give `tot` scaled down to `av` when it exceeds it
__device__ __forceinline__ double ex_factor(double tot, double av) {
    return tot > av ? av / (tot > 0.0 ? tot : 1.0) : 1.0;
}
// limit factor of unique 2-D cell k from the gross volume its links give to the 1-D side
__device__ __forceinline__ double ex_f2(const ExWs& w, const his_links& L, int k) {
    double tot = 0.0;
    for (int p = L.u_ptr[k]; p < L.u_ptr[k + 1]; ++p) {
        const int l = L.u_link[p];
        tot += w.sg[l] > 0.0 ? w.V[l] : 0.0;
    }
    return ex_factor(tot, L.limit * w.hu_[k] * L.area2);
}
// limit factor of 1-D cell c from the gross volume its links give to the 2-D side
__device__ __forceinline__ double ex_f1(const ExWs& w, const his_links& L, int c) {
    double tot = 0.0;
    for (int p = L.c_ptr[c]; p < L.c_ptr[c + 1]; ++p) {
        const int l = L.c_link[p];
        tot += w.sg[l] < 0.0 ? w.V[l] : 0.0;
    }
    return ex_factor(tot, w.av1[c]);
}

// Exchange phase 1 for member b, item r: link r (r < nl) -> free-weir volume capped by the
// level-equalising volume, and its sign; the first link of each 1-D cell also stores that cell's
// available donor volume; unique 2-D cell r - nl -> its donor depth.  Phase 2 reads only these, so
// its in-place updates of h and A never feed another item's limit.
template <typename R2>
__device__ __forceinline__ void ex_link_item(const ExchangeArgs& a, const int b, const int r, const double dt) {
    const his_links& L = a.L;
    const his_net1d& N = a.N;
    const R2* h = reinterpret_cast<const R2*>(a.h) + (int64_t)b * a.plane;
    ExWs w = ex_ws(a, b);
    if (r >= L.nl) {
        const int k = r - L.nl;
        w.hu_[k] = (double)h[L.uniq[k]];
        return;
    }
    const int l = r;
    const int c1 = L.cell1d[l];
    const Sec s = cell_sec(N, b, c1);
    const double h2 = (double)h[L.uniq[L.inv[l]]];
    const double eta2 = L.z2[l] + h2;
    const double A1 = a.A[(int64_t)b * N.nc + c1];
    const double y1 = sec_depth(s, A1);
    const double eta1 = N.z[(int64_t)b * N.bs + c1] + y1;
    const double up = tmax(eta1, eta2), dn = tmin(eta1, eta2);
    const double d_up = clamp0(up - L.zbank[l]), d_dn = clamp0(dn - L.zbank[l]);
    const double ratio = d_up > 0.0 ? d_dn / (d_up > 0.0 ? d_up : 1.0) : 0.0;
    const double q = L.weir_cw * L.length[l] * (d_up * sqrt(d_up)) * pow(clamp0(1.0 - ratio * sqrt(ratio)), 0.385);
    w.sg[l] = eta2 > eta1 ? 1.0 : -1.0;
    const double v = q * dt;
    const double plan1 = (y1 <= s.yb ? s.b + 2 * s.m * y1 : s.ts) * N.dx[c1];
    const double v_eq = (up - dn) / (1.0 / L.area2 + 1.0 / plan1);
    w.V[l] = tmin(v, L.relax * v_eq);
    if (L.c_link[L.c_ptr[c1]] == l) w.av1[c1] = L.limit * clamp0(A1 - bank_area(s)) * N.dx[c1];
}

// Single-phase exchange of unique 2-D cell k of member b, for link tables in which every 1-D cell
// has at most one link (one link per interior 1-D cell, as hydrointel/solver/coupling.py
// build_links makes them): the cell's links, both donor limits and both applications are this
// item's alone, so one thread does what the two phases do, with the same operations in the same
// order (identical results).  The workspace receives V, sign, 1-D donor volume and final V as the
// two-phase form leaves them.
template <typename R2>
__device__ __forceinline__ void ex_one_item(const ExchangeArgs& a, const int b, const int k, const double dt) {
    const his_links& L = a.L;
    const his_net1d& N = a.N;
    ExWs w = ex_ws(a, b);
    R2* h = reinterpret_cast<R2*>(a.h);
    const int64_t q = (int64_t)b * a.plane + L.uniq[k];
    const double h2 = (double)h[q];
    const int p0 = L.u_ptr[k], p1 = L.u_ptr[k + 1];
    double* A = a.A + (int64_t)b * N.nc;
    // phase 1: each link's free-weir volume capped by the level-equalising volume, its sign, and
    // its 1-D cell's available donor volume
    for (int p = p0; p < p1; ++p) {
        const int l = L.u_link[p];
        const int c1 = L.cell1d[l];
        const Sec s = cell_sec(N, b, c1);
        const double eta2 = L.z2[l] + h2;
        const double A1 = A[c1];
        const double y1 = sec_depth(s, A1);
        const double eta1 = N.z[(int64_t)b * N.bs + c1] + y1;
        const double up = tmax(eta1, eta2), dn = tmin(eta1, eta2);
        const double d_up = clamp0(up - L.zbank[l]), d_dn = clamp0(dn - L.zbank[l]);
        const double ratio = d_up > 0.0 ? d_dn / (d_up > 0.0 ? d_up : 1.0) : 0.0;
        const double qw = L.weir_cw * L.length[l] * (d_up * sqrt(d_up)) * pow(clamp0(1.0 - ratio * sqrt(ratio)), 0.385);
        w.sg[l] = eta2 > eta1 ? 1.0 : -1.0;
        const double v = qw * dt;
        const double plan1 = (y1 <= s.yb ? s.b + 2 * s.m * y1 : s.ts) * N.dx[c1];
        const double v_eq = (up - dn) / (1.0 / L.area2 + 1.0 / plan1);
        w.V[l] = tmin(v, L.relax * v_eq);
        w.av1[c1] = L.limit * clamp0(A1 - bank_area(s)) * N.dx[c1];
    }
    w.hu_[k] = h2;
    // phase 2: limit factors, final volumes, and the application to the 2-D cell and the 1-D cells
    const double fk = ex_f2(w, L, k);
    double dV = 0.0;
    for (int p = p0; p < p1; ++p) {
        const int l = L.u_link[p];
        const int c1 = L.cell1d[l];
        const double v = (w.sg[l] > 0.0 ? w.V[l] * fk : w.V[l] * ex_f1(w, L, c1)) * w.sg[l];
        w.Vf[l] = v;
        if (a.V) a.V[(int64_t)b * L.nl + l] = v;
        dV += v;
        double add = 0.0;
        add += v;
        A[c1] = A[c1] + add / N.dx[c1];
    }
    const double hn = h2 - dV / L.area2;
    const double keep = h2 > 0.0 ? tmin(hn / (h2 > 0.0 ? h2 : 1.0), 1.0) : 0.0;
    R2* hu = reinterpret_cast<R2*>(a.hu);
    R2* hv = reinterpret_cast<R2*>(a.hv);
    const R2 kp = (R2)keep;
    h[q] = (R2)hn;
    hu[q] = hu[q] * kp;
    hv[q] = hv[q] * kp;
}

// ex_one_item for a unique 2-D cell with exactly one link, split where dt enters: ex_one_pre does
// the loads and the arithmetic that do not depend on dt (the weir law, the level-equalising volume,
// the sign, the 1-D cell's donor volume), ex_one_post the rest.  Every operation is ex_one_item's,
// in its order.  Nothing else in the launch writes the cell's h, hu, hv or its 1-D cell's area
// (each belongs to this item alone), so reading them before dt is known reads the same values.
template <typename R2>
struct ExPre {
    int64_t q;
    int l, c1;
    double h2, A1, dx1, qw, v_eq, sg, av1;
    R2 hu, hv;
};

template <typename R2>
__device__ __forceinline__ void ex_one_pre(const ExchangeArgs& a, const int b, const int k, ExPre<R2>& e) {
    const his_links& L = a.L;
    const his_net1d& N = a.N;
    e.q = (int64_t)b * a.plane + L.uniq[k];
    e.l = L.u_link[L.u_ptr[k]];
    e.c1 = L.cell1d[e.l];
    const int l = e.l, c1 = e.c1;
    e.h2 = (double)reinterpret_cast<const R2*>(a.h)[e.q];
    e.hu = reinterpret_cast<const R2*>(a.hu)[e.q];
    e.hv = reinterpret_cast<const R2*>(a.hv)[e.q];
    const Sec s = cell_sec(N, b, c1);
    e.A1 = a.A[(int64_t)b * N.nc + c1];
    e.dx1 = N.dx[c1];
    const double eta2 = L.z2[l] + e.h2;
    const double y1 = sec_depth(s, e.A1);
    const double eta1 = N.z[(int64_t)b * N.bs + c1] + y1;
    const double up = tmax(eta1, eta2), dn = tmin(eta1, eta2);
    const double d_up = clamp0(up - L.zbank[l]), d_dn = clamp0(dn - L.zbank[l]);
    const double ratio = d_up > 0.0 ? d_dn / (d_up > 0.0 ? d_up : 1.0) : 0.0;
    e.qw = L.weir_cw * L.length[l] * (d_up * sqrt(d_up)) * pow(clamp0(1.0 - ratio * sqrt(ratio)), 0.385);
    e.sg = eta2 > eta1 ? 1.0 : -1.0;
    const double plan1 = (y1 <= s.yb ? s.b + 2 * s.m * y1 : s.ts) * e.dx1;
    e.v_eq = (up - dn) / (1.0 / L.area2 + 1.0 / plan1);
    e.av1 = L.limit * clamp0(e.A1 - bank_area(s)) * e.dx1;
}

template <typename R2>
__device__ __forceinline__ void ex_one_post(const ExchangeArgs& a, const int b, const int k, const ExPre<R2>& e,
                                            const double dt) {
    const his_links& L = a.L;
    ExWs w = ex_ws(a, b);
    const double V = tmin(e.qw * dt, L.relax * e.v_eq);
    w.sg[e.l] = e.sg;
    w.V[e.l] = V;
    w.av1[e.c1] = e.av1;
    w.hu_[k] = e.h2;
    // ex_f2 and ex_f1, each over its one link
    const double fk = ex_factor(0.0 + (e.sg > 0.0 ? V : 0.0), L.limit * e.h2 * L.area2);
    const double f1 = ex_factor(0.0 + (e.sg < 0.0 ? V : 0.0), e.av1);
    const double v = (e.sg > 0.0 ? V * fk : V * f1) * e.sg;
    w.Vf[e.l] = v;
    if (a.V) a.V[(int64_t)b * L.nl + e.l] = v;
    const double dV = 0.0 + v;
    a.A[(int64_t)b * a.N.nc + e.c1] = e.A1 + (0.0 + v) / e.dx1;
    const double hn = e.h2 - dV / L.area2;
    const double keep = e.h2 > 0.0 ? tmin(hn / (e.h2 > 0.0 ? e.h2 : 1.0), 1.0) : 0.0;
    const R2 kp = (R2)keep;
    reinterpret_cast<R2*>(a.h)[e.q] = (R2)hn;
    reinterpret_cast<R2*>(a.hu)[e.q] = e.hu * kp;
    reinterpret_cast<R2*>(a.hv)[e.q] = e.hv * kp;
}

// Exchange phase 2: net volume of unique 2-D cell r (r < nu) -> depth, momentum kept by
// min(h_new / h, 1); of 1-D cell r - nu -> area.  Each link's volume is scaled by its donor's limit
// factor, formed here from phase 1's volumes; sums run over the target's links in link order.
template <typename R2>
__device__ __forceinline__ void ex_apply_item(const ExchangeArgs& a, const int b, const int r) {
    const his_links& L = a.L;
    const his_net1d& N = a.N;
    ExWs w = ex_ws(a, b);
    if (r < L.nu) {
        const int k = r;
        const double fk = ex_f2(w, L, k);
        double dV = 0.0;
        for (int p = L.u_ptr[k]; p < L.u_ptr[k + 1]; ++p) {
            const int l = L.u_link[p];
            const double v = (w.sg[l] > 0.0 ? w.V[l] * fk : w.V[l] * ex_f1(w, L, L.cell1d[l])) * w.sg[l];
            w.Vf[l] = v;
            if (a.V) a.V[(int64_t)b * L.nl + l] = v;
            dV += v;
        }
        const double h0 = w.hu_[k];
        const double hn = h0 - dV / L.area2;
        const double keep = h0 > 0.0 ? tmin(hn / (h0 > 0.0 ? h0 : 1.0), 1.0) : 0.0;
        const int64_t q = (int64_t)b * a.plane + L.uniq[k];
        R2* h = reinterpret_cast<R2*>(a.h);
        R2* hu = reinterpret_cast<R2*>(a.hu);
        R2* hv = reinterpret_cast<R2*>(a.hv);
        const R2 kp = (R2)keep;
        h[q] = (R2)hn;
        hu[q] = hu[q] * kp;
        hv[q] = hv[q] * kp;
        return;
    }
    const int c = r - L.nu;
    const int p0 = L.c_ptr[c], p1 = L.c_ptr[c + 1];
    if (p0 == p1) return;
    const double fc = ex_f1(w, L, c);
    double add = 0.0;
    for (int p = p0; p < p1; ++p) {
        const int l = L.c_link[p];
        add += (w.sg[l] > 0.0 ? w.V[l] * ex_f2(w, L, L.inv[l]) : w.V[l] * fc) * w.sg[l];
    }
    double* A = a.A + (int64_t)b * N.nc;
    A[c] = A[c] + add / N.dx[c];
}

}  // namespace his
