// Per-member reduction of the 2-D step's tile partials into volumes, CFL rate and the float64
// ledger (the finalize), as a block-level device function: the standalone finalize kernel runs one
// block per member, and the engine appends the same blocks to the next step's first exchange
// launch so that the reduction needs no stream of its own.
//
// Each thread sums its tiles' stage-1 sums (slots 0..8), stage-2 sums (10..22), rate maximum (23)
// and depth maximum (24); each quantity is then reduced by a fixed butterfly within every wave and
// the waves' results are combined in wave order, so the result is bit-reproducible.  A depth or
// storage sum that is not finite flags the member with error code 3, a depth above the plausibility
// limit with error code 4 (hydrointel/solver/coupling.py:169-192, checked here after every step).
// Volumes follow hydrointel/solver/swe2d.py:_step_core 305-371: boundary volumes from the RK
// average of the two stages' edge fluxes (324-328), rain (331-337), infiltration and clipping.
#pragma once
#include "swe2d_desc.hpp"

namespace his {

// finite test on the bit pattern (exponent not all ones), whatever the unit's floating-point flags
__device__ __forceinline__ bool finite_bits(double x) {
    return ((unsigned long long)__double_as_longlong(x) & 0x7ff0000000000000ULL) != 0x7ff0000000000000ULL;
}

constexpr int FIN_NV = 25;      // reduced quantities per member: 9 + 13 sums (slot 9 unused), 2 maxima

template <int NT>
__device__ __forceinline__ void finalize_member(const FinalizeArgs& F, const int m, double (*red)[FIN_NV],
                                                double* tot) {
    constexpr int NV = FIN_NV;
    if (F.snap) {
        if (F.snap[2 * F.Btot + m] == 0.0) return;     // the member did not step
    } else if (F.active && !F.active[m]) {
        return;
    }
    if (F.only && !F.only[m]) return;
    int gi = 0;
    while (gi + 1 < F.ngroups && m >= F.member0[gi + 1]) ++gi;
    const int b = m - F.member0[gi];
    const int per = F.per[gi];
    const int64_t t0 = F.tile0[gi] + (int64_t)b * per;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    for (int t = threadIdx.x; t < per; t += NT) {
        const double* p = F.part + (t0 + t) * NP2;
        if (!F.stage_only)
#pragma unroll
            for (int k = 0; k <= P_NOUT; ++k) v[k] += p[k];
#pragma unroll
        for (int k = 0; k < P_NSUM; ++k) v[10 + k] += p[NP + k];
        v[23] = p[NP + P_RATE] > v[23] ? p[NP + P_RATE] : v[23];
        v[24] = p[NP + P_HMAX] > v[24] ? p[NP + P_HMAX] : v[24];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV - 2; ++k) v[k] = wave_sum(v[k]);
    v[NV - 2] = wave_max(v[NV - 2]);
    v[NV - 1] = wave_max(v[NV - 1]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[wv][k] = v[k];
    __syncthreads();
    if (threadIdx.x < NV) {
        const int k = threadIdx.x;
        double a = red[0][k];
        for (int w = 1; w < NT / 64; ++w) a = k >= NV - 2 ? (red[w][k] > a ? red[w][k] : a) : a + red[w][k];
        tot[k] = a;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double s1[P_NOUT + 1], s2[P_NSUM];
    for (int k = 0; k <= P_NOUT; ++k) s1[k] = tot[k];
    for (int k = 0; k < P_NSUM; ++k) s2[k] = tot[10 + k];
    const double rmax = tot[23];
    const double area = F.area[gi];
    if (F.error && F.error[m] == 0 && !(finite_bits(s2[P_VOLH]) && finite_bits(s2[P_VOLS]))) F.error[m] = 3;
    if (F.error && F.error[m] == 0 && !(tot[24] <= HIS_PLAUSIBLE_DEPTH_2D_M)) F.error[m] = 4;
    if (F.rate) F.rate[m] = rmax;
    if (F.volume) {
        F.volume[m] = s2[P_VOLH] * area;
        F.volume[F.Btot + m] = s2[P_VOLS] * area;
    }
    if (F.stage_only) return;
    const double dt = F.snap ? F.snap[m] : F.dt[m], dx = F.dx[gi], dy = F.dy[gi];
    const double q2i = (s2[P_WIN] + s2[P_EIN]) * dy + (s2[P_SIN] + s2[P_NIN]) * dx;
    const double q2o = (s2[P_WOUT] + s2[P_EOUT]) * dy + (s2[P_SOUT] + s2[P_NOUT]) * dx;
    double qin, qout, clip;
    if (F.order[gi] == 2) {
        const double q1i = (s1[P_WIN] + s1[P_EIN]) * dy + (s1[P_SIN] + s1[P_NIN]) * dx;
        const double q1o = (s1[P_WOUT] + s1[P_EOUT]) * dy + (s1[P_SOUT] + s1[P_NOUT]) * dx;
        qin = 0.5 * (q1i + q2i);
        qout = 0.5 * (q1o + q2o);
        clip = s1[P_CLIP] * area + s2[P_CLIP] * area;
    } else {
        qin = q2i;
        qout = q2o;
        clip = s2[P_CLIP] * area;
    }
    const double rain = F.snap ? F.snap[F.Btot + m] : F.pm[m];
    double rain_v;
    if (F.weighted[gi]) rain_v = s2[P_RAINW] * area;
    else {
        const double rd = F.rain_f32[gi] ? (double)((float)rain * (float)dt) : rain * dt;
        rain_v = rd * (F.cells[gi] * area);
    }
    const double inf_v = s2[P_INF] * area;
    const double bin = qin * dt, bout = qout * dt;
    if (F.vols) {
        F.vols[m] = bin;
        F.vols[F.Btot + m] = bout;
        F.vols[2 * F.Btot + m] = rain_v;
        F.vols[3 * F.Btot + m] = inf_v;
        F.vols[4 * F.Btot + m] = clip;
    }
    if (F.t_out && F.advance[gi]) {
        F.t_out[m] = F.t_step[m] + dt;
        F.steps[m] += 1;
    }
    if (F.acc) {
        F.acc[0 * F.Btot + m] += rain_v;
        F.acc[1 * F.Btot + m] += inf_v;
        F.acc[2 * F.Btot + m] += bin;
        F.acc[3 * F.Btot + m] += bout;
        F.acc[6 * F.Btot + m] += clip;
    }
    if (F.snap) F.snap[2 * F.Btot + m] = 0.0;
}

}  // namespace his
