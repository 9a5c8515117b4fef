// The head of a coupled step: dt and n_sub of each shared-dt group, every member's control
// entries and the step's forcing (hydrointel/solver/coupling.py:201-215 and engine.py:104-166).
// Used by the engine's own dt kernel and, fused, by the first weir-exchange kernel.
#pragma once
#include "net1d_dev.hpp"

namespace his {

struct DtGroup { int m0, m1, gi, pad; };

struct DtArgs {
    const DtGroup* groups;
    int ngroups_dt;
    double cfl, dt_min, dt_max;
    // per grid group
    int bc[HIS_MAX_GROUPS][4];
    his_series rain[HIS_MAX_GROUPS], edge[HIS_MAX_GROUPS][4];
    int member0[HIS_MAX_GROUPS];
    int has_net[HIS_MAX_GROUPS];
    double cfl_1d[HIS_MAX_GROUPS], dt_max_1d[HIS_MAX_GROUPS];
    his_control ctl;
    int32_t* active;
    double* t_step;            // (Btot) start time of the current step
    double* pm;                // (9, Btot)
    const double* rate2d;      // (Btot)
    const double* rate1d;      // (Btot)
    const int32_t* wet1d;      // (Btot)
    double* rate2d_reset;      // (Btot) or NULL: cleared for every member that steps, for the stage-2 atomic max
    int Btot;
};

struct HeadState {
    double dt;
    int go, nsub, clamp, err;
};

// What head_group_dt reads: this thread's share of the group's CFL rates, and the clock of the
// group's first member.  Issued apart from the reduction, so that a caller can keep other loads in
// flight with them.
struct HeadLoads {
    double r2, r1, w1;
    double t, t_limit, t_end;
    int err0;
};

__device__ inline void head_loads(const DtArgs& a, const DtGroup& G, HeadLoads& v) {
    const int gi = G.gi;
    v.r2 = 0.0; v.r1 = 0.0; v.w1 = 0.0;
    for (int m = G.m0 + threadIdx.x; m < G.m1; m += blockDim.x) {
        v.r2 = a.rate2d[m] > v.r2 ? a.rate2d[m] : v.r2;
        if (a.has_net[gi]) {
            v.r1 = a.rate1d[m] > v.r1 ? a.rate1d[m] : v.r1;
            v.w1 = a.wet1d[m] ? 1.0 : v.w1;
        }
    }
    const int m0 = G.m0;
    v.t = a.ctl.t[m0];
    v.t_limit = a.ctl.t_limit[m0];
    v.t_end = a.ctl.t_end[m0];
    v.err0 = a.ctl.error[m0];
}

// dt of one shared-dt group from its members' CFL rates, for every thread of the block (up to 256
// threads; two workgroup barriers).  Every block that evaluates a group gets the same values.
__device__ inline void head_reduce(const DtArgs& a, const DtGroup& G, const HeadLoads& v, double (*red)[3],
                                   HeadState& hs) {
    const int gi = G.gi;
    double r2 = wave_max(v.r2);
    double r1 = wave_max(v.r1);
    double w1 = wave_max(v.w1);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[w][0] = r2; red[w][1] = r1; red[w][2] = w1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            r2 = red[k][0] > r2 ? red[k][0] : r2;
            r1 = red[k][1] > r1 ? red[k][1] : r1;
            w1 = red[k][2] > w1 ? red[k][2] : w1;
        }
        const double t = v.t;
        const double lim = v.t_limit < v.t_end ? v.t_limit : v.t_end;
        const int act = t < lim - 1e-9;
        double dt = 0.0;
        int nsub = 0, clamp = 0, err = 0;
        if (act && !(r2 == r2 && r2 < 1e300 && r1 == r1 && r1 < 1e300)) err = 2;   // NaN or inf rate
        if (act && !err) {
            double d2 = r2 > 0.0 ? a.cfl / r2 : a.dt_max;
            if (d2 < a.dt_min || d2 > a.dt_max) {
                clamp = d2 < a.dt_min;
                d2 = d2 < a.dt_min ? a.dt_min : d2;
                d2 = d2 > a.dt_max ? a.dt_max : d2;
            }
            dt = d2 < lim - t ? d2 : lim - t;
            if (a.has_net[gi]) {
                double d1 = a.dt_max_1d[gi];
                if (w1 > 0.0) {
                    d1 = a.cfl_1d[gi] / r1;
                    d1 = d1 < a.dt_max_1d[gi] ? d1 : a.dt_max_1d[gi];
                }
                const double q = ceil(dt / d1 - 1e-9);
                if (q > (double)a.ctl.n_sub_limit) err = 1;
                else nsub = q < 1.0 ? 1 : (int)q;
            }
        }
        hs.dt = dt;
        hs.nsub = nsub;
        hs.clamp = clamp;
        hs.err = err;
        // a member flagged before stays stopped; a flag raised now stops the whole group.  Blocks of
        // one launch that evaluate the same group compute the same err, and a block writes a flag
        // only when err != 0, when go is false whatever err0 held.
        hs.go = act && !err && v.err0 == 0;
    }
    __syncthreads();
}

__device__ inline void head_group_dt(const DtArgs& a, const DtGroup& G, double (*red)[3], HeadState& hs) {
    HeadLoads v;
    head_loads(a, G, v);
    head_reduce(a, G, v, red, hs);
}

// member m's control entries for the step (one thread per member)
__device__ inline void head_member_ctl(const DtArgs& a, const int m, const HeadState& hs, double* rate_reset,
                                       const double t) {
    if (hs.err) a.ctl.error[m] = hs.err;
    a.active[m] = hs.go;
    if (!hs.go) return;
    const double dt = hs.dt;
    a.t_step[m] = t;
    a.ctl.dt[m] = dt;
    a.ctl.n_sub[m] = hs.nsub;
    if (hs.clamp) a.ctl.clamped[m] += 1;
    a.ctl.dt_min_seen[m] = dt < a.ctl.dt_min_seen[m] ? dt : a.ctl.dt_min_seen[m];
    a.ctl.dt_max_seen[m] = dt > a.ctl.dt_max_seen[m] ? dt : a.ctl.dt_max_seen[m];
    a.ctl.n_sub_max[m] = hs.nsub > a.ctl.n_sub_max[m] ? hs.nsub : a.ctl.n_sub_max[m];
    if (rate_reset) rate_reset[m] = 0.0;
}

// forcing slot of member m for the step: 0 rain at t, 1-4 W, E, S, N at t, 5-8 the same at t + dt
// (t: the member's clock, ctl.t[m])
__device__ inline void head_member_pm(const DtArgs& a, const int gi, const int m, const int slot, const double dt,
                                      const double t) {
    const int b = m - a.member0[gi];
    double v = 0.0;
    if (slot == 0) {
        v = series_eval(a.rain[gi], b, t);
    } else {
        const int e = (slot - 1) & 3;
        const bool valued = a.bc[gi][e] == HIS_BC_STAGE || a.bc[gi][e] == HIS_BC_DISCHARGE;
        if (valued) v = series_eval(a.edge[gi][e], b, slot <= 4 ? t : t + dt);
    }
    a.pm[slot * a.Btot + m] = v;
}

}  // namespace his
