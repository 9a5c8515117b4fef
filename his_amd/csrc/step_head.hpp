// The head of a coupled step: dt and n_sub of each shared-dt group, every member's control
// entries and the step's forcing (hydrointel/solver/coupling.py:201-215 and engine.py:104-166).
// Used by the engine's own dt kernel and, fused, by the first weir-exchange kernel.
#pragma once
#include "net1d_dev.hpp"

namespace his {

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

// What head_group_dt reads: this thread's share of the group's CFL rates, the clock of the group's
// first member, and the dt limits.  Issued apart from the reduction, so that a caller can keep
// other loads in flight with them.
struct HeadLoads {
    double r2, r1, w1;
    double t, t_limit, t_end;
    int err0, has_net, n_sub_limit;
    double cfl, dt_min, dt_max, cfl_1d, dt_max_1d;
};

__device__ inline void head_loads(const HeadIn& a, const DtGroup& G, HeadLoads& v) {
    const int gi = G.gi;
    v.has_net = a.has_net[gi];
    v.cfl = a.cfl; v.dt_min = a.dt_min; v.dt_max = a.dt_max;
    v.cfl_1d = a.cfl_1d[gi]; v.dt_max_1d = a.dt_max_1d[gi];
    v.n_sub_limit = a.n_sub_limit;
    v.r2 = 0.0; v.r1 = 0.0; v.w1 = 0.0;
    for (int m = G.m0 + threadIdx.x; m < G.m1; m += blockDim.x) {
        v.r2 = a.rate2d[m] > v.r2 ? a.rate2d[m] : v.r2;
        if (v.has_net) {
            v.r1 = a.rate1d[m] > v.r1 ? a.rate1d[m] : v.r1;
            v.w1 = a.wet1d[m] ? 1.0 : v.w1;
        }
    }
    const int m0 = G.m0;
    v.t = a.t[m0];
    v.t_limit = a.t_limit[m0];
    v.t_end = a.t_end[m0];
    v.err0 = a.error[m0];
}

// HeadIn of the engine's dt arguments
__host__ __device__ inline HeadIn head_in_of(const DtArgs& a) {
    HeadIn h{};
    h.rate2d = a.rate2d; h.rate1d = a.rate1d; h.wet1d = a.wet1d;
    h.t = a.ctl.t; h.t_limit = a.ctl.t_limit; h.t_end = a.ctl.t_end; h.error = a.ctl.error;
    h.cfl = a.cfl; h.dt_min = a.dt_min; h.dt_max = a.dt_max;
    for (int g = 0; g < HIS_MAX_GROUPS; ++g) {
        h.cfl_1d[g] = a.cfl_1d[g];
        h.dt_max_1d[g] = a.dt_max_1d[g];
        h.has_net[g] = a.has_net[g];
    }
    h.n_sub_limit = a.ctl.n_sub_limit;
    return h;
}

// dt of one shared-dt group from its members' CFL rates, for every thread of the block (up to 256
// threads; two workgroup barriers).  Every block that evaluates a group gets the same values.
__device__ inline void head_reduce(const HeadLoads& v, double (*red)[3], HeadState& hs) {
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
            double d2 = r2 > 0.0 ? v.cfl / r2 : v.dt_max;
            if (d2 < v.dt_min || d2 > v.dt_max) {
                clamp = d2 < v.dt_min;
                d2 = d2 < v.dt_min ? v.dt_min : d2;
                d2 = d2 > v.dt_max ? v.dt_max : d2;
            }
            dt = d2 < lim - t ? d2 : lim - t;
            if (v.has_net) {
                double d1 = v.dt_max_1d;
                if (w1 > 0.0) {
                    d1 = v.cfl_1d / r1;
                    d1 = d1 < v.dt_max_1d ? d1 : v.dt_max_1d;
                }
                const double q = ceil(dt / d1 - 1e-9);
                if (q > (double)v.n_sub_limit) err = 1;
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
    head_loads(head_in_of(a), G, v);
    head_reduce(v, red, hs);
}

// The control entries head_member_ctl updates from their old values, read ahead of it.
struct MemberSeen {
    double dt_min, dt_max;
    int n_sub_max, clamped;
};

__device__ inline MemberSeen member_seen(const DtArgs& a, const int m) {
    return MemberSeen{a.ctl.dt_min_seen[m], a.ctl.dt_max_seen[m], a.ctl.n_sub_max[m], a.ctl.clamped[m]};
}

// member m's control entries for the step (one thread per member; t = ctl.t[m], o = its entries)
__device__ inline void head_member_ctl(const DtArgs& a, const int m, const HeadState& hs, double* rate_reset,
                                       const double t, const MemberSeen& o) {
    if (hs.err) a.ctl.error[m] = hs.err;
    a.active[m] = hs.go;
    if (!hs.go) return;
    const double dt = hs.dt;
    a.t_step[m] = t;
    a.ctl.dt[m] = dt;
    a.ctl.n_sub[m] = hs.nsub;
    if (hs.clamp) a.ctl.clamped[m] = o.clamped + 1;
    a.ctl.dt_min_seen[m] = dt < o.dt_min ? dt : o.dt_min;
    a.ctl.dt_max_seen[m] = dt > o.dt_max ? dt : o.dt_max;
    a.ctl.n_sub_max[m] = hs.nsub > o.n_sub_max ? hs.nsub : o.n_sub_max;
    if (rate_reset) rate_reset[m] = 0.0;
}

// forcing slot of member m for the step: 0 rain at t, 1-4 W, E, S, N at t, 5-8 the same at t + dt;
// head_pm_value evaluates a slot at time tt (t, or t + dt for slots 5-8)
__device__ inline double head_pm_value(const DtArgs& a, const int gi, const int m, const int slot, const double tt) {
    const int b = m - a.member0[gi];
    if (slot == 0) return series_eval(a.rain[gi], b, tt);
    const int e = (slot - 1) & 3;
    const bool valued = a.bc[gi][e] == HIS_BC_STAGE || a.bc[gi][e] == HIS_BC_DISCHARGE;
    return valued ? series_eval(a.edge[gi][e], b, tt) : 0.0;
}

__device__ inline void head_member_pm(const DtArgs& a, const int gi, const int m, const int slot, const double dt,
                                      const double t) {
    a.pm[slot * a.Btot + m] = head_pm_value(a, gi, m, slot, slot <= 4 ? t : t + dt);
}

}  // namespace his
