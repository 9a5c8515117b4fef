// 1-D St. Venant network step and Villemonte weir exchange (float64, gfx950).
//
// One workgroup per member keeps the whole network in flight: the reference's
// gathers become per-face reads, and every index_add becomes a per-cell loop over
// that cell's faces (or links) in increasing index order -- the order torch's CPU
// index_add accumulates in -- so results are deterministic and follow the reference
// bit for bit wherever the libm functions agree.
//
// SWE1D._step_core  hydrointel/solver/swe1d.py:320-368 (+ _fill_ghosts 265-294,
//                   Sections 50-84, hll_1d 87-107, max_dt 253-262)
// exchange          hydrointel/solver/coupling.py:73-124
#include "finalize_dev.hpp"
#include "step_head.hpp"

#include <type_traits>

namespace his {

template <int NT, int NV>
__device__ __forceinline__ void block_sum_bcast(double (&v)[NV], double (*red)[8]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
        for (int k = 0; k < NV; ++k) red[w][k] = v[k];
    __syncthreads();
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int w2 = 0; w2 < NT / 64; ++w2) s += red[w2][k];
        v[k] = s;
    }
}

// per-member epilogue: rate of the current state (swe1d.py:253-262)
template <int NT>
__device__ void rate1d(const his_net1d& N, int b, const double* A, const double* Q, double* rate, int32_t* wet,
                       int m, double (*red)[8]) {
    double mx = 0.0, anyw = 0.0;
    for (int c = threadIdx.x; c < N.nc; c += NT) {
        if (N.kind[c] >= 2) continue;
        Sec s = cell_sec(N, b, c);
        double a = A[c], q = Q[c];
        double y = sec_depth(s, a);
        if (!(y > N.h_dry)) continue;
        anyw = 1.0;
        double cc = sqrt(N.g * a / sec_top(s, y));
        double u = vel1(s, a, q, N.h_dry);
        double r = (fabs(u) + cc) / N.dx[c];
        mx = r > mx ? r : mx;
    }
    mx = wave_max(mx);
    anyw = wave_max(anyw);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) { red[w][0] = mx; red[w][1] = anyw; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, bw = 0.0;
        for (int w2 = 0; w2 < NT / 64; ++w2) {
            a = red[w2][0] > a ? red[w2][0] : a;
            bw = red[w2][1] > bw ? red[w2][1] : bw;
        }
        if (rate) rate[m] = a;
        if (wet) wet[m] = bw > 0.0 ? 1 : 0;
    }
}

// The cells, faces and ghosts one workgroup steps: part p of the network's split, or all of them.
struct PartRange {
    int c0, c1, f0, f1, g0, g1;
    const int32_t *cell, *face, *ghost;     // NULL = identity (the whole network)
    __device__ __forceinline__ int c(int i) const { return cell ? cell[i] : i; }
    __device__ __forceinline__ int f(int i) const { return face ? face[i] : i; }
    __device__ __forceinline__ int g(int i) const { return ghost ? ghost[i] : i; }
};

__device__ __forceinline__ PartRange part_range(const his_net1d& N, int p) {
    PartRange r;
    if (N.nparts > 1) {
        r.c0 = N.p_cptr[p]; r.c1 = N.p_cptr[p + 1];
        r.f0 = N.p_fptr[p]; r.f1 = N.p_fptr[p + 1];
        r.g0 = N.p_gptr[p]; r.g1 = N.p_gptr[p + 1];
        r.cell = N.p_cell; r.face = N.p_face; r.ghost = N.p_ghost;
    } else {
        r.c0 = 0; r.c1 = N.nc; r.f0 = 0; r.f1 = N.nf; r.g0 = 0; r.g1 = N.ngh;
        r.cell = r.face = r.ghost = nullptr;
    }
    return r;
}

// All substeps of one coupled step for part p of member b's network (one workgroup).  A split
// network's parts share no face, so each part's workgroup runs every substep on its own; the
// member's ledger terms, CFL rate and wet flag are per-part partials that the last part to finish
// combines in part order (a counter per member, reset by that part), so results are deterministic.
template <int NT>
__device__ void swe1d_member(const his_net1d& N, double* __restrict__ Ag, double* __restrict__ Qg,
                             const Net1DDrive& d, double* __restrict__ wsg, const int b, const int p,
                             double (*red)[8]) {
    const int m = d.member0 + b;
    if (d.active && !d.active[m]) return;
    const int nc = N.nc, nf = N.nf;
    const int P = N.nparts > 1 ? N.nparts : 1;
    const PartRange R = part_range(N, p);
    double* A = Ag + (int64_t)b * nc;
    double* Q = Qg + (int64_t)b * nc;
    double* ws = wsg + (int64_t)b * (3 * nc + 4 * nf);
    double *y = ws, *u = ws + nc, *th = ws + 2 * nc;
    double *F1 = ws + 3 * nc, *F2 = F1 + nf, *pL = F2 + nf, *pR = pL + nf;
    const double g = N.g, h_dry = N.h_dry;
    // split networks: this part's ledger partials (b_in, b_out, clip, rate, and wet + 64 x implausible:
    // summed over at most 16 parts, the sum still tells both flags apart) and the member's counter
    double* led = wsg + (int64_t)N.B * (3 * nc + 4 * nf) + ((int64_t)b * P + p) * LED_PART;
    unsigned int* cnt = reinterpret_cast<unsigned int*>(wsg + (int64_t)N.B * (3 * nc + 4 * nf) +
                                                        (int64_t)N.B * P * LED_PART + b);
    double led_in = 0.0, led_out = 0.0, led_clip = 0.0;
    int nsub = 1;
    double dt_full = d.dt[m], t0 = 0.0;
    if (d.engine) {
        nsub = d.n_sub[m];
        t0 = d.t[m];
    }
    const double sub = d.engine ? dt_full / nsub : dt_full;
    if (d.engine && d.ex_vf && p == 0) {
        // gross exchanged volumes of this step's weir exchange (coupling.py:210-211), fixed order
        const double* vf = d.ex_vf + (int64_t)b * d.ex_stride;
        double part[2] = {0.0, 0.0};
        for (int l = threadIdx.x; l < d.ex_nl; l += NT) {
            part[0] += clamp0(vf[l]);
            part[1] += clamp0(-vf[l]);
        }
        block_sum_bcast<NT, 2>(part, red);
        if (threadIdx.x == 0) {
            d.acc[8 * d.Btot + m] += part[0];
            d.acc[9 * d.Btot + m] += part[1];
        }
        __syncthreads();
    }
    // CFL rate of the final state (swe1d.py:253-262), formed in the last substep's update
    const bool want_rate = d.rate || d.wet;
    double r_mx = 0.0, r_wet = 0.0;
    // plausibility guard of the final state (coupling.py:169-192): a channel level more than
    // HIS_PLAUSIBLE_RISE_1D_M above its bank crest, or not finite, flags the member (error code 5).
    // The crest is taken as bed + bankfull depth, which equals it wherever the bank stands 0.2 m or
    // more above the bed (the host's check at check times uses the crest itself)
    const bool guard = d.engine && d.error;
    double implaus = 0.0;
    for (int k = 0; k < nsub; ++k) {
        double q_in, eta;
        if (d.engine) {
            const double tk = t0 + k * sub;
            q_in = series_eval(d.inflow, b, tk);
            eta = series_eval(d.stage, b, tk);
        } else {
            q_in = d.pm[m];
            eta = d.pm[d.Btot + m];
        }
        const double dt = sub;
        // ghosts from their interior neighbours (interior cells are not written here)
        for (int ig = R.g0 + (int)threadIdx.x; ig < R.g1; ig += NT) {
            const int kg = R.g(ig);
            const int gi = N.g_idx[kg], src = N.g_src[kg];
            const double sg = N.g_sign[kg];
            const Sec s = cell_sec(N, b, gi);
            const double As = A[src], Qs = Q[src];
            const double ys = sec_depth(s, As);
            const double zg = N.z[(int64_t)b * N.bs + gi];
            const double un = sg * vel1(s, As, Qs, h_dry);
            const bool wet_i = ys >= h_dry;
            const double yst = clamp0(eta - zg);
            double Ast = sec_area(s, yst);
            const double ci = sqrt(g * As / sec_top(s, ys));
            const double cg = sqrt(g * Ast / sec_top(s, yst));
            double un_st = wet_i ? un + 2.0 * (ci - cg) : 0.0;
            if (wet_i && (un > ci)) { un_st = un; Ast = As; }
            if (!(sec_depth(s, Ast) >= h_dry)) un_st = 0.0;
            const double Q_st = sg * Ast * un_st;
            double yc = cbrt(q_in * q_in / (g * s.b * s.b));
            if (yc < h_dry) yc = h_dry;
            const double A_in = wet_i ? As : sec_area(s, yc);
            const double Q_in = -sg * q_in;
            const int kd = N.kind[gi];
            A[gi] = kd == 3 ? Ast : (kd == 2 ? A_in : As);
            Q[gi] = kd == 3 ? Q_st : (kd == 2 ? Q_in : -Qs);
        }
        __syncthreads();
        for (int ic = R.c0 + (int)threadIdx.x; ic < R.c1; ic += NT) {
            const int c = R.c(ic);
            const Sec s = cell_sec(N, b, c);
            y[c] = sec_depth(s, A[c]);
            u[c] = vel1(s, A[c], Q[c], h_dry);
        }
        __syncthreads();
        for (int jf = R.f0 + (int)threadIdx.x; jf < R.f1; jf += NT) {
            const int f = R.f(jf);
            const int l = N.fl[f], r = N.fr[f];
            const Sec fs = face_sec(N, b, f);
            const double zL = N.z[(int64_t)b * N.bs + l], zR = N.z[(int64_t)b * N.bs + r];
            const double zf = tmax(zL, zR);
            const double yL = clamp0(zL + y[l] - zf), yR = clamp0(zR + y[r] - zf);
            const double AL = sec_area(fs, yL), AR = sec_area(fs, yR);
            const double uL = u[l], uR = u[r];
            const double pl = g * sec_i1(fs, yL), pr = g * sec_i1(fs, yR);
            const double cL = sqrt(g * AL / sec_top(fs, yL)), cR = sqrt(g * AR / sec_top(fs, yR));
            const double QL = AL * uL, QR = AR * uR;
            const bool dryL = AL <= HIS_TINY, dryR = AR <= HIS_TINY;
            double SL = tmin(uL - cL, uR - cR), SR = tmax(uL + cL, uR + cR);
            if (dryL) { SL = uR - 2.0 * cR; SR = uR + cR; }
            if (dryR) { SL = uL - cL; SR = uL + 2.0 * cL; }
            const double FL2 = QL * uL + pl, FR2 = QR * uR + pr;
            double den = SR - SL;
            if (fabs(den) < HIS_TINY) den = 1.0;
            double f1, f2;
            if (SL >= 0.0) { f1 = QL; f2 = FL2; }
            else if (SR <= 0.0) { f1 = QR; f2 = FR2; }
            else {
                const double rden = 1.0 / den;
                f1 = (SR * QL - SL * QR + SL * SR * (AR - AL)) * rden;
                f2 = (SR * FL2 - SL * FR2 + SL * SR * (QR - QL)) * rden;
            }
            if (dryL && dryR) { f1 = 0.0; f2 = 0.0; }
            F1[f] = f1; F2[f] = f2; pL[f] = pl; pR[f] = pr;
        }
        __syncthreads();
        // conservative outflow limiter (donor = upwind cell; ghosts unlimited)
        for (int ic = R.c0 + (int)threadIdx.x; ic < R.c1; ic += NT) {
            const int c = R.c(ic);
            double out = 0.0;
            for (int p = N.lf_ptr[c]; p < N.lf_ptr[c + 1]; ++p) out += relu(F1[N.lf_face[p]]);
            for (int p = N.rf_ptr[c]; p < N.rf_ptr[c + 1]; ++p) out += relu(-F1[N.rf_face[p]]);
            const double vol = A[c] * N.dx[c];
            const double need = out * dt;
            const double t = need > vol ? vol / (need > 0.0 ? need : 1.0) : 1.0;
            th[c] = N.kind[c] >= 2 ? 1.0 : t;
        }
        __syncthreads();
        double part[5] = {0.0, 0.0, 0.0, 0.0, 0.0};   // clip, in(gl), in(gr), out(gl), out(gr)
        for (int jf = R.f0 + (int)threadIdx.x; jf < R.f1; jf += NT) {
            const int f = R.f(jf);
            const double f1 = F1[f];
            const double sc = f1 > 0.0 ? th[N.fl[f]] : th[N.fr[f]];
            const double f1s = f1 * sc;
            F1[f] = f1s;
            const bool gl = N.kind[N.fl[f]] >= 2, gr = N.kind[N.fr[f]] >= 2;
            if (gl) { part[1] += relu(f1s); part[3] += relu(-f1s); }
            if (gr) { part[2] += relu(-f1s); part[4] += relu(f1s); }
        }
        __syncthreads();
        for (int ic = R.c0 + (int)threadIdx.x; ic < R.c1; ic += NT) {
            const int c = R.c(ic);
            const int kd = N.kind[c];
            if (kd >= 2) continue;   // ghosts keep their (refilled) values
            // dA and dQ each summed over the left faces, then the right faces, in list order
            double dA = 0.0, dQ = 0.0;
            for (int p = N.lf_ptr[c]; p < N.lf_ptr[c + 1]; ++p) {
                const int f = N.lf_face[p];
                dA += -F1[f];
                dQ += -(F2[f] - pL[f]);
            }
            for (int p = N.rf_ptr[c]; p < N.rf_ptr[c + 1]; ++p) {
                const int f = N.rf_face[p];
                dA += F1[f];
                dQ += F2[f] - pR[f];
            }
            const double inv_dx = 1.0 / N.dx[c];
            const double rdx = dt * inv_dx;
            double A1 = A[c] + dA * rdx;
            double Q1 = Q[c] + dQ * rdx;
            const double neg = clamp0(-A1);
            part[0] += neg * N.dx[c];
            A1 = A1 + neg;
            const Sec s = cell_sec(N, b, c);
            const double y1 = sec_depth(s, A1);
            const double Ac = A1 < 1e-12 ? 1e-12 : A1;
            const double Rh = Ac / sec_perim(s, y1);
            const double Rc = Rh < h_dry ? h_dry : Rh;
            const double nm = N.nman[c];
            // Q / (1 + dt g n^2 |Q| / (A R^(4/3))) with one division: Q X / (X + dt g n^2 |Q|)
            const double X = Ac * (Rc * cbrt(Rc));
            Q1 = Q1 * X / (X + dt * g * (nm * nm) * fabs(Q1));
            if (!((y1 >= h_dry) && kd != 1)) Q1 = 0.0;
            A[c] = A1;
            Q[c] = Q1;
            // level above the bank crest: the depth above the section's bankfull depth, which is
            // the crest above the bed (at least 0.2 m), so no further loads are needed
            if (guard && k == nsub - 1 && !(y1 - s.yb <= HIS_PLAUSIBLE_RISE_1D_M)) implaus = 1.0;
            if (want_rate && k == nsub - 1 && y1 > h_dry) {
                r_wet = 1.0;
                const double cc = sqrt(g * A1 / sec_top(s, y1));
                const double uu = vel1(s, A1, Q1, h_dry);
                const double r = (fabs(uu) + cc) * inv_dx;
                r_mx = r > r_mx ? r : r_mx;
            }
        }
        block_sum_bcast<NT, 5>(part, red);
        if (threadIdx.x == 0) {
            const double bin = (part[1] + part[2]) * dt, bout = (part[3] + part[4]) * dt;
            if (P > 1) {
                led_in += bin;
                led_out += bout;
                led_clip += part[0];
            } else if (d.engine) {
                d.acc[4 * d.Btot + m] += bin;
                d.acc[5 * d.Btot + m] += bout;
                d.acc[7 * d.Btot + m] += part[0];
            } else if (d.vols) {
                d.vols[m] = bin;
                d.vols[d.Btot + m] = bout;
                d.vols[2 * d.Btot + m] = part[0];
            }
        }
        __syncthreads();
    }
    double rate_blk = 0.0, wet_blk = 0.0, bad_blk = 0.0;
    if (want_rate || guard) {
        r_mx = wave_max(r_mx);
        r_wet = wave_max(r_wet);
        implaus = wave_max(implaus);
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        if (lane == 0) { red[w][0] = r_mx; red[w][1] = r_wet; red[w][2] = implaus; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w2 = 0; w2 < NT / 64; ++w2) {
                rate_blk = red[w2][0] > rate_blk ? red[w2][0] : rate_blk;
                wet_blk = red[w2][1] > wet_blk ? red[w2][1] : wet_blk;
                bad_blk = red[w2][2] > bad_blk ? red[w2][2] : bad_blk;
            }
            if (P == 1) {
                if (want_rate && d.rate) d.rate[m] = rate_blk;
                if (want_rate && d.wet) d.wet[m] = wet_blk > 0.0 ? 1 : 0;
                if (guard && bad_blk > 0.0 && d.error[m] == 0) d.error[m] = 5;
            }
        }
    }
    if (P > 1 && threadIdx.x == 0) {
        led[0] = led_in; led[1] = led_out; led[2] = led_clip; led[3] = rate_blk;
        led[4] = (wet_blk > 0.0 ? 1.0 : 0.0) + (bad_blk > 0.0 ? 64.0 : 0.0);
        __threadfence();
        const unsigned int done = atomicAdd(cnt, 1u);
        if (done == (unsigned int)(P - 1)) {
            // every other part has published its partials (each fenced before its increment)
            __threadfence();
            const volatile double* all = wsg + (int64_t)N.B * (3 * nc + 4 * nf) + (int64_t)b * P * LED_PART;
            double s_in = 0.0, s_out = 0.0, s_clip = 0.0, s_rate = 0.0, s_flags = 0.0;
            for (int q = 0; q < P; ++q) {
                s_in += all[q * LED_PART + 0];
                s_out += all[q * LED_PART + 1];
                s_clip += all[q * LED_PART + 2];
                s_rate = all[q * LED_PART + 3] > s_rate ? all[q * LED_PART + 3] : s_rate;
                s_flags += all[q * LED_PART + 4];
            }
            const double s_wet = s_flags >= 64.0 ? s_flags - 64.0 * floor(s_flags / 64.0) : s_flags;
            if (guard && s_flags >= 64.0 && d.error[m] == 0) d.error[m] = 5;
            if (d.engine) {
                d.acc[4 * d.Btot + m] += s_in;
                d.acc[5 * d.Btot + m] += s_out;
                d.acc[7 * d.Btot + m] += s_clip;
            } else if (d.vols) {
                d.vols[m] = s_in;
                d.vols[d.Btot + m] = s_out;
                d.vols[2 * d.Btot + m] = s_clip;
            }
            if (want_rate) {
                if (d.rate) d.rate[m] = s_rate;
                if (d.wet) d.wet[m] = s_wet > 0.0 ? 1 : 0;
            }
            *cnt = 0u;                 // ready for the next launch
        }
    }
    if (d.engine && d.t_out && p == 0 && threadIdx.x == 0) {
        d.t_out[m] = t0 + dt_full;
        d.steps[m] += 1;
    }
}

// workgroup k of a network's launch = (member k / parts, part k % parts)
__device__ __forceinline__ int net_parts(const his_net1d& N) { return N.nparts > 1 ? N.nparts : 1; }

template <int NT>
__global__ void __launch_bounds__(NT) k_swe1d(const his_net1d N, double* __restrict__ Ag, double* __restrict__ Qg,
                                               const Net1DDrive d, double* __restrict__ wsg) {
    __shared__ double red[NT / 64][8];
    const int P = net_parts(N);
    swe1d_member<NT>(N, Ag, Qg, d, wsg, (int)blockIdx.x / P, (int)blockIdx.x % P, red);
}

// Every group's network in one launch: workgroups blk0[g] .. blk0[g+1]-1 are group g's (member, part)s.
template <int NT>
__global__ void __launch_bounds__(NT) k_swe1d_groups(const Swe1DGroups G) {
    __shared__ double red[NT / 64][8];
    int gi = 0;
    while (gi + 1 < G.ngroups && (int)blockIdx.x >= G.blk0[gi + 1]) ++gi;
    const int P = net_parts(G.N[gi]);
    const int k = (int)blockIdx.x - G.blk0[gi];
    swe1d_member<NT>(G.N[gi], G.A[gi], G.Q[gi], G.d[gi], G.ws[gi], k / P, k % P, red);
}

// clears the per-member counters of a split network's workspace (a kernel, not a memset, so a
// captured graph holds no memset node)
__global__ void k_clear_u32(unsigned int* p, int n, int stride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[(int64_t)i * stride] = 0u;
}

template <int NT>
__global__ void __launch_bounds__(NT) k_swe1d_rate(const his_net1d N, const double* Ag, const double* Qg,
                                                    double* rate, int32_t* wet, int member0, const int32_t* only) {
    __shared__ double red[NT / 64][8];
    const int b = blockIdx.x;
    if (only && !only[member0 + b]) return;
    rate1d<NT>(N, b, Ag + (int64_t)b * N.nc, Qg + (int64_t)b * N.nc, rate, wet, member0 + b, red);
}

// per member: y, u, theta (nc each), F1, F2, pL, pR (nf each); then for a split network the parts'
// ledger partials (LED_PART per part) and one counter slot per member
size_t swe1d_ws_doubles(const his_net1d& n) {
    const size_t P = n.nparts > 1 ? (size_t)n.nparts : 1;
    const size_t split = n.nparts > 1 ? LED_PART * P + 1 : 0;
    return (size_t)n.B * (3 * (size_t)n.nc + 4 * (size_t)n.nf + split);
}

unsigned int* swe1d_counters(const his_net1d& n, double* ws) {
    return reinterpret_cast<unsigned int*>(ws + (size_t)n.B * (3 * (size_t)n.nc + 4 * (size_t)n.nf) +
                                           (size_t)n.B * LED_PART * (size_t)n.nparts);
}

int clear_swe1d_counters(const his_net1d& n, double* ws, hipStream_t s) {
    if (n.nparts <= 1) return 0;
    hipLaunchKernelGGL(k_clear_u32, dim3((n.B + 255) / 256), dim3(256), 0, s, swe1d_counters(n, ws), n.B, 2);
    HIS_HIP(hipGetLastError());
    return 0;
}

// the most cells or faces one workgroup of the network steps (the descriptor's part width when the
// network is split, which only sizes the workgroup: every loop strides over its part)
int swe1d_part_width(const his_net1d& n) {
    if (n.nparts > 1 && n.part_width > 0) return n.part_width;
    return n.nc > n.nf ? n.nc : n.nf;
}

// Workgroup size for a network part of `width` cells or faces: one thread per cell and per face
// where the part fits in 1024, in whole waves.  Idle waves would still meet every barrier of every
// substep and hold wave slots the 2-D stages running beside the network could use.
static int swe1d_threads(int width) {
    int nt = 64;
    while (nt < width && nt < 1024) nt *= 2;
    return nt;
}

template <typename F>
static int with_threads(int nt, F&& launch) {
    switch (nt) {
    case 64: return launch(std::integral_constant<int, 64>{});
    case 128: return launch(std::integral_constant<int, 128>{});
    case 256: return launch(std::integral_constant<int, 256>{});
    case 512: return launch(std::integral_constant<int, 512>{});
    default: return launch(std::integral_constant<int, 1024>{});
    }
}

int launch_swe1d(const his_net1d& n, double* A, double* Q, const Net1DDrive& d, double* ws, int width,
                 hipStream_t s) {
    const int P = n.nparts > 1 ? n.nparts : 1;
    return with_threads(swe1d_threads(width), [&](auto nt) {
        constexpr int NT = decltype(nt)::value;
        hipLaunchKernelGGL(k_swe1d<NT>, dim3(n.B * P), dim3(NT), 0, s, n, A, Q, d, ws);
        HIS_HIP(hipGetLastError());
        return 0;
    });
}

int launch_swe1d_groups(const Swe1DGroups& g, hipStream_t s) {
    if (g.ngroups < 1) return 0;
    int nt = 64;
    for (int gi = 0; gi < g.ngroups; ++gi) nt = swe1d_threads(g.width[gi]) > nt ? swe1d_threads(g.width[gi]) : nt;
    return with_threads(nt, [&](auto t) {
        constexpr int NT = decltype(t)::value;
        hipLaunchKernelGGL(k_swe1d_groups<NT>, dim3(g.blk0[g.ngroups]), dim3(NT), 0, s, g);
        HIS_HIP(hipGetLastError());
        return 0;
    });
}

int launch_swe1d_rate(const his_net1d& n, const double* A, const double* Q, double* rate, int32_t* wet,
                      int member0, hipStream_t s, const int32_t* only) {
    hipLaunchKernelGGL(k_swe1d_rate<1024>, dim3(n.B), dim3(1024), 0, s, n, A, Q, rate, wet, member0, only);
    HIS_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------ exchange
//
// Three grid-wide kernels, every member and link / unique 2-D cell / 1-D cell in flight at once:
//   1. per link: free-weir + Villemonte volume capped by the level-equalising volume, its sign;
//      per unique 2-D cell: the donor depth h_u; per 1-D cell (by its first link): its available
//      donor volume
//   2. per unique 2-D cell: donor limit factors of its links -> net volume -> new depth, momentum
//      kept by min(h_new/h, 1); per 1-D cell: the same -> new area (sums run over each target's
//      links in link order; limits read only phase 1's outputs, so in-place updates do not feed them)
//   3. per member: gross volumes each way into the ledger (fixed-order block reduction)
// When every 1-D cell has at most one link, a linked 2-D cell's links and the cells they change
// belong to it alone: one item per unique 2-D cell then does phases 1 and 2 in the first launch
// (ex_one_item, or ex_one_pre / ex_one_post around the fused step head), with the same operations.

size_t exchange_ws_doubles(const his_links& l, const his_net1d& n) {
    return (size_t)n.B * (size_t)exchange_ws_stride(l, n);
}

// thread -> (group, item within the group) for a grid that spans every group's items
__device__ __forceinline__ int ex_group(const ExGroups& G, const int64_t* it0, int64_t it) {
    int gi = 0;
    while (gi + 1 < G.ngroups && it >= it0[gi + 1]) ++gi;
    return gi;
}

// member of link-phase item it, and whether it is that member's first item
__device__ __forceinline__ int ex_link_member(const ExGroups& G, int64_t it, bool& first) {
    const int gi = ex_group(G, G.it_links, it);
    const ExchangeArgs& a = G.a[gi];
    const int64_t local = it - G.it_links[gi];
    const int per = G.one_phase ? a.L.nu : a.L.nl + a.L.nu;
    const int b = (int)(local / per);
    first = local == (int64_t)b * per;
    return a.member0 + b;
}

// The step head fused with the single-phase exchange (one item per unique 2-D cell).  Both have a
// chain of dependent loads: the head's from the block's members to their dt group and its rates and
// clock, the item's from the cell to its link, its 1-D cell and that cell's section and area.  The
// item's dt-free part (ex_one_pre) is written between the head's loads and its reduction, so the two
// chains are in flight together; only ex_one_post waits for dt.
template <typename R2>
__device__ __forceinline__ void ex_head_one(const ExGroups& G, const int64_t blk, double (*red)[3], HeadState* hs) {
    const DtArgs& D = *G.head;
    const int64_t total = G.it_links[G.ngroups];
    const int64_t i0 = blk * blockDim.x;
    const int64_t i1 = i0 + blockDim.x - 1 < total - 1 ? i0 + blockDim.x - 1 : total - 1;
    bool firstA, firstB;
    const int mA = ex_link_member(G, i0, firstA);
    const int mB = ex_link_member(G, i1, firstB);
    // the block's members lie in at most two dt groups (checked when the engine is created)
    const int gA = G.dtg_of[mA], gB = G.dtg_of[mB];
    const bool two = gB != gA;
    // this thread's item; threads past the last item load the last item's data and write nothing
    const bool mine = i0 + (int64_t)threadIdx.x < total;
    const int64_t it = mine ? i0 + (int64_t)threadIdx.x : total - 1;
    const int gi = ex_group(G, G.it_links, it);
    const ExchangeArgs& a = G.a[gi];
    const int64_t loc = it - G.it_links[gi];
    const int b = (int)(loc / a.L.nu), k = (int)(loc - (int64_t)b * a.L.nu);
    const bool single = a.L.u_ptr[k + 1] - a.L.u_ptr[k] == 1;
    const DtGroup GA = D.groups[gA];
    const DtGroup GB = D.groups[gB];
    HeadLoads vA, vB;
    head_loads(D, GA, vA);
    if (two) head_loads(D, GB, vB);
    // the clocks of the members whose control entries and forcing this block writes
    const double tA = D.ctl.t[mA], tB = D.ctl.t[mB];
    ExPre<R2> e;
    ex_one_pre<R2>(a, b, k, e);        // a cell with several links redoes its item after the head
    head_reduce(D, GA, vA, red, hs[0]);
    if (two) head_reduce(D, GB, vB, red, hs[1]);
    // every member whose first item lies in this block: mA if its first item is i0, and the rest
    for (int mm = firstA ? mA : mA + 1; mm <= mB; ++mm) {
        const int j = mm < GA.m1 ? 0 : 1;
        const HeadState& h = hs[j];
        const int gim = (j == 0 ? GA : GB).gi;
        const double tm = mm == mA ? tA : (mm == mB ? tB : D.ctl.t[mm]);
        if (threadIdx.x == 0) head_member_ctl(D, mm, h, nullptr, tm);
        if (h.go && threadIdx.x >= 64 && threadIdx.x < 73) head_member_pm(D, gim, mm, threadIdx.x - 64, h.dt, tm);
    }
    if (!mine) return;
    const HeadState& h = a.member0 + b < GA.m1 ? hs[0] : hs[1];
    if (!h.go) return;
    if (single) ex_one_post<R2>(a, b, k, e, h.dt);
    else ex_one_item<R2>(a, b, k, h.dt);
}

template <typename R2>
__global__ void __launch_bounds__(256) k_ex_links(const ExGroups G) {
    __shared__ double red[4][3];
    __shared__ HeadState hs[2];
    __shared__ double fred[4][FIN_NV];
    __shared__ double ftot[FIN_NV];
    const int64_t total = G.it_links[G.ngroups];
    // the previous step's ledger reduction takes the first blocks, so it is dispatched first and
    // overlaps the link blocks
    const int nfin = G.fin ? G.fin_members : 0;
    if ((int)blockIdx.x < nfin) {
        finalize_member<256>(*G.fin, (int)blockIdx.x, fred, ftot);
        return;
    }
    const int64_t blk = (int64_t)blockIdx.x - nfin;
    if (G.head && G.one_phase) {
        ex_head_one<R2>(G, blk, red, hs);
        return;
    }
    int mA = -1, mB = -1, mA_end = 0;
    if (G.head) {
        // the block's members lie in at most two dt groups (checked when the engine is created):
        // evaluate their dt here
        const int64_t i0 = blk * blockDim.x;
        const int64_t i1 = i0 + blockDim.x - 1 < total - 1 ? i0 + blockDim.x - 1 : total - 1;
        bool firstA, firstB;
        mA = ex_link_member(G, i0, firstA);
        mB = ex_link_member(G, i1, firstB);
        const int gA = G.dtg_of[mA], gB = G.dtg_of[mB];
        mA_end = G.head->groups[gA].m1;
        head_group_dt(*G.head, G.head->groups[gA], red, hs[0]);
        if (gB != gA) head_group_dt(*G.head, G.head->groups[gB], red, hs[1]);
        // the block holding a member's first item writes its control entries and its forcing
        for (int mm = firstA ? mA : mA + 1; mm <= mB; ++mm) {
            const HeadState& h = hs[mm < mA_end ? 0 : 1];
            const int gim = G.head->groups[G.dtg_of[mm]].gi;
            const double tm = G.head->ctl.t[mm];
            if (threadIdx.x == 0) head_member_ctl(*G.head, mm, h, nullptr, tm);
            if (h.go && threadIdx.x >= 64 && threadIdx.x < 73) head_member_pm(*G.head, gim, mm, threadIdx.x - 64, h.dt, tm);
        }
    }
    int64_t it = blk * (int64_t)blockDim.x + threadIdx.x;
    if (it >= total) return;
    const int gi = ex_group(G, G.it_links, it);
    const ExchangeArgs& a = G.a[gi];
    it -= G.it_links[gi];
    const int per = G.one_phase ? a.L.nu : a.L.nl + a.L.nu;
    const int b = (int)(it / per), r = (int)(it - (int64_t)b * per);
    const int m = a.member0 + b;
    double dt;
    if (G.head) {
        const HeadState& h = m < mA_end ? hs[0] : hs[1];
        if (!h.go) return;
        dt = h.dt;
    } else {
        if (a.active && !a.active[m]) return;
        dt = a.dt[m];
    }
    if (G.one_phase) ex_one_item<R2>(a, b, r, dt);
    else ex_link_item<R2>(a, b, r, dt);
}

template <typename R2>
__global__ void __launch_bounds__(256) k_ex_apply(const ExGroups G) {
    int64_t it = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (it >= G.it_cells[G.ngroups]) return;
    const int gi = ex_group(G, G.it_cells, it);
    const ExchangeArgs& a = G.a[gi];
    it -= G.it_cells[gi];
    const int per = a.L.nu + a.N.nc;
    const int b = (int)(it / per), r = (int)(it - (int64_t)b * per);
    const int m = a.member0 + b;
    if (a.active && !a.active[m]) return;
    if (G.rate_reset && r == 0) G.rate_reset[m] = 0.0;   // every dt of this step has been taken
    ex_apply_item<R2>(a, b, r);
}

template <int NT>
__global__ void __launch_bounds__(NT) k_ex_ledger(const ExchangeArgs a) {
    __shared__ double red[NT / 64][8];
    const int b = blockIdx.x;
    const int m = a.member0 + b;
    if (a.active && !a.active[m]) return;
    ExWs w = ex_ws(a, b);
    double part[2] = {0.0, 0.0};
    for (int l = threadIdx.x; l < a.L.nl; l += NT) {
        const double v = w.Vf[l];
        part[0] += clamp0(v);
        part[1] += clamp0(-v);
    }
    block_sum_bcast<NT, 2>(part, red);
    if (threadIdx.x == 0) {
        a.acc[(int64_t)a.acc_slot_in * a.acc_stride + m] += part[0];
        a.acc[(int64_t)(a.acc_slot_in + 1) * a.acc_stride + m] += part[1];
    }
}

int launch_exchange_groups(ExGroups& G, hipStream_t s) {
    if (G.ngroups < 1) return 0;
    G.it_links[0] = G.it_cells[0] = 0;
    for (int gi = 0; gi < G.ngroups; ++gi) {
        const ExchangeArgs& a = G.a[gi];
        if (a.dtype2d != G.a[0].dtype2d) return HIS_E_ARG;
        G.it_links[gi + 1] = G.it_links[gi] + (int64_t)a.N.B * (G.one_phase ? a.L.nu : a.L.nl + a.L.nu);
        G.it_cells[gi + 1] = G.it_cells[gi] + (int64_t)a.N.B * (a.L.nu + a.N.nc);
    }
    auto blocks = [](int64_t n) { return dim3((unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1)); };
    dim3 g1 = blocks(G.it_links[G.ngroups]);
    const dim3 g2 = blocks(G.it_cells[G.ngroups]);
    if (G.fin) g1.x += (unsigned)G.fin_members;
    if (G.a[0].dtype2d == HIS_F64) hipLaunchKernelGGL(k_ex_links<double>, g1, dim3(256), 0, s, G);
    else hipLaunchKernelGGL(k_ex_links<float>, g1, dim3(256), 0, s, G);
    HIS_HIP(hipGetLastError());
    if (!G.one_phase) {
        if (G.a[0].dtype2d == HIS_F64) hipLaunchKernelGGL(k_ex_apply<double>, g2, dim3(256), 0, s, G);
        else hipLaunchKernelGGL(k_ex_apply<float>, g2, dim3(256), 0, s, G);
        HIS_HIP(hipGetLastError());
    }
    for (int gi = 0; gi < G.ngroups; ++gi) {
        const ExchangeArgs& a = G.a[gi];
        if (!a.acc) continue;
        hipLaunchKernelGGL(k_ex_ledger<256>, dim3(a.N.B), dim3(256), 0, s, a);
        HIS_HIP(hipGetLastError());
    }
    return 0;
}

int launch_exchange(const ExchangeArgs& a, hipStream_t s) {
    ExGroups G{};
    G.ngroups = 1;
    G.a[0] = a;
    return launch_exchange_groups(G, s);
}

// ------------------------------------------------------------------ plausibility

template <typename R2, int NT>
__global__ void __launch_bounds__(NT) k_plaus(const R2* h, int64_t plane, const his_net1d N, const double* Ag,
                                               double* out, int B) {
    __shared__ double red[NT / 64][8];
    const int b = blockIdx.x;
    double dmax = -1e300, rise = -1e300, nonfin = 0.0;
    for (int64_t i = threadIdx.x; i < plane; i += NT) {
        const double v = (double)h[(int64_t)b * plane + i];
        if (!isfinite(v)) nonfin = 1.0;
        dmax = v > dmax ? v : dmax;
    }
    if (Ag) {
        for (int c = threadIdx.x; c < N.nc; c += NT) {
            if (N.kind[c] >= 2) continue;
            const Sec s = cell_sec(N, b, c);
            const double a = Ag[(int64_t)b * N.nc + c];
            const double zb = N.z[(int64_t)b * N.bs + c];
            const double lev = zb + sec_depth(s, a);
            const double r = lev - N.zbank[c];
            if (!isfinite(r)) nonfin = 1.0;
            rise = r > rise ? r : rise;
        }
    }
    dmax = wave_max(dmax);
    rise = wave_max(rise);
    nonfin = wave_max(nonfin);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[w][0] = dmax; red[w][1] = rise; red[w][2] = nonfin; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a0 = -1e300, a1 = -1e300, a2 = 0.0;
        for (int w2 = 0; w2 < NT / 64; ++w2) {
            a0 = red[w2][0] > a0 ? red[w2][0] : a0;
            a1 = red[w2][1] > a1 ? red[w2][1] : a1;
            a2 = red[w2][2] > a2 ? red[w2][2] : a2;
        }
        out[b] = a2 > 0.0 ? __builtin_nan("") : a0;
        out[B + b] = a2 > 0.0 ? __builtin_nan("") : (Ag ? a1 : -1e9);
    }
}

int launch_plaus(int dtype, const void* h, int64_t plane, int B, const his_net1d* N, const double* A, double* out,
                 hipStream_t s) {
    his_net1d dummy{};
    const his_net1d& n = N ? *N : dummy;
    if (dtype == HIS_F64)
        hipLaunchKernelGGL((k_plaus<double, 256>), dim3(B), dim3(256), 0, s, (const double*)h, plane, n,
                           N ? A : nullptr, out, B);
    else
        hipLaunchKernelGGL((k_plaus<float, 256>), dim3(B), dim3(256), 0, s, (const float*)h, plane, n,
                           N ? A : nullptr, out, B);
    HIS_HIP(hipGetLastError());
    return 0;
}

}  // namespace his
