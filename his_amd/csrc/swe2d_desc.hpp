// Internal launch descriptors of the fused 2-D finite-volume kernels.
#pragma once
#include "common.hpp"

namespace his {

// Per-tile float64 partial sums, one block of NP slots per stage (stage-1 block, then stage-2 block).
enum {
    P_CLIP = 0,                 // sum of depth added by the h >= 0 clip
    P_WIN, P_EIN, P_SIN, P_NIN, // relu of inflow mass flux through the W, E, S, N boundary faces
    P_WOUT, P_EOUT, P_SOUT, P_NOUT,
    P_INF,                      // infiltrated depth
    P_RAINW,                    // rain depth added (rain_weight runs only)
    P_VOLH, P_VOLS,             // sum of h, sum of s_filled (new state)
    P_RATE,                     // CFL rate maximum (new state)
    P_HMAX,                     // largest depth of the new state (the plausibility guard)
    P_USED
};
constexpr int P_NSUM = P_RATE;  // stage-2 slots that are sums (the rest are maxima)
constexpr int NP = 16;          // slots per stage block
constexpr int NP2 = 2 * NP;     // slots per tile

template <typename R>
struct Group2D {
    int B, ny, nx, order, flux;
    int bc[4];
    int tiles_x, tiles_y, tile0, member0;
    // Dispatch order of the register march: strips in_s0..in_s1 and chunks in_r0..in_r1 are interior
    // (no ghost cell, no boundary row or column); the other `ring` tiles of a member touch an edge.
    // e0 / i0: the group's first index among all edge tiles / all interior tiles of the launch.
    int in_s0, in_s1, in_r0, in_r1, ring, e0, i0;
    R dx, dy, g, h_dry, kp_eps;
    double dxd, dyd, gd, t0, area;
    const R *z, *n, *S, *f0, *fc, *k, *rw;
    int64_t n_bs, S_bs, hz_bs, rw_bs;
    R *h, *hu, *hv, *s, *ic;      // state, updated in place by the last stage
    R *h1, *hu1, *hv1;            // intermediate stage state (or the order-1 result)
};

template <typename R>
struct Launch2D {
    Group2D<R> g[HIS_MAX_GROUPS];
    int ngroups, ntiles, Btot;
    int nedge;                    // tiles that touch a grid edge, over all groups
    int tx, ty;                   // tile shape used to build tiles_x / tiles_y
    const double *dt, *t, *pm;    // (Btot), (Btot), (9, Btot)
    const int32_t *active;        // (Btot) or NULL = every member steps
    double *part;                 // (ntiles, NP2)
    // (Btot) or NULL: the last stage also folds each tile's CFL-rate maximum into rate[m] with an
    // atomic max on the bit pattern (non-negative doubles order as unsigned integers), so the next
    // step's dt does not wait for the partials to be reduced
    unsigned long long *rate;
    // (3, Btot) or NULL: the last stage's first tile of each member records that member's dt, rain
    // rate and a 1 (stepped) for a ledger reduction that runs while the next step's dt is chosen
    double *snap;
    // (Btot) or NULL: stage 1's first tile of each stepping member clears rate_reset[m] (the rates
    // the last stage folds into with atomic max), after every reader of the previous rates has run
    double *rate_reset;
};

// Kernel shapes compiled for each precision; the index into this table is the tuning knob.
//   march = 0: a TX x TY tile per workgroup of NT threads, halo staged in LDS (k_stage)
//   march = 1: one wave marches up a strip of TX = 58 columns for TY rows with LDS row rings (k_march)
//   march = 2: the same walk with each lane's column held in registers, x-neighbours by shuffle (k_wave)
struct TileShape { int tx, ty, nt, march; };
// Chunk heights between 8 and 19 trade the march's warm-up rows (H + 6 row steps for H rows)
// against rounds of waves: on 400 x 325 grids H = 13 divides the rows exactly, and at 32 members
// H = 19 makes 4,032 waves, one round of the 4,096 wave slots (1,024 SIMDs x 4).
constexpr TileShape kShapesF32[] = {{58, 16, 64, 2}, {58, 19, 64, 2}, {58, 8, 64, 2}, {58, 16, 64, 1},
                                    {64, 8, 256, 0}, {64, 16, 256, 0}, {58, 10, 64, 2}, {58, 12, 64, 2},
                                    {58, 13, 64, 2}};
// float64 runs mostly at B = 1, where small tiles are what fills the chip
constexpr TileShape kShapesF64[] = {{58, 16, 64, 2}, {58, 32, 64, 2}, {58, 8, 64, 2}, {58, 16, 64, 1},
                                    {32, 8, 256, 0}, {64, 4, 256, 0}, {16, 8, 128, 0}, {32, 4, 128, 0},
                                    {58, 4, 64, 2}};
constexpr int kNumShapes = 9;
static_assert(sizeof(kShapesF32) / sizeof(TileShape) == kNumShapes, "float32 shape table");
static_assert(sizeof(kShapesF64) / sizeof(TileShape) == kNumShapes, "float64 shape table");

// host launchers, defined once per precision (swe2d_f32.hip, swe2d_f64.hip)
int launch_step(const Launch2D<float>& L, int shape, hipStream_t s);
int launch_step(const Launch2D<double>& L, int shape, hipStream_t s);
int launch_state_reduce(const Launch2D<float>& L, int shape, hipStream_t s);   // rate + volumes into stage-2 slots
int launch_state_reduce(const Launch2D<double>& L, int shape, hipStream_t s);
int launch_velocities(const Launch2D<float>& L, float* u, float* v, hipStream_t s);
int launch_velocities(const Launch2D<double>& L, double* u, double* v, hipStream_t s);
int launch_fold_max(const Launch2D<float>& L, int group, float* hmax, hipStream_t s);
int launch_fold_max(const Launch2D<double>& L, int group, double* hmax, hipStream_t s);

// Per member: reduce the member's tile partials in fixed order.
//   vols (5, Btot):  bnd_in, bnd_out, rain, infil, clip     (may be NULL)
//   acc  (10, Btot): += rain, infil, bnd_in, bnd_out at slots 0, 1, 2, 3 and clip at slot 6 (may be NULL)
//   rate (Btot), volume (2, Btot)                            (may be NULL)
struct FinalizeArgs {
    const double* part;
    const double *dt, *pm;
    const int32_t* active;
    double *vols, *acc, *rate, *volume;
    int Btot;
    int ngroups;
    int member0[HIS_MAX_GROUPS], B[HIS_MAX_GROUPS], tile0[HIS_MAX_GROUPS], per[HIS_MAX_GROUPS];
    int order[HIS_MAX_GROUPS], weighted[HIS_MAX_GROUPS];
    double dx[HIS_MAX_GROUPS], dy[HIS_MAX_GROUPS], area[HIS_MAX_GROUPS], cells[HIS_MAX_GROUPS];
    int rain_f32[HIS_MAX_GROUPS];        // rain * dt formed in float32 (the working dtype) before widening
    int stage_only;                      // 1 = only rate/volume (after a state reduction)
    const int32_t* only;                 // stage_only: members with only[m] == 0 are left untouched
    double* snap;                        // (3, Btot) or NULL: dt, rain and stepped flag of the step reduced
                                         // (recorded by the last stage; the flag is cleared here)
    // engine: members of groups without a 1-D network advance here, t_out = t_step + dt, steps += 1
    double* t_out;
    const double* t_step;
    int32_t* steps;
    int advance[HIS_MAX_GROUPS];
    // (Btot) or NULL: a member whose depth or storage sum is not finite gets error code 3 (unless it
    // carries one already).  The float32 kernels are built without NaN semantics, so a NaN cell can
    // drop out of a max; the float64 sums it enters cannot lose it.
    int32_t* error;
};
int launch_finalize(const FinalizeArgs& f, hipStream_t s);

// blockIdx -> logical tile so that each XCD (blocks are dealt round-robin over 8) walks a
// contiguous run of tiles and neighbouring tiles share that XCD's L2 for their halos
__host__ __device__ inline int xcd_tile(int p, int n) {
    int q = n >> 3, r = n & 7, x = p & 7;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + (p >> 3);
}

// blockIdx -> tile for the register march.  A tile that touches a grid edge takes about 1.5 times
// as long as an interior one (ghost layers, boundary sums), so blocks 0 .. nedge-1 take the edge
// tiles of every group and the rest the interior tiles, each class spread over the XCDs as
// xcd_tile does.  A slot that frees up gets the next block, so short interior tiles fill the end of
// the launch.  Only the order changes: tile numbering, partials and their reduction are the same.
template <typename R>
__host__ __device__ inline int edge_first_tile(const Launch2D<R>& L, int p, int& gi) {
    const bool edge = p < L.nedge;
    const int k = edge ? xcd_tile(p, L.nedge) : xcd_tile(p - L.nedge, L.ntiles - L.nedge);
    gi = 0;
    if (edge) {
        while (gi + 1 < L.ngroups && k >= L.g[gi + 1].e0) ++gi;
    } else {
        while (gi + 1 < L.ngroups && k >= L.g[gi + 1].i0) ++gi;
    }
    const Group2D<R>& G = L.g[gi];
    const int X = G.tiles_x, per = X * G.tiles_y;
    const int ns = G.in_s1 - G.in_s0 + 1;
    int b, r, s;
    if (edge) {
        const int q = k - G.e0;
        b = q / G.ring;
        int c = q - b * G.ring;
        const int below = G.in_r0 * X, above = (G.tiles_y - 1 - G.in_r1) * X;
        if (G.ring == per || c < below) {
            r = c / X; s = c % X;
        } else if ((c -= below) < above) {
            r = G.in_r1 + 1 + c / X; s = c % X;
        } else {
            c -= above;
            const int w = X - ns;            // edge strips beside the interior rows (in_s0 >= 1)
            r = G.in_r0 + c / w;
            const int cc = c % w;
            s = cc < G.in_s0 ? cc : G.in_s1 + 1 + (cc - G.in_s0);
        }
    } else {
        const int q = k - G.i0;
        const int inner = per - G.ring;
        b = q / inner;
        const int c = q - b * inner;
        r = G.in_r0 + c / ns;
        s = G.in_s0 + c % ns;
    }
    return G.tile0 + (b * G.tiles_y + r) * X + s;
}

// Tiles of each group for a tile shape, and the two dispatch classes of the register march: a tile
// is interior when its columns x0-3 .. x0+tx+2 and rows y0-3 .. y0+ty+2 lie inside the grid, which
// holds for strips 1 .. (nx-tx-3)/tx and chunks 1 .. (ny-ty-3)/ty (a rectangle, possibly empty).
template <typename R>
inline void tiles_for(Launch2D<R>& L, int tx, int ty) {
    int t = 0, ne = 0, ni = 0;
    L.tx = tx;
    L.ty = ty;
    for (int gi = 0; gi < L.ngroups; ++gi) {
        Group2D<R>& G = L.g[gi];
        G.tiles_x = (G.nx + tx - 1) / tx;
        G.tiles_y = (G.ny + ty - 1) / ty;
        G.tile0 = t;
        const int per = G.tiles_x * G.tiles_y;
        t += G.B * per;
        G.in_s0 = 1;
        G.in_s1 = G.nx - tx - 3 >= 0 ? (G.nx - tx - 3) / tx : 0;
        G.in_r0 = 1;
        G.in_r1 = G.ny - ty - 3 >= 0 ? (G.ny - ty - 3) / ty : 0;
        const int ns = G.in_s1 - G.in_s0 + 1, nr = G.in_r1 - G.in_r0 + 1;
        if (ns <= 0 || nr <= 0) {           // no interior tile: every tile is an edge tile
            G.in_s1 = 0; G.in_r1 = 0;
            G.ring = per;
        } else {
            G.ring = per - ns * nr;
        }
        G.e0 = ne;
        G.i0 = ni;
        ne += G.B * G.ring;
        ni += G.B * (per - G.ring);
    }
    L.ntiles = t;
    L.nedge = ne;
}

}  // namespace his
