// C ABI entry points and the device-resident coupled engine.
//
// The coupled step keeps the order of CoupledEngine.step (hydrointel/solver/coupling.py:201-227):
//   dt from the pre-exchange state (2-D CFL, clamped; never past the output time),
//   n_sub from the 1-D CFL, the weir exchange, the 2-D step with rain(t) and edge values at t and
//   t+dt, then n_sub 1-D substeps at t + k*dt/n_sub.
// Nothing crosses to the host: the CFL rates are reduced in the producing kernels' epilogues, the
// forcing is evaluated from device tables, and the float64 ledger lives on the device.  One step
// (or several) is captured into a hipGraph; the 1-D network branch runs beside the 2-D stages, since
// after the exchange the two models do not read each other until the next step.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "finalize_dev.hpp"
#include "step_head.hpp"

namespace his {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char* where) {
    set_error("%s: %s", where, hipGetErrorString(e));
    return HIS_E_HIP;
}

int launch_plaus(int dtype, const void* h, int64_t plane, int B, const his_net1d* N, const double* A, double* out,
                 hipStream_t s);

// ------------------------------------------------------------------ finalize

template <int NT>
__global__ void __launch_bounds__(NT) k_finalize(const FinalizeArgs F) {
    __shared__ double red[NT / 64][FIN_NV];
    __shared__ double tot[FIN_NV];
    finalize_member<NT>(F, blockIdx.x, red, tot);
}

int launch_finalize(const FinalizeArgs& f, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize<256>, dim3(f.Btot), dim3(256), 0, s, f);
    HIS_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------ building launches

// Tile-shape index per dtype (HIS_F64, HIS_F32).  The defaults are the shapes the autotuner picks
// on the 400 x 325 grids: 32 x 8 LDS tiles in float64, the register march in 8-row chunks in float32.
constexpr int kDefaultShape[2] = {4, 2};
static int g_shape[2] = {kDefaultShape[0], kDefaultShape[1]};

static int shape_of(int dtype, int* tx, int* ty) {
    const TileShape& t = dtype == HIS_F64 ? kShapesF64[g_shape[0]] : kShapesF32[g_shape[1]];
    *tx = t.tx;
    *ty = t.ty;
    return dtype == HIS_F64 ? g_shape[0] : g_shape[1];
}

static int check_grid(const his_grid2d* g) {
    if (!g) { set_error("grid is NULL"); return HIS_E_ARG; }
    if (g->dtype != HIS_F64 && g->dtype != HIS_F32) { set_error("dtype must be HIS_F64 or HIS_F32"); return HIS_E_ARG; }
    if (g->B < 1 || g->ny < 1 || g->nx < 1) { set_error("bad grid size B=%d ny=%d nx=%d", g->B, g->ny, g->nx); return HIS_E_ARG; }
    if (g->order != 1 && g->order != 2) { set_error("order must be 1 or 2"); return HIS_E_ARG; }
    for (int e = 0; e < 4; ++e)
        if (g->bc[e] < 0 || g->bc[e] > 3) { set_error("bad edge kind %d", g->bc[e]); return HIS_E_ARG; }
    if (!g->z || !g->manning) { set_error("z and manning are required"); return HIS_E_ARG; }
    if (g->f0 && (!g->fc || !g->k)) { set_error("Horton needs f0, fc and k"); return HIS_E_ARG; }
    return 0;
}

template <typename R>
static void fill_group(Group2D<R>& G, const his_grid2d& g, const his_state2d& st, R* tmp, int member0) {
    G.B = g.B; G.ny = g.ny; G.nx = g.nx; G.order = g.order; G.flux = g.flux;
    for (int e = 0; e < 4; ++e) G.bc[e] = g.bc[e];
    G.member0 = member0;
    G.dx = (R)g.dx; G.dy = (R)g.dy; G.g = (R)g.g; G.h_dry = (R)g.h_dry; G.kp_eps = (R)g.kp_eps;
    G.dxd = g.dx; G.dyd = g.dy; G.gd = g.g; G.t0 = g.horton_t0; G.area = g.dx * g.dy;
    G.z = (const R*)g.z; G.n = (const R*)g.manning; G.S = (const R*)g.storage;
    G.f0 = (const R*)g.f0; G.fc = (const R*)g.fc; G.k = (const R*)g.k; G.rw = (const R*)g.rain_weight;
    G.n_bs = g.manning_bs; G.S_bs = g.storage_bs; G.hz_bs = g.horton_bs; G.rw_bs = g.rain_weight_bs;
    G.h = (R*)st.h; G.hu = (R*)st.hu; G.hv = (R*)st.hv; G.s = (R*)st.s_filled; G.ic = (R*)st.infil_cum;
    const int64_t n = (int64_t)g.B * g.ny * g.nx;
    G.h1 = tmp; G.hu1 = tmp ? tmp + n : nullptr; G.hv1 = tmp ? tmp + 2 * n : nullptr;
}

static int64_t max_tiles(const his_grid2d& g) {
    const TileShape* T = g.dtype == HIS_F64 ? kShapesF64 : kShapesF32;
    int64_t best = 0;
    for (int k = 0; k < kNumShapes; ++k) {
        int64_t t = (int64_t)g.B * ((g.nx + T[k].tx - 1) / T[k].tx) * ((g.ny + T[k].ty - 1) / T[k].ty);
        best = t > best ? t : best;
    }
    return best;
}

static size_t esize(int dtype) { return dtype == HIS_F64 ? 8 : 4; }

static size_t ws2d_bytes(const his_grid2d& g) {
    const size_t tmp = align_up(3 * (size_t)g.B * g.ny * g.nx * esize(g.dtype));
    const size_t part = align_up((size_t)max_tiles(g) * NP2 * sizeof(double));
    return tmp + part;
}

static void fill_finalize(FinalizeArgs& F, int gi, const his_grid2d& g, int member0, int tile0, int per) {
    F.member0[gi] = member0; F.B[gi] = g.B; F.tile0[gi] = tile0; F.per[gi] = per;
    F.order[gi] = g.order; F.weighted[gi] = g.rain_weight != nullptr;
    F.dx[gi] = g.dx; F.dy[gi] = g.dy; F.area[gi] = g.dx * g.dy; F.cells[gi] = (double)((int64_t)g.nx * g.ny);
    F.rain_f32[gi] = g.dtype == HIS_F32;
}

template <typename R>
static int op_step(const his_grid2d* g, his_state2d st, const double* dt_t, const double* pm, double* vols,
                   double* rate, void* ws, hipStream_t s) {
    Launch2D<R> L{};
    L.ngroups = 1;
    L.Btot = g->B;
    R* tmp = (R*)ws;
    fill_group(L.g[0], *g, st, tmp, 0);
    int tx, ty;
    const int shape = shape_of(g->dtype, &tx, &ty);
    tiles_for(L, tx, ty);
    L.dt = dt_t; L.t = dt_t + g->B; L.pm = pm; L.active = nullptr;
    L.part = (double*)((char*)ws + align_up(3 * (size_t)g->B * g->ny * g->nx * sizeof(R)));
    int rc = launch_step(L, shape, s);
    if (rc) return rc;
    FinalizeArgs F{};
    F.part = L.part; F.dt = dt_t; F.pm = pm; F.vols = vols; F.rate = rate; F.Btot = g->B; F.ngroups = 1;
    fill_finalize(F, 0, *g, 0, 0, L.g[0].tiles_x * L.g[0].tiles_y);
    return launch_finalize(F, s);
}

template <typename R>
static int op_reduce(const his_grid2d* g, const void* h, const void* hu, const void* hv, const void* sfill,
                     double* rate, double* volume, void* ws, hipStream_t s) {
    Launch2D<R> L{};
    L.ngroups = 1;
    L.Btot = g->B;
    his_state2d st{const_cast<void*>(h), const_cast<void*>(hu), const_cast<void*>(hv), const_cast<void*>(sfill),
                   nullptr};
    fill_group(L.g[0], *g, st, (R*)nullptr, 0);
    int tx, ty;
    const int shape = shape_of(g->dtype, &tx, &ty);
    tiles_for(L, tx, ty);
    L.part = (double*)((char*)ws + align_up(3 * (size_t)g->B * g->ny * g->nx * sizeof(R)));
    int rc = launch_state_reduce(L, shape, s);
    if (rc) return rc;
    FinalizeArgs F{};
    F.part = L.part; F.rate = rate; F.volume = volume; F.Btot = g->B; F.ngroups = 1; F.stage_only = 1;
    fill_finalize(F, 0, *g, 0, 0, L.g[0].tiles_x * L.g[0].tiles_y);
    return launch_finalize(F, s);
}

}  // namespace his

using namespace his;

// ------------------------------------------------------------------ C ABI: 2-D

extern "C" size_t his_swe2d_workspace_bytes(const his_grid2d* g) {
    if (check_grid(g)) return 0;
    return ws2d_bytes(*g);
}

extern "C" int his_swe2d_step(const his_grid2d* g, his_state2d st, const double* dt_t, const double* pm,
                              double* vols, double* rate, void* ws, size_t ws_bytes, void* stream) {
    int rc = check_grid(g);
    if (rc) return rc;
    if (!st.h || !st.hu || !st.hv || !st.s_filled || !st.infil_cum || !dt_t || !pm || !vols) {
        set_error("his_swe2d_step: a state, control or output pointer is NULL");
        return HIS_E_ARG;
    }
    if (!ws || ws_bytes < ws2d_bytes(*g)) {
        set_error("his_swe2d_step: workspace of %zu bytes, %zu needed", ws_bytes, ws2d_bytes(*g));
        return HIS_E_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    return g->dtype == HIS_F64 ? op_step<double>(g, st, dt_t, pm, vols, rate, ws, s)
                               : op_step<float>(g, st, dt_t, pm, vols, rate, ws, s);
}

// ------------------------------------------------------------------ C ABI: 2-D step without allocation
//
// Scratch of his_swe2d_step_noalloc: one file-scope device array, a copy of which every GPU the
// library runs on holds from module load until the process ends.  In host code `scratch` names a
// host-side stand-in, not a device address, so it is never passed to a kernel or a copy: the device
// address comes from hipGetSymbolAddress, taken once per device with that device current and kept in
// g_scratch.  op_step writes every stage temporary and every tile partial before reading it, so
// nothing in the array has to start at zero, and a call never sees what an earlier one left.
static __device__ __align__(256) unsigned char scratch[HIS_SWE2D_NOALLOC_MAX_BYTES];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_of_current_device(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "his_swe2d_step_noalloc: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("his_swe2d_step_noalloc: device %d is beyond the %d devices the scratch table holds", dev,
                  kMaxDevices);
        return HIS_E_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "his_swe2d_step_noalloc: hipGetSymbolAddress(scratch)");
        if (!p) { set_error("his_swe2d_step_noalloc: no device address for the scratch array"); return HIS_E_HIP; }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return 0;
}
}  // namespace

extern "C" int his_swe2d_step_noalloc(const his_grid2d* g, his_state2d st, const double* dt_t, const double* pm,
                                      double* vols, double* rate, void* stream) {
    int rc = check_grid(g);
    if (rc) return rc;
    if (!st.h || !st.hu || !st.hv || !st.s_filled || !st.infil_cum || !dt_t || !pm || !vols) {
        set_error("his_swe2d_step_noalloc: a state, control or output pointer is NULL");
        return HIS_E_ARG;
    }
    const size_t need = ws2d_bytes(*g);
    if (need > (size_t)HIS_SWE2D_NOALLOC_MAX_BYTES) {
        set_error("his_swe2d_step_noalloc: this grid needs %zu bytes of scratch, more than the %zu the library "
                  "holds per GPU; use his_swe2d_step with a workspace", need, (size_t)HIS_SWE2D_NOALLOC_MAX_BYTES);
        return HIS_E_WORKSPACE;
    }
    void* ws = nullptr;
    rc = scratch_of_current_device(&ws);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    return g->dtype == HIS_F64 ? op_step<double>(g, st, dt_t, pm, vols, rate, ws, s)
                               : op_step<float>(g, st, dt_t, pm, vols, rate, ws, s);
}

extern "C" int his_swe2d_max_rate(const his_grid2d* g, const void* h, const void* hu, const void* hv, double* rate,
                                  void* ws, size_t ws_bytes, void* stream) {
    int rc = check_grid(g);
    if (rc) return rc;
    if (!h || !hu || !hv || !rate) { set_error("his_swe2d_max_rate: NULL pointer"); return HIS_E_ARG; }
    if (!ws || ws_bytes < ws2d_bytes(*g)) { set_error("his_swe2d_max_rate: workspace too small"); return HIS_E_WORKSPACE; }
    // the reduction also sums s_filled; h stands in for it here and that sum is discarded
    hipStream_t s = (hipStream_t)stream;
    return g->dtype == HIS_F64 ? op_reduce<double>(g, h, hu, hv, h, rate, nullptr, ws, s)
                               : op_reduce<float>(g, h, hu, hv, h, rate, nullptr, ws, s);
}

extern "C" int his_swe2d_volumes(const his_grid2d* g, const void* h, const void* s_filled, double* out, void* ws,
                                 size_t ws_bytes, void* stream) {
    int rc = check_grid(g);
    if (rc) return rc;
    if (!h || !s_filled || !out) { set_error("his_swe2d_volumes: NULL pointer"); return HIS_E_ARG; }
    if (!ws || ws_bytes < ws2d_bytes(*g)) { set_error("his_swe2d_volumes: workspace too small"); return HIS_E_WORKSPACE; }
    // the reduction also forms the CFL rate; velocities are taken as zero (hu = hv = h is not used)
    hipStream_t s = (hipStream_t)stream;
    return g->dtype == HIS_F64 ? op_reduce<double>(g, h, h, h, s_filled, nullptr, out, ws, s)
                               : op_reduce<float>(g, h, h, h, s_filled, nullptr, out, ws, s);
}

extern "C" int his_swe2d_velocities(const his_grid2d* g, const void* h, const void* hu, const void* hv, void* u,
                                    void* v, void* stream) {
    int rc = check_grid(g);
    if (rc) return rc;
    if (!h || !hu || !hv || !u || !v) { set_error("his_swe2d_velocities: NULL pointer"); return HIS_E_ARG; }
    his_state2d st{const_cast<void*>(h), const_cast<void*>(hu), const_cast<void*>(hv), nullptr, nullptr};
    hipStream_t s = (hipStream_t)stream;
    if (g->dtype == HIS_F64) {
        Launch2D<double> L{};
        L.ngroups = 1;
        fill_group(L.g[0], *g, st, (double*)nullptr, 0);
        return launch_velocities(L, (double*)u, (double*)v, s);
    }
    Launch2D<float> L{};
    L.ngroups = 1;
    fill_group(L.g[0], *g, st, (float*)nullptr, 0);
    return launch_velocities(L, (float*)u, (float*)v, s);
}

extern "C" int his_plausibility(const his_grid2d* g, const void* h, const his_net1d* net, const double* A,
                                double* out, void* stream) {
    int rc = check_grid(g);
    if (rc) return rc;
    if (!h || !out || (net && !A)) { set_error("his_plausibility: NULL pointer"); return HIS_E_ARG; }
    return launch_plaus(g->dtype, h, (int64_t)g->ny * g->nx, g->B, net, A, out, (hipStream_t)stream);
}

// ------------------------------------------------------------------ C ABI: 1-D and exchange

static int check_net(const his_net1d* n) {
    if (!n || n->B < 1 || n->nc < 1 || n->nf < 0 || n->ngh < 0) { set_error("bad network descriptor"); return HIS_E_ARG; }
    if (!n->z || !n->b || !n->yb || !n->ts || !n->m || !n->nman || !n->dx || !n->kind || !n->lf_ptr || !n->rf_ptr) {
        set_error("network descriptor has a NULL array");
        return HIS_E_ARG;
    }
    if (n->nf > 0 && (!n->fl || !n->fr || !n->fb || !n->fyb || !n->fts || !n->fm || !n->lf_face || !n->rf_face)) {
        set_error("network descriptor has a NULL face array");
        return HIS_E_ARG;
    }
    if (n->nparts > HIS_MAX_NET_PARTS) {
        set_error("network descriptor splits into %d parts, more than %d", n->nparts, HIS_MAX_NET_PARTS);
        return HIS_E_ARG;
    }
    if (n->nparts > 1 && (!n->p_cptr || !n->p_cell || !n->p_fptr || !n->p_face || !n->p_gptr || !n->p_ghost)) {
        set_error("network descriptor splits into %d parts but a part array is NULL", n->nparts);
        return HIS_E_ARG;
    }
    return 0;
}

extern "C" size_t his_swe1d_workspace_bytes(const his_net1d* n) {
    if (check_net(n)) return 0;
    return swe1d_ws_doubles(*n) * sizeof(double);
}

extern "C" int his_swe1d_step(const his_net1d* n, double* A, double* Q, const double* dt, const double* pm,
                              double* vols, void* ws, size_t ws_bytes, void* stream) {
    int rc = check_net(n);
    if (rc) return rc;
    if (!A || !Q || !dt || !pm || !vols) { set_error("his_swe1d_step: NULL pointer"); return HIS_E_ARG; }
    if (!ws || ws_bytes < swe1d_ws_doubles(*n) * sizeof(double)) {
        set_error("his_swe1d_step: workspace too small");
        return HIS_E_WORKSPACE;
    }
    Net1DDrive d{};
    d.dt = dt; d.pm = pm; d.vols = vols; d.Btot = n->B; d.member0 = 0;
    hipStream_t s = (hipStream_t)stream;
    if ((rc = clear_swe1d_counters(*n, (double*)ws, s))) return rc;
    return launch_swe1d(*n, A, Q, d, (double*)ws, swe1d_part_width(*n), s);
}

extern "C" int his_swe1d_max_rate(const his_net1d* n, const double* A, const double* Q, double* rate,
                                  int32_t* any_wet, void* stream) {
    int rc = check_net(n);
    if (rc) return rc;
    if (!A || !Q || !rate || !any_wet) { set_error("his_swe1d_max_rate: NULL pointer"); return HIS_E_ARG; }
    return launch_swe1d_rate(*n, A, Q, rate, any_wet, 0, (hipStream_t)stream, nullptr);
}

static int check_links(const his_links* l, const his_net1d* n) {
    if (!l || l->nl < 0 || l->nu < 0) { set_error("bad links descriptor"); return HIS_E_ARG; }
    if (l->nl > 0 && (!l->cell1d || !l->inv || !l->uniq || !l->zbank || !l->length || !l->z2 || !l->u_ptr ||
                      !l->u_link || !l->c_ptr || !l->c_link)) {
        set_error("links descriptor has a NULL array");
        return HIS_E_ARG;
    }
    return check_net(n);
}

extern "C" size_t his_exchange_workspace_bytes(const his_links* l, const his_net1d* n) {
    if (check_links(l, n)) return 0;
    return exchange_ws_doubles(*l, *n) * sizeof(double);
}

extern "C" int his_exchange(const his_links* l, const his_net1d* n, int32_t dtype2d, int64_t plane, void* h,
                            void* hu, void* hv, double* A, const double* dt, double* V, double* ex, void* ws,
                            size_t ws_bytes, void* stream) {
    int rc = check_links(l, n);
    if (rc) return rc;
    if (!h || !hu || !hv || !A || !dt) { set_error("his_exchange: NULL pointer"); return HIS_E_ARG; }
    if (dtype2d != HIS_F64 && dtype2d != HIS_F32) { set_error("his_exchange: bad dtype"); return HIS_E_ARG; }
    if (!ws || ws_bytes < exchange_ws_doubles(*l, *n) * sizeof(double)) {
        set_error("his_exchange: workspace too small");
        return HIS_E_WORKSPACE;
    }
    ExchangeArgs a{};
    a.L = *l; a.N = *n; a.dtype2d = dtype2d; a.plane = plane; a.h = h; a.hu = hu; a.hv = hv; a.A = A; a.dt = dt;
    a.V = V; a.acc = ex; a.acc_slot_in = 0; a.acc_stride = n->B; a.ws = (double*)ws; a.member0 = 0;
    return launch_exchange(a, (hipStream_t)stream);
}

// ------------------------------------------------------------------ engine

namespace his {

// One block per shared-dt group: choose dt and n_sub, mark members active, evaluate the forcing.
// Snapshots each member's step start time into t_step, which the step's kernels read while the
// 1-D kernel (or finalize, without a network) advances t.
__global__ void __launch_bounds__(256) k_dt(const DtArgs a) {
    const DtGroup G = a.groups[blockIdx.x];
    __shared__ double red[4][3];
    __shared__ HeadState hs;
    head_group_dt(a, G, red, hs);
    for (int m = G.m0 + threadIdx.x; m < G.m1; m += blockDim.x) head_member_ctl(a, m, hs, a.rate2d_reset, a.ctl.t[m]);
    if (!hs.go) return;
    const int nm = G.m1 - G.m0;
    for (int k = threadIdx.x; k < nm * 9; k += blockDim.x)
        head_member_pm(a, G.gi, G.m0 + k / 9, k % 9, hs.dt, a.ctl.t[G.m0 + k / 9]);
}

}  // namespace his

struct his_engine {
    his_engine_desc d;
    int dtype, shape, Btot;
    Launch2D<float> L32;
    Launch2D<double> L64;
    FinalizeArgs F;
    DtArgs D;
    int n_dtg;
    his_net1d net[HIS_MAX_GROUPS];
    his_links links[HIS_MAX_GROUPS];
    int has_net[HIS_MAX_GROUPS];
    double* ws1d[HIS_MAX_GROUPS];
    double* wsx[HIS_MAX_GROUPS];
    int32_t* active;
    double *pm, *rate2d, *rate1d, *volume, *t_step, *snap;
    int32_t* wet1d;
    hipStream_t cap, side;
    hipEvent_t ev_fork, ev_join;
    int async_fin;            // ledger reduction deferred into the next step's first exchange launch
    FinalizeArgs* fin_dev;    // device copy of F for that deferred reduction
    int fuse_head;            // dt and forcing evaluated inside the first exchange phase
    int one_first;            // capture the 1-D network before the 2-D stages (the 1-D branch is longer)
    int one_first_est;        // the same from the cost estimate at creation (his_engine_set_branch_order -1)
    int ex_one;               // every 1-D cell has at most one link: the exchange is one launch
    DtArgs* head_dev;         // device copy of D for the fused head
    int32_t* dtg_of;          // (Btot) device: dt group of each member
    hipGraphExec_t exec[3];   // one captured graph per entry of kGraphSteps
    int graphs_ready;
    void* step_max[HIS_MAX_GROUPS];   // per grid group: running maximum of h folded every step, or NULL
};

// Whether a step's ledger reduction is still owed.  With async_fin it is not launched after the
// step but carried by the next step's first exchange launch (extra blocks reading the tile
// partials and the stage-2 snapshot), so the step chain needs no other stream; a run or a
// captured graph ends with a standalone reduction of its last step.
struct FinChain {
    bool fin_due = false;
};

static int engine_close_chain(his_engine* e, hipStream_t s, FinChain& c) {
    if (!c.fin_due) return 0;
    c.fin_due = false;
    return launch_finalize(e->F, s);
}

// One coupled step on s: exchange (with the previous step's ledger reduction) -> 2-D stages, the
// 1-D network on `side` (forked after the exchange) -> join.  In a captured graph the exchange's
// first-captured dependent keeps its hardware queue and the other branch pays a cross-queue start,
// so the branch expected to be longer (one_first) is captured first.
static int engine_enqueue_step(his_engine* e, hipStream_t s, hipStream_t side, FinChain& chain) {
    int rc;
    if (!e->fuse_head) {
        hipLaunchKernelGGL(k_dt, dim3(e->n_dtg), dim3(e->n_dtg == 1 ? 256 : 64), 0, s, e->D);
        HIS_HIP(hipGetLastError());
    }
    // every group's weir exchange: one launch per phase
    ExGroups X{};
    Swe1DGroups W{};
    for (int gi = 0; gi < e->d.ngroups; ++gi) {
        if (!e->has_net[gi]) continue;
        const his_group& G = e->d.group[gi];
        const int m0 = e->dtype == HIS_F64 ? e->L64.g[gi].member0 : e->L32.g[gi].member0;
        ExchangeArgs& a = X.a[X.ngroups++];
        a.L = e->links[gi]; a.N = e->net[gi]; a.dtype2d = e->dtype; a.plane = (int64_t)G.grid.ny * G.grid.nx;
        a.h = G.state.h; a.hu = G.state.hu; a.hv = G.state.hv; a.A = G.A; a.dt = e->d.ctl.dt; a.V = nullptr;
        a.acc = nullptr; a.acc_slot_in = 8; a.acc_stride = e->Btot; a.active = e->active; a.ws = e->wsx[gi];
        a.member0 = m0;
        const int k = W.ngroups++;
        W.N[k] = e->net[gi]; W.A[k] = G.A; W.Q[k] = G.Q; W.ws[k] = e->ws1d[gi];
        W.width[k] = swe1d_part_width(e->net[gi]);
        W.blk0[k + 1] = W.blk0[k] + G.grid.B * (e->net[gi].nparts > 1 ? e->net[gi].nparts : 1);
        Net1DDrive& d = W.d[k];
        d.dt = e->d.ctl.dt; d.engine = 1; d.n_sub = e->d.ctl.n_sub; d.t = e->t_step;
        d.t_out = e->d.ctl.t; d.steps = e->d.ctl.steps;
        d.ex_vf = exchange_vf(e->wsx[gi], e->links[gi], e->net[gi], 0);
        d.ex_nl = e->links[gi].nl;
        d.ex_stride = (int)exchange_ws_stride(e->links[gi], e->net[gi]);
        d.inflow = G.inflow; d.stage = G.stage; d.acc = e->d.ctl.acc; d.active = e->active;
        d.rate = e->rate1d; d.wet = e->wet1d; d.Btot = e->Btot;
        d.member0 = m0;
        d.error = e->d.ctl.error;
    }
    X.one_phase = e->ex_one;
    if (e->fuse_head) {
        X.head = e->head_dev;
        X.dtg_of = e->dtg_of;
        // the second phase clears the rates once every block has read them; with a single phase
        // stage 1 of the 2-D step does it (Launch2D::rate_reset)
        X.rate_reset = e->ex_one ? nullptr : e->rate2d;
    }
    if (chain.fin_due) {                       // the previous step's ledger, beside the first phase
        X.fin = e->fin_dev;
        X.fin_members = e->Btot;
        chain.fin_due = false;
    }
    if ((rc = launch_exchange_groups(X, s))) return rc;
    const bool any_net = W.ngroups > 0;
    const bool fork = any_net && side;
    if (fork) HIS_HIP(hipEventRecord(e->ev_fork, s));
    // every group's 1-D network in one launch, on the side stream beside the 2-D stages
    auto one_d = [&]() -> int {
        if (fork) HIS_HIP(hipStreamWaitEvent(side, e->ev_fork, 0));
        return launch_swe1d_groups(W, fork ? side : s);
    };
    if (e->one_first && (rc = one_d())) return rc;
    rc = e->dtype == HIS_F64 ? launch_step(e->L64, e->shape, s) : launch_step(e->L32, e->shape, s);
    if (rc) return rc;
    for (int gi = 0; gi < e->d.ngroups; ++gi) {
        if (!e->step_max[gi]) continue;
        rc = e->dtype == HIS_F64 ? launch_fold_max(e->L64, gi, (double*)e->step_max[gi], s)
                                 : launch_fold_max(e->L32, gi, (float*)e->step_max[gi], s);
        if (rc) return rc;
    }
    if (!e->one_first && (rc = one_d())) return rc;
    if (e->async_fin) {
        // the next step's dt reads the rates the last stage folded in; the ledger reduction rides
        // on the next step's exchange (or closes the chain)
        chain.fin_due = true;
    } else if ((rc = launch_finalize(e->F, s))) {
        return rc;
    }
    if (fork) {
        HIS_HIP(hipEventRecord(e->ev_join, side));
        HIS_HIP(hipStreamWaitEvent(s, e->ev_join, 0));
    }
    return 0;
}

extern "C" size_t his_engine_workspace_bytes(const his_engine_desc* d) {
    if (!d || d->ngroups < 1 || d->ngroups > HIS_MAX_GROUPS) return 0;
    size_t tot = 0;
    int Btot = 0;
    int64_t tiles = 0;
    for (int gi = 0; gi < d->ngroups; ++gi) {
        const his_group& G = d->group[gi];
        if (check_grid(&G.grid)) return 0;
        tot += align_up(3 * (size_t)G.grid.B * G.grid.ny * G.grid.nx * esize(G.grid.dtype));
        tiles += max_tiles(G.grid);
        Btot += G.grid.B;
        if (G.net) {
            tot += align_up(swe1d_ws_doubles(*G.net) * 8);
            if (G.links) tot += align_up(exchange_ws_doubles(*G.links, *G.net) * 8);
        }
    }
    tot += align_up((size_t)tiles * NP2 * 8);
    tot += align_up((size_t)Btot * 4) + align_up((size_t)Btot * 9 * 8) + 3 * align_up((size_t)Btot * 8) +
           align_up((size_t)Btot * 2 * 8) + align_up((size_t)Btot * 4) + align_up((size_t)Btot * 3 * 8) +
           align_up((size_t)Btot * sizeof(DtGroup)) + align_up(sizeof(DtArgs)) + align_up(sizeof(FinalizeArgs)) +
           align_up((size_t)Btot * 4);
    return tot;
}

template <typename R>
static void engine_fill_launch(his_engine* e, Launch2D<R>& L, char*& p) {
    L.ngroups = e->d.ngroups;
    int m0 = 0;
    for (int gi = 0; gi < e->d.ngroups; ++gi) {
        const his_group& G = e->d.group[gi];
        R* tmp = (R*)p;
        p += align_up(3 * (size_t)G.grid.B * G.grid.ny * G.grid.nx * sizeof(R));
        fill_group(L.g[gi], G.grid, G.state, tmp, m0);
        m0 += G.grid.B;
    }
    L.Btot = m0;
    int tx, ty;
    e->shape = shape_of(e->dtype, &tx, &ty);
    tiles_for(L, tx, ty);
}

extern "C" int his_engine_create(const his_engine_desc* d, his_engine** out, void* stream) {
    if (!d || !out) { set_error("his_engine_create: NULL argument"); return HIS_E_ARG; }
    if (d->ngroups < 1 || d->ngroups > HIS_MAX_GROUPS) { set_error("1 to %d groups", HIS_MAX_GROUPS); return HIS_E_ARG; }
    const size_t need = his_engine_workspace_bytes(d);
    if (!need) return HIS_E_ARG;
    if (!d->ws || d->ws_bytes < need) {
        set_error("his_engine_create: workspace of %zu bytes, %zu needed", d->ws_bytes, need);
        return HIS_E_WORKSPACE;
    }
    const int dtype = d->group[0].grid.dtype;
    const int order = d->group[0].grid.order;
    for (int gi = 0; gi < d->ngroups; ++gi) {
        const his_group& G = d->group[gi];
        if (G.grid.dtype != dtype || G.grid.order != order) {
            set_error("every group of an engine must share dtype and spatial order");
            return HIS_E_ARG;
        }
        if ((G.net == nullptr) != (G.links == nullptr)) { set_error("net and links go together"); return HIS_E_ARG; }
        if (G.net && (G.net->B != G.grid.B || !G.A || !G.Q)) {
            set_error("group %d: network batch or A/Q do not match the grid", gi);
            return HIS_E_ARG;
        }
        if (G.dt_group_size < 1 || G.grid.B % G.dt_group_size) {
            set_error("group %d: dt_group_size %d must divide B=%d", gi, G.dt_group_size, G.grid.B);
            return HIS_E_ARG;
        }
        if (!G.state.h || !G.state.hu || !G.state.hv || !G.state.s_filled || !G.state.infil_cum) {
            set_error("group %d: state pointer NULL", gi);
            return HIS_E_ARG;
        }
    }
    const his_control& c = d->ctl;
    if (!c.t || !c.t_limit || !c.t_end || !c.dt || !c.n_sub || !c.steps || !c.clamped || !c.acc || !c.dt_min_seen ||
        !c.dt_max_seen || !c.n_sub_max || !c.error || c.n_sub_limit < 1) {
        set_error("his_engine_create: control block has a NULL array");
        return HIS_E_ARG;
    }
    his_engine* e = new his_engine();
    e->d = *d;
    e->dtype = dtype;
    char* p = (char*)d->ws;
    if (dtype == HIS_F64) engine_fill_launch(e, e->L64, p);
    else engine_fill_launch(e, e->L32, p);
    const int Btot = e->Btot = dtype == HIS_F64 ? e->L64.Btot : e->L32.Btot;
    const int ntiles = dtype == HIS_F64 ? e->L64.ntiles : e->L32.ntiles;
    double* part = (double*)p;
    int64_t tiles_cap = 0;
    for (int gi = 0; gi < d->ngroups; ++gi) tiles_cap += max_tiles(d->group[gi].grid);
    p += align_up((size_t)tiles_cap * NP2 * 8);
    (void)ntiles;
    for (int gi = 0; gi < d->ngroups; ++gi) {
        const his_group& G = d->group[gi];
        e->has_net[gi] = G.net != nullptr;
        if (G.net) {
            e->net[gi] = *G.net;
            e->links[gi] = *G.links;
            e->ws1d[gi] = (double*)p;
            p += align_up(swe1d_ws_doubles(*G.net) * 8);
            e->wsx[gi] = (double*)p;
            p += align_up(exchange_ws_doubles(*G.links, *G.net) * 8);
        }
    }
    e->active = (int32_t*)p; p += align_up((size_t)Btot * 4);
    e->pm = (double*)p; p += align_up((size_t)Btot * 9 * 8);
    e->rate2d = (double*)p; p += align_up((size_t)Btot * 8);
    e->rate1d = (double*)p; p += align_up((size_t)Btot * 8);
    e->t_step = (double*)p; p += align_up((size_t)Btot * 8);
    e->volume = (double*)p; p += align_up((size_t)Btot * 2 * 8);
    e->wet1d = (int32_t*)p; p += align_up((size_t)Btot * 4);
    e->snap = (double*)p; p += align_up((size_t)Btot * 3 * 8);
    DtGroup* dtg_dev = (DtGroup*)p; p += align_up((size_t)Btot * sizeof(DtGroup));
    e->head_dev = (DtArgs*)p; p += align_up(sizeof(DtArgs));
    e->fin_dev = (FinalizeArgs*)p; p += align_up(sizeof(FinalizeArgs));
    e->dtg_of = (int32_t*)p;
    // launches; with a network in every group the 1-D kernel advances t, so the ledger reduction is
    // off the step's critical path and the last stage publishes the CFL rates itself
    e->async_fin = 1;
    for (int gi = 0; gi < d->ngroups; ++gi) e->async_fin &= e->has_net[gi];
    unsigned long long* rate_bits = e->async_fin ? (unsigned long long*)e->rate2d : nullptr;
    if (dtype == HIS_F64) {
        e->L64.dt = c.dt; e->L64.t = e->t_step; e->L64.pm = e->pm; e->L64.active = e->active; e->L64.part = part;
        e->L64.rate = rate_bits;
        e->L64.snap = e->async_fin ? e->snap : nullptr;
    } else {
        e->L32.dt = c.dt; e->L32.t = e->t_step; e->L32.pm = e->pm; e->L32.active = e->active; e->L32.part = part;
        e->L32.rate = rate_bits;
        e->L32.snap = e->async_fin ? e->snap : nullptr;
    }
    FinalizeArgs& F = e->F;
    F = FinalizeArgs{};
    F.part = part; F.dt = c.dt; F.pm = e->pm; F.active = e->active; F.acc = c.acc;
    F.rate = e->async_fin ? nullptr : e->rate2d;
    F.snap = e->async_fin ? e->snap : nullptr;
    F.volume = e->volume; F.Btot = Btot; F.ngroups = d->ngroups;
    F.t_out = c.t; F.t_step = e->t_step; F.steps = c.steps; F.error = c.error;
    std::vector<DtGroup> dtg;
    for (int gi = 0; gi < d->ngroups; ++gi) {
        const his_group& G = d->group[gi];
        const int m0 = dtype == HIS_F64 ? e->L64.g[gi].member0 : e->L32.g[gi].member0;
        const int t0 = dtype == HIS_F64 ? e->L64.g[gi].tile0 : e->L32.g[gi].tile0;
        const int per = dtype == HIS_F64 ? e->L64.g[gi].tiles_x * e->L64.g[gi].tiles_y
                                         : e->L32.g[gi].tiles_x * e->L32.g[gi].tiles_y;
        fill_finalize(F, gi, G.grid, m0, t0, per);
        F.advance[gi] = G.net == nullptr;
        for (int k = 0; k < G.grid.B; k += G.dt_group_size) dtg.push_back(DtGroup{m0 + k, m0 + k + G.dt_group_size, gi, 0});
    }
    e->n_dtg = (int)dtg.size();
    DtArgs& D = e->D;
    D = DtArgs{};
    D.groups = dtg_dev; D.ngroups_dt = e->n_dtg; D.cfl = d->cfl; D.dt_min = d->dt_min; D.dt_max = d->dt_max;
    D.ctl = c; D.active = e->active; D.t_step = e->t_step; D.pm = e->pm; D.rate2d = e->rate2d; D.rate1d = e->rate1d; D.wet1d = e->wet1d;
    // single-phase exchange: second-order engines whose 1-D cells each have at most one link (the
    // rates it leaves to be cleared are cleared by stage 1, which a first-order step does not run)
    e->ex_one = order == 2;
    for (int gi = 0; gi < d->ngroups && e->ex_one; ++gi) {
        if (!e->has_net[gi]) continue;
        const int nc = e->net[gi].nc;
        std::vector<int32_t> cp((size_t)nc + 1);
        hipStream_t s0 = (hipStream_t)stream;
        if (hipMemcpyAsync(cp.data(), e->links[gi].c_ptr, cp.size() * 4, hipMemcpyDeviceToHost, s0) != hipSuccess ||
            hipStreamSynchronize(s0) != hipSuccess) {
            delete e;
            set_error("his_engine_create: reading the link table failed");
            return HIS_E_HIP;
        }
        for (int c = 0; c < nc; ++c)
            if (cp[c + 1] - cp[c] > 1) { e->ex_one = 0; break; }
    }
    // the fused head: every member has first-launch items (links and unique cells, or unique cells
    // alone with one phase), and the members of each block of 256 items lie in at most two dt groups
    e->fuse_head = e->async_fin;
    if (e->fuse_head) {
        std::vector<int32_t> gof(Btot, 0);
        for (int k = 0; k < (int)dtg.size(); ++k)
            for (int m = dtg[k].m0; m < dtg[k].m1; ++m) gof[m] = k;
        std::vector<int64_t> it0(1, 0);
        std::vector<int> per_g, m0_g;
        for (int gi = 0; gi < d->ngroups; ++gi) {
            const int per = e->ex_one ? e->links[gi].nu : e->links[gi].nl + e->links[gi].nu;
            if (per < 1) e->fuse_head = 0;
            per_g.push_back(per < 1 ? 1 : per);
            m0_g.push_back(dtype == HIS_F64 ? e->L64.g[gi].member0 : e->L32.g[gi].member0);
            it0.push_back(it0.back() + (int64_t)d->group[gi].grid.B * per);
        }
        auto member_of = [&](int64_t it) {
            int gi = 0;
            while (gi + 1 < d->ngroups && it >= it0[gi + 1]) ++gi;
            return m0_g[gi] + (int)((it - it0[gi]) / per_g[gi]);
        };
        const int64_t total = it0.back();
        for (int64_t i0 = 0; i0 < total && e->fuse_head; i0 += 256) {
            const int64_t i1 = i0 + 255 < total - 1 ? i0 + 255 : total - 1;
            if (gof[member_of(i1)] - gof[member_of(i0)] > 1) e->fuse_head = 0;
        }
    }
    D.rate2d_reset = e->async_fin && !e->fuse_head ? e->rate2d : nullptr;
    if (e->async_fin && e->ex_one) {
        if (dtype == HIS_F64) e->L64.rate_reset = e->rate2d;
        else e->L32.rate_reset = e->rate2d;
    }
    // which branch of a step is longer: a network part's workgroup takes ~15 us plus ~18 us per cell
    // or face each thread carries (a 4,693-cell network on 1,024 threads: ~100 us), whatever the
    // batch (all members and parts at once); the 2-D stages scale with the member-cells (~23 G/s in
    // float32, ~4 G/s in float64 on this part)
    {
        double cells = 0.0, est_1d_us = 0.0;
        for (int gi = 0; gi < d->ngroups; ++gi) {
            cells += (double)d->group[gi].grid.B * d->group[gi].grid.ny * d->group[gi].grid.nx;
            if (!d->group[gi].net) continue;
            const int w = swe1d_part_width(*d->group[gi].net);
            int nt = 64;
            while (nt < w && nt < 1024) nt *= 2;
            const double us = 15.0 + 18.0 * (double)((w + nt - 1) / nt);
            est_1d_us = us > est_1d_us ? us : est_1d_us;
        }
        const double est_2d_us = cells / (dtype == HIS_F64 ? 4.0e3 : 2.3e4);
        e->one_first = est_2d_us < est_1d_us;
        e->one_first_est = e->one_first;
    }
    D.Btot = Btot;
    for (int gi = 0; gi < d->ngroups; ++gi) {
        const his_group& G = d->group[gi];
        for (int k = 0; k < 4; ++k) { D.bc[gi][k] = G.grid.bc[k]; D.edge[gi][k] = G.edge[k]; }
        D.rain[gi] = G.rain;
        D.member0[gi] = dtype == HIS_F64 ? e->L64.g[gi].member0 : e->L32.g[gi].member0;
        D.has_net[gi] = G.net != nullptr;
        D.cfl_1d[gi] = G.net ? G.net->cfl_1d : 0.45;
        D.dt_max_1d[gi] = G.net ? G.net->dt_max : d->dt_max;
    }
    hipStream_t s = (hipStream_t)stream;
    hipError_t he = hipMemcpyAsync(dtg_dev, dtg.data(), dtg.size() * sizeof(DtGroup), hipMemcpyHostToDevice, s);
    if (he != hipSuccess) { delete e; return hip_fail(he, "hipMemcpyAsync(dt groups)"); }
    std::vector<int32_t> dtg_of(Btot, 0);
    for (int k = 0; k < (int)dtg.size(); ++k)
        for (int m = dtg[k].m0; m < dtg[k].m1; ++m) dtg_of[m] = k;
    he = hipMemcpyAsync(e->head_dev, &e->D, sizeof(DtArgs), hipMemcpyHostToDevice, s);
    if (he == hipSuccess) he = hipMemcpyAsync(e->dtg_of, dtg_of.data(), (size_t)Btot * 4, hipMemcpyHostToDevice, s);
    if (he == hipSuccess) he = hipMemcpyAsync(e->fin_dev, &e->F, sizeof(FinalizeArgs), hipMemcpyHostToDevice, s);
    if (he != hipSuccess) { delete e; return hip_fail(he, "hipMemcpyAsync(step head)"); }
    he = hipMemsetAsync(e->snap, 0, (size_t)Btot * 3 * 8, s);
    if (he != hipSuccess) { delete e; return hip_fail(he, "hipMemsetAsync(ledger snapshot)"); }
    for (int gi = 0; gi < d->ngroups; ++gi) {
        if (!e->has_net[gi]) continue;
        const int rc1 = clear_swe1d_counters(e->net[gi], e->ws1d[gi], s);
        if (rc1) { delete e; return rc1; }
    }
    he = hipStreamSynchronize(s);
    if (he != hipSuccess) { delete e; return hip_fail(he, "hipStreamSynchronize"); }
    if (hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming) != hipSuccess) {
        delete e;
        set_error("his_engine_create: stream/event creation failed");
        return HIS_E_HIP;
    }
    *out = e;
    int rc = his_engine_refresh(e, stream);
    if (rc) return rc;
    return 0;
}

static int engine_follow_shape(his_engine* e);

static int engine_refresh(his_engine* e, const int32_t* mask, hipStream_t s) {
    int rc = engine_follow_shape(e);
    if (rc) return rc;
    if (e->dtype == HIS_F64) rc = launch_state_reduce(e->L64, e->shape, s);
    else rc = launch_state_reduce(e->L32, e->shape, s);
    if (rc) return rc;
    FinalizeArgs F = e->F;
    F.stage_only = 1;
    F.active = nullptr;
    F.acc = nullptr;
    F.only = mask;
    F.rate = e->rate2d;
    F.snap = nullptr;
    if ((rc = launch_finalize(F, s))) return rc;
    for (int gi = 0; gi < e->d.ngroups; ++gi) {
        if (!e->has_net[gi]) continue;
        const his_group& G = e->d.group[gi];
        const int m0 = e->dtype == HIS_F64 ? e->L64.g[gi].member0 : e->L32.g[gi].member0;
        if ((rc = launch_swe1d_rate(e->net[gi], G.A, G.Q, e->rate1d, e->wet1d, m0, s, mask))) return rc;
    }
    return 0;
}

extern "C" int his_engine_refresh(his_engine* e, void* stream) {
    if (!e) { set_error("NULL engine"); return HIS_E_ARG; }
    return engine_refresh(e, nullptr, (hipStream_t)stream);
}

extern "C" int his_engine_refresh_members(his_engine* e, const int32_t* mask, void* stream) {
    if (!e || !mask) { set_error("his_engine_refresh_members: NULL argument"); return HIS_E_ARG; }
    return engine_refresh(e, mask, (hipStream_t)stream);
}

static void engine_drop_graphs(his_engine* e);

// Steps per captured graph.  Each graph ends with a standalone ledger reduction and its launch
// meets the queue empty, which costs about 20 us, so long runs replay the longest graph; the shorter
// ones take the remainder.
static constexpr int kGraphSteps[3] = {1, 8, 32};

static int engine_build_graphs(his_engine* e) {
    for (int k = 0; k < 3; ++k) {
        const int n = kGraphSteps[k];
        hipGraph_t g = nullptr;
        HIS_HIP(hipStreamBeginCapture(e->cap, hipStreamCaptureModeRelaxed));
        int rc = 0;
        FinChain chain;
        for (int i = 0; i < n && !rc; ++i) rc = engine_enqueue_step(e, e->cap, e->side, chain);
        if (!rc) rc = engine_close_chain(e, e->cap, chain);
        hipError_t he = hipStreamEndCapture(e->cap, &g);
        if (!rc && he != hipSuccess) rc = hip_fail(he, "hipStreamEndCapture");
        if (!rc) {
            he = hipGraphInstantiate(&e->exec[k], g, nullptr, nullptr, 0);
            if (he != hipSuccess) rc = hip_fail(he, "hipGraphInstantiate");
        }
        if (g) (void)hipGraphDestroy(g);
        if (rc) {
            for (int j = 0; j < k; ++j) (void)hipGraphExecDestroy(e->exec[j]);   // the graphs made above
            return rc;
        }
    }
    e->graphs_ready = 1;
    return 0;
}

// The tile shape is a library-wide setting (his_set_tile_shape, which the autotuner drives).  An
// engine follows it from its next step or refresh: the tiles, the per-tile partial layout and the
// finalize's tile ranges are rebuilt and the step graphs captured again.  Between calls no ledger
// reduction is pending (every run of steps ends with its own), so the layout can change there.
static int engine_follow_shape(his_engine* e) {
    int tx, ty;
    const int shape = shape_of(e->dtype, &tx, &ty);
    if (shape == e->shape) return 0;
    engine_drop_graphs(e);
    if (e->dtype == HIS_F64) tiles_for(e->L64, tx, ty);
    else tiles_for(e->L32, tx, ty);
    e->shape = shape;
    for (int gi = 0; gi < e->d.ngroups; ++gi) {
        e->F.tile0[gi] = e->dtype == HIS_F64 ? e->L64.g[gi].tile0 : e->L32.g[gi].tile0;
        e->F.per[gi] = e->dtype == HIS_F64 ? e->L64.g[gi].tiles_x * e->L64.g[gi].tiles_y
                                           : e->L32.g[gi].tiles_x * e->L32.g[gi].tiles_y;
    }
    const hipError_t he = hipMemcpy(e->fin_dev, &e->F, sizeof(FinalizeArgs), hipMemcpyHostToDevice);
    if (he != hipSuccess) return hip_fail(he, "engine tile shape: hipMemcpy(finalize arguments)");
    return 0;
}

extern "C" int his_engine_step(his_engine* e, int32_t n_steps, void* stream) {
    if (!e || n_steps < 0) { set_error("his_engine_step: bad argument"); return HIS_E_ARG; }
    hipStream_t s = (hipStream_t)stream;
    if (int rc = engine_follow_shape(e)) return rc;
    if (!e->d.use_graph) {
        FinChain chain;
        for (int i = 0; i < n_steps; ++i) {
            int rc = engine_enqueue_step(e, s, e->side, chain);
            if (rc) return rc;
        }
        return engine_close_chain(e, s, chain);
    }
    if (!e->graphs_ready) {
        int rc = engine_build_graphs(e);
        if (rc) return rc;
    }
    int left = n_steps;
    for (int k = 2; k >= 0; --k) {
        while (left >= kGraphSteps[k]) {
            HIS_HIP(hipGraphLaunch(e->exec[k], s));
            left -= kGraphSteps[k];
        }
    }
    return 0;
}

extern "C" int his_engine_fold_max(his_engine* e, int32_t group, void* hmax, void* stream) {
    if (!e || group < 0 || group >= e->d.ngroups || !hmax) { set_error("his_engine_fold_max: bad argument"); return HIS_E_ARG; }
    hipStream_t s = (hipStream_t)stream;
    if (e->dtype == HIS_F64) return launch_fold_max(e->L64, group, (double*)hmax, s);
    return launch_fold_max(e->L32, group, (float*)hmax, s);
}

// the captured steps no longer match the engine's settings: capture again at the next step
static void engine_drop_graphs(his_engine* e) {
    if (!e->graphs_ready) return;
    (void)hipDeviceSynchronize();   // a replay may still be running on the caller's stream
    for (int k = 0; k < 3; ++k) (void)hipGraphExecDestroy(e->exec[k]);
    e->graphs_ready = 0;
}

extern "C" int his_engine_track_step_max(his_engine* e, int32_t group, void* hmax) {
    if (!e || group < 0 || group >= e->d.ngroups) { set_error("his_engine_track_step_max: bad argument"); return HIS_E_ARG; }
    if (e->step_max[group] != hmax) {
        e->step_max[group] = hmax;
        engine_drop_graphs(e);
    }
    return 0;
}

extern "C" int his_engine_set_branch_order(his_engine* e, int32_t order) {
    if (!e || order < -1 || order > 1) { set_error("his_engine_set_branch_order: bad argument"); return HIS_E_ARG; }
    const int v = order < 0 ? e->one_first_est : order;
    if (v != e->one_first) {
        e->one_first = v;
        engine_drop_graphs(e);
    }
    return 0;
}

extern "C" int his_engine_destroy(his_engine* e) {
    if (!e) return 0;
    engine_drop_graphs(e);
    (void)hipStreamDestroy(e->cap);
    (void)hipStreamDestroy(e->side);
    (void)hipEventDestroy(e->ev_fork);
    (void)hipEventDestroy(e->ev_join);
    delete e;
    return 0;
}

extern "C" int his_set_tile_shape(int32_t dtype, int32_t shape) {
    if ((dtype != HIS_F64 && dtype != HIS_F32) || shape < -1 || shape >= kNumShapes) {
        set_error("his_set_tile_shape: bad dtype or shape");
        return HIS_E_ARG;
    }
    const int d = dtype == HIS_F64 ? 0 : 1;
    g_shape[d] = shape < 0 ? kDefaultShape[d] : shape;
    return 0;
}

extern "C" int his_tile_shape(int32_t dtype, int32_t shape, int32_t* tx, int32_t* ty, int32_t* variant) {
    if ((dtype != HIS_F64 && dtype != HIS_F32) || shape < 0 || shape >= kNumShapes || !tx || !ty) return HIS_E_ARG;
    const TileShape& t = dtype == HIS_F64 ? kShapesF64[shape] : kShapesF32[shape];
    *tx = t.tx;
    *ty = t.ty;
    if (variant) *variant = t.march;
    return 0;
}

extern "C" const char* his_last_error(void) { return g_err; }
extern "C" const char* his_version(void) { return "his_amd 0.1 gfx950"; }
