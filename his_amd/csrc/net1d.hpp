// Internal descriptors of the 1-D network and weir-exchange kernels.
#pragma once
#include "common.hpp"

namespace his {

// Where a 1-D launch takes its per-member dt, substep count and forcing from.
struct Net1DDrive {
    // per-op mode: one substep of dt[m] with pm (2, Btot) = q_in, eta; vols (3, Btot) written
    const double *dt, *pm;
    double *vols;
    // engine mode (engine != 0): n_sub[m] substeps of dt[m]/n_sub[m] from t[m], forcing from the
    // series, ledger acc (10, Btot) += bnd1d_in (4), bnd1d_out (5), clip1d (7)
    int engine;
    const int32_t *n_sub;
    const double *t;
    his_series inflow, stage;
    double *acc;
    const int32_t *active;
    // epilogue: CFL rate of the final state (may be NULL)
    double *rate;
    int32_t *wet;
    int Btot, member0;
    // engine mode: fold the exchange's signed link volumes (B, ex_stride) into acc slots 8 / 9 first,
    // and advance t_out[m] = t[m] + dt[m], steps[m] += 1 at the end
    const double *ex_vf;
    int ex_nl, ex_stride;
    double *t_out;
    int32_t *steps;
    // engine mode, or NULL: a member whose final channel state fails the plausibility guard gets
    // error code 5 (unless it carries one already)
    int32_t *error;
};

// doubles per part of a split network's ledger partials: b_in, b_out, clip, rate, flags (wet +
// 64 x implausible, summed over the parts: HIS_MAX_NET_PARTS < 64 keeps the two apart)
constexpr int LED_PART = 5;
static_assert(HIS_MAX_NET_PARTS < 64, "the parts' flag sum needs fewer than 64 parts");

// one workgroup per (member, part); width = the most cells or faces of any part (swe1d_part_width)
int launch_swe1d(const his_net1d& n, double* A, double* Q, const Net1DDrive& d, double* ws, int width,
                 hipStream_t s);
int swe1d_part_width(const his_net1d& n);
// a split network's workspace holds one counter per member that must be zero before its first step
int clear_swe1d_counters(const his_net1d& n, double* ws, hipStream_t s);
int launch_swe1d_rate(const his_net1d& n, const double* A, const double* Q, double* rate, int32_t* wet,
                      int member0, hipStream_t s, const int32_t* only);   // only: NULL or (Btot) member mask

struct ExchangeArgs {
    his_links L;
    his_net1d N;
    int dtype2d;
    int64_t plane;
    void *h, *hu, *hv;
    double* A;
    const double* dt;        // (Btot)
    double* V;               // (B, nl) or NULL
    double* acc;             // ledger (10, Btot): += exchanged in (8), out (9); or NULL
    int acc_slot_in;         // 8 for the ledger, 0 for a plain (2, B) output
    int acc_stride;          // Btot of that array
    const int32_t* active;
    double* ws;
    int member0;
};
int launch_exchange(const ExchangeArgs& a, hipStream_t s);

struct DtArgs;
struct FinalizeArgs;

// Several groups' exchanges in one launch per phase (every group shares the 2-D dtype).
struct ExGroups {
    int ngroups;
    int64_t it_links[HIS_MAX_GROUPS + 1], it_cells[HIS_MAX_GROUPS + 1];   // filled by the launcher
    ExchangeArgs a[HIS_MAX_GROUPS];
    // optional step head fused into the first phase (device copy of the engine's dt arguments and the
    // dt group of every member); each block then derives its members' dt itself and the block holding
    // a member's first link writes that member's control entries.  Needs every group's nl + nu >= 256.
    const DtArgs* head;
    const int32_t* dtg_of;
    double* rate_reset;      // (Btot) or NULL: cleared in the last phase for every member that steps
    // optional: the previous step's ledger reduction (device copy of its arguments), one extra
    // 256-thread block per member after the link blocks of the first phase
    const FinalizeArgs* fin;
    int fin_members;
    // every group's 1-D cells have at most one link each: the first launch does the whole exchange
    // (one item per unique 2-D cell) and no second phase runs
    int one_phase;
};
int launch_exchange_groups(ExGroups& g, hipStream_t s);

// Several groups' 1-D networks in one launch: workgroups blk0[g] .. blk0[g+1]-1 run group g, one per
// (member, part); width[g] = the most cells or faces of any of group g's parts.
struct Swe1DGroups {
    int ngroups;
    int blk0[HIS_MAX_GROUPS + 1];
    int width[HIS_MAX_GROUPS];
    his_net1d N[HIS_MAX_GROUPS];
    double *A[HIS_MAX_GROUPS], *Q[HIS_MAX_GROUPS], *ws[HIS_MAX_GROUPS];
    Net1DDrive d[HIS_MAX_GROUPS];
};
int launch_swe1d_groups(const Swe1DGroups& g, hipStream_t s);

size_t swe1d_ws_doubles(const his_net1d& n);
// doubles of exchange workspace per member: V, sign, final V (nl each), donor depth of each unique
// 2-D cell (nu), available donor volume of each 1-D cell (nc)
__host__ __device__ inline int64_t exchange_ws_stride(const his_links& l, const his_net1d& n) {
    return 3 * (int64_t)l.nl + (int64_t)l.nu + (int64_t)n.nc;
}
// signed link volumes of the last exchange for member b (float64, nl values) inside its workspace
inline const double* exchange_vf(double* ws, const his_links& l, const his_net1d& n, int b) {
    return ws + (size_t)b * (size_t)exchange_ws_stride(l, n) + 2 * (size_t)l.nl;
}
size_t exchange_ws_doubles(const his_links& l, const his_net1d& n);

}  // namespace his
