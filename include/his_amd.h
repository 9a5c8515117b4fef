/*
 * his_amd -- C ABI of the MI355X (gfx950) coupled flood-engine time step.
 *
 * Every entry point takes plain device pointers and sizes, enqueues work on the
 * given HIP stream and returns 0 on success or a negative HIS_E_* code (the text
 * of the last error is available from his_last_error()).  No entry point
 * allocates device memory: scratch comes in through a workspace pointer whose
 * size the matching *_workspace_bytes() call returns, except for
 * his_swe2d_step_noalloc, whose scratch is a fixed device array of the library.
 *
 * Layouts: 2-D fields are row-major (B, ny, nx) -- rows south to north, columns
 * west to east -- or (ny, nx) when shared by every member (member stride 0).
 * Per-member control values (dt, t, boundary values) are float64 device arrays.
 */
#ifndef HIS_AMD_H
#define HIS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HIS_F64 0
#define HIS_F32 1

#define HIS_BC_CLOSED 0
#define HIS_BC_TRANSMISSIVE 1
#define HIS_BC_STAGE 2
#define HIS_BC_DISCHARGE 3

/* Plausibility limits the engine checks on the device after every coupled step, as hydrointel's
 * check_plausible does at its check times (hydrointel/solver/coupling.py:169-192): the deepest 2-D
 * cell, and the highest channel level above its bank crest (interior 1-D cells), in metres. */
#define HIS_PLAUSIBLE_DEPTH_2D_M 10.0
#define HIS_PLAUSIBLE_RISE_1D_M 10.0

/* Most parts a network may be split into (his_net1d.nparts). */
#define HIS_MAX_NET_PARTS 63

#define HIS_E_ARG (-1)
#define HIS_E_HIP (-2)
#define HIS_E_WORKSPACE (-3)
#define HIS_E_STATE (-4)

#define HIS_MAX_GROUPS 4

/* ------------------------------------------------------------------ 2-D grid */

/* Closure of one SWE2D (hydrointel/solver/swe2d.py:114-145): bed, roughness,
 * retention capacity, Horton fields, rain weight, edge kinds and constants.
 * Fields carry their member stride: 0 = shared (ny, nx), ny*nx = per member. */
typedef struct his_grid2d {
    int32_t dtype;                 /* HIS_F64 or HIS_F32: z, fields and state */
    int32_t B, ny, nx;             /* members (1 when unbatched), rows, columns */
    int32_t order, flux;           /* spatial order 1|2; flux 0 = HLLC, 1 = HLL */
    int32_t bc[4];                 /* west, east, south, north: HIS_BC_* */
    double dx, dy, g, h_dry, kp_eps;
    double horton_t0;              /* Horton decay clock start (s) */
    const void *z;                 /* (ny, nx) */
    const void *manning; int64_t manning_bs;
    const void *storage; int64_t storage_bs;      /* NULL = no retention store */
    const void *f0; const void *fc; const void *k; int64_t horton_bs;   /* f0 NULL = no infiltration */
    const void *rain_weight; int64_t rain_weight_bs;                   /* NULL = uniform rain */
} his_grid2d;

/* The five state fields, (B, ny, nx) each, in the grid's dtype; updated in place. */
typedef struct his_state2d {
    void *h, *hu, *hv, *s_filled, *infil_cum;
} his_state2d;

size_t his_swe2d_workspace_bytes(const his_grid2d *g);

/* One SSP-RK2 (order 2) or Euler (order 1) step with sources and friction.
 * Replaces SWE2D._step_core (hydrointel/solver/swe2d.py:305-371).
 * dt_t: device (2, B) float64 = dt per member, then t per member.
 * pm:   device (9, B) float64 = rain rate, W,E,S,N boundary values at t, the same at t+dt
 *       (the reference's pm, solver/swe2d.py:282-300).
 * vols: device (5, B) float64 out = bnd_in, bnd_out, rain, infil, clip (StepVolumes, swe2d.py:104-110).
 * rate: device (B) float64 out, may be NULL = each member's CFL rate of the new state
 *       (SWE2D._max_rate, swe2d.py:179-184, evaluated per member). */
int his_swe2d_step(const his_grid2d *g, his_state2d st, const double *dt_t, const double *pm, double *vols,
                   double *rate, void *ws, size_t ws_bytes, void *stream);

/* his_swe2d_step with no workspace argument, allocating nothing (no hipMalloc, hipFree or allocating
 * library call).  No counterpart in hydrointel: it is for callers that cannot allocate device memory
 * -- code under HIP graph capture, tight serving loops, harnesses that forbid it.  Its scratch is one
 * fixed device array of HIS_SWE2D_NOALLOC_MAX_BYTES that every GPU the library runs on holds from the
 * first use until the process ends; a grid whose his_swe2d_workspace_bytes() exceeds it is refused
 * with HIS_E_WORKSPACE (400x325 grids fit up to 41 members in float32, 19 in float64).  Every call on
 * one GPU shares that GPU's array: calls that overlap on a GPU (two streams, two host threads) race,
 * so the caller serialises them.  The stream must belong to the current device.  Same arguments and
 * results as his_swe2d_step otherwise, and the same kernels. */
#define HIS_SWE2D_NOALLOC_MAX_BYTES (64u * 1024u * 1024u)
int his_swe2d_step_noalloc(const his_grid2d *g, his_state2d st, const double *dt_t, const double *pm,
                           double *vols, double *rate, void *stream);

/* Per-member maximum of (|u|+c)/dx + (|v|+c)/dy over wet cells.
 * Replaces SWE2D._max_rate (hydrointel/solver/swe2d.py:179-184); rate is device (B) float64. */
int his_swe2d_max_rate(const his_grid2d *g, const void *h, const void *hu, const void *hv, double *rate,
                       void *ws, size_t ws_bytes, void *stream);

/* Kurganov-Petrova velocities u, v (SWE2D.velocities, hydrointel/solver/swe2d.py:161-167). */
int his_swe2d_velocities(const his_grid2d *g, const void *h, const void *hu, const void *hv, void *u, void *v,
                         void *stream);

/* Float64 per-member sums of h and s_filled times the cell area
 * (SWE2D.volume / storage_volume, hydrointel/solver/swe2d.py:169-174); out is device (2, B). */
int his_swe2d_volumes(const his_grid2d *g, const void *h, const void *s_filled, double *out, void *ws,
                      size_t ws_bytes, void *stream);

/* ------------------------------------------------------------------ 1-D network */

/* Flattened network of SWE1D (hydrointel/solver/swe1d.py:176-231); all float64.
 * Cell arrays z, b, yb, ts carry member stride bs (0 or nc); face sections fb, fyb, fts stride fbs. */
typedef struct his_net1d {
    int32_t B, nc, nf, ngh;
    double g, h_dry, cfl_1d, dt_max;
    const double *z, *b, *yb, *ts; int64_t bs;
    const double *m, *nman, *dx;           /* (nc) shared */
    const double *fb, *fyb, *fts; int64_t fbs;
    const double *fm;                      /* (nf) shared */
    const int32_t *fl, *fr;                /* (nf) left / right cell of each face */
    const int32_t *kind;                   /* (nc) 0 reach, 1 junction, 2 inflow, 3 stage, 4 closed ghost */
    const int32_t *g_idx, *g_src;          /* (ngh) ghost cell, its interior neighbour */
    const double *g_sign;                  /* (ngh) +1 downstream end, -1 upstream end */
    /* faces grouped by cell in increasing face order (index_add order): faces with fl == c are
     * lf_face[lf_ptr[c] .. lf_ptr[c+1]), faces with fr == c are rf_face[rf_ptr[c] .. rf_ptr[c+1]) */
    const int32_t *lf_ptr, *lf_face, *rf_ptr, *rf_face;
    const double *zbank;                   /* (nc) shared bank crest, for the plausibility guard */
    /* Optional split of the network into nparts parts that no face joins (unions of whole connected
     * components), which separate workgroups then step without synchronising: part p owns the cells
     * p_cell[p_cptr[p] .. p_cptr[p+1]), the faces p_face[p_fptr[p] .. p_fptr[p+1]) and the ghosts
     * p_ghost[p_gptr[p] .. p_gptr[p+1]) (indices into g_idx), each list ascending.  nparts <= 1 =
     * one part, the whole network (the arrays may then be NULL); at most HIS_MAX_NET_PARTS. */
    int32_t nparts;
    int32_t part_width;                    /* the most cells or faces of any part: sizes the workgroup */
    const int32_t *p_cptr, *p_cell, *p_fptr, *p_face, *p_gptr, *p_ghost;
} his_net1d;

size_t his_swe1d_workspace_bytes(const his_net1d *n);

/* One 1-D step of every member.  Replaces SWE1D._step_core (hydrointel/solver/swe1d.py:320-368).
 * A, Q: device (B, nc) updated in place.  dt: device (B).  pm: device (2, B) = q_in, eta_sea.
 * vols: device (3, B) out = b_in, b_out, clip (Step1DVolumes, swe1d.py:169-173).
 * One workgroup per member and part; the cell and face arithmetic does not depend on the split. */
int his_swe1d_step(const his_net1d *n, double *A, double *Q, const double *dt, const double *pm, double *vols,
                   void *ws, size_t ws_bytes, void *stream);

/* Per-member maximum of (|u|+c)/dx over wet interior cells, and whether any is wet.
 * Replaces SWE1D.max_dt's reduction (hydrointel/solver/swe1d.py:253-262). */
int his_swe1d_max_rate(const his_net1d *n, const double *A, const double *Q, double *rate, int32_t *any_wet,
                       void *stream);

/* Per-member maximum depth and 1-D rise above the bank crest, for the plausibility guard
 * (CoupledEngine.check_plausible, hydrointel/solver/coupling.py:169-192); out is device (2, B). */
int his_plausibility(const his_grid2d *g, const void *h, const his_net1d *net, const double *A,
                     double *out, void *stream);

/* ------------------------------------------------------------------ weir exchange */

/* Links of build_links (hydrointel/solver/coupling.py:47-70) with CSR groupings in link order. */
typedef struct his_links {
    int32_t nl, nu;                 /* links, unique linked 2-D cells */
    const int32_t *cell1d, *inv;    /* (nl) 1-D cell, unique slot */
    const int64_t *uniq;            /* (nu) flat 2-D cell index (j * nx + i) */
    const double *zbank, *length, *z2;   /* (nl) */
    const int32_t *u_ptr, *u_link;  /* links of unique slot k: u_link[u_ptr[k] .. u_ptr[k+1]) */
    const int32_t *c_ptr, *c_link;  /* links of 1-D cell c:    c_link[c_ptr[c] .. c_ptr[c+1]) */
    double area2, weir_cw, limit, relax;
} his_links;

size_t his_exchange_workspace_bytes(const his_links *l, const his_net1d *n);

/* Move water between the models in place.  Replaces exchange (hydrointel/solver/coupling.py:73-124).
 * h, hu, hv: device (B, plane) in dtype2d; A: device (B, nc); dt: device (B).
 * V: device (B, nl) float64 out (signed, + = 2-D to 1-D), may be NULL.
 * ex: device (2, B) float64, may be NULL: += gross 2-D->1-D and 1-D->2-D volumes (coupling.py:210-211). */
int his_exchange(const his_links *l, const his_net1d *n, int32_t dtype2d, int64_t plane, void *h, void *hu,
                 void *hv, double *A, const double *dt, double *V, double *ex, void *ws, size_t ws_bytes,
                 void *stream);

/* ------------------------------------------------------------------ forcing series */

/* One per-member time series evaluated on the device (float64):
 *   kind 0 constant  v = vals[b]
 *   kind 1 piecewise linear on t0 + k*step, end values held  (hydrointel/engine.py:158-166)
 *   kind 2 piecewise constant blocks, zero outside             (hydrointel/engine.py:145-149)
 *   kind 3 harmonic base[b] + on[b] * sum_c amp[c] cos(w[c] (t + shift[b]) - phase[c])  (engine.py:151-156) */
typedef struct his_series {
    int32_t kind, K;
    double t0, step;
    const double *vals;            /* kind 0: (B); kinds 1, 2: (B, K) */
    const double *base, *shift;    /* kind 3: (B) */
    const int32_t *on;             /* kind 3: (B) */
    const double *amp, *w, *phase; /* kind 3: (K) */
} his_series;

/* ------------------------------------------------------------------ coupled engine */

/* One grid group: members that share a bed, a grid and a network topology. */
typedef struct his_group {
    his_grid2d grid;
    his_state2d state;
    const his_net1d *net;          /* NULL = 2-D only */
    const his_links *links;        /* NULL = 2-D only */
    double *A, *Q;                 /* (B, nc) */
    his_series rain, edge[4], inflow, stage;   /* edge kinds other than stage/discharge ignore edge[e] */
    int32_t dt_group_size;         /* members per shared-dt group: B = one dt (reference), 1 = per member */
} his_group;

/* Control block, one float64/int32 slot per member across all groups (member index = group
 * offset + b).  t, t_limit are read and written by the engine; the host sets them between runs. */
typedef struct his_control {
    double *t;          /* (Btot) simulated time */
    double *t_limit;    /* (Btot) next output time or t_end: a step never crosses it */
    double *t_end;      /* (Btot) members at or past t_end - 1e-9 stop */
    double *dt;         /* (Btot) last step's dt */
    int32_t *n_sub;     /* (Btot) last step's 1-D substeps */
    int32_t *steps;     /* (Btot) coupled steps taken */
    int32_t *clamped;   /* (Btot) dt clamped up to dt_min (SWE2D.n_dt_clamped) */
    double *acc;        /* (10, Btot) float64 ledger: rain, infiltration, bnd2d_in, bnd2d_out,
                           bnd1d_in, bnd1d_out, clip2d, clip1d, exchanged_in, exchanged_out */
    double *dt_min_seen, *dt_max_seen;   /* (Btot) */
    int32_t *n_sub_max; /* (Btot) */
    int32_t *error;     /* (Btot) 0 ok; 1 = 1-D substeps above n_sub_limit, 2 = non-finite CFL rate,
                           3 = non-finite 2-D depth or storage (found by the step's float64 sums; a
                           non-finite momentum reaches the depth through the next step's fluxes),
                           4 = a 2-D depth above HIS_PLAUSIBLE_DEPTH_2D_M after the step,
                           5 = a channel level more than HIS_PLAUSIBLE_RISE_1D_M above its bank (or not
                           finite) after the step: the member stops (like hydrointel's
                           NumericalInstabilityError and check_plausible, coupling.py:169-192) */
    int32_t n_sub_limit;
} his_control;

typedef struct his_engine_desc {
    int32_t ngroups;
    his_group group[HIS_MAX_GROUPS];
    his_control ctl;
    double cfl, dt_min, dt_max;
    void *ws;                      /* workspace of his_engine_workspace_bytes() */
    size_t ws_bytes;
    int32_t use_graph;             /* 1 = capture one coupled step into a hipGraph and replay it */
} his_engine_desc;

typedef struct his_engine his_engine;

size_t his_engine_workspace_bytes(const his_engine_desc *d);

/* Build an engine over caller-owned buffers.  Computes the initial CFL rates.
 * The coupled step is the order of CoupledEngine.step (hydrointel/solver/coupling.py:201-227):
 * dt from the pre-exchange state, n_sub, exchange, 2-D step, n_sub 1-D substeps. */
int his_engine_create(const his_engine_desc *d, his_engine **out, void *stream);

/* Advance every active member by n_steps coupled steps (members at t_limit or t_end idle).
 * Replaces the body of CoupledEngine.run's loop (hydrointel/solver/coupling.py:260-262). */
int his_engine_step(his_engine *e, int32_t n_steps, void *stream);

/* Recompute the CFL rates after the host has changed any state (set_state, exchange by hand). */
int his_engine_refresh(his_engine *e, void *stream);

/* The same for the members with mask[m] != 0 only (mask: device (Btot) int32); the others keep the
 * rates their last step left, so loading a new storm into one slot leaves every running storm's
 * step sequence exactly as it would be alone.  No counterpart in hydrointel, which runs one batch
 * at a time (hydrointel/api.py:274-324); this serves continuous batching. */
int his_engine_refresh_members(his_engine *e, const int32_t *mask, void *stream);

/* Fold each member's current depth into a running maximum (device, grid dtype), for the
 * depth_max of FloodResult taken over output times (hydrointel/api.py:239). */
int his_engine_fold_max(his_engine *e, int32_t group, void *hmax, void *stream);

/* From the next step on, fold group `group`'s depth into hmax (device, (B, ny, nx), grid dtype)
 * after the 2-D stages of every coupled step; hmax = NULL stops it.  The captured step graphs
 * are captured again at the next step.  No counterpart in hydrointel, whose depth_max is taken
 * over output times only (hydrointel/api.py:239): for callers that need the peak between outputs.
 * Cost: one pass over the group's depth per step. */
int his_engine_track_step_max(his_engine *e, int32_t group, void *hmax);

/* Which branch of a coupled step is captured first: 0 the 2-D stages, 1 the 1-D network,
 * -1 the engine's own estimate of the longer one (the default).  In a captured graph the first
 * branch keeps the exchange's hardware queue and the other pays a cross-queue start, so the longer
 * branch should go first; the estimate can miss, and a tuner can time both.  The captured step
 * graphs are captured again at the next step.  No counterpart in hydrointel (no graphs there). */
int his_engine_set_branch_order(his_engine *e, int32_t order);

int his_engine_destroy(his_engine *e);

/* ------------------------------------------------------------------ misc */

/* Tile shape of the 2-D kernels, per dtype: index into the compiled shapes; -1 restores the
 * default (float64: shape 4, 32 x 8 LDS tiles; float32: shape 2, register march in 8-row chunks).
 * The autotuner (his_amd/tuning.py) times each and stores the fastest per grid.
 * his_tile_shape reports a shape: tile or strip width tx, rows ty, and the kernel variant
 * (0 LDS-tiled, 1 LDS-ring march, 2 register march); variant may be NULL. */
int his_set_tile_shape(int32_t dtype, int32_t shape);
int his_tile_shape(int32_t dtype, int32_t shape, int32_t *tx, int32_t *ty, int32_t *variant);

const char *his_last_error(void);
const char *his_version(void);

#ifdef __cplusplus
}
#endif
#endif
