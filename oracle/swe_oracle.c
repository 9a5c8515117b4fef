/*
 * CPU oracle for the coupled flood-engine time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is a plain-C, float64 restatement of the reference solver, written
 * array-by-array in the same order as the reference's torch-eager operations so
 * that each intermediate is rounded the way the reference rounds it:
 *
 *   orc_swe2d_step       hydrointel/solver/swe2d.py:305-371 (_step_core), with
 *                        rhs 252-279, _sweep 196-234, riemann_flux 61-101,
 *                        minmod 57-58, _donor_scale 236-241, _assemble 243-250,
 *                        velocities 161-167; ghosts from solver/boundaries.py:56-128;
 *                        Horton rate from solver/infiltration.py:34-37.
 *   orc_swe2d_max_rate   hydrointel/solver/swe2d.py:179-184 (per member)
 *   orc_swe1d_step       hydrointel/solver/swe1d.py:320-368 (+ _fill_ghosts 265-294,
 *                        Sections 50-84, hll_1d 87-107)
 *   orc_swe1d_max_rate   hydrointel/solver/swe1d.py:253-262 (per member)
 *   orc_exchange         hydrointel/solver/coupling.py:73-124
 *
 * Only tests, the smoke check and the CPU-baseline leg of the benchmark use it,
 * always as the checker. It is never part of the shipped path.
 *
 * Transcendentals come from the C library (sqrt is correctly rounded in both;
 * pow/exp/cos may differ from torch's vectorised versions by an ulp), sums over
 * the grid run in row-major order (torch sums pairwise), and index_add runs in
 * index order exactly as torch's CPU index_add does.
 *
 * Optional OpenMP over members/rows/columns keeps every element's arithmetic
 * identical to the serial order (reductions are done serially afterwards).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#define SQRT2 1.4142135623730951
#define TINY 1e-14

enum { BC_CLOSED = 0, BC_TRANSMISSIVE = 1, BC_STAGE = 2, BC_DISCHARGE = 3 };

typedef struct {
    int32_t B, ny, nx;
    int32_t order, flux;          /* flux: 0 hllc, 1 hll */
    int32_t bc[4];                /* west, east, south, north */
    double dx, dy, g, h_dry, kp_eps;
    const double *z;              /* (ny, nx) shared by the batch */
    const double *n;  int64_t n_bs;   /* member stride: 0 (shared) or ny*nx */
    const double *S;  int64_t S_bs;   /* storage capacity, may be NULL */
    const double *f0, *fc, *k; int64_t hz_bs; double t0;   /* Horton, f0 NULL = off */
    const double *rw; int64_t rw_bs;  /* rain weight, may be NULL */
} orc_grid;

static inline double dmin(double a, double b) { return a < b ? a : b; }   /* torch.minimum (no NaN here) */
static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double relu(double a) { return a > 0.0 ? a : 0.0; }
static inline double clamp0(double a) { return a < 0.0 ? 0.0 : a; }       /* torch.clamp(min=0) */
static inline double sgnf(double a) { return a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : 0.0); }

static inline double minmod(double a, double b) {
    return (a * b > 0.0) ? sgnf(a) * dmin(fabs(a), fabs(b)) : 0.0;
}

/* swe2d.py:161-167 */
static inline void kp_vel(double h, double hu, double hv, double kp_eps, double h_dry, double *u, double *v) {
    double h4 = pow(h, 4.0);
    double den = sqrt(h4 + dmax(h4, kp_eps));
    if (h >= h_dry) {
        *u = SQRT2 * h * hu / den;
        *v = SQRT2 * h * hv / den;
    } else {
        *u = 0.0;
        *v = 0.0;
    }
}

/* swe2d.py:61-101 for one face */
static void riemann(double hL, double uL, double vL, double hR, double uR, double vR, double g, int hll,
                    double *F1o, double *F2o, double *F3o) {
    double cL = sqrt(g * hL), cR = sqrt(g * hR);
    int dryL = hL <= TINY, dryR = hR <= TINY;
    double u_s = 0.5 * (uL + uR) + cL - cR;
    double c_s = 0.5 * (cL + cR) + 0.25 * (uL - uR);
    double SL = dmin(uL - cL, u_s - c_s);
    double SR = dmax(uR + cR, u_s + c_s);
    if (dryL) { SL = uR - 2.0 * cR; SR = uR + cR; }
    if (dryR) { SL = uL - cL; SR = uL + 2.0 * cL; }
    int both = dryL && dryR;
    double qL = hL * uL, qR = hR * uR;
    double FL1 = qL, FR1 = qR;
    double hg = 0.5 * g;
    double FL2 = qL * uL + hg * hL * hL, FR2 = qR * uR + hg * hR * hR;
    double den = SR - SL;
    if (fabs(den) < TINY) den = 1.0;
    double H1 = (SR * FL1 - SL * FR1 + SL * SR * (hR - hL)) / den;
    double H2 = (SR * FL2 - SL * FR2 + SL * SR * (qR - qL)) / den;
    int left = SL >= 0.0, right = SR <= 0.0;
    double F1 = left ? FL1 : (right ? FR1 : H1);
    double F2 = left ? FL2 : (right ? FR2 : H2);
    double F3;
    if (!hll) {
        double num = SL * hR * (uR - SR) - SR * hL * (uL - SL);
        double dd = hR * (uR - SR) - hL * (uL - SL);
        double SM = (fabs(dd) < TINY) ? 0.5 * (uL + uR) : num / dd;
        F3 = F1 * (SM >= 0.0 ? vL : vR);
    } else {
        double FL3 = qL * vL, FR3 = qR * vR;
        double H3 = (SR * FL3 - SL * FR3 + SL * SR * (hR * vR - hL * vL)) / den;
        F3 = left ? FL3 : (right ? FR3 : H3);
    }
    if (both) { F1 = 0.0; F2 = 0.0; F3 = 0.0; }
    *F1o = F1; *F2o = F2; *F3o = F3;
}

/* boundaries.py:56-78: ghost state from one interior layer; un along the outward normal */
static void edge_ghost(int kind, double h, double un, double ut, double z, double val, double g, double h_dry,
                       double *hg_o, double *un_o, double *ut_o, double *z_o) {
    if (kind == BC_CLOSED) { *hg_o = h; *un_o = -un; *ut_o = ut; *z_o = z; return; }
    if (kind == BC_TRANSMISSIVE) { *hg_o = h; *un_o = un; *ut_o = ut; *z_o = z; return; }
    if (kind == BC_STAGE) {
        double hg = clamp0(val - z);
        double ci = sqrt(g * h), cg = sqrt(g * hg);
        int wet_i = h >= h_dry;
        double un_g = wet_i ? un + 2.0 * (ci - cg) : 0.0;
        int sup = wet_i && (un > ci);
        if (sup) { un_g = un; hg = h; }
        if (!(hg >= h_dry)) un_g = 0.0;
        double ut_g = un_g > 0.0 ? ut : 0.0;
        *hg_o = hg; *un_o = un_g; *ut_o = ut_g; *z_o = z;
        return;
    }
    /* discharge */
    double hc = pow(val * val / g, 1.0 / 3.0);
    if (hc < h_dry) hc = h_dry;
    double hg = (h >= h_dry) ? h : hc;
    *hg_o = hg; *un_o = -val / hg; *ut_o = 0.0; *z_o = z;
}

/* boundaries.py:81-128 for one line of n cells: fill padded arrays of length n + 2*ng.
 * q_un is the component normal to the sweep direction, q_ut the tangential one. */
static void pad_line(int n, int ng, const double *h, const double *q_un, const double *q_ut, const double *z,
                     int kind_lo, int kind_hi, double val_lo, double val_hi, double g, double h_dry,
                     double *hp, double *up, double *vp, double *zp) {
    for (int i = 0; i < n; ++i) { hp[ng + i] = h[i]; up[ng + i] = q_un[i]; vp[ng + i] = q_ut[i]; zp[ng + i] = z[i]; }
    for (int side = 0; side < 2; ++side) {
        int kind = side == 0 ? kind_lo : kind_hi;
        double val = side == 0 ? val_lo : val_hi;
        double sgn = side == 0 ? -1.0 : 1.0;
        int mirror = (kind == BC_CLOSED || kind == BC_TRANSMISSIVE);
        for (int k = 0; k < ng; ++k) {
            int src;
            if (mirror && n >= ng) src = side == 0 ? k : n - 1 - k;
            else src = side == 0 ? 0 : n - 1;
            double hg, ung, utg, zg;
            edge_ghost(kind, h[src], sgn * q_un[src], q_ut[src], z[src], val, g, h_dry, &hg, &ung, &utg, &zg);
            int p = side == 0 ? ng - 1 - k : ng + n + k;
            hp[p] = hg; up[p] = sgn * ung; vp[p] = utg; zp[p] = zg;
        }
    }
}

/* swe2d.py:196-234 on one padded line.  Outputs on n+1 faces: F1, F2, F3, corrL, corrR; Sc on n cells */
static void sweep_line(int n, int order, double g, double h_dry, int hll,
                       const double *hp, const double *up, const double *vp, const double *zp,
                       double *F1, double *F2, double *F3, double *cL, double *cR, double *Sc,
                       double *wk /* 8*(n+2) + 2*(n+4) scratch */) {
    int ng = order == 1 ? 1 : 2;
    int m = n + 2;                       /* reconstructed cells: padded ng-1 .. ng+n */
    double *Hm = wk, *Hp = wk + m, *Um = wk + 2 * m, *Up = wk + 3 * m;
    double *Vm = wk + 4 * m, *Vp = wk + 5 * m, *Zm = wk + 6 * m, *Zp = wk + 7 * m;
    if (order == 2) {
        for (int k = 0; k < m; ++k) {
            int p = ng - 1 + k;          /* padded index of this reconstructed cell (centre) */
            double H = hp[p], U = up[p], V = vp[p], Z = zp[p];
            double E = H + Z;
            double hmin = dmin(dmin(hp[p - 1], hp[p]), hp[p + 1]);
            double ok = (hmin >= h_dry) ? 1.0 : 0.0;
            double Em1 = hp[p - 1] + zp[p - 1], E0 = hp[p] + zp[p], Ep1 = hp[p + 1] + zp[p + 1];
            double sH = minmod(hp[p] - hp[p - 1], hp[p + 1] - hp[p]) * ok;
            double sE = minmod(E0 - Em1, Ep1 - E0) * ok;
            double sU = minmod(up[p] - up[p - 1], up[p + 1] - up[p]) * ok;
            double sV = minmod(vp[p] - vp[p - 1], vp[p + 1] - vp[p]) * ok;
            Hm[k] = H - 0.5 * sH; Hp[k] = H + 0.5 * sH;
            double Emk = E - 0.5 * sE, Epk = E + 0.5 * sE;
            Um[k] = U - 0.5 * sU; Up[k] = U + 0.5 * sU;
            Vm[k] = V - 0.5 * sV; Vp[k] = V + 0.5 * sV;
            Zm[k] = Emk - Hm[k]; Zp[k] = Epk - Hp[k];
        }
        for (int i = 0; i < n; ++i)
            Sc[i] = -0.5 * g * (Hm[i + 1] + Hp[i + 1]) * (Zp[i + 1] - Zm[i + 1]);
    } else {
        for (int k = 0; k < m; ++k) {
            Hm[k] = Hp[k] = hp[k]; Um[k] = Up[k] = up[k]; Vm[k] = Vp[k] = vp[k]; Zm[k] = Zp[k] = zp[k];
        }
    }
    double hg = 0.5 * g;
    for (int f = 0; f <= n; ++f) {
        double hL = Hp[f], uL = Up[f], vL = Vp[f], zL = Zp[f];
        double hR = Hm[f + 1], uR = Um[f + 1], vR = Vm[f + 1], zR = Zm[f + 1];
        double zf = dmax(zL, zR);
        double hLs = clamp0(hL + zL - zf);
        double hRs = clamp0(hR + zR - zf);
        riemann(hLs, uL, vL, hRs, uR, vR, g, hll, &F1[f], &F2[f], &F3[f]);
        cL[f] = hg * (hL * hL - hLs * hLs);
        cR[f] = hg * (hR * hR - hRs * hRs);
    }
}

typedef struct {
    double *u, *v;                         /* ny*nx */
    double *F1x, *F2x, *F3x, *cLx, *cRx, *Scx;   /* ny*(nx+1), Sc ny*nx */
    double *F1y, *F2y, *F3y, *cLy, *cRy, *Scy;   /* nx*(ny+1), Sc nx*ny (column-major lines) */
    double *theta;
} rhs_work;

/* swe2d.py:252-279 for one member.  vals: boundary values for W,E,S,N at this stage */
static void rhs_member(const orc_grid *G, const double *h, const double *hu, const double *hv, const double *vals,
                       double dt, double *dh, double *dhu, double *dhv, double *qin, double *qout, rhs_work *W) {
    int ny = G->ny, nx = G->nx, order = G->order, ng = order == 1 ? 1 : 2;
    int hll = G->flux == 1;
    double g = G->g, h_dry = G->h_dry;
    const double *z = G->z;
    for (int i = 0; i < ny * nx; ++i) kp_vel(h[i], hu[i], hv[i], G->kp_eps, h_dry, &W->u[i], &W->v[i]);
    int L = (nx > ny ? nx : ny) + 2 * ng + 4;
    double *buf = (double *)malloc(sizeof(double) * (size_t)(8 * L + 8 * L + 16 * L));
    double *lh = buf, *lu = buf + L, *lv = buf + 2 * L, *lz = buf + 3 * L;
    double *hp = buf + 4 * L, *upd = buf + 5 * L, *vpd = buf + 6 * L, *zp = buf + 7 * L;
    double *wk = buf + 8 * L;
    /* x sweeps: rows, west/east ghosts, normal = u */
    for (int j = 0; j < ny; ++j) {
        const double *hr = h + (size_t)j * nx;
        pad_line(nx, ng, hr, W->u + (size_t)j * nx, W->v + (size_t)j * nx, z + (size_t)j * nx,
                 G->bc[0], G->bc[1], vals[0], vals[1], g, h_dry, hp, upd, vpd, zp);
        size_t fo = (size_t)j * (nx + 1);
        sweep_line(nx, order, g, h_dry, hll, hp, upd, vpd, zp, W->F1x + fo, W->F2x + fo, W->F3x + fo,
                   W->cLx + fo, W->cRx + fo, W->Scx + (size_t)j * nx, wk);
    }
    /* y sweeps: columns, south/north ghosts, normal = v */
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            size_t c = (size_t)j * nx + i;
            lh[j] = h[c]; lu[j] = W->v[c]; lv[j] = W->u[c]; lz[j] = z[c];
        }
        pad_line(ny, ng, lh, lu, lv, lz, G->bc[2], G->bc[3], vals[2], vals[3], g, h_dry, hp, upd, vpd, zp);
        size_t fo = (size_t)i * (ny + 1);
        sweep_line(ny, order, g, h_dry, hll, hp, upd, vpd, zp, W->F1y + fo, W->F2y + fo, W->F3y + fo,
                   W->cLy + fo, W->cRy + fo, W->Scy + (size_t)i * ny, wk);
    }
    free(buf);
    /* positivity limiter */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double *Fx = W->F1x + (size_t)j * (nx + 1);
            const double *Gy = W->F1y + (size_t)i * (ny + 1);
            double out = (relu(Fx[i + 1]) + relu(-Fx[i])) / G->dx + (relu(Gy[j + 1]) + relu(-Gy[j])) / G->dy;
            double need = out * dt;
            double hc = h[(size_t)j * nx + i];
            W->theta[(size_t)j * nx + i] = need > hc ? hc / (need > 0.0 ? need : 1.0) : 1.0;
        }
    /* donor scaling + assembly (x) */
    for (int j = 0; j < ny; ++j) {
        double *F1 = W->F1x + (size_t)j * (nx + 1), *F3 = W->F3x + (size_t)j * (nx + 1);
        const double *th = W->theta + (size_t)j * nx;
        for (int f = 0; f <= nx; ++f) {
            double tl = f == 0 ? 1.0 : th[f - 1], tr = f == nx ? 1.0 : th[f];
            double s = F1[f] > 0.0 ? tl : tr;
            F1[f] = F1[f] * s;
            F3[f] = F3[f] * s;
        }
    }
    for (int i = 0; i < nx; ++i) {
        double *F1 = W->F1y + (size_t)i * (ny + 1), *F3 = W->F3y + (size_t)i * (ny + 1);
        for (int f = 0; f <= ny; ++f) {
            double tl = f == 0 ? 1.0 : W->theta[(size_t)(f - 1) * nx + i];
            double tr = f == ny ? 1.0 : W->theta[(size_t)f * nx + i];
            double s = F1[f] > 0.0 ? tl : tr;
            F1[f] = F1[f] * s;
            F3[f] = F3[f] * s;
        }
    }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t fx = (size_t)j * (nx + 1) + i, fy = (size_t)i * (ny + 1) + j;
            double a1 = -(W->F1x[fx + 1] - W->F1x[fx]) / G->dx;
            double a2 = -((W->F2x[fx + 1] + W->cLx[fx + 1]) - (W->F2x[fx] + W->cRx[fx])) / G->dx;
            if (order == 2) a2 = a2 + W->Scx[(size_t)j * nx + i] / G->dx;
            double a3 = -(W->F3x[fx + 1] - W->F3x[fx]) / G->dx;
            double b1 = -(W->F1y[fy + 1] - W->F1y[fy]) / G->dy;
            double b2 = -((W->F2y[fy + 1] + W->cLy[fy + 1]) - (W->F2y[fy] + W->cRy[fy])) / G->dy;
            if (order == 2) b2 = b2 + W->Scy[(size_t)i * ny + j] / G->dy;
            double b3 = -(W->F3y[fy + 1] - W->F3y[fy]) / G->dy;
            size_t c = (size_t)j * nx + i;
            dh[c] = a1 + b1;
            dhu[c] = a2 + b3;
            dhv[c] = a3 + b2;
        }
    /* gross boundary rates, swe2d.py:273-278 */
    double sw_in = 0, se_in = 0, ss_in = 0, sn_in = 0, sw_out = 0, se_out = 0, ss_out = 0, sn_out = 0;
    for (int j = 0; j < ny; ++j) {
        double w = W->F1x[(size_t)j * (nx + 1)], e = W->F1x[(size_t)j * (nx + 1) + nx];
        sw_in += relu(w); se_in += relu(-e); sw_out += relu(-w); se_out += relu(e);
    }
    for (int i = 0; i < nx; ++i) {
        double so = W->F1y[(size_t)i * (ny + 1)], no = W->F1y[(size_t)i * (ny + 1) + ny];
        ss_in += relu(so); sn_in += relu(-no); ss_out += relu(-so); sn_out += relu(no);
    }
    *qin = (sw_in + se_in) * G->dy + (ss_in + sn_in) * G->dx;
    *qout = (sw_out + se_out) * G->dy + (ss_out + sn_out) * G->dx;
}

static rhs_work alloc_work(int ny, int nx) {
    rhs_work W;
    size_t c = (size_t)ny * nx, fx = (size_t)ny * (nx + 1), fy = (size_t)nx * (ny + 1);
    W.u = calloc(c, 8); W.v = calloc(c, 8); W.theta = calloc(c, 8);
    W.F1x = calloc(fx, 8); W.F2x = calloc(fx, 8); W.F3x = calloc(fx, 8); W.cLx = calloc(fx, 8); W.cRx = calloc(fx, 8);
    W.Scx = calloc(c, 8);
    W.F1y = calloc(fy, 8); W.F2y = calloc(fy, 8); W.F3y = calloc(fy, 8); W.cLy = calloc(fy, 8); W.cRy = calloc(fy, 8);
    W.Scy = calloc(c, 8);
    return W;
}

static void free_work(rhs_work *W) {
    free(W->u); free(W->v); free(W->theta); free(W->F1x); free(W->F2x); free(W->F3x); free(W->cLx); free(W->cRx);
    free(W->Scx); free(W->F1y); free(W->F2y); free(W->F3y); free(W->cLy); free(W->cRy); free(W->Scy);
}

/* swe2d.py:179-184, per member maximum */
void orc_swe2d_max_rate(const orc_grid *G, const double *h, const double *hu, const double *hv, double *rate) {
    size_t plane = (size_t)G->ny * G->nx;
    for (int b = 0; b < G->B; ++b) {
        double m = 0.0;
        for (size_t i = 0; i < plane; ++i) {
            size_t c = b * plane + i;
            double u, v;
            kp_vel(h[c], hu[c], hv[c], G->kp_eps, G->h_dry, &u, &v);
            double cc = sqrt(G->g * h[c]);
            double r = (h[c] > G->h_dry) ? (fabs(u) + cc) / G->dx + (fabs(v) + cc) / G->dy : 0.0;
            if (r > m) m = r;
        }
        rate[b] = m;
    }
}

/*
 * swe2d.py:305-371.  Inputs are (B, ny, nx); pm is (9, B): rain, W,E,S,N at t, W,E,S,N at t+dt.
 * vols is (5, B): bnd_in, bnd_out, rain, infil, clip.
 */
void orc_swe2d_step(const orc_grid *G, const double *h0, const double *hu0, const double *hv0, const double *s0,
                    const double *ic0, double dt, double t, const double *pm,
                    double *h, double *hu, double *hv, double *s, double *ic, double *vols) {
    int B = G->B, ny = G->ny, nx = G->nx;
    size_t plane = (size_t)ny * nx;
    double area = G->dx * G->dy;
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < B; ++b) {
        rhs_work W = alloc_work(ny, nx);
        double *d0 = malloc(plane * 8), *d1 = malloc(plane * 8), *d2 = malloc(plane * 8);
        double *h1 = malloc(plane * 8), *hu1 = malloc(plane * 8), *hv1 = malloc(plane * 8);
        const double *H0 = h0 + b * plane, *HU0 = hu0 + b * plane, *HV0 = hv0 + b * plane;
        double vals0[4], vals1[4];
        for (int e = 0; e < 4; ++e) { vals0[e] = pm[(1 + e) * B + b]; vals1[e] = pm[(5 + e) * B + b]; }
        double rain = pm[b];
        double qin, qout, qin2, qout2;
        rhs_member(G, H0, HU0, HV0, vals0, dt, d0, d1, d2, &qin, &qout, &W);
        double clip = 0.0;
        for (size_t i = 0; i < plane; ++i) {
            h1[i] = H0[i] + dt * d0[i]; hu1[i] = HU0[i] + dt * d1[i]; hv1[i] = HV0[i] + dt * d2[i];
        }
        for (size_t i = 0; i < plane; ++i) { double ng_ = clamp0(-h1[i]); clip += ng_; h1[i] = h1[i] + ng_; }
        clip = clip * area;
        if (G->order == 2) {
            rhs_member(G, h1, hu1, hv1, vals1, dt, d0, d1, d2, &qin2, &qout2, &W);
            for (size_t i = 0; i < plane; ++i) {
                h1[i] = 0.5 * (H0[i] + h1[i] + dt * d0[i]);
                hu1[i] = 0.5 * (HU0[i] + hu1[i] + dt * d1[i]);
                hv1[i] = 0.5 * (HV0[i] + hv1[i] + dt * d2[i]);
            }
            qin = 0.5 * (qin + qin2);
            qout = 0.5 * (qout + qout2);
            double c2 = 0.0;
            for (size_t i = 0; i < plane; ++i) { double ng_ = clamp0(-h1[i]); c2 += ng_; h1[i] = h1[i] + ng_; }
            clip = clip + c2 * area;
        }
        double bin = qin * dt, bout = qout * dt;
        /* rainfall */
        double rain_v;
        if (!G->rw) {
            double rd = rain * dt;
            for (size_t i = 0; i < plane; ++i) h1[i] = h1[i] + rd;
            rain_v = rd * ((double)((int64_t)nx * ny) * area);
        } else {
            const double *rw = G->rw + b * G->rw_bs;
            double acc = 0.0;
            for (size_t i = 0; i < plane; ++i) { double add = rain * dt * rw[i]; h1[i] = h1[i] + add; acc += add; }
            rain_v = acc * area;
        }
        double *S = s + b * plane, *IC = ic + b * plane;
        const double *S0 = s0 + b * plane, *IC0 = ic0 + b * plane;
        for (size_t i = 0; i < plane; ++i) { S[i] = S0[i]; IC[i] = IC0[i]; }
        double inf_v = 0.0;
        if (G->f0) {
            const double *f0 = G->f0 + b * G->hz_bs, *fc = G->fc + b * G->hz_bs, *kk = G->k + b * G->hz_bs;
            double tt = clamp0(t - G->t0);
            double acc = 0.0;
            for (size_t i = 0; i < plane; ++i) {
                double rate = fc[i] + (f0[i] - fc[i]) * exp(-kk[i] * tt);
                double pot = rate * dt;
                double take = dmin(pot, h1[i]);
                double safe = h1[i] > 0.0 ? h1[i] : 1.0;
                double scale = h1[i] > 0.0 ? (h1[i] - take) / safe : 0.0;
                h1[i] = h1[i] - take;
                hu1[i] = hu1[i] * scale; hv1[i] = hv1[i] * scale;
                double taken = take;
                if (G->S) {
                    double ts = dmin(pot - take, S[i]);
                    S[i] = S[i] - ts;
                    taken = taken + ts;
                }
                IC[i] = IC[i] + taken;
                acc += taken;
            }
            inf_v = acc * area;
        }
        if (G->S) {
            const double *Sc = G->S + b * G->S_bs;
            for (size_t i = 0; i < plane; ++i) {
                double room = clamp0(Sc[i] - S[i]);
                double fill = dmin(room, h1[i]);
                double safe = h1[i] > 0.0 ? h1[i] : 1.0;
                double scale = h1[i] > 0.0 ? (h1[i] - fill) / safe : 0.0;
                h1[i] = h1[i] - fill;
                hu1[i] = hu1[i] * scale; hv1[i] = hv1[i] * scale;
                S[i] = S[i] + fill;
            }
        }
        /* semi-implicit friction */
        const double *nn = G->n + b * G->n_bs;
        double *H = h + b * plane, *HU = hu + b * plane, *HV = hv + b * plane;
        for (size_t i = 0; i < plane; ++i) {
            double u, v;
            kp_vel(h1[i], hu1[i], hv1[i], G->kp_eps, G->h_dry, &u, &v);
            double speed = sqrt(u * u + v * v);
            double hf = h1[i] < G->h_dry ? G->h_dry : h1[i];
            double denom = 1.0 + dt * G->g * nn[i] * nn[i] * speed / pow(hf, 4.0 / 3.0);
            u = u / denom; v = v / denom;
            int wet = h1[i] >= G->h_dry;
            H[i] = h1[i];
            HU[i] = wet ? h1[i] * u : 0.0;
            HV[i] = wet ? h1[i] * v : 0.0;
        }
        vols[0 * B + b] = bin; vols[1 * B + b] = bout; vols[2 * B + b] = rain_v;
        vols[3 * B + b] = inf_v; vols[4 * B + b] = clip;
        free(d0); free(d1); free(d2); free(h1); free(hu1); free(hv1);
        free_work(&W);
    }
}

/* ------------------------------------------------------------------ 1-D network */

typedef struct {
    int32_t B, nc, nf, ngh;
    double g, h_dry;
    /* cell sections; member stride bs (0 or nc) for b, yb, ts, z */
    const double *z, *b, *yb, *ts; int64_t bs;
    const double *m, *nman, *dx;           /* shared (nc) */
    /* face sections; member stride fbs (0 or nf) for fb, fyb, fts */
    const double *fb, *fyb, *fts; int64_t fbs;
    const double *fm;                      /* shared (nf) */
    const int32_t *fl, *fr;                /* (nf) */
    const int32_t *kind;                   /* (nc): 0 reach, 1 junction, 2 inflow, 3 stage, 4 closed */
    const int32_t *g_idx, *g_src;          /* (ngh) */
    const double *g_sign;                  /* (ngh) */
} orc_net;

typedef struct { double b, m, yb, ts; } sec_t;

static inline double sec_bank_area(sec_t s) { return s.yb * (s.b + s.m * s.yb); }
static inline double sec_area(sec_t s, double y) {
    double yt = dmin(y, s.yb);
    double at = yt * (s.b + s.m * yt);
    return at + s.ts * clamp0(y - s.yb);
}
static inline double sec_top(sec_t s, double y) { return y <= s.yb ? s.b + 2.0 * s.m * y : s.ts; }
static inline double sec_i1(sec_t s, double y) {
    double yt = dmin(y, s.yb);
    double above = clamp0(y - s.yb);
    double i_t = s.b * yt * yt / 2.0 + s.m * (yt * yt * yt) / 3.0;
    double at = yt * (s.b + s.m * yt);
    return i_t + at * above + s.ts * above * above / 2.0;
}
static inline double sec_perim(sec_t s, double y) {
    double yt = dmin(y, s.yb);
    return s.b + 2.0 * yt * sqrt(1.0 + s.m * s.m);
}
static inline double sec_depth(sec_t s, double A) {
    A = clamp0(A);
    double ab = sec_bank_area(s);
    double y_t = 2.0 * A / (s.b + sqrt(s.b * s.b + 4.0 * s.m * A));
    return A <= ab ? y_t : s.yb + (A - ab) / s.ts;
}

static inline sec_t cell_sec(const orc_net *N, int bm, int c) {
    sec_t s = { N->b[bm * N->bs + c], N->m[c], N->yb[bm * N->bs + c], N->ts[bm * N->bs + c] };
    return s;
}
static inline sec_t face_sec(const orc_net *N, int bm, int f) {
    sec_t s = { N->fb[bm * N->fbs + f], N->fm[f], N->fyb[bm * N->fbs + f], N->fts[bm * N->fbs + f] };
    return s;
}

static inline double vel1(sec_t s, double A, double Q, double h_dry) {
    double y = sec_depth(s, A);
    return y >= h_dry ? Q / (A < 1e-12 ? 1e-12 : A) : 0.0;
}

/* swe1d.py:253-262 per member: rate[b] = max over wet interior cells, wet[b] = any wet */
void orc_swe1d_max_rate(const orc_net *N, const double *A, const double *Q, double *rate, int32_t *any_wet) {
    for (int bm = 0; bm < N->B; ++bm) {
        double m = 0.0; int w = 0;
        for (int c = 0; c < N->nc; ++c) {
            if (N->kind[c] >= 2) continue;
            sec_t s = cell_sec(N, bm, c);
            double a = A[(size_t)bm * N->nc + c], q = Q[(size_t)bm * N->nc + c];
            double y = sec_depth(s, a);
            if (!(y > N->h_dry)) continue;
            w = 1;
            double cc = sqrt(N->g * a / sec_top(s, y));
            double u = vel1(s, a, q, N->h_dry);
            double r = (fabs(u) + cc) / N->dx[c];
            if (r > m) m = r;
        }
        rate[bm] = m; any_wet[bm] = w;
    }
}

static double hll1(double AL, double QL, double pL, double cL, double AR, double QR, double pR, double cR,
                   double uL, double uR, double *F2o) {
    int dryL = AL <= TINY, dryR = AR <= TINY;
    double SL = dmin(uL - cL, uR - cR);
    double SR = dmax(uL + cL, uR + cR);
    if (dryL) { SL = uR - 2.0 * cR; SR = uR + cR; }
    if (dryR) { SL = uL - cL; SR = uL + 2.0 * cL; }
    double FL1 = QL, FR1 = QR, FL2 = QL * uL + pL, FR2 = QR * uR + pR;
    double den = SR - SL;
    if (fabs(den) < TINY) den = 1.0;
    double H1 = (SR * FL1 - SL * FR1 + SL * SR * (AR - AL)) / den;
    double H2 = (SR * FL2 - SL * FR2 + SL * SR * (QR - QL)) / den;
    int left = SL >= 0.0, right = SR <= 0.0;
    double F1 = left ? FL1 : (right ? FR1 : H1);
    double F2 = left ? FL2 : (right ? FR2 : H2);
    if (dryL && dryR) { F1 = 0.0; F2 = 0.0; }
    *F2o = F2;
    return F1;
}

/*
 * swe1d.py:320-368 for every member; A, Q (B, nc) are read, A1, Q1 written; q_in and eta (B);
 * vols (3, B): b_in, b_out, clip.
 */
void orc_swe1d_step(const orc_net *N, const double *A0, const double *Q0, double dt, const double *q_in,
                    const double *eta, double *A1o, double *Q1o, double *vols) {
    int nc = N->nc, nf = N->nf, B = N->B;
    double g = N->g, h_dry = N->h_dry;
    double *A = malloc(nc * 8), *Q = malloc(nc * 8), *y = malloc(nc * 8), *u = malloc(nc * 8);
    double *F1 = malloc(nf * 8), *F2 = malloc(nf * 8), *pL = malloc(nf * 8), *pR = malloc(nf * 8);
    double *out = malloc(nc * 8), *th = malloc(nc * 8), *dA = malloc(nc * 8), *dQ = malloc(nc * 8);
    for (int bm = 0; bm < B; ++bm) {
        const double *Ab = A0 + (size_t)bm * nc, *Qb = Q0 + (size_t)bm * nc;
        memcpy(A, Ab, nc * 8); memcpy(Q, Qb, nc * 8);
        /* ghosts, swe1d.py:265-294 (all ghosts read the pre-fill state) */
        double *Ag = malloc((N->ngh + 1) * 8), *Qg = malloc((N->ngh + 1) * 8);
        for (int k = 0; k < N->ngh; ++k) {
            int gi = N->g_idx[k], src = N->g_src[k];
            double sg = N->g_sign[k];
            sec_t s = cell_sec(N, bm, gi);
            double As = Ab[src], Qs = Qb[src];
            double ys = sec_depth(s, As);
            double zg = N->z[bm * N->bs + gi];
            double un = sg * vel1(s, As, Qs, h_dry);
            int wet_i = ys >= h_dry;
            double yst = clamp0(eta[bm] - zg);
            double Ast = sec_area(s, yst);
            double ci = sqrt(g * As / sec_top(s, ys));
            double cg = sqrt(g * Ast / sec_top(s, yst));
            double un_st = wet_i ? un + 2.0 * (ci - cg) : 0.0;
            int sup = wet_i && (un > ci);
            if (sup) { un_st = un; Ast = As; }
            if (!(sec_depth(s, Ast) >= h_dry)) un_st = 0.0;
            double Q_st = sg * Ast * un_st;
            double yc = pow(q_in[bm] * q_in[bm] / (g * s.b * s.b), 1.0 / 3.0);
            if (yc < h_dry) yc = h_dry;
            double A_in = wet_i ? As : sec_area(s, yc);
            double Q_in = -sg * q_in[bm];
            int kd = N->kind[gi];
            Ag[k] = kd == 3 ? Ast : (kd == 2 ? A_in : As);
            Qg[k] = kd == 3 ? Q_st : (kd == 2 ? Q_in : -Qs);
        }
        for (int k = 0; k < N->ngh; ++k) { A[N->g_idx[k]] = Ag[k]; Q[N->g_idx[k]] = Qg[k]; }
        free(Ag); free(Qg);
        for (int c = 0; c < nc; ++c) { sec_t s = cell_sec(N, bm, c); y[c] = sec_depth(s, A[c]); u[c] = vel1(s, A[c], Q[c], h_dry); }
        for (int f = 0; f < nf; ++f) {
            int l = N->fl[f], r = N->fr[f];
            sec_t fs = face_sec(N, bm, f);
            double zL = N->z[bm * N->bs + l], zR = N->z[bm * N->bs + r];
            double zf = dmax(zL, zR);
            double yL = clamp0(zL + y[l] - zf), yR = clamp0(zR + y[r] - zf);
            double AL = sec_area(fs, yL), AR = sec_area(fs, yR);
            double uL = u[l], uR = u[r];
            pL[f] = g * sec_i1(fs, yL); pR[f] = g * sec_i1(fs, yR);
            double cL = sqrt(g * AL / sec_top(fs, yL)), cR = sqrt(g * AR / sec_top(fs, yR));
            F1[f] = hll1(AL, AL * uL, pL[f], cL, AR, AR * uR, pR[f], cR, uL, uR, &F2[f]);
        }
        for (int c = 0; c < nc; ++c) out[c] = 0.0;
        for (int f = 0; f < nf; ++f) out[N->fl[f]] += relu(F1[f]);
        for (int f = 0; f < nf; ++f) out[N->fr[f]] += relu(-F1[f]);
        for (int c = 0; c < nc; ++c) {
            double vol = A[c] * N->dx[c];
            double need = out[c] * dt;
            double t = need > vol ? vol / (need > 0.0 ? need : 1.0) : 1.0;
            th[c] = N->kind[c] >= 2 ? 1.0 : t;
        }
        for (int f = 0; f < nf; ++f) { double s = F1[f] > 0.0 ? th[N->fl[f]] : th[N->fr[f]]; F1[f] = F1[f] * s; }
        for (int c = 0; c < nc; ++c) { dA[c] = 0.0; dQ[c] = 0.0; }
        for (int f = 0; f < nf; ++f) dA[N->fl[f]] += -F1[f];
        for (int f = 0; f < nf; ++f) dA[N->fr[f]] += F1[f];
        for (int f = 0; f < nf; ++f) dQ[N->fl[f]] += -(F2[f] - pL[f]);
        for (int f = 0; f < nf; ++f) dQ[N->fr[f]] += F2[f] - pR[f];
        double clip = 0.0, bin = 0.0, bout = 0.0;
        double *A1 = A1o + (size_t)bm * nc, *Q1 = Q1o + (size_t)bm * nc;
        for (int c = 0; c < nc; ++c) {
            A1[c] = A[c] + dt * dA[c] / N->dx[c];
            Q1[c] = Q[c] + dt * dQ[c] / N->dx[c];
            double ng_ = clamp0(-A1[c]);
            clip += ng_ * N->dx[c] * (N->kind[c] >= 2 ? 0.0 : 1.0);
            A1[c] = A1[c] + ng_;
        }
        double i1 = 0.0, i2 = 0.0, o1 = 0.0, o2 = 0.0;
        for (int f = 0; f < nf; ++f) {
            int gl = N->kind[N->fl[f]] >= 2, gr = N->kind[N->fr[f]] >= 2;
            i1 += gl ? relu(F1[f]) : 0.0; i2 += gr ? relu(-F1[f]) : 0.0;
            o1 += gl ? relu(-F1[f]) : 0.0; o2 += gr ? relu(F1[f]) : 0.0;
        }
        bin = (i1 + i2) * dt; bout = (o1 + o2) * dt;
        for (int c = 0; c < nc; ++c) {
            sec_t s = cell_sec(N, bm, c);
            double y1 = sec_depth(s, A1[c]);
            double Ac = A1[c] < 1e-12 ? 1e-12 : A1[c];
            double R = Ac / sec_perim(s, y1);
            double Rc = R < h_dry ? h_dry : R;
            double denom = 1.0 + dt * g * (N->nman[c] * N->nman[c]) * fabs(Q1[c]) / (Ac * pow(Rc, 4.0 / 3.0));
            Q1[c] = Q1[c] / denom;
            if (!((y1 >= h_dry) && N->kind[c] != 1)) Q1[c] = 0.0;
            if (N->kind[c] >= 2) { A1[c] = A[c]; Q1[c] = Q[c]; }
        }
        vols[0 * B + bm] = bin; vols[1 * B + bm] = bout; vols[2 * B + bm] = clip;
    }
    free(A); free(Q); free(y); free(u); free(F1); free(F2); free(pL); free(pR);
    free(out); free(th); free(dA); free(dQ);
}

/* ------------------------------------------------------------------ exchange */

typedef struct {
    int32_t nl, nu;                /* links, unique linked 2-D cells */
    const int32_t *cell1d;         /* (nl) */
    const int32_t *inv;            /* (nl) -> unique slot */
    const int64_t *uniq;           /* (nu) flat 2-D cell index */
    const double *zbank, *length, *z2;   /* (nl) */
    double area2, weir_cw, limit, relax;
} orc_links;

/*
 * coupling.py:73-124.  h, hu, hv (B, ny*nx) and A (B, nc) are updated in place; V (B, nl) out.
 */
void orc_exchange(const orc_links *L, const orc_net *N, int64_t plane, double *h, double *hu, double *hv,
                  double *A, double dt, double *V) {
    int nl = L->nl, nu = L->nu, nc = N->nc;
    double *hu_ = malloc(nu * 8), *tot2 = malloc(nu * 8), *f2 = malloc(nu * 8), *dV2 = malloc(nu * 8);
    double *tot1 = malloc(nc * 8), *f1 = malloc(nc * 8), *addA = malloc(nc * 8), *y = malloc(nc * 8);
    double *sign = malloc(nl * 8), *Vl = malloc(nl * 8);
    for (int bm = 0; bm < N->B; ++bm) {
        double *H = h + bm * plane, *HU = hu + bm * plane, *HV = hv + bm * plane, *Ab = A + (size_t)bm * nc;
        for (int k = 0; k < nu; ++k) hu_[k] = H[L->uniq[k]];
        for (int c = 0; c < nc; ++c) y[c] = sec_depth(cell_sec(N, bm, c), Ab[c]);
        for (int l = 0; l < nl; ++l) {
            int c1 = L->cell1d[l];
            sec_t s = cell_sec(N, bm, c1);
            double h2 = hu_[L->inv[l]];
            double eta2 = L->z2[l] + h2;
            double y1 = y[c1];
            double eta1 = N->z[bm * N->bs + c1] + y1;
            double up = dmax(eta1, eta2), dn = dmin(eta1, eta2);
            double d_up = clamp0(up - L->zbank[l]), d_dn = clamp0(dn - L->zbank[l]);
            double ratio = d_up > 0.0 ? d_dn / (d_up > 0.0 ? d_up : 1.0) : 0.0;
            double q = L->weir_cw * L->length[l] * pow(d_up, 1.5) * pow(clamp0(1.0 - pow(ratio, 1.5)), 0.385);
            sign[l] = eta2 > eta1 ? 1.0 : -1.0;
            double v = q * dt;
            double plan1 = (y1 <= s.yb ? s.b + 2 * s.m * y1 : s.ts) * N->dx[c1];
            double v_eq = (up - dn) / (1.0 / L->area2 + 1.0 / plan1);
            Vl[l] = dmin(v, L->relax * v_eq);
        }
        for (int k = 0; k < nu; ++k) tot2[k] = 0.0;
        for (int l = 0; l < nl; ++l) tot2[L->inv[l]] += sign[l] > 0.0 ? Vl[l] : 0.0;
        for (int k = 0; k < nu; ++k) {
            double av = L->limit * hu_[k] * L->area2;
            f2[k] = tot2[k] > av ? av / (tot2[k] > 0.0 ? tot2[k] : 1.0) : 1.0;
        }
        for (int c = 0; c < nc; ++c) tot1[c] = 0.0;
        for (int l = 0; l < nl; ++l) tot1[L->cell1d[l]] += sign[l] < 0.0 ? Vl[l] : 0.0;
        for (int c = 0; c < nc; ++c) {
            sec_t s = cell_sec(N, bm, c);
            double av = L->limit * clamp0(Ab[c] - sec_bank_area(s)) * N->dx[c];
            f1[c] = tot1[c] > av ? av / (tot1[c] > 0.0 ? tot1[c] : 1.0) : 1.0;
        }
        for (int l = 0; l < nl; ++l) {
            double w = sign[l] > 0.0 ? Vl[l] * f2[L->inv[l]] : Vl[l] * f1[L->cell1d[l]];
            Vl[l] = w * sign[l];
            V[(size_t)bm * nl + l] = Vl[l];
        }
        for (int k = 0; k < nu; ++k) dV2[k] = 0.0;
        for (int l = 0; l < nl; ++l) dV2[L->inv[l]] += Vl[l];
        for (int k = 0; k < nu; ++k) {
            double hn = hu_[k] - dV2[k] / L->area2;
            double keep = hu_[k] > 0.0 ? dmin(hn / (hu_[k] > 0.0 ? hu_[k] : 1.0), 1.0) : 0.0;
            int64_t c = L->uniq[k];
            H[c] = hn;
            HU[c] = HU[c] * keep;
            HV[c] = HV[c] * keep;
        }
        for (int c = 0; c < nc; ++c) addA[c] = 0.0;
        for (int l = 0; l < nl; ++l) addA[L->cell1d[l]] += Vl[l];
        for (int c = 0; c < nc; ++c) Ab[c] = Ab[c] + addA[c] / N->dx[c];
    }
    free(hu_); free(tot2); free(f2); free(dV2); free(tot1); free(f1); free(addA); free(y); free(sign); free(Vl);
}
