/*
 * wql_oracle.c -- FP64 CPU restatement of the WaveQLab3D time-stepping hot
 * path (TEST INFRASTRUCTURE: the checker for the HIP path, never the product).
 *
 * Every function follows a reference routine cited as file:line (relative to
 * the reference's src/).  Operation order inside each expression follows the
 * Fortran source so that FP64 results agree to rounding.  Indices here are
 * 0-based (reference index i <-> i-1 here).
 *
 * Parity status (what pins each function; the tests live in tests/):
 *   ora_metrics      golden metrics/J/normals of the reference build (curv_sw.npz)
 *                    + free-stream identity + exactness on linear maps
 *   ora_set_rates    golden DF after set_rates (cart_sw, curv_sw, curv_rs, pml_locked)
 *                    + SBP identity / polynomial exactness of the closure table;
 *                    upwind / upwind-DRP operators: golden DF of seven operators
 *                    (up2_cart, up4_curv, up6_pml, up9_curv, drp3_curv,
 *                    drp66_cart, drp679_curv) + the upwind SBP identity and
 *                    consistency of every table row (all fifteen operators)
 *   ora_boundary,
 *   ora_interface    golden DF after the SAT terms, golden DS/DW (same cases)
 *                    + closed-form welded/locked checks
 *   ora_step         golden trajectories (all six cases incl. PML, R&S, plasticity)
 *   ora_hat          slip-weakening closed form, R&S residual |f(V)| ~ 0
 */
#include "wql_oracle.h"
#include "wql_sbp_ops.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(b, i, j, k) ((size_t)(i) + (size_t)(b)->nq * ((size_t)(j) + (size_t)(b)->nr * (size_t)(k)))
#define NPTS(b) ((size_t)(b)->nq * (b)->nr * (b)->ns)

static const double PI = 3.141592653589793;

/* ------------------------------------------------------------------------ */
/* SBP 6th-order first-derivative closure (JU_xJU_yJU_z6.f90:77-298 and the
 * identical table in metrics.f90:16-160).  Rows 1..6 as (column, coefficient)
 * pairs in the order the reference sums them; rows n-5..n mirror with sign
 * flip (JU..:188-298).  Central row: JU..:57-70. */
typedef struct { int n; int col[8]; double c[8]; } sbp_row;

static const sbp_row LEFT[6] = {
    {6, {6, 5, 4, 3, 2, 1}, {0.036577936277544, 0.104488069284042, -0.450398306578272,
                             -0.141512858744873, 2.033378678700676, -1.582533518939116}},
    {5, {6, 5, 4, 3, 1}, {-0.014903449191300, -0.069112065532624, 0.258816087376832,
                          0.287258622978251, -0.462059195631158}},
    {5, {6, 5, 4, 2, 1}, {-0.018129342917256, -0.022902190275815, 0.606235523609147,
                          -0.636451095137907, 0.071247104721830}},
    {6, {7, 6, 5, 3, 2, 1}, {0.013435342414630, -0.051642265516119, 0.520262285050482,
                             -0.306681191361148, -0.290087484386815, 0.114713313798970}},
    {7, {8, 7, 6, 4, 3, 2, 1}, {0.018281071473911, -0.164529643265203, 0.769199413962647,
                                -0.707905442575989, 0.015764336127392, 0.105400944933782,
                                -0.036210680656541}},
    {8, {9, 8, 7, 5, 4, 3, 2, 1}, {0.016437980868017, -0.147941827812150, 0.739709139060752,
                                   -0.691649024426814, 0.063183694641876, 0.011220896474665,
                                   0.020437334208704, -0.011398193015050}},
};

/* Right-boundary rows are written in the reference with ascending column
 * order (JU..:188-298): row n-5 sums n-8, n-7, n-6, n-4, n-3, n-2, n-1, n. */
static const sbp_row RIGHT[6] = { /* offsets from n (column = n + off), rows n-5 .. n */
    {8, {-8, -7, -6, -4, -3, -2, -1, 0}, {-0.016437980868017, 0.147941827812150, -0.739709139060752,
                                         0.691649024426814, -0.063183694641876, -0.011220896474665,
                                         -0.020437334208704, 0.011398193015050}},
    {7, {-7, -6, -5, -3, -2, -1, 0}, {-0.018281071473911, 0.164529643265203, -0.769199413962647,
                                      0.707905442575989, -0.015764336127392, -0.105400944933782,
                                      0.036210680656541}},
    {6, {-6, -5, -4, -2, -1, 0}, {-0.013435342414630, 0.051642265516119, -0.520262285050482,
                                  0.306681191361148, 0.290087484386815, -0.114713313798970}},
    {5, {-5, -4, -3, -1, 0}, {0.018129342917256, 0.022902190275815, -0.606235523609147,
                              0.636451095137907, -0.071247104721830}},
    {5, {-5, -4, -3, -2, 0}, {0.014903449191300, 0.069112065532624, -0.258816087376832,
                              -0.287258622978251, 0.462059195631158}},
    {6, {-5, -4, -3, -2, -1, 0}, {-0.036577936277544, -0.104488069284042, 0.450398306578272,
                                  0.141512858744873, -2.033378678700676, 1.582533518939116}},
};

/* central row: -1/60, 0.15, -0.75, (0), 0.75, -0.15, 1/60 at offsets -3..3 */
static const int CEN_OFF[6] = {-3, -2, -1, 1, 2, 3};
static double cen_coef(int k)
{
    switch (k) {
    case 0: return -1.0 / 60.0;
    case 1: return 0.15;
    case 2: return -0.75;
    case 3: return 0.75;
    case 4: return -0.15;
    default: return 1.0 / 60.0;
    }
}

/* Stencil (0-based indices + coefficients, in reference summation order) for
 * derivative row i of an n-point line.  Returns the count. */
static int stencil(int i, int n, int *idx, double *c)
{
    int r = i + 1; /* 1-based row */
    if (r >= 7 && r <= n - 6) {
        for (int k = 0; k < 6; k++) { idx[k] = i + CEN_OFF[k]; c[k] = cen_coef(k); }
        return 6;
    }
    if (r <= 6) {
        const sbp_row *R = &LEFT[r - 1];
        for (int k = 0; k < R->n; k++) { idx[k] = R->col[k] - 1; c[k] = R->c[k]; }
        return R->n;
    }
    const sbp_row *R = &RIGHT[r - (n - 5)];
    for (int k = 0; k < R->n; k++) { idx[k] = n - 1 + R->col[k]; c[k] = R->c[k]; }
    return R->n;
}

double ora_d1(const double *u, int n, int i, double h)
{
    int idx[8]; double c[8];
    int m = stencil(i, n, idx, c);
    double s = c[0] * u[idx[0]];
    for (int k = 1; k < m; k++) s = s + c[k] * u[idx[k]];
    return (1.0 / h) * s;
}

/* ------------------------------------------------------------------------ */
/* metrics: grid.f90:396-469 (derivatives) with D6x/D6y/D6z (metrics.f90:16-490)
 * which divide the closure sum by h; face frames: unit_normals.f90:8-168,
 * gen_orth_vectors 232-266, Gram_Schmidt 269-282. */

static void d6(const ora_block *b, const double *X, int dir, int i, int j, int k, double out[3])
{
    int n = dir == 0 ? b->nq : dir == 1 ? b->nr : b->ns;
    double h = 1.0 / (double)(n - 1);
    int pos = dir == 0 ? i : dir == 1 ? j : k;
    int idx[8]; double c[8];
    int m = stencil(pos, n, idx, c);
    size_t NP = NPTS(b);
    for (int f = 0; f < 3; f++) {
        double s = 0.0;
        for (int t = 0; t < m; t++) {
            int ii = i, jj = j, kk = k;
            if (dir == 0) ii = idx[t]; else if (dir == 1) jj = idx[t]; else kk = idx[t];
            double v = c[t] * X[f * NP + IDX(b, ii, jj, kk)];
            s = (t == 0) ? v : s + v;
        }
        out[f] = (1.0 / h) * s;
    }
}

static void gram_schmidt(const double y[3], double z[3])
{
    double a = y[0] * z[0] + y[1] * z[1] + y[2] * z[2];
    for (int t = 0; t < 3; t++) z[t] = z[t] - a * y[t];
    double s = 1.0 / sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    for (int t = 0; t < 3; t++) z[t] = s * z[t];
}

static int gen_orth(const double n[3], double l[3], double m[3])
{
    const double tol = 1.0e-12;
    m[0] = 0.0; m[1] = 1.0; m[2] = 0.0;
    if (fabs(sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) - 1.0) >= tol) return -1;
    double d1 = sqrt((n[0] - m[0]) * (n[0] - m[0]) + (n[1] - m[1]) * (n[1] - m[1]) + (n[2] - m[2]) * (n[2] - m[2]));
    double d2 = sqrt((n[0] + m[0]) * (n[0] + m[0]) + (n[1] + m[1]) * (n[1] + m[1]) + (n[2] + m[2]) * (n[2] + m[2]));
    if (!(d1 >= tol && d2 >= tol)) { m[0] = 0.0; m[1] = 0.0; m[2] = 1.0; }
    gram_schmidt(n, m);
    l[0] = n[1] * m[2] - n[2] * m[1];
    l[1] = n[2] * m[0] - n[0] * m[2];
    l[2] = n[0] * m[1] - n[1] * m[0];
    return 0;
}

static size_t face_pts(const ora_block *b, int side)
{
    if (side < 2) return (size_t)b->nr * b->ns;
    if (side < 4) return (size_t)b->nq * b->ns;
    return (size_t)b->nq * b->nr;
}

int ora_metrics(ora_block *b)
{
    size_t NP = NPTS(b);
    double *g = b->metric;
    for (int k = 0; k < b->ns; k++)
        for (int j = 0; j < b->nr; j++)
            for (int i = 0; i < b->nq; i++) {
                double Xq[3], Xr[3], Xs[3];
                d6(b, b->X, 0, i, j, k, Xq);
                d6(b, b->X, 1, i, j, k, Xr);
                d6(b, b->X, 2, i, j, k, Xs);
                size_t p = IDX(b, i, j, k);
                double J = Xq[0] * (Xr[1] * Xs[2] - Xr[2] * Xs[1])
                         - Xq[1] * (Xr[0] * Xs[2] - Xr[2] * Xs[0])
                         + Xq[2] * (Xr[0] * Xs[1] - Xr[1] * Xs[0]);
                double iJ = 1.0 / J;
                g[0 * NP + p] = iJ * (Xr[1] * Xs[2] - Xr[2] * Xs[1]);
                g[1 * NP + p] = iJ * (Xq[2] * Xs[1] - Xq[1] * Xs[2]);
                g[2 * NP + p] = iJ * (Xq[1] * Xr[2] - Xq[2] * Xr[1]);
                g[3 * NP + p] = iJ * (Xr[2] * Xs[0] - Xr[0] * Xs[2]);
                g[4 * NP + p] = iJ * (Xq[0] * Xs[2] - Xq[2] * Xs[0]);
                g[5 * NP + p] = iJ * (Xq[2] * Xr[0] - Xq[0] * Xr[2]);
                g[6 * NP + p] = iJ * (Xr[0] * Xs[1] - Xr[1] * Xs[0]);
                g[7 * NP + p] = iJ * (Xq[1] * Xs[0] - Xq[0] * Xs[1]);
                g[8 * NP + p] = iJ * (Xq[0] * Xr[1] - Xq[1] * Xr[0]);
                g[9 * NP + p] = J;
            }
    /* face frames: n = grad(xi)/|grad(xi)| at the face, xi = q, r or s */
    for (int side = 0; side < 6; side++) {
        int dir = side / 2;
        size_t FP = face_pts(b, side);
        double *nm = b->nml[side];
        if (!nm) continue;
        int na = dir == 0 ? b->nr : b->nq;
        int nb2 = dir == 2 ? b->nr : b->ns;
        for (int kk = 0; kk < nb2; kk++)
            for (int jj = 0; jj < na; jj++) {
                int i, j, k;
                if (dir == 0) { i = (side == 0) ? 0 : b->nq - 1; j = jj; k = kk; }
                else if (dir == 1) { j = (side == 2) ? 0 : b->nr - 1; i = jj; k = kk; }
                else { k = (side == 4) ? 0 : b->ns - 1; i = jj; j = kk; }
                size_t p = IDX(b, i, j, k);
                double ax = g[(0 + dir) * NP + p], ay = g[(3 + dir) * NP + p], az = g[(6 + dir) * NP + p];
                double norm = sqrt(pow(ax, 2.0) + pow(ay, 2.0) + pow(az, 2.0));
                double n[3] = {ax / norm, ay / norm, az / norm}, l[3], m[3];
                if (gen_orth(n, l, m)) return -1;
                size_t fp = (size_t)jj + (size_t)na * kk;
                for (int t = 0; t < 3; t++) {
                    nm[(0 + t) * FP + fp] = n[t];
                    nm[(3 + t) * FP + fp] = m[t];
                    nm[(6 + t) * FP + fp] = l[t];
                }
            }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* volume operator */

/* one axis contribution of JJU_x6 (JU_xJU_yJU_z6.f90:12-812) for metric
 * component `mc` (0..8 = metricx/y/z(1..3)) along direction dir at (i,j,k):
 * velocity fields use X_x at the centre times the closure sum, stresses sum
 * c*F*J*X_x at the stencil points. */
static void jju_axis(const ora_block *b, int dir, int mc, int i, int j, int k, double out[9])
{
    int n = dir == 0 ? b->nq : dir == 1 ? b->nr : b->ns;
    int pos = dir == 0 ? i : dir == 1 ? j : k;
    int idx[8]; double c[8];
    int m = stencil(pos, n, idx, c);
    size_t NP = NPTS(b);
    const double *F = b->F, *G = b->metric;
    size_t pc = IDX(b, i, j, k);
    for (int f = 0; f < 9; f++) {
        double s = 0.0;
        for (int t = 0; t < m; t++) {
            int ii = i, jj = j, kk = k;
            if (dir == 0) ii = idx[t]; else if (dir == 1) jj = idx[t]; else kk = idx[t];
            size_t p = IDX(b, ii, jj, kk);
            double v;
            if (f < 3) v = c[t] * F[f * NP + p];
            else v = c[t] * F[f * NP + p] * G[9 * NP + p] * G[mc * NP + p];
            s = (t == 0) ? v : s + v;
        }
        out[f] = (f < 3) ? G[mc * NP + pc] * s : s;
    }
}

/* JJU_x6 for metric family fam (0: metricx, 1: metricy, 2: metricz):
 * Ju = (1/hq) Jq + (1/hr) Jr + (1/hs) Js  (JU..:808) */
static void jju6(const ora_block *b, int fam, int i, int j, int k, double U[9])
{
    double hq = 1.0 / (double)(b->nq - 1), hr = 1.0 / (double)(b->nr - 1), hs = 1.0 / (double)(b->ns - 1);
    double a[9], bb[9], cc[9];
    jju_axis(b, 0, 3 * fam + 0, i, j, k, a);
    jju_axis(b, 1, 3 * fam + 1, i, j, k, bb);
    jju_axis(b, 2, 3 * fam + 2, i, j, k, cc);
    for (int f = 0; f < 9; f++) U[f] = (1.0 / hq) * a[f] + (1.0 / hr) * bb[f] + (1.0 / hs) * cc[f];
}

/* PML slab geometry: slab `side` covers the owned range restricted to the
 * layer (pml.f90:11-168).  Returns linear index into the slab array or -1. */
static long slab_index(const ora_block *b, int side, int i, int j, int k)
{
    int np = b->npml;
    int lo[3] = {0, 0, 0}, hi[3] = {b->nq - 1, b->nr - 1, b->ns - 1};
    int dir = side / 2;
    if (side % 2 == 0) hi[dir] = np - 1;
    else lo[dir] = (dir == 0 ? b->nq : dir == 1 ? b->nr : b->ns) - np;
    int p[3] = {i, j, k};
    for (int t = 0; t < 3; t++) if (p[t] < lo[t] || p[t] > hi[t]) return -1;
    long n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1;
    return (long)(p[0] - lo[0]) + n0 * ((long)(p[1] - lo[1]) + n1 * (long)(p[2] - lo[2]));
}

size_t ora_slab_points(const ora_block *b, int side)
{
    int dir = side / 2;
    size_t n[3] = {(size_t)b->nq, (size_t)b->nr, (size_t)b->ns};
    n[dir] = (size_t)b->npml;
    return n[0] * n[1] * n[2];
}

/* PML: RHS_Interior.f90:1782-2201, order z-low, z-high, y-low, y-high, x-low, x-high */
static void pml_point(ora_block *b, double DU[9], const double Ux[9], const double Uy[9],
                      const double Uz[9], int i, int j, int k, double rhoJ_inv, double l2m)
{
    size_t NP = NPTS(b);
    size_t p = IDX(b, i, j, k);
    double lam = b->M[p], mu = b->M[NP + p];
    double J = b->metric[9 * NP + p];
    const double Rf = 1.0e-3, cfs = 1.0;
    double d0 = 6.0 / (2.0 * 1.0) * log(1.0 / Rf);
    int np = b->npml;
    static const int order[6] = {4, 5, 2, 3, 0, 1};
    for (int o = 0; o < 6; o++) {
        int side = order[o];
        if (!b->pml[side]) continue;
        long s = slab_index(b, side, i, j, k);
        if (s < 0) continue;
        int dir = side / 2;
        int pos = (dir == 0 ? i : dir == 1 ? j : k) + 1; /* 1-based */
        int n = dir == 0 ? b->nq : dir == 1 ? b->nr : b->ns;
        double d;
        if (side == 4) d = d0 * fabs((double)(pos - np) / (double)np);          /* 1855 */
        else if (side % 2 == 0) d = d0 * fabs((double)(pos - (1 + np)) / (double)np); /* 1973, 2092 */
        else d = d0 * fabs((double)(pos - (n - np)) / (double)np);              /* 1913, 2033, 2150 */
        size_t SP = ora_slab_points(b, side);
        double *Q = b->Q[side], *DQ = b->DQ[side];
        double q[9];
        for (int f = 0; f < 9; f++) q[f] = Q[f * SP + s];
        DU[0] = DU[0] - d * J * q[0] * rhoJ_inv;
        DU[1] = DU[1] - d * J * q[1] * rhoJ_inv;
        DU[2] = DU[2] - d * J * q[2] * rhoJ_inv;
        DU[3] = DU[3] - d * (l2m * q[3] + lam * q[4] + lam * q[5]);
        DU[4] = DU[4] - d * (lam * q[3] + l2m * q[4] + lam * q[5]);
        DU[5] = DU[5] - d * (lam * q[3] + lam * q[4] + l2m * q[5]);
        DU[6] = DU[6] - mu * d * q[6];
        DU[7] = DU[7] - mu * d * q[7];
        DU[8] = DU[8] - mu * d * q[8];
        double src[9];
        if (dir == 2) {
            src[0] = Uz[7] / J; src[1] = Uz[8] / J; src[2] = Uz[5] / J;
            src[3] = 0.0; src[4] = 0.0; src[5] = Uz[2]; src[6] = 0.0; src[7] = Uz[0]; src[8] = Uz[1];
        } else if (dir == 1) {
            src[0] = Uy[6] / J; src[1] = Uy[4] / J; src[2] = Uy[8] / J;
            src[3] = 0.0; src[4] = Uy[1]; src[5] = 0.0; src[6] = Uy[0]; src[7] = 0.0; src[8] = Uy[2];
        } else {
            src[0] = Ux[3] / J; src[1] = Ux[6] / J; src[2] = Ux[7] / J;
            src[3] = Ux[0]; src[4] = 0.0; src[5] = 0.0; src[6] = Ux[1]; src[7] = Ux[2]; src[8] = 0.0;
        }
        for (int f = 0; f < 9; f++) DQ[f * SP + s] = DQ[f * SP + s] + src[f] - (d + cfs) * q[f];
    }
}

/* ------------------------------------------------------------------------ */
/* anelastic response: GSLS memory variables */

int ora_anelastic_init(ora_block *b, double c, double weight_exp, double fref)
{
    static const double w1[4] = {1.6126, 0.6255, 0.6382, 1.5969}; /* weight_exp ~ 0 (material.f90:87-90) */
    static const double w2[4] = {0.0336, 0.6873, 0.8767, 1.5202}; /* weight_exp ~ 0.6 (101-104) */
    const double *w;
    if (weight_exp < 0.01) w = w1;
    else if (weight_exp > 0.59 && weight_exp < 0.61) w = w2;
    else return -1;
    size_t NP = NPTS(b);
    double *M = b->M, *Qs = b->qinv, *Qp = b->qinv + NP;
    for (size_t p = 0; p < NP; p++) { /* 64-71 */
        Qs[p] = 1.0 / (c * sqrt(M[NP + p] / M[2 * NP + p]));
        Qp[p] = 0.5 * Qs[p];
    }
    const double FAC = 1.0;
    double taumin = 1.0 / (2.0 * PI * 15.0) * 1.0 * FAC, taumax = 1.0 / (2.0 * PI * 0.08) * 200.0 * FAC;
    for (int k = 1; k <= 4; k++) { /* 82-85 */
        b->an_tau[k - 1] = exp(log(taumin) + (2.0 * k - 1.0) / 16.0 * (log(taumax) - log(taumin)));
        b->an_w[k - 1] = w[k - 1];
    }
    double wref = 2.0 * PI * fref, facex = 1.0;
    for (size_t p = 0; p < NP; p++) { /* 108-140: unrelaxed moduli */
        double val_S = 0.0, val_P = 0.0;
        for (int l = 0; l < 4; l++) {
            double den_S = (pow(wref, 2.0) * pow(b->an_tau[l], 2.0) + 1.0) * (1.0 / Qs[p]) * facex;
            double den_P = (pow(wref, 2.0) * pow(b->an_tau[l], 2.0) + 1.0) * (1.0 / Qp[p]) * facex;
            val_S = val_S + b->an_w[l] / den_S;
            val_P = val_P + b->an_w[l] / den_P;
        }
        double lam = M[p], mu = M[NP + p], rho = M[2 * NP + p];
        double vs = sqrt(mu / rho), vp = sqrt((lam + 2.0 * mu) / rho);
        double mu_S = rho * pow(vs, 2.0) / (1.0 - val_S), mu_P = rho * pow(vp, 2.0) / (1.0 - val_P);
        vs = sqrt(mu_S / rho);
        vp = sqrt(mu_P / rho);
        M[NP + p] = pow(vs, 2.0) * rho;
        M[p] = pow(vp, 2.0) * rho - 2.0 * M[NP + p];
    }
    b->anelastic = 1;
    return 0;
}

/* PML damping and auxiliary fields of the slabs holding node (i,j,k), per
 * direction (pml_damping_and_q, RHS_Interior.f90:2287-2387); returns 0 when
 * no slab holds it (point_in_pml, 2225-2285) */
static int pml_dq(const ora_block *b, int i, int j, int k, double d[3], double Q[3][9])
{
    const double d0 = 6.0 / (2.0 * 1.0) * log(1.0 / 1.0e-3);
    int in = 0, np = b->npml, pos[3] = {i + 1, j + 1, k + 1}, n[3] = {b->nq, b->nr, b->ns};
    for (int a = 0; a < 3; a++) {
        d[a] = 0.0;
        for (int f = 0; f < 9; f++) Q[a][f] = 0.0;
        for (int side = 2 * a; side <= 2 * a + 1; side++) {
            if (!b->pml[side]) continue;
            long s = slab_index(b, side, i, j, k);
            if (s < 0) continue;
            in = 1;
            if (side == 4) d[a] = d0 * fabs((double)(pos[a] - np) / (double)np);
            else if (side % 2 == 0) d[a] = d0 * fabs((double)(pos[a] - (1 + np)) / (double)np);
            else d[a] = d0 * fabs((double)(pos[a] - (n[a] - np)) / (double)np);
            size_t SP = ora_slab_points(b, side);
            for (int f = 0; f < 9; f++) Q[a][f] = b->Q[side][f * SP + s];
        }
    }
    return in;
}

/* memory-variable terms of one node: out[c] -= sum_l eta(c,l) for the six
 * stress rates (out[c] at stride `st`), Deta += rates (apply_anelastic_point,
 * RHS_Interior.f90:2390-2441; in PML slabs the strains net of the auxiliary
 * fields, apply_anelastic_point_pml 2444-2509; the interior routines' form
 * JU_xJU_yJU_z6.f90:16317-16353 is the same arithmetic) */
static void anelastic_point(ora_block *b, int i, int j, int k, const double Ux[9], const double Uy[9],
                            const double Uz[9], double *out, size_t st, int pml_form)
{
    size_t NP = NPTS(b), p = IDX(b, i, j, k);
    double lam = b->M[p], mu = b->M[NP + p], Qs = b->qinv[p], Qp = b->qinv[NP + p];
    double e[6];
    double d[3], Q[3][9];
    if (pml_form && pml_dq(b, i, j, k, d, Q)) {
        e[0] = Ux[0] - d[0] * Q[0][3];
        e[1] = Uy[1] - d[1] * Q[1][4];
        e[2] = Uz[2] - d[2] * Q[2][5];
        e[3] = (Uy[0] + Ux[1]) - d[0] * Q[0][6] - d[1] * Q[1][6];
        e[4] = (Uz[0] + Ux[2]) - d[0] * Q[0][7] - d[2] * Q[2][7];
        e[5] = (Uz[1] + Uy[2]) - d[1] * Q[1][8] - d[2] * Q[2][8];
    } else {
        e[0] = Ux[0]; e[1] = Uy[1]; e[2] = Uz[2];
        e[3] = Uy[0] + Ux[1]; e[4] = Uz[0] + Ux[2]; e[5] = Uz[1] + Uy[2];
    }
    double *eta = b->eta, *deta = b->deta;
    for (int c = 0; c < 6; c++) {
        size_t q = (size_t)(4 * c) * NP + p;
        out[c * st] = out[c * st] - (eta[q] + eta[q + NP] + eta[q + 2 * NP] + eta[q + 3 * NP]);
    }
    double tr = e[0] + e[1] + e[2];
    for (int l = 0; l < 4; l++) {
        double w = b->an_w[l], tau = b->an_tau[l];
        for (int c = 0; c < 3; c++) {
            size_t q = (size_t)(4 * c + l) * NP + p;
            deta[q] = deta[q] + (((w * 2.0 * mu * Qs) * e[c] + (w * ((lam + 2.0 * mu) * Qp - 2.0 * mu * Qs)) * tr)
                                 - eta[q]) / tau;
        }
        for (int c = 3; c < 6; c++) {
            size_t q = (size_t)(4 * c + l) * NP + p;
            deta[q] = deta[q] + (w * mu * Qs * e[c] - eta[q]) / tau;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* upwind and upwind-DRP operators (tables in wql_sbp_ops.h) */

const void *ora_sbp_table(int k) { return k >= 0 && k < WQL_SBP_NOPS ? (const void *)&WQL_SBP_OPS[k] : NULL; }

int ora_sbp_op(int fd_type, int order)
{
    if (fd_type == 0) return -1;
    const char *fam = fd_type == 1 ? "upwind" : fd_type == 2 ? "upwind_drp" : "";
    for (int k = 0; k < WQL_SBP_NOPS; k++)
        if (strcmp(WQL_SBP_OPS[k].family, fam) == 0 && WQL_SBP_OPS[k].order == order) return k;
    return -2;
}

/* taps of node pos on a line of n nodes (band routine rows, JJU_x<o>_upwind
 * JU..:815-16122): closure row at either face, else the printed interior row */
static int op_taps(const wql_sbp_half *h, int pos, int n, int *idx, double *c)
{
    int m = 0;
    if (pos < h->nl) {
        for (int t = 0; t < WQL_SBP_MAXW; t++)
            if (h->left[pos][t] != 0.0) { idx[m] = t; c[m] = h->left[pos][t]; m++; }
    } else if (pos >= n - h->nr) {
        int r = n - 1 - pos;
        for (int t = 0; t < WQL_SBP_MAXW; t++)
            if (h->right[r][t] != 0.0) { idx[m] = n - 1 - t; c[m] = h->right[r][t]; m++; }
    } else {
        for (int o = h->lo; o <= h->hi; o++) { idx[m] = pos + o; c[m] = h->band[o - h->lo]; m++; }
    }
    return m;
}

/* one direction of the band routine for metric component mc: velocities
 * X_x(centre) * sum c F (forward operator), stresses sum c X_x J F at the
 * taps (backward operator) */
static void op_axis(const ora_block *b, const wql_sbp_op *op, int dir, int mc, int i, int j, int k, double out[9])
{
    int n = dir == 0 ? b->nq : dir == 1 ? b->nr : b->ns;
    int pos = dir == 0 ? i : dir == 1 ? j : k;
    size_t NP = NPTS(b);
    const double *F = b->F, *G = b->metric;
    for (int grp = 0; grp < 2; grp++) {
        int idx[WQL_SBP_MAXW]; double c[WQL_SBP_MAXW];
        int m = op_taps(grp == 0 ? &op->v[dir] : &op->s[dir], pos, n, idx, c);
        for (int f = grp == 0 ? 0 : 3; f < (grp == 0 ? 3 : 9); f++) {
            double s = 0.0;
            for (int t = 0; t < m; t++) {
                int ii = i, jj = j, kk = k;
                if (dir == 0) ii = idx[t]; else if (dir == 1) jj = idx[t]; else kk = idx[t];
                size_t p = IDX(b, ii, jj, kk);
                double v = grp == 0 ? c[t] * F[f * NP + p] : c[t] * G[mc * NP + p] * G[9 * NP + p] * F[f * NP + p];
                s = (t == 0) ? v : s + v;
            }
            out[f] = grp == 0 ? G[mc * NP + IDX(b, i, j, k)] * s : s;
        }
    }
}

/* RHS_near_boundaries with the operator's band routine
 * (RHS_Interior.f90:144-1506): Ju = (1/hq) Jq + (1/hr) Jr + (1/hs) Js */
static void op_band(const ora_block *b, const wql_sbp_op *op, int fam, int i, int j, int k, double U[9])
{
    double hq = 1.0 / (double)(b->nq - 1), hr = 1.0 / (double)(b->nr - 1), hs = 1.0 / (double)(b->ns - 1);
    double a[9], bb[9], cc[9];
    op_axis(b, op, 0, 3 * fam + 0, i, j, k, a);
    op_axis(b, op, 1, 3 * fam + 1, i, j, k, bb);
    op_axis(b, op, 2, 3 * fam + 2, i, j, k, cc);
    for (int f = 0; f < 9; f++) U[f] = (1.0 / hq) * a[f] + (1.0 / hr) * bb[f] + (1.0 / hs) * cc[f];
}

/* interior routine JJU_x<o>_interior_upwind[_drp] (JU..:16451-20922): one
 * loop over the stencil taps, the three directions added per tap, with the
 * exact rational coefficients over h; the 'cartesian' form takes the
 * centre's diagonal metric and no J */
static void op_interior(const ora_block *b, const wql_sbp_op *op, int i, int j, int k, double DFx[9], double DFy[9],
                        double DFz[9])
{
    size_t NP = NPTS(b);
    const double *F = b->F, *G = b->metric, *J = G + 9 * NP;
    double h[3] = {1.0 / (double)(b->nq - 1), 1.0 / (double)(b->nr - 1), 1.0 / (double)(b->ns - 1)};
    size_t p = IDX(b, i, j, k);
    double *D[3] = {DFx, DFy, DFz};
    for (int f = 0; f < 9; f++) DFx[f] = DFy[f] = DFz[f] = 0.0;
    int nt = op->v[0].hi - op->v[0].lo + 1;
    for (int t = 0; t < nt; t++) {
        size_t pv[3], ps[3];
        double cv[3], cs[3];
        for (int d = 0; d < 3; d++) {
            int ov = op->v[d].lo + t, os = op->s[d].lo + t;
            pv[d] = IDX(b, i + (d == 0) * ov, j + (d == 1) * ov, k + (d == 2) * ov);
            ps[d] = IDX(b, i + (d == 0) * os, j + (d == 1) * os, k + (d == 2) * os);
            cv[d] = op->v[d].exact[t] / h[d];
            cs[d] = op->s[d].exact[t] / h[d];
        }
        for (int fam = 0; fam < 3; fam++) {
            double *Df = D[fam];
            const double *g0 = G + (3 * fam + 0) * NP, *g1 = G + (3 * fam + 1) * NP, *g2 = G + (3 * fam + 2) * NP;
            if (b->cartesian) {
                const double *gd = G + (3 * fam + fam) * NP;
                for (int f = 0; f < 3; f++) Df[f] = Df[f] + gd[p] * cv[fam] * F[f * NP + pv[fam]];
                for (int f = 3; f < 9; f++) Df[f] = Df[f] + gd[p] * cs[fam] * F[f * NP + ps[fam]];
            } else {
                for (int f = 0; f < 3; f++)
                    Df[f] = Df[f] + g0[p] * cv[0] * F[f * NP + pv[0]] + g1[p] * cv[1] * F[f * NP + pv[1]]
                                  + g2[p] * cv[2] * F[f * NP + pv[2]];
                for (int f = 3; f < 9; f++)
                    Df[f] = Df[f] + cs[0] * F[f * NP + ps[0]] * g0[ps[0]] * J[ps[0]]
                                  + cs[1] * F[f * NP + ps[1]] * g1[ps[1]] * J[ps[1]]
                                  + cs[2] * F[f * NP + ps[2]] * g2[ps[2]] * J[ps[2]];
            }
        }
    }
}

static void set_rates_op(ora_block *b)
{
    const wql_sbp_op *op = &WQL_SBP_OPS[b->sbp];
    size_t NP = NPTS(b);
    int nx = b->nq, ny = b->nr, nz = b->ns;
    /* band width per order (RHS_Interior.f90:144-1506), widened to npml on a PML side */
    int w = op->band;
    int ix = b->pml[0] && b->npml > w ? b->npml : w, fx = b->pml[1] && b->npml > w ? b->npml : w;
    int iy = b->pml[2] && b->npml > w ? b->npml : w, fy = b->pml[3] && b->npml > w ? b->npml : w;
    int iz = b->pml[4] && b->npml > w ? b->npml : w, fz = b->pml[5] && b->npml > w ? b->npml : w;
    const double *G = b->metric, *M = b->M;
    double *DF = b->DF;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int x = i + 1, y = j + 1, z = k + 1;
                size_t p = IDX(b, i, j, k);
                int interior = (iz + 1 <= z && z <= nz - fz) && (iy + 1 <= y && y <= ny - fy) &&
                               (ix + 1 <= x && x <= nx - fx);
                if (interior) {
                    double DFx[9], DFy[9], DFz[9];
                    op_interior(b, op, i, j, k, DFx, DFy, DFz);
                    double a1 = b->cartesian ? 1.0 / M[2 * NP + p] : 1.0 / (M[2 * NP + p] * G[9 * NP + p]);
                    double a2 = 2.0 * M[NP + p] + M[p], a3 = M[p], a4 = M[NP + p];
                    DF[0 * NP + p] = DF[0 * NP + p] + a1 * (DFx[3] + DFy[6] + DFz[7]);
                    DF[1 * NP + p] = DF[1 * NP + p] + a1 * (DFx[6] + DFy[4] + DFz[8]);
                    DF[2 * NP + p] = DF[2 * NP + p] + a1 * (DFx[7] + DFy[8] + DFz[5]);
                    DF[3 * NP + p] = DF[3 * NP + p] + a2 * DFx[0] + a3 * (DFy[1] + DFz[2]);
                    DF[4 * NP + p] = DF[4 * NP + p] + a2 * DFy[1] + a3 * (DFx[0] + DFz[2]);
                    DF[5 * NP + p] = DF[5 * NP + p] + a2 * DFz[2] + a3 * (DFx[0] + DFy[1]);
                    DF[6 * NP + p] = DF[6 * NP + p] + a4 * (DFy[0] + DFx[1]);
                    DF[7 * NP + p] = DF[7 * NP + p] + a4 * (DFz[0] + DFx[2]);
                    DF[8 * NP + p] = DF[8 * NP + p] + a4 * (DFz[1] + DFy[2]);
                    if (b->anelastic) anelastic_point(b, i, j, k, DFx, DFy, DFz, DF + 3 * NP + p, NP, 0);
                } else {
                    double Ux[9], Uy[9], Uz[9], DU[9];
                    op_band(b, op, 0, i, j, k, Ux);
                    op_band(b, op, 1, i, j, k, Uy);
                    op_band(b, op, 2, i, j, k, Uz);
                    double rhoJ_inv = 1.0 / (M[2 * NP + p] * G[9 * NP + p]);
                    double l2m = M[p] + 2.0 * M[NP + p];
                    double lam = M[p], mu = M[NP + p];
                    DU[0] = (Ux[3] + Uy[6] + Uz[7]) * rhoJ_inv;
                    DU[1] = (Ux[6] + Uy[4] + Uz[8]) * rhoJ_inv;
                    DU[2] = (Ux[7] + Uy[8] + Uz[5]) * rhoJ_inv;
                    DU[3] = l2m * Ux[0] + lam * (Uy[1] + Uz[2]);
                    DU[4] = l2m * Uy[1] + lam * (Ux[0] + Uz[2]);
                    DU[5] = l2m * Uz[2] + lam * (Ux[0] + Uy[1]);
                    DU[6] = mu * (Uy[0] + Ux[1]);
                    DU[7] = mu * (Uz[0] + Ux[2]);
                    DU[8] = mu * (Uz[1] + Uy[2]);
                    if (b->anelastic) anelastic_point(b, i, j, k, Ux, Uy, Uz, DU + 3, 1, 1);
                    pml_point(b, DU, Ux, Uy, Uz, i, j, k, rhoJ_inv, l2m);
                    for (int f = 0; f < 9; f++) DF[f * NP + p] = DF[f * NP + p] + DU[f];
                }
            }
}

void ora_set_rates(ora_block *b)
{
    if (b->sbp >= 0) {
        set_rates_op(b);
        return;
    }
    size_t NP = NPTS(b);
    int nx = b->nq, ny = b->nr, nz = b->ns;
    double hx = 1.0 / (double)(nx - 1), hy = 1.0 / (double)(ny - 1), hz = 1.0 / (double)(nz - 1);
    int ix = 6, iy = 6, iz = 6, fx = 6, fy = 6, fz = 6;
    /* band width: RHS_Interior.f90:196-211, JU..:16211-16226 */
    if (b->pml[4]) iz = b->npml > 6 ? b->npml : 6;
    if (b->pml[5]) fz = b->npml > 6 ? b->npml : 6;
    if (b->pml[2]) iy = b->npml > 6 ? b->npml : 6;
    if (b->pml[3]) fy = b->npml > 6 ? b->npml : 6;
    if (b->pml[0]) ix = b->npml > 6 ? b->npml : 6;
    if (b->pml[1]) fx = b->npml > 6 ? b->npml : 6;
    const double c3[3] = {0.75, -0.15, 1.0 / 60.0};
    double cx[3], cy[3], cz[3];
    for (int s = 0; s < 3; s++) { cx[s] = c3[s] / hx; cy[s] = c3[s] / hy; cz[s] = c3[s] / hz; }
    const double *F = b->F, *G = b->metric, *M = b->M;
    double *DF = b->DF;
    /* every node writes only its own rates (and PML auxiliary rates) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int x = i + 1, y = j + 1, z = k + 1;
                size_t p = IDX(b, i, j, k);
                int interior = (iz + 1 <= z && z <= nz - fz) && (iy + 1 <= y && y <= ny - fy) &&
                               (ix + 1 <= x && x <= nx - fx);
                if (interior) {
                    /* JJU_x6_interior JU..:16232-16314 (curvilinear), 16359-16403 (cartesian) */
                    double DFx[9] = {0}, DFy[9] = {0}, DFz[9] = {0};
                    for (int s = 1; s <= 3; s++) {
                        size_t xp = IDX(b, i + s, j, k), xm = IDX(b, i - s, j, k);
                        size_t yp = IDX(b, i, j + s, k), ym = IDX(b, i, j - s, k);
                        size_t zp = IDX(b, i, j, k + s), zm = IDX(b, i, j, k - s);
                        if (b->cartesian) {
                            for (int f = 0; f < 9; f++) {
                                DFx[f] = DFx[f] + G[0 * NP + p] * cx[s - 1] * (F[f * NP + xp] - F[f * NP + xm]);
                                DFy[f] = DFy[f] + G[4 * NP + p] * cy[s - 1] * (F[f * NP + yp] - F[f * NP + ym]);
                                DFz[f] = DFz[f] + G[8 * NP + p] * cz[s - 1] * (F[f * NP + zp] - F[f * NP + zm]);
                            }
                        } else {
                            for (int fam = 0; fam < 3; fam++) {
                                double *D = fam == 0 ? DFx : fam == 1 ? DFy : DFz;
                                const double *g0 = G + (3 * fam + 0) * NP, *g1 = G + (3 * fam + 1) * NP,
                                             *g2 = G + (3 * fam + 2) * NP, *J = G + 9 * NP;
                                for (int f = 0; f < 3; f++)
                                    D[f] = D[f] + g0[p] * cx[s - 1] * (F[f * NP + xp] - F[f * NP + xm])
                                                + g1[p] * cy[s - 1] * (F[f * NP + yp] - F[f * NP + ym])
                                                + g2[p] * cz[s - 1] * (F[f * NP + zp] - F[f * NP + zm]);
                                for (int f = 3; f < 9; f++)
                                    D[f] = D[f]
                                         + cx[s - 1] * (F[f * NP + xp] * (J[xp] * g0[xp]) - F[f * NP + xm] * (J[xm] * g0[xm]))
                                         + cy[s - 1] * (F[f * NP + yp] * (J[yp] * g1[yp]) - F[f * NP + ym] * (J[ym] * g1[ym]))
                                         + cz[s - 1] * (F[f * NP + zp] * (J[zp] * g2[zp]) - F[f * NP + zm] * (J[zm] * g2[zm]));
                            }
                        }
                    }
                    double a1 = b->cartesian ? 1.0 / M[2 * NP + p] : 1.0 / (M[2 * NP + p] * G[9 * NP + p]);
                    double a2 = 2.0 * M[NP + p] + M[p], a3 = M[p], a4 = M[NP + p];
                    DF[0 * NP + p] = DF[0 * NP + p] + a1 * (DFx[3] + DFy[6] + DFz[7]);
                    DF[1 * NP + p] = DF[1 * NP + p] + a1 * (DFx[6] + DFy[4] + DFz[8]);
                    DF[2 * NP + p] = DF[2 * NP + p] + a1 * (DFx[7] + DFy[8] + DFz[5]);
                    DF[3 * NP + p] = DF[3 * NP + p] + a2 * DFx[0] + a3 * (DFy[1] + DFz[2]);
                    DF[4 * NP + p] = DF[4 * NP + p] + a2 * DFy[1] + a3 * (DFx[0] + DFz[2]);
                    DF[5 * NP + p] = DF[5 * NP + p] + a2 * DFz[2] + a3 * (DFx[0] + DFy[1]);
                    DF[6 * NP + p] = DF[6 * NP + p] + a4 * (DFy[0] + DFx[1]);
                    DF[7 * NP + p] = DF[7 * NP + p] + a4 * (DFz[0] + DFx[2]);
                    DF[8 * NP + p] = DF[8 * NP + p] + a4 * (DFz[1] + DFy[2]);
                    if (b->anelastic) anelastic_point(b, i, j, k, DFx, DFy, DFz, DF + 3 * NP + p, NP, 0);
                } else {
                    /* RHS_near_boundaries RHS_Interior.f90:221-268 */
                    double Ux[9], Uy[9], Uz[9], DU[9];
                    jju6(b, 0, i, j, k, Ux);
                    jju6(b, 1, i, j, k, Uy);
                    jju6(b, 2, i, j, k, Uz);
                    double rhoJ_inv = 1.0 / (M[2 * NP + p] * G[9 * NP + p]);
                    double l2m = M[p] + 2.0 * M[NP + p];
                    double lam = M[p], mu = M[NP + p];
                    DU[0] = (Ux[3] + Uy[6] + Uz[7]) * rhoJ_inv;
                    DU[1] = (Ux[6] + Uy[4] + Uz[8]) * rhoJ_inv;
                    DU[2] = (Ux[7] + Uy[8] + Uz[5]) * rhoJ_inv;
                    DU[3] = l2m * Ux[0] + lam * (Uy[1] + Uz[2]);
                    DU[4] = l2m * Uy[1] + lam * (Ux[0] + Uz[2]);
                    DU[5] = l2m * Uz[2] + lam * (Ux[0] + Uy[1]);
                    DU[6] = mu * (Uy[0] + Ux[1]);
                    DU[7] = mu * (Uz[0] + Ux[2]);
                    DU[8] = mu * (Uz[1] + Uy[2]);
                    if (b->anelastic) anelastic_point(b, i, j, k, Ux, Uy, Uz, DU + 3, 1, 1);
                    pml_point(b, DU, Ux, Uy, Uz, i, j, k, rhoJ_inv, l2m);
                    for (int f = 0; f < 9; f++) DF[f * NP + p] = DF[f * NP + p] + DU[f];
                }
            }
}

/* ------------------------------------------------------------------------ */
/* outer boundaries: BoundaryConditions.f90:50-775 (BC_Lx..BC_Rz), BCm/BCp
 * 777-918, structure RHS_Interior.f90:2203-2223, application 1516-1703 */
static void bc_point(double ubc[9], const double u[9], double rho, double mu, double lambda,
                     const double l[3], const double m[3], const double n[3], double norm,
                     double Zs, double Zp, double r, int plus)
{
    double T[3];
    T[0] = n[0] * u[3] + n[1] * u[6] + n[2] * u[7];
    T[1] = n[0] * u[6] + n[1] * u[4] + n[2] * u[8];
    T[2] = n[0] * u[7] + n[1] * u[8] + n[2] * u[5];
    double vl = l[0] * u[0] + l[1] * u[1] + l[2] * u[2];
    double vm = m[0] * u[0] + m[1] * u[1] + m[2] * u[2];
    double vn = n[0] * u[0] + n[1] * u[1] + n[2] * u[2];
    double Tl = l[0] * T[0] + l[1] * T[1] + l[2] * T[2];
    double Tm = m[0] * T[0] + m[1] * T[1] + m[2] * T[2];
    double Tn = n[0] * T[0] + n[1] * T[1] + n[2] * T[2];
    double pl, pm, pn, ql, qm, qn;
    if (plus) {
        pl = 0.5 * ((1.0 - r) * Zs * vl + (1.0 + r) * Tl);
        pm = 0.5 * ((1.0 - r) * Zs * vm + (1.0 + r) * Tm);
        pn = 0.5 * ((1.0 - r) * Zp * vn + (1.0 + r) * Tn);
        ql = 0.5 * ((1.0 - r) * vl + (1.0 + r) / Zs * Tl);
        qm = 0.5 * ((1.0 - r) * vm + (1.0 + r) / Zs * Tm);
        qn = 0.5 * ((1.0 - r) * vn + (1.0 + r) / Zp * Tn);
    } else {
        pl = 0.5 * ((1.0 - r) * Zs * vl - (1.0 + r) * Tl);
        pm = 0.5 * ((1.0 - r) * Zs * vm - (1.0 + r) * Tm);
        pn = 0.5 * ((1.0 - r) * Zp * vn - (1.0 + r) * Tn);
        ql = 0.5 * ((1.0 - r) * vl - (1.0 + r) / Zs * Tl);
        qm = 0.5 * ((1.0 - r) * vm - (1.0 + r) / Zs * Tm);
        qn = 0.5 * ((1.0 - r) * vn - (1.0 + r) / Zp * Tn);
    }
    double Fx = l[0] * pl + m[0] * pm + n[0] * pn;
    double Fy = l[1] * pl + m[1] * pm + n[1] * pn;
    double Fz = l[2] * pl + m[2] * pm + n[2] * pn;
    double F_x = l[0] * ql + m[0] * qm + n[0] * qn;
    double F_y = l[1] * ql + m[1] * qm + n[1] * qn;
    double F_z = l[2] * ql + m[2] * qm + n[2] * qn;
    double sg = plus ? 1.0 : -1.0;
    ubc[0] = norm * Fx;
    ubc[1] = norm * Fy;
    ubc[2] = norm * Fz;
    if (plus) {
        ubc[3] = norm * n[0] * F_x;
        ubc[4] = norm * n[1] * F_y;
        ubc[5] = norm * n[2] * F_z;
        ubc[6] = norm * (n[1] * F_x + n[0] * F_y);
        ubc[7] = norm * (n[2] * F_x + n[0] * F_z);
        ubc[8] = norm * (n[2] * F_y + n[1] * F_z);
    } else {
        ubc[3] = -norm * n[0] * F_x;
        ubc[4] = -norm * n[1] * F_y;
        ubc[5] = -norm * n[2] * F_z;
        ubc[6] = -norm * (n[1] * F_x + n[0] * F_y);
        ubc[7] = -norm * (n[2] * F_x + n[0] * F_z);
        ubc[8] = -norm * (n[2] * F_y + n[1] * F_z);
    }
    (void)sg;
}

static void structure(double Ux[9], const double Uq[9], double rho, double mu, double lam)
{
    Ux[0] = 1.0 / rho * Uq[0];
    Ux[1] = 1.0 / rho * Uq[1];
    Ux[2] = 1.0 / rho * Uq[2];
    Ux[3] = (2.0 * mu + lam) * Uq[3] + lam * (Uq[4] + Uq[5]);
    Ux[4] = (2.0 * mu + lam) * Uq[4] + lam * (Uq[3] + Uq[5]);
    Ux[5] = (2.0 * mu + lam) * Uq[5] + lam * (Uq[3] + Uq[4]);
    Ux[6] = mu * Uq[6];
    Ux[7] = mu * Uq[7];
    Ux[8] = mu * Uq[8];
}

#define TAU0 (1.0 / (0.315949074074074)) /* block.f90:191 */
/* SAT penalty of the block's operator (block.f90:191-255) */
static double tau0_of(const ora_block *b) { return b->sbp < 0 ? TAU0 : WQL_SBP_OPS[b->sbp].tau0; }

void ora_boundary(ora_block *b)
{
    size_t NP = NPTS(b);
    for (int side = 0; side < 6; side++) {
        int code = b->bc[side];
        if (code <= 0) continue;
        int dir = side / 2, plus = side % 2;
        int n = dir == 0 ? b->nq : dir == 1 ? b->nr : b->ns;
        double h = 1.0 / (double)(n - 1);
        double r = code == 1 ? 0.0 : code == 2 ? 1.0 : -1.0;
        int na = dir == 0 ? b->nr : b->nq;
        int nb2 = dir == 2 ? b->nr : b->ns;
        size_t FP = face_pts(b, side);
        const double *nm = b->nml[side];
        for (int kk = 0; kk < nb2; kk++)
            for (int jj = 0; jj < na; jj++) {
                int i, j, k;
                if (dir == 0) { i = plus ? b->nq - 1 : 0; j = jj; k = kk; }
                else if (dir == 1) { j = plus ? b->nr - 1 : 0; i = jj; k = kk; }
                else { k = plus ? b->ns - 1 : 0; i = jj; j = kk; }
                size_t p = IDX(b, i, j, k), fp = (size_t)jj + (size_t)na * kk;
                double rho = b->M[2 * NP + p], mu = b->M[NP + p], lambda = b->M[p];
                double c_p = sqrt((2.0 * mu + lambda) / rho), c_s = sqrt(mu / rho);
                double Zp = rho * c_p, Zs = rho * c_s;
                double ax = b->metric[(0 + dir) * NP + p], ay = b->metric[(3 + dir) * NP + p],
                       az = b->metric[(6 + dir) * NP + p];
                double norm = sqrt(ax * ax + ay * ay + az * az);
                double nn[3], mm[3], ll[3], u[9], ubc[9], Ux[9];
                for (int t = 0; t < 3; t++) {
                    nn[t] = nm[(0 + t) * FP + fp]; mm[t] = nm[(3 + t) * FP + fp]; ll[t] = nm[(6 + t) * FP + fp];
                }
                for (int f = 0; f < 9; f++) u[f] = b->F[f * NP + p] - 0.0 * 0.0;
                bc_point(ubc, u, rho, mu, lambda, ll, mm, nn, norm, Zs, Zp, r, plus);
                structure(Ux, ubc, rho, mu, lambda);
                for (int f = 0; f < 9; f++) b->DF[f * NP + p] = b->DF[f * NP + p] - tau0_of(b) / h * Ux[f];
                if (b->pml[side]) {
                    long s = slab_index(b, side, i, j, k);
                    size_t SP = ora_slab_points(b, side);
                    for (int f = 0; f < 9; f++)
                        b->DQ[side][f * SP + s] = b->DQ[side][f * SP + s] - tau0_of(b) / h * ubc[f];
                }
            }
    }
}

/* ------------------------------------------------------------------------ */
/* fault: prestress, strength, friction laws (Interface_Condition.f90) */

static double boxcar(double x, double w) /* 2438-2456 */
{
    const double tol = 100.0 * 2.220446049250313e-16;
    if (-w + tol < x && x < w - tol) return 1.0;
    if (fabs(-w - x) <= tol || fabs(x - w) <= tol) return 0.5;
    return 0.0;
}

static double smooth_boxcar(double x, double w, double wr) /* 2459-2476 */
{
    if (-w <= x && x <= w) return 1.0;
    if (fabs(x) > w && fabs(x) < w + wr) return 0.5 * (1.0 + tanh(wr / (fabs(x) - w - wr) + wr / (fabs(x) - w)));
    return 0.0;
}

/* TPV26/27 background stress tensor at depth y (initial_stress_condition.f90:26-76) */
static void tpv26_stress(double y, double s0[6])
{
    double Omega = 0.0;
    if (y <= 15.0) Omega = 1.0;
    else if (y >= 15.0 && y <= 20.0) Omega = (20.0 - y) / 5.0;
    else if (y >= 20.0) Omega = 0.0;
    double b11 = 0.926793, b33 = 1.073206, b13 = -0.169029;
    double Pf = 9.8 * y;
    double sigma22 = -2.67 * 9.8 * y;
    s0[0] = Omega * (b33 * (sigma22 + Pf) - Pf) + (1.0 - Omega) * sigma22 + Pf;
    s0[1] = sigma22 + Pf;
    s0[2] = Omega * (b11 * (sigma22 + Pf) - Pf) + (1.0 - Omega) * sigma22 + Pf;
    s0[3] = 0.0;
    s0[4] = Omega * (b13 * (sigma22 + Pf));
    s0[5] = 0.0;
}

/* prestress 1802-2435: TPV5 (1822-1830) as stress_mode 0 with patches,
 * TPV28 (2049-2079) / TPV101-102 (2324-2385) / 'rate-weakening' (2387-2420)
 * as stress_mode 1, TPV26/27 (2008-2048) as stress_mode 2 */
static void prestress(const ora_fault *f, const double x[3], double t, const double l[3],
                      const double m[3], const double n[3], double *Tn, double *Tm, double *Tl)
{
    if (f->stress_mode == 0) {
        *Tn = f->T0[0]; *Tm = f->T0[1]; *Tl = f->T0[2];
        for (int k = 0; k < f->npatch; k++) {
            const double *P = f->patch[k];
            *Tl = *Tl + P[4] * boxcar(x[1] - P[0], P[2]) * boxcar(x[2] - P[1], P[3]);
        }
    } else {
        double st[6];
        if (f->stress_mode == 2) tpv26_stress(x[1], st);
        else for (int c = 0; c < 6; c++) st[c] = f->s0[c];
        const double *s = st;
        double Tx = n[0] * s[0] + n[1] * s[3] + n[2] * s[4];
        double Ty = n[0] * s[3] + n[1] * s[1] + n[2] * s[5];
        double Tz = n[0] * s[4] + n[1] * s[5] + n[2] * s[2];
        *Tn = n[0] * Tx + n[1] * Ty + n[2] * Tz;
        *Tm = m[0] * Tx + m[1] * Ty + m[2] * Tz;
        *Tl = l[0] * Tx + l[1] * Ty + l[2] * Tz;
    }
    if (f->nuc == 1) { /* TPV28 2069-2077 */
        double r = sqrt(pow(x[1] - f->nucp[0], 2) + pow(x[2] - f->nucp[1], 2));
        double tn;
        if (r <= f->nucp[2]) tn = f->nucp[4];
        else if (r >= f->nucp[2] && r <= f->nucp[2] + f->nucp[3])
            tn = 0.5 * f->nucp[4] * (1.0 + cos(PI * (r - f->nucp[2]) / f->nucp[3]));
        else tn = 0.0;
        *Tl = *Tl + tn;
    } else if (f->nuc == 2) { /* TPV101 2341-2351 */
        double r = sqrt(pow(x[1] - f->nucp[0], 2) + pow(x[2] - f->nucp[1], 2));
        double R = f->nucp[2], T = f->nucp[4];
        double Fv = 0.0, Gv = 0.0;
        if (r < R) Fv = exp(pow(r, 2) / (pow(r, 2) - R * R));
        if (t > 0.0 && t < T) Gv = exp(pow(t - T, 2) / (t * (t - 2.0 * T)));
        if (t >= T) Gv = 1.0;
        *Tl = *Tl + f->nucp[3] * Fv * Gv;
    } else if (f->nuc == 3) { /* 'rate-weakening' 2405-2420 */
        double tau_pulse = f->nucp[3] * *Tn;
        double r0 = f->nucp[2];
        double r = sqrt(pow(x[1] - f->nucp[0], 2) + pow(x[2] - f->nucp[1], 2));
        double Fv = exp(-(pow(r / r0, 2)) / 2.0);
        double Gv = 0.0;
        if (t > 0.0) Gv = 1.0;
        *Tl = *Tl + 2.0 * tau_pulse * Fv * Gv;
    }
}

/* Tau_strength TPV5 1283-1299 / TPV28 1459-1474 (sw_law 0); TPV26/27
 * 1376-1458 with Frictional_Cohesion 2491-2530 and Time_of_Forced_Rupture
 * 2625-2660 (sw_law 1) */
static double tau_strength(const ora_fault *f, double sigma_n, double D, const double x[3], double t)
{
    if (f->sw_law == 1) {
        double fs = f->fs, fd = f->fd, Dc = f->Dc, t0 = f->frc[4], Vs = f->frc[3];
        double r = sqrt(pow(x[1] - f->frc[0], 2) + pow(x[2] - f->frc[1], 2));
        double depth = x[1], C0, tfr, f1, f2 = 0.0;
        if (depth <= f->coh[2]) C0 = f->coh[0] + f->coh[1] * (f->coh[2] - depth);
        else C0 = f->coh[0];
        double r_crit = f->frc[2];
        if (r < r_crit - 1.0e-9)
            tfr = r / (0.7 * Vs) + 0.081 * r_crit / (0.7 * Vs) * (1.0 / (1.0 - pow(r / r_crit, 2)) - 1.0);
        else
            tfr = 1.0e9;
        f1 = fmin(D / Dc, 1.0);
        if (t < tfr) f2 = 0.0;
        else if (t >= tfr && t < tfr + t0) f2 = (t - tfr) / t0;
        else if (t >= tfr + t0) f2 = 1.0;
        fs = fs + f->bar_big * (1.0 - boxcar(x[1], f->bar_wy) * boxcar(x[2], f->bar_wz));
        double fr = fs - (fs - fd) * fmax(f1, f2);
        return C0 + fr * sigma_n;
    }
    double fs = f->fs + f->bar_big * (1.0 - boxcar(x[1], f->bar_wy) * boxcar(x[2], f->bar_wz));
    double fr = fs - (fs - f->fd) * fmin(D, f->Dc) / f->Dc;
    return sigma_n * fr;
}

static void solve_sw(double *v1, double *v2, double *Vel, double *tau1, double *tau2,
                     double phi_1, double phi_2, double eta, double tau_str) /* 1147-1164 */
{
    double Phi = sqrt(phi_1 * phi_1 + phi_2 * phi_2);
    double V = (Phi - tau_str) / eta;
    *Vel = V;
    *v1 = phi_1 / (eta + tau_str / V);
    *v2 = phi_2 / (eta + tau_str / V);
    *tau1 = phi_1 - eta * *v1;
    *tau2 = phi_2 - eta * *v2;
}

static double rs_f(double V, double Phi, double eta, double sigma_n, double psi, double V0, double a)
{ /* 1262 */
    return V + a * sigma_n / eta * asinh(0.5 * V / V0 * exp(psi / a)) - Phi / eta;
}

static double regula_falsi(double V, double Phi, double eta, double sigma_n, double psi, double V0, double a)
{ /* 1170-1239 */
    double err = 1.0, tol = 1.0e-12;
    double Vl = 0.0, Vr = Phi / eta;
    int k = 1, maxit = 5000;
    double fv = rs_f(V, Phi, eta, sigma_n, psi, V0, a);
    double fl = rs_f(Vl, Phi, eta, sigma_n, psi, V0, a);
    double fr = rs_f(Vr, Phi, eta, sigma_n, psi, V0, a);
    if (fabs(fv) <= tol) return V;
    if (fabs(fl) <= tol) return Vl;
    if (fabs(fr) <= tol) return Vr;
    while (k <= maxit && err >= tol) {
        double V1 = V;
        if (fv * fl > tol) {
            Vl = V;
            V = Vr - (Vr - Vl) * fr / (fr - fl);
            fl = fv;
        } else if (fv * fr > tol) {
            Vr = V;
            V = Vr - (Vr - Vl) * fr / (fr - fl);
            fr = fv;
        }
        fv = rs_f(V, Phi, eta, sigma_n, psi, V0, a);
        err = fabs(V - V1);
        k = k + 1;
    }
    return V;
}

/* the library's fast slip-rate mode (rs_solver 1, no counterpart in the
 * reference): Newton's method on the same f(V) inside the bracket
 * [0, Phi/eta], retried in ln V and then bisected when a step leaves the
 * bracket; stops when a step moves V by at most 1e-13 V */
static double newton_rs(double V, double Phi, double eta, double sigma_n, double psi, double V0, double a)
{
    double lo = 0.0, hi = Phi / eta;
    if (!(hi > 0.0)) return 0.0;
    if (!(V > lo && V < hi)) V = 0.5 * hi;
    double e = exp(psi / a) / (2.0 * V0), c = a * sigma_n / eta, rhs = Phi / eta;
    for (int k = 0; k < 100; k++) {
        double z = V * e;
        double fv = V + c * asinh(z) - rhs;
        if (fv == 0.0) return V;
        if (fv < 0.0) lo = V;
        else hi = V;
        double df = 1.0 + c * e / sqrt(1.0 + z * z);
        double Vn = V - fv / df;
        if (!(Vn > lo && Vn < hi)) {
            Vn = V * exp(-fv / (V * df));
            if (!(Vn > lo && Vn < hi)) Vn = lo > 0.0 ? sqrt(lo * hi) : 0.5 * hi;
        }
        if (fabs(Vn - V) <= 1.0e-13 * Vn) return Vn;
        V = Vn;
    }
    return V;
}

static double rs_a(const ora_fault *f, const double x[3])
{ /* 386-400 */
    return f->rs_a + f->rs_da * (1.0 - smooth_boxcar(x[1], f->rs_w, f->rs_wr) * smooth_boxcar(x[2], f->rs_w, f->rs_wr));
}

/* hat variables for one fault node (Interface_RHS 62-452, Welded 516-637,
 * Slip_Weakening 745-916, rate_state 918-1067); u, v in (n,m,l) frame */
int ora_hat(const ora_fault *f, const double u[6], const double v[6], const double mat1[3],
            const double mat2[3], const double x[3], const double nrm[9], double slip,
            double psi, double t, double uo[6], double vo[6], double *Vel)
{
    double lambda1 = mat1[0], mu1 = mat1[1], rho1 = mat1[2];
    double lambda2 = mat2[0], mu2 = mat2[1], rho2 = mat2[2];
    double c1_p = sqrt((2.0 * mu1 + lambda1) / rho1), c1_s = sqrt(mu1 / rho1);
    double c2_p = sqrt((2.0 * mu2 + lambda2) / rho2), c2_s = sqrt(mu2 / rho2);
    double Z1_s = rho1 * c1_s, Z2_s = rho2 * c2_s, Z1_p = rho1 * c1_p, Z2_p = rho2 * c2_p;
    const double *n = nrm, *m = nrm + 3, *l = nrm + 6;
    double p_n = Z1_p * u[0] - u[3], p_m = Z1_s * u[1] - u[4], p_l = Z1_s * u[2] - u[5];
    double q_n = Z2_p * v[0] + v[3], q_m = Z2_s * v[1] + v[4], q_l = Z2_s * v[2] + v[5];
    double eta_s, eta_p;
    *Vel = 0.0;
    if (f->law == ORA_LOCKED) {
        eta_p = 1.0 / (1.0 / Z1_p + 1.0 / Z2_p);
        eta_s = 1.0 / (1.0 / Z1_s + 1.0 / Z2_s);
        double phi_n = eta_p * (q_n / Z2_p - p_n / Z1_p);
        double phi_m = eta_s * (q_m / Z2_s - p_m / Z1_s);
        double phi_l = eta_s * (q_l / Z2_s - p_l / Z1_s);
        uo[0] = (q_n - phi_n) / Z2_p; uo[3] = phi_n;
        vo[0] = (p_n + phi_n) / Z1_p; vo[3] = phi_n;
        uo[2] = (q_l - phi_l) / Z2_s; uo[5] = phi_l;
        vo[2] = (p_l + phi_l) / Z1_s; vo[5] = phi_l;
        uo[1] = (q_m - phi_m) / Z2_s; uo[4] = phi_m;
        vo[1] = (p_m + phi_m) / Z1_s; vo[4] = phi_m;
        return 0;
    }
    if (Z1_s <= 1.0e-9 || Z2_s <= 1.0e-9) return -1;
    eta_s = 1.0 / (1.0 / Z1_s + 1.0 / Z2_s);
    eta_p = 1.0 / (1.0 / Z1_p + 1.0 / Z2_p);
    double T0_n, T0_m, T0_l;
    prestress(f, x, t, l, m, n, &T0_n, &T0_m, &T0_l);
    double phi_n = eta_p * (q_n / Z2_p - p_n / Z1_p);
    double phi_m = eta_s * (q_m / Z2_s - p_m / Z1_s);
    double phi_l = eta_s * (q_l / Z2_s - p_l / Z1_s);
    double u_n_hat = (q_n - phi_n) / Z2_p, Tu_n_hat = phi_n;
    double v_n_hat = (p_n + phi_n) / Z1_p, Tv_n_hat = phi_n;
    double sigma_n = fmax(0.0, -(T0_n + phi_n));
    double vv_m, vv_l, tau_m, tau_l, Tu_m_hat, Tu_l_hat, Tv_m_hat, Tv_l_hat;
    if (f->law == ORA_SLIP_WEAKENING) {
        double tau_lock = sqrt(pow(T0_l + phi_l, 2) + pow(T0_m + phi_m, 2));
        double tau_str = tau_strength(f, sigma_n, slip, x, t);
        if (tau_lock > tau_str) {
            solve_sw(&vv_m, &vv_l, Vel, &tau_m, &tau_l, T0_m + phi_m, T0_l + phi_l, eta_s, tau_str);
            if (*Vel < 0.0) return -2;
            Tu_m_hat = tau_m - T0_m; Tu_l_hat = tau_l - T0_l;
            Tv_m_hat = tau_m - T0_m; Tv_l_hat = tau_l - T0_l;
        } else {
            Tu_m_hat = phi_m; Tu_l_hat = phi_l; Tv_m_hat = phi_m; Tv_l_hat = phi_l;
            vv_m = 0.0; vv_l = 0.0; *Vel = 0.0;
        }
    } else {
        /* rate_state_friction 1017-1043, solve 1120-1145 */
        double a = rs_a(f, x);
        double V0 = f->rs_V0;
        tau_m = 0.0; tau_l = 0.0;
        vv_m = v[1] - u[1]; vv_l = v[2] - u[2];
        double phi_1 = T0_m + phi_m, phi_2 = T0_l + phi_l;
        double Phi = sqrt(phi_1 * phi_1 + phi_2 * phi_2);
        double V = sqrt(vv_m * vv_m + vv_l * vv_l);
        if (V > Phi) V = 0.5 * Phi / eta_s;
        V = f->rs_solver == 1 ? newton_rs(V, Phi, eta_s, sigma_n, psi, V0, a)
                              : regula_falsi(V, Phi, eta_s, sigma_n, psi, V0, a);
        if (V < 0.0) return -2;
        *Vel = V;
        double fv = sigma_n * a * asinh(0.5 * V / V0 * exp(psi / a));
        vv_m = phi_1 / (eta_s + fv / V);
        vv_l = phi_2 / (eta_s + fv / V);
        tau_m = phi_1 - eta_s * vv_m;
        tau_l = phi_2 - eta_s * vv_l;
        Tu_m_hat = tau_m - T0_m; Tu_l_hat = tau_l - T0_l;
        Tv_m_hat = tau_m - T0_m; Tv_l_hat = tau_l - T0_l;
    }
    double v_m_hat = (Tu_m_hat + p_m) / (Z1_s) + vv_m;
    double v_l_hat = (Tu_l_hat + p_l) / (Z1_s) + vv_l;
    double u_m_hat = (q_m - Tv_m_hat) / (Z2_s) - vv_m;
    double u_l_hat = (q_l - Tv_l_hat) / (Z2_s) - vv_l;
    uo[0] = u_n_hat; vo[0] = v_n_hat;
    uo[1] = u_m_hat; uo[2] = u_l_hat;
    vo[1] = v_m_hat; vo[2] = v_l_hat;
    uo[3] = Tu_n_hat; vo[3] = Tv_n_hat;
    uo[4] = Tu_m_hat; uo[5] = Tu_l_hat;
    vo[4] = Tv_m_hat; vo[5] = Tv_l_hat;
    return 0;
}

/* Couple_Interface_x CouplingForcing.f90:9-467 for block ib (1: left block,
 * its face q = nq (ib==2 in the reference); 2: right block, face q = 1),
 * plus the SAT application of Impose_Interface_Condition RHS_Interior.f90:1750-1772 */
int ora_interface(ora_domain *d, int ib, double t, int stage)
{
    ora_block *B1 = &d->b[0], *B2 = &d->b[1];
    ora_block *B = (ib == 1) ? B1 : B2;
    ora_iface *I = &d->I[ib - 1];
    int side = (ib == 1) ? 1 : 0;            /* face of B carrying the fault */
    int xend = (ib == 1) ? B->nq - 1 : 0;
    int nr = B->nr, ns = B->ns;
    size_t FP = (size_t)nr * ns;
    size_t NP1 = NPTS(B1), NP2 = NPTS(B2), NP = NPTS(B);
    const double *nml = B->nml[side];
    double hx = 1.0 / (double)(B->nq - 1);
    int err = 0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < ns; k++)
        for (int j = 0; j < nr; j++) {
            size_t fp = (size_t)j + (size_t)nr * k;
            size_t pu = IDX(B1, B1->nq - 1, j, k), pv = IDX(B2, 0, j, k);
            double u[9], v[9];
            for (int f = 0; f < 9; f++) { u[f] = B1->F[f * NP1 + pu]; v[f] = B2->F[f * NP2 + pv]; }
            double nn[3], mm[3], ll[3];
            for (int c = 0; c < 3; c++) {
                nn[c] = nml[(0 + c) * FP + fp]; mm[c] = nml[(3 + c) * FP + fp]; ll[c] = nml[(6 + c) * FP + fp];
            }
            /* tractions + velocities 105-191, rotation 470-524 */
            double ur[6], vr[6];
            ur[0] = u[0]; ur[1] = u[1]; ur[2] = u[2];
            ur[3] = nn[0] * (u[3]) + nn[1] * (u[6]) + nn[2] * (u[7]);
            ur[4] = nn[0] * (u[6]) + nn[1] * (u[4]) + nn[2] * (u[8]);
            ur[5] = nn[0] * (u[7]) + nn[1] * (u[8]) + nn[2] * (u[5]);
            vr[0] = v[0]; vr[1] = v[1]; vr[2] = v[2];
            vr[3] = nn[0] * (v[3]) + nn[1] * (v[6]) + nn[2] * (v[7]);
            vr[4] = nn[0] * (v[6]) + nn[1] * (v[4]) + nn[2] * (v[8]);
            vr[5] = nn[0] * (v[7]) + nn[1] * (v[8]) + nn[2] * (v[5]);
            double uL[6], vL[6];
            for (int g = 0; g < 2; g++) {
                const double *src = g == 0 ? ur : vr;
                double *dst = g == 0 ? uL : vL;
                double a1 = src[0], a2 = src[1], a3 = src[2], T1 = src[3], T2 = src[4], T3 = src[5];
                dst[0] = nn[0] * a1 + nn[1] * a2 + nn[2] * a3;
                dst[1] = mm[0] * a1 + mm[1] * a2 + mm[2] * a3;
                dst[2] = ll[0] * a1 + ll[1] * a2 + ll[2] * a3;
                dst[3] = nn[0] * T1 + nn[1] * T2 + nn[2] * T3;
                dst[4] = mm[0] * T1 + mm[1] * T2 + mm[2] * T3;
                dst[5] = ll[0] * T1 + ll[1] * T2 + ll[2] * T3;
            }
            size_t pm1 = IDX(B1, B1->nq - 1, j, k), pm2 = IDX(B2, 0, j, k);
            double mat1[3] = {B1->M[pm1], B1->M[NP1 + pm1], B1->M[2 * NP1 + pm1]};
            double mat2[3] = {B2->M[pm2], B2->M[NP2 + pm2], B2->M[2 * NP2 + pm2]};
            size_t px = IDX(B, xend, j, k);
            double x[3] = {B->X[px], B->X[NP + px], B->X[2 * NP + px]};
            double frame[9] = {nn[0], nn[1], nn[2], mm[0], mm[1], mm[2], ll[0], ll[1], ll[2]};
            double uh[6], vh[6], Vel;
            double slip = I->S[3 * FP + fp];
            double psi = I->W[fp];
            int rc = ora_hat(&d->fault, uL, vL, mat1, mat2, x, frame, slip, psi, t, uh, vh, &Vel);
            if (rc) {
#pragma omp atomic write
                err = rc;
                continue;
            }
            double slip_rate = sqrt(pow(vh[1] - uh[1], 2) + pow(vh[2] - uh[2], 2)); /* 425, 470 */
            if (d->fault.law == ORA_RATE_STATE) { /* 427-450 */
                const ora_fault *f = &d->fault;
                double W = I->W[fp], L0 = f->rs_L, f0 = f->rs_f0, b = f->rs_b, V0 = f->rs_V0;
                if (f->rs_law == 0) { /* TPV101/102 431-436 */
                    I->DW[fp] = I->DW[fp] + b * V0 / L0 * exp(-(W - f0) / b) - Vel * b / L0;
                } else {
                    double a = rs_a(f, x);
                    double Fss = f0 - (b - a) * log(Vel / V0);
                    if (f->rs_law == 1) { /* 'rate-weakening' 438-444 */
                        if (f->rs_Vw > 0.0) {
                            double V_w = f->rs_Vw + f->rs_dVw * (1.0 - smooth_boxcar(x[1], f->rs_w, f->rs_wr)
                                                                       * smooth_boxcar(x[2], f->rs_w, f->rs_wr));
                            double q = Vel / V_w, q8 = (q * q) * (q * q) * ((q * q) * (q * q));
                            Fss = f->rs_fw + (Fss - f->rs_fw) / pow(1.0 + q8, 0.125);
                        }
                        I->DW[fp] = I->DW[fp] - Vel / L0 * (a * asinh(Vel / (2.0 * V0) * exp(W / a)) - Fss);
                    } else { /* 'KDN' 446-449 */
                        I->DW[fp] = I->DW[fp] - (1.0 / L0) * slip_rate * (asinh(Vel / (2.0 * V0) * exp(W / a)) - Fss);
                    }
                }
            }
            for (int c = 0; c < 3; c++) I->DS[c * FP + fp] = I->DS[c * FP + fp] + vh[c] - uh[c];
            I->DS[3 * FP + fp] = I->DS[3 * FP + fp] + slip_rate;
            if (stage == 1) { /* 480-509 and CouplingForcing.f90:213-228, 584-605 */
                double T0n = 0.0, T0m = 0.0, T0l = 0.0;
                if (d->fault.law != ORA_LOCKED)
                    prestress(&d->fault, x, t, ll, mm, nn, &T0n, &T0m, &T0l);
                for (int c = 0; c < 3; c++) {
                    I->Svel[c * FP + fp] = vh[c] - uh[c];
                    I->Uhat[c * FP + fp] = vh[c] - uh[c];
                    I->Vhat[c * FP + fp] = vh[c] - uh[c];
                }
                I->Uhat[3 * FP + fp] = uh[3] + T0n; I->Vhat[3 * FP + fp] = vh[3] + T0n;
                I->Uhat[4 * FP + fp] = uh[4] + T0m; I->Vhat[4 * FP + fp] = vh[4] + T0m;
                I->Uhat[5 * FP + fp] = uh[5] + T0l; I->Vhat[5 * FP + fp] = vh[5] + T0l;
                if (sqrt(pow(I->Svel[1 * FP + fp], 2) + pow(I->Svel[2 * FP + fp], 2)) > 1e-3)
                    I->trup[fp] = fmin(I->trup[fp], t);
            }
            /* rotate back 527-581 */
            double uX[6], vX[6], uH[6], vH[6];
            const double *srcs[4] = {uL, vL, uh, vh};
            double *dsts[4] = {uX, vX, uH, vH};
            for (int g = 0; g < 4; g++) {
                const double *s = srcs[g];
                double *o = dsts[g];
                for (int c = 0; c < 3; c++) {
                    o[c] = nn[c] * s[0] + mm[c] * s[1] + ll[c] * s[2];
                    o[3 + c] = nn[c] * s[3] + mm[c] * s[4] + ll[c] * s[5];
                }
            }
            /* SAT forcing 242-463 (metric of the local block for both sides) */
            size_t pe = IDX(B, xend, j, k);
            double q_x = B->metric[0 * NP + pe], q_y = B->metric[3 * NP + pe], q_z = B->metric[6 * NP + pe];
            double norm = sqrt(q_x * q_x + q_y * q_y + q_z * q_z);
            double n_1 = nn[0], n_2 = nn[1], n_3 = nn[2], m_1 = mm[0], m_2 = mm[1], m_3 = mm[2];
            double l_1 = ll[0], l_2 = ll[1], l_3 = ll[2];
            double bc[9];
            if (ib == 1) {
                double Zp = mat1[2] * sqrt((2.0 * mat1[1] + mat1[0]) / mat1[2]);
                double Zs = mat1[2] * sqrt(mat1[1] / mat1[2]);
                double m11 = n_1 * n_1 * Zp + n_2 * n_2 * Zs + n_3 * n_3 * Zs;
                double m12 = n_1 * m_1 * Zp + n_2 * m_2 * Zs + n_3 * m_3 * Zs;
                double m13 = n_1 * l_1 * Zp + n_2 * l_2 * Zs + n_3 * l_3 * Zs;
                double m21 = n_1 * m_1 * Zp + n_2 * m_2 * Zs + n_3 * m_3 * Zs;
                double m22 = m_1 * m_1 * Zp + m_2 * m_2 * Zs + m_3 * m_3 * Zs;
                double m23 = m_1 * l_1 * Zp + m_2 * l_2 * Zs + m_3 * l_3 * Zs;
                double m31 = l_1 * n_1 * Zp + l_2 * n_2 * Zs + l_3 * n_3 * Zs;
                double m32 = l_1 * m_1 * Zp + l_2 * m_2 * Zs + l_3 * m_3 * Zs;
                double m33 = l_1 * l_1 * Zp + l_2 * l_2 * Zs + l_3 * l_3 * Zs;
                double p11 = n_1 * n_1 * 1.0 / Zp + n_2 * n_2 * 1.0 / Zs + n_3 * n_3 * 1.0 / Zs;
                double p12 = n_1 * m_1 * 1.0 / Zp + n_2 * m_2 * 1.0 / Zs + n_3 * m_3 * 1.0 / Zs;
                double p13 = n_1 * l_1 * 1.0 / Zp + n_2 * l_2 * 1.0 / Zs + n_3 * l_3 * 1.0 / Zs;
                double p21 = n_1 * m_1 * 1.0 / Zp + n_2 * m_2 * 1.0 / Zs + n_3 * m_3 * 1.0 / Zs;
                double p22 = m_1 * m_1 * 1.0 / Zp + m_2 * m_2 * 1.0 / Zs + m_3 * m_3 * 1.0 / Zs;
                double p23 = m_1 * l_1 * 1.0 / Zp + m_2 * l_2 * 1.0 / Zs + m_3 * l_3 * 1.0 / Zs;
                double p31 = l_1 * n_1 * 1.0 / Zp + l_2 * n_2 * 1.0 / Zs + l_3 * n_3 * 1.0 / Zs;
                double p32 = l_1 * m_1 * 1.0 / Zp + l_2 * m_2 * 1.0 / Zs + l_3 * m_3 * 1.0 / Zs;
                double p33 = l_1 * l_1 * 1.0 / Zp + l_2 * l_2 * 1.0 / Zs + l_3 * l_3 * 1.0 / Zs;
                double p1h = 0.5 * (m11 * uH[0] + m12 * uH[1] + m13 * uH[2] + uH[3]);
                double p2h = 0.5 * (m21 * uH[0] + m22 * uH[1] + m23 * uH[2] + uH[4]);
                double p3h = 0.5 * (m31 * uH[0] + m32 * uH[1] + m33 * uH[2] + uH[5]);
                double p1 = 0.5 * (m11 * uX[0] + m12 * uX[1] + m13 * uX[2] + uX[3]);
                double p2 = 0.5 * (m21 * uX[0] + m22 * uX[1] + m23 * uX[2] + uX[4]);
                double p3 = 0.5 * (m31 * uX[0] + m32 * uX[1] + m33 * uX[2] + uX[5]);
                double lam = mat1[0], mu = mat1[1], rho = mat1[2];
                double e1 = p11 * (p1 - p1h) + p12 * (p2 - p2h) + p13 * (p3 - p3h);
                double e2 = p21 * (p1 - p1h) + p22 * (p2 - p2h) + p23 * (p3 - p3h);
                double e3 = p31 * (p1 - p1h) + p32 * (p2 - p2h) + p33 * (p3 - p3h);
                bc[0] = norm / rho * (p1 - p1h);
                bc[1] = norm / rho * (p2 - p2h);
                bc[2] = norm / rho * (p3 - p3h);
                bc[3] = (2.0 * mu + lam) * q_x * e1 + lam * q_y * e2 + lam * q_z * e3;
                bc[4] = (2.0 * mu + lam) * q_y * e2 + lam * q_z * e3 + lam * q_x * e1;
                bc[5] = (2.0 * mu + lam) * q_z * e3 + lam * q_y * e2 + lam * q_x * e1;
                bc[6] = mu * (q_y * e1 + q_x * e2);
                bc[7] = mu * (q_z * e1 + q_x * e3);
                bc[8] = mu * (q_z * e2 + q_y * e3);
            } else {
                double Zp = mat2[2] * sqrt((2.0 * mat2[1] + mat2[0]) / mat2[2]);
                double Zs = mat2[2] * sqrt(mat2[1] / mat2[2]);
                double n11 = n_1 * n_1 * Zp + n_2 * n_2 * Zs + n_3 * n_3 * Zs;
                double n12 = n_1 * m_1 * Zp + n_2 * m_2 * Zs + n_3 * m_3 * Zs;
                double n13 = n_1 * l_1 * Zp + n_2 * l_2 * Zs + n_3 * l_3 * Zs;
                double n21 = n_1 * m_1 * Zp + n_2 * m_2 * Zs + n_3 * m_3 * Zs;
                double n22 = m_1 * m_1 * Zp + m_2 * m_2 * Zs + m_3 * m_3 * Zs;
                double n23 = m_1 * l_1 * Zp + m_2 * l_2 * Zs + m_3 * l_3 * Zs;
                double n31 = l_1 * n_1 * Zp + l_2 * n_2 * Zs + l_3 * n_3 * Zs;
                double n32 = l_1 * m_1 * Zp + l_2 * m_2 * Zs + l_3 * m_3 * Zs;
                double n33 = l_1 * l_1 * Zp + l_2 * l_2 * Zs + l_3 * l_3 * Zs;
                double q11 = n_1 * n_1 * 1.0 / Zp + n_2 * n_2 * 1.0 / Zs + n_3 * n_3 * 1.0 / Zs;
                double q12 = n_1 * m_1 * 1.0 / Zp + n_2 * m_2 * 1.0 / Zs + n_3 * m_3 * 1.0 / Zs;
                double q13 = n_1 * l_1 * 1.0 / Zp + n_2 * l_2 * 1.0 / Zs + n_3 * l_3 * 1.0 / Zs;
                double q21 = n_1 * m_1 * 1.0 / Zp + n_2 * m_2 * 1.0 / Zs + n_3 * m_3 * 1.0 / Zs;
                double q22 = m_1 * m_1 * 1.0 / Zp + m_2 * m_2 * 1.0 / Zs + m_3 * m_3 * 1.0 / Zs;
                double q23 = m_1 * l_1 * 1.0 / Zp + m_2 * l_2 * 1.0 / Zs + m_3 * l_3 * 1.0 / Zs;
                double q31 = l_1 * n_1 * 1.0 / Zp + l_2 * n_2 * 1.0 / Zs + l_3 * n_3 * 1.0 / Zs;
                double q32 = l_1 * m_1 * 1.0 / Zp + l_2 * m_2 * 1.0 / Zs + l_3 * m_3 * 1.0 / Zs;
                double q33 = l_1 * l_1 * 1.0 / Zp + l_2 * l_2 * 1.0 / Zs + l_3 * l_3 * 1.0 / Zs;
                double q1h = 0.5 * (n11 * vH[0] + n12 * vH[1] + n13 * vH[2] - vH[3]);
                double q2h = 0.5 * (n21 * vH[0] + n22 * vH[1] + n23 * vH[2] - vH[4]);
                double q3h = 0.5 * (n31 * vH[0] + n32 * vH[1] + n33 * vH[2] - vH[5]);
                double q1 = 0.5 * (n11 * vX[0] + n12 * vX[1] + n13 * vX[2] - vX[3]);
                double q2 = 0.5 * (n21 * vX[0] + n22 * vX[1] + n23 * vX[2] - vX[4]);
                double q3 = 0.5 * (n31 * vX[0] + n32 * vX[1] + n33 * vX[2] - vX[5]);
                double lam = mat2[0], mu = mat2[1], rho = mat2[2];
                double e1 = q11 * (q1 - q1h) + q12 * (q2 - q2h) + q13 * (q3 - q3h);
                double e2 = q21 * (q1 - q1h) + q22 * (q2 - q2h) + q23 * (q3 - q3h);
                double e3 = q31 * (q1 - q1h) + q32 * (q2 - q2h) + q33 * (q3 - q3h);
                bc[0] = norm / rho * (q1 - q1h);
                bc[1] = norm / rho * (q2 - q2h);
                bc[2] = norm / rho * (q3 - q3h);
                bc[3] = -(2.0 * mu + lam) * q_x * e1 - lam * q_y * e2 - lam * q_z * e3;
                bc[4] = -(2.0 * mu + lam) * q_y * e2 - lam * q_z * e3 - lam * q_x * e1;
                bc[5] = -(2.0 * mu + lam) * q_z * e3 - lam * q_y * e2 - lam * q_x * e1;
                bc[6] = -mu * (q_y * e1 + q_x * e2);
                bc[7] = -mu * (q_z * e1 + q_x * e3);
                bc[8] = -mu * (q_z * e2 + q_y * e3);
            }
            for (int f = 0; f < 9; f++) B->DF[f * NP + pe] = B->DF[f * NP + pe] - tau0_of(B) / hx * bc[f];
        }
    return err;
}

/* ------------------------------------------------------------------------ */
/* plasticity: plastic.f90:11-151 (update_fields_plastic), 153-217
 * (plastic_flow2), 221-290 (plastic_flow), 293-352 (invariants/stress/yield);
 * initial stress: initial_stress_condition.f90:8-120 */
static void s0_at(const ora_domain *d, const double x[3], double s0[6])
{
    if (d->s0_kind == 1) tpv26_stress(x[1], s0); /* TPV26 26-50 */
    else for (int c = 0; c < 6; c++) s0[c] = d->fault.s0[c];
}

static void invariants(const double sa[6], double *tau, double *sigma)
{
    *sigma = (sa[0] + sa[1] + sa[2]) / 3.0;
    double J2 = (pow(sa[0] - sa[1], 2) + pow(sa[1] - sa[2], 2) + pow(sa[2] - sa[0], 2)) / 6.0 +
                pow(sa[3], 2) + pow(sa[4], 2) + pow(sa[5], 2);
    *tau = sqrt(J2);
}

static void plastic_update(ora_domain *d, ora_block *b)
{
    size_t NP = NPTS(b);
    int nx = b->nq, ny = b->nr, nz = b->ns;
    int ix = 1, fx = nx, iy = 1, fy = ny, iz = 1, fz = nz, np = b->npml;
    /* PML exclusion 67-113 */
    if (b->pml[4]) iz = np + 1;
    if (b->pml[5]) fz = nz - np;
    if (b->pml[2]) iy = np + 1;
    if (b->pml[3]) fy = ny - np;
    if (b->pml[0]) ix = np + 1;
    if (b->pml[1]) fx = nx - np;
    const double delta[6] = {1, 1, 1, 0, 0, 0};
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = iz - 1; k < fz; k++)
        for (int j = iy - 1; j < fy; j++)
            for (int i = ix - 1; i < fx; i++) {
                size_t p = IDX(b, i, j, k);
                double x[3] = {b->X[p], b->X[NP + p], b->X[2 * NP + p]};
                double s0[6], s[6], sa[6], sd[6], tau, sigma;
                s0_at(d, x, s0);
                for (int c = 0; c < 6; c++) { s[c] = b->F[(3 + c) * NP + p]; sa[c] = s[c] + s0[c]; }
                invariants(sa, &tau, &sigma);
                if (d->plastic == 1) {
                    double c = d->pl_c, nu = d->pl_nu, Tv = d->pl_Tv;
                    double Ys = fmax(0.0, c * cos(atan(nu)) - (sigma) * sin(atan(nu)));
                    double Yf = tau - Ys;
                    /* no accumulator in the reference's plastic_flow2 (plastic.f90:153-217);
                     * the build keeps one in P: lambda = dgamma / dt, gammap += dgamma with
                     * dgamma = (1 - r) tau / G, the deviator's shear-stress drop over G */
                    if (Yf <= 0.0) { if (b->P) b->P[p] = 0.0; continue; }
                    double r = exp(-d->dt / Tv) + (1.0 - exp(-d->dt / Tv)) * Ys / tau;
                    for (int c2 = 0; c2 < 6; c2++) sd[c2] = sa[c2] - sigma * delta[c2];
                    for (int c2 = 0; c2 < 6; c2++) sa[c2] = sigma * delta[c2] + r * sd[c2];
                    if (b->P) {
                        double G = b->M[NP + p], dg = (1.0 - r) * tau / G;
                        b->P[p] = dg / d->dt;
                        b->P[NP + p] = b->P[NP + p] + dg;
                    }
                } else {
                    double mu = d->pl_mu, beta = d->pl_beta, eta = d->pl_eta;
                    double G = b->M[NP + p], K = b->M[p] + 2.0 / 3.0 * b->M[NP + p];
                    double Y = tau + mu * sigma;
                    /* plastic_flow(s, gammap=P(:,:,:,2), lambda=P(:,:,:,1), ...) plastic.f90:135 */
                    if (Y <= 0.0) { if (b->P) b->P[p] = 0.0; continue; }
                    for (int c2 = 0; c2 < 6; c2++) sd[c2] = sa[c2] - sigma * delta[c2];
                    double lambda = (tau + mu * sigma) / (eta + d->dt * (G + mu * beta * K));
                    if (b->P) { b->P[p] = lambda; b->P[NP + p] = b->P[NP + p] + d->dt * lambda; }
                    double tau_n = tau - d->dt * lambda * G;
                    sigma = sigma - d->dt * lambda * beta * K;
                    for (int c2 = 0; c2 < 6; c2++) sd[c2] = sd[c2] * tau_n / (tau_n + d->dt * lambda * G);
                    for (int c2 = 0; c2 < 6; c2++) sa[c2] = sigma * delta[c2] + sd[c2];
                }
                for (int c2 = 0; c2 < 6; c2++) b->F[(3 + c2) * NP + p] = sa[c2] - s0[c2];
            }
}

void ora_plastic(ora_domain *d)
{ /* the return map of the last RK stage on both blocks (domain.f90:986-993) */
    if (!d->plastic) return;
    plastic_update(d, &d->b[0]);
    plastic_update(d, &d->b[1]);
}

void ora_initial_stress(const ora_domain *d, const double x[3], double s0[6])
{ /* background stress tensor of the plastic update at one point */
    s0_at(d, x, s0);
}

/* ------------------------------------------------------------------------ */
void ora_rk_coeffs(double *A, double *B, double *C)
{ /* time_step.f90:53-62 */
    A[0] = 0.0;
    A[1] = -567301805773.0 / 1357537059087.0;
    A[2] = -2404267990393.0 / 2016746695238.0;
    A[3] = -3550918686646.0 / 2091501179385.0;
    A[4] = -1275806237668.0 / 842570457699.0;
    B[0] = 1432997174477.0 / 9575080441755.0;
    B[1] = 5161836677717.0 / 13612068292357.0;
    B[2] = 1720146321549.0 / 2090206949498.0;
    B[3] = 3134564353537.0 / 4481467310338.0;
    B[4] = 2277821191437.0 / 14882151754819.0;
    C[0] = 0.0;
    C[1] = 1432997174477.0 / 9575080441755.0;
    C[2] = 2526269341429.0 / 6820363962896.0;
    C[3] = 2006345519317.0 / 3224310063776.0;
    C[4] = 2802321613138.0 / 2924317926251.0;
    C[5] = 1.0;
}

void ora_init_state(ora_domain *d)
{ /* Init_Vel_State 2722-2850: TPV101/102 branch (steady state, rs_init 0),
   * 'rate-weakening' branch (constant state, rs_init 1) */
    const ora_fault *f = &d->fault;
    if (f->law != ORA_RATE_STATE) return;
    for (int ib = 1; ib <= 2; ib++) {
        ora_block *B = &d->b[ib - 1];
        size_t NP = NPTS(B);
        double aa = ib == 1 ? -1.0 : 1.0;
        double vz = copysign(0.5 * f->rs_Vin, aa);
        for (size_t p = 0; p < NP; p++) B->F[2 * NP + p] = vz;
        int side = ib == 1 ? 1 : 0, xend = ib == 1 ? B->nq - 1 : 0;
        size_t FP = (size_t)B->nr * B->ns;
        for (int k = 0; k < B->ns; k++)
            for (int j = 0; j < B->nr; j++) {
                size_t fp = (size_t)j + (size_t)B->nr * k, px = IDX(B, xend, j, k);
                double x[3] = {B->X[px], B->X[NP + px], B->X[2 * NP + px]};
                const double *nm = B->nml[side];
                double n[3], m[3], l[3];
                for (int c = 0; c < 3; c++) { n[c] = nm[c * FP + fp]; m[c] = nm[(3 + c) * FP + fp]; l[c] = nm[(6 + c) * FP + fp]; }
                if (f->rs_init == 1) { /* 2820-2850 */
                    d->I[ib - 1].W[fp] = f->rs_W0;
                    continue;
                }
                double T0n, T0m, T0l;
                prestress(f, x, 0.0, l, m, n, &T0n, &T0m, &T0l);
                double a = rs_a(f, x);
                double sigma_n = -T0n;
                double theta = f->rs_L / f->rs_V0 *
                               exp(((a * log(2.0 * sinh(T0l / (a * sigma_n)))) - f->rs_f0 - a * log(f->rs_Vin / f->rs_V0)) / f->rs_b);
                d->I[ib - 1].W[fp] = f->rs_f0 + f->rs_b * log(f->rs_V0 / f->rs_L * theta);
            }
    }
}

/* time_step_RK time_step.f90:68-150 with the 2-rank domain semantics of
 * domain.f90:586-994 */
/* ---- moment-tensor point sources (moment_tensor.f90:958-1270) ---------- */

static double stf(int type, double t, double t0, double w) /* source_time, 958-1017 */
{
    const double pi = 3.141592653589793;
    switch (type) {
    case 0: return (w / sqrt(2.0 * pi)) * exp((-(w * w) * ((t - t0) * (t - t0))) / 2.0);
    case 1: return (t / (w * w)) * exp(-t / w);
    case 2: return (w * w) * t * exp(-w * t);
    case 3: return t < t0 ? (1.0 - cos(2.0 * pi * t / t0)) / t0 : 0.0;
    case 4: return (t >= 0.0 && t <= w) ? 1.0 - cos(2.0 * pi * t / w) : 0.0;
    case 5: return 1.0 - (1.0 + t / w) * exp(-t / w);
    default: return t <= w ? t - w / (2.0 * pi) * sin(2.0 * pi * t / w) : w;
    }
}

/* singular_source (1083-1270): 1-D discrete delta at node i (0-based) */
static double singular(int i, int i0, double h, double a, int order, int alpha)
{
    int k = i - i0;
    if (order == 2) {
        double beta = (1.0 - 2.0 * a) / 4.0 + 0.25, gam = (2.0 * a - 1.0) / 16.0, gap = (2.0 * a + 1.0) / 16.0;
        if (k == -2) return (-gam) / h;
        if (k == -1) return (beta + 3.0 * gam - gap) / h;
        if (k == 0) return (1.0 - a - 2.0 * beta - 3.0 * gam + 3.0 * gap) / h;
        if (k == 1) return (a + beta + gam - 3.0 * gap) / h;
        if (k == 2) return gap / h;
        return 0.0;
    }
    static const double tab[4][9] = {
        {-0.01171875, -0.03125, 0.046875, 0.28125, 0.4296875, 0.28125, 0.046875, -0.03125, -0.01171875},
        {-0.00246875, -0.02275, -0.014125, 0.16575, 0.4171875, 0.38675, 0.117875, -0.02975, -0.01846875},
        {-0.00728125, -0.028, 0.014875, 0.224, 0.4265625, 0.336, 0.081375, -0.032, -0.01553125},
        {-0.0095625, -0.02990625, 0.03053125, 0.25284375, 0.42890625, 0.30903125, 0.06384375, -0.03196875,
         -0.01371875}};
    if (k < -4 || k > 4) return 0.0;
    int row = alpha == 0 ? 0 : alpha == 25 ? 1 : alpha == 50 ? 2 : 3;
    return tab[row][k + 4] / h;
}

void ora_sources(ora_block *b, double t) /* set_moment_tensor, 1020-1080 */
{
    const int nq = b->nq, nr = b->nr, ns = b->ns;
    const size_t NP = NPTS(b);
    const int w = b->src_order == 2 ? 2 : 4;
    for (int s = 0; s < b->nsrc; s++) {
        const ora_source *S = &b->src[s];
        const double A = stf(S->type, t, S->t_init, S->duration);
        /* nodes outside the support get delta = 0 (an exact no-op there) */
        for (int k = S->node[2] - w; k <= S->node[2] + w; k++)
            for (int j = S->node[1] - w; j <= S->node[1] + w; j++)
                for (int i = S->node[0] - w; i <= S->node[0] + w; i++) {
                    if (i < 0 || i >= nq || j < 0 || j >= nr || k < 0 || k >= ns) continue;
                    size_t p = IDX(b, i, j, k);
                    double hx = (1.0 / ((double)nq - 1.0)) / b->metric[0 * NP + p];
                    double hy = (1.0 / ((double)nr - 1.0)) / b->metric[4 * NP + p];
                    double hz = (1.0 / ((double)ns - 1.0)) / b->metric[8 * NP + p];
                    double ax = (S->x[0] - S->node_x[0]) / hx;
                    double ay = (S->x[1] - S->node_x[1]) / hy;
                    double az = (S->x[2] - S->node_x[2]) / hz;
                    double dx = singular(i, S->node[0], hx, ax, b->src_order, b->src_alpha);
                    double dy = singular(j, S->node[1], hy, ay, b->src_order, b->src_alpha);
                    double dz = singular(k, S->node[2], hz, az, b->src_order, b->src_alpha);
                    double delta = dx * dy * dz;
                    for (int c = 0; c < 6; c++) b->DF[(3 + c) * NP + p] = b->DF[(3 + c) * NP + p] - A * S->m[c] * delta;
                }
    }
}

void ora_locate_sources(const ora_block *b, int n, const double *xyz, ora_source *src)
{
    const int nq = b->nq, nr = b->nr, ns = b->ns;
    const size_t NP = NPTS(b);
    const double *X = b->X;
    double lo[3], hi[3];
    for (int c = 0; c < 3; c++) {
        lo[c] = hi[c] = X[c * NP];
        for (size_t p = 1; p < NP; p++) {
            if (X[c * NP + p] < lo[c]) lo[c] = X[c * NP + p];
            if (X[c * NP + p] > hi[c]) hi[c] = X[c * NP + p];
        }
    }
    const size_t last = IDX(b, nq - 1, nr - 1, ns - 1);
    for (int s = 0; s < n; s++) {
        const double *x = xyz + 3 * s;
        int f[3] = {-1, -1, -1};
        if (lo[0] <= x[0] && x[0] <= hi[0] && lo[1] <= x[1] && x[1] <= hi[1] && lo[2] <= x[2] && x[2] <= hi[2]) {
            double mindist = 1.0e8;
            for (int k = 0; k < ns; k++)
                for (int j = 0; j < nr; j++)
                    for (int i = 0; i < nq; i++) {
                        size_t p = IDX(b, i, j, k);
                        double v0 = X[p] - x[0], v1 = X[NP + p] - x[1], v2 = X[2 * NP + p] - x[2];
                        double dist = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                        if (dist <= mindist) { f[0] = i; f[1] = j; f[2] = k; mindist = dist; }
                    }
        }
        int inb = X[0] <= x[0] && x[0] <= X[last] && X[NP] <= x[1] && x[1] <= X[NP + last] &&
                  X[2 * NP] <= x[2] && x[2] <= X[2 * NP + last];
        for (int c = 0; c < 3; c++) {
            src[s].x[c] = x[c];
            if (inb && f[0] >= 0) {
                src[s].node[c] = f[c];
                src[s].node_x[c] = X[c * NP + IDX(b, f[0], f[1], f[2])];
            } else {
                src[s].node[c] = -2;
                src[s].node_x[c] = -1.0;
            }
        }
    }
}

int ora_step(ora_domain *d, int nsteps)
{
    double A[5], Bc[5], C[6];
    ora_rk_coeffs(A, Bc, C);
    for (int n = 0; n < nsteps; n++) {
        double t0 = d->t;
        for (int stage = 1; stage <= 5; stage++) {
            double a = A[stage - 1];
            for (int ib = 0; ib < 2; ib++) {
                ora_block *b = &d->b[ib];
                size_t NP = NPTS(b);
#pragma omp parallel for schedule(static)
                for (size_t p = 0; p < 9 * NP; p++) b->DF[p] = a * b->DF[p];
                if (b->anelastic) /* scale_rates_interior, fields.f90:61-68 */
                    for (size_t p = 0; p < 24 * NP; p++) b->deta[p] = a * b->deta[p];
                for (int s = 0; s < 6; s++)
                    if (b->pml[s]) {
                        size_t SP = ora_slab_points(b, s);
                        for (size_t p = 0; p < 9 * SP; p++) b->DQ[s][p] = a * b->DQ[s][p];
                    }
                ora_iface *I = &d->I[ib];
                size_t FP = (size_t)b->nr * b->ns;
                for (size_t p = 0; p < 4 * FP; p++) I->DS[p] = a * I->DS[p];
                for (size_t p = 0; p < FP; p++) I->DW[p] = a * I->DW[p];
            }
            ora_set_rates(&d->b[0]);
            ora_set_rates(&d->b[1]);
            ora_sources(&d->b[0], d->t);
            ora_sources(&d->b[1], d->t);
            for (int ib = 1; ib <= 2; ib++) {
                ora_boundary(&d->b[ib - 1]);
                int rc = ora_interface(d, ib, d->t, stage);
                if (rc) return rc;
            }
            d->t = t0 + C[stage] * d->dt;
            double bdt = Bc[stage - 1] * d->dt;
            for (int ib = 0; ib < 2; ib++) {
                ora_block *b = &d->b[ib];
                size_t NP = NPTS(b);
#pragma omp parallel for schedule(static)
                for (size_t p = 0; p < 9 * NP; p++) b->F[p] = b->F[p] + bdt * b->DF[p];
                if (b->anelastic) /* update_fields_interior, fields.f90:122-129 */
                    for (size_t p = 0; p < 24 * NP; p++) b->eta[p] = b->eta[p] + bdt * b->deta[p];
                for (int s = 0; s < 6; s++)
                    if (b->pml[s]) {
                        size_t SP = ora_slab_points(b, s);
                        for (size_t p = 0; p < 9 * SP; p++) b->Q[s][p] = b->Q[s][p] + bdt * b->DQ[s][p];
                    }
                ora_iface *I = &d->I[ib];
                size_t FP = (size_t)b->nr * b->ns;
                for (size_t p = 0; p < 4 * FP; p++) I->S[p] = I->S[p] + bdt * I->DS[p];
                for (size_t p = 0; p < FP; p++) I->W[p] = I->W[p] + bdt * I->DW[p];
            }
            if (stage == 5 && d->plastic) {
                plastic_update(d, &d->b[0]);
                plastic_update(d, &d->b[1]);
            }
        }
    }
    return 0;
}
