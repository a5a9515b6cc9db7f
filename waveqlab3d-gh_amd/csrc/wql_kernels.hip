// wql_kernels.hip -- gfx950 kernels of the WaveQLab3D time-stepping path.
//
// Per RK stage (time_step.f90:109-147) the host enqueues:
//   k_boundary       every outer-face node of both blocks: characteristic SAT
//                    penalties (1516-1703) into a compact per-node array, and
//                    the fault SAT + friction solve (1706-1779,
//                    CouplingForcing.f90, Interface_Condition.f90)
//   k_volume_fused   interior + q/r closure rows: LDS-tiled q/r stencils with
//                    an s-marching register queue; fused DF = A*DF + L(F) +
//                    penalty, F_out = F + B*dt*DF               (HBM-bound)
//   k_volume_split   the six s-closure planes at each s face (role-split
//                    march in closure mode)
//   k_volume_generic the PML slabs, per node in the reference's boundary-band
//                    form (RHS_Interior.f90:221-268) with the PML auxiliary
//                    update (1782-2201)
//   k_plastic        last stage only: viscoplastic return map (plastic.f90)
#include "wql_device.h"
#include "wql_models.h"

#include <math.h>

namespace wql {

__constant__ SbpTable c_sbp;
// The closure rows as dense weights over the 9 planes next to a face:
// c_sbw*[r][p] for low rows r = 0..5 over planes 0..8, and for high rows
// r = 6..11 (row n-12+r) over planes n-9..n-1.  Filled with c_sbp.
__constant__ double c_sbw64[12][9];
__constant__ float c_sbw32[12][9];

// ----------------------------------------------------------------------------
// SBP stencil taps for row i of an n-node line (absolute indices)
__device__ __forceinline__ int sbp_taps(int i, int n, int *idx, double *c)
{
    if (i >= 6 && i <= n - 7) {
        idx[0] = i - 3; c[0] = -1.0 / 60.0;
        idx[1] = i - 2; c[1] = 0.15;
        idx[2] = i - 1; c[2] = -0.75;
        idx[3] = i + 1; c[3] = 0.75;
        idx[4] = i + 2; c[4] = -0.15;
        idx[5] = i + 3; c[5] = 1.0 / 60.0;
        return 6;
    }
    const SbpRow &R = (i < 6) ? c_sbp.left[i] : c_sbp.right[i - (n - 6)];
    for (int t = 0; t < R.n; t++) { idx[t] = i + R.off[t]; c[t] = R.c[t]; }
    return R.n;
}

// ----------------------------------------------------------------------------
// init: metrics and Jacobian in FP64 from node coordinates (grid.f90:396-469)
__global__ void k_metrics(int nq, int nr, int ns, const double *__restrict__ X, double *__restrict__ G)
{
    long NP = (long)nq * nr * ns;
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    int i = (int)(p % nq), j = (int)((p / nq) % nr), k = (int)(p / ((long)nq * nr));
    double D[3][3];  // D[axis][component] = dX_comp / d axis
    for (int a = 0; a < 3; a++) {
        int n = a == 0 ? nq : a == 1 ? nr : ns;
        int pos = a == 0 ? i : a == 1 ? j : k;
        int idx[8]; double c[8];
        int m = sbp_taps(pos, n, idx, c);
        double h = 1.0 / (double)(n - 1);
        for (int f = 0; f < 3; f++) {
            double s = 0.0;
            for (int t = 0; t < m; t++) {
                int ii = i, jj = j, kk = k;
                if (a == 0) ii = idx[t]; else if (a == 1) jj = idx[t]; else kk = idx[t];
                double v = c[t] * X[f * NP + pidx(nq, nr, ii, jj, kk)];
                s = (t == 0) ? v : s + v;
            }
            D[a][f] = (1.0 / h) * s;
        }
    }
    const double *Xq = D[0], *Xr = D[1], *Xs = D[2];
    double J = Xq[0] * (Xr[1] * Xs[2] - Xr[2] * Xs[1]) - Xq[1] * (Xr[0] * Xs[2] - Xr[2] * Xs[0]) +
               Xq[2] * (Xr[0] * Xs[1] - Xr[1] * Xs[0]);
    double iJ = 1.0 / J;
    G[0 * NP + p] = iJ * (Xr[1] * Xs[2] - Xr[2] * Xs[1]);
    G[1 * NP + p] = iJ * (Xq[2] * Xs[1] - Xq[1] * Xs[2]);
    G[2 * NP + p] = iJ * (Xq[1] * Xr[2] - Xq[2] * Xr[1]);
    G[3 * NP + p] = iJ * (Xr[2] * Xs[0] - Xr[0] * Xs[2]);
    G[4 * NP + p] = iJ * (Xq[0] * Xs[2] - Xq[2] * Xs[0]);
    G[5 * NP + p] = iJ * (Xq[2] * Xr[0] - Xq[0] * Xr[2]);
    G[6 * NP + p] = iJ * (Xr[0] * Xs[1] - Xr[1] * Xs[0]);
    G[7 * NP + p] = iJ * (Xq[1] * Xs[0] - Xq[0] * Xs[1]);
    G[8 * NP + p] = iJ * (Xq[0] * Xr[1] - Xq[1] * Xr[0]);
    G[9 * NP + p] = J;
}

// init: face frames (unit_normals.f90:8-168, gen_orth_vectors 232-266,
// Gram_Schmidt 269-282) and face copies of metric / material / coordinates
__global__ void k_frames(int nq, int nr, int ns, int side, const double *__restrict__ G,
                         const double *__restrict__ M, const double *__restrict__ X,
                         double *__restrict__ frame, double *__restrict__ fgm, double *__restrict__ fmat,
                         double *__restrict__ fx, int *__restrict__ err)
{
    int dir = side / 2;
    int na = dir == 0 ? nr : nq;
    int nb = dir == 2 ? nr : ns;
    long FP = (long)na * nb;
    long fp = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (fp >= FP) return;
    int jj = (int)(fp % na), kk = (int)(fp / na);
    int i, j, k;
    if (dir == 0) { i = (side == 0) ? 0 : nq - 1; j = jj; k = kk; }
    else if (dir == 1) { j = (side == 2) ? 0 : nr - 1; i = jj; k = kk; }
    else { k = (side == 4) ? 0 : ns - 1; i = jj; j = kk; }
    long NP = (long)nq * nr * ns;
    long p = pidx(nq, nr, i, j, k);
    double ax = G[(0 + dir) * NP + p], ay = G[(3 + dir) * NP + p], az = G[(6 + dir) * NP + p];
    double norm = sqrt(ax * ax + ay * ay + az * az);
    double n[3] = {ax / norm, ay / norm, az / norm};
    double m[3] = {0.0, 1.0, 0.0};
    const double tol = 1.0e-12;
    if (fabs(sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) - 1.0) >= tol) err[0] = 1;
    double d1 = sqrt((n[0] - m[0]) * (n[0] - m[0]) + (n[1] - m[1]) * (n[1] - m[1]) + (n[2] - m[2]) * (n[2] - m[2]));
    double d2 = sqrt((n[0] + m[0]) * (n[0] + m[0]) + (n[1] + m[1]) * (n[1] + m[1]) + (n[2] + m[2]) * (n[2] + m[2]));
    if (!(d1 >= tol && d2 >= tol)) { m[0] = 0.0; m[1] = 0.0; m[2] = 1.0; }
    double a = n[0] * m[0] + n[1] * m[1] + n[2] * m[2];
    for (int t = 0; t < 3; t++) m[t] = m[t] - a * n[t];
    double s = 1.0 / sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    for (int t = 0; t < 3; t++) m[t] = s * m[t];
    double l[3] = {n[1] * m[2] - n[2] * m[1], n[2] * m[0] - n[0] * m[2], n[0] * m[1] - n[1] * m[0]};
    for (int t = 0; t < 3; t++) {
        frame[(0 + t) * FP + fp] = n[t];
        frame[(3 + t) * FP + fp] = m[t];
        frame[(6 + t) * FP + fp] = l[t];
        fmat[t * FP + fp] = M[t * NP + p];
        if (fx) fx[t * FP + fp] = X[t * NP + p];
    }
    fgm[0 * FP + fp] = ax;
    fgm[1 * FP + fp] = ay;
    fgm[2 * FP + fp] = az;
}

// storage conversion FP64 -> real
template <typename real>
__global__ void k_convert(long n, const double *__restrict__ src, real *__restrict__ dst)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = (real)src[p];
}

// ----------------------------------------------------------------------------
// volume: generic per-node path (reference boundary-band form)

struct Box {
    int blk;
    int x0, x1, y0, y1, z0, z1;
    long start;  // prefix of node counts
};
struct BoxList {
    int nbox;
    Box box[16];
    long total;
};

template <typename real>
struct Pair {
    Blk<real> b[2];
};

template <typename real>
__device__ __forceinline__ double gmet(const Blk<real> &B, int c, long p)
{
    return B.cart ? B.gc[c] : (double)B.G[(long)c * B.NP + p];
}

template <typename real>
__device__ void node_tail(const Blk<real> &B, int i, int j, int k, long p, real (&U)[3][NF], const real (&fin)[NF],
                          real Jc, real A, real bdt);

template <typename real>
__device__ void volume_node_generic(const Blk<real> &B, int i, int j, int k, real A, real bdt)
{
    const long NP = B.NP;
    const long p = pidx(B.px, B.nr, i, j, k);
    real U[3][NF];  // Ux, Uy, Uz (JU_xJU_yJU_z6.f90:12-812, summed as in 808)
    const int nn[3] = {B.nq, B.nr, B.ns};
    const int pos[3] = {i, j, k};
    for (int a = 0; a < 3; a++) {
        int idx[8]; double c[8];
        int m = sbp_taps(pos[a], nn[a], idx, c);
        real sv[3] = {0, 0, 0};
        real ss[3][6];
        for (int fam = 0; fam < 3; fam++)
            for (int f = 0; f < 6; f++) ss[fam][f] = 0;
        // fixed trip count (closure rows have at most eight taps), so the
        // tap tables stay in registers and the taps' loads issue together
#pragma unroll
        for (int t = 0; t < 8; t++) {
            if (t >= m) break;
            int ii = i, jj = j, kk = k;
            if (a == 0) ii = idx[t]; else if (a == 1) jj = idx[t]; else kk = idx[t];
            long q = pidx(B.px, B.nr, ii, jj, kk);
            real ct = (real)c[t];
            real J = (real)gmet(B, 9, q);
            real mt[3] = {(real)gmet(B, 0 + a, q), (real)gmet(B, 3 + a, q), (real)gmet(B, 6 + a, q)};
            for (int f = 0; f < 3; f++) {
                real v = ct * B.Fin[f * NP + q];
                sv[f] = (t == 0) ? v : sv[f] + v;
            }
            for (int f = 0; f < 6; f++) {
                real cf = ct * B.Fin[(3 + f) * NP + q] * J;
                for (int fam = 0; fam < 3; fam++) {
                    real v = cf * mt[fam];
                    ss[fam][f] = (t == 0) ? v : ss[fam][f] + v;
                }
            }
        }
        real ih = (real)(1.0 / (1.0 / (double)(nn[a] - 1)));
        for (int fam = 0; fam < 3; fam++) {
            real mc = (real)gmet(B, 3 * fam + a, p);
            for (int f = 0; f < 3; f++) {
                real v = ih * (mc * sv[f]);
                U[fam][f] = (a == 0) ? v : U[fam][f] + v;
            }
            for (int f = 0; f < 6; f++) {
                real v = ih * ss[fam][f];
                U[fam][3 + f] = (a == 0) ? v : U[fam][3 + f] + v;
            }
        }
    }
    real fin[NF];
    for (int f = 0; f < NF; f++) fin[f] = B.Fin[f * NP + p];
    node_tail<real>(B, i, j, k, p, U, fin, (real)gmet(B, 9, p), A, bdt);
}

// Rates and stage output of one node from its derivative terms U[fam][f]:
// U[fam][0..2] = d(v_f)/d(x_fam) and U[fam][3+f'] = the flux derivative of
// stress f' in family fam.  Only the nine stress terms of the momentum
// equations are read (Ux: sxx sxy sxz, Uy: syy sxy syz, Uz: szz sxz syz).
// Hooke's law and the divergence (RHS_Interior.f90:221-268), the PML
// damping and auxiliary update of every PML slab holding the node
// (1782-2201), the boundary penalty, DF = A DF + rate, F_out = F + B dt DF.
// fin: the node's stage input, Jc: its Jacobian.
template <typename real>
__device__ void node_tail(const Blk<real> &B, int i, int j, int k, long p, real (&U)[3][NF], const real (&fin)[NF],
                          real Jc, real A, real bdt)
{
    const long NP = B.NP;
    const int nn[3] = {B.nq, B.nr, B.ns};
    const int pos[3] = {i, j, k};
    const real *Ux = U[0], *Uy = U[1], *Uz = U[2];
    real lam, mu, rho;
    if (B.M) { lam = B.M[p]; mu = B.M[NP + p]; rho = B.M[2 * NP + p]; }
    else { lam = (real)B.lam; mu = (real)B.mu; rho = (real)B.rho; }
    real rhoJ_inv = (real)1 / (rho * Jc);
    real l2m = lam + (real)2 * mu;
    real DU[NF];
    DU[0] = (Ux[3] + Uy[6] + Uz[7]) * rhoJ_inv;
    DU[1] = (Ux[6] + Uy[4] + Uz[8]) * rhoJ_inv;
    DU[2] = (Ux[7] + Uy[8] + Uz[5]) * rhoJ_inv;
    DU[3] = l2m * Ux[0] + lam * (Uy[1] + Uz[2]);
    DU[4] = l2m * Uy[1] + lam * (Ux[0] + Uz[2]);
    DU[5] = l2m * Uz[2] + lam * (Ux[0] + Uy[1]);
    DU[6] = mu * (Uy[0] + Ux[1]);
    DU[7] = mu * (Uz[0] + Ux[2]);
    DU[8] = mu * (Uz[1] + Uy[2]);
    const double d0 = 6.0 / (2.0 * 1.0) * log(1.0 / 1.0e-3);
    if (B.eta) {
        // anelastic memory variables (apply_anelastic_point, RHS_Interior.f90:
        // 2390-2441): DU -= sum over mechanisms, Deta = A Deta + rate, and the
        // stage update eta += B dt Deta in place (a node's memory variables are
        // read by that node only).  In PML slabs the strains are taken net of
        // the auxiliary fields (apply_anelastic_point_pml, 2444-2509,
        // pml_damping_and_q 2287-2387); off the slabs d = 0 and Q = 0 leave
        // them as they are.
        real dd[3] = {0, 0, 0}, qx[3] = {0, 0, 0}, qy[3] = {0, 0, 0}, qz[3] = {0, 0, 0};
        for (int side = 0; side < 6; side++) {
            if (!B.pml[side]) continue;
            const int dir = side / 2, np = B.npml, n = nn[dir];
            const int p1 = pos[dir] + 1;
            if (side % 2 == 0 ? p1 > np : p1 < n - np + 1) continue;
            int lo[3] = {0, 0, 0}, hi[3] = {B.nq - 1, B.nr - 1, B.ns - 1};
            if (side % 2 == 0) hi[dir] = np - 1; else lo[dir] = n - np;
            const long n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1, n2 = hi[2] - lo[2] + 1;
            const long s = (long)(pos[0] - lo[0]) + n0 * ((long)(pos[1] - lo[1]) + n1 * (long)(pos[2] - lo[2]));
            const long SP = n0 * n1 * n2;
            double dv;
            if (side == 4) dv = d0 * fabs((double)(p1 - np) / (double)np);
            else if (side % 2 == 0) dv = d0 * fabs((double)(p1 - (1 + np)) / (double)np);
            else dv = d0 * fabs((double)(p1 - (n - np)) / (double)np);
            dd[dir] = (real)dv;
            const real *Qs = B.Q[side];
            if (dir == 0) { qx[0] = Qs[3 * SP + s]; qx[1] = Qs[6 * SP + s]; qx[2] = Qs[7 * SP + s]; }
            else if (dir == 1) { qy[0] = Qs[4 * SP + s]; qy[1] = Qs[6 * SP + s]; qy[2] = Qs[8 * SP + s]; }
            else { qz[0] = Qs[5 * SP + s]; qz[1] = Qs[7 * SP + s]; qz[2] = Qs[8 * SP + s]; }
        }
        real e[6];
        e[0] = Ux[0] - dd[0] * qx[0];
        e[1] = Uy[1] - dd[1] * qy[0];
        e[2] = Uz[2] - dd[2] * qz[0];
        e[3] = (Uy[0] + Ux[1]) - dd[0] * qx[1] - dd[1] * qy[1];
        e[4] = (Uz[0] + Ux[2]) - dd[0] * qx[2] - dd[2] * qz[1];
        e[5] = (Uz[1] + Uy[2]) - dd[1] * qy[2] - dd[2] * qz[2];
        const real Qsi = B.qinv[p], Qpi = B.qinv[NP + p];
        const real tr = e[0] + e[1] + e[2];
        real et[24];
        for (int q = 0; q < 24; q++) et[q] = B.eta[q * NP + p];
        for (int c = 0; c < 6; c++) DU[3 + c] = DU[3 + c] - (et[4 * c] + et[4 * c + 1] + et[4 * c + 2] + et[4 * c + 3]);
        for (int l = 0; l < 4; l++) {
            const real w = (real)B.an_w[l], tau = (real)B.an_tau[l];
            for (int c = 0; c < 6; c++) {
                const int q = 4 * c + l;
                real rate;
                if (c < 3)
                    rate = (((w * (real)2 * mu * Qsi) * e[c] + (w * ((lam + (real)2 * mu) * Qpi - (real)2 * mu * Qsi)) * tr)
                            - et[q]) / tau;
                else
                    rate = (w * mu * Qsi * e[c] - et[q]) / tau;
                const real de = A * B.deta[q * NP + p] + rate;
                B.deta[q * NP + p] = de;
                B.eta[q * NP + p] = et[q] + bdt * de;
            }
        }
    }
    // PML (RHS_Interior.f90:1782-2201), order z-low, z-high, y-low, y-high, x-low, x-high
    const int order[6] = {4, 5, 2, 3, 0, 1};
    const real cfs = (real)1;
    for (int o = 0; o < 6; o++) {
        int side = order[o];
        if (!B.pml[side]) continue;
        int dir = side / 2, np = B.npml;
        int lo[3] = {0, 0, 0}, hi[3] = {B.nq - 1, B.nr - 1, B.ns - 1};
        int n = nn[dir];
        if (side % 2 == 0) hi[dir] = np - 1; else lo[dir] = n - np;
        if (pos[0] < lo[0] || pos[0] > hi[0] || pos[1] < lo[1] || pos[1] > hi[1] || pos[2] < lo[2] || pos[2] > hi[2])
            continue;
        long n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1, n2 = hi[2] - lo[2] + 1;
        long s = (long)(pos[0] - lo[0]) + n0 * ((long)(pos[1] - lo[1]) + n1 * (long)(pos[2] - lo[2]));
        long SP = n0 * n1 * n2;
        int p1 = pos[dir] + 1;
        double dd;
        if (side == 4) dd = d0 * fabs((double)(p1 - np) / (double)np);
        else if (side % 2 == 0) dd = d0 * fabs((double)(p1 - (1 + np)) / (double)np);
        else dd = d0 * fabs((double)(p1 - (n - np)) / (double)np);
        real d = (real)dd;
        real *Q = B.Q[side], *DQ = B.DQ[side];
        real q[NF];
        for (int f = 0; f < NF; f++) q[f] = Q[f * SP + s];
        DU[0] = DU[0] - d * Jc * q[0] * rhoJ_inv;
        DU[1] = DU[1] - d * Jc * q[1] * rhoJ_inv;
        DU[2] = DU[2] - d * Jc * q[2] * rhoJ_inv;
        DU[3] = DU[3] - d * (l2m * q[3] + lam * q[4] + lam * q[5]);
        DU[4] = DU[4] - d * (lam * q[3] + l2m * q[4] + lam * q[5]);
        DU[5] = DU[5] - d * (lam * q[3] + lam * q[4] + l2m * q[5]);
        DU[6] = DU[6] - mu * d * q[6];
        DU[7] = DU[7] - mu * d * q[7];
        DU[8] = DU[8] - mu * d * q[8];
        real src[NF];
        if (dir == 2) {
            src[0] = Uz[7] / Jc; src[1] = Uz[8] / Jc; src[2] = Uz[5] / Jc;
            src[3] = 0; src[4] = 0; src[5] = Uz[2]; src[6] = 0; src[7] = Uz[0]; src[8] = Uz[1];
        } else if (dir == 1) {
            src[0] = Uy[6] / Jc; src[1] = Uy[4] / Jc; src[2] = Uy[8] / Jc;
            src[3] = 0; src[4] = Uy[1]; src[5] = 0; src[6] = Uy[0]; src[7] = 0; src[8] = Uy[2];
        } else {
            src[0] = Ux[3] / Jc; src[1] = Ux[6] / Jc; src[2] = Ux[7] / Jc;
            src[3] = Ux[0]; src[4] = 0; src[5] = 0; src[6] = Ux[1]; src[7] = Ux[2]; src[8] = 0;
        }
        // this side's boundary-face penalty (written by the boundary kernel
        // before this launch), for a node on that face
        const real *pq = nullptr;
        if (B.PQ[side] && B.bc[side] > 0 && pos[dir] == (side % 2 == 0 ? 0 : n - 1)) {
            long fpi;
            if (dir == 0) fpi = j + (long)B.nr * k;
            else if (dir == 1) fpi = i + (long)B.nq * k;
            else fpi = i + (long)B.nq * j;
            pq = B.PQ[side] + fpi;
        }
        const long FPs = dir == 0 ? (long)B.nr * B.ns : dir == 1 ? (long)B.nq * B.ns : (long)B.nq * B.nr;
        for (int f = 0; f < NF; f++) {
            real dq = A * DQ[f * SP + s];
            dq = dq + src[f] - (d + cfs) * q[f];
            if (pq) dq = dq + pq[f * FPs];
            DQ[f * SP + s] = dq;
            Q[f * SP + s] = q[f] + bdt * dq;
        }
    }
    const long bi = B.satb ? bnd_index(B, i, j, k) : -1;
    real *fq = nullptr;
    if (B.FQ && (i == 0 || i == B.nq - 1))
        fq = B.FQ + (i == 0 ? 0L : 9L * B.nr * B.ns) + j + (long)B.nr * k;
    for (int f = 0; f < NF; f++) {
        real df = A * B.DF[f * NP + p];
        df = df + DU[f];
        if (bi >= 0) df = df + B.satb[f * B.nbnd + bi];
        B.DF[f * NP + p] = df;
        const real fo = fin[f] + bdt * df;
        B.Fout[f * NP + p] = fo;
        if (fq) fq[(long)f * B.nr * B.ns] = fo;
    }
}

template <typename real>
__global__ void __launch_bounds__(256) k_volume_generic(Pair<real> P, BoxList L, real A0, real A1, real bdt)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.total) return;
    int bi = 0;
    for (int b = 1; b < L.nbox; b++)
        if (t >= L.box[b].start) bi = b;
    const Box &bx = L.box[bi];
    long r = t - bx.start;
    int w = bx.x1 - bx.x0, h = bx.y1 - bx.y0;
    int i = bx.x0 + (int)(r % w);
    int j = bx.y0 + (int)((r / w) % h);
    int k = bx.z0 + (int)(r / ((long)w * h));
    const Blk<real> &B = P.b[bx.blk];
    volume_node_generic<real>(B, i, j, k, bx.blk == 0 ? A0 : A1, bdt);
}

// ----------------------------------------------------------------------------
// volume: the upwind and upwind-DRP operators ('upwind' / 'upwind_drp')
//
// One thread per node of the whole block.  The operator is a table
// (wql_sbp_ops.h): per direction, the closure rows at both faces and the
// interior row, for the velocities (forward operator Dp) and for the stress
// fluxes J a_x sigma (backward operator Dm).  A node whose three directions
// are all interior (band width per order, widened to npml on PML sides)
// takes the form of the interior routine JJU_x<o>_interior_upwind[_drp]
// (JU_xJU_yJU_z6.f90:16451-20922): the exact rational coefficients, one
// loop over the taps adding the three directions per tap; with Cartesian
// metrics the diagonal metric of the node and no J.  Every other node
// takes the band routine JJU_x<o>_upwind[_drp] (815-16122, through
// RHS_near_boundaries, RHS_Interior.f90:144-1506): per direction the row's
// taps, velocities X_x(node) sum c v, stresses sum c X_x J sigma, and
// Ju = (1/hq) Jq + (1/hr) Jr + (1/hs) Js.  Hooke's law, the PML terms, the
// boundary penalty and the stage update follow in node_tail.

// calls fn(node, coefficient) for each tap of node pos on a line of n
// nodes: a closure row or the printed interior row (no tap arrays, which
// would sit in scratch memory)
template <typename real, typename Fn>
__device__ __forceinline__ void op_taps(const wql_sbp_half &h, int pos, int n, Fn fn)
{
    if (pos < h.nl) {
        for (int t = 0; t < WQL_SBP_MAXW; t++) {
            const double v = h.left[pos][t];
            if (v != 0.0) fn(t, (real)v);
        }
    } else if (pos >= n - h.nr) {
        const int r = n - 1 - pos;
        for (int t = 0; t < WQL_SBP_MAXW; t++) {
            const double v = h.right[r][t];
            if (v != 0.0) fn(n - 1 - t, (real)v);
        }
    } else {
        for (int o = h.lo; o <= h.hi; o++) fn(pos + o, (real)h.band[o - h.lo]);
    }
}

template <typename real>
__device__ void volume_node_op(const Blk<real> &B, int i, int j, int k, real A, real bdt)
{
    if (!B.sbp) {
        // the sixth-order central operator at every node (an anelastic block
        // on the per-node path)
        volume_node_generic<real>(B, i, j, k, A, bdt);
        return;
    }
    const long NP = B.NP;
    const long p = pidx(B.px, B.nr, i, j, k);
    const int nn[3] = {B.nq, B.nr, B.ns};
    const int pos[3] = {i, j, k};
    const wql_sbp_half *V = B.sbp, *S = B.sbp + 3;
    real U[3][NF];
    const int w = B.sbw;
    bool interior = true;
    for (int a = 0; a < 3; a++) {
        const int lo = B.pml[2 * a] && B.npml > w ? B.npml : w;
        const int hi = B.pml[2 * a + 1] && B.npml > w ? B.npml : w;
        interior = interior && pos[a] >= lo && pos[a] <= nn[a] - 1 - hi;
    }
    real Jc;
    if (interior) {
        for (int fam = 0; fam < 3; fam++)
            for (int f = 0; f < NF; f++) U[fam][f] = 0;
        real gp[9];
        for (int c = 0; c < 9; c++) gp[c] = (real)gmet(B, c, p);
        real ih[3];
        for (int a = 0; a < 3; a++) ih[a] = (real)(1.0 / (1.0 / (double)(nn[a] - 1)));
        const int nt = V[0].hi - V[0].lo + 1;
        for (int t = 0; t < nt; t++) {
            long pv[3], ps[3];
            real cv[3], cs[3];
            for (int a = 0; a < 3; a++) {
                const int ov = V[a].lo + t, os = S[a].lo + t;
                int q[3] = {i, j, k};
                q[a] += ov;
                pv[a] = pidx(B.px, B.nr, q[0], q[1], q[2]);
                q[a] += os - ov;
                ps[a] = pidx(B.px, B.nr, q[0], q[1], q[2]);
                cv[a] = (real)V[a].exact[t] * ih[a];
                cs[a] = (real)S[a].exact[t] * ih[a];
            }
            for (int fam = 0; fam < 3; fam++) {
                if (B.cart) {
                    const real g = gp[3 * fam + fam];
                    for (int f = 0; f < 3; f++) U[fam][f] = U[fam][f] + g * cv[fam] * B.Fin[f * NP + pv[fam]];
                    for (int f = 3; f < NF; f++) U[fam][f] = U[fam][f] + g * cs[fam] * B.Fin[f * NP + ps[fam]];
                } else {
                    for (int f = 0; f < 3; f++) {
                        real v = U[fam][f];
                        for (int a = 0; a < 3; a++) v = v + gp[3 * fam + a] * cv[a] * B.Fin[f * NP + pv[a]];
                        U[fam][f] = v;
                    }
                    real gJ[3];
                    for (int a = 0; a < 3; a++) gJ[a] = (real)gmet(B, 3 * fam + a, ps[a]) * (real)gmet(B, 9, ps[a]);
                    for (int f = 3; f < NF; f++) {
                        real v = U[fam][f];
                        for (int a = 0; a < 3; a++) v = v + cs[a] * B.Fin[f * NP + ps[a]] * gJ[a];
                        U[fam][f] = v;
                    }
                }
            }
        }
        Jc = B.cart ? (real)1 : (real)gmet(B, 9, p);
    } else {
#pragma unroll 1
        for (int a = 0; a < 3; a++) {
            // node index along direction a -> its linear index
            const long base = p - (long)pos[a] * (a == 0 ? 1L : a == 1 ? (long)B.px : (long)B.px * B.nr);
            const long step = a == 0 ? 1L : a == 1 ? (long)B.px : (long)B.px * B.nr;
            real sv[3] = {0, 0, 0};
            op_taps<real>(V[a], pos[a], nn[a], [&](int node, real cf) {
                const long qq = base + (long)node * step;
#pragma unroll
                for (int f = 0; f < 3; f++) sv[f] = sv[f] + cf * B.Fin[f * NP + qq];
            });
            real ss[3][6];
#pragma unroll
            for (int fam = 0; fam < 3; fam++)
#pragma unroll
                for (int f = 0; f < 6; f++) ss[fam][f] = 0;
            op_taps<real>(S[a], pos[a], nn[a], [&](int node, real cf) {
                const long qq = base + (long)node * step;
                const real J = (real)gmet(B, 9, qq);
                real cm[3];
#pragma unroll
                for (int fam = 0; fam < 3; fam++) cm[fam] = cf * (real)gmet(B, 3 * fam + a, qq) * J;
#pragma unroll
                for (int f = 0; f < 6; f++) {
                    const real sg = B.Fin[(3 + f) * NP + qq];
#pragma unroll
                    for (int fam = 0; fam < 3; fam++) ss[fam][f] = ss[fam][f] + cm[fam] * sg;
                }
            });
            const real ih = (real)(1.0 / (1.0 / (double)(nn[a] - 1)));
            for (int fam = 0; fam < 3; fam++) {
                const real mc = (real)gmet(B, 3 * fam + a, p);
                for (int f = 0; f < 3; f++) {
                    const real v = ih * (mc * sv[f]);
                    U[fam][f] = a == 0 ? v : U[fam][f] + v;
                }
                for (int f = 0; f < 6; f++) {
                    const real v = ih * ss[fam][f];
                    U[fam][3 + f] = a == 0 ? v : U[fam][3 + f] + v;
                }
            }
        }
        Jc = (real)gmet(B, 9, p);
    }
    real fin[NF];
    for (int f = 0; f < NF; f++) fin[f] = B.Fin[f * NP + p];
    node_tail<real>(B, i, j, k, p, U, fin, Jc, A, bdt);
}

// three waves per SIMD in FP32 (a few spilled registers cost less than the
// third wave gains: the per-node loads are L2-latency bound), two in FP64
// (three spill 260 bytes a lane and run slower)
template <typename real>
__global__ void __launch_bounds__(256, sizeof(real) == 4 ? 3 : 2)
    k_volume_op(Pair<real> P, BoxList L, real A0, real A1, real bdt)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.total) return;
    int bi = 0;
    for (int b = 1; b < L.nbox; b++)
        if (t >= L.box[b].start) bi = b;
    const Box &bx = L.box[bi];
    long r = t - bx.start;
    int w = bx.x1 - bx.x0, h = bx.y1 - bx.y0;
    int i = bx.x0 + (int)(r % w);
    int j = bx.y0 + (int)((r / w) % h);
    int k = bx.z0 + (int)(r / ((long)w * h));
    volume_node_op<real>(P.b[bx.blk], i, j, k, bx.blk == 0 ? A0 : A1, bdt);
}

// init_anelastic_properties (material.f90:11-145) on the FP64 material M
// [3][nd] (lambda, mu, rho): Q_s^-1 = 1/(c v_s), Q_p^-1 = Q_s^-1 / 2 into
// qinv [2][nd], then the unrelaxed lambda and mu written back into M.  tau,
// w: relaxation times and weights of the four mechanisms; wref = 2 pi fref.
__global__ void k_anelastic_init(long nd, double *M, double *qinv, double c, double wref, double t0, double t1,
                                 double t2, double t3, double w0, double w1, double w2, double w3)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nd) return;
    const double tau[4] = {t0, t1, t2, t3}, w[4] = {w0, w1, w2, w3};
    const double lam = M[p], mu = M[nd + p], rho = M[2 * nd + p];
    const double Qs = 1.0 / (c * sqrt(mu / rho)), Qp = 0.5 * Qs;
    qinv[p] = Qs;
    qinv[nd + p] = Qp;
    double val_S = 0.0, val_P = 0.0;
    for (int l = 0; l < 4; l++) {
        const double den_S = ((wref * wref) * (tau[l] * tau[l]) + 1.0) * (1.0 / Qs) * 1.0;
        const double den_P = ((wref * wref) * (tau[l] * tau[l]) + 1.0) * (1.0 / Qp) * 1.0;
        val_S = val_S + w[l] / den_S;
        val_P = val_P + w[l] / den_P;
    }
    double vs = sqrt(mu / rho), vp = sqrt((lam + 2.0 * mu) / rho);
    const double mu_S = rho * (vs * vs) / (1.0 - val_S), mu_P = rho * (vp * vp) / (1.0 - val_P);
    vs = sqrt(mu_S / rho);
    vp = sqrt(mu_P / rho);
    M[nd + p] = (vs * vs) * rho;
    M[p] = (vp * vp) * rho - 2.0 * M[nd + p];
}

// ----------------------------------------------------------------------------
// volume: main tiled kernel
//
// A thread block owns a TX x TY tile of (q, r) columns and marches a chunk of
// s planes.  Velocities and the q-/r-direction stress fluxes
//     f_a,i = J (a_x s_xi + a_y s_yi + a_z s_zi),  a = q, r
// of the current plane sit in LDS with a 3-node halo; the s-direction data
// (velocities and s-fluxes of planes z-3..z+3) ride in a register queue.
// Boundary closure rows in q and r read the same LDS tile: the tiling keeps
// all six closure rows and their taps inside one tile (first tile starts at
// the domain edge, last tile is right-aligned, TX, TY >= 6).

struct TileGeom {
    int blk;
    int xa0, xa1, ya0, ya1, za0, za1;  // node range of this kernel in the block
    int ntx, nty, nzc, zc;             // tiles in x, y; z chunks and chunk length
    long first;                        // first logical tile id of this block
};
struct TileList {
    int n;
    TileGeom g[4];
    long total;
    // 0: q-tiles fastest, then r-tiles, then s chunks
    // 1: s chunks fastest, then q-tiles, then r-tiles
    // 2: groups of gw q-tile columns; inside a group s chunks outermost,
    //    then r-tiles, then q-tiles.  Consecutive logical ids land on one
    //    XCD, so each XCD walks its group plane by plane and the s-neighbour
    //    planes and q/r halos of a tile are still in that XCD's L2.
    // 3: s chunks in pairs (2m, 2m+1) of one tile at consecutive ids, then
    //    q-tiles, r-tiles, pairs.  The fused march runs the lower chunk of a
    //    pair downwards and the upper one upwards, so both start at their
    //    shared boundary and load its warm-up planes together (one XCD, L2).
    int order;
    int gw;
};

// logical tile id -> (q tile, r tile, s chunk or plane)
__device__ __forceinline__ void tile_coords(const TileList &TL, const TileGeom &T, long lid, int nz, int &tx, int &ty,
                                            int &tz)
{
    if (TL.order == 0) {
        tx = (int)(lid % T.ntx);
        ty = (int)((lid / T.ntx) % T.nty);
        tz = (int)(lid / ((long)T.ntx * T.nty));
    } else if (TL.order == 1) {
        tz = (int)(lid % nz);
        tx = (int)((lid / nz) % T.ntx);
        ty = (int)(lid / ((long)nz * T.ntx));
    } else if (TL.order == 3) {
        const long layer = (long)T.ntx * T.nty;
        const long pr = lid / (2 * layer);
        const long rem = lid - pr * 2 * layer;
        if (2 * pr + 1 >= nz) {  // last chunk of an odd count: unpaired
            tx = (int)(rem % T.ntx);
            ty = (int)(rem / T.ntx);
            tz = (int)(2 * pr);
        } else {
            const long t2 = rem >> 1;
            tx = (int)(t2 % T.ntx);
            ty = (int)(t2 / T.ntx);
            tz = (int)(2 * pr + (rem & 1));
        }
    } else {
        const int gw = TL.gw;
        const int ngx = (T.ntx + gw - 1) / gw;
        const long full = (long)gw * T.nty * nz;
        int g = (int)(lid / full);
        int w = gw;
        long rem;
        if (g >= ngx - 1) {
            g = ngx - 1;
            w = T.ntx - g * gw;
            rem = lid - (long)g * full;
        } else {
            rem = lid - (long)g * full;
        }
        const long pl = (long)w * T.nty;
        tz = (int)(rem / pl);
        const int r2 = (int)(rem - (long)tz * pl);
        ty = r2 / w;
        tx = g * gw + (r2 - ty * w);
    }
}

template <int T>
__device__ __forceinline__ void tile_range(int t, int nt, int a0, int a1, int nfull, int &start, int &c0, int &c1)
{
    // start: first node held by the tile; [c0,c1): nodes computed by it.
    // The last tile owns the six closure rows at a face (cut) and must hold
    // every tap they reach (nodes n-9..n-1, the first three through its
    // halo): it starts at its natural position when that is at most n-6,
    // else it is right-aligned at a1 - T.
    int cut = (a1 == nfull) ? nfull - 6 : a1;
    if (t < nt - 1) {
        start = a0 + t * T;
        c0 = start;
        c1 = min(start + T, cut);
        if (c1 < c0) c1 = c0;
    } else {
        const int nat = a0 + t * T;
        start = nat <= cut ? nat : max(a0, a1 - T);
        c0 = min(nat, cut);
        if (c0 < a0) c0 = a0;
        c1 = a1;
    }
}

// ----------------------------------------------------------------------------
// Raw buffer loads and stores of the tiled volume kernels: the per-thread
// byte offset (one VGPR per node column) is shared by all fields and the
// field offset rides in an SGPR, so no 64-bit address arithmetic is spent
// per access.  A block's arrays must each span < 2 GiB (checked by
// wql_create).
template <typename real>
__device__ __forceinline__ real bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(real) == 4) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
    } else {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
    }
}

template <typename real>
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, real v)
{
    if constexpr (sizeof(real) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 0);
    } else {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)voff, (int)soff, 0);
    }
}

// non-temporal forms (cache policy NT): for data no other workgroup reads in
// this launch (the rates and the stage output), so that it does not evict the
// lines the neighbouring tiles' halo loads hit in L2
template <typename real>
__device__ __forceinline__ real bload_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(real) == 4) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 2));
    } else {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 2));
    }
}

template <typename real>
__device__ __forceinline__ void bstore_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, real v)
{
    if constexpr (sizeof(real) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 2);
    } else {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)voff, (int)soff, 2);
    }
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const void *p, unsigned bytes)
{
    // raw (stride 0) buffer over [p, p + bytes): out-of-range loads return 0
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}

// Dense closure weights in LDS: row r of the 12 closure rows over the nine
// nodes next to its face (c_sbw32 / c_sbw64), so a closure node is nine
// independent LDS reads and FMAs with no per-lane tap loop.
template <typename real>
struct SbpDense {
    real w[12][9];
};

template <typename real>
__device__ __forceinline__ void sbd_to_lds(SbpDense<real> &S, int tid, int nthreads)
{
    for (int e = tid; e < 12 * 9; e += nthreads) {
        if constexpr (sizeof(real) == 4) S.w[e / 9][e % 9] = c_sbw32[e / 9][e % 9];
        else S.w[e / 9][e % 9] = c_sbw64[e / 9][e % 9];
    }
}

// first derivative at local index l of an LDS line (element stride st) whose
// global index is g on an n-node line; closure rows 0..5 use nodes 0..8,
// rows n-6..n-1 nodes n-9..n-1
template <typename real>
__device__ __forceinline__ real d1_dense(const SbpDense<real> &S, const real *v, int st, int l, int g, int n, real ih)
{
    if (g >= 6 && g <= n - 7) {
        return ih * ((real)0.75 * (v[(l + 1) * st] - v[(l - 1) * st]) +
                     (real)(-0.15) * (v[(l + 2) * st] - v[(l - 2) * st]) +
                     (real)(1.0 / 60.0) * (v[(l + 3) * st] - v[(l - 3) * st]));
    }
    const bool lo = g < 6;
    const int r = lo ? g : g - n + 12;
    const int b = lo ? l - g : l - g + n - 9;
    real acc = 0;
#pragma unroll
    for (int t = 0; t < 9; t++) acc = acc + S.w[r][t] * v[(b + t) * st];
    return ih * acc;
}

// ----------------------------------------------------------------------------
// Role-split pipelined march.  The velocity rates depend on the stresses only
// and the stress rates on the velocities only, so each tile is processed by
// two workgroups with consecutive ids (one XCD, dispatched together): role 0
// marches the stress fluxes and writes DF/F_out of vx vy vz, role 1 marches
// the velocities and writes DF/F_out of the six stresses.  Each carries half
// of the fused kernel's state (one 7-plane s-queue of 3 values), which keeps
// occupancy high; the metric and field lines both roles touch are fetched
// from HBM once and hit L2 for the second role.
// minimum waves per SIMD asked of the register allocator: three for FP32
// (<= 168 VGPRs), two for FP64 (<= 256 registers incl. AGPRs)
template <typename real>
constexpr int split_min_waves() { return sizeof(real) == 4 ? 3 : 2; }

template <typename real, int TX, int TY, bool CURV, bool HET, int SB>
__global__ void __launch_bounds__(TX * TY, split_min_waves<real>()) k_volume_split(Pair<real> P, TileList TL, real A0, real A1, real bdt)
{
    constexpr int WX = TX + 6, WY = TY + 6;
    constexpr unsigned RS = sizeof(real);
    // role 0: flux tiles sfq/sfr; role 1: velocity tile sv (same storage)
    constexpr int NQ = 2 * 3 * TY * WX, NR = 2 * 3 * WY * TX, NV = 2 * 3 * WY * WX;
    constexpr int NL = (NQ + NR) > NV ? (NQ + NR) : NV;
    __shared__ real lds[NL];
    real(*sfq)[3][TY][WX] = reinterpret_cast<real(*)[3][TY][WX]>(lds);
    real(*sfr)[3][WY][TX] = reinterpret_cast<real(*)[3][WY][TX]>(lds + NQ);
    real(*sv)[3][WY][WX] = reinterpret_cast<real(*)[3][WY][WX]>(lds);
    __shared__ SbpDense<real> sbd;

    const long nwg = (long)gridDim.x;
    const long per = (nwg + 7) / 8;
    const long logical2 = (long)(blockIdx.x % 8) * per + (long)(blockIdx.x / 8);
    const int role = (int)(logical2 & 1);
    const long logical = logical2 >> 1;
    if (logical >= TL.total) return;
    int gi = 0;
    for (int k = 1; k < TL.n; k++)
        if (logical >= TL.g[k].first) gi = k;
    const TileGeom &T = TL.g[gi];
    const Blk<real> &B = P.b[T.blk];
    const real A = T.blk == 0 ? A0 : A1;
    // the first RK stage has A = 0: its rates are not read (A * DF = 0)
    const bool rdA = A != (real)0;
    long lid = logical - T.first;
    int tx_i, ty_i, zc_i;
    tile_coords(TL, T, lid, T.nzc, tx_i, ty_i, zc_i);

    int xs, xc0, xc1, ys, yc0, yc1;
    tile_range<TX>(tx_i, T.ntx, T.xa0, T.xa1, B.nq, xs, xc0, xc1);
    tile_range<TY>(ty_i, T.nty, T.ya0, T.ya1, B.nr, ys, yc0, yc1);
    // chunks of balanced length (lengths differ by at most one plane)
    const int span = T.za1 - T.za0;
    const int z0 = T.za0 + (int)((long)zc_i * span / T.nzc);
    const int z1 = T.za0 + (int)((long)(zc_i + 1) * span / T.nzc);

    const int lx = threadIdx.x, ly = threadIdx.y;
    const int gx = xs + lx, gy = ys + ly;
    const bool own = (gx >= xc0 && gx < xc1 && gy >= yc0 && gy < yc1);
    const bool inb = (gx < B.nq && gy < B.nr);
    const unsigned NPb = (unsigned)B.NP * RS;
    const unsigned planeb = (unsigned)(B.px * B.nr) * RS;
    const unsigned colb = (unsigned)(gx + B.px * gy) * RS;
    // boundary-node index of this column (fused boundary penalties): a q/r
    // face column has index bbase + bstride*z; inner columns are boundary
    // nodes only on the s faces (closure mode), at bsface (+ one s face)
    int bbase = -1, bstride = 0, bsface = -1;
    if (B.satb && own) {
        const int nq = B.nq, nr = B.nr, ns = B.ns;
        const int nqf = 2 * nr * ns, nrf = 2 * (nq - 2) * ns;
        if (gx == 0 || gx == nq - 1) { bbase = (gx == 0 ? 0 : nr * ns) + gy; bstride = nr; }
        else if (gy == 0 || gy == nr - 1) { bbase = nqf + (gy == 0 ? 0 : (nq - 2) * ns) + gx - 1; bstride = nq - 2; }
        else bsface = nqf + nrf + (gx - 1) + (nq - 2) * (gy - 1);
    }
    auto bnd = [&](int z) -> int {
        if (bbase >= 0) return bbase + bstride * z;
        if (SB != 0 && bsface >= 0) {
            if (z == 0) return bsface;
            if (z == B.ns - 1) return bsface + (B.nq - 2) * (B.nr - 2);
        }
        return -1;
    };
    const real ihq = (real)(B.nq - 1), ihr = (real)(B.nr - 1), ihs = (real)(B.ns - 1);
    // q-face buffer: an own column on a q face also writes its stage output
    // there (byte offset of its node at plane 0; -1 elsewhere)
    const unsigned FPqb = (unsigned)(B.nr * B.ns) * RS;
    const __amdgpu_buffer_rsrc_t rQ = brsrc(B.FQ ? (const void *)B.FQ : (const void *)B.Fin, B.FQ ? 18u * FPqb : 0u);
    int fqo = -1;
    if (B.FQ && own && (gx == 0 || gx == B.nq - 1)) fqo = (int)((gx == 0 ? 0u : 9u * FPqb) + (unsigned)gy * RS);

    const __amdgpu_buffer_rsrc_t rF = brsrc(B.Fin, 9u * NPb);
    const __amdgpu_buffer_rsrc_t rD = brsrc(B.DF, 9u * NPb);
    const __amdgpu_buffer_rsrc_t rO = brsrc(B.Fout, 9u * NPb);
    const __amdgpu_buffer_rsrc_t rG = brsrc(CURV ? (const void *)B.G : (const void *)B.Fin, CURV ? 10u * NPb : 0u);
    const __amdgpu_buffer_rsrc_t rM = brsrc(HET ? (const void *)B.M : (const void *)B.Fin, HET ? 3u * NPb : 0u);
    auto G = [&](unsigned o, unsigned c) -> real { return CURV ? bload<real>(rG, o, c * NPb) : (real)B.gc[c]; };

    // halo node of this thread (x halo: 6*TY nodes, then y halo: 6*TX nodes)
    const int tid = ly * TX + lx;
    sbd_to_lds<real>(sbd, tid, TX * TY);  // published by the first barrier of the march
    int hk = 0, hcx = 0, hcy = 0;
    unsigned hcolb = 0;
    if (tid < 6 * TY) {
        int hx = tid % 6, hy = tid / 6;
        hcx = hx < 3 ? hx : TX + hx;
        hcy = hy;
        int ggx = xs - 3 + hcx, ggy = ys + hy;
        if (ggx >= 0 && ggx < B.nq && ggy < B.nr) { hk = 1; hcolb = (unsigned)(ggx + B.px * ggy) * RS; }
    } else if (tid < 6 * TY + 6 * TX) {
        int t2 = tid - 6 * TY;
        int hx = t2 % TX, hy = t2 / TX;
        hcy = hy < 3 ? hy : TY + hy;
        hcx = hx;
        int ggx = xs + hx, ggy = ys - 3 + hcy;
        if (ggx < B.nq && ggy >= 0 && ggy < B.nr) { hk = 2; hcolb = (unsigned)(ggx + B.px * ggy) * RS; }
    }
    const unsigned hax = hk == 2 ? 1u : 0u;

    auto fluxv = [&](const real sg[6], real J, real ax, real ay, real az, real fl[3]) {
        real Jx = J * ax, Jy = J * ay, Jz = J * az;
        fl[0] = Jx * sg[0] + Jy * sg[3] + Jz * sg[4];
        fl[1] = Jx * sg[3] + Jy * sg[1] + Jz * sg[5];
        fl[2] = Jx * sg[4] + Jy * sg[5] + Jz * sg[2];
    };
    const real c1 = (real)0.75, c2 = (real)(-0.15), c3 = (real)(1.0 / 60.0);
    // SB = 0: interior chunk, s-queue of 7 planes; SB = 1: the six closure
    // planes at an s face (low if the chunk starts at plane 0), with all 9
    // planes they reach held in registers (planes 0..8, resp. ns-9..ns-1)
    const bool slow = z0 == 0;
    const int sbase = slow ? 0 : B.ns - 9;
    auto wrow = [&](int z) -> int { return slow ? z : 6 + z - (B.ns - 6); };
    auto sw = [&](int r, int p) -> real {
        if constexpr (sizeof(real) == 4) return c_sbw32[r][p];
        else return c_sbw64[r][p];
    };

    if (role == 0) {
        // ---------------- velocity rates: div of the stress fluxes / (rho J)
        struct PF {
            real s[6], g[7];   // centre stress; J, q_x q_y q_z, r_x r_y r_z
            real v[3], df[3];  // centre velocity and rates (update)
            real hs[6], hg[4]; // halo stress; J + axis metrics
            real ls[6], lg[4]; // lead (z+3) stress; J + s metrics
        };
        auto prefetch = [&](PF &R, int z) {
            if (inb) {
                const unsigned o = colb + planeb * (unsigned)z;
#pragma unroll
                for (int f = 0; f < 6; f++) R.s[f] = bload<real>(rF, o, (3 + f) * NPb);
                R.g[0] = G(o, 9); R.g[1] = G(o, 0); R.g[2] = G(o, 3); R.g[3] = G(o, 6);
                R.g[4] = G(o, 1); R.g[5] = G(o, 4); R.g[6] = G(o, 7);
                if (SB == 0) {
                    const unsigned ol = o + 3u * planeb;
#pragma unroll
                    for (int f = 0; f < 6; f++) R.ls[f] = bload<real>(rF, ol, (3 + f) * NPb);
                    R.lg[0] = G(ol, 9); R.lg[1] = G(ol, 2); R.lg[2] = G(ol, 5); R.lg[3] = G(ol, 8);
                }
            }
            if (own) {
                const unsigned o = colb + planeb * (unsigned)z;
#pragma unroll
                for (int f = 0; f < 3; f++) { R.v[f] = bload<real>(rF, o, f * NPb); R.df[f] = rdA ? bload<real>(rD, o, f * NPb) : (real)0; }
            }
            if (hk) {
                const unsigned o = hcolb + planeb * (unsigned)z;
#pragma unroll
                for (int f = 0; f < 6; f++) R.hs[f] = bload<real>(rF, o, (3 + f) * NPb);
                R.hg[0] = G(o, 9); R.hg[1] = G(o, 0 + hax); R.hg[2] = G(o, 3 + hax); R.hg[3] = G(o, 6 + hax);
            }
        };
        constexpr int NQF = SB == 0 ? 7 : 9;
        real qf[NQF][3];
#pragma unroll
        for (int a = 0; a < NQF; a++)
#pragma unroll
            for (int f = 0; f < 3; f++) qf[a][f] = 0;
        if (SB != 0 && inb) {
            // every plane the closure rows reach, loads all in flight together
#pragma unroll
            for (int a = 0; a < 9; a++) {
                const unsigned o = colb + planeb * (unsigned)(sbase + a);
                real sg[6];
#pragma unroll
                for (int f = 0; f < 6; f++) sg[f] = bload<real>(rF, o, (3 + f) * NPb);
                fluxv(sg, G(o, 9), G(o, 2), G(o, 5), G(o, 8), qf[a]);
            }
        }
        if (SB == 0 && inb) {
            // warm the queue with planes z0-3 .. z0+2, one plane at a time
#pragma unroll 1
            for (int a = 1; a < 7; a++) {
                const unsigned o = colb + planeb * (unsigned)(z0 - 4 + a);
                real sg[6];
#pragma unroll
                for (int f = 0; f < 6; f++) sg[f] = bload<real>(rF, o, (3 + f) * NPb);
#pragma unroll
                for (int q = 0; q < 6; q++)
#pragma unroll
                    for (int f = 0; f < 3; f++) qf[q][f] = qf[q + 1][f];
                fluxv(sg, G(o, 9), G(o, 2), G(o, 5), G(o, 8), qf[NQF - 1]);
            }
        }
        auto body = [&](PF &cur, PF &nxt, int z) {
            const int b = z & 1;
            real cv[3], cdf[3], J = 0;
            if (inb) {
                real fq[3], fr[3];
                J = cur.g[0];
                fluxv(cur.s, J, cur.g[1], cur.g[2], cur.g[3], fq);
                fluxv(cur.s, J, cur.g[4], cur.g[5], cur.g[6], fr);
#pragma unroll
                for (int f = 0; f < 3; f++) { sfq[b][f][ly][lx + 3] = fq[f]; sfr[b][f][ly + 3][lx] = fr[f]; }
                if (SB == 0) {
#pragma unroll
                    for (int a = 0; a < NQF - 1; a++)
#pragma unroll
                        for (int f = 0; f < 3; f++) qf[a][f] = qf[a + 1][f];
                    fluxv(cur.ls, cur.lg[0], cur.lg[1], cur.lg[2], cur.lg[3], qf[NQF - 1]);
                }
            }
            if (own) {
#pragma unroll
                for (int f = 0; f < 3; f++) { cv[f] = cur.v[f]; cdf[f] = cur.df[f]; }
            }
            if (hk) {
                real fl[3];
                fluxv(cur.hs, cur.hg[0], cur.hg[1], cur.hg[2], cur.hg[3], fl);
                if (hk == 1) {
#pragma unroll
                    for (int f = 0; f < 3; f++) sfq[b][f][hcy][hcx] = fl[f];
                } else {
#pragma unroll
                    for (int f = 0; f < 3; f++) sfr[b][f][hcy][hcx] = fl[f];
                }
            }
            __syncthreads();
            // unconditional: past the chunk's last plane this reads plane z1
            // (and z1+3 for the lead), which lie inside the block because
            // the march stops six planes short of the s faces; buffer loads
            // return 0 for anything beyond the arrays
            prefetch(nxt, z + 1);
            if (own) {
                const unsigned o = colb + planeb * (unsigned)z;
                const int bi = bnd(z);
                real rho = HET ? bload<real>(rM, o, 2u * NPb) : (real)B.rho;
                real rhoJ_inv = (real)1 / (rho * J);
#pragma unroll
                for (int f = 0; f < 3; f++) {
                    real a = d1_dense<real>(sbd, &sfq[b][f][ly][0], 1, lx + 3, gx, B.nq, ihq);
                    real bb = d1_dense<real>(sbd, &sfr[b][f][0][lx], TX, ly + 3, gy, B.nr, ihr);
                    real c;
                    if constexpr (SB == 0) {
                        c = ihs * (c1 * (qf[4][f] - qf[2][f]) + c2 * (qf[5][f] - qf[1][f]) + c3 * (qf[6][f] - qf[0][f]));
                    } else {
                        const int r = wrow(z);
                        real acc = 0;
#pragma unroll
                        for (int q = 0; q < 9; q++) acc = acc + sw(r, q) * qf[q][f];
                        c = ihs * acc;
                    }
                    real df = A * cdf[f] + (a + bb + c) * rhoJ_inv;
                    if (bi >= 0) df = df + B.satb[(long)f * B.nbnd + bi];
                    bstore<real>(rD, o, f * NPb, df);
                    const real vo = cv[f] + bdt * df;
                    bstore<real>(rO, o, f * NPb, vo);
                    if (fqo >= 0) bstore<real>(rQ, (unsigned)fqo + (unsigned)(B.nr * z) * RS, f * FPqb, vo);
                }
            }
        };
        PF pa, pb;
        prefetch(pa, z0);
        int z = z0;
        for (; z + 1 < z1; z += 2) {
            body(pa, pb, z);
            body(pb, pa, z + 1);
        }
        if (z < z1) body(pa, pb, z);
    } else {
        // ---------------- stress rates: Hooke's law on the velocity gradient
        struct PF {
            real v[3];          // centre velocity
            real g[9];          // centre metrics (no J)
            real s[6], df[6];   // centre stress and rates (update)
            real hv[3], lv[3];  // halo and lead (z+3) velocity
        };
        auto prefetch = [&](PF &R, int z) {
            if (inb) {
                const unsigned o = colb + planeb * (unsigned)z;
#pragma unroll
                for (int f = 0; f < 3; f++) R.v[f] = bload<real>(rF, o, f * NPb);
                if (SB == 0) {
                    const unsigned ol = o + 3u * planeb;
#pragma unroll
                    for (int f = 0; f < 3; f++) R.lv[f] = bload<real>(rF, ol, f * NPb);
                }
            }
            if (own) {
                const unsigned o = colb + planeb * (unsigned)z;
#pragma unroll
                for (int c = 0; c < 9; c++) R.g[c] = G(o, c);
#pragma unroll
                for (int f = 0; f < 6; f++) {
                    R.s[f] = bload<real>(rF, o, (3 + f) * NPb);
                    R.df[f] = rdA ? bload<real>(rD, o, (3 + f) * NPb) : (real)0;
                }
            }
            if (hk) {
                const unsigned o = hcolb + planeb * (unsigned)z;
#pragma unroll
                for (int f = 0; f < 3; f++) R.hv[f] = bload<real>(rF, o, f * NPb);
            }
        };
        constexpr int NQV = SB == 0 ? 7 : 9;
        real qv[NQV][3];
#pragma unroll
        for (int a = 0; a < NQV; a++)
#pragma unroll
            for (int f = 0; f < 3; f++) qv[a][f] = 0;
        if (SB != 0 && inb) {
#pragma unroll
            for (int a = 0; a < 9; a++) {
                const unsigned o = colb + planeb * (unsigned)(sbase + a);
#pragma unroll
                for (int f = 0; f < 3; f++) qv[a][f] = bload<real>(rF, o, f * NPb);
            }
        }
        if (SB == 0 && inb) {
#pragma unroll 1
            for (int a = 1; a < 7; a++) {
                const unsigned o = colb + planeb * (unsigned)(z0 - 4 + a);
#pragma unroll
                for (int q = 0; q < NQV - 1; q++)
#pragma unroll
                    for (int f = 0; f < 3; f++) qv[q][f] = qv[q + 1][f];
#pragma unroll
                for (int f = 0; f < 3; f++) qv[NQV - 1][f] = bload<real>(rF, o, f * NPb);
            }
        }
        auto body = [&](PF &cur, PF &nxt, int z) {
            const int b = z & 1;
            real cg[9], cs[6], cdf[6];
            if (inb) {
#pragma unroll
                for (int f = 0; f < 3; f++) sv[b][f][ly + 3][lx + 3] = cur.v[f];
                if (SB == 0) {
#pragma unroll
                    for (int a = 0; a < NQV - 1; a++)
#pragma unroll
                        for (int f = 0; f < 3; f++) qv[a][f] = qv[a + 1][f];
#pragma unroll
                    for (int f = 0; f < 3; f++) qv[NQV - 1][f] = cur.lv[f];
                }
            }
            if (own) {
#pragma unroll
                for (int c = 0; c < 9; c++) cg[c] = cur.g[c];
#pragma unroll
                for (int f = 0; f < 6; f++) { cs[f] = cur.s[f]; cdf[f] = cur.df[f]; }
            }
            if (hk) {
                if (hk == 1) {
#pragma unroll
                    for (int f = 0; f < 3; f++) sv[b][f][hcy + 3][hcx] = cur.hv[f];
                } else {
#pragma unroll
                    for (int f = 0; f < 3; f++) sv[b][f][hcy][hcx + 3] = cur.hv[f];
                }
            }
            __syncthreads();
            // unconditional: past the chunk's last plane this reads plane z1
            // (and z1+3 for the lead), which lie inside the block because
            // the march stops six planes short of the s faces; buffer loads
            // return 0 for anything beyond the arrays
            prefetch(nxt, z + 1);
            if (own) {
                const unsigned o = colb + planeb * (unsigned)z;
                real Ux[3], Uy[3], Uz[3];
#pragma unroll
                for (int f = 0; f < 3; f++) {
                    real dq = d1_dense<real>(sbd, &sv[b][f][ly + 3][0], 1, lx + 3, gx, B.nq, ihq);
                    real dr = d1_dense<real>(sbd, &sv[b][f][0][lx + 3], WX, ly + 3, gy, B.nr, ihr);
                    real ds;
                    if constexpr (SB == 0) {
                        ds = ihs * (c1 * (qv[4][f] - qv[2][f]) + c2 * (qv[5][f] - qv[1][f]) + c3 * (qv[6][f] - qv[0][f]));
                    } else {
                        const int r = wrow(z);
                        real acc = 0;
#pragma unroll
                        for (int q = 0; q < 9; q++) acc = acc + sw(r, q) * qv[q][f];
                        ds = ihs * acc;
                    }
                    Ux[f] = cg[0] * dq + cg[1] * dr + cg[2] * ds;
                    Uy[f] = cg[3] * dq + cg[4] * dr + cg[5] * ds;
                    Uz[f] = cg[6] * dq + cg[7] * dr + cg[8] * ds;
                }
                real lam, mu;
                if (HET) { lam = bload<real>(rM, o, 0u); mu = bload<real>(rM, o, NPb); }
                else { lam = (real)B.lam; mu = (real)B.mu; }
                real l2m = lam + (real)2 * mu;
                real DU[6];
                DU[0] = l2m * Ux[0] + lam * (Uy[1] + Uz[2]);
                DU[1] = l2m * Uy[1] + lam * (Ux[0] + Uz[2]);
                DU[2] = l2m * Uz[2] + lam * (Ux[0] + Uy[1]);
                DU[3] = mu * (Uy[0] + Ux[1]);
                DU[4] = mu * (Uz[0] + Ux[2]);
                DU[5] = mu * (Uz[1] + Uy[2]);
                const int bi = bnd(z);
#pragma unroll
                for (int f = 0; f < 6; f++) {
                    real df = A * cdf[f] + DU[f];
                    if (bi >= 0) df = df + B.satb[(long)(3 + f) * B.nbnd + bi];
                    bstore<real>(rD, o, (3 + f) * NPb, df);
                    const real so = cs[f] + bdt * df;
                    bstore<real>(rO, o, (3 + f) * NPb, so);
                    if (fqo >= 0) bstore<real>(rQ, (unsigned)fqo + (unsigned)(B.nr * z) * RS, (3 + f) * FPqb, so);
                }
            }
        };
        PF pa, pb;
        prefetch(pa, z0);
        int z = z0;
        for (; z + 1 < z1; z += 2) {
            body(pa, pb, z);
            body(pb, pa, z + 1);
        }
        if (z < z1) body(pa, pb, z);
    }
}

// ----------------------------------------------------------------------------
// Fused single-pass march.  One workgroup per tile and s-chunk forms the
// rates of all nine fields, and every value of the stage input is fetched
// once per chunk: plane p enters as the lead of the s-stencil (stresses, J
// and the s metrics -> s-flux queue; velocities -> velocity queue) and, with
// HOLD, its stresses and J wait in a four-plane register queue until p is
// the centre plane three iterations later, so the centre plane only adds
// its q/r metrics and rates.  The q/r stencils read LDS tiles of the centre
// plane's q-fluxes, r-fluxes and velocities (3-node halo), double buffered
// with one barrier per plane.  The chunk's warm-up (lead planes z0-3 ..
// z0+2) issues all of its loads together.  Arithmetic per node is the
// role-split kernel's, in the same order.
// Occupancy asked of the register allocator: FP32 re-reading the centre
// plane fits three waves per SIMD (<= 168 VGPRs), holding it needs two; FP64
// runs one wave per SIMD with AGPRs (two would spill heavily).
// With the anelastic response (AN) FP32 asks for two waves per SIMD: the
// memory-variable update adds its loads to the live prefetch buffers.
template <typename real, bool HOLD, bool AN>
constexpr int fused_min_waves() { return sizeof(real) == 4 ? (HOLD || AN ? 2 : 3) : 1; }

template <typename real, int TX, int TY, bool CURV, bool HET, bool HOLD, bool AN>
__global__ void __launch_bounds__(TX * TY, (fused_min_waves<real, HOLD, AN>()))
    k_volume_fused(Pair<real> P, TileList TL, real A0, real A1, real bdt)
{
    constexpr int WX = TX + 6, WY = TY + 6;
    constexpr unsigned RS = sizeof(real);
    __shared__ real sfq[2][3][TY][WX];
    __shared__ real sfr[2][3][WY][TX];
    __shared__ real sv[2][3][WY][WX];
    __shared__ SbpDense<real> sbd;

    // consecutive logical tiles on one XCD
    const long nwg = (long)gridDim.x;
    const long per = (nwg + 7) / 8;
    const long logical = (long)(blockIdx.x % 8) * per + (long)(blockIdx.x / 8);
    if (logical >= TL.total) return;
    int gi = 0;
    for (int k = 1; k < TL.n; k++)
        if (logical >= TL.g[k].first) gi = k;
    const TileGeom &T = TL.g[gi];
    const Blk<real> &B = P.b[T.blk];
    const real A = T.blk == 0 ? A0 : A1;
    // the first RK stage has A = 0: its rates are not read (A * DF = 0)
    const bool rdA = A != (real)0;
    long lid = logical - T.first;
    int tx_i, ty_i, zc_i;
    tile_coords(TL, T, lid, T.nzc, tx_i, ty_i, zc_i);

    int xs, xc0, xc1, ys, yc0, yc1;
    tile_range<TX>(tx_i, T.ntx, T.xa0, T.xa1, B.nq, xs, xc0, xc1);
    tile_range<TY>(ty_i, T.nty, T.ya0, T.ya1, B.nr, ys, yc0, yc1);
    // chunks of balanced length (lengths differ by at most one plane)
    const int span = T.za1 - T.za0;
    const int z0 = T.za0 + (int)((long)zc_i * span / T.nzc);
    const int z1 = T.za0 + (int)((long)(zc_i + 1) * span / T.nzc);
    // march direction: the lower chunk of a pair (tile order 3) runs down
    // from its top plane, every other chunk up from its bottom plane; the
    // queues hold planes in march order, so the s-derivative takes the sign
    // of the direction (an exact negation)
    const int dir = (TL.order == 3 && (zc_i & 1) == 0 && zc_i + 1 < T.nzc) ? -1 : 1;
    const int s0 = dir > 0 ? z0 : z1 - 1;
    const real sgn = (real)dir;

    const int lx = threadIdx.x, ly = threadIdx.y;
    const int gx = xs + lx, gy = ys + ly;
    const bool own = (gx >= xc0 && gx < xc1 && gy >= yc0 && gy < yc1);
    const bool inb = (gx < B.nq && gy < B.nr);
    const unsigned NPb = (unsigned)B.NP * RS;
    const unsigned planeb = (unsigned)(B.px * B.nr) * RS;
    const unsigned colb = (unsigned)(gx + B.px * gy) * RS;
    // boundary-node index of a q/r face column (fused boundary penalties);
    // interior chunks hold no s-face node
    int bbase = -1, bstride = 0;
    if (B.satb && own) {
        const int nq = B.nq, nr = B.nr, ns = B.ns;
        const int nqf = 2 * nr * ns;
        if (gx == 0 || gx == nq - 1) { bbase = (gx == 0 ? 0 : nr * ns) + gy; bstride = nr; }
        else if (gy == 0 || gy == nr - 1) { bbase = nqf + (gy == 0 ? 0 : (nq - 2) * ns) + gx - 1; bstride = nq - 2; }
    }
    const real ihq = (real)(B.nq - 1), ihr = (real)(B.nr - 1), ihs = (real)(B.ns - 1);
    // q-face buffer: an own column on a q face also writes its stage output
    // there (byte offset of its node at plane 0; -1 elsewhere).  Compiled
    // out of the metric-holding variants, which run at the register limit.
    const unsigned FPqb = (unsigned)(B.nr * B.ns) * RS;
    const __amdgpu_buffer_rsrc_t rQ = brsrc(B.FQ ? (const void *)B.FQ : (const void *)B.Fin, B.FQ ? 18u * FPqb : 0u);
    int fqo = -1;
    if (!HOLD && B.FQ && own && (gx == 0 || gx == B.nq - 1))
        fqo = (int)((gx == 0 ? 0u : 9u * FPqb) + (unsigned)gy * RS);

    const __amdgpu_buffer_rsrc_t rF = brsrc(B.Fin, 9u * NPb);
    const __amdgpu_buffer_rsrc_t rD = brsrc(B.DF, 9u * NPb);
    const __amdgpu_buffer_rsrc_t rO = brsrc(B.Fout, 9u * NPb);
    const __amdgpu_buffer_rsrc_t rG = brsrc(CURV ? (const void *)B.G : (const void *)B.Fin, CURV ? 10u * NPb : 0u);
    const __amdgpu_buffer_rsrc_t rM = brsrc(HET ? (const void *)B.M : (const void *)B.Fin, HET ? 3u * NPb : 0u);
    // anelastic memory variables, their rates and Q_s^-1, Q_p^-1 (AN only)
    const __amdgpu_buffer_rsrc_t rE = brsrc(AN ? (const void *)B.eta : (const void *)B.Fin, AN ? 24u * NPb : 0u);
    const __amdgpu_buffer_rsrc_t rDE = brsrc(AN ? (const void *)B.deta : (const void *)B.Fin, AN ? 24u * NPb : 0u);
    const __amdgpu_buffer_rsrc_t rQi = brsrc(AN ? (const void *)B.qinv : (const void *)B.Fin, AN ? 2u * NPb : 0u);
    // metric c of the node at byte offset o.  A field that is constant over
    // the block (bit gconst) is loaded through a zero-length buffer (0, and no
    // memory access); K(v, c) then selects its constant.  K is applied where
    // a value is consumed, not at the prefetch, so that the loads of the next
    // plane stay in flight under the current plane's work.  The metric-holding
    // variants, at the register limit, load every field (G holds the
    // constants' values too).
    constexpr bool KF = CURV && !HOLD;
    const __amdgpu_buffer_rsrc_t rG0 = brsrc(B.Fin, 0u);
    auto G = [&](unsigned o, unsigned c) -> real {
        if constexpr (KF)
            return bload<real>(((B.gconst >> c) & 1u) ? rG0 : rG, o, c * NPb);
        else if constexpr (CURV)
            return bload<real>(rG, o, c * NPb);
        else
            return (real)B.gc[c];
    };
    auto K = [&](real v, unsigned c) -> real {
        if constexpr (KF)
            return ((B.gconst >> c) & 1u) ? (real)B.gc[c] : v;
        else
            return v;
    };
    // a field index that differs between lanes (the halo threads' flux
    // axis): a plain load
    auto GH = [&](unsigned o, unsigned c) -> real { return CURV ? bload<real>(rG, o, c * NPb) : (real)B.gc[c]; };

    // halo node of this thread (x halo: 6*TY nodes, then y halo: 6*TX nodes)
    const int tid = ly * TX + lx;
    sbd_to_lds<real>(sbd, tid, TX * TY);  // published by the first barrier of the march
    int hk = 0, hcx = 0, hcy = 0;
    unsigned hcolb = 0;
    if (tid < 6 * TY) {
        int hx = tid % 6, hy = tid / 6;
        hcx = hx < 3 ? hx : TX + hx;
        hcy = hy;
        int ggx = xs - 3 + hcx, ggy = ys + hy;
        if (ggx >= 0 && ggx < B.nq && ggy < B.nr) { hk = 1; hcolb = (unsigned)(ggx + B.px * ggy) * RS; }
    } else if (tid < 6 * TY + 6 * TX) {
        int t2 = tid - 6 * TY;
        int hx = t2 % TX, hy = t2 / TX;
        hcy = hy < 3 ? hy : TY + hy;
        hcx = hx;
        int ggx = xs + hx, ggy = ys - 3 + hcy;
        if (ggx < B.nq && ggy >= 0 && ggy < B.nr) { hk = 2; hcolb = (unsigned)(ggx + B.px * ggy) * RS; }
    }
    const unsigned hax = hk == 2 ? 1u : 0u;

    auto fluxv = [&](const real sg[6], real J, real ax, real ay, real az, real fl[3]) {
        real Jx = J * ax, Jy = J * ay, Jz = J * az;
        fl[0] = Jx * sg[0] + Jy * sg[3] + Jz * sg[4];
        fl[1] = Jx * sg[3] + Jy * sg[1] + Jz * sg[5];
        fl[2] = Jx * sg[4] + Jy * sg[5] + Jz * sg[2];
    };
    const real c1 = (real)0.75, c2 = (real)(-0.15), c3 = (real)(1.0 / 60.0);

    struct PF {
        real ls[6], lg[4], lv[3];  // lead plane c+3: stresses; J, s_x s_y s_z; velocities
        real cs[6], cJ;            // centre stresses and J (loaded when not held)
        real gm[6], gs[3];         // centre q_x q_y q_z r_x r_y r_z; s_x s_y s_z (not held)
        real df[9], mat[3];        // centre rates; lambda mu rho (heterogeneous)
        real hs[6], hg[4], hv[3];  // halo node: stresses; J + axis metrics; velocities
    };
    auto prefetch = [&](PF &R, int c) {
        if (inb) {
            const unsigned ol = colb + planeb * (unsigned)(c + 3 * dir);
#pragma unroll
            for (int f = 0; f < 6; f++) R.ls[f] = bload<real>(rF, ol, (3 + f) * NPb);
            R.lg[0] = G(ol, 9); R.lg[1] = G(ol, 2); R.lg[2] = G(ol, 5); R.lg[3] = G(ol, 8);
#pragma unroll
            for (int f = 0; f < 3; f++) R.lv[f] = bload<real>(rF, ol, f * NPb);
            const unsigned o = colb + planeb * (unsigned)c;
            if (!HOLD) {
#pragma unroll
                for (int f = 0; f < 6; f++) R.cs[f] = bload<real>(rF, o, (3 + f) * NPb);
                R.cJ = G(o, 9);
            }
            R.gm[0] = G(o, 0); R.gm[1] = G(o, 3); R.gm[2] = G(o, 6);
            R.gm[3] = G(o, 1); R.gm[4] = G(o, 4); R.gm[5] = G(o, 7);
            if (!HOLD) { R.gs[0] = G(o, 2); R.gs[1] = G(o, 5); R.gs[2] = G(o, 8); }
        }
        if (own) {
            const unsigned o = colb + planeb * (unsigned)c;
#pragma unroll
            for (int f = 0; f < 9; f++) R.df[f] = rdA ? bload_nt<real>(rD, o, f * NPb) : (real)0;
            if (HET) {
#pragma unroll
                for (int f = 0; f < 3; f++) R.mat[f] = bload<real>(rM, o, f * NPb);
            }
        }
        if (hk) {
            const unsigned o = hcolb + planeb * (unsigned)c;
#pragma unroll
            for (int f = 0; f < 6; f++) R.hs[f] = bload<real>(rF, o, (3 + f) * NPb);
            R.hg[0] = G(o, 9); R.hg[1] = GH(o, 0 + hax); R.hg[2] = GH(o, 3 + hax); R.hg[3] = GH(o, 6 + hax);
#pragma unroll
            for (int f = 0; f < 3; f++) R.hv[f] = bload<real>(rF, o, f * NPb);
        }
    };

    // queues: fs/vq hold planes c-3..c+3, sm/hq planes c..c+3 (with HOLD;
    // without it the centre plane re-reads its stresses, J and s-metrics,
    // which keeps the kernel within 168 VGPRs)
    real fs[7][3], vq[7][3], sm[HOLD ? 4 : 1][3], hq[HOLD ? 4 : 1][7];
#pragma unroll
    for (int a = 0; a < 7; a++)
#pragma unroll
        for (int f = 0; f < 3; f++) { fs[a][f] = 0; vq[a][f] = 0; }
#pragma unroll
    for (int a = 0; a < (HOLD ? 4 : 1); a++)
#pragma unroll
        for (int f = 0; f < 3; f++) sm[a][f] = 0;
#pragma unroll
    for (int a = 0; a < (HOLD ? 4 : 1); a++)
#pragma unroll
        for (int f = 0; f < 7; f++) hq[a][f] = 0;
    if (inb) {
        // warm-up: lead planes s0-3 .. s0+2 (march order), every load in
        // flight together
        real ws[6][6], wg[6][4], wv[6][3];
#pragma unroll
        for (int a = 0; a < 6; a++) {
            const unsigned o = colb + planeb * (unsigned)(s0 + dir * (a - 3));
#pragma unroll
            for (int f = 0; f < 6; f++) ws[a][f] = bload<real>(rF, o, (3 + f) * NPb);
            wg[a][0] = G(o, 9); wg[a][1] = G(o, 2); wg[a][2] = G(o, 5); wg[a][3] = G(o, 8);
#pragma unroll
            for (int f = 0; f < 3; f++) wv[a][f] = bload<real>(rF, o, f * NPb);
        }
#pragma unroll
        for (int a = 0; a < 6; a++) {
            wg[a][0] = K(wg[a][0], 9); wg[a][1] = K(wg[a][1], 2); wg[a][2] = K(wg[a][2], 5); wg[a][3] = K(wg[a][3], 8);
        }
#pragma unroll
        for (int a = 0; a < 6; a++) {
            fluxv(ws[a], wg[a][0], wg[a][1], wg[a][2], wg[a][3], fs[a + 1]);
#pragma unroll
            for (int f = 0; f < 3; f++) vq[a + 1][f] = wv[a][f];
        }
#pragma unroll
        for (int a = 3; a < 6; a++) {
            if constexpr (HOLD) {
#pragma unroll
                for (int f = 0; f < 3; f++) sm[a - 2][f] = wg[a][1 + f];
#pragma unroll
                for (int f = 0; f < 6; f++) hq[a - 2][f] = ws[a][f];
                hq[a - 2][6] = wg[a][0];
            }
        }
    }

    auto body = [&](PF &cur, PF &nxt, int c) {
        const int b = c & 1;
        // plane c+3 enters the queues
#pragma unroll
        for (int a = 0; a < 6; a++)
#pragma unroll
            for (int f = 0; f < 3; f++) { fs[a][f] = fs[a + 1][f]; vq[a][f] = vq[a + 1][f]; }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if constexpr (HOLD) {
#pragma unroll
                for (int f = 0; f < 3; f++) sm[a][f] = sm[a + 1][f];
#pragma unroll
                for (int f = 0; f < 7; f++) hq[a][f] = hq[a + 1][f];
            }
        }
        real sc[6], Jc = 0;
        if (inb) {
            const real lg[4] = {K(cur.lg[0], 9), K(cur.lg[1], 2), K(cur.lg[2], 5), K(cur.lg[3], 8)};
            fluxv(cur.ls, lg[0], lg[1], lg[2], lg[3], fs[6]);
#pragma unroll
            for (int f = 0; f < 3; f++) vq[6][f] = cur.lv[f];
            if constexpr (HOLD) {
#pragma unroll
                for (int f = 0; f < 3; f++) sm[3][f] = lg[1 + f];
#pragma unroll
                for (int f = 0; f < 6; f++) hq[3][f] = cur.ls[f];
                hq[3][6] = lg[0];
            }
#pragma unroll
            for (int f = 0; f < 6; f++) sc[f] = HOLD ? hq[0][f] : cur.cs[f];
            Jc = HOLD ? hq[0][6] : K(cur.cJ, 9);
            real fq[3], fr[3];
            fluxv(sc, Jc, K(cur.gm[0], 0), K(cur.gm[1], 3), K(cur.gm[2], 6), fq);
            fluxv(sc, Jc, K(cur.gm[3], 1), K(cur.gm[4], 4), K(cur.gm[5], 7), fr);
#pragma unroll
            for (int f = 0; f < 3; f++) {
                sfq[b][f][ly][lx + 3] = fq[f];
                sfr[b][f][ly + 3][lx] = fr[f];
                sv[b][f][ly + 3][lx + 3] = vq[3][f];
            }
        }
        real cdf[9], cgm[6], csm[3], cmat[3];
        if (own) {
#pragma unroll
            for (int f = 0; f < 9; f++) cdf[f] = cur.df[f];
#pragma unroll
            for (int f = 0; f < 6; f++) cgm[f] = K(cur.gm[f], f < 3 ? 3 * f : 3 * (f - 3) + 1);
#pragma unroll
            for (int f = 0; f < 3; f++) csm[f] = HOLD ? sm[0][f] : K(cur.gs[f], 3 * f + 2);
            if (HET) {
#pragma unroll
                for (int f = 0; f < 3; f++) cmat[f] = cur.mat[f];
            }
        }
        if (hk) {
            real fl[3];
            fluxv(cur.hs, K(cur.hg[0], 9), cur.hg[1], cur.hg[2], cur.hg[3], fl);
            if (hk == 1) {
#pragma unroll
                for (int f = 0; f < 3; f++) { sfq[b][f][hcy][hcx] = fl[f]; sv[b][f][hcy + 3][hcx] = cur.hv[f]; }
            } else {
#pragma unroll
                for (int f = 0; f < 3; f++) { sfr[b][f][hcy][hcx] = fl[f]; sv[b][f][hcy][hcx + 3] = cur.hv[f]; }
            }
        }
        __syncthreads();
        // unconditional: past the chunk's last plane this reads one plane
        // beyond the chunk (and three more for the lead), which lie inside
        // the block because the march stops six planes short of the s faces;
        // buffer loads return 0 for anything beyond the arrays
        prefetch(nxt, c + dir);
        if (own) {
            const unsigned o = colb + planeb * (unsigned)c;
            const int bi = bbase >= 0 ? bbase + bstride * c : -1;
            // velocity rates: divergence of the stress fluxes / (rho J)
            real rho = HET ? cmat[2] : (real)B.rho;
            real rhoJ_inv = (real)1 / (rho * Jc);
#pragma unroll
            for (int f = 0; f < 3; f++) {
                real a = d1_dense<real>(sbd, &sfq[b][f][ly][0], 1, lx + 3, gx, B.nq, ihq);
                real bb = d1_dense<real>(sbd, &sfr[b][f][0][lx], TX, ly + 3, gy, B.nr, ihr);
                real cc = sgn * (ihs * (c1 * (fs[4][f] - fs[2][f]) + c2 * (fs[5][f] - fs[1][f]) + c3 * (fs[6][f] - fs[0][f])));
                real df = A * cdf[f] + (a + bb + cc) * rhoJ_inv;
                if (bi >= 0) df = df + B.satb[(long)f * B.nbnd + bi];
                bstore_nt<real>(rD, o, f * NPb, df);
                const real vo = vq[3][f] + bdt * df;
                bstore_nt<real>(rO, o, f * NPb, vo);
                if (fqo >= 0) bstore<real>(rQ, (unsigned)fqo + (unsigned)(B.nr * c) * RS, f * FPqb, vo);
            }
            // stress rates: Hooke's law on the velocity gradient
            const real sx = CURV ? csm[0] : (real)B.gc[2];
            const real sy = CURV ? csm[1] : (real)B.gc[5];
            const real sz = CURV ? csm[2] : (real)B.gc[8];
            real Ux[3], Uy[3], Uz[3];
#pragma unroll
            for (int f = 0; f < 3; f++) {
                real dq = d1_dense<real>(sbd, &sv[b][f][ly + 3][0], 1, lx + 3, gx, B.nq, ihq);
                real dr = d1_dense<real>(sbd, &sv[b][f][0][lx + 3], WX, ly + 3, gy, B.nr, ihr);
                real ds = sgn * (ihs * (c1 * (vq[4][f] - vq[2][f]) + c2 * (vq[5][f] - vq[1][f]) + c3 * (vq[6][f] - vq[0][f])));
                Ux[f] = cgm[0] * dq + cgm[3] * dr + sx * ds;
                Uy[f] = cgm[1] * dq + cgm[4] * dr + sy * ds;
                Uz[f] = cgm[2] * dq + cgm[5] * dr + sz * ds;
            }
            real lam, mu;
            if (HET) { lam = cmat[0]; mu = cmat[1]; }
            else { lam = (real)B.lam; mu = (real)B.mu; }
            real l2m = lam + (real)2 * mu;
            real DU[6];
            DU[0] = l2m * Ux[0] + lam * (Uy[1] + Uz[2]);
            DU[1] = l2m * Uy[1] + lam * (Ux[0] + Uz[2]);
            DU[2] = l2m * Uz[2] + lam * (Ux[0] + Uy[1]);
            DU[3] = mu * (Uy[0] + Ux[1]);
            DU[4] = mu * (Uz[0] + Ux[2]);
            DU[5] = mu * (Uz[1] + Uy[2]);
            if constexpr (AN) {
                // the anelastic memory variables of the node (apply_anelastic_point,
                // RHS_Interior.f90:2390-2441; the march covers no PML slab, so the
                // strains are the velocity gradient's): DU -= sum over the four
                // mechanisms, Deta = A Deta + rate, eta += B dt Deta in place,
                // one stress component at a time
                const real e[6] = {Ux[0], Uy[1], Uz[2], Uy[0] + Ux[1], Uz[0] + Ux[2], Uz[1] + Uy[2]};
                const real tr = e[0] + e[1] + e[2];
                const real Qsi = bload<real>(rQi, o, 0), Qpi = bload<real>(rQi, o, NPb);
                const real wa = (real)2 * mu * Qsi, wb = (lam + (real)2 * mu) * Qpi - (real)2 * mu * Qsi;
                const real wc = mu * Qsi;
#pragma unroll
                for (int f = 0; f < 6; f++) {
                    real et[4], de0[4];
#pragma unroll
                    for (int l = 0; l < 4; l++) {
                        et[l] = bload_nt<real>(rE, o, (unsigned)(4 * f + l) * NPb);
                        de0[l] = rdA ? bload_nt<real>(rDE, o, (unsigned)(4 * f + l) * NPb) : (real)0;
                    }
                    DU[f] = DU[f] - (et[0] + et[1] + et[2] + et[3]);
#pragma unroll
                    for (int l = 0; l < 4; l++) {
                        const real w = (real)B.an_w[l], tau = (real)B.an_tau[l];
                        const real rate = f < 3 ? (((w * wa) * e[f] + (w * wb) * tr) - et[l]) / tau
                                                : (w * wc * e[f] - et[l]) / tau;
                        const real de = A * de0[l] + rate;
                        bstore_nt<real>(rDE, o, (unsigned)(4 * f + l) * NPb, de);
                        bstore_nt<real>(rE, o, (unsigned)(4 * f + l) * NPb, et[l] + bdt * de);
                    }
                }
            }
#pragma unroll
            for (int f = 0; f < 6; f++) {
                real df = A * cdf[3 + f] + DU[f];
                if (bi >= 0) df = df + B.satb[(long)(3 + f) * B.nbnd + bi];
                bstore_nt<real>(rD, o, (3 + f) * NPb, df);
                const real so = sc[f] + bdt * df;
                bstore_nt<real>(rO, o, (3 + f) * NPb, so);
                if (fqo >= 0) bstore<real>(rQ, (unsigned)fqo + (unsigned)(B.nr * c) * RS, (3 + f) * FPqb, so);
            }
        }
    };
    PF pa, pb;
    prefetch(pa, s0);
    const int nz = z1 - z0;
    // one plane per iteration, not unrolled: unrolling by two keeps both
    // prefetch buffers and two copies of the queues live and roughly
    // doubles the register demand; the copies below are plain moves
#pragma unroll 1
    for (int m = 0; m < nz; m++) {
        body(pa, pb, s0 + dir * m);
        pa = pb;
    }
}

// ----------------------------------------------------------------------------
// volume: the upwind / upwind-DRP operators (and the anelastic response),
// tiled s-march
//
// A thread block owns a TX x TY tile of (q, r) columns and marches a chunk
// of the planes whose s row is the operator's interior row (band..ns-1-band,
// widened to npml on PML sides), over the q/r nodes outside the PML slabs.
// Per plane the velocities and the q- and r-fluxes
//     f_a,i = J (a_x s_xi + a_y s_yi + a_z s_zi),  a = q, r
// of the tile and an H-node halo sit in LDS, H the operator's largest reach
// (every closure row's taps lie within H nodes of the row, so no tile needs
// an alignment with the faces); the velocities and s-fluxes of planes
// c-HS..c+HS ride in a register queue.  A node applies its q and r rows (a
// closure row at a face, the interior row elsewhere) to LDS and the interior
// s row to the queue: the forward operator Dp to the velocities, the backward
// operator Dm to the fluxes, with the interior routine's rationals at nodes
// interior in all three directions and the band routine's printed row
// elsewhere (the forms of JJU_x<o>_interior_upwind and JJU_x<o>_upwind,
// JU_xJU_yJU_z6.f90:16451-20922, 815-16122; the linearity of the operator
// lets the three flux families of a direction be summed before the
// derivative).  The velocity gradient (chain rule with the node's metrics)
// and the three flux divergences go to node_tail (Hooke's law, anelastic
// terms, boundary penalty, stage update).  The s-band planes and the PML
// slabs run on k_volume_op.

// one row of an operator in LDS: 0..15 closure-left taps (node t), 16..31
// closure-right taps (node n-1-t), 32..47 the printed interior row, 48..63
// the interior routine's rationals (offsets lo..hi)
struct OptRows {
    int nl, nr, lo, hi;
};

// FP32 at three waves per SIMD (168 VGPRs and about 200 spilled bytes a
// lane): 18% shorter DRP-6 steps than at two; FP64 one wave
template <typename real, int TX, int TY, int H>
__global__ void __launch_bounds__(TX * TY, (sizeof(real) == 4 ? 3 : 1))
    k_volume_opt(Pair<real> P, TileList TL, real A0, real A1, real bdt)
{
    constexpr int HS = H < 5 ? H : 5;  // interior rows reach at most 5 planes
    constexpr int QL = 2 * HS + 1;
    constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
    constexpr int NT = TX * TY;
    __shared__ real sv[3][WY][WX];    // velocities (corners unused)
    __shared__ real sfq[3][TY][WX];   // q-fluxes
    __shared__ real sfr[3][WY][TX];   // r-fluxes
    __shared__ real tab[6][8][16];    // v q/r/s then s q/r/s: [0..7] left rows
    __shared__ real tabr[6][8][16];   // right rows
    __shared__ real tint[6][2][16];   // interior row: printed (0), rationals (1)
    __shared__ OptRows trows[6];

    const long nwg = (long)gridDim.x;
    const long per = (nwg + 7) / 8;
    const long logical = (long)(blockIdx.x % 8) * per + (long)(blockIdx.x / 8);
    const int tid = threadIdx.y * TX + threadIdx.x;
    if (logical >= TL.total) return;
    int gi = 0;
    for (int k = 1; k < TL.n; k++)
        if (logical >= TL.g[k].first) gi = k;
    const TileGeom &T = TL.g[gi];
    const Blk<real> &B = P.b[T.blk];
    const real A = T.blk == 0 ? A0 : A1;
    const long lid = logical - T.first;
    const int tx_i = (int)(lid % T.ntx), ty_i = (int)((lid / T.ntx) % T.nty), zc_i = (int)(lid / ((long)T.ntx * T.nty));

    // the operator's tables into LDS
    for (int e = tid; e < 6 * 8 * 16; e += NT) {
        const int h = e / 128, r = (e / 16) % 8, t = e % 16;
        tab[h][r][t] = (real)B.sbp[h].left[r][t];
        tabr[h][r][t] = (real)B.sbp[h].right[r][t];
    }
    for (int e = tid; e < 6 * 16; e += NT) {
        const int h = e / 16, t = e % 16;
        tint[h][0][t] = (real)B.sbp[h].band[t];
        tint[h][1][t] = (real)B.sbp[h].exact[t];
    }
    if (tid < 6) {
        trows[tid].nl = B.sbp[tid].nl; trows[tid].nr = B.sbp[tid].nr;
        trows[tid].lo = B.sbp[tid].lo; trows[tid].hi = B.sbp[tid].hi;
    }

    const int xs = T.xa0 + tx_i * TX, ys = T.ya0 + ty_i * TY;
    const int span = T.za1 - T.za0;
    const int z0 = T.za0 + (int)((long)zc_i * span / T.nzc);
    const int z1 = T.za0 + (int)((long)(zc_i + 1) * span / T.nzc);
    const int lx = threadIdx.x, ly = threadIdx.y;
    const int gx = xs + lx, gy = ys + ly;
    const bool own = gx < T.xa1 && gy < T.ya1;
    const long NP = B.NP;
    const int nq = B.nq, nr = B.nr, ns = B.ns;
    // interior in q and r (the tile's planes are interior in s)
    const int w = B.sbw;
    const int ix0 = B.pml[0] && B.npml > w ? B.npml : w, ix1 = B.pml[1] && B.npml > w ? B.npml : w;
    const int iy0 = B.pml[2] && B.npml > w ? B.npml : w, iy1 = B.pml[3] && B.npml > w ? B.npml : w;
    const bool interior = gx >= ix0 && gx <= nq - 1 - ix1 && gy >= iy0 && gy <= nr - 1 - iy1;
    const real ihq = (real)(nq - 1), ihr = (real)(nr - 1), ihs = (real)(ns - 1);

    auto met = [&](int c, long p) -> real { return (real)gmet(B, c, p); };
    auto flux = [&](long p, int a, real fl[3]) {  // J (a_x s_x. + a_y s_y. + a_z s_z.), a = 0 q, 1 r, 2 s
        const real J = met(9, p);
        const real jx = J * met(a, p), jy = J * met(3 + a, p), jz = J * met(6 + a, p);
        const real sxx = B.Fin[3 * NP + p], syy = B.Fin[4 * NP + p], szz = B.Fin[5 * NP + p];
        const real sxy = B.Fin[6 * NP + p], sxz = B.Fin[7 * NP + p], syz = B.Fin[8 * NP + p];
        fl[0] = jx * sxx + jy * sxy + jz * sxz;
        fl[1] = jx * sxy + jy * syy + jz * syz;
        fl[2] = jx * sxz + jy * syz + jz * szz;
    };
    // the queue: planes c-HS..c+HS of the own column
    real vq[QL][3], fs[QL][3];
#pragma unroll
    for (int a = 0; a < QL; a++)
#pragma unroll
        for (int f = 0; f < 3; f++) { vq[a][f] = 0; fs[a][f] = 0; }
    auto lead = [&](int plane, real v[3], real fl[3]) {
        const long p = pidx(B.px, nr, gx, gy, plane);
#pragma unroll
        for (int f = 0; f < 3; f++) v[f] = B.Fin[f * NP + p];
        flux(p, 2, fl);
    };
    if (own)
        for (int a = 0; a < QL - 1; a++) lead(z0 - HS + a, vq[a + 1], fs[a + 1]);
    __syncthreads();

    const int nwin = TX * TY + 2 * H * TY + 2 * H * TX;
    for (int c = z0; c < z1; c++) {
        // plane c+HS enters the queue
#pragma unroll
        for (int a = 0; a < QL - 1; a++)
#pragma unroll
            for (int f = 0; f < 3; f++) { vq[a][f] = vq[a + 1][f]; fs[a][f] = fs[a + 1][f]; }
        if (own) lead(c + HS, vq[QL - 1], fs[QL - 1]);
        // the plane's window: the tile, its q halo (TY rows), its r halo (TX columns)
        for (int e = tid; e < nwin; e += NT) {
            int wx, wy;
            if (e < TX * TY) { wx = H + e % TX; wy = H + e / TX; }
            else if (e < TX * TY + 2 * H * TY) {
                const int k = e - TX * TY;
                wx = k % (2 * H); wx = wx < H ? wx : TX + wx; wy = H + k / (2 * H);
            } else {
                const int k = e - TX * TY - 2 * H * TY;
                wx = H + k % TX; wy = k / TX; wy = wy < H ? wy : TY + wy;
            }
            const int qx = xs - H + wx, ry = ys - H + wy;
            if (qx < 0 || qx >= nq || ry < 0 || ry >= nr) continue;
            const long p = pidx(B.px, nr, qx, ry, c);
#pragma unroll
            for (int f = 0; f < 3; f++) sv[f][wy][wx] = B.Fin[f * NP + p];
            real fl[3];
            if (wy >= H && wy < H + TY) {
                flux(p, 0, fl);
#pragma unroll
                for (int f = 0; f < 3; f++) sfq[f][wy - H][wx] = fl[f];
            }
            if (wx >= H && wx < H + TX) {
                flux(p, 1, fl);
#pragma unroll
                for (int f = 0; f < 3; f++) sfr[f][wy][wx - H] = fl[f];
            }
        }
        __syncthreads();
        if (own) {
            // q and r: rows at the node's position from LDS
            real dv[3][3], df[3][3];  // [direction][field]
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const int pos = a == 0 ? gx : gy, n = a == 0 ? nq : nr;
                const int base = a == 0 ? xs - H : ys - H;   // window coordinate of node 0
#pragma unroll
                for (int grp = 0; grp < 2; grp++) {
                    const int h = grp * 3 + a;
                    const OptRows R = trows[h];
                    real s3[3] = {0, 0, 0};
                    auto tap = [&](int node, real cf) {
                        const int wpos = node - base;
#pragma unroll
                        for (int f = 0; f < 3; f++) {
                            real v;
                            if (grp == 0) v = a == 0 ? sv[f][H + ly][wpos] : sv[f][wpos][H + lx];
                            else v = a == 0 ? sfq[f][ly][wpos] : sfr[f][wpos][lx];
                            s3[f] = s3[f] + cf * v;
                        }
                    };
                    if (pos < R.nl) {
                        for (int t = 0; t < 16; t++) {
                            const real cf = tab[h][pos][t];
                            if (cf != (real)0) tap(t, cf);
                        }
                    } else if (pos >= n - R.nr) {
                        const int r = n - 1 - pos;
                        for (int t = 0; t < 16; t++) {
                            const real cf = tabr[h][r][t];
                            if (cf != (real)0) tap(n - 1 - t, cf);
                        }
                    } else {
                        const int set = interior ? 1 : 0;
                        for (int o = R.lo; o <= R.hi; o++) tap(pos + o, tint[h][set][o - R.lo]);
                    }
#pragma unroll
                    for (int f = 0; f < 3; f++) {
                        if (grp == 0) dv[a][f] = s3[f];
                        else df[a][f] = s3[f];
                    }
                }
            }
            // s: the interior row from the queue (plane c + o is queue slot HS + o)
#pragma unroll
            for (int grp = 0; grp < 2; grp++) {
                const int h = grp * 3 + 2;
                const OptRows R = trows[h];
                const int set = interior ? 1 : 0;
                real s3[3] = {0, 0, 0};
#pragma unroll
                for (int o = -HS; o <= HS; o++) {
                    if (o < R.lo || o > R.hi) continue;
                    const real cf = tint[h][set][o - R.lo];
#pragma unroll
                    for (int f = 0; f < 3; f++) s3[f] = s3[f] + cf * (grp == 0 ? vq[HS + o][f] : fs[HS + o][f]);
                }
#pragma unroll
                for (int f = 0; f < 3; f++) {
                    if (grp == 0) dv[2][f] = s3[f];
                    else df[2][f] = s3[f];
                }
            }
            const long p = pidx(B.px, nr, gx, gy, c);
            const real ih[3] = {ihq, ihr, ihs};
            real U[3][NF];
#pragma unroll
            for (int fam = 0; fam < 3; fam++) {
#pragma unroll
                for (int f = 0; f < 3; f++)
                    U[fam][f] = met(3 * fam, p) * (ih[0] * dv[0][f]) + met(3 * fam + 1, p) * (ih[1] * dv[1][f]) +
                                met(3 * fam + 2, p) * (ih[2] * dv[2][f]);
#pragma unroll
                for (int f = 3; f < NF; f++) U[fam][f] = 0;
            }
            // the flux divergences in the slots node_tail sums for the velocity rates
            U[0][3] = ih[0] * df[0][0] + ih[1] * df[1][0] + ih[2] * df[2][0];
            U[0][6] = ih[0] * df[0][1] + ih[1] * df[1][1] + ih[2] * df[2][1];
            U[0][7] = ih[0] * df[0][2] + ih[1] * df[1][2] + ih[2] * df[2][2];
            real fin[NF];
#pragma unroll
            for (int f = 0; f < NF; f++) fin[f] = B.Fin[f * NP + p];
            node_tail<real>(B, gx, gy, c, p, U, fin, met(9, p), A, bdt);
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------
// boundary: outer-face SAT terms and the fault interface
//
// FP64 throughout and without FMA contraction, so that the friction solve
// rounds as the reference's scalar code does (the Regula Falsi iteration
// count depends on last-bit differences near its 1e-12 step tolerance).
#pragma clang fp contract(off)

__device__ __forceinline__ double boxcar(double x, double w)  // Interface_Condition.f90:2438-2456
{
    const double tol = 100.0 * 2.220446049250313e-16;
    if (-w + tol < x && x < w - tol) return 1.0;
    if (fabs(-w - x) <= tol || fabs(x - w) <= tol) return 0.5;
    return 0.0;
}

__device__ __forceinline__ double smooth_boxcar(double x, double w, double wr)  // 2459-2476
{
    if (-w <= x && x <= w) return 1.0;
    if (fabs(x) > w && fabs(x) < w + wr) return 0.5 * (1.0 + tanh(wr / (fabs(x) - w - wr) + wr / (fabs(x) - w)));
    return 0.0;
}

// prestress (Interface_Condition.f90:1802-2435)
// TPV26/27 background stress at depth y (initial_stress_condition.f90:26-76):
// effective vertical stress -(rho - rho_w) g y, the tectonic ratios b11, b33,
// b13 above 15 km tapered linearly to isotropic at 20 km
__device__ void tpv26_stress(double y, double s0[6])
{
    double Omega = 0.0;
    if (y <= 15.0) Omega = 1.0;
    else if (y >= 15.0 && y <= 20.0) Omega = (20.0 - y) / 5.0;
    else if (y >= 20.0) Omega = 0.0;
    const double b11 = 0.926793, b33 = 1.073206, b13 = -0.169029;
    const double Pf = 9.8 * y, s22 = -2.67 * 9.8 * y;
    s0[0] = Omega * (b33 * (s22 + Pf) - Pf) + (1.0 - Omega) * s22 + Pf;
    s0[1] = s22 + Pf;
    s0[2] = Omega * (b11 * (s22 + Pf) - Pf) + (1.0 - Omega) * s22 + Pf;
    s0[3] = 0.0;
    s0[4] = Omega * (b13 * (s22 + Pf));
    s0[5] = 0.0;
}

// TPV29/30 background stress at depth y (initial_stress_condition.f90:87-112):
// the TPV26 form with the taper between 17 and 22 km and its own ratios
__device__ void tpv29_stress(double y, double s0[6])
{
    double Omega = 0.0;
    if (y <= 17.0) Omega = 1.0;
    else if (y >= 17.0 && y <= 22.0) Omega = (22.0 - y) / 5.0;
    else if (y >= 22.0) Omega = 0.0;
    const double b11 = 1.025837, b33 = 0.974162, b13 = -0.158649;
    const double Pf = 9.8 * y, s22 = -(2.67 * 9.8 * y);
    s0[0] = Omega * (b33 * (s22 + Pf) - Pf) + (1.0 - Omega) * s22 + Pf;
    s0[1] = s22 + Pf;
    s0[2] = Omega * (b11 * (s22 + Pf) - Pf) + (1.0 - Omega) * s22 + Pf;
    s0[3] = 0.0;
    s0[4] = Omega * (b13 * (s22 + Pf));
    s0[5] = 0.0;
}

__device__ void prestress(const FaultParams &f, const double x[3], double t, const double l[3], const double m[3],
                          const double n[3], double &Tn, double &Tm, double &Tl, const double *kt = nullptr)
{
    if (f.key) {  // a keyed problem: the node's static prestress, tabled at create
        Tn = kt[0]; Tm = kt[1]; Tl = kt[2];
        return;
    }
    if (f.stress_mode == 0) {
        Tn = f.T0[0]; Tm = f.T0[1]; Tl = f.T0[2];
        for (int k = 0; k < f.npatch; k++)
            Tl = Tl + f.patch[k][4] * boxcar(x[1] - f.patch[k][0], f.patch[k][2]) * boxcar(x[2] - f.patch[k][1], f.patch[k][3]);
    } else {
        double st[6];
        if (f.stress_mode == 2) tpv26_stress(x[1], st);
        else for (int c = 0; c < 6; c++) st[c] = f.s0[c];
        const double *s = st;
        double Tx = n[0] * s[0] + n[1] * s[3] + n[2] * s[4];
        double Ty = n[0] * s[3] + n[1] * s[1] + n[2] * s[5];
        double Tz = n[0] * s[4] + n[1] * s[5] + n[2] * s[2];
        Tn = n[0] * Tx + n[1] * Ty + n[2] * Tz;
        Tm = m[0] * Tx + m[1] * Ty + m[2] * Tz;
        Tl = l[0] * Tx + l[1] * Ty + l[2] * Tz;
    }
    if (f.nuc == 1) {
        double dy = x[1] - f.nucp[0], dz = x[2] - f.nucp[1];
        double r = sqrt(dy * dy + dz * dz);
        double tn = 0.0;
        if (r <= f.nucp[2]) tn = f.nucp[4];
        else if (r >= f.nucp[2] && r <= f.nucp[2] + f.nucp[3])
            tn = 0.5 * f.nucp[4] * (1.0 + cos(3.141592653589793 * (r - f.nucp[2]) / f.nucp[3]));
        Tl = Tl + tn;
    } else if (f.nuc == 2) {
        double dy = x[1] - f.nucp[0], dz = x[2] - f.nucp[1];
        double r = sqrt(dy * dy + dz * dz);
        double R = f.nucp[2], T = f.nucp[4];
        double Fv = 0.0, Gv = 0.0;
        if (r < R) Fv = exp(r * r / (r * r - R * R));
        if (t > 0.0 && t < T) Gv = exp((t - T) * (t - T) / (t * (t - 2.0 * T)));
        if (t >= T) Gv = 1.0;
        Tl = Tl + f.nucp[3] * Fv * Gv;
    } else if (f.nuc == 3) {  // 'rate-weakening' (2387-2420): Gaussian pulse scaled by the normal traction
        const double tau_pulse = f.nucp[3] * Tn;
        const double r0 = f.nucp[2];
        double dy = x[1] - f.nucp[0], dz = x[2] - f.nucp[1];
        double r = sqrt(dy * dy + dz * dz);
        double Fv = exp(-((r / r0) * (r / r0)) / 2.0);
        double Gv = t > 0.0 ? 1.0 : 0.0;
        Tl = Tl + 2.0 * tau_pulse * Fv * Gv;
    }
}

// direct-effect parameter a at fault point x (Interface_Condition.f90:386-416)
__device__ __forceinline__ double rs_a_at(const FaultParams &f, const double x[3])
{
    return f.rs_a + f.rs_da * (1.0 - smooth_boxcar(x[1], f.rs_w, f.rs_wr) * smooth_boxcar(x[2], f.rs_w, f.rs_wr));
}

// shear strength of the slip-weakening laws (Tau_strength, 1268-1797)
__device__ double tau_strength(const FaultParams &f, double sigma_n, double D, const double x[3], double t,
                               const double *kt = nullptr)
{
    if (f.key) {
        // a keyed problem: fs, fd, Dc and C0 (barriers included) and the
        // forced-rupture time of the node, tabled at create.  TPV10, TPV31-34
        // and SCITS2016 weaken linearly in slip; TPV29/30, TPV36/37 and
        // wasatch100 also by forced rupture over t0 = 0.5 s.  sigma_n >= 0
        // here, so the reference's max(sigma_n, 0) is sigma_n.
        const double fs = kt[3], fd = kt[4], Dc = kt[5], C0 = kt[6], tf = kt[7];
        double fr;
        if (f.key == WQL_KEY_TPV29 || f.key == WQL_KEY_TPV36 || f.key == WQL_KEY_TPV37 || f.key == WQL_KEY_WASATCH100) {
            const double t0 = 0.5;
            const double f1 = fmin(D / Dc, 1.0);
            double f2;
            if (t < tf) f2 = 0.0;
            else if (t < tf + t0) f2 = (t - tf) / t0;
            else f2 = 1.0;
            fr = fs - (fs - fd) * fmax(f1, f2);
        } else {
            fr = fs - (fs - fd) * fmin(D, Dc) / Dc;
        }
        return C0 + fr * sigma_n;
    }
    if (f.sw_law == 1) {  // TPV26/27 (1376-1458)
        const double yc = f.frc[0], zc = f.frc[1], r_crit = f.frc[2], Vs = f.frc[3], t0 = f.frc[4];
        const double dy = x[1] - yc, dz = x[2] - zc;
        const double r = sqrt(dy * dy + dz * dz);
        // Frictional_Cohesion (2491-2530), depth = y
        const double C0 = x[1] <= f.coh[2] ? f.coh[0] + f.coh[1] * (f.coh[2] - x[1]) : f.coh[0];
        // Time_of_Forced_Rupture (2625-2660)
        double tf;
        if (r < r_crit - 1.0e-9) {
            const double q = r / r_crit;
            tf = r / (0.7 * Vs) + 0.081 * r_crit / (0.7 * Vs) * (1.0 / (1.0 - q * q) - 1.0);
        } else {
            tf = 1.0e9;
        }
        const double f1 = fmin(D / f.Dc, 1.0);
        double f2;
        if (t < tf) f2 = 0.0;
        else if (t < tf + t0) f2 = (t - tf) / t0;
        else f2 = 1.0;
        const double fs = f.fs + f.bar_big * (1.0 - boxcar(x[1], f.bar_wy) * boxcar(x[2], f.bar_wz));
        const double fr = fs - (fs - f.fd) * fmax(f1, f2);
        return C0 + fr * sigma_n;
    }
    const double fs = f.fs + f.bar_big * (1.0 - boxcar(x[1], f.bar_wy) * boxcar(x[2], f.bar_wz));
    const double fr = fs - (fs - f.fd) * fmin(D, f.Dc) / f.Dc;
    return sigma_n * fr;
}

// ---------------------------------------------------------------------------
// keyed slip-weakening problems (wql_fault_params.key): the static fault
// fields of one block's fault face, once at create, from the node's
// coordinates x (x[1] depth, x[2] along strike), its frame (n, m, l) and the
// block's material there.  Every expression is the reference's, in its
// operation order (FP64, no contraction).

// Frictional_Cohesion (Interface_Condition.f90:2491-2580) at depth d
__device__ __forceinline__ double key_cohesion(int key, double d)
{
    switch (key) {
    case WQL_KEY_TPV29: case WQL_KEY_WASATCH100:
        return d <= 4.0 ? 0.4 + 0.2 * (4.0 - d) : 0.4;
    case WQL_KEY_TPV31: case WQL_KEY_TPV32: case WQL_KEY_TPV34: case WQL_KEY_SCITS2016:
        return d <= 2.4 ? 0.425 * (2.4 - d) : 0.0;
    case WQL_KEY_TPV36:
        return d <= 8.0 ? 0.5 * (8.0 - d) : 0.0;
    case WQL_KEY_TPV37:
        return d <= 8.0 ? 1.875 * (8.0 - d) : 0.0;
    default:
        return 0.0;
    }
}

// Time_of_Forced_Rupture (2583-2660): r_crit 4 km; thr and big are the
// per-problem cut-off offset and "never" time
__device__ __forceinline__ double key_forced_rupture(double r, double Vs, double thr, double big)
{
    const double r_crit = 4.0;
    if (r < r_crit - thr) {
        const double q = r / r_crit;
        return r / (0.7 * Vs) + 0.081 * r_crit / (0.7 * Vs) * (1.0 / (1.0 - q * q) - 1.0);
    }
    return big;
}

// tractions of a stress tensor on the face frame (prestress 1802-2435)
__device__ __forceinline__ void key_tractions(const double s[6], const double n[3], const double m[3],
                                              const double l[3], double &Tn, double &Tm, double &Tl)
{
    const double Tx = n[0] * s[0] + n[1] * s[3] + n[2] * s[4];
    const double Ty = n[0] * s[3] + n[1] * s[1] + n[2] * s[5];
    const double Tz = n[0] * s[4] + n[1] * s[5] + n[2] * s[2];
    Tn = n[0] * Tx + n[1] * Ty + n[2] * Tz;
    Tm = m[0] * Tx + m[1] * Ty + m[2] * Tz;
    Tl = l[0] * Tx + l[1] * Ty + l[2] * Tz;
}

// cosine-tapered nucleation disk of prestress (TPV31/32/34, TPV33,
// SCITS2016): a for r <= r1, ha (1 + cos(pi (r - r1) / w)) for r1 <= r <= r2,
// both times s (the shear modulus over u0; 1 for TPV33), as
// (a s) / u0 -- the reference's order -- when s_num is given
__device__ __forceinline__ double key_nucleation(double r, double r1, double w, double r2, double a, double ha,
                                                 double mu, double u0)
{
    const double pi = 3.141592653589793;
    if (r <= r1) return u0 > 0.0 ? a * mu / u0 : a;
    if (r >= r1 && r <= r2) {
        const double v = ha * (1.0 + cos(pi * (r - r1) / w));
        return u0 > 0.0 ? v * mu / u0 : v;
    }
    return 0.0;
}

__global__ void k_fault_tables(int key, long FP, const double *__restrict__ fx, const double *__restrict__ frame,
                               const double *__restrict__ fmat, double *__restrict__ tab)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= FP) return;
    const double pi = 3.141592653589793;
    double x[3], n[3], m[3], l[3];
    for (int c = 0; c < 3; c++) {
        x[c] = fx[c * FP + t];
        n[c] = frame[c * FP + t]; m[c] = frame[(3 + c) * FP + t]; l[c] = frame[(6 + c) * FP + t];
    }
    double Tn = 0.0, Tm = 0.0, Tl = 0.0, fs = 0.0, fd = 0.0, Dc = 1.0, C0 = 0.0, tf = 1.0e9;
    double s0[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double u0 = 32.03812032;
    switch (key) {
    case WQL_KEY_TPV10: {  // 1591-1610, 1882-1912: 60-degree normal fault
        const double B = boxcar(x[1] * (2.0 / sqrt(3.0)), 15.0) * boxcar(x[2] - 0.0, 15.0);
        fs = 0.760 + (1.0e4 - 0.760) * (1.0 - B);
        fd = 0.448; Dc = 0.50;
        C0 = 0.2 + 1.0e3 * (1.0 - B);
        Tn = -(7.378 * x[1]) * (2.0 / sqrt(3.0));
        Tm = -(0.55 * Tn);
        Tl = 0.0;
        Tm = Tm + (-((0.760 + 0.0057) * Tn) + 0.2 - Tm) * boxcar(x[1] * (2.0 / sqrt(3.0)) - 12.0, 1.5) *
                      boxcar(x[2] - 0.0, 1.5);
        break;
    }
    case WQL_KEY_TPV29: {  // 1477-1518, 2081-2101; initial_stress_condition.f90:87-112
        fs = 0.18; fd = 0.12; Dc = 0.3;
        const double Vs = 3.464;
        const double dy = x[1] - 10.0, dz = x[2] + 5.0;
        const double r = sqrt(dy * dy + dz * dz);
        C0 = key_cohesion(key, x[1]) + 1.0e10 * (1.0 - boxcar(x[1], 20.0) * boxcar(x[2], 20.0));
        tf = key_forced_rupture(r, Vs, 1.0e-9, 1.0e9);
        tpv29_stress(x[1], s0);
        key_tractions(s0, n, m, l, Tn, Tm, Tl);
        break;
    }
    case WQL_KEY_TPV31: case WQL_KEY_TPV32: case WQL_KEY_TPV34: case WQL_KEY_SCITS2016: {
        // 1520-1590 and 2102-2135, 2169-2240; the shear modulus of the fault
        // node: the TPV31 / TPV32 depth profile (Interface_RHS 133-254), the
        // block's own for TPV34 and SCITS2016 (255-280)
        double mu;
        if (key == WQL_KEY_TPV31 || key == WQL_KEY_TPV32) {
            double vp = 0.0, vs = 0.0, rho = 0.0;
            layered_medium(key == WQL_KEY_TPV31 ? WQL_MEDIUM_TPV31 : WQL_MEDIUM_TPV32, x[0], x[1], vp, vs, rho);
            mu = rho * (vs * vs);
        } else {
            mu = fmat[FP + t];
        }
        fd = 0.45; Dc = 0.18;
        if (key == WQL_KEY_TPV34) {
            fs = 0.58 + 1.0e10 * (1.0 - boxcar(x[0], 15.0) * boxcar(x[1], 15.0));
            C0 = key_cohesion(key, x[0]);
        } else if (key == WQL_KEY_SCITS2016) {
            fs = 0.58 + 1.0e10 * (1.0 - boxcar(x[1] - 5.0, 2.0) * boxcar(x[2] - 20.0, 3.0));
            fd = 0.35; Dc = 0.09;
            C0 = key_cohesion(key, x[1]);
        } else {
            fs = 0.58 + 1.0e10 * (1.0 - boxcar(x[1], 15.0) * boxcar(x[2], 15.0));
            C0 = key_cohesion(key, x[1]);
        }
        s0[0] = -(60.0 * mu) / u0;
        s0[2] = -(60.0 * mu) / u0;
        s0[4] = 30.0 * mu / u0;
        key_tractions(s0, n, m, l, Tn, Tm, Tl);
        if (key == WQL_KEY_SCITS2016) {
            const double dy = x[1] - 4.5, dz = x[2] - 20.0;
            const double r = sqrt(dy * dy + dz * dz);
            Tl = Tl + key_nucleation(r, 1.4, 0.6, 2.0, 14.95, 7.475, mu, u0);
        } else {
            const double dy = x[1] - 7.5, dz = x[2];
            const double r = sqrt(dy * dy + dz * dz);
            Tl = Tl + key_nucleation(r, 1.4, 0.6, 2.0, 4.95, 2.475, mu, u0);
        }
        break;
    }
    case WQL_KEY_TPV33: {  // 1538-1553, 2136-2168; initial_stress_condition.f90:133-158
        fs = 0.55 + 1.0e10 * (1.0 - boxcar(x[0], 10.0) * boxcar(x[1] + 4.0, 8.0));
        fd = 0.45; Dc = 0.18; C0 = 0.0;
        double Rx, Ry;
        if (x[2] < -9.8) Rx = (-x[2] - 9.8) / 10.0;
        else if (x[2] > 1.1) Rx = (x[2] - 1.1) / 10.0;
        else Rx = 0.0;
        if (x[1] < 2.3) Ry = (-x[1] + 2.3) / 10.0;
        else if (x[1] > 8.0) Ry = (x[1] - 8.0) / 10.0;
        else Ry = 0.0;
        const double Rt = fmin(1.0, sqrt(Rx * Rx + Ry * Ry));
        s0[0] = -60.0;
        s0[4] = 30.0 * (1.0 - Rt);
        key_tractions(s0, n, m, l, Tn, Tm, Tl);
        const double dy = x[1] - 6.0, dz = x[2] + 6.0;
        const double r = sqrt(dy * dy + dz * dz);
        Tl = Tl + key_nucleation(r, 0.55, 0.25, 0.8, 3.15, 1.575, 1.0, 0.0);
        break;
    }
    case WQL_KEY_TPV36: case WQL_KEY_TPV37: {  // 1612-1690, 1914-1920: 15-degree thrust
        fs = 0.575; fd = 0.450; Dc = 0.180;
        const double Vs = 3.464;
        const double sd = sin(15.0 / 180.0 * pi);
        const double d = x[1] / sd;
        const double r = sqrt((d - 18.0) * (d - 18.0) + x[2] * x[2]);
        C0 = key_cohesion(key, d) + 1.0e10 * (1.0 - boxcar(d, 28.0) * boxcar(x[2], 15.0));
        tf = key_forced_rupture(r, Vs, 10.0e-10, 10.0e10);
        Tn = -(4.24 * x[1]) / sd;
        Tm = -(2.12 * x[1]) / sd;
        Tl = 0.0;
        break;
    }
    case WQL_KEY_WASATCH100: {  // 1692-1742, 1922-1990; initial_stress_condition.f90:222-244
        const double d = x[1] + 2.149;
        fs = 0.6;
        if (d <= 4.0) fd = 0.44 + (0.6 - 0.44) * 0.25 * (4.0 - d);
        else if (d > 4.0 && d <= 17.0) fd = 0.44;
        else if (d > 17.0 && d <= 18.0) fd = 0.44 + (0.6 - 0.44) * (d - 17.0);
        else if (d > 18.0) fd = 0.6;
        if (d <= 4.0) Dc = 0.3 + 0.2 * (4.0 - d);
        else if (d >= 4.0) Dc = 0.3;
        const double Vs = 3.464;
        const double a = 1.3054 * (d - 14.0), b = x[2] - 38.0;
        const double r = sqrt(a * a + b * b);
        const double B = boxcar(d, 18.0) * boxcar(x[2] - 38.0, 18.6);
        fs = fs + 1.0e10 * (1.0 - B);
        C0 = key_cohesion(key, d) + 1.0e10 * (1.0 - B);
        tf = key_forced_rupture(r, Vs, 1.0e-9, 1.0e9);
        const double Pf = 9.8 * (x[1] + 2.149), s22 = -(2.7 * 9.8) * (x[1] + 2.149);
        s0[0] = s22 + Pf - 0.4 * s22;
        s0[1] = s22 + Pf;
        s0[2] = (s0[0] + s0[1]) * 0.5;
        key_tractions(s0, n, m, l, Tn, Tm, Tl);
        break;
    }
    default:
        break;
    }
    const double v[8] = {Tn, Tm, Tl, fs, fd, Dc, C0, tf};
    for (int c = 0; c < 8; c++) tab[c * FP + t] = v[c];
}

__device__ __forceinline__ double rs_f(double V, double Phi, double eta, double sn, double psi, double V0, double a)
{  // Interface_Condition.f90:1262
    return V + a * sn / eta * asinh(0.5 * V / V0 * exp(psi / a)) - Phi / eta;
}

// Regula Falsi of Interface_Condition.f90:1170-1239, kept step for step
__device__ double regula_falsi(double V, double Phi, double eta, double sn, double psi, double V0, double a)
{
    double err = 1.0, tol = 1.0e-12;
    double Vl = 0.0, Vr = Phi / eta;
    int k = 1;
    const int maxit = 5000;
    double fv = rs_f(V, Phi, eta, sn, psi, V0, a);
    double fl = rs_f(Vl, Phi, eta, sn, psi, V0, a);
    double fr = rs_f(Vr, Phi, eta, sn, psi, V0, a);
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
        fv = rs_f(V, Phi, eta, sn, psi, V0, a);
        err = fabs(V - V1);
        k = k + 1;
    }
    return V;
}

// Fast slip-rate solve: the root of the same f(V) (1241-1265) by Newton's
// method inside the bracket [0, Phi/eta] that regula falsi starts from.  f is
// increasing in V; every iterate updates the bracket by the sign of f.  A
// Newton step in V that leaves the bracket is retried in ln V (where f is
// nearly linear once the asinh argument is large), and bisected if that
// leaves it too.  Stops when a step moves V by at most 1e-13 V.
__device__ double newton_rs(double V, double Phi, double eta, double sn, double psi, double V0, double a)
{
    double lo = 0.0, hi = Phi / eta;
    if (!(hi > 0.0)) return 0.0;
    if (!(V > lo && V < hi)) V = 0.5 * hi;
    const double e = exp(psi / a) / (2.0 * V0);
    const double c = a * sn / eta;
    const double rhs = Phi / eta;
    for (int k = 0; k < 100; k++) {
        const double z = V * e;
        const double fv = V + c * asinh(z) - rhs;
        if (fv == 0.0) return V;
        if (fv < 0.0) lo = V;
        else hi = V;
        const double df = 1.0 + c * e / sqrt(1.0 + z * z);
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

// hat variables of one fault node (Interface_RHS 62-452; Welded 516-637;
// Slip_Weakening 745-916; rate_state 918-1067); returns 0 or an error code
__device__ int hat_vars(const FaultParams &f, const double u[6], const double v[6], const double m1[3],
                        const double m2[3], const double x[3], const double nn[3], const double mm[3],
                        const double ll[3], double slip, double psi, double t, double uo[6], double vo[6],
                        double &Vel, double T0[3], const double *kt)
{
    T0[0] = T0[1] = T0[2] = 0.0;
    double c1p = sqrt((2.0 * m1[1] + m1[0]) / m1[2]), c1s = sqrt(m1[1] / m1[2]);
    double c2p = sqrt((2.0 * m2[1] + m2[0]) / m2[2]), c2s = sqrt(m2[1] / m2[2]);
    double Z1s = m1[2] * c1s, Z2s = m2[2] * c2s, Z1p = m1[2] * c1p, Z2p = m2[2] * c2p;
    double p_n = Z1p * u[0] - u[3], p_m = Z1s * u[1] - u[4], p_l = Z1s * u[2] - u[5];
    double q_n = Z2p * v[0] + v[3], q_m = Z2s * v[1] + v[4], q_l = Z2s * v[2] + v[5];
    Vel = 0.0;
    if (f.law == 0) {
        double eta_p = 1.0 / (1.0 / Z1p + 1.0 / Z2p);
        double eta_s = 1.0 / (1.0 / Z1s + 1.0 / Z2s);
        double phi_n = eta_p * (q_n / Z2p - p_n / Z1p);
        double phi_m = eta_s * (q_m / Z2s - p_m / Z1s);
        double phi_l = eta_s * (q_l / Z2s - p_l / Z1s);
        uo[0] = (q_n - phi_n) / Z2p; uo[3] = phi_n;
        vo[0] = (p_n + phi_n) / Z1p; vo[3] = phi_n;
        uo[2] = (q_l - phi_l) / Z2s; uo[5] = phi_l;
        vo[2] = (p_l + phi_l) / Z1s; vo[5] = phi_l;
        uo[1] = (q_m - phi_m) / Z2s; uo[4] = phi_m;
        vo[1] = (p_m + phi_m) / Z1s; vo[4] = phi_m;
        return 0;
    }
    if (Z1s <= 1.0e-9 || Z2s <= 1.0e-9) return 2;
    double eta_s = 1.0 / (1.0 / Z1s + 1.0 / Z2s);
    double eta_p = 1.0 / (1.0 / Z1p + 1.0 / Z2p);
    double T0n, T0m, T0l;
    prestress(f, x, t, ll, mm, nn, T0n, T0m, T0l, kt);
    T0[0] = T0n; T0[1] = T0m; T0[2] = T0l;
    double phi_n = eta_p * (q_n / Z2p - p_n / Z1p);
    double phi_m = eta_s * (q_m / Z2s - p_m / Z1s);
    double phi_l = eta_s * (q_l / Z2s - p_l / Z1s);
    double u_n_hat = (q_n - phi_n) / Z2p, Tu_n_hat = phi_n;
    double v_n_hat = (p_n + phi_n) / Z1p, Tv_n_hat = phi_n;
    double sigma_n = fmax(0.0, -(T0n + phi_n));
    double vv_m, vv_l, Tu_m, Tu_l, Tv_m, Tv_l;
    if (f.law == 1) {
        double tau_lock = sqrt((T0l + phi_l) * (T0l + phi_l) + (T0m + phi_m) * (T0m + phi_m));
        double tau_str = tau_strength(f, sigma_n, slip, x, t, kt);
        if (tau_lock > tau_str) {
            double phi_1 = T0m + phi_m, phi_2 = T0l + phi_l;
            double Phi = sqrt(phi_1 * phi_1 + phi_2 * phi_2);
            double V = (Phi - tau_str) / eta_s;
            if (V < 0.0) return 3;
            Vel = V;
            vv_m = phi_1 / (eta_s + tau_str / V);
            vv_l = phi_2 / (eta_s + tau_str / V);
            double tau_m = phi_1 - eta_s * vv_m, tau_l = phi_2 - eta_s * vv_l;
            Tu_m = tau_m - T0m; Tu_l = tau_l - T0l;
            Tv_m = tau_m - T0m; Tv_l = tau_l - T0l;
        } else {
            Tu_m = phi_m; Tu_l = phi_l; Tv_m = phi_m; Tv_l = phi_l;
            vv_m = 0.0; vv_l = 0.0; Vel = 0.0;
        }
    } else {
        double a = rs_a_at(f, x);
        double V0 = f.rs_V0;
        vv_m = v[1] - u[1];
        vv_l = v[2] - u[2];
        double phi_1 = T0m + phi_m, phi_2 = T0l + phi_l;
        double Phi = sqrt(phi_1 * phi_1 + phi_2 * phi_2);
        double V = sqrt(vv_m * vv_m + vv_l * vv_l);
        if (V > Phi) V = 0.5 * Phi / eta_s;
        V = f.rs_solver == 1 ? newton_rs(V, Phi, eta_s, sigma_n, psi, V0, a)
                             : regula_falsi(V, Phi, eta_s, sigma_n, psi, V0, a);
        if (V < 0.0) return 3;
        Vel = V;
        double fv = sigma_n * a * asinh(0.5 * V / V0 * exp(psi / a));
        vv_m = phi_1 / (eta_s + fv / V);
        vv_l = phi_2 / (eta_s + fv / V);
        double tau_m = phi_1 - eta_s * vv_m, tau_l = phi_2 - eta_s * vv_l;
        Tu_m = tau_m - T0m; Tu_l = tau_l - T0l;
        Tv_m = tau_m - T0m; Tv_l = tau_l - T0l;
    }
    uo[0] = u_n_hat; vo[0] = v_n_hat;
    uo[1] = (q_m - Tv_m) / Z2s - vv_m;
    uo[2] = (q_l - Tv_l) / Z2s - vv_l;
    vo[1] = (Tu_m + p_m) / Z1s + vv_m;
    vo[2] = (Tu_l + p_l) / Z1s + vv_l;
    uo[3] = Tu_n_hat; vo[3] = Tv_n_hat;
    uo[4] = Tu_m; uo[5] = Tu_l;
    vo[4] = Tv_m; vo[5] = Tv_l;
    return 0;
}

// characteristic boundary SAT of one node (BoundaryConditions.f90:777-918)
// followed by `structure` (RHS_Interior.f90:2203-2223)
__device__ void bc_sat(const double u[9], double rho, double mu, double lambda, const double l[3],
                       const double m[3], const double n[3], double norm, double r, int plus,
                       double ubc[9], double Ux[9])
{
    double c_p = sqrt((2.0 * mu + lambda) / rho), c_s = sqrt(mu / rho);
    double Zp = rho * c_p, Zs = rho * c_s;
    double T0 = n[0] * u[3] + n[1] * u[6] + n[2] * u[7];
    double T1 = n[0] * u[6] + n[1] * u[4] + n[2] * u[8];
    double T2 = n[0] * u[7] + n[1] * u[8] + n[2] * u[5];
    double vl = l[0] * u[0] + l[1] * u[1] + l[2] * u[2];
    double vm = m[0] * u[0] + m[1] * u[1] + m[2] * u[2];
    double vn = n[0] * u[0] + n[1] * u[1] + n[2] * u[2];
    double Tl = l[0] * T0 + l[1] * T1 + l[2] * T2;
    double Tm = m[0] * T0 + m[1] * T1 + m[2] * T2;
    double Tn = n[0] * T0 + n[1] * T1 + n[2] * T2;
    double s = plus ? 1.0 : -1.0;
    double pl = 0.5 * ((1.0 - r) * Zs * vl + s * ((1.0 + r) * Tl));
    double pm = 0.5 * ((1.0 - r) * Zs * vm + s * ((1.0 + r) * Tm));
    double pn = 0.5 * ((1.0 - r) * Zp * vn + s * ((1.0 + r) * Tn));
    double ql = 0.5 * ((1.0 - r) * vl + s * ((1.0 + r) / Zs * Tl));
    double qm = 0.5 * ((1.0 - r) * vm + s * ((1.0 + r) / Zs * Tm));
    double qn = 0.5 * ((1.0 - r) * vn + s * ((1.0 + r) / Zp * Tn));
    double Fx = l[0] * pl + m[0] * pm + n[0] * pn;
    double Fy = l[1] * pl + m[1] * pm + n[1] * pn;
    double Fz = l[2] * pl + m[2] * pm + n[2] * pn;
    double F_x = l[0] * ql + m[0] * qm + n[0] * qn;
    double F_y = l[1] * ql + m[1] * qm + n[1] * qn;
    double F_z = l[2] * ql + m[2] * qm + n[2] * qn;
    ubc[0] = norm * Fx;
    ubc[1] = norm * Fy;
    ubc[2] = norm * Fz;
    double sn = s * norm;
    ubc[3] = sn * n[0] * F_x;
    ubc[4] = sn * n[1] * F_y;
    ubc[5] = sn * n[2] * F_z;
    ubc[6] = sn * (n[1] * F_x + n[0] * F_y);
    ubc[7] = sn * (n[2] * F_x + n[0] * F_z);
    ubc[8] = sn * (n[2] * F_y + n[1] * F_z);
    Ux[0] = 1.0 / rho * ubc[0];
    Ux[1] = 1.0 / rho * ubc[1];
    Ux[2] = 1.0 / rho * ubc[2];
    Ux[3] = (2.0 * mu + lambda) * ubc[3] + lambda * (ubc[4] + ubc[5]);
    Ux[4] = (2.0 * mu + lambda) * ubc[4] + lambda * (ubc[3] + ubc[5]);
    Ux[5] = (2.0 * mu + lambda) * ubc[5] + lambda * (ubc[3] + ubc[4]);
    Ux[6] = mu * ubc[6];
    Ux[7] = mu * ubc[7];
    Ux[8] = mu * ubc[8];
}

// fault SAT for block `me` (0: minus side, face q = nq-1; 1: plus side,
// face q = 0) at fault node (j, k): CouplingForcing.f90:9-467 with
// Interface_RHS; adds -tau0/h_q * bc to sat[9]
template <typename real>
__device__ int fault_node(const Pair<real> &P, const Iface &I, const FaultParams &fp, int me, int j, int k,
                          double t, int snap, double A, double bdt, double sat[9], double tau0, int fq_read)
{
    const Blk<real> &B1 = P.b[0], &B2 = P.b[1];
    const Blk<real> &B = P.b[me];
    const int side = me == 0 ? 1 : 0;
    const int nr = B.nr;
    const long FP = (long)nr * B.ns;
    const long fpi = (long)j + (long)nr * k;
    const long pu = pidx(B1.px, B1.nr, B1.nq - 1, j, k), pv = pidx(B2.px, B2.nr, 0, j, k);
    // every load of the node is issued here, before any store: each thread
    // owns index fpi of every interface array, but the compiler cannot
    // prove the arrays disjoint and would otherwise wait for each store
    // before the next load
    double u[9], v[9];
    if (fq_read) {
        // block 0's face q = nq-1 is the second half of its buffer, block
        // 1's face q = 0 the first half of its buffer
        const long FP1 = (long)B1.nr * B1.ns, FP2 = (long)B2.nr * B2.ns;
        const real *fu = B1.FQ + 9 * FP1 + j + (long)B1.nr * k, *fv = B2.FQ + j + (long)B2.nr * k;
        for (int f = 0; f < 9; f++) { u[f] = (double)fu[f * FP1]; v[f] = (double)fv[f * FP2]; }
    } else {
        for (int f = 0; f < 9; f++) { u[f] = (double)B1.Fin[f * B1.NP + pu]; v[f] = (double)B2.Fin[f * B2.NP + pv]; }
    }
    const double *fr = B.frame[side];
    double nn[3], mm[3], ll[3];
    for (int c = 0; c < 3; c++) { nn[c] = fr[c * FP + fpi]; mm[c] = fr[(3 + c) * FP + fpi]; ll[c] = fr[(6 + c) * FP + fpi]; }
    const double *fm1 = B1.fmat[1], *fm2 = B2.fmat[0];
    double mat1[3] = {fm1[fpi], fm1[FP + fpi], fm1[2 * FP + fpi]};
    double mat2[3] = {fm2[fpi], fm2[FP + fpi], fm2[2 * FP + fpi]};
    double x[3] = {B.fx[fpi], B.fx[FP + fpi], B.fx[2 * FP + fpi]};
    const double *gm = B.fgm[side];
    const double q_x = gm[fpi], q_y = gm[FP + fpi], q_z = gm[2 * FP + fpi];
    double S0[4], DS0[4];
    for (int c = 0; c < 4; c++) { S0[c] = I.S[c * FP + fpi]; DS0[c] = I.DS[c * FP + fpi]; }
    const double psi = I.W[fpi], DW0 = I.DW[fpi];
    const double trup0 = snap ? I.trup[fpi] : 0.0, psr0 = snap ? I.psr[fpi] : 0.0;
    // tractions and rotation into (n, m, l) (CouplingForcing.f90:105-191, 470-524)
    double ur[6], vr[6];
    ur[0] = u[0]; ur[1] = u[1]; ur[2] = u[2];
    ur[3] = nn[0] * u[3] + nn[1] * u[6] + nn[2] * u[7];
    ur[4] = nn[0] * u[6] + nn[1] * u[4] + nn[2] * u[8];
    ur[5] = nn[0] * u[7] + nn[1] * u[8] + nn[2] * u[5];
    vr[0] = v[0]; vr[1] = v[1]; vr[2] = v[2];
    vr[3] = nn[0] * v[3] + nn[1] * v[6] + nn[2] * v[7];
    vr[4] = nn[0] * v[6] + nn[1] * v[4] + nn[2] * v[8];
    vr[5] = nn[0] * v[7] + nn[1] * v[8] + nn[2] * v[5];
    double uL[6], vL[6];
    for (int g = 0; g < 2; g++) {
        const double *s = g == 0 ? ur : vr;
        double *d = g == 0 ? uL : vL;
        d[0] = nn[0] * s[0] + nn[1] * s[1] + nn[2] * s[2];
        d[1] = mm[0] * s[0] + mm[1] * s[1] + mm[2] * s[2];
        d[2] = ll[0] * s[0] + ll[1] * s[1] + ll[2] * s[2];
        d[3] = nn[0] * s[3] + nn[1] * s[4] + nn[2] * s[5];
        d[4] = mm[0] * s[3] + mm[1] * s[4] + mm[2] * s[5];
        d[5] = ll[0] * s[3] + ll[1] * s[4] + ll[2] * s[5];
    }
    double slip = S0[3];
    double uh[6], vh[6], Vel, T0[3];
    double kt[8];
    if (fp.key)
        for (int c = 0; c < 8; c++) kt[c] = I.ftab[c * FP + fpi];
    int rc = hat_vars(fp, uL, vL, mat1, mat2, x, nn, mm, ll, slip, psi, t, uh, vh, Vel, T0, kt);
    if (rc) return rc;
    // fault state rates (Interface_Condition.f90:431-436, 470-476) and RK update
    double slip_rate = sqrt((vh[1] - uh[1]) * (vh[1] - uh[1]) + (vh[2] - uh[2]) * (vh[2] - uh[2]));
    {
        double dw = A * DW0;
        if (fp.law == 2) {
            const double b = fp.rs_b, V0 = fp.rs_V0, L0 = fp.rs_L, f0 = fp.rs_f0;
            if (fp.rs_law == 0) {  // aging law (431-436)
                dw = dw + b * V0 / L0 * exp(-(psi - f0) / b) - Vel * b / L0;
            } else {
                const double a = rs_a_at(fp, x);
                double Fss = f0 - (b - a) * log(Vel / V0);
                if (fp.rs_law == 1) {  // slip law with flash weakening (438-444)
                    if (fp.rs_Vw > 0.0) {
                        const double Vw = fp.rs_Vw + fp.rs_dVw * (1.0 - smooth_boxcar(x[1], fp.rs_w, fp.rs_wr) *
                                                                            smooth_boxcar(x[2], fp.rs_w, fp.rs_wr));
                        const double q = Vel / Vw, q2 = q * q, q4 = q2 * q2, q8 = q4 * q4;
                        Fss = fp.rs_fw + (Fss - fp.rs_fw) / pow(1.0 + q8, 0.125);
                    }
                    dw = dw - Vel / L0 * (a * asinh(Vel / (2.0 * V0) * exp(psi / a)) - Fss);
                } else {  // KDN (446-449): slip_rate weighting, asinh without a
                    dw = dw - (1.0 / L0) * slip_rate * (asinh(Vel / (2.0 * V0) * exp(psi / a)) - Fss);
                }
            }
        }
        I.DW[fpi] = dw;
        I.W[fpi] = psi + bdt * dw;
    }
    for (int c = 0; c < 4; c++) {
        double ds = A * DS0[c];
        ds = ds + (c < 3 ? vh[c] - uh[c] : slip_rate);
        I.DS[c * FP + fpi] = ds;
        I.S[c * FP + fpi] = S0[c] + bdt * ds;
    }
    if (!isfinite(slip_rate) || !isfinite(uh[4]) || !isfinite(uh[5])) return 4;
    if (snap) {  // stage 1: 480-509, CouplingForcing.f90:213-228, 584-605
        // prestress of this node and time, as formed in hat_vars (zero for
        // the locked law)
        const double T0n = T0[0], T0m = T0[1], T0l = T0[2];
        double sv[3];
        for (int c = 0; c < 3; c++) {
            sv[c] = vh[c] - uh[c];
            I.Svel[c * FP + fpi] = sv[c];
            I.Uhat[c * FP + fpi] = sv[c];
            I.Vhat[c * FP + fpi] = sv[c];
        }
        I.Uhat[3 * FP + fpi] = uh[3] + T0n; I.Vhat[3 * FP + fpi] = vh[3] + T0n;
        I.Uhat[4 * FP + fpi] = uh[4] + T0m; I.Vhat[4 * FP + fpi] = vh[4] + T0m;
        I.Uhat[5 * FP + fpi] = uh[5] + T0l; I.Vhat[5 * FP + fpi] = vh[5] + T0l;
        const double srate = sqrt(sv[1] * sv[1] + sv[2] * sv[2]);
        if (srate > 1e-3) I.trup[fpi] = fmin(trup0, t);
        I.psr[fpi] = fmax(psr0, srate);
    }
    // rotate back (527-581) and SAT forcing (242-463)
    double uX[6], vX[6], uH[6], vH[6];
    const double *srcs[4] = {uL, vL, uh, vh};
    double *dsts[4] = {uX, vX, uH, vH};
    for (int g = 0; g < 4; g++)
        for (int c = 0; c < 3; c++) {
            dsts[g][c] = nn[c] * srcs[g][0] + mm[c] * srcs[g][1] + ll[c] * srcs[g][2];
            dsts[g][3 + c] = nn[c] * srcs[g][3] + mm[c] * srcs[g][4] + ll[c] * srcs[g][5];
        }
    double norm = sqrt(q_x * q_x + q_y * q_y + q_z * q_z);
    const double *mat = me == 0 ? mat1 : mat2;
    double lam = mat[0], mu = mat[1], rho = mat[2];
    double Zp = rho * sqrt((2.0 * mu + lam) / rho), Zs = rho * sqrt(mu / rho);
    double n_1 = nn[0], n_2 = nn[1], n_3 = nn[2], m_1 = mm[0], m_2 = mm[1], m_3 = mm[2];
    double l_1 = ll[0], l_2 = ll[1], l_3 = ll[2];
    double M11 = n_1 * n_1 * Zp + n_2 * n_2 * Zs + n_3 * n_3 * Zs;
    double M12 = n_1 * m_1 * Zp + n_2 * m_2 * Zs + n_3 * m_3 * Zs;
    double M13 = n_1 * l_1 * Zp + n_2 * l_2 * Zs + n_3 * l_3 * Zs;
    double M22 = m_1 * m_1 * Zp + m_2 * m_2 * Zs + m_3 * m_3 * Zs;
    double M23 = m_1 * l_1 * Zp + m_2 * l_2 * Zs + m_3 * l_3 * Zs;
    double M31 = l_1 * n_1 * Zp + l_2 * n_2 * Zs + l_3 * n_3 * Zs;
    double M32 = l_1 * m_1 * Zp + l_2 * m_2 * Zs + l_3 * m_3 * Zs;
    double M33 = l_1 * l_1 * Zp + l_2 * l_2 * Zs + l_3 * l_3 * Zs;
    double P11 = n_1 * n_1 * 1.0 / Zp + n_2 * n_2 * 1.0 / Zs + n_3 * n_3 * 1.0 / Zs;
    double P12 = n_1 * m_1 * 1.0 / Zp + n_2 * m_2 * 1.0 / Zs + n_3 * m_3 * 1.0 / Zs;
    double P13 = n_1 * l_1 * 1.0 / Zp + n_2 * l_2 * 1.0 / Zs + n_3 * l_3 * 1.0 / Zs;
    double P22 = m_1 * m_1 * 1.0 / Zp + m_2 * m_2 * 1.0 / Zs + m_3 * m_3 * 1.0 / Zs;
    double P23 = m_1 * l_1 * 1.0 / Zp + m_2 * l_2 * 1.0 / Zs + m_3 * l_3 * 1.0 / Zs;
    double P31 = l_1 * n_1 * 1.0 / Zp + l_2 * n_2 * 1.0 / Zs + l_3 * n_3 * 1.0 / Zs;
    double P32 = l_1 * m_1 * 1.0 / Zp + l_2 * m_2 * 1.0 / Zs + l_3 * m_3 * 1.0 / Zs;
    double P33 = l_1 * l_1 * 1.0 / Zp + l_2 * l_2 * 1.0 / Zs + l_3 * l_3 * 1.0 / Zs;
    double bc[9];
    if (me == 0) {
        double p1h = 0.5 * (M11 * uH[0] + M12 * uH[1] + M13 * uH[2] + uH[3]);
        double p2h = 0.5 * (M12 * uH[0] + M22 * uH[1] + M23 * uH[2] + uH[4]);
        double p3h = 0.5 * (M31 * uH[0] + M32 * uH[1] + M33 * uH[2] + uH[5]);
        double p1 = 0.5 * (M11 * uX[0] + M12 * uX[1] + M13 * uX[2] + uX[3]);
        double p2 = 0.5 * (M12 * uX[0] + M22 * uX[1] + M23 * uX[2] + uX[4]);
        double p3 = 0.5 * (M31 * uX[0] + M32 * uX[1] + M33 * uX[2] + uX[5]);
        double e1 = P11 * (p1 - p1h) + P12 * (p2 - p2h) + P13 * (p3 - p3h);
        double e2 = P12 * (p1 - p1h) + P22 * (p2 - p2h) + P23 * (p3 - p3h);
        double e3 = P31 * (p1 - p1h) + P32 * (p2 - p2h) + P33 * (p3 - p3h);
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
        double q1h = 0.5 * (M11 * vH[0] + M12 * vH[1] + M13 * vH[2] - vH[3]);
        double q2h = 0.5 * (M12 * vH[0] + M22 * vH[1] + M23 * vH[2] - vH[4]);
        double q3h = 0.5 * (M31 * vH[0] + M32 * vH[1] + M33 * vH[2] - vH[5]);
        double q1 = 0.5 * (M11 * vX[0] + M12 * vX[1] + M13 * vX[2] - vX[3]);
        double q2 = 0.5 * (M12 * vX[0] + M22 * vX[1] + M23 * vX[2] - vX[4]);
        double q3 = 0.5 * (M31 * vX[0] + M32 * vX[1] + M33 * vX[2] - vX[5]);
        double e1 = P11 * (q1 - q1h) + P12 * (q2 - q2h) + P13 * (q3 - q3h);
        double e2 = P12 * (q1 - q1h) + P22 * (q2 - q2h) + P23 * (q3 - q3h);
        double e3 = P31 * (q1 - q1h) + P32 * (q2 - q2h) + P33 * (q3 - q3h);
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
    double hq = 1.0 / (double)(B.nq - 1);
    for (int f = 0; f < 9; f++) sat[f] = sat[f] - tau0 / hq * bc[f];
    return 0;
}

struct BoundaryArgs {
    FaultParams fault;
    Iface I[2];
    long nodes[2];        // boundary nodes per block
    double t, A[2], bdt, tau0;
    const double *tdev;   // non-null: stage time = *tdev + tinc (step time kept on the device)
    double tinc;
    // 0: apply every boundary and fault penalty to DF / F_out and the PML
    //    faces' penalties to DQ / Q (after the volume kernels); 1: before the
    //    volume kernels, write each boundary node's penalty to satb and the
    //    PML faces' penalties to the blocks' PQ arrays (the volume and
    //    PML-slab kernels add them) and run the fault
    int mode;
    void *satb[2];        // [9][nbnd] per block, in the stage's real type
    int stage;
    int rates_only;       // 1: RHS evaluation without the RK update (no F_out / S / W write, and no
                          //    stage-1 snapshot of Svel / Uhat / Vhat / trup / peak slip rate)
    int fq_read;          // 1: read the stage input of q-face nodes from the blocks' q-face buffers
    int *flag;
};

template <typename real>
__global__ void __launch_bounds__(128) k_boundary(Pair<real> P, BoundaryArgs a)
{
    // dispatch order: the two fault faces first (block 0's +q face, block
    // 1's -q face: their threads run the friction solve, the longest
    // dependency chains), then the other boundary nodes of block 0 and of
    // block 1.  t becomes the block-local boundary-node index.
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nodes[0] + a.nodes[1]) return;
    int me;
    {
        const long F0 = (long)P.b[0].nr * P.b[0].ns, F1 = (long)P.b[1].nr * P.b[1].ns;
        if (t < F0) { me = 0; t = F0 + t; }
        else if (t < F0 + F1) { me = 1; t = t - F0; }
        else {
            t -= F0 + F1;
            if (t < a.nodes[0] - F0) { me = 0; t = t < F0 ? t : t + F0; }
            else { me = 1; t = F1 + (t - (a.nodes[0] - F0)); }
        }
    }
    const Blk<real> &B = P.b[me];
    const int nq = B.nq, nr = B.nr, ns = B.ns;
    // unique enumeration: q-faces, then r-faces without q edges, then s-faces
    int i, j, k;
    long nqf = 2L * nr * ns, nrf = 2L * (nq - 2) * ns;
    if (t < nqf) {
        i = (t < (long)nr * ns) ? 0 : nq - 1;
        long r = t % ((long)nr * ns);
        j = (int)(r % nr); k = (int)(r / nr);
    } else if (t < nqf + nrf) {
        long r = t - nqf;
        j = (r < (long)(nq - 2) * ns) ? 0 : nr - 1;
        r = r % ((long)(nq - 2) * ns);
        i = 1 + (int)(r % (nq - 2)); k = (int)(r / (nq - 2));
    } else {
        long r = t - nqf - nrf;
        k = (r < (long)(nq - 2) * (nr - 2)) ? 0 : ns - 1;
        r = r % ((long)(nq - 2) * (nr - 2));
        i = 1 + (int)(r % (nq - 2)); j = 1 + (int)(r / (nq - 2));
    }
    const long NP = B.NP;
    const long p = pidx(B.px, nr, i, j, k);
    double u[9];
    if (a.fq_read && (i == 0 || i == nq - 1)) {
        const long FPq = (long)nr * ns;
        const real *fq = B.FQ + (i == 0 ? 0 : 9 * FPq) + j + (long)nr * k;
        for (int f = 0; f < 9; f++) u[f] = (double)fq[f * FPq];
    } else {
        for (int f = 0; f < 9; f++) u[f] = (double)B.Fin[f * NP + p];
    }
    double sat[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int pos[3] = {i, j, k};
    const int nn[3] = {nq, nr, ns};
    const double A = a.A[me];
    for (int side = 0; side < 6; side++) {
        int dir = side / 2;
        int on = (side % 2 == 0) ? (pos[dir] == 0) : (pos[dir] == nn[dir] - 1);
        if (!on || B.bc[side] <= 0) continue;
        long fpi;
        long FP;
        if (dir == 0) { fpi = j + (long)nr * k; FP = (long)nr * ns; }
        else if (dir == 1) { fpi = i + (long)nq * k; FP = (long)nq * ns; }
        else { fpi = i + (long)nq * j; FP = (long)nq * nr; }
        const double *fr = B.frame[side], *gm = B.fgm[side], *fm = B.fmat[side];
        double n[3], m[3], l[3];
        for (int c = 0; c < 3; c++) { n[c] = fr[c * FP + fpi]; m[c] = fr[(3 + c) * FP + fpi]; l[c] = fr[(6 + c) * FP + fpi]; }
        double ax = gm[fpi], ay = gm[FP + fpi], az = gm[2 * FP + fpi];
        double norm = sqrt(ax * ax + ay * ay + az * az);
        int code = B.bc[side];
        double r = code == 1 ? 0.0 : code == 2 ? 1.0 : -1.0;
        double ubc[9], Ux[9];
        bc_sat(u, fm[2 * FP + fpi], fm[FP + fpi], fm[fpi], l, m, n, norm, r, side % 2, ubc, Ux);
        double h = 1.0 / (double)(nn[dir] - 1);
        for (int f = 0; f < 9; f++) sat[f] = sat[f] - a.tau0 / h * Ux[f];
        if (B.pml[side] && a.mode == 1) {
            // the PML-slab kernel adds these to DQ (RHS_Interior.f90:1574-1700)
            for (int f = 0; f < 9; f++) B.PQ[side][f * FP + fpi] = (real)(-(a.tau0 / h) * ubc[f]);
        } else if (B.pml[side]) {  // RHS_Interior.f90:1574-1700: SAT on the PML auxiliary rates
            int np = B.npml;
            int lo[3] = {0, 0, 0}, hi[3] = {nq - 1, nr - 1, ns - 1};
            if (side % 2 == 0) hi[dir] = np - 1; else lo[dir] = nn[dir] - np;
            long n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1, n2 = hi[2] - lo[2] + 1;
            long s = (long)(i - lo[0]) + n0 * ((long)(j - lo[1]) + n1 * (long)(k - lo[2]));
            long SP = n0 * n1 * n2;
            for (int f = 0; f < 9; f++) {
                real dsat = (real)(-(a.tau0 / h) * ubc[f]);
                real dq = B.DQ[side][f * SP + s];
                B.DQ[side][f * SP + s] = dq + dsat;
                if (!a.rates_only) B.Q[side][f * SP + s] = B.Q[side][f * SP + s] + (real)a.bdt * dsat;
            }
        }
    }
    // fault face (domain.f90:600-609): block 0's face q = nq-1, block 1's face q = 0
    int on_fault = (me == 0) ? (i == nq - 1 && B.bc[1] == 0) : (i == 0 && B.bc[0] == 0);
    if (on_fault) {
        const double ts = a.tdev ? *a.tdev + a.tinc : a.t;
        // the stage-1 snapshot (slip rate, hat variables, rupture time, peak
        // slip rate) belongs to a step; a rates-only evaluation leaves it
        const int snap = a.stage == 1 && !a.rates_only;
        int rc = fault_node<real>(P, a.I[me], a.fault, me, j, k, ts, snap, A, a.rates_only ? 0.0 : a.bdt,
                                  sat, a.tau0, a.fq_read);
        if (rc) a.flag[0] = rc;
    }
    if (a.mode == 1) {
        real *sb = (real *)a.satb[me];
        const long nb = a.nodes[me];
        for (int f = 0; f < 9; f++) sb[f * nb + t] = (real)sat[f];
        return;
    }
    for (int f = 0; f < 9; f++) {
        real df = B.DF[f * NP + p] + (real)sat[f];
        B.DF[f * NP + p] = df;
        if (!a.rates_only) B.Fout[f * NP + p] = B.Fin[f * NP + p] + (real)a.bdt * df;
    }
}

#pragma clang fp contract(fast)
// ----------------------------------------------------------------------------
// plasticity (plastic.f90:11-151, 153-217, 221-290, 293-352); initial stress
// initial_stress_condition.f90:8-120
struct PlasticArgs {
    int model;          // 1 plastic_flow2, 2 plastic_flow
    double c, nu, Tv, mu, beta, eta, dt;
    int s0_kind;
    double s0[6];
    int x0[2], x1[2], y0[2], y1[2], z0[2], z1[2];
    long n[2];
};

template <typename real>
__global__ void k_plastic(Pair<real> P, PlasticArgs a)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int me = 0;
    if (t >= a.n[0]) { me = 1; t -= a.n[0]; }
    if (t >= a.n[me]) return;
    const Blk<real> &B = P.b[me];
    int w = a.x1[me] - a.x0[me], h = a.y1[me] - a.y0[me];
    int i = a.x0[me] + (int)(t % w), j = a.y0[me] + (int)((t / w) % h), k = a.z0[me] + (int)(t / ((long)w * h));
    const long NP = B.NP;
    const long p = pidx(B.px, B.nr, i, j, k);
    double s0[6];
    if (a.s0_kind == 1) {
        tpv26_stress(B.Y[p], s0);
    } else if (a.s0_kind == 2) {
        tpv29_stress(B.Y[p], s0);
    } else {
        for (int c = 0; c < 6; c++) s0[c] = a.s0[c];
    }
    double sa[6];
    for (int c = 0; c < 6; c++) sa[c] = (double)B.Fout[(3 + c) * NP + p] + s0[c];
    double sigma = (sa[0] + sa[1] + sa[2]) / 3.0;
    double J2 = ((sa[0] - sa[1]) * (sa[0] - sa[1]) + (sa[1] - sa[2]) * (sa[1] - sa[2]) +
                 (sa[2] - sa[0]) * (sa[2] - sa[0])) / 6.0 + sa[3] * sa[3] + sa[4] * sa[4] + sa[5] * sa[5];
    double tau = sqrt(J2);
    const double delta[6] = {1, 1, 1, 0, 0, 0};
    if (a.model == 1) {
        double Ys = fmax(0.0, a.c * cos(atan(a.nu)) - sigma * sin(atan(a.nu)));
        if (tau - Ys <= 0.0) { if (B.P) B.P[p] = 0.0; return; }
        double r = exp(-a.dt / a.Tv) + (1.0 - exp(-a.dt / a.Tv)) * Ys / tau;
        for (int c = 0; c < 6; c++) sa[c] = sigma * delta[c] + r * (sa[c] - sigma * delta[c]);
        // plastic strain accumulator (FP64): equivalent plastic strain increment
        // dgamma = (1 - r) tau / G, G = shear modulus
        if (B.P) {
            double G = B.M ? (double)B.M[NP + p] : B.mu;
            double dg = (1.0 - r) * tau / G;
            B.P[p] = dg / a.dt;
            B.P[NP + p] = B.P[NP + p] + dg;
        }
    } else {
        double Y = tau + a.mu * sigma;
        if (Y <= 0.0) { if (B.P) B.P[p] = 0.0; return; }
        double G = B.M ? (double)B.M[NP + p] : B.mu;
        double lam_e = B.M ? (double)B.M[p] : B.lam;
        double K = lam_e + 2.0 / 3.0 * G;
        double sd[6];
        for (int c = 0; c < 6; c++) sd[c] = sa[c] - sigma * delta[c];
        double lambda = (tau + a.mu * sigma) / (a.eta + a.dt * (G + a.mu * a.beta * K));
        if (B.P) { B.P[p] = lambda; B.P[NP + p] = B.P[NP + p] + a.dt * lambda; }
        double taun = tau - a.dt * lambda * G;
        sigma = sigma - a.dt * lambda * a.beta * K;
        for (int c = 0; c < 6; c++) sa[c] = sigma * delta[c] + sd[c] * taun / (taun + a.dt * lambda * G);
    }
    for (int c = 0; c < 6; c++) B.Fout[(3 + c) * NP + p] = (real)(sa[c] - s0[c]);
    if (B.FQ && (i == 0 || i == B.nq - 1)) {
        real *fq = B.FQ + (i == 0 ? 0L : 9L * B.nr * B.ns) + j + (long)B.nr * k;
        for (int c = 0; c < 6; c++) fq[(long)(3 + c) * B.nr * B.ns] = (real)(sa[c] - s0[c]);
    }
}

// rate-and-state initial state (Interface_Condition.f90:2722-2806)
template <typename real>
__global__ void k_init_state(Pair<real> P, FaultParams f, Iface I0, Iface I1)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int me = 0;
    long n0 = (long)P.b[0].nr * P.b[0].ns;
    if (t >= n0) { me = 1; t -= n0; }
    const Blk<real> &B = P.b[me];
    long FP = (long)B.nr * B.ns;
    if (t >= FP) return;
    int side = me == 0 ? 1 : 0;
    const double *fr = B.frame[side];
    double n[3], m[3], l[3], x[3];
    for (int c = 0; c < 3; c++) {
        n[c] = fr[c * FP + t]; m[c] = fr[(3 + c) * FP + t]; l[c] = fr[(6 + c) * FP + t];
        x[c] = B.fx[c * FP + t];
    }
    if (f.rs_init == 1) {  // constant state ('rate-weakening', 2820-2850)
        (me == 0 ? I0 : I1).W[t] = f.rs_W0;
        return;
    }
    double T0n, T0m, T0l;
    prestress(f, x, 0.0, l, m, n, T0n, T0m, T0l);
    double a = rs_a_at(f, x);
    double sn = -T0n;
    double theta = f.rs_L / f.rs_V0 *
                   exp(((a * log(2.0 * sinh(T0l / (a * sn)))) - f.rs_f0 - a * log(f.rs_Vin / f.rs_V0)) / f.rs_b);
    (me == 0 ? I0 : I1).W[t] = f.rs_f0 + f.rs_b * log(f.rs_V0 / f.rs_L * theta);
}

// ----------------------------------------------------------------------------
// fault-plane statistics of one block's interface copy, one workgroup:
// out = {max slip, max current slip rate, max peak slip rate,
//        ruptured nodes (trup < 1e9), mean slip over ruptured nodes, earliest trup}
// Each wave reduces with cross-lane shuffles (64 lanes), waves combine in LDS.
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void __launch_bounds__(1024) k_fault_stats(Iface I, long FP, double *out)
{
    __shared__ double part[6][16];
    double smax = 0.0, vmax = 0.0, pmax = 0.0, nrup = 0.0, srup = 0.0, tmin = 1.0e9;
    for (long p = threadIdx.x; p < FP; p += blockDim.x) {
        double sl = I.S[3 * FP + p];
        double v1 = I.Svel[FP + p], v2 = I.Svel[2 * FP + p];
        smax = fmax(smax, sl);
        vmax = fmax(vmax, sqrt(v1 * v1 + v2 * v2));
        pmax = fmax(pmax, I.psr[p]);
        double tr = I.trup[p];
        if (tr < 1.0e9) { nrup += 1.0; srup += sl; tmin = fmin(tmin, tr); }
    }
    smax = wave_max(smax); vmax = wave_max(vmax); pmax = wave_max(pmax);
    nrup = wave_sum(nrup); srup = wave_sum(srup); tmin = wave_min(tmin);
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64, nw = blockDim.x / 64;
    if (lane == 0) {
        part[0][w] = smax; part[1][w] = vmax; part[2][w] = pmax;
        part[3][w] = nrup; part[4][w] = srup; part[5][w] = tmin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 1.0e9};
        for (int k = 0; k < nw; k++) {
            r[0] = fmax(r[0], part[0][k]); r[1] = fmax(r[1], part[1][k]); r[2] = fmax(r[2], part[2][k]);
            r[3] += part[3][k]; r[4] += part[4][k]; r[5] = fmin(r[5], part[5][k]);
        }
        out[0] = r[0]; out[1] = r[1]; out[2] = r[2]; out[3] = r[3];
        out[4] = r[3] > 0.0 ? r[4] / r[3] : 0.0;
        out[5] = r[5];
    }
}

// dense [nf][ns][nr][nq] FP64 -> padded [nf][ns][nr][px] volume array
template <typename real>
__global__ void k_pad(int nf, int nq, int nr, int ns, int px, const double *__restrict__ src, real *__restrict__ dst)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long rows = (long)nf * nr * ns;
    if (t >= rows * px) return;
    const long row = t / px;
    const int i = (int)(t - row * px);
    dst[t] = i < nq ? (real)src[row * nq + i] : (real)0;
}

// padded [nf][ns][nr][px] volume array -> dense [nf][ns][nr][nq] FP64
template <typename real>
__global__ void k_unpad(int nf, int nq, int nr, int ns, int px, const real *__restrict__ src, double *__restrict__ dst)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long rows = (long)nf * nr * ns;
    if (t >= rows * nq) return;
    const long row = t / nq;
    const int i = (int)(t - row * nq);
    dst[t] = (double)src[row * px + i];
}

// min and max of n values (out[0] min, out[1] max; the caller seeds them
// with +inf and -inf): the spread of a metric field at set-up
__global__ void __launch_bounds__(256) k_range(long n, const double *__restrict__ a, double *out)
{
    __shared__ double smin[256], smax[256];
    double lo = INFINITY, hi = -INFINITY;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        const double v = a[p];
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMin(&out[0], smin[0]);
        atomicMax(&out[1], smax[0]);
    }
}

// refresh the q-face buffer of a block from its fields: [2][9][nr ns]
template <typename real>
__global__ void k_fq_gather(int nq, int nr, int ns, int px, long NP, const real *__restrict__ F, real *__restrict__ FQ)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long FP = (long)nr * ns;
    if (t >= 18 * FP) return;
    const long sf = t / FP, rem = t - sf * FP;
    const int side = (int)(sf / 9), f = (int)(sf % 9);
    const int j = (int)(rem % nr), k = (int)(rem / nr);
    const int i = side == 0 ? 0 : nq - 1;
    FQ[t] = F[(long)f * NP + i + (long)px * ((long)j + (long)nr * k)];
}

// the nine fields of the q-face i of a block in FP64, [field][s][r] (the
// layout of the reference's fault output records)
template <typename real>
__global__ void k_qface(int i, int nr, int ns, int px, long NP, const real *__restrict__ F, double *__restrict__ out)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long FP = (long)nr * ns;
    if (t >= 9 * FP) return;
    const long f = t / FP, rem = t - f * FP;
    const int j = (int)(rem % nr), k = (int)(rem / nr);
    out[t] = (double)F[f * NP + i + (long)px * ((long)j + (long)nr * k)];
}

// ----------------------------------------------------------------------------
// moment-tensor point sources (moment_tensor.f90:958-1270)
// (FP64, no FMA contraction: the reference's operation order)

// source time function at t (source_time, 958-1017)
__device__ double source_time_fn(int type, double t, double t0, double w)
{
#pragma clang fp contract(off)
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

// one 1-D discrete delta weight at node i for a source at node i0 with
// offset a (in units of h) (singular_source, 1083-1270): five nodes for
// order 2, the nine-node table of alpha for order 4
__device__ double singular_weight(int i, int i0, double h, double a, int order, int alpha)
{
#pragma clang fp contract(off)
    const int k = i - i0;
    if (order == 2) {
        if (k < -2 || k > 2) return 0.0;
        const double beta = (1.0 - 2.0 * a) / 4.0 + 0.25;
        const double gam = (2.0 * a - 1.0) / 16.0;
        const double gap = (2.0 * a + 1.0) / 16.0;
        switch (k) {
        case -2: return (-gam) / h;
        case -1: return (beta + 3.0 * gam - gap) / h;
        case 0: return (1.0 - a - 2.0 * beta - 3.0 * gam + 3.0 * gap) / h;
        case 1: return (a + beta + gam - 3.0 * gap) / h;
        default: return gap / h;
        }
    }
    if (k < -4 || k > 4) return 0.0;
    static constexpr double W[4][9] = {
        {-0.01171875, -0.03125, 0.046875, 0.28125, 0.4296875, 0.28125, 0.046875, -0.03125, -0.01171875},
        {-0.00246875, -0.02275, -0.014125, 0.16575, 0.4171875, 0.38675, 0.117875, -0.02975, -0.01846875},
        {-0.00728125, -0.028, 0.014875, 0.224, 0.4265625, 0.336, 0.081375, -0.032, -0.01553125},
        {-0.0095625, -0.02990625, 0.03053125, 0.25284375, 0.42890625, 0.30903125, 0.06384375, -0.03196875,
         -0.01371875}};
    const int row = alpha == 0 ? 0 : alpha == 25 ? 1 : alpha == 50 ? 2 : 3;
    return W[row][k + 4] / h;
}

struct SourceArgs {
    const SourceDev *src;
    int n, order, alpha;
    int lo[3], hi[3];      // node box holding every source's support (hi exclusive)
    double t;
    const double *tdev;    // non-null: stage time = *tdev + tinc
    double tinc;
    double bdt;
    // the Gaussian-mollified form (set_moment_tensor_smooth): every source
    // spreads over the whole block; X [3][ns][nr][nq] dense node coordinates
    int smooth;
    const double *X;
};

// DF -= A M delta at the nodes of the box, sources in list order, and the
// stage output and q-face buffer follow (F_out = F_in + B dt DF)
template <typename real>
__global__ void k_sources(Blk<real> B, SourceArgs a)
{
#pragma clang fp contract(off)
    const int bx = a.hi[0] - a.lo[0], by = a.hi[1] - a.lo[1];
    const long nbox = (long)bx * by * (a.hi[2] - a.lo[2]);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbox) return;
    const int i = a.lo[0] + (int)(t % bx), j = a.lo[1] + (int)((t / bx) % by), k = a.lo[2] + (int)(t / ((long)bx * by));
    const long p = pidx(B.px, B.nr, i, j, k);
    const double ts = a.tdev ? *a.tdev + a.tinc : a.t;
    // local spacing (moment_tensor.f90:1044-1047): h / (q_x, r_y, s_z)
    const double hx = (1.0 / (double)(B.nq - 1)) / gmet(B, 0, p);
    const double hy = (1.0 / (double)(B.nr - 1)) / gmet(B, 4, p);
    const double hz = (1.0 / (double)(B.ns - 1)) / gmet(B, 8, p);
    double df[6];
    for (int c = 0; c < 6; c++) df[c] = (double)B.DF[(3 + c) * B.NP + p];
    const double df0[6] = {df[0], df[1], df[2], df[3], df[4], df[5]};
    bool hit = false;
    if (a.smooth) {
        // set_moment_tensor_smooth (moment_tensor.f90:1400-1475): a normal
        // density in each coordinate with standard deviation twice the local
        // spacing (backward differences of X along q, Y along r, Z along s;
        // forward at the first node), with the reference's pi = 3.14159265359
        const long N = (long)B.nq * B.nr * B.ns, rs = B.nq, ss = (long)B.nq * B.nr;
        const long d = i + rs * j + ss * k;
        const double *X = a.X, *Y = a.X + N, *Z = a.X + 2 * N;
        const double hx = i == 0 ? X[d + 1] - X[d] : X[d] - X[d - 1];
        const double hy = j == 0 ? Y[d + rs] - Y[d] : Y[d] - Y[d - rs];
        const double hz = k == 0 ? Z[d + ss] - Z[d] : Z[d] - Z[d - ss];
        const double sx = 2.0 * hx, sy = 2.0 * hy, sz = 2.0 * hz;
        const double r2pi = sqrt(2.0 * 3.14159265359);
        for (int s = 0; s < a.n; s++) {
            const SourceDev &S = a.src[s];
            const double ux = (X[d] - S.x[0]) / sx, uy = (Y[d] - S.x[1]) / sy, uz = (Z[d] - S.x[2]) / sz;
            const double dx = exp(-(0.5 * (ux * ux))) / (sx * r2pi);
            const double dy = exp(-(0.5 * (uy * uy))) / (sy * r2pi);
            const double dz = exp(-(0.5 * (uz * uz))) / (sz * r2pi);
            const double A = source_time_fn(S.type, ts, S.t_init, S.duration);
            const double delta = dx * dy * dz;
            if (delta == 0.0) continue;
            hit = true;
            for (int c = 0; c < 6; c++) df[c] = df[c] - A * S.m[c] * delta;
        }
    }
    for (int s = 0; s < (a.smooth ? 0 : a.n); s++) {
        const SourceDev &S = a.src[s];
        const double ax = (S.x[0] - S.node_x[0]) / hx;
        const double ay = (S.x[1] - S.node_x[1]) / hy;
        const double az = (S.x[2] - S.node_x[2]) / hz;
        const double A = source_time_fn(S.type, ts, S.t_init, S.duration);
        const double dx = singular_weight(i, S.node[0], hx, ax, a.order, a.alpha);
        const double dy = singular_weight(j, S.node[1], hy, ay, a.order, a.alpha);
        const double dz = singular_weight(k, S.node[2], hz, az, a.order, a.alpha);
        const double delta = dx * dy * dz;
        if (delta == 0.0) continue;
        hit = true;
        for (int c = 0; c < 6; c++) df[c] = df[c] - A * S.m[c] * delta;
    }
    if (!hit) return;
    real *fq = nullptr;
    if (B.FQ && (i == 0 || i == B.nq - 1))
        fq = B.FQ + (i == 0 ? 0L : 9L * B.nr * B.ns) + j + (long)B.nr * k;
    for (int c = 0; c < 6; c++) {
        const long q = (3 + c) * B.NP + p;
        B.DF[q] = (real)df[c];
        const real fo = B.Fout[q] + (real)(a.bdt * (df[c] - df0[c]));
        B.Fout[q] = fo;
        if (fq) fq[(long)(3 + c) * B.nr * B.ns] = fo;
    }
}
// velocities at the station nodes (write_seismogram, seismogram.f90:460-478):
// out[n][3] = vx vy vz of node off[n]
template <typename real>
__global__ void k_station_gather(int n, const long *__restrict__ off, long NP, const real *__restrict__ F,
                                 double *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    out[t] = (double)F[(long)(t % 3) * NP + off[t / 3]];
}

// step time advance of a captured step graph: t = t + inc, the same FP64
// operation as the host-side clock
__global__ void k_advance_time(double *t, double inc)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) t[0] = t[0] + inc;
}

__global__ void k_set_time(double *t, double v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) t[0] = v;
}

template <typename real>
__global__ void k_fill(long n, real *__restrict__ a, real v)
{
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) a[p] = v;
}

}  // namespace wql
