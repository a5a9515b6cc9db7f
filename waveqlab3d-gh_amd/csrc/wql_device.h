// wql_device.h -- device-side data structures and small helpers shared by the
// kernels of libwql (gfx950).  Host code fills these structs; kernels take
// them by value.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wql_sbp_ops.h"

namespace wql {

constexpr int NF = 9;  // vx vy vz sxx syy szz sxy sxz syz (RHS_Interior.f90:241-251)

// 6th-order SBP first-derivative closure (JU_xJU_yJU_z6.f90:77-298,
// metrics.f90:16-160): boundary rows 0..5 and their mirror rows n-6..n-1, as
// tap offsets relative to the row and coefficients, in the order the
// reference sums them.  Filled once per process by wql_create.
struct SbpRow {
    int n;
    int off[8];
    double c[8];
};
struct SbpTable {
    SbpRow left[6];
    SbpRow right[6];
};

// Fault description, identical in meaning to wql_fault_params.
struct FaultParams {
    int law;
    int stress_mode;
    double T0[3];
    double s0[6];
    int npatch;
    double patch[4][5];
    int nuc;
    double nucp[5];
    double fs, fd, Dc;
    double bar_wy, bar_wz, bar_big;
    double rs_a, rs_da, rs_w, rs_wr, rs_b, rs_L, rs_V0, rs_f0, rs_Vin;
    int rs_law;              // 0 aging, 1 slip law + flash weakening, 2 KDN
    double rs_fw, rs_Vw, rs_dVw;
    int rs_init;             // 0 steady state from the prestress, 1 constant rs_W0
    double rs_W0;
    int sw_law;              // 0 linear slip weakening, 1 TPV26/27 cohesion + forced rupture
    double coh[3];
    double frc[5];
    int rs_solver;           // 0 regula falsi (reference), 1 bracketed Newton
    int key;                 // keyed problem whose fault fields are per-node tables (0: none)
};

// One moment-tensor point source, identical in meaning to wql_source.
struct SourceDev {
    int type;
    double duration, t_init;
    double m[6];
    double x[3];
    int node[3];
    double node_x[3];
};

// One block as the kernels see it.
template <typename real>
struct Blk {
    int nq, nr, ns;
    int px;                  // row pitch of the volume arrays (>= nq, rows 128-byte aligned)
    long NP;                 // elements per field: px*nr*ns
    const real *Fin;         // [9][NP] stage input
    real *Fout;              // [9][NP] stage output
    real *DF;                // [9][NP] low-storage RK rates
    // volume arrays are [field][s][r][q] with q rows padded to px elements
    const real *G;           // [10][NP] metrics + J (curvilinear); nullptr for Cartesian
    const real *M;           // [3][NP] lambda mu rho (heterogeneous); nullptr if homogeneous
    double gc[10];           // constant metrics + J (Cartesian; constant fields of G)
    unsigned gconst;         // bit c: metric field c of G holds gc[c] at every node
    double lam, mu, rho;     // homogeneous material
    int cart;                // 1: constant metrics
    int bc[6];
    int pml[6];
    int npml;
    real *Q[6], *DQ[6];      // PML slabs [9][slab nodes]
    real *PQ[6];             // PML face penalties [9][face nodes] of the stage (written by the
                             // boundary kernel, added to DQ by the PML-slab kernel); nullptr = unused
    // face data (FP64): per side, nodes in face layout (q: j + nr*k, r: i + nq*k, s: i + nq*j)
    const double *frame[6];  // [9][face] n m l
    const double *fgm[6];    // [3][face] metric components of the face axis (a_x, a_y, a_z)
    const double *fmat[6];   // [3][face] lambda mu rho
    const double *fx;        // [3][nr*ns] coordinates of the fault face (q-face carrying bc==0)
    double *P;               // [2][NP] plastic strain rate / accumulated plastic strain (FP64)
    const double *Y;         // [NP] y-coordinate (plastic depth-dependent background stress)
    // boundary penalty of this stage per boundary node, added by the volume
    // kernels when they form DF (nullptr: the boundary kernel applies it)
    const real *satb;        // [9][nbnd]
    long nbnd;               // boundary nodes: 2 nr ns + 2 (nq-2) ns + 2 (nq-2)(nr-2)
    // the nine fields on the q faces, [2][9][nr ns] (face q = 0, then q = nq-1;
    // node j + nr k): written with the stage output by the volume kernels,
    // read as the next stage's input by the boundary kernel (nullptr: unused)
    real *FQ;
    // upwind / upwind-DRP operator (device copy of WQL_SBP_OPS[k].v[0..2]
    // then .s[0..2]) and its band width; nullptr: the sixth-order central one
    const wql_sbp_half *sbp;
    int sbw;
    // anelastic response: memory variables and their rates [24][NP]
    // (component c of sxx..syz, mechanism l at field 4c + l), Q_s^-1 and
    // Q_p^-1 [2][NP], relaxation times and weights; eta == nullptr: elastic
    real *eta, *deta;
    const real *qinv;
    double an_tau[4], an_w[4];
};

// index of node (i, j, k) in the boundary-node enumeration (q faces, then r
// faces without q edges, then s faces without q/r edges), -1 inside
template <typename real>
__device__ __forceinline__ long bnd_index(const Blk<real> &B, int i, int j, int k)
{
    const int nq = B.nq, nr = B.nr, ns = B.ns;
    const long nqf = 2L * nr * ns, nrf = 2L * (nq - 2) * ns;
    if (i == 0) return (long)j + (long)nr * k;
    if (i == nq - 1) return (long)nr * ns + j + (long)nr * k;
    if (j == 0) return nqf + (i - 1) + (long)(nq - 2) * k;
    if (j == nr - 1) return nqf + (long)(nq - 2) * ns + (i - 1) + (long)(nq - 2) * k;
    if (k == 0) return nqf + nrf + (i - 1) + (long)(nq - 2) * (j - 1);
    if (k == ns - 1) return nqf + nrf + (long)(nq - 2) * (nr - 2) + (i - 1) + (long)(nq - 2) * (j - 1);
    return -1;
}

// One block's copy of the interface state (iface.f90:13-48), FP64.
struct Iface {
    double *S, *DS;          // [4][nr*ns]
    double *W, *DW;          // [nr*ns]
    double *trup;            // [nr*ns]
    double *Svel;            // [3][nr*ns]
    double *Uhat, *Vhat;     // [6][nr*ns]
    double *psr;             // [nr*ns] peak tangential slip rate over the run (stage-1 samples)
    // keyed problems: [8][nr*ns] static prestress T0_n, T0_m, T0_l, strength
    // fs, fd, Dc, cohesion C0 (barriers included) and the time of forced
    // rupture of this block's fault face (null otherwise)
    double *ftab;
};

__device__ __forceinline__ long pidx(int nq, int nr, int i, int j, int k)
{
    return (long)i + (long)nq * ((long)j + (long)nr * (long)k);
}

}  // namespace wql
