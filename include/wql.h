/*
 * wql.h -- C ABI of the MI355X WaveQLab3D time-stepping library (libwql.so).
 *
 * The library replaces the reference's per-step loop
 *     call time_step_RK(D, D%dt, RK, n)          (main.f90:85, time_step.f90:68-150)
 * for a two-block domain joined by one frictional fault (domain.f90:243-244).
 * The whole block state lives in HBM between calls; the host hands over plain
 * arrays in the reference's column-major layout (q fastest, then r, s, field:
 * datatypes.f90, mpi3dcomm.f90:736) restricted to the owned nodes.
 *
 * Every entry returns 0 on success or a negative WQL_E* code; the message of
 * the last failure on the calling thread is wql_last_error().  Calls on one
 * context are not reentrant.  No host pointer is retained after a call.
 */
#ifndef WQL_H
#define WQL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WQL_ABI_VERSION 4

enum {
    WQL_OK = 0,
    WQL_E_ARG = -1,        /* invalid descriptor or argument */
    WQL_E_NOMEM = -2,      /* device allocation failed / over the HBM budget */
    WQL_E_HIP = -3,        /* HIP runtime error */
    WQL_E_NONFINITE = -4,  /* NaN/Inf detected in the fault state */
    WQL_E_SOLVER = -5,     /* friction solve failed (negative slip rate) */
    WQL_E_STATE = -6       /* call out of order */
};

enum { WQL_FP64 = 0, WQL_FP32 = 1 };              /* volume precision */
enum { WQL_LOCKED = 0, WQL_SLIP_WEAKENING = 1, WQL_RATE_STATE = 2 };

/* One block: replaces the fields of block_type the hot path reads
 * (datatypes.f90:118-135; block.f90:27-269 fills them). */
typedef struct {
    int nq, nr, ns;        /* owned nodes per direction (each >= 13) */
    int bc[6];             /* Lx Rx Ly Ry Lz Rz: 0 fault, 1 absorbing, 2 free surface, 3 clamped
                              (BoundaryConditions.f90:15-48) */
    int pml[6];            /* PML slab per side (block.f90:160-178) */
    int npml;              /* PML thickness in nodes (pml.f90:11) */
    int cartesian;         /* 1: uniform Cartesian block (type_of_mesh='cartesian') */
    const double *X;       /* [3][ns][nr][nq] node coordinates (grid.f90:13-395) */
    const double *M;       /* [3][ns][nr][nq] lambda, mu, rho; NULL -> homogeneous rho_s_p */
    double rho_s_p[3];     /* rho, cs, cp (material.f90:181-188) */
} wql_block_desc;

/* Fault law and prestress (Interface_Condition.f90:1268-2476 keyed tables,
 * made explicit).  Same meaning as the oracle's ora_fault. */
typedef struct {
    int law;               /* WQL_LOCKED / WQL_SLIP_WEAKENING / WQL_RATE_STATE */
    int stress_mode;       /* 0: tractions T0 + boxcar patches; 1: tensor s0 on the face frame;
                              2: TPV26/27 depth-dependent tensor on the face frame
                              (prestress 2008-2048, initial_stress_condition.f90:26-76) */
    double T0[3];
    double s0[6];
    int npatch;
    double patch[4][5];    /* (yc, zc, wy, wz, amplitude) added to T0_l */
    int nuc;               /* 0 none, 1 cosine disk (yc,zc,r1,width,amp), 2 smooth bump (yc,zc,R,amp,T),
                              3 Gaussian pulse (yc,zc,r0,c,-): T0_l += 2 c T0_n exp(-(r/r0)^2/2) for t > 0
                              ('rate-weakening', 2387-2420) */
    double nucp[5];
    double fs, fd, Dc;
    double bar_wy, bar_wz, bar_big;
    /* rate and state: a = rs_a + rs_da (1 - B(y) B(z)), B the smooth boxcar of
     * half-width rs_w and ramp rs_wr (386-416) */
    double rs_a, rs_da, rs_w, rs_wr, rs_b, rs_L, rs_V0, rs_f0, rs_Vin;
    /* ---- ABI 2 ---- */
    int rs_law;            /* state evolution: 0 aging law (TPV101/102, 431-436), 1 slip law with
                              flash weakening ('rate-weakening', 438-444), 2 KDN (446-449) */
    double rs_fw;          /* flash weakening: Fss <- fw + (Fss - fw) / (1 + (V/Vw)^8)^(1/8) */
    double rs_Vw, rs_dVw;  /* Vw = rs_Vw + rs_dVw (1 - B(y) B(z)); rs_Vw <= 0: no flash weakening */
    int rs_init;           /* wql_init_state: 0 steady state from the prestress and rs_Vin
                              (2756-2819), 1 constant state rs_W0, zero background velocity (2820-2850) */
    double rs_W0;
    int sw_law;            /* slip-weakening strength: 0 sigma_n (fs - (fs-fd) min(D,Dc)/Dc)
                              (1283-1299); 1 TPV26/27: C0 + sigma_n (fs - (fs-fd) max(min(D/Dc,1), f2))
                              with cohesion C0 and forced rupture f2 (1376-1458) */
    double coh[3];         /* C0 = coh[0] + coh[1] (coh[2] - y) for y <= coh[2], else coh[0] (2491-2530) */
    double frc[5];         /* forced rupture (yc, zc, r_crit, Vs, t0) (2625-2660) */
    int rs_solver;         /* slip-rate solve: 0 regula falsi, the reference's iteration step for step
                              (1170-1239); 1 bracketed Newton, |step| <= 1e-12 |V| + 1e-300 */
    /* ---- ABI 4 ---- */
    int key;               /* WQL_KEY_*: a slip-weakening problem the reference keys by name whose
                              prestress, strength, cohesion and forced-rupture time the library
                              evaluates at every fault node at create (prestress 1802-2435,
                              Tau_strength 1268-1797, Frictional_Cohesion 2491-2580,
                              Time_of_Forced_Rupture 2583-2660, the fault-node material of
                              Interface_RHS 100-260, initial_stress_tensor in
                              initial_stress_condition.f90:87-240); needs law WQL_SLIP_WEAKENING,
                              and then no other prestress or strength field is read.  0: none */
} wql_fault_params;

/* wql_fault_params.key: TPV30 is TPV29's fault, TPV32 is TPV31's with its own
 * fault-node material */
enum {
    WQL_KEY_NONE = 0, WQL_KEY_TPV10 = 1, WQL_KEY_TPV29 = 2, WQL_KEY_TPV31 = 3, WQL_KEY_TPV32 = 4,
    WQL_KEY_TPV33 = 5, WQL_KEY_TPV34 = 6, WQL_KEY_SCITS2016 = 7, WQL_KEY_TPV36 = 8, WQL_KEY_TPV37 = 9,
    WQL_KEY_WASATCH100 = 10
};

enum { WQL_FD_TRADITIONAL = 0, WQL_FD_UPWIND = 1, WQL_FD_UPWIND_DRP = 2 };

typedef struct {
    int precision;         /* WQL_FP64, or WQL_FP32: volume fields/metrics FP32, fault state FP64 */
    double dt;             /* time step (domain.f90:222-235) */
    double t0;             /* initial time */
    int plastic;           /* 0 elastic, 1 Drucker-Prager viscoplastic ('scec', plastic.f90:153-217),
                              2 default model (plastic.f90:221-290) */
    double pl_c, pl_nu, pl_Tv;
    double pl_mu, pl_beta, pl_eta;
    int s0_kind;           /* plastic background stress: 0 constant fault.s0, 1 TPV26/27 depth profile,
                              2 TPV29/30 depth profile (initial_stress_condition.f90:26-112) */
    wql_fault_params fault;
    /* ---- ABI 3 ---- */
    int fd_type;           /* &problem_list fd_type (domain.f90:106, 135-136): WQL_FD_TRADITIONAL (the
                              sixth-order central operator, JU_xJU_yJU_z6.f90:12-812, 16141-16449),
                              WQL_FD_UPWIND or WQL_FD_UPWIND_DRP: forward operator Dp on the velocities,
                              backward operator Dm on the stress fluxes (815-16122, 16451-20922) */
    int fd_order;          /* &problem_list order: 2..9 (upwind), 3..7, 66, 679 (upwind_drp); it also sets
                              the SAT penalty tau0 and the band width (block.f90:103-255,
                              RHS_Interior.f90:133-1514).  Ignored for WQL_FD_TRADITIONAL */
    int anelastic;         /* 1: response = 'anelastic' (material.f90:11-145): four GSLS mechanisms per
                              stress component, memory variables in every volume kernel
                              (RHS_Interior.f90:2390-2532); not together with plastic */
    double an_c;           /* &anelastic_list c: Q_s = c v_s (v_s in km/s), Q_p = Q_s / 2 */
    double an_weight_exp;  /* &anelastic_list weight_exp: < 0.01 or in (0.59, 0.61), the two weight sets
                              the reference defines (material.f90:80-106) */
    double an_fref;        /* &anelastic_list fref: reference frequency of the unrelaxed moduli */
} wql_problem;

/* The fault law, prestress and plasticity settings the reference keys by
 * name: &problem_list's problem, coupling, response and plastic_model, and
 * btp%mu_beta_eta (Interface_Condition.f90:350-450, 1268-2476, 2722-2850;
 * initial_stress_condition.f90; plastic.f90:174-182; plastic_material.f90:
 * 38-44) for TPV5, TPV28, TPV26/27 (slip weakening), TPV101/102 and
 * 'rate-weakening' (rate and state), 'locked' and the 'scec' / default
 * plasticity models; response 'anelastic' sets anelastic with the
 * &anelastic_list defaults (c 1, weight_exp 0, fref 1), which the caller
 * replaces by the list's values.  Fills out->fault and the plastic and
 * anelastic fields; keeps out->precision, dt and t0 (the operator fields
 * are zeroed: the caller sets fd_type and fd_order).  WQL_E_ARG for keys
 * outside that set (among them 'KDN', which the reference cannot run: it
 * has no prestress). */
int wql_problem_keys(const char *problem, const char *coupling, const char *response, const char *plastic_model,
                     const double mu_beta_eta[3], wql_problem *out);

typedef struct wql_ctx wql_ctx;

/* Node coordinates of one block, generated on the device: the uniform grid
 * of cartesian_grid_3d (grid.f90:213-240) or the transfinite interpolation
 * of Curve_Grid_3D3 (grid.f90:471-1907; faces Gen_Surf_Mesh 2418-2455,
 * volume Gen3_Curve_Mesh 2290-2416) with planar q faces, the
 * 'analytical_test_problem' bump (775-831) or the 'analytical_tpv28',
 * 'analytical_tpv10', 'analytical_tpv36' fault profiles (625-773), plus
 * optional offset fields of one q face (a fault surface, such as the x column
 * of a 'read_from_memomry_fractal' file times lc or rc, 832-925) and of the
 * r = 0 face (topography), the surfaces the reference reads from files
 * through the same interpolation. */
typedef struct {
    int nq, nr, ns;
    double a[3], b[3];     /* btp%aqrs, btp%bqrs: corners of the block */
    int mesh;              /* 0 uniform ('cartesian'), 1 transfinite ('curvilinear') */
    int profile;           /* q faces: 0 planar, 1 'analytical_test_problem', 2 'analytical_tpv28',
                              3 'analytical_tpv10', 4 'analytical_tpv36', each weighted by lc / rc */
    double lc, rc;         /* btp%lc, btp%rc */
    const double *fault_x; /* optional [ns][nr] x offsets of the q face fault_side (mesh 1, profile 0) */
    int fault_side;        /* 0: q = 0 face, 1: q = nq-1 face */
    const double *top_y;   /* optional [ns][nq] y offsets of the r = 0 face (mesh 1) */
} wql_grid_desc;
/* X[3][ns][nr][nq] (the reference's X(q,r,s,1:3)) */
int wql_grid_block(const wql_grid_desc *grid, double *X);

/* Gen3_Curve_Mesh (grid.f90:2290-2416) on the device from given faces:
 * X[3][ns][nr][nq] from faces[3][...], per component x, y, z the q = 0 and
 * q = nq-1 faces [ns][nr], the r = 0 and r = nr-1 faces [ns][nq], the s = 0
 * and s = ns-1 faces [nr][nq].  The faces of the topography branch of
 * Curve_Grid_3D3 (use_topography, grid.f90:1349-1860) are built from
 * surface data (the topography and fault-profile files); the volume
 * interpolation, the O(n^3) part, runs here. */
int wql_grid_faces(int nq, int nr, int ns, const double *faces, double *X);

/* init_material's media keyed by problem name (material.f90:179-690): for
 * 'LOH1', 'LOH1_Harmonic', 'TPV31', 'TPV32', 'TPV33' and 'OKLAHOMA' fills
 * M[3][ns][nr][nq] (lambda, mu, rho, as wql_block_desc.M) from the node
 * coordinates X[3][ns][nr][nq] and returns 1; returns 0 for any other name
 * (the homogeneous rho_s_p medium) without touching M.  Host only: no
 * device call. */
int wql_problem_material(const char *problem, int nq, int nr, int ns, const double *X, double *M);

/* replaces init_domain's device-relevant part (domain.f90:31-564): uploads
 * geometry/material, computes metrics and face frames on the device, zeroes
 * the fields.  blocks[0] is the block on the minus side of the fault. */
int wql_create(const wql_problem *prob, const wql_block_desc blocks[2], wql_ctx **out);
int wql_destroy(wql_ctx *ctx);
const char *wql_last_error(void);
int wql_abi_version(void);

/* device memory the context holds, bytes */
size_t wql_device_bytes(const wql_ctx *ctx);
/* bytes a context for this problem would need (admission control) */
size_t wql_estimate_bytes(const wql_problem *prob, const wql_block_desc blocks[2]);

/* Init_Vel_State (Interface_Condition.f90:2722-2850): state variable and
 * background velocity for rate-and-state faults. */
int wql_init_state(wql_ctx *ctx);

/* field upload/download, F[9][ns][nr][nq] in the reference field order
 * vx vy vz sxx syy szz sxy sxz syz (RHS_Interior.f90:241-251) */
int wql_set_fields(wql_ctx *ctx, int block, const double *F);
int wql_get_fields(wql_ctx *ctx, int block, double *F);
int wql_get_rates(wql_ctx *ctx, int block, double *DF);
/* anelastic memory variables eta (what 0) or their rates (what 1) of a
 * block, [24][ns][nr][nq]: stress component sxx syy szz sxy sxz syz
 * (M%eta4..eta9, datatypes.f90:50-51) times mechanism 1..4.  WQL_E_ARG
 * for a problem without anelastic response. */
int wql_set_anelastic(wql_ctx *ctx, int block, const double *eta);
int wql_get_anelastic(wql_ctx *ctx, int block, int what, double *out);
/* wql_get_fields without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  It stages through a fixed
 * per-GPU device array of 256 MiB that the library holds until the process
 * ends, so a block of at most 3,728,270 nodes (9 x 8-byte fields); larger
 * blocks get WQL_E_ARG.  All calls on one GPU share that array: calls that
 * overlap on one GPU (from several threads) race. */
int wql_get_fields_noalloc(wql_ctx *ctx, int block, double *F);
/* PML auxiliary fields of one slab [9][slab nodes] (pml.f90:11-168) */
int wql_set_pml(wql_ctx *ctx, int block, int side, const double *Q);
int wql_get_pml(wql_ctx *ctx, int block, int side, double *Q);
int wql_get_pml_rates(wql_ctx *ctx, int block, int side, double *DQ);
/* metrics [10][ns][nr][nq] = q_x r_x s_x q_y r_y s_y q_z r_z s_z J (grid.f90:396-469) */
int wql_get_metrics(wql_ctx *ctx, int block, double *metric);
/* face frame of one side [9][face nodes] = n, m, l (unit_normals.f90:8-168) */
int wql_get_frame(wql_ctx *ctx, int block, int side, double *nml);

/* The nine fields of both blocks on the fault: Uface is block 0's +q face,
 * Vface block 1's -q face, each [9][ns][nr] in FP64.  These are the arrays
 * write_fault_output hands to the fault writer, F(pq1,:,:,1:9) and
 * F(mq2,:,:,1:9) (domain.f90:741-779, fault_output.f90:72-95).  Either
 * pointer may be NULL. */
int wql_get_fault_faces(wql_ctx *ctx, double *Uface, double *Vface);

/* fault state of one block's copy (iface.f90:13-48): S[4], W[1], trup[1], Svel[3],
 * Uhat[6], Vhat[6] on [nr*ns] (q-face layout [f][s][r]); any pointer may be NULL */
int wql_set_fault(wql_ctx *ctx, int block, const double *S, const double *W);
int wql_get_fault(wql_ctx *ctx, int block, double *S, double *W, double *trup, double *Svel,
                  double *Uhat, double *Vhat);
int wql_get_fault_rates(wql_ctx *ctx, int block, double *DS, double *DW);
/* plastic state of one block [2][ns][nr][nq] in FP64, the reference's
 * P%P(:,:,:,1:2) (datatypes.f90 block_plastic, plastic_material.f90:19-60):
 * field 0 the plastic strain rate lambda of the last update, field 1 the
 * accumulated plastic strain gammap (plastic.f90:269-270).  Model 1
 * ('scec', plastic_flow2) has no accumulator in the reference; there the
 * library keeps lambda = (1 - r) tau / (G dt) and gammap += dt lambda.
 * WQL_E_STATE if the problem is elastic. */
int wql_get_plastic(wql_ctx *ctx, int block, double *P);
/* moment-tensor point sources (moment_tensor.f90:82-1270): each source adds
 * -A(t) M delta to the stress rates of the nodes around its nearest node,
 * delta the product of three 1-D discrete deltas (singular_source,
 * moment_tensor.f90:1083-1270: five nodes with the source's offset for
 * order 2, nine nodes from a table chosen by `alpha` for order 4), A the
 * source time function at the stage time (source_time, 958-1017).  Applied
 * in every RK stage after the volume rates (set_rates, domain.f90:908-951). */
enum {
    WQL_STF_GAUSSIAN = 0,          /* 'gaussian' */
    WQL_STF_LOH_DISCONTINUITY = 1, /* 'LOH_discontinuity' */
    WQL_STF_BRUNE = 2,             /* 'Brune' */
    WQL_STF_COS_BELL = 3,          /* 'Cos-Bell' */
    WQL_STF_COSINE = 4,            /* 'cosine' */
    WQL_STF_LOH_MOMENT = 5,        /* 'LOH_moment_time_history' */
    WQL_STF_SINE_MOMENT = 6        /* 'sine_moment_time_history' */
};
typedef struct {
    int type;              /* WQL_STF_* */
    double duration;       /* w of the source time function */
    double t_init;         /* t0 */
    double m[6];           /* mXX mYY mZZ mXY mXZ mYZ */
    double x[3];           /* location */
    int node[3];           /* 0-based nearest node; -2 in each component where the block does not
                              hold the source (the reference's index -1, kept: its 1-D deltas
                              still reach the block's first nodes) */
    double node_x[3];      /* coordinates of that node (-1 where none, as the reference) */
} wql_source;
/* init_moment_tensor's location step (moment_tensor.f90:280-372 with
 * Find_Coordinates_moment, seismogram.f90:652-732) on the host, no device
 * call: the nearest node of X[3][ns][nr][nq] to each of xyz[n][3] if the
 * location lies within the block's coordinate ranges and between its first
 * and last nodes; src[s].node and .node_x are filled, the other fields are
 * left as given */
int wql_locate_sources(const double *X, int nq, int nr, int ns, int n, const double *xyz, wql_source *src);
/* the sources of one block (n = 0: none); order 2 or 4; alpha (order 4:
 * 0, 25, 50 or 100) is the first source's, as the reference uses
 * alpha(1) for all of a block's sources */
int wql_set_sources(wql_ctx *ctx, int block, int n, const wql_source *src, int order, int alpha);
/* the Gaussian-mollified form of the same sources (mollify_source = T,
 * set_moment_tensor_smooth, moment_tensor.f90:1400-1475; chosen in set_rates,
 * domain.f90:926-953): every source of the block's list -- all of the
 * unified list in both blocks -- adds -A(t) M dx dy dz at every node, dx a
 * normal density in x about the source with standard deviation twice the
 * node's spacing along q (likewise y along r, z along s).  Only type,
 * duration, t_init, m and x of src are read; X [3][ns][nr][nq] is the block's
 * node coordinates, copied to the device.  Replaces the block's point
 * sources (and wql_set_sources replaces these). */
int wql_set_sources_smooth(wql_ctx *ctx, int block, int n, const wql_source *src, const double *X);

/* station seismograms (seismogram.f90:12-527).  wql_find_stations is
 * Find_Coordinates (seismogram.f90:531-650) on the host, no device call:
 * for each station (x, y, z) of xyz[n][3] the node of X[3][ns][nr][nq]
 * nearest to it among those within half the local grid diagonal, the last
 * such node in (s, r, q) order on ties, as 0-based (i, j, k) in ijk[n][3],
 * or -1 where the station lies off the block.  wql_set_stations records the
 * stations' nodes of a block (n = 0 clears them); wql_get_stations returns
 * v[n][3] = vx vy vz at them from the current fields, the values
 * write_seismogram prints (seismogram.f90:460-478). */
int wql_find_stations(const double *X, int nq, int nr, int ns, int n, const double *xyz, int *ijk);
int wql_set_stations(wql_ctx *ctx, int block, int n, const int *ijk);
int wql_get_stations(wql_ctx *ctx, int block, double *v);
/* peak tangential slip rate per fault node over the run, sampled where the
 * reference samples Svel (stage 1, CouplingForcing.f90:213-228) [nr*ns] */
int wql_get_fault_peak(wql_ctx *ctx, int block, double *psr);
/* fault-plane statistics on the device: {max slip, max slip rate, max peak
 * slip rate, ruptured nodes (trup set), mean slip over them, earliest trup} */
int wql_fault_stats(wql_ctx *ctx, int block, double stats[6]);

/* One right-hand-side evaluation at time t with zero incoming rates:
 * what = 0: volume operator only (set_rates, domain.f90:908-924)
 * what = 1: plus boundary and fault SAT terms (enforce_bound_iface_conditions,
 *           domain.f90:586-611) at RK stage `stage` */
int wql_eval_rates(wql_ctx *ctx, double t, int what, int stage);

/* nsteps LSRK(5,4) steps (time_step.f90:68-150); enqueued on the context's
 * stream, returns without waiting. */
int wql_step(wql_ctx *ctx, int nsteps);
/* one RK stage of the current step (the body of time_step_RK's stage loop,
 * time_step.f90:109-147), for stage-by-stage comparison: stages 1..5 in
 * order (WQL_E_STATE otherwise); after stage 5 the step is complete and
 * wql_time advances by dt.  Between stages wql_get_fields / wql_get_rates
 * return the stage's output and rates.  wql_step refuses to start while a
 * step is part done. */
int wql_stage(wql_ctx *ctx, int stage);
/* wait for the context's stream; reports NaN/Inf found in the fault state */
int wql_sync(wql_ctx *ctx);
double wql_time(const wql_ctx *ctx);

/* launch shape of the volume kernel (autotuner hook): tile variant index and
 * z-chunk length; returns WQL_E_ARG for unknown variants */
int wql_set_tile(wql_ctx *ctx, int variant, int zchunk);
int wql_get_tile(const wql_ctx *ctx, int *variant, int *zchunk);
/* dispatch order of the volume tiles: 0 q-tiles fastest, 1 s-chunks fastest,
 * 2 per-XCD groups of `group` q-tile columns walked plane by plane */
int wql_set_tile_order(wql_ctx *ctx, int order, int group);
int wql_get_tile_order(const wql_ctx *ctx, int *order, int *group);
int wql_tile_variants(void);

/* capture one LSRK step into a HIP graph and replay it (1) or launch kernels
 * directly (0) */
int wql_use_graph(wql_ctx *ctx, int on);

/* average device time (ms) of the main volume kernel (the tiled interior
 * kernel, or the per-node kernel of the upwind / DRP operators and the
 * anelastic response) over the last wql_step call when profiling is on (HIP
 * events on the context's stream) */
int wql_profile(wql_ctx *ctx, int on);
double wql_volume_kernel_ms(wql_ctx *ctx);

/* ---- serving: continuous batching of independent scenarios ------------
 * No counterpart in the reference (one run per process there).  A server
 * admits submitted scenarios first-fit by device footprint
 * (wql_estimate_bytes) within its HBM budget and a cap on active
 * scenarios, keeps each admitted scenario's state on the device, runs it on
 * its own stream with a captured step graph in chunks of steps_per_launch,
 * and retires it when its steps are done. */
typedef struct wql_server wql_server;
/* hbm_budget 0: 90% of the device memory free at creation */
int wql_server_create(int max_active, size_t hbm_budget, int steps_per_launch, wql_server **out);
int wql_server_destroy(wql_server *srv);
/* copies the descriptors and arrays; *id identifies the scenario */
int wql_server_submit(wql_server *srv, const wql_problem *prob, const wql_block_desc blocks[2], int nsteps,
                      int *id);
/* non-blocking progress: retires finished scenarios, launches the next
 * chunk of every active scenario whose chunk ended, admits queued ones, and
 * reports retired (finished or failed) ids in order into done[], at most
 * max_done per call (the return value is their number; the rest are
 * reported by later calls) */
int wql_server_poll(wql_server *srv, int *done, int max_done);
/* block 1's fault copy of a retired scenario: S[4][nr*ns], trup, peak slip
 * rate [nr*ns], wql_fault_stats; frees the record (a failed scenario returns
 * its error here) */
int wql_server_result(wql_server *srv, int id, double *S, double *trup, double *psr, double stats[6]);
/* {queued, active, device bytes in use (active and pooled contexts), budget} */
int wql_server_status(const wql_server *srv, long out[4]);
/* the context pool: {contexts created, contexts reused, contexts freed,
 * contexts pooled now}.  A retired scenario's context is kept and cleared
 * for the next scenario with the same blocks (shapes, descriptors,
 * coordinates and material), precision and plasticity model; the least
 * recently used pooled contexts are freed when the budget needs the room. */
int wql_server_pool_stats(const wql_server *srv, long out[4]);
/* admission order of queued scenarios: 0 arrival order (the default),
 * 1 shortest job first with aging (key: nodes x steps / (1 + seconds
 * waited)).  No counterpart in the reference. */
int wql_server_set_policy(wql_server *srv, int policy);
/* launch shape of the main volume kernel (as wql_set_tile / _order) for
 * scenarios whose blocks are n[0] x n[1] x n[2] in `precision` on a
 * Cartesian (cartesian = 1) or curvilinear grid, applied when such a
 * scenario is admitted; the autotuner's per-shape choices are registered
 * here.  No counterpart in the reference. */
int wql_server_set_shape(wql_server *srv, const int n[3], int precision, int cartesian, int variant, int zchunk,
                         int order, int group);

#ifdef __cplusplus
}
#endif
#endif
