// wql_api.hip -- C ABI of libwql (declared in include/wql.h): context
// lifetime, uploads/downloads, and the per-stage launch sequence.
#include "wql.h"
#include "wql_kernels.hip"
#include "wql_grid.hip"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace wql;

// Staging memory of the allocation-free volume download
// (wql_get_fields_noalloc): a file-scope device array of 256 MiB, of which
// every GPU the library runs on holds its own copy until the process ends.
// In host code `scratch` is only a symbol: its address on a device comes
// from hipGetSymbolAddress, taken once per device (wql_create, with the
// device current) and kept in g_scratch.
constexpr size_t SCRATCH_DOUBLES = (size_t)32 << 20;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch[MAX_DEVICES] = {};

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(WQL_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// LSRK(5,4) of Carpenter & Kennedy (time_step.f90:53-62)
const double RK_A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                        -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double RK_B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                        1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                        2277821191437.0 / 14882151754819.0};
const double RK_C[6] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                        2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0, 1.0};

const double TAU0 = 1.0 / (0.315949074074074);  // block.f90:191

// closure table rows (JU_xJU_yJU_z6.f90:77-298) as (1-based column, coefficient)
struct RowSrc { int n; int col[8]; double c[8]; };
const RowSrc LEFT_SRC[6] = {
    {6, {6, 5, 4, 3, 2, 1}, {0.036577936277544, 0.104488069284042, -0.450398306578272, -0.141512858744873, 2.033378678700676, -1.582533518939116}},
    {5, {6, 5, 4, 3, 1}, {-0.014903449191300, -0.069112065532624, 0.258816087376832, 0.287258622978251, -0.462059195631158}},
    {5, {6, 5, 4, 2, 1}, {-0.018129342917256, -0.022902190275815, 0.606235523609147, -0.636451095137907, 0.071247104721830}},
    {6, {7, 6, 5, 3, 2, 1}, {0.013435342414630, -0.051642265516119, 0.520262285050482, -0.306681191361148, -0.290087484386815, 0.114713313798970}},
    {7, {8, 7, 6, 4, 3, 2, 1}, {0.018281071473911, -0.164529643265203, 0.769199413962647, -0.707905442575989, 0.015764336127392, 0.105400944933782, -0.036210680656541}},
    {8, {9, 8, 7, 5, 4, 3, 2, 1}, {0.016437980868017, -0.147941827812150, 0.739709139060752, -0.691649024426814, 0.063183694641876, 0.011220896474665, 0.020437334208704, -0.011398193015050}},
};
// right rows n-5..n: columns as offsets from n
const RowSrc RIGHT_SRC[6] = {
    {8, {-8, -7, -6, -4, -3, -2, -1, 0}, {-0.016437980868017, 0.147941827812150, -0.739709139060752, 0.691649024426814, -0.063183694641876, -0.011220896474665, -0.020437334208704, 0.011398193015050}},
    {7, {-7, -6, -5, -3, -2, -1, 0}, {-0.018281071473911, 0.164529643265203, -0.769199413962647, 0.707905442575989, -0.015764336127392, -0.105400944933782, 0.036210680656541}},
    {6, {-6, -5, -4, -2, -1, 0}, {-0.013435342414630, 0.051642265516119, -0.520262285050482, 0.306681191361148, 0.290087484386815, -0.114713313798970}},
    {5, {-5, -4, -3, -1, 0}, {0.018129342917256, 0.022902190275815, -0.606235523609147, 0.636451095137907, -0.071247104721830}},
    {5, {-5, -4, -3, -2, 0}, {0.014903449191300, 0.069112065532624, -0.258816087376832, -0.287258622978251, 0.462059195631158}},
    {6, {-5, -4, -3, -2, -1, 0}, {-0.036577936277544, -0.104488069284042, 0.450398306578272, 0.141512858744873, -2.033378678700676, 1.582533518939116}},
};

SbpTable make_table()
{
    SbpTable T;
    memset(&T, 0, sizeof T);
    for (int r = 0; r < 6; r++) {
        T.left[r].n = LEFT_SRC[r].n;
        for (int t = 0; t < LEFT_SRC[r].n; t++) {
            T.left[r].off[t] = (LEFT_SRC[r].col[t] - 1) - r;
            T.left[r].c[t] = LEFT_SRC[r].c[t];
        }
        // row index n-6+r (0-based), column n-1+off  ->  relative off + 5 - r
        T.right[r].n = RIGHT_SRC[r].n;
        for (int t = 0; t < RIGHT_SRC[r].n; t++) {
            T.right[r].off[t] = RIGHT_SRC[r].col[t] + 5 - r;
            T.right[r].c[t] = RIGHT_SRC[r].c[t];
        }
    }
    return T;
}

struct DevBlock {
    int nq, nr, ns;
    int px;                 // row pitch of the volume arrays
    long NP;                // px*nr*ns (elements per field on the device)
    long ND;                // nq*nr*ns (dense, as the host sees a field)
    void *satb = nullptr;   // [9][nbnd] real: boundary penalties of the stage
    long nbnd = 0;          // boundary nodes
    void *fq = nullptr;     // [2][9][nr ns] real: the nine fields on the q faces (q = 0, q = nq-1)
    int cart;
    int bc[6], pml[6], npml;
    double lam, mu, rho;
    double gc[10];
    unsigned gconst = 0;    // bit c: metric field c of G is constant, value gc[c]
    void *F[2] = {nullptr, nullptr};
    void *DF = nullptr;
    void *G = nullptr;      // [10][NP] real (curvilinear)
    void *M = nullptr;      // [3][NP] real (heterogeneous)
    void *Q[6] = {}, *DQ[6] = {};
    void *pq[6] = {};       // [9][face nodes] real: PML face penalties of the stage
    void *eta = nullptr, *deta = nullptr;  // [24][NP] real: anelastic memory variables and rates
    void *qinv = nullptr;   // [2][NP] real: Q_s^-1, Q_p^-1
    long SP[6] = {};
    double *frame[6] = {}, *fgm[6] = {}, *fmat[6] = {};
    double *fx = nullptr;
    double *P = nullptr, *Y = nullptr;
    // stations (wql_set_stations): node offsets [nst] and velocity staging [nst][3]
    int nst = 0, st_cap = 0;
    long *st_off = nullptr;
    double *st_v = nullptr;
    // moment-tensor sources (wql_set_sources) and the node box of their supports
    int nsrc = 0, src_cap = 0, src_order = 2, src_alpha = 0;
    int src_lo[3] = {0, 0, 0}, src_hi[3] = {0, 0, 0};
    SourceDev *src = nullptr;
    // the mollified form (wql_set_sources_smooth): node coordinates [3][ND]
    int src_smooth = 0;
    double *Xs = nullptr;
};

}  // namespace

struct wql_ctx {
    int prec;
    hipStream_t stream = nullptr;
    DevBlock b[2];
    Iface I[2];
    FaultParams fault;
    wql_problem prob;
    double t, dt;
    int cur = 0;
    // launch shape of the main volume kernel (set per precision in
    // wql_create: the fastest measured on the 2 x 201x101x101 curvilinear
    // problem)
    int variant = 0, zchunk = 6, order = 0, group = 2;
    int use_graph = 0;
    int stage_next = 1;     // next stage of a step run stage by stage (wql_stage)
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int *d_flag = nullptr;
    double *d_t = nullptr;
    // persistent device staging of the small per-output downloads: both
    // fault faces [9][nr ns] each, and the six fault statistics
    double *d_faces = nullptr, *d_stats = nullptr;
    // dense FP64 staging of the volume uploads/downloads, allocated on first
    // use and grown to the largest transfer (freed by wql_destroy)
    double *d_stage = nullptr;
    size_t stage_elems = 0;
    int dev = 0;            // device the context lives on
    size_t bytes = 0;
    std::vector<void *> allocs;
    int profile = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // side stream: the s-closure, PML-slab and per-node operator launches of
    // a stage run there beside the main volume launch (fork and join events;
    // 5-17% shorter steps on the operator path, 8% with the anelastic
    // response, 1-3% with PML slabs)
    hipStream_t side = nullptr;
    hipEvent_t evf = nullptr, evj = nullptr;
    int overlap = 0;
    double vol_ms = 0.0;
    long vol_launches = 0;
    // spatial operator: -1 the sixth-order central one, else WQL_SBP_OPS[sbp]
    // (its v and s tables copied to d_sbp); tau0 the SAT penalty of its norm
    int sbp = -1;
    double tau0 = 0.0;
    wql_sbp_half *d_sbp = nullptr;
    // anelastic response: relaxation times and weights of the four mechanisms
    double an_tau[4] = {0, 0, 0, 0}, an_w[4] = {0, 0, 0, 0};
    // tiled operator kernel: halo (the operator's largest reach: 3, 5 or 7)
    // and s-chunk length
    int opt_h = 7, opt_zc = 12;
    // band width of the tabled operator (d_sbp != nullptr)
    int band = 6;
};

namespace {

int dalloc(wql_ctx *c, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(WQL_E_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    c->allocs.push_back(*p);
    c->bytes += bytes;
    return 0;
}

template <typename real>
Blk<real> blk_view(const wql_ctx *c, int ib, int cur)
{
    const DevBlock &d = c->b[ib];
    Blk<real> B;
    memset(&B, 0, sizeof B);
    B.nq = d.nq; B.nr = d.nr; B.ns = d.ns; B.px = d.px; B.NP = d.NP;
    B.Fin = (const real *)d.F[cur];
    B.Fout = (real *)d.F[1 - cur];
    B.DF = (real *)d.DF;
    B.G = (const real *)d.G;
    B.M = (const real *)d.M;
    for (int k = 0; k < 10; k++) B.gc[k] = d.gc[k];
    B.gconst = d.gconst;
    B.lam = d.lam; B.mu = d.mu; B.rho = d.rho;
    B.cart = d.cart;
    for (int s = 0; s < 6; s++) {
        B.bc[s] = d.bc[s]; B.pml[s] = d.pml[s];
        B.Q[s] = (real *)d.Q[s]; B.DQ[s] = (real *)d.DQ[s];
        B.frame[s] = d.frame[s]; B.fgm[s] = d.fgm[s]; B.fmat[s] = d.fmat[s];
    }
    B.npml = d.npml;
    B.fx = d.fx;
    B.P = d.P;
    B.Y = d.Y;
    B.satb = nullptr;
    B.nbnd = d.nbnd;
    B.sbp = c->d_sbp;
    B.sbw = c->d_sbp ? c->band : 0;
    B.eta = (real *)d.eta;
    B.deta = (real *)d.deta;
    B.qinv = (const real *)d.qinv;
    for (int l = 0; l < 4; l++) { B.an_tau[l] = c->an_tau[l]; B.an_w[l] = c->an_w[l]; }
    return B;
}

// node ranges of the main kernel (all nodes not in a PML slab and not within
// six nodes of an s boundary)
void main_range(const DevBlock &d, int r[6])
{
    int pz0 = d.pml[4] ? std::max(d.npml, 6) : 6, pz1 = d.pml[5] ? std::max(d.npml, 6) : 6;
    int px0 = d.pml[0] ? std::max(d.npml, 6) : 0, px1 = d.pml[1] ? std::max(d.npml, 6) : 0;
    int py0 = d.pml[2] ? std::max(d.npml, 6) : 0, py1 = d.pml[3] ? std::max(d.npml, 6) : 0;
    r[0] = std::min(px0, d.nq); r[1] = std::max(d.nq - px1, r[0]);
    r[2] = std::min(py0, d.nr); r[3] = std::max(d.nr - py1, r[2]);
    r[4] = std::min(pz0, d.ns); r[5] = std::max(d.ns - pz1, r[4]);
}

// Launch variants of the interior volume kernel: the fused single-pass march
// re-reading the centre plane's stresses and J (FP32 default) or holding the
// lead plane in registers until it is the centre (FP64 default), on 32x8,
// 64x8 and 32x16 tiles, and the role-split march (two workgroups per tile,
// one per rate group) on 32x8 tiles.
enum { FAM_FUSED = 0, FAM_HOLD = 1, FAM_SPLIT = 2 };
const int NVARIANT = 7;
const int VARIANT_FAM[NVARIANT] = {FAM_FUSED, FAM_HOLD, FAM_FUSED, FAM_HOLD, FAM_FUSED, FAM_HOLD, FAM_SPLIT};
const int VARIANT_TX[NVARIANT] = {32, 32, 64, 64, 32, 32, 32};
const int VARIANT_TY[NVARIANT] = {8, 8, 8, 8, 16, 16, 8};

// the volume kernels address every array of a block with 32-bit buffer
// offsets: the largest (metrics + J, 10 fields) must span < 2 GiB
bool fits_buffers(long NP, size_t elem) { return (double)NP * 10.0 * (double)elem < 2147483647.0; }

TileList tile_list(const wql_ctx *c)
{
    TileList L;
    memset(&L, 0, sizeof L);
    int TX = VARIANT_TX[c->variant], TY = VARIANT_TY[c->variant];
    long first = 0;
    for (int ib = 0; ib < 2; ib++) {
        int r[6];
        main_range(c->b[ib], r);
        if (r[1] <= r[0] || r[3] <= r[2] || r[5] <= r[4]) continue;
        TileGeom &g = L.g[L.n];
        g.blk = ib;
        g.xa0 = r[0]; g.xa1 = r[1]; g.ya0 = r[2]; g.ya1 = r[3]; g.za0 = r[4]; g.za1 = r[5];
        g.ntx = (r[1] - r[0] + TX - 1) / TX;
        g.nty = (r[3] - r[2] + TY - 1) / TY;
        g.zc = std::max(1, c->zchunk);
        g.nzc = (r[5] - r[4] + g.zc - 1) / g.zc;
        g.first = first;
        first += (long)g.ntx * g.nty * g.nzc;
        L.n++;
    }
    L.total = first;
    L.order = c->order;
    L.gw = std::max(1, c->group);
    return L;
}

// s-band planes: the six closure planes at each s face that has no PML, as
// one 6-plane chunk per tile (per_plane = 0), one plane per tile
// (per_plane = 1) or chunks of per_plane planes (per_plane > 1)
TileList band_list(const wql_ctx *c, int per_plane)
{
    TileList L;
    memset(&L, 0, sizeof L);
    const int TX = 32, TY = 8;
    long first = 0;
    for (int ib = 0; ib < 2; ib++) {
        const DevBlock &d = c->b[ib];
        int r[6];
        main_range(d, r);
        if (r[1] <= r[0] || r[3] <= r[2]) continue;
        // without PML main_range leaves exactly planes [0, 6) and [ns-6, ns)
        int seg[2][2] = {{0, d.pml[4] ? 0 : r[4]}, {d.pml[5] ? d.ns : r[5], d.ns}};
        for (int k = 0; k < 2; k++) {
            if (seg[k][1] <= seg[k][0]) continue;
            TileGeom &g = L.g[L.n++];
            g.blk = ib;
            g.xa0 = r[0]; g.xa1 = r[1]; g.ya0 = r[2]; g.ya1 = r[3];
            g.za0 = seg[k][0]; g.za1 = seg[k][1];
            g.ntx = (r[1] - r[0] + TX - 1) / TX;
            g.nty = (r[3] - r[2] + TY - 1) / TY;
            const int len = g.za1 - g.za0;
            g.zc = per_plane == 0 ? len : per_plane;
            g.nzc = (len + g.zc - 1) / g.zc;
            g.first = first;
            first += (long)g.ntx * g.nty * g.nzc;
        }
    }
    L.total = first;
    L.order = 0;
    L.gw = 1;
    return L;
}

// nodes of the per-node generic kernel: the PML slabs (including their
// boundary band rows), and with the anelastic response also the s-closure
// planes of the faces without PML (which k_volume_split does not cover then)
BoxList box_list(const wql_ctx *c)
{
    BoxList L;
    memset(&L, 0, sizeof L);
    long start = 0;
    auto add = [&](int ib, int x0, int x1, int y0, int y1, int z0, int z1) {
        if (x1 <= x0 || y1 <= y0 || z1 <= z0) return;
        Box &b = L.box[L.nbox++];
        b.blk = ib; b.x0 = x0; b.x1 = x1; b.y0 = y0; b.y1 = y1; b.z0 = z0; b.z1 = z1;
        b.start = start;
        start += (long)(x1 - x0) * (y1 - y0) * (z1 - z0);
    };
    for (int ib = 0; ib < 2; ib++) {
        const DevBlock &d = c->b[ib];
        int r[6];
        main_range(d, r);
        int z0 = d.pml[4] ? r[4] : 0, z1 = d.pml[5] ? r[5] : d.ns;
        add(ib, 0, d.nq, 0, d.nr, 0, z0);          // s-low PML slab
        add(ib, 0, d.nq, 0, d.nr, z1, d.ns);       // s-high PML slab
        add(ib, 0, r[0], 0, d.nr, z0, z1);         // q-low slab
        add(ib, r[1], d.nq, 0, d.nr, z0, z1);      // q-high slab
        add(ib, r[0], r[1], 0, r[2], z0, z1);      // r-low slab
        add(ib, r[0], r[1], r[3], d.nr, z0, z1);   // r-high slab
        if (c->prob.anelastic && r[1] > r[0] && r[3] > r[2]) {
            if (!d.pml[4]) add(ib, r[0], r[1], r[2], r[3], 0, r[4]);       // s-low closure planes
            if (!d.pml[5]) add(ib, r[0], r[1], r[2], r[3], r[5], d.ns);    // s-high closure planes
        }
    }
    L.total = start;
    return L;
}

// the PML slabs of one stage (after the interior kernels): per node, one
// thread each
template <typename real>
void launch_pml(wql_ctx *c, Pair<real> &P, real A, real bdt, hipStream_t st)
{
    BoxList BL = box_list(c);
    if (BL.total > 0)
        k_volume_generic<real><<<(unsigned)((BL.total + 255) / 256), 256, 0, st>>>(P, BL, A, A, bdt);
}

// problems on the operator path: the upwind / upwind-DRP operators (the
// sixth-order kernels do not take their tables)
bool per_node(const wql_ctx *c) { return c->sbp >= 0; }

// node ranges of the tiled operator kernel k_volume_opt in one block: q and
// r outside the PML slabs, s the planes whose s row is the interior row and
// at least HS planes from each s face (the s queue of the march); r[6] =
// x0 x1 y0 y1 z0 z1, empty when the block has none
void opt_range(const wql_ctx *c, const DevBlock &d, int r[6])
{
    const int w = c->band, hs = std::min(c->opt_h, 5);
    r[0] = d.pml[0] ? d.npml : 0; r[1] = d.nq - (d.pml[1] ? d.npml : 0);
    r[2] = d.pml[2] ? d.npml : 0; r[3] = d.nr - (d.pml[3] ? d.npml : 0);
    r[4] = std::max(hs, d.pml[4] ? std::max(d.npml, w) : w);
    r[5] = d.ns - std::max(hs, d.pml[5] ? std::max(d.npml, w) : w);
    if (r[1] <= r[0] || r[3] <= r[2] || r[5] <= r[4]) r[1] = r[0] = r[3] = r[2] = r[5] = r[4] = 0;
}

// the tiled operator kernel's tiles: 32 x 8 columns, s chunks of opt_zc planes
TileList opt_tiles(const wql_ctx *c)
{
    TileList L;
    memset(&L, 0, sizeof L);
    long first = 0;
    for (int ib = 0; ib < 2; ib++) {
        int r[6];
        opt_range(c, c->b[ib], r);
        if (r[1] <= r[0]) continue;
        TileGeom &g = L.g[L.n++];
        g.blk = ib;
        g.xa0 = r[0]; g.xa1 = r[1]; g.ya0 = r[2]; g.ya1 = r[3]; g.za0 = r[4]; g.za1 = r[5];
        g.ntx = (r[1] - r[0] + 31) / 32;
        g.nty = (r[3] - r[2] + 7) / 8;
        g.zc = std::max(1, c->opt_zc);
        g.nzc = (r[5] - r[4] + g.zc - 1) / g.zc;
        g.first = first;
        first += (long)g.ntx * g.nty * g.nzc;
    }
    L.total = first;
    return L;
}

// the operator path: k_volume_opt on the tiles of opt_tiles, k_volume_op
// (one thread per node) on every other node (s-band planes, PML slabs)
template <typename real>
void launch_op(wql_ctx *c, Pair<real> &P, real A, real bdt, int part, hipStream_t st)
{
    BoxList L;
    memset(&L, 0, sizeof L);
    long start = 0;
    auto add = [&](int ib, int x0, int x1, int y0, int y1, int z0, int z1) {
        if (x1 <= x0 || y1 <= y0 || z1 <= z0) return;
        Box &b = L.box[L.nbox++];
        b.blk = ib; b.x0 = x0; b.x1 = x1; b.y0 = y0; b.y1 = y1; b.z0 = z0; b.z1 = z1;
        b.start = start;
        start += (long)(x1 - x0) * (y1 - y0) * (z1 - z0);
    };
    TileList TL;
    memset(&TL, 0, sizeof TL);
    if (c->d_sbp) TL = opt_tiles(c);
    for (int ib = 0; ib < 2; ib++) {
        const DevBlock &d = c->b[ib];
        int r[6] = {0, 0, 0, 0, 0, 0};
        if (c->d_sbp) opt_range(c, d, r);
        if (r[1] <= r[0]) {
            add(ib, 0, d.nq, 0, d.nr, 0, d.ns);
            continue;
        }
        add(ib, 0, d.nq, 0, d.nr, 0, r[4]);           // s-low planes
        add(ib, 0, d.nq, 0, d.nr, r[5], d.ns);        // s-high planes
        add(ib, 0, r[0], 0, d.nr, r[4], r[5]);        // q-low PML slab
        add(ib, r[1], d.nq, 0, d.nr, r[4], r[5]);     // q-high PML slab
        add(ib, r[0], r[1], 0, r[2], r[4], r[5]);     // r-low PML slab
        add(ib, r[0], r[1], r[3], d.nr, r[4], r[5]);  // r-high PML slab
    }
    L.total = start;
    if (TL.total > 0 && (part & 1)) {
        // a multiple of 8 workgroups: the kernel's XCD remap of block ids is a
        // bijection onto [0, grid) only then
        const dim3 blk(32, 8), grd((unsigned)(((TL.total + 7) / 8) * 8));
        if (c->opt_h == 3) k_volume_opt<real, 32, 8, 3><<<grd, blk, 0, st>>>(P, TL, A, A, bdt);
        else if (c->opt_h == 5) k_volume_opt<real, 32, 8, 5><<<grd, blk, 0, st>>>(P, TL, A, A, bdt);
        else k_volume_opt<real, 32, 8, 7><<<grd, blk, 0, st>>>(P, TL, A, A, bdt);
    }
    if (L.total > 0 && (part & 2))
        k_volume_op<real><<<(unsigned)((L.total + 255) / 256), 256, 0, st>>>(P, L, A, A, bdt);
}

// the moment-tensor sources of both blocks at the stage time (after the
// volume kernels, before the plastic return map)
template <typename real>
void launch_sources(wql_ctx *c, Pair<real> &P, double t, const double *tdev, double tinc, double bdt)
{
    for (int ib = 0; ib < 2; ib++) {
        const DevBlock &d = c->b[ib];
        if (d.nsrc == 0) continue;
        SourceArgs a;
        memset(&a, 0, sizeof a);
        a.src = d.src;
        a.n = d.nsrc;
        a.order = d.src_order;
        a.alpha = d.src_alpha;
        long nbox = 1;
        for (int k = 0; k < 3; k++) {
            a.lo[k] = d.src_lo[k];
            a.hi[k] = d.src_hi[k];
            nbox *= (long)(a.hi[k] - a.lo[k]);
        }
        if (nbox <= 0) continue;
        a.t = t;
        a.tdev = tdev;
        a.tinc = tinc;
        a.bdt = bdt;
        a.smooth = d.src_smooth;
        a.X = d.Xs;
        k_sources<real><<<(unsigned)((nbox + 255) / 256), 256, 0, c->stream>>>(P.b[ib], a);
    }
}

template <typename real, int TX, int TY>
void launch_main_v(const wql_ctx *c, Pair<real> &P, const TileList &TL, real A0, real A1, real bdt)
{
    dim3 blk(TX, TY);
    dim3 grd((unsigned)(((TL.total + 7) / 8) * 8));
    const bool curv = !c->b[0].cart || !c->b[1].cart;
    const bool het = c->b[0].M || c->b[1].M;
    const int fam = VARIANT_FAM[c->variant];
    if constexpr (TX == 32 && TY == 8) {
        // the anelastic response: 32 x 8 tiles only (wql_set_tile maps the
        // other shapes onto them)
        if (c->prob.anelastic) {
            if (fam == FAM_HOLD) {
                if (curv && het) k_volume_fused<real, 32, 8, true, true, true, true><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
                else if (curv) k_volume_fused<real, 32, 8, true, false, true, true><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
                else if (het) k_volume_fused<real, 32, 8, false, true, true, true><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
                else k_volume_fused<real, 32, 8, false, false, true, true><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
            } else {
                if (curv && het) k_volume_fused<real, 32, 8, true, true, false, true><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
                else if (curv) k_volume_fused<real, 32, 8, true, false, false, true><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
                else if (het) k_volume_fused<real, 32, 8, false, true, false, true><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
                else k_volume_fused<real, 32, 8, false, false, false, true><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
            }
            return;
        }
    }
    if (fam == FAM_HOLD) {
        if (curv && het) k_volume_fused<real, TX, TY, true, true, true, false><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
        else if (curv) k_volume_fused<real, TX, TY, true, false, true, false><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
        else if (het) k_volume_fused<real, TX, TY, false, true, true, false><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
        else k_volume_fused<real, TX, TY, false, false, true, false><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
    } else if (fam == FAM_FUSED) {
        if (curv && het) k_volume_fused<real, TX, TY, true, true, false, false><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
        else if (curv) k_volume_fused<real, TX, TY, true, false, false, false><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
        else if (het) k_volume_fused<real, TX, TY, false, true, false, false><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
        else k_volume_fused<real, TX, TY, false, false, false, false><<<grd, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
    } else {
        dim3 grd2((unsigned)(((2 * TL.total + 7) / 8) * 8));
        if (curv && het) k_volume_split<real, TX, TY, true, true, 0><<<grd2, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
        else if (curv) k_volume_split<real, TX, TY, true, false, 0><<<grd2, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
        else if (het) k_volume_split<real, TX, TY, false, true, 0><<<grd2, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
        else k_volume_split<real, TX, TY, false, false, 0><<<grd2, blk, 0, c->stream>>>(P, TL, A0, A1, bdt);
    }
}

template <typename real>
void launch_main(const wql_ctx *c, Pair<real> &P, const TileList &TL, real A0, real A1, real bdt)
{
    const int tx = VARIANT_TX[c->variant], ty = VARIANT_TY[c->variant];
    if (tx == 64) launch_main_v<real, 64, 8>(c, P, TL, A0, A1, bdt);
    else if (ty == 16) launch_main_v<real, 32, 16>(c, P, TL, A0, A1, bdt);
    else launch_main_v<real, 32, 8>(c, P, TL, A0, A1, bdt);
}

template <typename real>
void launch_band(const wql_ctx *c, Pair<real> &P, real A, real bdt, hipStream_t st)
{
    // the six closure planes at each s face without PML: role-split march in
    // closure mode, one 6-plane chunk per face and tile (32 x 8 tiles)
    if (c->prob.anelastic) return;  // the per-node kernel takes these planes (box_list)
    const bool curv = !c->b[0].cart || !c->b[1].cart;
    const bool het = c->b[0].M || c->b[1].M;
    dim3 blk(32, 8);
    TileList TB = band_list(c, 0);
    if (TB.total <= 0) return;
    dim3 grd((unsigned)(((2 * TB.total + 7) / 8) * 8));
    if (curv && het) k_volume_split<real, 32, 8, true, true, 1><<<grd, blk, 0, st>>>(P, TB, A, A, bdt);
    else if (curv) k_volume_split<real, 32, 8, true, false, 1><<<grd, blk, 0, st>>>(P, TB, A, A, bdt);
    else if (het) k_volume_split<real, 32, 8, false, true, 1><<<grd, blk, 0, st>>>(P, TB, A, A, bdt);
    else k_volume_split<real, 32, 8, false, false, 1><<<grd, blk, 0, st>>>(P, TB, A, A, bdt);
}

// the boundary penalties are folded into the volume kernels: every kernel
// of the stage (the interior march, its s-closure mode and the PML-slab
// kernel) adds a boundary node's penalty when it forms DF
template <typename real>
bool fused_faces(const wql_ctx *c)
{
    for (int ib = 0; ib < 2; ib++)
        if (!c->b[ib].satb) return false;
    return true;
}

// the q-face buffers are in use with the fused penalties and the fused
// volume variants that re-read the centre plane's metrics; every
// kernel that writes the stage output of a q-face node then also writes the
// buffer.  The metric-holding variants (the FP64 default) run at the
// register limit: there the buffer's offset and the predicated stores cost
// the volume kernel more (14 us a stage on the 2x201x101x101 rupture) than
// the coalesced reads save the boundary kernel.
template <typename real>
bool fq_on(const wql_ctx *c)
{
    return VARIANT_FAM[c->variant] == FAM_FUSED && fused_faces<real>(c);
}

// refresh the q-face buffers from the current fields (after host updates)
template <typename real>
void gather_fq(wql_ctx *c)
{
    if (!fq_on<real>(c)) return;
    for (int ib = 0; ib < 2; ib++) {
        const DevBlock &d = c->b[ib];
        const long n = 18L * d.nr * d.ns;
        k_fq_gather<real><<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(d.nq, d.nr, d.ns, d.px, d.NP,
                                                                              (const real *)d.F[c->cur], (real *)d.fq);
    }
}

template <typename real>
void launch_stage(wql_ctx *c, int cur, double A, double bdt, double t_stage, int stage, int rates_only,
                  int do_plastic, const double *tdev = nullptr, double tinc = 0.0)
{
    Pair<real> P;
    P.b[0] = blk_view<real>(c, 0, cur);
    P.b[1] = blk_view<real>(c, 1, cur);
    BoundaryArgs a;
    memset(&a, 0, sizeof a);
    a.fault = c->fault;
    a.I[0] = c->I[0];
    a.I[1] = c->I[1];
    for (int ib = 0; ib < 2; ib++) a.nodes[ib] = c->b[ib].nbnd;
    a.t = t_stage;
    a.tdev = tdev;
    a.tinc = tinc;
    a.A[0] = a.A[1] = A;
    a.bdt = bdt;
    a.tau0 = c->tau0;
    a.stage = stage;
    a.rates_only = rates_only;
    a.flag = c->d_flag;
    unsigned gb = (unsigned)((a.nodes[0] + a.nodes[1] + 127) / 128);
    // fused boundary penalties: the role-split kernels (and the s-closure and
    // PML-slab kernels launched with them) add each boundary node's penalty
    // when they form DF, so the boundary kernel runs first and only writes
    // the compact penalty array (plus the fault); PML penalties follow the
    // volume kernels as before
    const bool fused = fused_faces<real>(c);
    bool any_pml = false;
    for (int ib = 0; ib < 2; ib++)
        for (int sd = 0; sd < 6; sd++) any_pml = any_pml || c->b[ib].pml[sd];
    // q-face buffers (fused variants): the boundary kernel reads the stage
    // input of the q-face nodes from them, coalesced, and the volume kernels
    // write the stage output of those nodes into them
    const bool fq = fq_on<real>(c);
    if (fq)
        for (int ib = 0; ib < 2; ib++) P.b[ib].FQ = (real *)c->b[ib].fq;
    // PML face penalties: with the fused penalties the boundary kernel writes
    // them to compact per-face arrays and the PML-slab kernel adds them to
    // DQ, so no second boundary launch follows the volume kernels
    if (fused && any_pml)
        for (int ib = 0; ib < 2; ib++)
            for (int sd = 0; sd < 6; sd++) P.b[ib].PQ[sd] = (real *)c->b[ib].pq[sd];
    if (fused) {
        a.mode = 1;
        a.satb[0] = c->b[0].satb;
        a.satb[1] = c->b[1].satb;
        a.fq_read = fq ? 1 : 0;
        k_boundary<real><<<gb, 128, 0, c->stream>>>(P, a);
        a.fq_read = 0;
        for (int ib = 0; ib < 2; ib++) P.b[ib].satb = (const real *)c->b[ib].satb;
    }
    // mixed grids: a Cartesian block inside a curvilinear pair still has no G
    TileList TL = tile_list(c);
    // the launches beside the main one: s-closure planes and PML slabs, or on
    // the operator path the per-node kernel; with overlap on they run on the
    // side stream, forked after the boundary kernel and joined before the
    // sources.  Profiling (per-launch kernel times) keeps one stream.
    const bool ov = c->overlap != 0 && c->side && !c->profile;
    auto rest = [&](hipStream_t st) {
        if (per_node(c)) {
            launch_op<real>(c, P, (real)A, (real)bdt, 2, st);
        } else {
            launch_band<real>(c, P, (real)A, (real)bdt, st);
            launch_pml<real>(c, P, (real)A, (real)bdt, st);
        }
    };
    if (ov) {
        hipEventRecord(c->evf, c->stream);
        hipStreamWaitEvent(c->side, c->evf, 0);
    }
    if (per_node(c) || TL.total > 0) {
        if (c->profile && !c->use_graph) hipEventRecord(c->ev0, c->stream);
        if (per_node(c)) launch_op<real>(c, P, (real)A, (real)bdt, ov ? 1 : 3, c->stream);
        else launch_main<real>(c, P, TL, (real)A, (real)A, (real)bdt);
        if (c->profile && !c->use_graph) {
            hipEventRecord(c->ev1, c->stream);
            hipEventSynchronize(c->ev1);
            float ms = 0.f;
            hipEventElapsedTime(&ms, c->ev0, c->ev1);
            c->vol_ms += ms;
            c->vol_launches++;
        }
    }
    if (ov) {
        rest(c->side);
        hipEventRecord(c->evj, c->side);
        hipStreamWaitEvent(c->stream, c->evj, 0);
    } else if (!per_node(c)) {
        rest(c->stream);
    }
    launch_sources<real>(c, P, t_stage, tdev, tinc, bdt);
    if (!fused) {
        a.mode = 0;
        k_boundary<real><<<gb, 128, 0, c->stream>>>(P, a);
    }
    if (do_plastic && c->prob.plastic) {
        PlasticArgs pa;
        memset(&pa, 0, sizeof pa);
        pa.model = c->prob.plastic;
        pa.c = c->prob.pl_c; pa.nu = c->prob.pl_nu; pa.Tv = c->prob.pl_Tv;
        pa.mu = c->prob.pl_mu; pa.beta = c->prob.pl_beta; pa.eta = c->prob.pl_eta;
        pa.dt = c->dt;
        pa.s0_kind = c->prob.s0_kind;
        for (int k = 0; k < 6; k++) pa.s0[k] = c->prob.fault.s0[k];
        long tot = 0;
        for (int ib = 0; ib < 2; ib++) {
            const DevBlock &d = c->b[ib];
            int np = d.npml;
            pa.x0[ib] = d.pml[0] ? np : 0; pa.x1[ib] = d.pml[1] ? d.nq - np : d.nq;
            pa.y0[ib] = d.pml[2] ? np : 0; pa.y1[ib] = d.pml[3] ? d.nr - np : d.nr;
            pa.z0[ib] = d.pml[4] ? np : 0; pa.z1[ib] = d.pml[5] ? d.ns - np : d.ns;
            pa.n[ib] = (long)std::max(0, pa.x1[ib] - pa.x0[ib]) * std::max(0, pa.y1[ib] - pa.y0[ib]) *
                       std::max(0, pa.z1[ib] - pa.z0[ib]);
            tot += pa.n[ib];
        }
        // plastic acts on the stage output
        Pair<real> Q = P;
        if (tot > 0) k_plastic<real><<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(Q, pa);
    }
}

template <typename real>
int enqueue_step(wql_ctx *c, int cur, double t0)
{
    for (int s = 0; s < 5; s++) {
        launch_stage<real>(c, cur, RK_A[s], RK_B[s] * c->dt, t0 + RK_C[s] * c->dt, s + 1, 0, s == 4);
        cur = 1 - cur;
    }
    return 0;
}

// the step as a graph body: stage times read from c->d_t, which the last node
// advances by one step
template <typename real>
int enqueue_step_graph(wql_ctx *c, int cur)
{
    for (int s = 0; s < 5; s++) {
        launch_stage<real>(c, cur, RK_A[s], RK_B[s] * c->dt, 0.0, s + 1, 0, s == 4, c->d_t, RK_C[s] * c->dt);
        cur = 1 - cur;
    }
    k_advance_time<<<1, 64, 0, c->stream>>>(c->d_t, RK_C[5] * c->dt);
    return 0;
}

void drop_graphs(wql_ctx *c)
{
    for (int k = 0; k < 2; k++)
        if (c->gexec[k]) {
            hipGraphExecDestroy(c->gexec[k]);
            c->gexec[k] = nullptr;
        }
}

int check_launch();

// Put a context back to the state wql_create leaves, for a new problem on
// the same blocks (geometry, material, precision and plasticity model):
// fields, rates, PML auxiliary fields, plastic state and interface state
// are cleared on the context's stream (no allocation, no host wait), the
// fault description, dt and t0 are taken from prob, and captured graphs
// (which hold the old fault description) are dropped.  Metrics, face
// frames and material are kept.
template <typename real>
void reset_typed(wql_ctx *c)
{
    auto fill = [&](void *p, long n, real v) {
        if (p && n > 0) k_fill<real><<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(n, (real *)p, v);
    };
    auto fill64 = [&](double *p, long n, double v) {
        if (p && n > 0) k_fill<double><<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(n, p, v);
    };
    for (int ib = 0; ib < 2; ib++) {
        DevBlock &d = c->b[ib];
        fill(d.F[0], 9 * d.NP, (real)0);
        fill(d.F[1], 9 * d.NP, (real)0);
        fill(d.DF, 9 * d.NP, (real)0);
        fill(d.eta, 24 * d.NP, (real)0);
        fill(d.deta, 24 * d.NP, (real)0);
        for (int s = 0; s < 6; s++) {
            fill(d.Q[s], 9 * d.SP[s], (real)0);
            fill(d.DQ[s], 9 * d.SP[s], (real)0);
        }
        fill64(d.P, 2 * d.NP, 0.0);
        const long FP = (long)d.nr * d.ns;
        Iface &I = c->I[ib];
        fill64(I.S, 4 * FP, 0.0);
        fill64(I.DS, 4 * FP, 0.0);
        fill64(I.W, FP, 0.0);
        fill64(I.DW, FP, 0.0);
        fill64(I.trup, FP, 1.0e9);  // iface.f90:41
        fill64(I.Svel, 3 * FP, 0.0);
        fill64(I.Uhat, 6 * FP, 0.0);
        fill64(I.Vhat, 6 * FP, 0.0);
        fill64(I.psr, FP, 0.0);
    }
    k_fill<int><<<1, 64, 0, c->stream>>>(1, c->d_flag, 0);
}

int valid_problem(const wql_problem *prob);
int sbp_index(const wql_problem *prob);
void fault_tables(wql_ctx *c);

int ctx_reset(wql_ctx *c, const wql_problem *prob)
{
    if (int rc = valid_problem(prob)) return rc;
    if (prob->precision != c->prec || prob->plastic != c->prob.plastic || prob->s0_kind != c->prob.s0_kind)
        return fail(WQL_E_ARG, "a context is reused only for the same precision and plasticity model");
    if (sbp_index(prob) != c->sbp) return fail(WQL_E_ARG, "a context is reused only for the same spatial operator");
    if (prob->fault.key != c->prob.fault.key)
        return fail(WQL_E_ARG, "a context is reused only for the same keyed fault problem");
    if (prob->anelastic != c->prob.anelastic ||
        (prob->anelastic && (prob->an_c != c->prob.an_c || prob->an_weight_exp != c->prob.an_weight_exp ||
                             prob->an_fref != c->prob.an_fref)))
        return fail(WQL_E_ARG, "a context is reused only for the same anelastic response (its moduli are set at create)");
    drop_graphs(c);
    c->prob = *prob;
    memcpy(&c->fault, &prob->fault, sizeof(FaultParams));
    c->dt = prob->dt;
    c->t = prob->t0;
    c->cur = 0;
    c->stage_next = 1;
    c->vol_ms = 0.0;
    c->vol_launches = 0;
    if (c->prec == WQL_FP64) reset_typed<double>(c);
    else reset_typed<float>(c);
    fault_tables(c);
    return check_launch();
}

int check_launch()
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(WQL_E_HIP, "kernel launch: %s", hipGetErrorString(e));
    return 0;
}

// Metric fields that are constant over a curvilinear block.  Grids that are
// affine along some axes (a planar fault, a flat bottom, topography only on
// the free surface) have metric fields that are constant, or zero, up to the
// round-off of the FP64 SBP derivatives (about 1e-14 of the metric scale).
// A field whose values are all equal is marked constant in either precision.
// In FP32 a field whose spread is within 1e-10 of the metric scale (the
// largest |q_x| .. |s_z|; |J| for J) is set to its midpoint: a change far
// below FP32 rounding (6e-8), written into G so that every kernel sees the
// same values.  The fused volume kernel skips the loads of constant fields.
template <typename real>
int constant_metrics(wql_ctx *c, DevBlock &d, const double *G64)
{
    const long ND = d.ND;
    double *rng = nullptr;
    if (hipMalloc(&rng, 20 * sizeof(double)) != hipSuccess) return fail(WQL_E_NOMEM, "hipMalloc(metric range)");
    double seed[20];
    for (int k = 0; k < 10; k++) { seed[2 * k] = INFINITY; seed[2 * k + 1] = -INFINITY; }
    double r[20];
    hipError_t e = hipMemcpyAsync(rng, seed, sizeof seed, hipMemcpyHostToDevice, c->stream);
    const unsigned g = (unsigned)std::min<long>(1024, (ND + 255) / 256);
    if (e == hipSuccess)
        for (int k = 0; k < 10; k++) k_range<<<g, 256, 0, c->stream>>>(ND, G64 + k * ND, rng + 2 * k);
    if (e == hipSuccess) e = hipMemcpyAsync(r, rng, sizeof r, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(rng);
    if (e != hipSuccess) return fail(WQL_E_HIP, "metric range: %s", hipGetErrorString(e));
    if (int rc = check_launch()) return rc;
    double scale = 0.0;
    for (int k = 0; k < 9; k++) scale = std::max(scale, std::max(fabs(r[2 * k]), fabs(r[2 * k + 1])));
    const double jscale = std::max(fabs(r[18]), fabs(r[19]));
    d.gconst = 0;
    for (int k = 0; k < 10; k++) {
        const double lo = r[2 * k], hi = r[2 * k + 1];
        if (!(lo <= hi)) continue;  // NaN: leave the field to the kernels
        if (lo == hi) {
            d.gconst |= 1u << k;
            d.gc[k] = lo;
        } else if (sizeof(real) == 4 && hi - lo <= 1e-10 * (k < 9 ? scale : jscale)) {
            const real v = (real)(0.5 * (lo + hi));
            d.gconst |= 1u << k;
            d.gc[k] = (double)v;
            k_fill<real><<<(unsigned)((d.NP + 255) / 256), 256, 0, c->stream>>>(d.NP, (real *)d.G + k * d.NP, v);
        }
    }
    return check_launch();
}

template <typename real>
int create_typed(wql_ctx *c, const wql_block_desc blocks[2])
{
    static bool table_done = false;
    if (!table_done) {
        SbpTable T = make_table();
        HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_sbp), &T, sizeof T));
        double w64[12][9];
        float w32[12][9];
        memset(w64, 0, sizeof w64);
        for (int r = 0; r < 6; r++) {
            for (int t = 0; t < T.left[r].n; t++) w64[r][r + T.left[r].off[t]] += T.left[r].c[t];
            // row n-6+r; plane (n-6+r)+off relative to n-9
            for (int t = 0; t < T.right[r].n; t++) w64[6 + r][3 + r + T.right[r].off[t]] += T.right[r].c[t];
        }
        for (int r = 0; r < 12; r++)
            for (int p = 0; p < 9; p++) w32[r][p] = (float)w64[r][p];
        HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_sbw64), w64, sizeof w64));
        HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_sbw32), w32, sizeof w32));
        table_done = true;
    }
    // A pair is curvilinear when either block is, and heterogeneous when
    // either block carries M: the volume kernels are instantiated once for
    // the pair, so a Cartesian block beside a curvilinear one keeps its
    // (uniform) metrics in G like a curvilinear block, and a homogeneous
    // block beside a heterogeneous one stores its constant lambda, mu, rho
    // in M.
    const bool pair_curv = !blocks[0].cartesian || !blocks[1].cartesian;
    const bool pair_het = blocks[0].M || blocks[1].M;
    int *d_err = nullptr;
    if (int rc = dalloc(c, (void **)&d_err, sizeof(int))) return rc;
    HIPCHK(hipMemset(d_err, 0, sizeof(int)));
    for (int ib = 0; ib < 2; ib++) {
        const wql_block_desc &bd = blocks[ib];
        DevBlock &d = c->b[ib];
        d.nq = bd.nq; d.nr = bd.nr; d.ns = bd.ns;
        // q rows padded to a multiple of 128 bytes: every row of a 32-wide
        // tile then starts on a cache line
        const int lane_elems = 128 / (int)sizeof(real);
        d.px = (bd.nq + lane_elems - 1) / lane_elems * lane_elems;
        d.NP = (long)d.px * bd.nr * bd.ns;
        d.ND = (long)bd.nq * bd.nr * bd.ns;
        d.cart = pair_curv ? 0 : 1;
        for (int s = 0; s < 6; s++) { d.bc[s] = bd.bc[s]; d.pml[s] = bd.pml[s] ? 1 : 0; }
        d.npml = bd.npml;
        const long NP = d.NP, ND = d.ND;
        double *X = nullptr, *M64 = nullptr, *G64 = nullptr;
        HIPCHK(hipMalloc(&X, 3 * ND * sizeof(double)));
        HIPCHK(hipMalloc(&M64, 3 * ND * sizeof(double)));
        HIPCHK(hipMalloc(&G64, 10 * ND * sizeof(double)));
        HIPCHK(hipMemcpy(X, bd.X, 3 * ND * sizeof(double), hipMemcpyHostToDevice));
        double rho = bd.rho_s_p[0], cs = bd.rho_s_p[1], cp = bd.rho_s_p[2];
        d.lam = rho * (cp * cp - 2.0 * (cs * cs));
        d.mu = rho * (cs * cs);
        d.rho = rho;
        unsigned g = (unsigned)((ND + 255) / 256);
        if (bd.M) {
            HIPCHK(hipMemcpy(M64, bd.M, 3 * ND * sizeof(double), hipMemcpyHostToDevice));
        } else {
            k_fill<double><<<g, 256, 0, c->stream>>>(ND, M64, d.lam);
            k_fill<double><<<g, 256, 0, c->stream>>>(ND, M64 + ND, d.mu);
            k_fill<double><<<g, 256, 0, c->stream>>>(ND, M64 + 2 * ND, d.rho);
        }
        k_metrics<<<g, 256, 0, c->stream>>>(d.nq, d.nr, d.ns, X, G64);
        if (int rc = check_launch()) return rc;
        if (c->prob.anelastic) {
            // init_anelastic_properties (material.f90:11-145): Q^-1 from the
            // material, then the unrelaxed lambda and mu replace it before the
            // face frames, the faces' material and the volume kernels read it
            double *q64 = nullptr;
            HIPCHK(hipMalloc(&q64, 2 * ND * sizeof(double)));
            const double *t = c->an_tau, *w = c->an_w;
            k_anelastic_init<<<g, 256, 0, c->stream>>>(ND, M64, q64, c->prob.an_c, 2.0 * 3.141592653589793 * c->prob.an_fref,
                                                       t[0], t[1], t[2], t[3], w[0], w[1], w[2], w[3]);
            if (int rc = dalloc(c, &d.qinv, 2 * NP * sizeof(real))) return rc;
            k_pad<real><<<(unsigned)((2 * NP + 255) / 256), 256, 0, c->stream>>>(2, d.nq, d.nr, d.ns, d.px, q64, (real *)d.qinv);
            for (void **a : {&d.eta, &d.deta}) {
                if (int rc = dalloc(c, a, 24 * NP * sizeof(real))) return rc;
                k_fill<real><<<(unsigned)((24 * NP + 255) / 256), 256, 0, c->stream>>>(24 * NP, (real *)*a, (real)0);
            }
            HIPCHK(hipStreamSynchronize(c->stream));
            if (!pair_het) {  // uniform material stays uniform: its unrelaxed moduli
                HIPCHK(hipMemcpy(&d.lam, M64, sizeof(double), hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(&d.mu, M64 + ND, sizeof(double), hipMemcpyDeviceToHost));
            }
            hipFree(q64);
            if (int rc = check_launch()) return rc;
        }
        for (int side = 0; side < 6; side++) {
            int dir = side / 2;
            long FP = dir == 0 ? (long)d.nr * d.ns : dir == 1 ? (long)d.nq * d.ns : (long)d.nq * d.nr;
            if (int rc = dalloc(c, (void **)&d.frame[side], 9 * FP * sizeof(double))) return rc;
            if (int rc = dalloc(c, (void **)&d.fgm[side], 3 * FP * sizeof(double))) return rc;
            if (int rc = dalloc(c, (void **)&d.fmat[side], 3 * FP * sizeof(double))) return rc;
            bool fault_side = (ib == 0 && side == 1) || (ib == 1 && side == 0);
            if (fault_side)
                if (int rc = dalloc(c, (void **)&d.fx, 3 * FP * sizeof(double))) return rc;
            k_frames<<<(unsigned)((FP + 127) / 128), 128, 0, c->stream>>>(d.nq, d.nr, d.ns, side, G64, M64, X,
                                                                          d.frame[side], d.fgm[side], d.fmat[side],
                                                                          fault_side ? d.fx : nullptr, d_err);
        }
        if (c->prob.plastic && c->prob.s0_kind >= 1) {
            if (int rc = dalloc(c, (void **)&d.Y, NP * sizeof(double))) return rc;
            k_pad<double><<<(unsigned)((NP + 255) / 256), 256, 0, c->stream>>>(1, d.nq, d.nr, d.ns, d.px, X + ND, d.Y);
        }
        if (c->prob.plastic) {
            if (int rc = dalloc(c, (void **)&d.P, 2 * NP * sizeof(double))) return rc;
            k_fill<double><<<(unsigned)((2 * NP + 255) / 256), 256, 0, c->stream>>>(2 * NP, d.P, 0.0);
        }
        if (!d.cart) {
            if (int rc = dalloc(c, &d.G, 10 * NP * sizeof(real))) return rc;
            k_pad<real><<<(unsigned)((10 * NP + 255) / 256), 256, 0, c->stream>>>(10, d.nq, d.nr, d.ns, d.px, G64, (real *)d.G);
            if (int rc = constant_metrics<real>(c, d, G64)) return rc;
        } else {
            // metrics of a uniform block: the value at the block centre
            long pc = (long)(d.nq / 2) + (long)d.nq * ((d.nr / 2) + (long)d.nr * (d.ns / 2));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int k = 0; k < 10; k++)
                HIPCHK(hipMemcpy(&d.gc[k], G64 + k * ND + pc, sizeof(double), hipMemcpyDeviceToHost));
        }
        if (pair_het) {
            if (int rc = dalloc(c, &d.M, 3 * NP * sizeof(real))) return rc;
            k_pad<real><<<(unsigned)((3 * NP + 255) / 256), 256, 0, c->stream>>>(3, d.nq, d.nr, d.ns, d.px, M64, (real *)d.M);
        }
        for (int k = 0; k < 2; k++) {
            if (int rc = dalloc(c, &d.F[k], 9 * NP * sizeof(real))) return rc;
            k_fill<real><<<(unsigned)((9 * NP + 255) / 256), 256, 0, c->stream>>>(9 * NP, (real *)d.F[k], (real)0);
        }
        if (int rc = dalloc(c, &d.DF, 9 * NP * sizeof(real))) return rc;
        d.nbnd = 2L * d.nr * d.ns + 2L * (d.nq - 2) * d.ns + 2L * (d.nq - 2) * (d.nr - 2);
        if (int rc = dalloc(c, &d.satb, 9 * d.nbnd * sizeof(real))) return rc;
        if (int rc = dalloc(c, &d.fq, 18L * d.nr * d.ns * sizeof(real))) return rc;
        k_fill<real><<<(unsigned)((9 * NP + 255) / 256), 256, 0, c->stream>>>(9 * NP, (real *)d.DF, (real)0);
        for (int s = 0; s < 6; s++) {
            if (!d.pml[s]) continue;
            int dir = s / 2;
            long n3[3] = {d.nq, d.nr, d.ns};
            n3[dir] = d.npml;
            d.SP[s] = n3[0] * n3[1] * n3[2];
            if (int rc = dalloc(c, &d.Q[s], 9 * d.SP[s] * sizeof(real))) return rc;
            if (int rc = dalloc(c, &d.DQ[s], 9 * d.SP[s] * sizeof(real))) return rc;
            const long FPs = dir == 0 ? (long)d.nr * d.ns : dir == 1 ? (long)d.nq * d.ns : (long)d.nq * d.nr;
            if (int rc = dalloc(c, &d.pq[s], 9 * FPs * sizeof(real))) return rc;
            unsigned gq = (unsigned)((9 * d.SP[s] + 255) / 256);
            k_fill<real><<<gq, 256, 0, c->stream>>>(9 * d.SP[s], (real *)d.Q[s], (real)0);
            k_fill<real><<<gq, 256, 0, c->stream>>>(9 * d.SP[s], (real *)d.DQ[s], (real)0);
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        if (int rc = check_launch()) return rc;
        hipFree(X);
        hipFree(M64);
        hipFree(G64);
    }
    int herr = 0;
    HIPCHK(hipMemcpy(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) return fail(WQL_E_ARG, "face frame: metric gradient is not a unit vector after normalisation");
    // interface state per block copy
    for (int ib = 0; ib < 2; ib++) {
        long FP = (long)c->b[ib].nr * c->b[ib].ns;
        Iface &I = c->I[ib];
        double **arrs[9] = {&I.S, &I.DS, &I.W, &I.DW, &I.trup, &I.Svel, &I.Uhat, &I.Vhat, &I.psr};
        int nf[9] = {4, 4, 1, 1, 1, 3, 6, 6, 1};
        for (int k = 0; k < 9; k++) {
            if (int rc = dalloc(c, (void **)arrs[k], nf[k] * FP * sizeof(double))) return rc;
            double v = (k == 4) ? 1.0e9 : 0.0;  // trup = 1e9 (iface.f90:41)
            k_fill<double><<<(unsigned)((nf[k] * FP + 255) / 256), 256, 0, c->stream>>>(nf[k] * FP, *arrs[k], v);
        }
        if (c->fault.key)
            if (int rc = dalloc(c, (void **)&I.ftab, 8 * FP * sizeof(double))) return rc;
    }
    fault_tables(c);
    if (int rc = dalloc(c, (void **)&c->d_flag, sizeof(int))) return rc;
    HIPCHK(hipMemset(c->d_flag, 0, sizeof(int)));
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_launch();
}

// the per-node fault fields of a keyed problem (k_fault_tables), from each
// block's fault-face coordinates, frame and material
void fault_tables(wql_ctx *c)
{
    if (!c->fault.key) return;
    for (int ib = 0; ib < 2; ib++) {
        const DevBlock &d = c->b[ib];
        const int side = ib == 0 ? 1 : 0;
        const long FP = (long)d.nr * d.ns;
        k_fault_tables<<<(unsigned)((FP + 127) / 128), 128, 0, c->stream>>>(c->fault.key, FP, d.fx, d.frame[side],
                                                                          d.fmat[side], c->I[ib].ftab);
    }
}

long face_nodes(const DevBlock &d, int side)
{
    int dir = side / 2;
    return dir == 0 ? (long)d.nr * d.ns : dir == 1 ? (long)d.nq * d.ns : (long)d.nq * d.nr;
}

int copy_field(wql_ctx *c, const void *dev, double *host, long n)
{
    if (c->prec == WQL_FP64) {
        HIPCHK(hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> tmp(n);
        HIPCHK(hipMemcpyAsync(tmp.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (long k = 0; k < n; k++) host[k] = tmp[k];
    }
    return 0;
}

int put_field(wql_ctx *c, void *dev, const double *host, long n)
{
    if (c->prec == WQL_FP64) {
        HIPCHK(hipMemcpyAsync(dev, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> tmp(host, host + n);
        HIPCHK(hipMemcpyAsync(dev, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return 0;
}

// volume arrays between the dense host layout [nf][ns][nr][nq] (FP64) and
// the padded device layout [nf][ns][nr][px]: one contiguous copy through a
// dense FP64 staging buffer and a (un)padding kernel on the context's stream
// the context's dense staging buffer, at least n doubles (allocated once
// and grown; a transfer then costs no hipMalloc / hipFree, which would
// synchronise the device)
int staging(wql_ctx *c, size_t n, double **out)
{
    if (c->stage_elems < n) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (c->d_stage) {
            hipFree(c->d_stage);
            c->bytes -= c->stage_elems * sizeof(double);
            c->d_stage = nullptr;
            c->stage_elems = 0;
        }
        if (hipMalloc(&c->d_stage, n * sizeof(double)) != hipSuccess) {
            c->d_stage = nullptr;
            return fail(WQL_E_NOMEM, "hipMalloc(staging, %zu bytes)", n * sizeof(double));
        }
        c->stage_elems = n;
        c->bytes += n * sizeof(double);
    }
    *out = c->d_stage;
    return 0;
}

// padded device volume -> dense FP64 device buffer -> host
template <typename real>
int download_vol(wql_ctx *c, const DevBlock &d, const void *dev, double *host, int nf, double *stage)
{
    const long n = (long)nf * d.ND;
    k_unpad<real><<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(nf, d.nq, d.nr, d.ns, d.px, (const real *)dev,
                                                                       stage);
    hipError_t e = hipMemcpyAsync(host, stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(WQL_E_HIP, "volume download: %s", hipGetErrorString(e));
    return check_launch();
}

template <typename real>
int copy_vol_t(wql_ctx *c, const DevBlock &d, const void *dev, double *host, int nf)
{
    double *stage = nullptr;
    if (int rc = staging(c, (size_t)nf * d.ND, &stage)) return rc;
    return download_vol<real>(c, d, dev, host, nf, stage);
}

template <typename real>
int put_vol_t(wql_ctx *c, const DevBlock &d, void *dev, const double *host, int nf)
{
    const long n = (long)nf * d.ND;
    double *stage = nullptr;
    if (int rc = staging(c, (size_t)n, &stage)) return rc;
    hipError_t e = hipMemcpyAsync(stage, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        const long np = (long)nf * d.nr * d.ns * d.px;
        k_pad<real><<<(unsigned)((np + 255) / 256), 256, 0, c->stream>>>(nf, d.nq, d.nr, d.ns, d.px, stage,
                                                                         (real *)dev);
        e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) return fail(WQL_E_HIP, "volume upload: %s", hipGetErrorString(e));
    return check_launch();
}

int copy_vol(wql_ctx *c, const DevBlock &d, const void *dev, double *host, int nf)
{
    return c->prec == WQL_FP64 ? copy_vol_t<double>(c, d, dev, host, nf) : copy_vol_t<float>(c, d, dev, host, nf);
}

int put_vol(wql_ctx *c, const DevBlock &d, void *dev, const double *host, int nf)
{
    return c->prec == WQL_FP64 ? put_vol_t<double>(c, d, dev, host, nf) : put_vol_t<float>(c, d, dev, host, nf);
}

// device address of `scratch` on device dev (the current device), once
int scratch_setup(int dev)
{
    if (dev < 0 || dev >= MAX_DEVICES) return fail(WQL_E_ARG, "device %d beyond %d", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[dev]) return 0;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(WQL_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[dev] = (double *)p;
    return 0;
}

double *scratch_ptr(int dev)
{
    if (dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch[dev];
}

// the operator of &problem_list fd_type / order (block.f90:101-255): -1 the
// sixth-order central one, else the index in WQL_SBP_OPS, -2 if undefined
int sbp_index(const wql_problem *prob)
{
    if (prob->fd_type == WQL_FD_TRADITIONAL) return -1;
    const char *fam = prob->fd_type == WQL_FD_UPWIND ? "upwind" : prob->fd_type == WQL_FD_UPWIND_DRP ? "upwind_drp" : "";
    for (int k = 0; k < WQL_SBP_NOPS; k++)
        if (strcmp(WQL_SBP_OPS[k].family, fam) == 0 && WQL_SBP_OPS[k].order == prob->fd_order) return k;
    return -2;
}

int valid_problem(const wql_problem *prob)
{
    if (prob->precision != WQL_FP64 && prob->precision != WQL_FP32) return fail(WQL_E_ARG, "bad precision");
    if (sbp_index(prob) < -1)
        return fail(WQL_E_ARG, "no operator for fd_type %d order %d (upwind: 2-9; upwind_drp: 3-7, 66, 679)",
                    prob->fd_type, prob->fd_order);
    if (prob->anelastic) {
        if (prob->plastic) return fail(WQL_E_ARG, "a response is either plastic or anelastic");
        if (!(prob->an_c > 0) || !(prob->an_fref > 0))
            return fail(WQL_E_ARG, "anelastic c and fref must be positive");
        if (!(prob->an_weight_exp < 0.01 || (prob->an_weight_exp > 0.59 && prob->an_weight_exp < 0.61)))
            return fail(WQL_E_ARG, "anelastic weight_exp %g has no weight set (the reference defines < 0.01 "
                        "and 0.6, material.f90:80-106)", prob->an_weight_exp);
    }
    if (!(prob->dt > 0)) return fail(WQL_E_ARG, "dt must be positive");
    const wql_fault_params &f = prob->fault;
    if (f.law < 0 || f.law > 2 || f.npatch < 0 || f.npatch > 4 || f.stress_mode < 0 || f.stress_mode > 2 ||
        f.nuc < 0 || f.nuc > 3 || f.rs_law < 0 || f.rs_law > 2 || f.rs_init < 0 || f.rs_init > 1 ||
        f.sw_law < 0 || f.sw_law > 1 || f.rs_solver < 0 || f.rs_solver > 1)
        return fail(WQL_E_ARG, "bad fault description");
    if (f.key < 0 || f.key > WQL_KEY_WASATCH100) return fail(WQL_E_ARG, "bad fault key %d", f.key);
    if (f.key && f.law != WQL_SLIP_WEAKENING) return fail(WQL_E_ARG, "a keyed fault problem is slip weakening");
    if (f.law == WQL_SLIP_WEAKENING && !f.key && !(f.Dc > 0)) return fail(WQL_E_ARG, "slip weakening needs Dc > 0");
    if (f.law == WQL_RATE_STATE && !(f.rs_b > 0 && f.rs_L > 0 && f.rs_V0 > 0 && f.rs_a > 0))
        return fail(WQL_E_ARG, "rate and state needs a, b, L, V0 > 0");
    if (prob->plastic < 0 || prob->plastic > 2 || prob->s0_kind < 0 || prob->s0_kind > 2)
        return fail(WQL_E_ARG, "bad plasticity description");
    return 0;
}

int valid_block(const wql_block_desc &b)
{
    if (b.nq < 13 || b.nr < 13 || b.ns < 13) return fail(WQL_E_ARG, "block needs at least 13 nodes per direction");
    if (!b.X) return fail(WQL_E_ARG, "block coordinates X are required");
    for (int s = 0; s < 6; s++)
        if (b.bc[s] < 0 || b.bc[s] > 3) return fail(WQL_E_ARG, "boundary code %d out of range", b.bc[s]);
    for (int s = 0; s < 6; s++)
        if (b.pml[s] && (b.npml < 1 || b.npml > (s < 2 ? b.nq : s < 4 ? b.nr : b.ns)))
            return fail(WQL_E_ARG, "PML thickness %d does not fit the block", b.npml);
    if (!b.M && (b.rho_s_p[0] <= 0 || b.rho_s_p[1] <= 0 || b.rho_s_p[2] <= 0))
        return fail(WQL_E_ARG, "rho_s_p must be positive");
    return 0;
}

}  // namespace

// ============================================================================
extern "C" {

const char *wql_last_error(void) { return g_err.c_str(); }
int wql_abi_version(void) { return WQL_ABI_VERSION; }
int wql_tile_variants(void) { return NVARIANT; }

size_t wql_estimate_bytes(const wql_problem *prob, const wql_block_desc blocks[2])
{
    size_t r = prob && prob->precision == WQL_FP32 ? 4 : 8, tot = 0;
    for (int ib = 0; ib < 2; ib++) {
        const wql_block_desc &b = blocks[ib];
        const size_t lane = 128 / r;
        size_t NP = (b.nq + lane - 1) / lane * lane * (size_t)b.nr * b.ns;  // padded q rows
        size_t ND = (size_t)b.nq * b.nr * b.ns;                // dense
        size_t v = 27 * NP * r;                                // F x2, DF
        if (!blocks[0].cartesian || !blocks[1].cartesian) v += 10 * NP * r;  // metrics + J (pair curvilinear)
        if (blocks[0].M || blocks[1].M) v += 3 * NP * r;       // lambda mu rho (pair heterogeneous)
        if (prob && prob->plastic) v += 3 * NP * 8;            // strain accumulators, y
        if (prob && prob->anelastic) v += 50 * NP * r;         // memory variables, rates, Q^-1
        size_t init = 16 * ND * 8;                             // dense X, M64, G64 during wql_create
        for (int s = 0; s < 6; s++)
            if (b.pml[s]) {
                size_t n3[3] = {(size_t)b.nq, (size_t)b.nr, (size_t)b.ns};
                n3[s / 2] = (size_t)b.npml;
                v += 18 * n3[0] * n3[1] * n3[2] * r;
            }
        size_t faces = 2 * ((size_t)b.nr * b.ns + (size_t)b.nq * b.ns + (size_t)b.nq * b.nr) * 18 * 8;
        // peak during wql_create: runtime arrays and the FP64 set-up arrays
        tot += v + init + faces + 25 * (size_t)b.nr * b.ns * 8;
    }
    return tot;
}

int wql_create(const wql_problem *prob, const wql_block_desc blocks[2], wql_ctx **out)
{
    if (!prob || !blocks || !out) return fail(WQL_E_ARG, "null argument");
    *out = nullptr;
    if (prob->precision != WQL_FP64 && prob->precision != WQL_FP32) return fail(WQL_E_ARG, "bad precision");
    if (!(prob->dt > 0)) return fail(WQL_E_ARG, "dt must be positive");
    for (int ib = 0; ib < 2; ib++)
        if (int rc = valid_block(blocks[ib])) return rc;
    if (blocks[0].nr != blocks[1].nr || blocks[0].ns != blocks[1].ns)
        return fail(WQL_E_ARG, "blocks must share the fault face size (nr, ns)");
    if (blocks[0].bc[1] != 0 || blocks[1].bc[0] != 0)
        return fail(WQL_E_ARG, "the fault joins block 0 side Rx and block 1 side Lx (bc code 0)");
    if (int rc = valid_problem(prob)) return rc;
    const int sbp = sbp_index(prob);
    if (sbp >= 0)
        for (int ib = 0; ib < 2; ib++) {
            // every closure row of the operator and an interior row in each direction
            const int need = 2 * WQL_SBP_OPS[sbp].band + 1;
            if (blocks[ib].nq < need || blocks[ib].nr < need || blocks[ib].ns < need)
                return fail(WQL_E_ARG, "block %d: the %s operator of order %d needs %d nodes per direction", ib,
                            WQL_SBP_OPS[sbp].family, WQL_SBP_OPS[sbp].order, need);
        }
    for (int ib = 0; ib < 2; ib++) {
        const size_t elem = prob->precision == WQL_FP32 ? 4 : 8;
        const long lane = 128 / (long)elem;
        const long NP = (blocks[ib].nq + lane - 1) / lane * lane * (long)blocks[ib].nr * blocks[ib].ns;
        if (!fits_buffers(NP, elem))
            return fail(WQL_E_ARG, "block %d (%d x %d x %d) exceeds the 32-bit buffer offsets of the volume kernels "
                        "(10 fields x padded nodes x %zu bytes must stay below 2 GiB)",
                        ib, blocks[ib].nq, blocks[ib].nr, blocks[ib].ns, elem);
    }
    wql_ctx *c = new wql_ctx();
    c->prec = prob->precision;
    if (hipGetDevice(&c->dev) != hipSuccess) {
        delete c;
        return fail(WQL_E_HIP, "hipGetDevice");
    }
    if (int rc = scratch_setup(c->dev)) {
        delete c;
        return rc;
    }
    // fused single-pass march on 32x8 tiles (measured fastest on MI355X for
    // both precisions): FP32 re-reads the centre plane at three waves per
    // SIMD with 6-plane s chunks; FP64 holds the lead plane in registers,
    // 12-plane chunks
    c->variant = prob->precision == WQL_FP64 ? 1 : 0;
    c->zchunk = prob->precision == WQL_FP64 ? 12 : 6;
    c->prob = *prob;
    c->dt = prob->dt;
    c->t = prob->t0;
    c->sbp = sbp;
    c->tau0 = sbp >= 0 ? WQL_SBP_OPS[sbp].tau0 : TAU0;
    // the operator kernels' tables of an upwind / DRP operator (the
    // sixth-order operator with the anelastic response runs on the per-node
    // kernel's own sixth-order rows: as a table on the tiled kernel it was no
    // faster in FP32, 8.24 against 8.33 ms per step on 2 x 201x101x101, and
    // slower in FP64, 15.6 against 11.0)
    wql_sbp_half halves[6];
    const bool tabled = sbp >= 0;
    if (tabled) {
        for (int a = 0; a < 3; a++) { halves[a] = WQL_SBP_OPS[sbp].v[a]; halves[3 + a] = WQL_SBP_OPS[sbp].s[a]; }
        c->band = WQL_SBP_OPS[sbp].band;
    }
    if (tabled) {
        // the largest distance between a row and its taps, over every row
        int reach = 0;
        for (int a = 0; a < 6; a++)
            for (const wql_sbp_half *h : {&halves[a]}) {
                reach = std::max(reach, std::max(-h->lo, h->hi));
                for (int r = 0; r < h->nl; r++)
                    for (int t = 0; t < WQL_SBP_MAXW; t++)
                        if (h->left[r][t] != 0.0) reach = std::max(reach, std::abs(t - r));
                for (int r = 0; r < h->nr; r++)
                    for (int t = 0; t < WQL_SBP_MAXW; t++)
                        if (h->right[r][t] != 0.0) reach = std::max(reach, std::abs(t - r));
            }
        c->opt_h = reach <= 3 ? 3 : reach <= 5 ? 5 : 7;
    }
    if (prob->anelastic) {
        // relaxation times and weights of the four mechanisms (material.f90:80-106)
        static const double w1[4] = {1.6126, 0.6255, 0.6382, 1.5969}, w2[4] = {0.0336, 0.6873, 0.8767, 1.5202};
        const double pi = 3.141592653589793, FAC = 1.0;
        const double taumin = 1.0 / (2.0 * pi * 15.0) * 1.0 * FAC, taumax = 1.0 / (2.0 * pi * 0.08) * 200.0 * FAC;
        for (int k = 1; k <= 4; k++) {
            c->an_tau[k - 1] = std::exp(std::log(taumin) + (2.0 * k - 1.0) / 16.0 * (std::log(taumax) - std::log(taumin)));
            c->an_w[k - 1] = (prob->an_weight_exp < 0.01 ? w1 : w2)[k - 1];
        }
    }
    static_assert(sizeof(FaultParams) == sizeof(wql_fault_params), "fault layout");
    memcpy(&c->fault, &prob->fault, sizeof(FaultParams));
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(WQL_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    int rc = c->prec == WQL_FP64 ? create_typed<double>(c, blocks) : create_typed<float>(c, blocks);
    if (rc) {
        std::string msg = g_err;
        wql_destroy(c);
        g_err = msg;
        return rc;
    }
    // small grids: shorter s chunks until the main kernel has enough
    // workgroups to keep every CU busy (1.5 per CU at three waves per SIMD
    // in FP32, 0.75 per CU at one wave in FP64).  On 2x31x61x61 this picks
    // 2-plane chunks in FP32 (0.430 -> 0.313 ms per step) and 4-plane chunks
    // in FP64 (0.671 -> 0.445 ms); the large grids keep the defaults.
    {
        int dev = 0, ncu = 0;
        hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
            ncu = 256;
        const long want = c->prec == WQL_FP64 ? 3L * ncu / 4 : 3L * ncu / 2;
        while (c->zchunk > 1 && tile_list(c).total < want) c->zchunk--;
    }
    if (hipMalloc(&c->d_t, sizeof(double)) != hipSuccess) {
        wql_destroy(c);
        return fail(WQL_E_NOMEM, "hipMalloc(time)");
    }
    c->allocs.push_back(c->d_t);
    if (tabled) {
        // the operator's velocity (Dp) and stress (Dm) tables, per direction
        if (int rc2 = dalloc(c, (void **)&c->d_sbp, sizeof halves)) {
            wql_destroy(c);
            return rc2;
        }
        if (hipMemcpy(c->d_sbp, halves, sizeof halves, hipMemcpyHostToDevice) != hipSuccess) {
            wql_destroy(c);
            return fail(WQL_E_HIP, "hipMemcpy(operator table)");
        }
    }
    {
        const size_t nf = 9 * ((size_t)c->b[0].nr * c->b[0].ns + (size_t)c->b[1].nr * c->b[1].ns);
        if (hipMalloc(&c->d_faces, nf * sizeof(double)) != hipSuccess ||
            (c->allocs.push_back(c->d_faces), hipMalloc(&c->d_stats, 6 * sizeof(double)) != hipSuccess)) {
            wql_destroy(c);
            return fail(WQL_E_NOMEM, "hipMalloc(download staging)");
        }
        c->allocs.push_back(c->d_stats);
    }
    hipEventCreate(&c->ev0);
    hipEventCreate(&c->ev1);
    // the side stream is on where it shortens the step: the operator path,
    // the anelastic response and PML slabs (on the plain sixth-order path the
    // s-closure launch gains nothing beside the main kernel); WQL_OVERLAP=0
    // in the environment runs every launch of a stage on the context's one
    // stream
    bool any_pml = false;
    for (int ib = 0; ib < 2; ib++)
        for (int sd = 0; sd < 6; sd++) any_pml = any_pml || c->b[ib].pml[sd];
    c->overlap = (c->sbp >= 0 || c->prob.anelastic || any_pml) ? 1 : 0;
    if (const char *ov = getenv("WQL_OVERLAP")) c->overlap = c->overlap && atoi(ov) != 0;
    if (c->overlap) {
        if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->evf, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->evj, hipEventDisableTiming) != hipSuccess) {
            wql_destroy(c);
            return fail(WQL_E_HIP, "side stream");
        }
    }
    *out = c;
    return 0;
}

int wql_destroy(wql_ctx *c)
{
    if (!c) return 0;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (int k = 0; k < 2; k++)
        if (c->gexec[k]) hipGraphExecDestroy(c->gexec[k]);
    for (void *p : c->allocs) hipFree(p);
    if (c->d_stage) hipFree(c->d_stage);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->side) hipStreamSynchronize(c->side);
    if (c->evf) hipEventDestroy(c->evf);
    if (c->evj) hipEventDestroy(c->evj);
    if (c->side) hipStreamDestroy(c->side);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

size_t wql_device_bytes(const wql_ctx *c) { return c ? c->bytes : 0; }
double wql_time(const wql_ctx *c) { return c ? c->t : 0.0; }

int wql_init_state(wql_ctx *c)
{
    if (!c) return fail(WQL_E_ARG, "null context");
    if (c->fault.law != WQL_RATE_STATE) return 0;
    // background slip velocity of Init_Vel_State (Interface_Condition.f90:2758-2761)
    double vz[2] = {std::copysign(0.5 * c->fault.rs_Vin, -1.0), std::copysign(0.5 * c->fault.rs_Vin, 1.0)};
    for (int ib = 0; ib < 2; ib++) {
        DevBlock &d = c->b[ib];
        unsigned g = (unsigned)((d.NP + 255) / 256);
        if (c->prec == WQL_FP64)
            k_fill<double><<<g, 256, 0, c->stream>>>(d.NP, (double *)d.F[c->cur] + 2 * d.NP, vz[ib]);
        else
            k_fill<float><<<g, 256, 0, c->stream>>>(d.NP, (float *)d.F[c->cur] + 2 * d.NP, (float)vz[ib]);
    }
    long n = (long)c->b[0].nr * c->b[0].ns + (long)c->b[1].nr * c->b[1].ns;
    unsigned g = (unsigned)((n + 127) / 128);
    if (c->prec == WQL_FP64) {
        Pair<double> P;
        P.b[0] = blk_view<double>(c, 0, c->cur);
        P.b[1] = blk_view<double>(c, 1, c->cur);
        k_init_state<double><<<g, 128, 0, c->stream>>>(P, c->fault, c->I[0], c->I[1]);
    } else {
        Pair<float> P;
        P.b[0] = blk_view<float>(c, 0, c->cur);
        P.b[1] = blk_view<float>(c, 1, c->cur);
        k_init_state<float><<<g, 128, 0, c->stream>>>(P, c->fault, c->I[0], c->I[1]);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_launch();
}

int wql_set_fields(wql_ctx *c, int ib, const double *F)
{
    if (!c || ib < 0 || ib > 1 || !F) return fail(WQL_E_ARG, "bad argument");
    return put_vol(c, c->b[ib], c->b[ib].F[c->cur], F, 9);
}

int wql_get_fields(wql_ctx *c, int ib, double *F)
{
    if (!c || ib < 0 || ib > 1 || !F) return fail(WQL_E_ARG, "bad argument");
    return copy_vol(c, c->b[ib], c->b[ib].F[c->cur], F, 9);
}

int wql_get_fields_noalloc(wql_ctx *c, int ib, double *F)
{
    if (!c || ib < 0 || ib > 1 || !F) return fail(WQL_E_ARG, "bad argument");
    const DevBlock &d = c->b[ib];
    const size_t n = 9 * (size_t)d.ND;
    if (n > SCRATCH_DOUBLES)
        return fail(WQL_E_ARG, "block %d: %zu bytes of fields exceed the %zu bytes of the allocation-free staging",
                    ib, n * sizeof(double), SCRATCH_DOUBLES * sizeof(double));
    double *stage = scratch_ptr(c->dev);
    if (!stage) return fail(WQL_E_STATE, "no staging address for device %d", c->dev);
    return c->prec == WQL_FP64 ? download_vol<double>(c, d, d.F[c->cur], F, 9, stage)
                               : download_vol<float>(c, d, d.F[c->cur], F, 9, stage);
}

int wql_get_rates(wql_ctx *c, int ib, double *DF)
{
    if (!c || ib < 0 || ib > 1 || !DF) return fail(WQL_E_ARG, "bad argument");
    return copy_vol(c, c->b[ib], c->b[ib].DF, DF, 9);
}

int wql_set_anelastic(wql_ctx *c, int ib, const double *eta)
{
    if (!c || ib < 0 || ib > 1 || !eta) return fail(WQL_E_ARG, "bad argument");
    if (!c->b[ib].eta) return fail(WQL_E_ARG, "the problem has no anelastic response");
    return put_vol(c, c->b[ib], c->b[ib].eta, eta, 24);
}

int wql_get_anelastic(wql_ctx *c, int ib, int what, double *out)
{
    if (!c || ib < 0 || ib > 1 || !out || what < 0 || what > 1) return fail(WQL_E_ARG, "bad argument");
    if (!c->b[ib].eta) return fail(WQL_E_ARG, "the problem has no anelastic response");
    return copy_vol(c, c->b[ib], what == 0 ? c->b[ib].eta : c->b[ib].deta, out, 24);
}

int wql_set_pml(wql_ctx *c, int ib, int side, const double *Q)
{
    if (!c || ib < 0 || ib > 1 || side < 0 || side > 5 || !Q || !c->b[ib].pml[side])
        return fail(WQL_E_ARG, "bad argument or no PML on that side");
    return put_field(c, c->b[ib].Q[side], Q, 9 * c->b[ib].SP[side]);
}

int wql_get_pml(wql_ctx *c, int ib, int side, double *Q)
{
    if (!c || ib < 0 || ib > 1 || side < 0 || side > 5 || !Q || !c->b[ib].pml[side])
        return fail(WQL_E_ARG, "bad argument or no PML on that side");
    return copy_field(c, c->b[ib].Q[side], Q, 9 * c->b[ib].SP[side]);
}

int wql_get_pml_rates(wql_ctx *c, int ib, int side, double *DQ)
{
    if (!c || ib < 0 || ib > 1 || side < 0 || side > 5 || !DQ || !c->b[ib].pml[side])
        return fail(WQL_E_ARG, "bad argument or no PML on that side");
    return copy_field(c, c->b[ib].DQ[side], DQ, 9 * c->b[ib].SP[side]);
}

int wql_get_metrics(wql_ctx *c, int ib, double *metric)
{
    if (!c || ib < 0 || ib > 1 || !metric) return fail(WQL_E_ARG, "bad argument");
    const DevBlock &d = c->b[ib];
    if (d.cart) {
        for (int k = 0; k < 10; k++)
            for (long p = 0; p < d.ND; p++) metric[k * d.ND + p] = d.gc[k];
        return 0;
    }
    return copy_vol(c, d, d.G, metric, 10);
}

int wql_get_frame(wql_ctx *c, int ib, int side, double *nml)
{
    if (!c || ib < 0 || ib > 1 || side < 0 || side > 5 || !nml) return fail(WQL_E_ARG, "bad argument");
    long FP = face_nodes(c->b[ib], side);
    HIPCHK(hipMemcpy(nml, c->b[ib].frame[side], 9 * FP * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int wql_set_fault(wql_ctx *c, int ib, const double *S, const double *W)
{
    if (!c || ib < 0 || ib > 1) return fail(WQL_E_ARG, "bad argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    long FP = (long)c->b[ib].nr * c->b[ib].ns;
    if (S) HIPCHK(hipMemcpy(c->I[ib].S, S, 4 * FP * sizeof(double), hipMemcpyHostToDevice));
    if (W) HIPCHK(hipMemcpy(c->I[ib].W, W, FP * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int wql_get_fault(wql_ctx *c, int ib, double *S, double *W, double *trup, double *Svel, double *Uhat, double *Vhat)
{
    if (!c || ib < 0 || ib > 1) return fail(WQL_E_ARG, "bad argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    long FP = (long)c->b[ib].nr * c->b[ib].ns;
    const Iface &I = c->I[ib];
    if (S) HIPCHK(hipMemcpy(S, I.S, 4 * FP * sizeof(double), hipMemcpyDeviceToHost));
    if (W) HIPCHK(hipMemcpy(W, I.W, FP * sizeof(double), hipMemcpyDeviceToHost));
    if (trup) HIPCHK(hipMemcpy(trup, I.trup, FP * sizeof(double), hipMemcpyDeviceToHost));
    if (Svel) HIPCHK(hipMemcpy(Svel, I.Svel, 3 * FP * sizeof(double), hipMemcpyDeviceToHost));
    if (Uhat) HIPCHK(hipMemcpy(Uhat, I.Uhat, 6 * FP * sizeof(double), hipMemcpyDeviceToHost));
    if (Vhat) HIPCHK(hipMemcpy(Vhat, I.Vhat, 6 * FP * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int wql_get_fault_faces(wql_ctx *c, double *Uface, double *Vface)
{
    if (!c) return fail(WQL_E_ARG, "bad argument");
    double *dst[2] = {Uface, Vface};
    for (int ib = 0; ib < 2; ib++) {
        if (!dst[ib]) continue;
        const DevBlock &d = c->b[ib];
        const long n = 9L * d.nr * d.ns;
        double *stage = c->d_faces + (ib == 0 ? 0 : 9L * c->b[0].nr * c->b[0].ns);
        // block 0's fault face is q = nq-1, block 1's is q = 0
        const int i = ib == 0 ? d.nq - 1 : 0;
        const unsigned g = (unsigned)((n + 255) / 256);
        if (c->prec == WQL_FP64)
            k_qface<double><<<g, 256, 0, c->stream>>>(i, d.nr, d.ns, d.px, d.NP, (const double *)d.F[c->cur], stage);
        else
            k_qface<float><<<g, 256, 0, c->stream>>>(i, d.nr, d.ns, d.px, d.NP, (const float *)d.F[c->cur], stage);
        hipError_t e = hipMemcpyAsync(dst[ib], stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return fail(WQL_E_HIP, "fault face download: %s", hipGetErrorString(e));
        if (int rc = check_launch()) return rc;
    }
    return 0;
}

int wql_locate_sources(const double *X, int nq, int nr, int ns, int n, const double *xyz, wql_source *src)
{
    if (!X || nq < 2 || nr < 2 || ns < 2 || n < 0 || (n > 0 && (!xyz || !src))) return fail(WQL_E_ARG, "bad argument");
    const long N = (long)nq * nr * ns;
    auto at = [&](int c, int i, int j, int k) { return X[c * N + i + (long)nq * ((long)j + (long)nr * k)]; };
    double lo[3], hi[3];
    for (int c = 0; c < 3; c++) {
        lo[c] = hi[c] = X[c * N];
        for (long p = 1; p < N; p++) {
            lo[c] = std::min(lo[c], X[c * N + p]);
            hi[c] = std::max(hi[c], X[c * N + p]);
        }
    }
    for (int s = 0; s < n; s++) {
        const double x1 = xyz[3 * s], y1 = xyz[3 * s + 1], z1 = xyz[3 * s + 2];
        int fi = -1, fj = -1, fk = -1;
        // Find_Coordinates_moment (seismogram.f90:652-732): the nearest node,
        // the last one in (s, r, q) order on ties, for a location inside the
        // block's coordinate ranges
        if (lo[0] <= x1 && x1 <= hi[0] && lo[1] <= y1 && y1 <= hi[1] && lo[2] <= z1 && z1 <= hi[2]) {
            double mindist = 1.0e8;
            for (int k = 0; k < ns; k++)
                for (int j = 0; j < nr; j++)
                    for (int i = 0; i < nq; i++) {
                        const double v0 = at(0, i, j, k) - x1, v1 = at(1, i, j, k) - y1, v2 = at(2, i, j, k) - z1;
                        const double dist = std::sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                        if (dist <= mindist) {
                            fi = i; fj = j; fk = k;
                            mindist = dist;
                        }
                    }
        }
        // init_moment_tensor (moment_tensor.f90:305-322): kept only when the
        // location lies between the block's first and last nodes
        const bool inb = at(0, 0, 0, 0) <= x1 && x1 <= at(0, nq - 1, nr - 1, ns - 1) && at(1, 0, 0, 0) <= y1 &&
                         y1 <= at(1, nq - 1, nr - 1, ns - 1) && at(2, 0, 0, 0) <= z1 && z1 <= at(2, nq - 1, nr - 1, ns - 1);
        if (inb && fi >= 0) {
            src[s].node[0] = fi; src[s].node[1] = fj; src[s].node[2] = fk;
            src[s].node_x[0] = at(0, fi, fj, fk);
            src[s].node_x[1] = at(1, fi, fj, fk);
            src[s].node_x[2] = at(2, fi, fj, fk);
        } else {
            for (int c = 0; c < 3; c++) { src[s].node[c] = -2; src[s].node_x[c] = -1.0; }
        }
    }
    return 0;
}

int wql_set_sources(wql_ctx *c, int ib, int n, const wql_source *src, int order, int alpha)
{
    if (!c || ib < 0 || ib > 1 || n < 0 || (n > 0 && !src)) return fail(WQL_E_ARG, "bad argument");
    if (order != 2 && order != 4) return fail(WQL_E_ARG, "moment tensor order %d: 2 or 4", order);
    if (order == 4 && alpha != 0 && alpha != 25 && alpha != 50 && alpha != 100)
        return fail(WQL_E_ARG, "moment tensor alpha %d: 0, 25, 50 or 100 with order 4", alpha);
    DevBlock &d = c->b[ib];
    std::vector<SourceDev> h(n);
    const int w = order == 2 ? 2 : 4;
    const int nn[3] = {d.nq, d.nr, d.ns};
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int s = 0; s < n; s++) {
        if (src[s].type < WQL_STF_GAUSSIAN || src[s].type > WQL_STF_SINE_MOMENT)
            return fail(WQL_E_ARG, "source %d: unknown source time function %d", s, src[s].type);
        SourceDev &o = h[s];
        o.type = src[s].type;
        o.duration = src[s].duration;
        o.t_init = src[s].t_init;
        for (int k = 0; k < 6; k++) o.m[k] = src[s].m[k];
        for (int k = 0; k < 3; k++) {
            o.x[k] = src[s].x[k];
            o.node[k] = src[s].node[k];
            o.node_x[k] = src[s].node_x[k];
            if (o.node[k] < -2 || o.node[k] >= nn[k])
                return fail(WQL_E_ARG, "source %d: node index %d outside block %d", s, o.node[k], ib + 1);
            const int a0 = std::max(0, o.node[k] - w), a1 = std::min(nn[k], o.node[k] + w + 1);
            if (s == 0 || a0 < lo[k]) lo[k] = a0;
            if (s == 0 || a1 > hi[k]) hi[k] = a1;
        }
    }
    if (n > d.src_cap) {
        if (int rc = dalloc(c, (void **)&d.src, n * sizeof(SourceDev))) return rc;
        d.src_cap = n;
    }
    drop_graphs(c);
    d.nsrc = n;
    d.src_smooth = 0;
    d.src_order = order;
    d.src_alpha = alpha;
    for (int k = 0; k < 3; k++) { d.src_lo[k] = lo[k]; d.src_hi[k] = hi[k]; }
    if (n > 0) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(d.src, h.data(), n * sizeof(SourceDev), hipMemcpyHostToDevice));
    }
    return 0;
}

int wql_set_sources_smooth(wql_ctx *c, int ib, int n, const wql_source *src, const double *X)
{
    if (!c || ib < 0 || ib > 1 || n < 0 || (n > 0 && (!src || !X))) return fail(WQL_E_ARG, "bad argument");
    DevBlock &d = c->b[ib];
    std::vector<SourceDev> h(n);
    for (int s = 0; s < n; s++) {
        if (src[s].type < WQL_STF_GAUSSIAN || src[s].type > WQL_STF_SINE_MOMENT)
            return fail(WQL_E_ARG, "source %d: unknown source time function %d", s, src[s].type);
        SourceDev &o = h[s];
        o.type = src[s].type;
        o.duration = src[s].duration;
        o.t_init = src[s].t_init;
        for (int k = 0; k < 6; k++) o.m[k] = src[s].m[k];
        for (int k = 0; k < 3; k++) {
            o.x[k] = src[s].x[k];
            o.node[k] = -1;
            o.node_x[k] = -1.0;
        }
    }
    if (n > d.src_cap) {
        if (int rc = dalloc(c, (void **)&d.src, n * sizeof(SourceDev))) return rc;
        d.src_cap = n;
    }
    if (n > 0 && !d.Xs)
        if (int rc = dalloc(c, (void **)&d.Xs, 3 * (size_t)d.ND * sizeof(double))) return rc;
    drop_graphs(c);
    d.nsrc = n;
    d.src_smooth = 1;
    d.src_order = 2;
    d.src_alpha = 0;
    d.src_lo[0] = d.src_lo[1] = d.src_lo[2] = 0;
    d.src_hi[0] = d.nq; d.src_hi[1] = d.nr; d.src_hi[2] = d.ns;
    if (n > 0) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(d.src, h.data(), n * sizeof(SourceDev), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d.Xs, X, 3 * (size_t)d.ND * sizeof(double), hipMemcpyHostToDevice));
    }
    return 0;
}

int wql_set_stations(wql_ctx *c, int ib, int n, const int *ijk)
{
    if (!c || ib < 0 || ib > 1 || n < 0 || (n > 0 && !ijk)) return fail(WQL_E_ARG, "bad argument");
    DevBlock &d = c->b[ib];
    std::vector<long> off(n);
    for (int s = 0; s < n; s++) {
        const int i = ijk[3 * s], j = ijk[3 * s + 1], k = ijk[3 * s + 2];
        if (i < 0 || i >= d.nq || j < 0 || j >= d.nr || k < 0 || k >= d.ns)
            return fail(WQL_E_ARG, "station %d: node (%d, %d, %d) outside block %d", s, i, j, k, ib + 1);
        off[s] = (long)i + (long)d.px * ((long)j + (long)d.nr * k);
    }
    if (n > d.st_cap) {
        if (int rc = dalloc(c, (void **)&d.st_off, n * sizeof(long))) return rc;
        if (int rc = dalloc(c, (void **)&d.st_v, 3L * n * sizeof(double))) return rc;
        d.st_cap = n;
    }
    d.nst = n;
    if (n > 0) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(d.st_off, off.data(), n * sizeof(long), hipMemcpyHostToDevice));
    }
    return 0;
}

int wql_get_stations(wql_ctx *c, int ib, double *v)
{
    if (!c || ib < 0 || ib > 1 || (c->b[ib].nst > 0 && !v)) return fail(WQL_E_ARG, "bad argument");
    const DevBlock &d = c->b[ib];
    if (d.nst == 0) return 0;
    const unsigned g = (unsigned)((3 * d.nst + 255) / 256);
    if (c->prec == WQL_FP64)
        k_station_gather<double><<<g, 256, 0, c->stream>>>(d.nst, d.st_off, d.NP, (const double *)d.F[c->cur], d.st_v);
    else
        k_station_gather<float><<<g, 256, 0, c->stream>>>(d.nst, d.st_off, d.NP, (const float *)d.F[c->cur], d.st_v);
    hipError_t e = hipMemcpyAsync(v, d.st_v, 3L * d.nst * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(WQL_E_HIP, "station download: %s", hipGetErrorString(e));
    return check_launch();
}

// Find_Coordinates (seismogram.f90:531-650): per station the node nearest
// to (x, y, z) among those within half the local grid diagonal (spacings to
// the next node in q, r and s, or to the previous one at the last node),
// the last such node in (s, r, q) loop order on ties; -1 where none is
int wql_find_stations(const double *X, int nq, int nr, int ns, int n, const double *xyz, int *ijk)
{
#pragma clang fp contract(off)
    if (!X || nq < 2 || nr < 2 || ns < 2 || n < 0 || (n > 0 && (!xyz || !ijk))) return fail(WQL_E_ARG, "bad argument");
    const long N = (long)nq * nr * ns;
    auto at = [&](int c, int i, int j, int k) { return X[c * N + i + (long)nq * ((long)j + (long)nr * k)]; };
    for (int st = 0; st < n; st++) {
        const double x1 = xyz[3 * st], y1 = xyz[3 * st + 1], z1 = xyz[3 * st + 2];
        double mindist = 1.0e8;
        int fi = -1, fj = -1, fk = -1;
        for (int k = 0; k < ns; k++)
            for (int j = 0; j < nr; j++)
                for (int i = 0; i < nq; i++) {
                    const double hx = i < nq - 1 ? at(0, i + 1, j, k) - at(0, i, j, k) : at(0, nq - 1, j, k) - at(0, nq - 2, j, k);
                    const double hy = j < nr - 1 ? at(1, i, j + 1, k) - at(1, i, j, k) : at(1, i, nr - 1, k) - at(1, i, nr - 2, k);
                    const double hz = k < ns - 1 ? at(2, i, j, k + 1) - at(2, i, j, k) : at(2, i, j, ns - 1) - at(2, i, j, ns - 2);
                    const double v0 = at(0, i, j, k) - x1, v1 = at(1, i, j, k) - y1, v2 = at(2, i, j, k) - z1;
                    const double dist = std::sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                    const double dist0 = std::sqrt(hx * hx + hy * hy + hz * hz);
                    if (dist <= mindist && dist <= 0.5 * dist0) {
                        fi = i; fj = j; fk = k;
                        mindist = dist;
                    }
                }
        ijk[3 * st] = fi;
        ijk[3 * st + 1] = fj;
        ijk[3 * st + 2] = fk;
    }
    return 0;
}

int wql_get_fault_peak(wql_ctx *c, int ib, double *psr)
{
    if (!c || ib < 0 || ib > 1 || !psr) return fail(WQL_E_ARG, "bad argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    long FP = (long)c->b[ib].nr * c->b[ib].ns;
    HIPCHK(hipMemcpy(psr, c->I[ib].psr, FP * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int wql_fault_stats(wql_ctx *c, int ib, double stats[6])
{
    if (!c || ib < 0 || ib > 1 || !stats) return fail(WQL_E_ARG, "bad argument");
    long FP = (long)c->b[ib].nr * c->b[ib].ns;
    double *d_out = c->d_stats;
    k_fault_stats<<<1, 1024, 0, c->stream>>>(c->I[ib], FP, d_out);
    hipError_t e = hipMemcpyAsync(stats, d_out, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(WQL_E_HIP, "fault stats: %s", hipGetErrorString(e));
    return check_launch();
}

int wql_get_fault_rates(wql_ctx *c, int ib, double *DS, double *DW)
{
    if (!c || ib < 0 || ib > 1) return fail(WQL_E_ARG, "bad argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    long FP = (long)c->b[ib].nr * c->b[ib].ns;
    if (DS) HIPCHK(hipMemcpy(DS, c->I[ib].DS, 4 * FP * sizeof(double), hipMemcpyDeviceToHost));
    if (DW) HIPCHK(hipMemcpy(DW, c->I[ib].DW, FP * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// the slip-weakening problems with per-node fault fields (wql_fault_params.key)
int fault_key(const std::string &pk)
{
    static const struct { const char *name; int key; } keys[] = {
        {"TPV10", WQL_KEY_TPV10}, {"TPV29", WQL_KEY_TPV29}, {"TPV30", WQL_KEY_TPV29}, {"TPV31", WQL_KEY_TPV31},
        {"TPV32", WQL_KEY_TPV32}, {"TPV33", WQL_KEY_TPV33}, {"TPV34", WQL_KEY_TPV34},
        {"SCITS2016", WQL_KEY_SCITS2016}, {"TPV36", WQL_KEY_TPV36}, {"TPV37", WQL_KEY_TPV37},
        {"wasatch100", WQL_KEY_WASATCH100}};
    for (const auto &k : keys)
        if (pk == k.name) return k.key;
    return WQL_KEY_NONE;
}

int wql_problem_keys(const char *problem, const char *coupling, const char *response, const char *plastic_model,
                     const double mu_beta_eta[3], wql_problem *out)
{
    if (!problem || !coupling || !out) return fail(WQL_E_ARG, "null argument");
    const std::string pk = problem, ck = coupling, rk = response ? response : "elastic",
                      mk = plastic_model ? plastic_model : "default";
    wql_problem p;
    memset(&p, 0, sizeof p);
    p.precision = out->precision;
    p.dt = out->dt;
    p.t0 = out->t0;
    // plasticity constants of plastic_flow2 (plastic.f90:174-182)
    p.pl_c = 1.36; p.pl_nu = 0.1934; p.pl_Tv = 0.03;
    wql_fault_params &f = p.fault;
    f.Dc = 1.0; f.bar_wy = 1e30; f.bar_wz = 1e30; f.rs_wr = 1.0; f.rs_L = 1.0; f.rs_V0 = 1.0;
    f.frc[2] = f.frc[3] = f.frc[4] = 1.0;
    auto patch = [&](int k, double yc, double zc, double wy, double wz, double amp) {
        f.patch[k][0] = yc; f.patch[k][1] = zc; f.patch[k][2] = wy; f.patch[k][3] = wz; f.patch[k][4] = amp;
    };
    if (ck == "locked") {
        f.law = WQL_LOCKED;
    } else if (ck == "slip-weakening_friction" && pk == "TPV5") {  // Interface_Condition.f90:1283-1299, 1822-1830
        f.law = WQL_SLIP_WEAKENING;
        f.stress_mode = 0;
        f.T0[0] = -120.0; f.T0[1] = 0.0; f.T0[2] = 70.0;
        f.npatch = 3;
        patch(0, 7.5, -7.5, 1.5, 1.5, 8.0);
        patch(1, 7.5, 0.0, 1.5, 1.5, 11.6);
        patch(2, 7.5, 7.5, 1.5, 1.5, -8.0);
        f.fs = 0.677; f.fd = 0.525; f.Dc = 0.4;
        f.bar_wy = 15.0; f.bar_wz = 15.0; f.bar_big = 1.0e10;
    } else if (ck == "slip-weakening_friction" && pk == "TPV28") {  // 1459-1474, 2049-2079
        f.law = WQL_SLIP_WEAKENING;
        f.stress_mode = 1;
        const double s0[6] = {-60.0, 0.0, 60.0, 0.0, 29.38, 0.0};  // initial_stress_condition.f90:78-84
        memcpy(f.s0, s0, sizeof s0);
        f.nuc = 1;
        const double np[5] = {7.5, 0.0, 1.4, 0.6, 11.6};
        memcpy(f.nucp, np, sizeof np);
        f.fs = 0.677; f.fd = 0.373; f.Dc = 0.4;
        f.bar_wy = 15.0; f.bar_wz = 18.0; f.bar_big = 1.0e10;
    } else if (ck == "slip-weakening_friction" && (pk == "TPV26" || pk == "TPV27")) {  // 1376-1458, 2008-2048
        f.law = WQL_SLIP_WEAKENING;
        f.stress_mode = 2;
        f.fs = 0.18; f.fd = 0.12; f.Dc = 0.3;
        f.bar_wy = 20.0; f.bar_wz = 20.0; f.bar_big = 1.0e10;
        f.sw_law = 1;
        f.coh[0] = 0.4; f.coh[1] = 0.72; f.coh[2] = 5.0;
        f.frc[0] = 10.0; f.frc[1] = -5.0; f.frc[2] = 4.0; f.frc[3] = 3.464; f.frc[4] = 0.5;
    } else if (ck == "slip-weakening_friction" && fault_key(pk) != WQL_KEY_NONE) {
        // the keyed problems whose fault fields k_fault_tables evaluates per node
        f.law = WQL_SLIP_WEAKENING;
        f.key = fault_key(pk);
    } else if (ck == "rate-and-state_friction" && (pk == "TPV101" || pk == "TPV102")) {
        // 386-400, 433-436, 2324-2385, 2756-2819, initial_stress_condition.f90:114-120
        f.law = WQL_RATE_STATE;
        f.stress_mode = 1;
        const double s0[6] = {-120.0, 0.0, -120.0, 0.0, 75.0, 0.0};
        memcpy(f.s0, s0, sizeof s0);
        f.nuc = 2;
        const double np[5] = {7.5, 0.0, 3.0, 25.0, 1.0};
        memcpy(f.nucp, np, sizeof np);
        f.rs_a = 0.008; f.rs_da = 0.008; f.rs_w = 15.0; f.rs_wr = 3.0; f.rs_b = 0.012; f.rs_L = 0.02;
        f.rs_V0 = 1.0e-6; f.rs_f0 = 0.6; f.rs_Vin = 1.0e-12;
    } else if (ck == "rate-and-state_friction" && pk == "rate-weakening") {
        // 410-416, 438-444, 2387-2420, 2820-2850, initial_stress_condition.f90:184-192 (with its pi constant)
        f.law = WQL_RATE_STATE;
        f.stress_mode = 1;
        const double pi = 3.14159265359, psi = 50.0 / 180.0 * pi;
        const double s1 = -126.0, s5 = 42.0;
        const double s3 = (1.0 - 2.0 * s5 / (s1 * tan(2.0 * psi))) * s1;
        f.s0[0] = s1; f.s0[1] = 0.5 * (s1 + s3); f.s0[2] = s3; f.s0[3] = 0.0; f.s0[4] = s5; f.s0[5] = 0.0;
        f.nuc = 3;
        f.nucp[0] = 8.0; f.nucp[1] = 0.0; f.nucp[2] = 3.0 * 1.5 * 0.3; f.nucp[3] = -0.2429; f.nucp[4] = 0.0;
        f.rs_a = 0.016; f.rs_da = 0.016; f.rs_w = 20.0; f.rs_wr = 2.5; f.rs_b = 0.02; f.rs_L = 2.0 * 0.2572;
        f.rs_V0 = 1.0e-6; f.rs_f0 = 0.7; f.rs_Vin = 0.0;
        f.rs_law = 1; f.rs_fw = 0.13; f.rs_Vw = 0.17; f.rs_dVw = 5.0;
        f.rs_init = 1; f.rs_W0 = 0.4367;
    } else if (ck == "rate-and-state_friction" && pk == "KDN") {
        return fail(WQL_E_ARG, "problem 'KDN' has no prestress in the reference (Interface_Condition.f90:2425-2427); "
                               "describe the fault explicitly (law WQL_RATE_STATE, rs_law 2)");
    } else {
        return fail(WQL_E_ARG, "unsupported problem/coupling: %s/%s", problem, coupling);
    }
    if (rk == "plastic") {
        p.plastic = mk == "scec" ? 1 : 2;
        // the plastic background stress is the problem's initial stress
        // tensor whatever the coupling (initial_stress_condition.f90:8-254)
        if (pk == "TPV26" || pk == "TPV27") {
            p.s0_kind = 1;
        } else if (pk == "TPV29" || pk == "TPV30") {
            p.s0_kind = 2;
        } else if (pk == "TPV28") {
            const double s0[6] = {-60.0, 0.0, 60.0, 0.0, 29.38, 0.0};
            memcpy(f.s0, s0, sizeof s0);
        } else if (pk == "TPV101" || pk == "TPV102") {
            const double s0[6] = {-120.0, 0.0, -120.0, 0.0, 75.0, 0.0};
            memcpy(f.s0, s0, sizeof s0);
        } else if (pk == "rate-weakening") {
            const double pi = 3.14159265359, psi = 50.0 / 180.0 * pi;
            const double s1 = -126.0, s5 = 42.0;
            const double s3 = (1.0 - 2.0 * s5 / (s1 * tan(2.0 * psi))) * s1;
            f.s0[0] = s1; f.s0[1] = 0.5 * (s1 + s3); f.s0[2] = s3; f.s0[3] = 0.0; f.s0[4] = s5; f.s0[5] = 0.0;
        } else {
            return fail(WQL_E_ARG, "no initial stress tensor for problem '%s' (initial_stress_condition.f90:250)",
                        problem);
        }
        if (pk == "TPV30") { p.pl_c = 1.18; p.pl_nu = 0.1680; p.pl_Tv = 0.05; }
        if (p.plastic == 2 && mu_beta_eta) {  // plastic_material.f90:38-44
            p.pl_mu = mu_beta_eta[0]; p.pl_beta = mu_beta_eta[1]; p.pl_eta = mu_beta_eta[2];
        }
    } else if (rk == "anelastic") {
        // the &anelastic_list defaults (material.f90:32-34); the host sets
        // the list's values
        p.anelastic = 1; p.an_c = 1.0; p.an_weight_exp = 0.0; p.an_fref = 1.0;
    } else if (rk != "elastic") {
        return fail(WQL_E_ARG, "unsupported response '%s'", rk.c_str());
    }
    *out = p;
    return 0;
}

int wql_grid_faces(int nq, int nr, int ns, const double *faces, double *X)
{
    if (!faces || !X) return fail(WQL_E_ARG, "null argument");
    if (nq < 2 || nr < 2 || ns < 2) return fail(WQL_E_ARG, "a grid needs two nodes per direction");
    const long N = (long)nq * nr * ns;
    const long FQ = (long)nr * ns, FS = (long)nq * nr, FR = (long)nq * ns;
    const long NF = 3 * 2 * (FQ + FR + FS);
    double *d = nullptr;
    if (hipMalloc(&d, (size_t)(3 * N + NF) * sizeof(double)) != hipSuccess) return fail(WQL_E_NOMEM, "hipMalloc(grid)");
    GridArgs g;
    memset(&g, 0, sizeof g);
    g.nq = nq; g.nr = nr; g.ns = ns;
    g.X = d;
    g.faces = d + 3 * N;
    hipError_t e = hipMemcpy(d + 3 * N, faces, (size_t)NF * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_grid_volume<<<(unsigned)((3 * N + 255) / 256), 256>>>(g);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(X, d, (size_t)(3 * N) * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(WQL_E_HIP, "grid faces: %s", hipGetErrorString(e));
    return 0;
}

int wql_problem_material(const char *problem, int nq, int nr, int ns, const double *X, double *M)
{
    if (!problem) return fail(WQL_E_ARG, "null argument");
    static const struct { const char *name; int model; } media[] = {
        {"LOH1", WQL_MEDIUM_LOH1}, {"LOH1_Harmonic", WQL_MEDIUM_LOH1_HARMONIC}, {"TPV31", WQL_MEDIUM_TPV31},
        {"TPV32", WQL_MEDIUM_TPV32}, {"TPV33", WQL_MEDIUM_TPV33}, {"OKLAHOMA", WQL_MEDIUM_OKLAHOMA}};
    int model = WQL_MEDIUM_NONE;
    for (const auto &m : media)
        if (strcmp(problem, m.name) == 0) model = m.model;
    if (model == WQL_MEDIUM_NONE) return 0;
    if (!X || !M || nq < 1 || nr < 1 || ns < 1) return fail(WQL_E_ARG, "bad argument");
    const long N = (long)nq * nr * ns;
    for (long p = 0; p < N; p++) {
        double vp, vs, rho;
        if (!layered_medium(model, X[p], X[N + p], vp, vs, rho))
            return fail(WQL_E_ARG, "node %ld: coordinate outside every layer of the %s medium", p, problem);
        M[p] = rho * (vp * vp - 2.0 * (vs * vs));  // lambda (material.f90: rho (Vp**2 - 2 Vs**2))
        M[N + p] = rho * (vs * vs);                // mu
        M[2 * N + p] = rho;
    }
    return 1;
}

int wql_grid_block(const wql_grid_desc *gd, double *X)
{
    if (!gd || !X) return fail(WQL_E_ARG, "null argument");
    if (gd->nq < 2 || gd->nr < 2 || gd->ns < 2) return fail(WQL_E_ARG, "a grid needs two nodes per direction");
    if (gd->mesh < 0 || gd->mesh > 1 || gd->profile < 0 || gd->profile > 4 || gd->fault_side < 0 ||
        gd->fault_side > 1)
        return fail(WQL_E_ARG, "bad grid description");
    const long N = (long)gd->nq * gd->nr * gd->ns;
    const long FQ = (long)gd->nr * gd->ns, FS = (long)gd->nq * gd->nr, FR = (long)gd->nq * gd->ns;
    const size_t work = (size_t)(3 * N + 4 * FQ + 4 * FS + 4 * FR + FQ + FR);
    double *d = nullptr;
    if (hipMalloc(&d, work * sizeof(double)) != hipSuccess) return fail(WQL_E_NOMEM, "hipMalloc(grid)");
    GridArgs g;
    memset(&g, 0, sizeof g);
    g.nq = gd->nq; g.nr = gd->nr; g.ns = gd->ns;
    for (int k = 0; k < 3; k++) { g.a[k] = gd->a[k]; g.b[k] = gd->b[k]; }
    g.lc = gd->lc; g.rc = gd->rc;
    g.profile = gd->profile;
    g.fault_side = gd->fault_side;
    g.X = d;
    g.qf = d + 3 * N;
    g.sf = g.qf + 4 * FQ;
    g.rf = g.sf + 4 * FS;
    double *fx = g.rf + 4 * FR, *ty = fx + FQ;
    hipError_t e = hipSuccess;
    if (gd->mesh == 1 && gd->fault_x) {
        e = hipMemcpy(fx, gd->fault_x, FQ * sizeof(double), hipMemcpyHostToDevice);
        g.fault_x = fx;
    }
    if (e == hipSuccess && gd->mesh == 1 && gd->top_y) {
        e = hipMemcpy(ty, gd->top_y, FR * sizeof(double), hipMemcpyHostToDevice);
        g.top_y = ty;
    }
    if (e == hipSuccess) {
        if (gd->mesh == 0) {
            k_grid_cartesian<<<(unsigned)((3 * N + 255) / 256), 256>>>(g.nq, g.nr, g.ns, g.a[0], g.a[1], g.a[2], g.b[0],
                                                                        g.b[1], g.b[2], g.X);
        } else {
            k_grid_qfaces<<<(unsigned)((FQ + 255) / 256), 256>>>(g);
            k_grid_sfaces<<<(unsigned)((2 * FS + 255) / 256), 256>>>(g);
            k_grid_rfaces<<<(unsigned)((2 * FR + 255) / 256), 256>>>(g);
            k_grid_volume<<<(unsigned)((3 * N + 255) / 256), 256>>>(g);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(X, g.X, 3 * N * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(WQL_E_HIP, "grid: %s", hipGetErrorString(e));
    return 0;
}

int wql_get_plastic(wql_ctx *c, int ib, double *P)
{
    if (!c || ib < 0 || ib > 1 || !P) return fail(WQL_E_ARG, "bad argument");
    const DevBlock &d = c->b[ib];
    if (!d.P) return fail(WQL_E_STATE, "elastic problem: no plastic state");
    return copy_vol_t<double>(c, d, d.P, P, 2);
}

int wql_eval_rates(wql_ctx *c, double t, int what, int stage)
{
    if (!c || what < 0 || what > 1) return fail(WQL_E_ARG, "bad argument");
    for (int ib = 0; ib < 2; ib++) {
        DevBlock &d = c->b[ib];
        unsigned g = (unsigned)((9 * d.NP + 255) / 256);
        if (c->prec == WQL_FP64) k_fill<double><<<g, 256, 0, c->stream>>>(9 * d.NP, (double *)d.DF, 0.0);
        else k_fill<float><<<g, 256, 0, c->stream>>>(9 * d.NP, (float *)d.DF, 0.0f);
        for (int s = 0; s < 6; s++)
            if (d.pml[s]) {
                unsigned gq = (unsigned)((9 * d.SP[s] + 255) / 256);
                if (c->prec == WQL_FP64) k_fill<double><<<gq, 256, 0, c->stream>>>(9 * d.SP[s], (double *)d.DQ[s], 0.0);
                else k_fill<float><<<gq, 256, 0, c->stream>>>(9 * d.SP[s], (float *)d.DQ[s], 0.0f);
            }
        long FP = (long)d.nr * d.ns;
        k_fill<double><<<(unsigned)((4 * FP + 255) / 256), 256, 0, c->stream>>>(4 * FP, c->I[ib].DS, 0.0);
        k_fill<double><<<(unsigned)((FP + 255) / 256), 256, 0, c->stream>>>(FP, c->I[ib].DW, 0.0);
    }
    // volume kernels with A = 0 and a zero update step; boundary in rates-only mode
    if (what == 0) {
        // run the volume kernels only: boundary kernel skipped by using a copy of the
        // launch sequence without it
        Pair<double> Pd;
        Pair<float> Pf;
        if (c->prec == WQL_FP64) {
            Pd.b[0] = blk_view<double>(c, 0, c->cur);
            Pd.b[1] = blk_view<double>(c, 1, c->cur);
        } else {
            Pf.b[0] = blk_view<float>(c, 0, c->cur);
            Pf.b[1] = blk_view<float>(c, 1, c->cur);
        }
        TileList TL = tile_list(c);
        if (c->prec == WQL_FP64) {
            if (per_node(c)) {
                launch_op<double>(c, Pd, 0.0, 0.0, 3, c->stream);
            } else {
                if (TL.total > 0) launch_main<double>(c, Pd, TL, 0.0, 0.0, 0.0);
                launch_band<double>(c, Pd, 0.0, 0.0, c->stream);
                launch_pml<double>(c, Pd, 0.0, 0.0, c->stream);
            }
            launch_sources<double>(c, Pd, t, nullptr, 0.0, 0.0);
        } else {
            if (per_node(c)) {
                launch_op<float>(c, Pf, 0.f, 0.f, 3, c->stream);
            } else {
                if (TL.total > 0) launch_main<float>(c, Pf, TL, 0.f, 0.f, 0.f);
                launch_band<float>(c, Pf, 0.f, 0.f, c->stream);
                launch_pml<float>(c, Pf, 0.f, 0.f, c->stream);
            }
            launch_sources<float>(c, Pf, t, nullptr, 0.0, 0.0);
        }
    } else {
        if (c->prec == WQL_FP64) {
            gather_fq<double>(c);
            launch_stage<double>(c, c->cur, 0.0, 0.0, t, stage, 1, 0);
        } else {
            gather_fq<float>(c);
            launch_stage<float>(c, c->cur, 0.0, 0.0, t, stage, 1, 0);
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_launch();
}

int wql_stage(wql_ctx *c, int stage)
{
    if (!c || stage < 1 || stage > 5) return fail(WQL_E_ARG, "bad argument");
    if (stage != c->stage_next)
        return fail(WQL_E_STATE, "RK stage %d requested, stage %d of the step is next", stage, c->stage_next);
    const int s = stage - 1;
    if (c->prec == WQL_FP64) {
        gather_fq<double>(c);
        launch_stage<double>(c, c->cur, RK_A[s], RK_B[s] * c->dt, c->t + RK_C[s] * c->dt, stage, 0, stage == 5);
    } else {
        gather_fq<float>(c);
        launch_stage<float>(c, c->cur, RK_A[s], RK_B[s] * c->dt, c->t + RK_C[s] * c->dt, stage, 0, stage == 5);
    }
    c->cur = 1 - c->cur;
    if (stage == 5) {
        c->t = c->t + RK_C[5] * c->dt;
        c->stage_next = 1;
    } else {
        c->stage_next = stage + 1;
    }
    return check_launch();
}

int wql_step(wql_ctx *c, int nsteps)
{
    if (!c || nsteps < 0) return fail(WQL_E_ARG, "bad argument");
    if (c->stage_next != 1)
        return fail(WQL_E_STATE, "a step run by wql_stage is at stage %d; finish it first", c->stage_next);
    if (nsteps > 0) {
        if (c->prec == WQL_FP64) gather_fq<double>(c);
        else gather_fq<float>(c);
    }
    const bool graph = c->use_graph && !c->profile;
    if (graph && nsteps > 0) {
        // the device clock starts from the host clock; one graph per
        // ping-pong parity, captured on first use
        k_set_time<<<1, 64, 0, c->stream>>>(c->d_t, c->t);
        for (int k = 0; k < 2; k++) {
            if (c->gexec[k]) continue;
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            if (c->prec == WQL_FP64) enqueue_step_graph<double>(c, k);
            else enqueue_step_graph<float>(c, k);
            HIPCHK(hipStreamEndCapture(c->stream, &g));
            hipError_t e = hipGraphInstantiate(&c->gexec[k], g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) {
                c->gexec[k] = nullptr;
                return fail(WQL_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
            }
        }
    }
    for (int n = 0; n < nsteps; n++) {
        if (graph) HIPCHK(hipGraphLaunch(c->gexec[c->cur], c->stream));
        else if (c->prec == WQL_FP64) enqueue_step<double>(c, c->cur, c->t);
        else enqueue_step<float>(c, c->cur, c->t);
        c->cur = 1 - c->cur;  // five stages: odd number of buffer swaps
        c->t = c->t + RK_C[5] * c->dt;
        if (int rc = check_launch()) return rc;
    }
    return 0;
}

int wql_sync(wql_ctx *c)
{
    if (!c) return fail(WQL_E_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    int flag = 0;
    HIPCHK(hipMemcpy(&flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost));
    if (flag == 3) return fail(WQL_E_SOLVER, "friction solve produced a negative slip rate");
    if (flag == 4) return fail(WQL_E_NONFINITE, "non-finite value in the fault state");
    if (flag) return fail(WQL_E_SOLVER, "fault error code %d", flag);
    return 0;
}

int wql_set_tile(wql_ctx *c, int variant, int zchunk)
{
    if (!c || variant < 0 || variant >= NVARIANT || zchunk < 1) return fail(WQL_E_ARG, "bad tile variant");
    drop_graphs(c);
    // the anelastic response runs on 32 x 8 tiles: the fused or held march
    if (c->prob.anelastic) variant = VARIANT_FAM[variant] == FAM_HOLD ? 1 : 0;
    c->variant = variant;
    c->zchunk = zchunk;
    return 0;
}

int wql_get_tile(const wql_ctx *c, int *variant, int *zchunk)
{
    if (!c || !variant || !zchunk) return fail(WQL_E_ARG, "null argument");
    *variant = c->variant;
    *zchunk = c->zchunk;
    return 0;
}

int wql_set_tile_order(wql_ctx *c, int order, int group)
{
    if (!c || order < 0 || order > 3 || group < 1) return fail(WQL_E_ARG, "bad tile order");
    drop_graphs(c);
    c->order = order;
    c->group = group;
    return 0;
}

int wql_get_tile_order(const wql_ctx *c, int *order, int *group)
{
    if (!c || !order || !group) return fail(WQL_E_ARG, "null argument");
    *order = c->order;
    *group = c->group;
    return 0;
}

int wql_use_graph(wql_ctx *c, int on)
{
    if (!c) return fail(WQL_E_ARG, "null context");
    c->use_graph = on ? 1 : 0;
    if (!on) drop_graphs(c);
    return 0;
}

int wql_profile(wql_ctx *c, int on)
{
    if (!c) return fail(WQL_E_ARG, "null context");
    c->profile = on ? 1 : 0;
    c->vol_ms = 0.0;
    c->vol_launches = 0;
    return 0;
}

double wql_volume_kernel_ms(wql_ctx *c)
{
    if (!c || c->vol_launches == 0) return 0.0;
    return c->vol_ms / (double)c->vol_launches;
}

}  // extern "C"

#include "wql_serve.hip"
