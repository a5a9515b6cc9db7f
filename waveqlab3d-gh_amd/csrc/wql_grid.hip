// wql_grid.hip -- block grids on the device (part of libwql; included by
// wql_api.hip): the uniform grid of cartesian_grid_3d (grid.f90:213-240) and
// the transfinite-interpolation grid of Curve_Grid_3D3 (grid.f90:471-1907)
// for planar or profiled q faces ('analytical_test_problem', 775-831;
// 'analytical_tpv28', 'analytical_tpv10', 'analytical_tpv36', 625-773),
// with optional caller-supplied fault-face and free-surface offset
// fields (the surfaces the reference reads from files through the same
// interpolation, 832-1430).
//
// Face interpolation Gen_Surf_Mesh (2418-2455) and volume interpolation
// Gen3_Curve_Mesh (2290-2416) run one node per thread in FP64 without FMA
// contraction, so that every coordinate is the same rounded expression as in
// the host formulation.
#pragma clang fp contract(off)

namespace wql {

// 2-D transfinite interpolation on an (m1 x m2) face, node (i, j) -> i + m1 j;
// L, R are the edges at i = 0 and i = m1-1 (indexed by j), B, T the edges at
// j = 0 and j = m2-1 (indexed by i)
__device__ __forceinline__ double surf_tfi(int m1, int m2, int i, int j, const double *L, const double *R,
                                           const double *B, const double *T)
{
    const double q = (double)i * (1.0 / (double)(m1 - 1));
    const double r = (double)j * (1.0 / (double)(m2 - 1));
    return (1.0 - q) * L[j] + q * R[j] + (1.0 - r) * B[i] + r * T[i] - (1.0 - q) * (1.0 - r) * L[0] -
           q * (1.0 - r) * R[0] - r * (1.0 - q) * T[0] - (r * q) * T[m1 - 1];
}

// an edge of m nodes from lo to hi
__device__ __forceinline__ double edge_at(double lo, double hi, int m, int i)
{
    const double h = (hi - lo) / (double)(m - 1);
    return lo + (double)i * h;
}

struct GridArgs {
    int nq, nr, ns;
    double a[3], b[3];
    double lc, rc;
    int profile;              // 0 planar q faces, 1 'analytical_test_problem', 2 'analytical_tpv28',
                              // 3 'analytical_tpv10', 4 'analytical_tpv36'
    const double *fault_x;    // [ns][nr] x offsets of one q face (or null)
    int fault_side;           // 0: q = 0 face, 1: q = nq-1 face
    const double *top_y;      // [ns][nq] y offsets of the r = 0 face (or null)
    // all six faces given (wql_grid_faces): per component x, y, z the q = 0,
    // q = nq-1 faces [ns][nr], the r = 0, r = nr-1 faces [ns][nq], the s = 0,
    // s = ns-1 faces [nr][nq] (or null: the faces of the kernels above)
    const double *faces;
    // work arrays (device)
    double *qf;               // [4][nr ns]: Yleft (= Yright), Zleft (= Zright), Xleft, Xright; node j + nr k
    double *sf;               // [2][2][nq nr]: X, Y of the front (k = 0) and back (k = ns-1) faces; node i + nq j
    double *rf;               // [2][2][nq ns]: X, Z of the top (j = 0) and bottom (j = nr-1) faces; node i + nq k
    double *X;                // [3][ns][nr][nq]
};

// q faces (grid.f90:779-831)
__global__ void k_grid_qfaces(GridArgs g)
{
    const int my = g.nr, mz = g.ns;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)my * mz) return;
    const int j = (int)(t % my), k = (int)(t / my);
    const double ay = g.a[1], by = g.b[1], az = g.a[2], bz = g.b[2];
    const double hy = (by - ay) / (double)(my - 1), hz = (bz - az) / (double)(mz - 1);
    const double yj = ay + (double)j * hy, y0 = ay, yl = ay + (double)(my - 1) * hy;
    const double zk = az + (double)k * hz, z0 = az, zl = az + (double)(mz - 1) * hz;
    // Y: L/R edges (along k) constant ay / by, B/T edges (along j) the y line
    const double q1 = (double)j * (1.0 / (double)(my - 1)), r1 = (double)k * (1.0 / (double)(mz - 1));
    const double Y = (1.0 - q1) * ay + q1 * by + (1.0 - r1) * yj + r1 * yj - (1.0 - q1) * (1.0 - r1) * ay -
                     q1 * (1.0 - r1) * by - r1 * (1.0 - q1) * y0 - (r1 * q1) * yl;
    // Z: L/R edges the z line, B/T edges constant az / bz
    const double Z = (1.0 - q1) * zk + q1 * zk + (1.0 - r1) * az + r1 * bz - (1.0 - q1) * (1.0 - r1) * z0 -
                     q1 * (1.0 - r1) * z0 - r1 * (1.0 - q1) * bz - (r1 * q1) * bz;
    double XL, XR;
    if (g.profile >= 2) {
        // the formula profiles set the face nodes' y and z directly
        // (grid.f90:636-639) and add rc f, lc f to the planar faces; the
        // dipping ones evaluate f when either weight is positive
        const double PI = 3.141592653589793;
        const double y = yj, z = zk;
        double f = 0.0;
        if (g.profile == 2) {  // TPV28 hills (627-656)
            const double r1 = sqrt((z + 10.5) * (z + 10.5) + (y - 7.5) * (y - 7.5));
            const double r2 = sqrt((z - 10.5) * (z - 10.5) + (y - 7.5) * (y - 7.5));
            if (r1 <= 3.0) f = -0.3 * (1.0 + cos(PI * r1 / 3.0));
            if (r2 <= 3.0) f = -0.3 * (1.0 + cos(PI * r2 / 3.0));
        } else if (g.rc > 0.0 || g.lc > 0.0) {
            if (g.profile == 3) {  // TPV10, 60 degrees (658-713)
                const double ang = 60.0 / 180.0 * PI;
                if (y <= 15.0) f = (1.0 / tan(ang)) * (y - 0.5 * (ay + by));
                else f = (1.0 / tan(ang)) * (15.0 - 0.5 * (ay + by)) + 0.5 * atan(4.0 * (y - 15.0)) * exp(-5.0 * (y - 15.0));
            } else {  // TPV36, 15 degrees (715-773)
                const double ang = 15.0 / 180.0 * PI;
                if (y <= 8.0) f = 1.0 / tan(ang) * y - 15.0;
                else f = 1.0 / tan(ang) * 8.0 - 15.0 + 0.25 * atan(4.0 * (y - 8.0)) * exp(-0.0 * (y - 8.0));
            }
        }
        const long FP = (long)my * mz;
        g.qf[t] = y;
        g.qf[FP + t] = z;
        g.qf[2 * FP + t] = g.a[0] + g.rc * f;
        g.qf[3 * FP + t] = g.b[0] + g.lc * f;
        return;
    }
    if (g.profile == 1) {
        const double PI = 3.141592653589793;
        const double cy = 0.5 * (ay + by), cz = 0.5 * (az + bz);
        XL = g.a[0] + 0.1 * g.rc * exp(-((1.0 * Y - cy) * (1.0 * Y - cy) + 1.0 * ((Z - cz) * (Z - cz))) / 0.025) +
             0.1 * g.rc * sin(2.0 / (by - ay) * PI * Y) * sin(2.0 / (bz - az) * PI * Z);
        XR = g.b[0] + 0.1 * g.lc * exp(-(1.0 * ((Y - cy) * (Y - cy)) + 1.0 * ((Z - cz) * (Z - cz))) / 0.025) +
             0.1 * g.lc * sin(2.0 / (by - ay) * PI * Y) * sin(2.0 / (bz - az) * PI * Z);
    } else {
        XL = g.a[0];
        XR = g.b[0];
        if (g.fault_x) {
            const double d = g.fault_x[j + (long)my * k];
            if (g.fault_side == 0) XL = XL + d;
            else XR = XR + d;
        }
    }
    const long FP = (long)my * mz;
    g.qf[t] = Y;
    g.qf[FP + t] = Z;
    g.qf[2 * FP + t] = XL;
    g.qf[3 * FP + t] = XR;
}

// s faces (front k = 0, back k = ns-1; grid.f90:930-1041)
__global__ void k_grid_sfaces(GridArgs g)
{
    const int mx = g.nq, my = g.nr, mz = g.ns;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2L * mx * my) return;
    const int side = (int)(t / ((long)mx * my));
    const long r = t - (long)side * mx * my;
    const int i = (int)(r % mx), j = (int)(r / mx);
    const int k = side == 0 ? 0 : mz - 1;
    const long FP = (long)my * mz;
    const double *Yq = g.qf, *Xl = g.qf + 2 * FP, *Xr = g.qf + 3 * FP;
    // edges of this face: L/R along j (q faces at column k), B/T along i
    auto XLj = [&](int jj) { return Xl[jj + (long)my * k]; };
    auto XRj = [&](int jj) { return Xr[jj + (long)my * k]; };
    auto YLj = [&](int jj) { return Yq[jj + (long)my * k]; };
    const double qq = (double)i * (1.0 / (double)(mx - 1)), rr = (double)j * (1.0 / (double)(my - 1));
    const double XB = edge_at(XLj(0), XRj(0), mx, i), XT = edge_at(XLj(my - 1), XRj(my - 1), mx, i);
    const double XT0 = edge_at(XLj(my - 1), XRj(my - 1), mx, 0), XTl = edge_at(XLj(my - 1), XRj(my - 1), mx, mx - 1);
    const double X = (1.0 - qq) * XLj(j) + qq * XRj(j) + (1.0 - rr) * XB + rr * XT - (1.0 - qq) * (1.0 - rr) * XLj(0) -
                     qq * (1.0 - rr) * XRj(0) - rr * (1.0 - qq) * XT0 - (rr * qq) * XTl;
    const double YB = edge_at(YLj(0), YLj(0), mx, i), YT = edge_at(YLj(my - 1), YLj(my - 1), mx, i);
    const double YT0 = edge_at(YLj(my - 1), YLj(my - 1), mx, 0), YTl = edge_at(YLj(my - 1), YLj(my - 1), mx, mx - 1);
    const double Y = (1.0 - qq) * YLj(j) + qq * YLj(j) + (1.0 - rr) * YB + rr * YT - (1.0 - qq) * (1.0 - rr) * YLj(0) -
                     qq * (1.0 - rr) * YLj(0) - rr * (1.0 - qq) * YT0 - (rr * qq) * YTl;
    g.sf[(long)side * 2 * mx * my + r] = X;
    g.sf[(long)side * 2 * mx * my + (long)mx * my + r] = Y;
}

// r faces (top j = 0, bottom j = nr-1; grid.f90:1044-1158), plus the
// free-surface offsets on the top face
__global__ void k_grid_rfaces(GridArgs g)
{
    const int mx = g.nq, my = g.nr, mz = g.ns;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2L * mx * mz) return;
    const int side = (int)(t / ((long)mx * mz));
    const long r = t - (long)side * mx * mz;
    const int i = (int)(r % mx), k = (int)(r / mx);
    const int j = side == 0 ? 0 : my - 1;
    const long FP = (long)my * mz;
    const double *Zq = g.qf + FP, *Xl = g.qf + 2 * FP, *Xr = g.qf + 3 * FP;
    auto XLk = [&](int kk) { return Xl[j + (long)my * kk]; };
    auto XRk = [&](int kk) { return Xr[j + (long)my * kk]; };
    auto ZLk = [&](int kk) { return Zq[j + (long)my * kk]; };
    const double qq = (double)i * (1.0 / (double)(mx - 1)), rr = (double)k * (1.0 / (double)(mz - 1));
    const double XB = edge_at(XLk(0), XRk(0), mx, i), XT = edge_at(XLk(mz - 1), XRk(mz - 1), mx, i);
    const double XT0 = edge_at(XLk(mz - 1), XRk(mz - 1), mx, 0), XTl = edge_at(XLk(mz - 1), XRk(mz - 1), mx, mx - 1);
    const double X = (1.0 - qq) * XLk(k) + qq * XRk(k) + (1.0 - rr) * XB + rr * XT - (1.0 - qq) * (1.0 - rr) * XLk(0) -
                     qq * (1.0 - rr) * XRk(0) - rr * (1.0 - qq) * XT0 - (rr * qq) * XTl;
    // Z: B/T edges are the front / back faces' constant z
    const double az = g.a[2], bz = g.b[2];
    const double Z = (1.0 - qq) * ZLk(k) + qq * ZLk(k) + (1.0 - rr) * az + rr * bz - (1.0 - qq) * (1.0 - rr) * ZLk(0) -
                     qq * (1.0 - rr) * ZLk(0) - rr * (1.0 - qq) * bz - (rr * qq) * bz;
    g.rf[(long)side * 2 * mx * mz + r] = X;
    g.rf[(long)side * 2 * mx * mz + (long)mx * mz + r] = Z;
}

// volume interpolation, one component (Gen3_Curve_Mesh); the face values of
// component c at node (i, j, k)
__global__ void k_grid_volume(GridArgs g)
{
    const int mx = g.nq, my = g.nr, mz = g.ns;
    const long N = (long)mx * my * mz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * N) return;
    const int c = (int)(t / N);
    const long p = t - (long)c * N;
    const int i = (int)(p % mx), j = (int)((p / mx) % my), k = (int)(p / ((long)mx * my));
    const long FQ = (long)my * mz, FS = (long)mx * my, FR = (long)mx * mz;
    const double ay = g.a[1], by = g.b[1], az = g.a[2], bz = g.b[2];
    // the given faces (wql_grid_faces) of component c
    const double *F = g.faces ? g.faces + (long)c * 2 * (FQ + FR + FS) : nullptr;
    // face arrays of component c
    auto L = [&](int jj, int kk) -> double {  // q = 0 face
        const long o = jj + (long)my * kk;
        if (F) return F[o];
        return c == 0 ? g.qf[2 * FQ + o] : c == 1 ? g.qf[o] : g.qf[FQ + o];
    };
    auto R = [&](int jj, int kk) -> double {  // q = nq-1 face
        const long o = jj + (long)my * kk;
        if (F) return F[FQ + o];
        return c == 0 ? g.qf[3 * FQ + o] : c == 1 ? g.qf[o] : g.qf[FQ + o];
    };
    auto Tp = [&](int ii, int kk) -> double {  // r = 0 face
        const long o = ii + (long)mx * kk;
        if (F) return F[2 * FQ + o];
        if (c == 0) return g.rf[o];
        if (c == 2) return g.rf[FR + o];
        return g.top_y ? ay + g.top_y[o] : ay;
    };
    auto Bt = [&](int ii, int kk) -> double {  // r = nr-1 face
        const long o = ii + (long)mx * kk;
        if (F) return F[2 * FQ + FR + o];
        if (c == 0) return g.rf[2 * FR + o];
        if (c == 2) return g.rf[3 * FR + o];
        return by;
    };
    auto Fr = [&](int ii, int jj) -> double {  // s = 0 face
        const long o = ii + (long)mx * jj;
        if (F) return F[2 * FQ + 2 * FR + o];
        return c == 0 ? g.sf[o] : c == 1 ? g.sf[FS + o] : az;
    };
    auto Bk = [&](int ii, int jj) -> double {  // s = ns-1 face
        const long o = ii + (long)mx * jj;
        if (F) return F[2 * FQ + 2 * FR + FS + o];
        return c == 0 ? g.sf[2 * FS + o] : c == 1 ? g.sf[3 * FS + o] : bz;
    };
    const double q = (double)i * (1.0 / (double)(mx - 1));
    const double r = (double)j * (1.0 / (double)(my - 1));
    const double s = (double)k * (1.0 / (double)(mz - 1));
    const double U = (1.0 - q) * L(j, k) + q * R(j, k);
    const double V = (1.0 - r) * Tp(i, k) + r * Bt(i, k);
    const double W = (1.0 - s) * Fr(i, j) + s * Bk(i, j);
    const double UW = (1.0 - q) * ((1.0 - r) * L(0, k) + r * L(my - 1, k)) + q * ((1.0 - r) * R(0, k) + r * R(my - 1, k));
    const double UV = (1.0 - r) * ((1.0 - s) * Tp(i, 0) + s * Tp(i, mz - 1)) + r * ((1.0 - s) * Bt(i, 0) + s * Bt(i, mz - 1));
    const double VW = (1.0 - s) * ((1.0 - q) * Fr(0, j) + q * Fr(mx - 1, j)) + s * ((1.0 - q) * Bk(0, j) + q * Bk(mx - 1, j));
    const double UVW1 = (1.0 - q) * ((1.0 - r) * ((1.0 - s) * Tp(0, 0) + s * Tp(0, mz - 1)) +
                                     r * ((1.0 - s) * Bt(0, 0) + s * Bt(0, mz - 1)));
    const double UVW2 = q * ((1.0 - r) * ((1.0 - s) * Tp(mx - 1, 0) + s * Tp(mx - 1, mz - 1)) +
                             r * ((1.0 - s) * Bt(mx - 1, 0) + s * Bt(mx - 1, mz - 1)));
    g.X[t] = U + V + W - UV - UW - VW + UVW1 + UVW2;
}

// uniform grid (cartesian_grid_3d)
__global__ void k_grid_cartesian(int nq, int nr, int ns, double a0, double a1, double a2, double b0, double b1,
                                 double b2, double *X)
{
    const long N = (long)nq * nr * ns;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * N) return;
    const int c = (int)(t / N);
    const long p = t - (long)c * N;
    const int i = (int)(p % nq), j = (int)((p / nq) % nr), k = (int)(p / ((long)nq * nr));
    const int idx = c == 0 ? i : c == 1 ? j : k;
    const int m = c == 0 ? nq : c == 1 ? nr : ns;
    const double lo = c == 0 ? a0 : c == 1 ? a1 : a2, hi = c == 0 ? b0 : c == 1 ? b1 : b2;
    const double h = (hi - lo) / (double)(m - 1);
    X[t] = lo + (double)idx * h;
}

}  // namespace wql

#pragma clang fp contract(fast)
