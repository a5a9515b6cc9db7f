// wql_models.h -- the layered media the reference keys by problem name
// (part of libwql; host and device).
//
// init_material's hard-coded models (material.f90:189-690): the P- and
// S-wave speeds and the density at a node, chosen by its x coordinate
// (LOH1, LOH1_Harmonic, TPV33, OKLAHOMA) or its depth y (TPV31, TPV32); the
// node's lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2.  The fault routine
// Interface_RHS takes the shear modulus of TPV31 and TPV32 from the same
// depth profiles (Interface_Condition.f90:133-254).  Each chain of
// comparisons is the reference's, in its order; a coordinate no branch
// covers (NaN) leaves the outputs unset and returns 0.
#pragma once

namespace wql {

enum { WQL_MEDIUM_NONE = 0, WQL_MEDIUM_LOH1, WQL_MEDIUM_LOH1_HARMONIC, WQL_MEDIUM_TPV31, WQL_MEDIUM_TPV32,
       WQL_MEDIUM_TPV33, WQL_MEDIUM_OKLAHOMA };

__host__ __device__ inline int layered_medium(int model, double x, double y, double &vp, double &vs, double &rho)
{
#define WQL_SET(a, b, c) do { vp = (a); vs = (b); rho = (c); return 1; } while (0)
    switch (model) {
    case WQL_MEDIUM_LOH1:  // 189-233
        if (x < 0.9999) WQL_SET(4.0, 2.0, 2.6);
        else if (x >= 0.9999 && x <= 1.0001) WQL_SET(5.0, 2.732, 2.65);
        else if (x > 1.0001) WQL_SET(6.0, 3.464, 2.7);
        return 0;
    case WQL_MEDIUM_LOH1_HARMONIC:  // 234-276: harmonic means on the interface plane
        if (x < 0.9999) WQL_SET(4.0, 2.0, 2.6);
        else if (x >= 0.9999 && x <= 1.0001)
            WQL_SET(2.0 * 4.0 * 6.0 / (4.0 + 6.0), 2.0 * 2.0 * 3.464 / (2.0 + 3.464), 2.0 * 2.7 * 2.6 / (2.7 + 2.6));
        else if (x > 1.0001) WQL_SET(6.0, 3.464, 2.7);
        return 0;
    case WQL_MEDIUM_TPV31:  // 277-361
        if (y < 2.3999) WQL_SET(4.05, 2.25, 2.58);
        else if (y >= 2.3999 && y <= 2.4001) WQL_SET(4.25, 2.4, 2.59);
        else if (y > 2.4001 && y < 4.9999)
            WQL_SET(4.45 + 0.75 * (y - 2.4) / 2.6, 2.55 + 0.5 * (y - 2.4) / 2.6, 2.6 + 0.02 * (y - 2.4) / 2.6);
        else if (y >= 4.9999 && y <= 5.0001) WQL_SET(5.475, 3.25, 2.67);
        else if (y > 5.0001 && y < 9.9999) WQL_SET(5.75, 3.45, 2.72);
        else if (y >= 9.9999 && y <= 10.0001) WQL_SET(6.125, 3.625, 2.76);
        else if (y > 10.0001) WQL_SET(6.5, 3.8, 3.0);
        return 0;
    case WQL_MEDIUM_TPV32:  // 362-476
        if (y <= 0.5) WQL_SET(2.2 + 0.8 * y / 0.5, 1.05 + 0.35 * y / 0.5, 2.2 + 0.25 * y / 0.5);
        else if (y > 0.5 && y <= 1.0)
            WQL_SET(3.0 + 0.6 * (y - 0.5) / 0.5, 1.4 + 0.55 * (y - 0.5) / 0.5, 2.45 + 0.1 * (y - 0.5) / 0.5);
        else if (y > 1.0 && y <= 1.6)
            WQL_SET(3.6 + 0.8 * (y - 1.0) / 0.6, 1.95 + 0.55 * (y - 1.0) / 0.6, 2.55 + 0.05 * (y - 1.0) / 0.6);
        else if (y > 1.6 && y <= 2.4) WQL_SET(4.4 + 0.4 * (y - 1.6) / 0.8, 2.5 + 0.3 * (y - 1.6) / 0.8, 2.6);
        else if (y > 2.4 && y <= 3.6)
            WQL_SET(4.8 + 0.45 * (y - 2.4) / 1.2, 2.8 + 0.3 * (y - 2.4) / 1.2, 2.6 + 0.02 * (y - 2.4) / 1.2);
        else if (y > 3.6 && y <= 5.0)
            WQL_SET(5.25 + 0.25 * (y - 3.6) / 1.4, 3.1 + 0.15 * (y - 3.6) / 1.4, 2.62 + 0.03 * (y - 3.6) / 1.4);
        else if (y > 5.0 && y <= 9.0)
            WQL_SET(5.5 + 0.25 * (y - 5.0) / 4.0, 3.25 + 0.2 * (y - 5.0) / 4.0, 2.65 + 0.07 * (y - 5.0) / 4.0);
        else if (y > 9.0 && y <= 11.0)
            WQL_SET(5.75 + 0.35 * (y - 9.0) / 2.0, 3.45 + 0.25 * (y - 9.0) / 2.0, 2.72 + 0.03 * (y - 9.0) / 2.0);
        else if (y > 11.0 && y <= 15.0)
            WQL_SET(6.1 + 0.2 * (y - 11.0) / 4.0, 3.6 + 0.1 * (y - 11.0) / 4.0, 2.75 + 0.15 * (y - 11.0) / 4.0);
        else if (y > 15.0) WQL_SET(6.3, 3.7, 2.9);
        return 0;
    case WQL_MEDIUM_TPV33:  // 477-522: low-velocity zone |x| < 0.8
        if (x <= -0.8) WQL_SET(5.626, 3.248, 2.67);
        else if (x >= 0.8) WQL_SET(6.0, 3.464, 2.67);
        else if (x > -0.8 && x < 0.8) WQL_SET(3.75, 2.165, 2.67);
        return 0;
    case WQL_MEDIUM_OKLAHOMA:  // 523-690
        if (x < 1.4999) WQL_SET(2.77, 1.49, 2.169);
        else if (x >= 1.4999 && x <= 1.5001) WQL_SET(3.74, 2.01, 2.418);
        else if (x > 1.5001 && x < 2.2999) WQL_SET(5.76, 3.1, 2.667);
        else if (x >= 2.2999 && x <= 2.3001) WQL_SET(5.755, 3.26, 2.666);
        else if (x > 2.3001 && x < 5.2999) WQL_SET(5.75, 3.423, 2.665);
        else if (x >= 5.2999 && x <= 5.3001) WQL_SET(5.96, 3.54, 2.711);
        else if (x > 5.3001 && x < 8.2999) WQL_SET(6.18, 3.65, 2.756);
        else if (x >= 8.2999 && x <= 8.3001) WQL_SET(6.20, 3.63, 2.762);
        else if (x > 8.3001 && x < 11.2999) WQL_SET(6.23, 3.60, 2.767);
        else if (x >= 11.2999 && x <= 11.3001) WQL_SET(6.26, 3.64, 2.776);
        else if (x > 11.3001 && x < 20.2999) WQL_SET(6.30, 3.67, 2.784);
        else if (x >= 20.2999 && x <= 20.3001) WQL_SET(6.54, 3.80, 2.848);
        else if (x > 20.3001 && x < 39.9999) WQL_SET(6.80, 3.93, 2.911);
        else if (x >= 39.9999 && x <= 40.0001) WQL_SET(7.39, 4.25, 3.119);
        else if (x > 40.0001) WQL_SET(8.10, 4.62, 3.326);
        return 0;
    default:
        return 0;
    }
#undef WQL_SET
}

}  // namespace wql
