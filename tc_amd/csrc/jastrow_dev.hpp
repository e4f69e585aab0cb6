// jastrow_dev.hpp -- Jastrow kernels u(G) and tc_2body(G) in FP64, usable on host
// and device.
//
// Restates Jastrow::uk / tc_2body (reference src/jastrow_functions.cpp:6-51).
// RPA part (uk_RPA / tc_2body_RPA, :53-97):
//   u(G)        = 4 pi A / (G^2 (F^2 G^2 + 1)),            G=0: -4 pi A F^2
//   tc_2body(G) = 4 pi/G^2 - 4 pi A/(1 + F^2 G^2)
//                 + (pi A^2/F^2 / G) [pi F^2 G^2 + (4 + 2 F^2 G^2) atan(F G/2)
//                                     - (4 + 4 F^2 G^2) atan(F G)],   G=0: -2 pi A^2 / F
// and with no RPA part (all |A| <= 1e-8): u = 0, tc_2body = 4 pi/G^2 (0 at G=0).
// A polynomial part (deg_poly_ > 0, jastrow_poly_dev.hpp) adds u_poly to u and
// turns tc_2body into tc_2body_POLY (no RPA part) or, for deg_poly_ > 1,
// tc_2body_RPAPOLY.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "jastrow_poly_dev.hpp"

namespace tcx {

struct JastrowPair {
    double A, F;   // RPA parameters of one spin pair
    int zero;      // Jastrow::is_A_zero_
    int dA = 0;    // 1: return the derivatives d/dA at fixed F instead (adjoint of the A-dependence)
    int pdeg = 0;  // deg_poly_ (0: no polynomial part)
    const double* poly = nullptr;  // the pair's polynomial block (jastrow_poly_dev.hpp), device or host memory
};

// kPoly = false leaves the polynomial part out at compile time (for kernels launched only
// when deg_poly_ == 0: no call, no stack).
template <bool kPoly = true>
__host__ __device__ __forceinline__ double jastrow_uk(const JastrowPair& j, double G2)
{
    constexpr double fourpi = 4.0 * 3.14159265358979323846;
    const double F2 = j.F * j.F;
    if (j.dA) return G2 < 1e-12 ? -fourpi * F2 : fourpi / (G2 * (G2 * F2 + 1.0));  // u is linear in A
    double v = 0.0;
    if (!j.zero) v = G2 < 1e-12 ? -fourpi * j.A * F2 : fourpi * j.A / (G2 * (G2 * F2 + 1.0));
    if constexpr (kPoly)
        if (j.pdeg > 0) v += jpoly::uk_poly(j.poly, G2);
    return v;
}

template <bool kPoly = true>
__host__ __device__ __forceinline__ double jastrow_tc2body(const JastrowPair& j, double G2)
{
    constexpr double pi = 3.14159265358979323846;
    constexpr double fourpi = 4.0 * pi;
    if (j.dA) {  // d/dA of the A and A^2 terms below (the 4 pi/G^2 Coulomb part is A-independent)
        const double F2 = j.F * j.F;
        const double dpaf = 2.0 * pi * j.A / F2;
        double d;
        if (G2 < 1e-12) {
            d = -2.0 * dpaf * j.F;
        } else {
            const double G1 = sqrt(G2);
            const double FG = j.F * G1, F2G2 = F2 * G2;
            d = -fourpi / (1.0 + F2G2) +
                (dpaf / G1) * (pi * F2G2 + (4.0 + 2.0 * F2G2) * atan(0.5 * FG) - (4.0 + 4.0 * F2G2) * atan(FG));
        }
        if constexpr (kPoly)
            if (!j.zero && j.pdeg > 1) d = jpoly::tc2body_rpapoly(j.poly, G2, d, 1);
        return d;
    }
    if (j.zero) {
        if constexpr (kPoly)
            if (j.pdeg > 0) return jpoly::tc2body_poly(j.poly, G2);
        return G2 > 1e-12 ? fourpi / G2 : 0.0;
    }
    const double F2 = j.F * j.F;
    const double paf = pi * j.A * j.A / F2;
    double rpa;
    if (G2 < 1e-12) {
        rpa = -2.0 * paf * j.F;
    } else {
        const double G1 = sqrt(G2);
        const double FG = j.F * G1, F2G2 = F2 * G2;
        rpa = fourpi / G2 - fourpi * j.A / (1.0 + F2G2) +
              (paf / G1) * (pi * F2G2 + (4.0 + 2.0 * F2G2) * atan(0.5 * FG) - (4.0 + 4.0 * F2G2) * atan(FG));
    }
    if constexpr (kPoly)
        if (j.pdeg > 1) return jpoly::tc2body_rpapoly(j.poly, G2, rpa, 0);
    return rpa;
}

}  // namespace tcx
