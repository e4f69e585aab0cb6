// jastrow_dev.hpp -- RPA Jastrow kernels in FP64, usable on host and device.
//
// Restates Jastrow::uk / uk_RPA / tc_2body / tc_2body_RPA
// (reference src/jastrow_functions.cpp:6-97) for the RPA-only case:
//   u(G)        = 4 pi A / (G^2 (F^2 G^2 + 1)),            G=0: -4 pi A F^2
//   tc_2body(G) = 4 pi/G^2 - 4 pi A/(1 + F^2 G^2)
//                 + (pi A^2/F^2 / G) [pi F^2 G^2 + (4 + 2 F^2 G^2) atan(F G/2)
//                                     - (4 + 4 F^2 G^2) atan(F G)],   G=0: -2 pi A^2 / F
// and with no Jastrow (all |A| <= 1e-8): u = 0, tc_2body = 4 pi/G^2 (0 at G=0).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace tcx {

struct JastrowRPA {
    double A, F;   // one spin pair
    int zero;      // Jastrow::is_A_zero_
    int dA = 0;    // 1: return the derivatives d/dA at fixed F instead (adjoint of the A-dependence)
};

__host__ __device__ __forceinline__ double jastrow_uk(const JastrowRPA& j, double G2)
{
    constexpr double fourpi = 4.0 * 3.14159265358979323846;
    const double F2 = j.F * j.F;
    if (j.dA) return G2 < 1e-12 ? -fourpi * F2 : fourpi / (G2 * (G2 * F2 + 1.0));  // u is linear in A
    if (j.zero) return 0.0;
    if (G2 < 1e-12) return -fourpi * j.A * F2;
    return fourpi * j.A / (G2 * (G2 * F2 + 1.0));
}

__host__ __device__ __forceinline__ double jastrow_tc2body(const JastrowRPA& j, double G2)
{
    constexpr double pi = 3.14159265358979323846;
    constexpr double fourpi = 4.0 * pi;
    if (j.dA) {  // d/dA of the A and A^2 terms below (the 4 pi/G^2 Coulomb part is A-independent)
        const double F2 = j.F * j.F;
        const double dpaf = 2.0 * pi * j.A / F2;
        if (G2 < 1e-12) return -2.0 * dpaf * j.F;
        const double G1 = sqrt(G2);
        const double FG = j.F * G1, F2G2 = F2 * G2;
        return -fourpi / (1.0 + F2G2) +
               (dpaf / G1) * (pi * F2G2 + (4.0 + 2.0 * F2G2) * atan(0.5 * FG) - (4.0 + 4.0 * F2G2) * atan(FG));
    }
    if (j.zero) return G2 > 1e-12 ? fourpi / G2 : 0.0;
    const double F2 = j.F * j.F;
    const double paf = pi * j.A * j.A / F2;
    if (G2 < 1e-12) return -2.0 * paf * j.F;
    const double G1 = sqrt(G2);
    const double FG = j.F * G1, F2G2 = F2 * G2;
    return fourpi / G2 - fourpi * j.A / (1.0 + F2G2) +
           (paf / G1) * (pi * F2G2 + (4.0 + 2.0 * F2G2) * atan(0.5 * FG) - (4.0 + 4.0 * F2G2) * atan(FG));
}

}  // namespace tcx
