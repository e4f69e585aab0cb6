// jastrow_poly_dev.hpp -- polynomial and RPA+polynomial Jastrow kernels, usable on
// host and device.
//
// u_poly(r) = -sum_i c_i r^i (1 - r/L)^C for r < L, expanded into
// sum_i c_int[i] r^i (Jastrow::c_poly_internal_).  Its Fourier transforms
// (reference src/jastrow_functions.cpp:99-413):
//   J[n]  = 4 pi int_0^L r^(n+1) sin(G r)/(G r) dr                 (calculate_J_integ)
//   J2[n] = (4 pi/G) int_0^L r^n sin(G r) exp(-r/F) dr              (calculate_J2_integ)
//   J3    = (4 pi/G) int_0^(L/F) sin(F G s) [exp(-s) + (exp(-s)-1)/s] ds  (J3_integ)
//   u_poly(G)          = sum_i c_int[i] J[i+1]
//   tc_2body_POLY(G)   = 4 pi/G^2 - G^2 u_poly(G) - sum_{i,j>=1} i j c_i c_j J[i+j-1]
//   tc_2body_RPAPOLY(G) = tc_2body_RPA(G) - G^2 u_poly(G) - (same double sum)
//                        + sum_{i>=2} 2 i c_i A (J[i-2] - J2[i-1]/F - J2[i-2]) - 2 c_1 A J3
// The reference sums the small-G series of J (G^2 L^2 < 1000, 75 terms, whose
// terms cancel by up to 14 digits) and the double sum in 50-digit decimal
// arithmetic; here both use double-double (~32 digits), which yields the same
// doubles.  Everything else follows the reference's double arithmetic in its
// order, with FMA contraction off, so that rounding matches a plain C build.
//
// Each spin pair's parameters live in one block of doubles (layout below),
// built on the host from the reference's set-up routines
// (src/jastrow_initialize_polynomial.cpp:6-354).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

namespace tcx {
namespace jpoly {

enum : int {
    kDeg = 0,      // deg_poly_
    kC = 1,        // C_damp_poly_
    kL = 2,        // L_poly_
    kA = 3,        // A_long_ (RPA+polynomial only)
    kF = 4,        // F_long_
    kFinv = 5,     // 1/F
    kExpLF = 6,    // exp(-L/F)
    kNr = 7,       // num_rmesh_J3_integ_ (0 without the RPA+polynomial set-up)
    kCoef = 16,    // c_poly_internal_[0 .. deg+C)
    kMaxCoef = 19,
    kPowL = kCoef + kMaxCoef,  // FourPI_power_L_[0 .. 2 deg + 6)
    kMaxPowL = 38,
    kM = kPowL + kMaxPowL,     // M_integ_[0 .. 30)
    kMaxM = 30,
    kMesh = kM + kMaxM,        // rmesh_J3_integ_[nr], then integrand_stored_J3_integ_[nr]
};
constexpr int kMaxDeg = 16;                // Jastrow::max_deg_poly_
constexpr double kThresholdTaylor = 1000.0;  // threshold_Taylor_
constexpr int kMaxDegTaylor = 75;            // max_deg_Taylor_
constexpr double kThresholdTaylor2 = 1.5;    // threshold_Taylor2_
constexpr int kMaxDegTaylor2 = 30;           // max_deg_Taylor2_
constexpr int kMaxJ = 2 * (kMaxDeg + 3);

inline size_t block_doubles(int nr) { return (size_t)kMesh + 2 * (size_t)nr; }

// ------------------------------------------------------------------ double-double
struct dd {
    double hi, lo;
};

__host__ __device__ inline dd two_sum(double a, double b)
{
#pragma clang fp contract(off)
    const double s = a + b;
    const double bb = s - a;
    return dd{s, (a - (s - bb)) + (b - bb)};
}

__host__ __device__ inline dd quick_two_sum(double a, double b)
{
#pragma clang fp contract(off)
    const double s = a + b;
    return dd{s, b - (s - a)};
}

__host__ __device__ inline dd dd_add(dd a, dd b)
{
#pragma clang fp contract(off)
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}

__host__ __device__ inline dd dd_add_d(dd a, double b)
{
#pragma clang fp contract(off)
    dd s = two_sum(a.hi, b);
    s.lo += a.lo;
    return quick_two_sum(s.hi, s.lo);
}

__host__ __device__ inline dd dd_mul_d(dd a, double b)
{
#pragma clang fp contract(off)
    const double p = a.hi * b;
    const double e = fma(a.hi, b, -p);
    return quick_two_sum(p, fma(a.lo, b, e));
}

__host__ __device__ inline dd dd_div_d(dd a, double b)
{
#pragma clang fp contract(off)
    const double q1 = a.hi / b;
    const double p = q1 * b;
    const double pe = fma(q1, b, -p);
    dd s = two_sum(a.hi, -p);
    s.lo -= pe;
    s.lo += a.lo;
    const double q2 = (s.hi + s.lo) / b;
    return quick_two_sum(q1, q2);
}

__host__ __device__ inline dd dd_recip(double m)
{
#pragma clang fp contract(off)
    const double h = 1.0 / m;
    const double r = fma(-h, m, 1.0);  // 1 - h m, exact
    return quick_two_sum(h, r / m);
}

// ------------------------------------------------------------------ integrals
// J[n] for n in [n0, nJ), n0 = 0 or 1 (calculate_J_integ, jastrow_functions.cpp:99-186)
__host__ __device__ inline void J_integ(const double* blk, double G2, double* J, int nJ, int n0)
{
#pragma clang fp contract(off)
    const double L = blk[kL];
    const double G2L2 = G2 * L * L;
    if (G2L2 < kThresholdTaylor) {
        for (int n = n0; n < nJ; ++n) {
            dd t{0.0, 0.0};
            for (int j = 0; j < kMaxDegTaylor; ++j) {
                const int dg = 2 * (kMaxDegTaylor - j);
                t = dd_div_d(dd_mul_d(dd{-t.hi, -t.lo}, G2L2), (double)((dg + 1) * dg));
                t = dd_add(t, dd_recip((double)(n + dg)));
            }
            J[n] = t.hi;
        }
    } else {
        const double GL = sqrt(G2L2);
        const double cs = cos(GL), sn = sin(GL);
        for (int n = n0; n < nJ; ++n) {
            double d = -1.0 / G2L2;
            int m = n;
            double coeff = d;
            for (int j = 0; j < n / 2; ++j) {
                d *= -m * (m - 1) / G2L2;
                m -= 2;
                coeff += d;
            }
            double v = coeff * cs;
            if (n % 2 == 0) v -= d;
            if (n > 0) {
                d = n / (G2L2 * GL);
                m = n - 1;
                coeff = d;
                for (int j = 0; j < (n - 1) / 2; ++j) {
                    d *= -m * (m - 1) / G2L2;
                    m -= 2;
                    coeff += d;
                }
                v += coeff * sn;
            }
            J[n] = v;
        }
    }
    for (int n = n0; n < nJ; ++n) J[n] *= blk[kPowL + n + 2];
}

// J2[0 .. nJ2) (calculate_J2_integ, jastrow_functions.cpp:333-382)
__host__ __device__ inline void J2_integ(const double* blk, double G1, double* J2, int nJ2)
{
#pragma clang fp contract(off)
    constexpr double fourpi = 4.0 * 3.14159265358979323846;
    const double L = blk[kL];
    const double G2L2 = (G1 * L) * (G1 * L);
    if (G2L2 < kThresholdTaylor2) {
        for (int n = 0; n < nJ2; ++n) {
            double acc = 0.0, coeff = 1.0;
            for (int j = 0; j < kMaxDegTaylor2 / 2; ++j) {
                const int idx = n + 2 * j + 1;  // M_integ_ beyond index 29 is zero
                acc += coeff * (idx < kMaxM ? blk[kM + idx] : 0.0);
                coeff *= (-G2L2) / (double)((2 * j + 3) * (2 * j + 2));
            }
            J2[n] = acc * blk[kPowL + n + 1];
        }
    } else {
        const double a = blk[kFinv], b = -G1;  // ctmp1 = 1/(1/F - iG)
        const double den = a * a + b * b;
        const double t_re = a / den, t_im = -b / den;
        const double e_re = blk[kExpLF] * cos(G1 * L), e_im = blk[kExpLF] * sin(G1 * L);
        const double u_re = fourpi * t_re, u_im = fourpi * t_im;
        const double v_re = 1.0 - e_re, v_im = -e_im;
        double c_re = u_re * v_re - u_im * v_im;
        double c_im = u_re * v_im + u_im * v_re;
        J2[0] = c_im / G1;
        for (int n = 1; n < nJ2; ++n) {
            const double w_re = n * c_re - e_re * blk[kPowL + n];
            const double w_im = n * c_im - e_im * blk[kPowL + n];
            const double r_re = t_re * w_re - t_im * w_im;
            const double r_im = t_re * w_im + t_im * w_re;
            c_re = r_re;
            c_im = r_im;
            J2[n] = c_im / G1;
        }
    }
}

// J3_integ by the stored Simpson rule (jastrow_functions.cpp:386-411)
__host__ __device__ inline double J3_integ(const double* blk, double G1)
{
#pragma clang fp contract(off)
    const int nr = (int)blk[kNr];
    const double F = blk[kF];
    const double FG = F * G1;
    const double* r = blk + kMesh;
    const double* w = blk + kMesh + nr;
    double acc = 0.0;
    for (int i = 0; i < nr; ++i) {
        const double FGr = FG * r[i];
        if (FGr < 0.1) {
            double s = 0.0, term = F * r[i];
            for (int t = 0; t < 3; ++t) {
                s += term;
                term *= -FGr * FGr / (double)((2 * t + 3) * (2 * t + 2));
            }
            acc += w[i] * s;
        } else {
            acc += w[i] * sin(FGr) / G1;
        }
    }
    return acc;
}

// ------------------------------------------------------------------ kernels
// uk_POLY (jastrow_functions.cpp:188-208)
__host__ __device__ inline __attribute__((noinline)) double uk_poly(const double* blk, double G2)
{
#pragma clang fp contract(off)
    const int deg = (int)blk[kDeg], nc = deg + (int)blk[kC];
    double J[kMaxDeg + 4];
    J_integ(blk, G2, J, deg + 4, 1);
    double v = 0.0;
    for (int i = 0; i < nc; ++i) v += blk[kCoef + i] * J[i + 1];
    return v;
}

// -sum_{i,j>=1} i j c_i c_j J[i+j-1], the double products summed in double-double
__host__ __device__ inline double grad_u_sq(const double* blk, const double* J, int nc)
{
#pragma clang fp contract(off)
    dd acc{0.0, 0.0};
    for (int i = 1; i < nc; ++i)
        for (int j = 1; j < nc; ++j) {
            const double term = (double)(i * j) * blk[kCoef + i] * blk[kCoef + j] * J[i + j - 1];
            acc = dd_add_d(acc, -term);
        }
    return acc.hi;
}

// tc_2body_POLY (jastrow_functions.cpp:210-258)
__host__ __device__ inline __attribute__((noinline)) double tc2body_poly(const double* blk, double G2)
{
#pragma clang fp contract(off)
    constexpr double fourpi = 4.0 * 3.14159265358979323846;
    const int nc = (int)blk[kDeg] + (int)blk[kC];
    double J[kMaxJ];
    J_integ(blk, G2, J, 2 * nc - 2, 1);
    double s1 = 0.0;
    if (G2 > 1e-12) s1 += fourpi / G2;
    double s2 = 0.0;
    for (int i = 0; i < nc; ++i) s2 -= G2 * blk[kCoef + i] * J[i + 1];
    return s1 + s2 + grad_u_sq(blk, J, nc);
}

// tc_2body_RPAPOLY (jastrow_functions.cpp:260-329) given rpa = tc_2body_RPA(G).
// dA: d/dA at fixed F (rpa is then d tc_2body_RPA/dA): the cross terms are linear in A.
__host__ __device__ inline __attribute__((noinline)) double tc2body_rpapoly(const double* blk, double G2, double rpa, int dA)
{
#pragma clang fp contract(off)
    const int nc = (int)blk[kDeg] + (int)blk[kC];
    const double A = dA ? 1.0 : blk[kA], F = blk[kF];
    double J[kMaxJ], J2[kMaxDeg + 3];
    J_integ(blk, G2, J, 2 * nc - 2, 0);
    const double G1 = sqrt(G2);
    J2_integ(blk, G1, J2, nc - 1);
    double s4 = 0.0;
    for (int i = 2; i < nc; ++i) s4 += 2 * i * blk[kCoef + i] * A * (J[i - 2] - (J2[i - 1] / F + J2[i - 2]));
    s4 -= 2 * blk[kCoef + 1] * A * J3_integ(blk, G1);
    if (dA) return rpa + s4;
    double s2 = 0.0;
    for (int i = 0; i < nc; ++i) s2 -= G2 * blk[kCoef + i] * J[i + 1];
    return rpa + s2 + grad_u_sq(blk, J, nc) + s4;
}

}  // namespace jpoly
}  // namespace tcx
