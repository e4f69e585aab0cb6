/*
 * jastrow_poly.c -- CPU restatement of TC++'s polynomial and RPA+polynomial
 * Jastrow kernels u(G) and tc_2body(G).
 *
 * TEST INFRASTRUCTURE ONLY (part of liboracle.so; see oracle.h).
 *
 * Follows reference src/jastrow_initialize_polynomial.cpp (parameter set-up)
 * and src/jastrow_functions.cpp:99-413 (evaluation).  The reference evaluates
 * the small-G Taylor series of J_integ and the double sum of -(grad u)^2 in
 * boost::multiprecision::cpp_dec_float_50 (50 decimal digits); Boost is not in
 * this image, so those two places use IEEE binary128 (__float128, 34 digits),
 * which carries the same series to the same double result: the series' worst
 * cancellation (G^2 L^2 just below 1000) costs about 14 digits.  Everything
 * the reference does in double is done here in double, in the same order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define PI 3.14159265358979323846
#define FOURPI (4.0 * PI)

/* Jastrow constructor constants, src/include/jastrow.hpp:117-122 */
#define MAX_DEG_POLY 16
#define THRESHOLD_TAYLOR 1000.0
#define MAX_DEG_TAYLOR 75
#define THRESHOLD_TAYLOR2 1.5
#define MAX_DEG_TAYLOR2 30

typedef __float128 quad;

/* Jastrow::set_polynomial_parameters (jastrow_initialize_polynomial.cpp:6-101),
 * set_derived_polynomial_parameters (:103-118), set_derived_RPA_polynomial_parameters
 * (:121-155) and prepare_for_J3_integ (:157-354). */
int orc_jastrow_poly_init(orc_jastrow_poly *jp, int deg, int C, const double L[4], const double *c_input,
                          int cusp, int rpa, const double A[4], const double F[4], const double recip[9],
                          const int n[3])
{
    memset(jp, 0, sizeof(*jp));
    if (deg < 0 || deg > MAX_DEG_POLY) return -1;                 /* :15-16 */
    if (C != 2 && C != 3) return -1;                              /* :17 */
    for (int p = 0; p < 4; ++p)
        if (!(L[p] >= 1e-8)) return -1;                           /* :22-25 */
    if (deg == 0 && cusp) return -1;                              /* :43-46 */
    jp->deg = deg;
    jp->C = C;
    jp->rpa = rpa;
    for (int p = 0; p < 4; ++p) {
        jp->L[p] = L[p];
        jp->A[p] = A[p];
        jp->F[p] = F[p];
        for (int i = 0; i < deg; ++i) jp->c_input[p][i] = c_input[p * deg + i];
    }
    /* cusp condition on the lowest coefficient, :47-60 */
    if (deg > 0) {
        for (int p = 0; p < 4; ++p) {
            const int s1 = p / 2, s2 = p % 2;
            double cusp_const = s1 == s2 ? 0.25 : 0.5;
            if (!cusp) cusp_const = 0.0;
            const double higher = deg == 1 ? 0.0 : jp->c_input[p][1];
            jp->c_input[p][0] = (L[p] / (double)C) * (higher - cusp_const);
        }
    }
    /* c_poly_internal: damping (1 - r/L)^C expanded, with the minus sign, :62-100 */
    for (int p = 0; p < 4; ++p) {
        const double Lp = L[p];
        for (int i = 0; i < deg + C; ++i) jp->c_int[p][i] = 0.0;
        for (int i = 0; i < deg; ++i) {
            const double c = jp->c_input[p][i];
            if (C == 3) {
                jp->c_int[p][i] -= c;
                jp->c_int[p][i + 1] += 3.0 * c / Lp;
                jp->c_int[p][i + 2] -= 3.0 * c / (Lp * Lp);
                jp->c_int[p][i + 3] += c / (Lp * Lp * Lp);
            } else {
                jp->c_int[p][i] -= c;
                jp->c_int[p][i + 1] += 2.0 * c / Lp;
                jp->c_int[p][i + 2] -= c / (Lp * Lp);
            }
        }
    }
    /* 4 pi L^i, :103-118 */
    for (int p = 0; p < 4; ++p) {
        jp->fourpi_L[p][0] = FOURPI;
        for (int i = 1; i < 2 * deg + 6; ++i) jp->fourpi_L[p][i] = L[p] * jp->fourpi_L[p][i - 1];
    }
    if (!rpa || deg == 0) return 0;
    /* exp(-L/F) and M_integ[n] = L^-n int_0^L exp(-r/F) r^n dr by upward recursion, :121-150.
     * The array is sized MAX_DEG_TAYLOR2 + deg + 1 but filled only below MAX_DEG_TAYLOR2;
     * the rest stays zero, and J2_integ's Taylor branch reads those zeros. */
    for (int p = 0; p < 4; ++p) {
        if (!(fabs(F[p]) > 1e-8)) return -1;
        jp->exp_LF[p] = exp(-L[p] / F[p]);
        jp->M[p][0] = F[p] * (1.0 - jp->exp_LF[p]);
        for (int i = 1; i < MAX_DEG_TAYLOR2; ++i)
            jp->M[p][i] = F[p] * (-jp->exp_LF[p] + i * jp->M[p][i - 1] / L[p]);
    }
    if (deg > 1) {
        /* prepare_for_J3_integ, :157-204: |G| of every FFT-grid G at Gamma */
        const long N = (long)n[0] * n[1] * n[2];
        double *Glist = malloc(sizeof(double) * N);
        for (long e = 0; e < N; ++e) {
            const int i1 = (int)(e % n[0]), i2 = (int)((e / n[0]) % n[1]), i3 = (int)(e / ((long)n[0] * n[1]));
            const int m1 = i1 <= n[0] / 2 ? i1 : i1 - n[0];
            const int m2 = i2 <= n[1] / 2 ? i2 : i2 - n[1];
            const int m3 = i3 <= n[2] / 2 ? i3 : i3 - n[2];
            double g2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double g = recip[d] * m1 + recip[3 + d] * m2 + recip[6 + d] * m3;
                g2 += g * g;
            }
            Glist[e] = sqrt(g2);
        }
        double *I_old = malloc(sizeof(double) * 4 * N), *I_new = malloc(sizeof(double) * 4 * N);
        int nr = 101;
        orc_jastrow_poly_set_rmesh(jp, nr);
        for (int p = 0; p < 4; ++p)
            for (long e = 0; e < N; ++e) I_old[p * N + e] = orc_J3_integ(jp, Glist[e], p);
        int small = 0;
        for (int it = 0; it < 15; ++it) {
            const int nr_new = 2 * nr - 1;
            orc_jastrow_poly_set_rmesh(jp, nr_new);
            double err = 0.0;
            for (int p = 0; p < 4; ++p)
                for (long e = 0; e < N; ++e) {
                    I_new[p * N + e] = orc_J3_integ(jp, Glist[e], p);
                    const double d = fabs(I_old[p * N + e] - I_new[p * N + e]);
                    if (err < d) err = d;
                }
            if (err < 1e-6) {                                     /* :189-193 */
                small = 1;
                break;
            }
            nr = nr_new;
            double *t = I_old;
            I_old = I_new;
            I_new = t;
        }
        free(Glist);
        free(I_old);
        free(I_new);
        if (!small) return -2;                                    /* :200 */
        orc_jastrow_poly_set_rmesh(jp, nr);                       /* :203-204: the coarser mesh */
    }
    return 0;
}

void orc_jastrow_poly_free(orc_jastrow_poly *jp)
{
    free(jp->rmesh);
    free(jp->integrand);
    jp->rmesh = jp->integrand = NULL;
    jp->nr = 0;
}

/* Jastrow::set_arrays_J3_integ (:295-354): Simpson weights times 4 pi dr/3 times
 * [exp(-r) + (exp(-r) - 1)/r] on r in [0, L/F] (r in units of F). */
void orc_jastrow_poly_set_rmesh(orc_jastrow_poly *jp, int nr)
{
    free(jp->rmesh);
    free(jp->integrand);
    jp->nr = nr;
    jp->rmesh = malloc(sizeof(double) * 4 * nr);
    jp->integrand = malloc(sizeof(double) * 4 * nr);
    for (int p = 0; p < 4; ++p) {
        const double dr = jp->L[p] / jp->F[p] / (double)(nr - 1);
        const double scale = (dr / 3.0) * FOURPI;
        double *r = jp->rmesh + p * nr, *w = jp->integrand + p * nr;
        for (int i = 0; i < nr; ++i) r[i] = i * dr;
        w[0] = 0.0;
        for (int i = 1; i < nr; ++i) {
            const double ex = exp(-r[i]);
            const double simpson = (i == 0 || i == nr - 1) ? 1.0 : (i % 2 == 0 ? 2.0 : 4.0);
            w[i] = ((ex - 1.0) / r[i] + ex) * (simpson * scale);
        }
    }
}

/* Jastrow::J3_integ (jastrow_functions.cpp:386-411):
 * (4 pi / G) int_0^{L/F} sin(F G r) [exp(-r) + (exp(-r) - 1)/r] dr by the stored Simpson rule. */
double orc_J3_integ(const orc_jastrow_poly *jp, double G1, int p)
{
    const double F = jp->F[p];
    const double FG = F * G1;
    const double *r = jp->rmesh + p * jp->nr, *w = jp->integrand + p * jp->nr;
    double acc = 0.0;
    for (int i = 0; i < jp->nr; ++i) {
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

/* Jastrow::calculate_J_integ (jastrow_functions.cpp:99-186):
 * J[n] = 4 pi int_0^L r^{n+1} sin(G r)/(G r) dr for n in [n0, nJ). */
void orc_J_integ(const orc_jastrow_poly *jp, double G2, int p, double *J, int nJ, int zeroth)
{
    const double L = jp->L[p];
    const double G2L2 = G2 * L * L;
    const int n0 = zeroth ? 0 : 1;
    if (G2L2 < THRESHOLD_TAYLOR) {
        /* 1/(n+2) - (GL)^2/(3!(n+4)) + (GL)^4/(5!(n+6)) - ..., Horner from the 75th term */
        for (int n = n0; n < nJ; ++n) {
            quad t = 0;
            for (int j = 0; j < MAX_DEG_TAYLOR; ++j) {
                const int dg = 2 * (MAX_DEG_TAYLOR - j);
                t = -t * (quad)G2L2 / (quad)((dg + 1) * dg) + (quad)1 / (quad)(n + dg);
            }
            J[n] = (double)t;
        }
    } else {
        const double GL = sqrt(G2L2);
        const double cs = cos(GL), sn = sin(GL);
        for (int n = n0; n < nJ; ++n) {
            double d = -1.0 / G2L2;                               /* cosine term */
            int m = n;
            double coeff = d;
            for (int j = 0; j < n / 2; ++j) {
                d *= -m * (m - 1) / G2L2;
                m -= 2;
                coeff += d;
            }
            J[n] = coeff * cs;
            if (n % 2 == 0) J[n] -= d;                            /* constant term */
            if (n > 0) {                                          /* sine term */
                d = n / (G2L2 * GL);
                m = n - 1;
                coeff = d;
                for (int j = 0; j < (n - 1) / 2; ++j) {
                    d *= -m * (m - 1) / G2L2;
                    m -= 2;
                    coeff += d;
                }
                J[n] += coeff * sn;
            }
        }
    }
    for (int n = n0; n < nJ; ++n) J[n] *= jp->fourpi_L[p][n + 2];
}

/* Jastrow::calculate_J2_integ (jastrow_functions.cpp:333-382):
 * J2[n] = (4 pi/G) Im int_0^L r^n exp(iGr) exp(-r/F) dr. */
void orc_J2_integ(const orc_jastrow_poly *jp, double G1, int p, double *J2, int nJ2)
{
    const double L = jp->L[p];
    const double G2L2 = (G1 * L) * (G1 * L);
    if (G2L2 < THRESHOLD_TAYLOR2) {
        for (int n = 0; n < nJ2; ++n) {
            double acc = 0.0, coeff = 1.0;
            for (int j = 0; j < MAX_DEG_TAYLOR2 / 2; ++j) {
                const int idx = n + 2 * j + 1;
                acc += coeff * (idx < MAX_DEG_TAYLOR2 ? jp->M[p][idx] : 0.0);
                coeff *= (-G2L2) / (double)((2 * j + 3) * (2 * j + 2));
            }
            J2[n] = acc * jp->fourpi_L[p][n + 1];
        }
    } else {
        /* ctmp1 = 1/(1/F - iG), e2 = exp(-L/F) exp(iGL) */
        const double a = 1.0 / jp->F[p], b = -G1;
        const double den = a * a + b * b;
        const double t_re = a / den, t_im = -b / den;
        const double e_re = jp->exp_LF[p] * cos(G1 * L), e_im = jp->exp_LF[p] * sin(G1 * L);
        double c_re, c_im;
        {   /* 4 pi ctmp1 (1 - e2) */
            const double u_re = FOURPI * t_re, u_im = FOURPI * t_im;
            const double v_re = 1.0 - e_re, v_im = -e_im;
            c_re = u_re * v_re - u_im * v_im;
            c_im = u_re * v_im + u_im * v_re;
        }
        J2[0] = c_im / G1;
        for (int n = 1; n < nJ2; ++n) {
            const double w_re = n * c_re - e_re * jp->fourpi_L[p][n];
            const double w_im = n * c_im - e_im * jp->fourpi_L[p][n];
            const double r_re = t_re * w_re - t_im * w_im;
            const double r_im = t_re * w_im + t_im * w_re;
            c_re = r_re;
            c_im = r_im;
            J2[n] = c_im / G1;
        }
    }
}

/* Jastrow::uk_POLY (jastrow_functions.cpp:188-208) */
double orc_uk_poly(const orc_jastrow_poly *jp, double G2, int p)
{
    double J[MAX_DEG_POLY + 4];
    orc_J_integ(jp, G2, p, J, jp->deg + 4, 0);
    double v = 0.0;
    for (int i = 0; i < jp->deg + jp->C; ++i) v += jp->c_int[p][i] * J[i + 1];
    return v;
}

/* -sum_{i,j>=1} i j c_i c_j J[i+j-1], accumulated beyond double (:247-256, :301-311) */
static double grad_u_squared(const orc_jastrow_poly *jp, int p, const double *J)
{
    const int nc = jp->deg + jp->C;
    quad acc = 0;
    for (int i = 1; i < nc; ++i)
        for (int j = 1; j < nc; ++j) {
            const double term = (double)(i * j) * jp->c_int[p][i] * jp->c_int[p][j] * J[i + j - 1];
            acc -= (quad)term;
        }
    return (double)acc;
}

/* Jastrow::tc_2body_POLY (jastrow_functions.cpp:210-258) */
double orc_tc2body_poly(const orc_jastrow_poly *jp, double G2, int p)
{
    double J[2 * (MAX_DEG_POLY + 3)];
    const int nc = jp->deg + jp->C;
    orc_J_integ(jp, G2, p, J, 2 * nc - 2, 0);
    double s1 = 0.0;
    if (G2 > 1e-12) s1 += FOURPI / G2;
    double s2 = 0.0;
    for (int i = 0; i < nc; ++i) s2 -= G2 * jp->c_int[p][i] * J[i + 1];
    return s1 + s2 + grad_u_squared(jp, p, J);
}

/* Jastrow::tc_2body_RPAPOLY (jastrow_functions.cpp:260-329); rpa = tc_2body_RPA(G). */
double orc_tc2body_rpapoly(const orc_jastrow_poly *jp, double G2, int p, double rpa)
{
    double J[2 * (MAX_DEG_POLY + 3)], J2[MAX_DEG_POLY + 3];
    const int nc = jp->deg + jp->C;
    const double A = jp->A[p], F = jp->F[p];
    orc_J_integ(jp, G2, p, J, 2 * nc - 2, 1);
    double s2 = 0.0;
    for (int i = 0; i < nc; ++i) s2 -= G2 * jp->c_int[p][i] * J[i + 1];
    const double s3 = grad_u_squared(jp, p, J);
    const double G1 = sqrt(G2);
    orc_J2_integ(jp, G1, p, J2, jp->deg + jp->C - 1);
    double s4 = 0.0;
    for (int i = 2; i < nc; ++i)
        s4 += 2 * i * jp->c_int[p][i] * A * (J[i - 2] - (J2[i - 1] / F + J2[i - 2]));
    s4 -= 2 * jp->c_int[p][1] * A * orc_J3_integ(jp, G1, p);
    return rpa + s2 + s3 + s4;
}
