/*
 * oracle.c -- FP64 CPU restatement of TC++'s pair-FFT operator apply.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Every function below restates
 * the reference algorithm from reading its source; the reference file:line
 * it follows is cited at each function.  The loop nests deliberately keep
 * the reference's structure (orbital FFTs recomputed inside the pair loop,
 * OpenMP over q, accumulation under a critical section) because this code
 * is also the CPU baseline that the GPU path is timed against.
 *
 * Parity pins (tests/): orc_fft3d against orc_dft3d_naive and numpy.fft;
 * orc_hf2x / orc_tc2x against a brute-force G-space convolution (no FFT) on
 * tiny grids, the plane-wave closed form of the exchange operator, and
 * operator invariants (Hermiticity, A=0 => tc2x == hf2x, BITC(chi=phi) ==
 * TC, quadratic dependence on lambda*u); orc_tc3a1 / orc_tc3b1 against a
 * brute-force G-space restatement and the uniform-density zero.
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static const double PI = 3.14159265358979323846; /* header.hpp:43 */
#define FOURPI (4.0 * PI)                          /* header.hpp:44 */

/* ------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time, any length.          */
/* ------------------------------------------------------------------ */

static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    for (int p = 2; p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

/* tw[j] = exp(sign * 2*pi*i * j / n0) for j < n0 */
static void make_twiddles(cplx *tw, int n0, int sign)
{
    for (int j = 0; j < n0; ++j) {
        double a = 2.0 * PI * (double)j / (double)n0;
        tw[j] = cos(a) + sign * I * sin(a);
    }
}

/* out[k] = sum_t in[t*is] * W_n^{t k}, W_n = tw[step] with n*step = n0 */
static void fft_rec(const cplx *in, int is, cplx *out, int n, int step, int n0, const cplx *tw)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n);
    int m = n / p;
    for (int r = 0; r < p; ++r)
        fft_rec(in + (long)r * is, is * p, out + r * m, m, step * p, n0, tw);
    cplx t[64], y[64];
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r)
            t[r] = out[r * m + k] * tw[((long)r * k * step) % n0];
        for (int s = 0; s < p; ++s) {
            cplx acc = 0;
            for (int r = 0; r < p; ++r)
                acc += t[r] * tw[((long)r * s * m * step) % n0];
            y[s] = acc;
        }
        for (int s = 0; s < p; ++s) out[k + s * m] = y[s];
    }
}

void orc_fft3d(double *grid_d, const int n[3], int sign)
{
    cplx *grid = (cplx *)grid_d;
    const int n1 = n[0], n2 = n[1], n3 = n[2];
    const long N = (long)n1 * n2 * n3;
    int nmax = n1 > n2 ? n1 : n2;
    if (n3 > nmax) nmax = n3;
    cplx *tw = malloc(sizeof(cplx) * nmax);
    cplx *line = malloc(sizeof(cplx) * nmax);
    cplx *res = malloc(sizeof(cplx) * nmax);
    /* axis 1 (contiguous) */
    make_twiddles(tw, n1, sign);
    for (long b = 0; b < (long)n2 * n3; ++b) {
        cplx *row = grid + b * n1;
        fft_rec(row, 1, res, n1, 1, n1, tw);
        memcpy(row, res, sizeof(cplx) * n1);
    }
    /* axis 2 */
    make_twiddles(tw, n2, sign);
    for (int i3 = 0; i3 < n3; ++i3)
        for (int i1 = 0; i1 < n1; ++i1) {
            cplx *base = grid + (long)i3 * n1 * n2 + i1;
            for (int i2 = 0; i2 < n2; ++i2) line[i2] = base[(long)i2 * n1];
            fft_rec(line, 1, res, n2, 1, n2, tw);
            for (int i2 = 0; i2 < n2; ++i2) base[(long)i2 * n1] = res[i2];
        }
    /* axis 3 */
    make_twiddles(tw, n3, sign);
    for (long b = 0; b < (long)n1 * n2; ++b) {
        cplx *base = grid + b;
        for (int i3 = 0; i3 < n3; ++i3) line[i3] = base[(long)i3 * n1 * n2];
        fft_rec(line, 1, res, n3, 1, n3, tw);
        for (int i3 = 0; i3 < n3; ++i3) base[(long)i3 * n1 * n2] = res[i3];
    }
    if (sign < 0) /* FFT_forward divides by N: plane_wave_basis.cpp:279 */
        for (long i = 0; i < N; ++i) grid[i] /= (double)N;
    free(tw); free(line); free(res);
}

void orc_dft3d_naive(const double *in_d, double *out_d, const int n[3], int sign)
{
    const cplx *in = (const cplx *)in_d;
    cplx *out = (cplx *)out_d;
    const long N = (long)n[0] * n[1] * n[2];
    for (long o = 0; o < N; ++o) {
        int o1 = o % n[0], o2 = (o / n[0]) % n[1], o3 = o / ((long)n[0] * n[1]);
        cplx acc = 0;
        for (long s = 0; s < N; ++s) {
            int s1 = s % n[0], s2 = (s / n[0]) % n[1], s3 = s / ((long)n[0] * n[1]);
            double ph = 2.0 * PI * ((double)((long)o1 * s1 % n[0]) / n[0] +
                                    (double)((long)o2 * s2 % n[1]) / n[1] +
                                    (double)((long)o3 * s3 % n[2]) / n[2]);
            acc += in[s] * (cos(ph) + sign * I * sin(ph));
        }
        out[o] = sign < 0 ? acc / (double)N : acc;
    }
}

/* ------------------------------------------------------------------ */
/* Grid helpers                                                         */
/* ------------------------------------------------------------------ */

/* plane_wave_basis.cpp:205-216 (Miller decode) and the Cartesian product
 * reciprocal_vectors().transpose() * m (calc_hamiltonian_hf2x.cpp:35-36). */
static void grid_gcart(const orc_problem *p, double *G /* N*3 */)
{
    const int n1 = p->n[0], n2 = p->n[1], n3 = p->n[2];
    const long N = (long)n1 * n2 * n3;
    for (long ipw = 0; ipw < N; ++ipw) {
        int i1 = ipw % n1, i2 = (ipw / n1) % n2, i3 = ipw / ((long)n1 * n2);
        int m1 = i1 <= n1 / 2 ? i1 : i1 - n1;
        int m2 = i2 <= n2 / 2 ? i2 : i2 - n2;
        int m3 = i3 <= n3 / 2 ? i3 : i3 - n3;
        for (int d = 0; d < 3; ++d)
            G[ipw * 3 + d] = p->recip[0 * 3 + d] * m1 + p->recip[1 * 3 + d] * m2 + p->recip[2 * 3 + d] * m3;
    }
}

static void kcart(const orc_problem *p, const double *kf, double *kc)
{
    for (int d = 0; d < 3; ++d)
        kc[d] = p->recip[0 * 3 + d] * kf[0] + p->recip[1 * 3 + d] * kf[1] + p->recip[2 * 3 + d] * kf[2];
}

static long pw_offset(const orc_problem *p, int ik)
{
    long off = 0;
    for (int k = 0; k < ik; ++k) off += p->npw[k];
    return off;
}

static long occ_offset(const orc_problem *p, int ik) /* in orbitals */
{
    long off = 0;
    for (int k = 0; k < ik; ++k) off += p->nocc[k];
    return off;
}

static long occ_coef_offset(const orc_problem *p, int ik) /* in complex numbers */
{
    long off = 0;
    for (int k = 0; k < ik; ++k) off += (long)p->nocc[k] * p->npw[k];
    return off;
}

static long app_coef_offset(const orc_problem *p, const int *nb, int ik)
{
    long off = 0;
    for (int k = 0; k < ik; ++k) off += (long)nb[k] * p->npw[k];
    return off;
}

/* get_orbital_FFTgrid, identity symmetry: plane_wave_basis.cpp:221-271 */
static void scatter(const orc_problem *p, int ik, const cplx *c, cplx *grid, long N)
{
    memset(grid, 0, sizeof(cplx) * N);
    const int *gi = p->gindex + pw_offset(p, ik);
    for (int l = 0; l < p->npw[ik]; ++l) grid[gi[l]] = c[l];
}

static cplx zdotc(const cplx *a, const cplx *b, int n) /* Eigen a.dot(b) = sum conj(a) b */
{
    cplx s = 0;
    for (int i = 0; i < n; ++i) s += conj(a[i]) * b[i];
    return s;
}

/* ------------------------------------------------------------------ */
/* Zero-weight k-points (BAND mode, fake SCF)                           */
/* ------------------------------------------------------------------ */

static int zero_weight(const orc_problem *p, int ik) { return p->kweight && p->kweight[ik] < 1e-8; }

/* k is on the weighted mesh (modulo reciprocal lattice vectors):
 * tc3a3b3b4b6.cpp:396-410, potentials_coulomb.cpp:271-287 */
static int if_match(const orc_problem *p, int ik)
{
    for (int iq = 0; iq < p->nk; ++iq) {
        if (zero_weight(p, iq)) continue;
        double kq[3];
        for (int d = 0; d < 3; ++d) {
            kq[d] = p->kfrac[3 * ik + d] - p->kfrac[3 * iq + d];
            while (kq[d] < -1e-8) kq[d] += 1.0;
            while (kq[d] > 1.0 - 1e-8) kq[d] -= 1.0;
        }
        if (kq[0] * kq[0] + kq[1] * kq[1] + kq[2] * kq[2] < 1e-8) return 1;
    }
    return 0;
}

/* The states the divergence terms at k project on (the ibandk loops, e.g.
 * hf2x.cpp:166-199): the occupied orbitals at a weighted k (their fillings, old bands
 * included), the given states with their return_band_filling values at a zero-weight k. */
typedef struct {
    int n;
    const double *fill;
    const cplx *phi, *chi;
} div_set;

static div_set div_states(const orc_problem *p, int ik)
{
    div_set s;
    if (zero_weight(p, ik)) {
        long fo = 0, co = 0;
        for (int k = 0; k < ik; ++k) {
            fo += p->zw_nstate ? p->zw_nstate[k] : 0;
            co += (long)(p->zw_nstate ? p->zw_nstate[k] : 0) * p->npw[k];
        }
        s.n = p->zw_nstate ? p->zw_nstate[ik] : 0;
        s.fill = p->zw_fill + fo;
        s.phi = (const cplx *)p->zw_phi + co;
        s.chi = (p->method == 2 && p->zw_chi) ? (const cplx *)p->zw_chi + co : s.phi;
    } else {
        long fo = 0, co = 0;
        for (int k = 0; k < ik; ++k) {
            fo += p->nocc[k];
            co += (long)p->nocc[k] * p->npw[k];
        }
        s.n = p->nocc[ik];
        s.fill = p->filling + fo;
        s.phi = (const cplx *)p->phi_occ + co;
        s.chi = (p->method == 2 && p->chi_occ) ? (const cplx *)p->chi_occ + co : s.phi;
    }
    return s;
}

static void kvaux_of(const orc_problem *p, int ik, double out[3])
{
    for (int d = 0; d < 3; ++d) out[d] = p->sum_kvaux ? p->sum_kvaux[3 * ik + d] : 0.0;
}

/* (k + G)_d at plane wave l of k (Cartesian) */
static void kg_of(const orc_problem *p, int ik, int l, double out[3])
{
    double kc[3];
    kcart(p, p->kfrac + 3 * ik, kc);
    const int e = p->gindex[pw_offset(p, ik) + l];
    const int n1 = p->n[0], n2 = p->n[1], n3 = p->n[2];
    const int i1 = e % n1, i2 = (e / n1) % n2, i3 = e / (n1 * n2);
    const int m1 = i1 <= n1 / 2 ? i1 : i1 - n1, m2 = i2 <= n2 / 2 ? i2 : i2 - n2, m3 = i3 <= n3 / 2 ? i3 : i3 - n3;
    for (int d = 0; d < 3; ++d)
        out[d] = kc[d] + p->recip[0 * 3 + d] * m1 + p->recip[1 * 3 + d] * m2 + p->recip[2 * 3 + d] * m3;
}

/* ------------------------------------------------------------------ */
/* Jastrow, jastrow_functions.cpp:6-97 (polynomial: jastrow_poly.c)     */
/* ------------------------------------------------------------------ */

double orc_uk(const orc_problem *p, double G2, int s1, int s2)
{
    const int poly = p->poly && p->poly->deg > 0;
    double v = 0.0;
    if (!p->jastrow_A_zero) {                             /* uk :9-13 */
        const double A = p->A[s1 * 2 + s2], F = p->F[s1 * 2 + s2];
        const double F2 = F * F;
        if (G2 < 1e-12) v += -FOURPI * A * F2;            /* uk_RPA :63-67 */
        else v += FOURPI * A / (G2 * (G2 * F2 + 1.0));    /* :70 */
    }
    if (poly) v += orc_uk_poly(p->poly, G2, s1 * 2 + s2); /* :14-17 */
    return v;
}

static double tc_2body_rpa(const orc_problem *p, double G2, int s1, int s2)
{
    const double A = p->A[s1 * 2 + s2], F = p->F[s1 * 2 + s2];
    const double F2 = F * F;
    const double PIAAF2 = PI * A * A / F2;                /* PI_AAFinv2_long_, jastrow_initialize.cpp:107 */
    if (G2 < 1e-12) return -2.0 * PIAAF2 * F;             /* tc_2body_RPA :79-85 */
    const double G1 = sqrt(G2);
    const double FG1 = F * G1;
    const double F2G2 = F2 * G2;
    return FOURPI / G2 - FOURPI * A / (1.0 + F2G2)
         + (PIAAF2 / G1) * (PI * F2G2 + (4.0 + 2.0 * F2G2) * atan(FG1 / 2.0)
                            - (4.0 + 4.0 * F2G2) * atan(FG1)); /* :88-95 */
}

double orc_tc_2body(const orc_problem *p, double G2, int s1, int s2)
{
    const int deg = p->poly ? p->poly->deg : 0;
    if (p->jastrow_A_zero) {                              /* tc_2body :22-39 */
        if (deg > 0) return orc_tc2body_poly(p->poly, G2, s1 * 2 + s2);
        if (G2 > 1e-12) return FOURPI / G2;
        return 0.0;
    }
    if (deg > 1)                                          /* :41-46 */
        return orc_tc2body_rpapoly(p->poly, G2, s1 * 2 + s2, tc_2body_rpa(p, G2, s1, s2));
    return tc_2body_rpa(p, G2, s1, s2);
}

/* ------------------------------------------------------------------ */
/* Divergence correction sums, potentials_coulomb.cpp:60-161           */
/* ------------------------------------------------------------------ */

void orc_sum_of_vaux(const double recip[9], double volume, int nk, const double *kfrac,
                     double *alpha_out, double *sum_out)
{
    double alpha = cbrt(volume) / (2.0 * PI);
    alpha = alpha * alpha;                                  /* :69-71 */
    for (int ik = 0; ik < nk; ++ik) {
        double sum = 0.0;
        int if_match = 0;
        for (int iq = 0; iq < nk; ++iq) {
            double kq[3];
            for (int d = 0; d < 3; ++d) {
                kq[d] = kfrac[ik * 3 + d] - kfrac[iq * 3 + d];
                while (kq[d] < -1e-8) kq[d] += 1.0;         /* :92-96 */
                while (kq[d] > 1.0 - 1e-8) kq[d] -= 1.0;
            }
            int match = (kq[0] * kq[0] + kq[1] * kq[1] + kq[2] * kq[2]) < 1e-8;
            if (match) if_match = 1;
            if (!match) {                                   /* nshell = 0: :109-114 */
                double c[3];
                for (int d = 0; d < 3; ++d)
                    c[d] = recip[0 * 3 + d] * kq[0] + recip[1 * 3 + d] * kq[1] + recip[2 * 3 + d] * kq[2];
                double kq2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
                sum += exp(-alpha * kq2) / kq2;
            }
            for (int nshell = 1;; ++nshell) {               /* :115-136 */
                double shell = 0.0;
                for (int gx = -nshell; gx <= nshell; ++gx)
                    for (int gy = -nshell; gy <= nshell; ++gy)
                        for (int gz = -nshell; gz <= nshell; ++gz) {
                            if (gx != nshell && -gx != nshell && gy != nshell && -gy != nshell &&
                                gz != nshell && -gz != nshell) continue;
                            double m[3] = {kq[0] + gx, kq[1] + gy, kq[2] + gz}, c[3];
                            for (int d = 0; d < 3; ++d)
                                c[d] = recip[0 * 3 + d] * m[0] + recip[1 * 3 + d] * m[1] + recip[2 * 3 + d] * m[2];
                            double g2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
                            shell += exp(-alpha * g2) / g2;
                        }
                sum += shell;
                if (shell < 1e-20) break;
            }
        }
        sum *= FOURPI;                                      /* :139 */
        if (if_match) sum -= FOURPI * alpha;                /* :149-160 (k on the SCF mesh) */
        sum_out[ik] = sum;
    }
    *alpha_out = alpha;
}

/* set_sum_of_Vaux_each and set_sum_of_kVaux_each (potentials_coulomb.cpp:60-254) with the
 * q-sum over the weighted k only (:86, :198): the shells of every (k, q) until a shell's
 * sum (squared norm for the vector) drops below 1e-20; the q = k term of the nshell = 0
 * shell is left out when k matches q; -4 pi alpha for k on the mesh (:149-160); the vector
 * sum is zero there (:187-191). */
void orc_sum_of_vaux_w(const double recip[9], double volume, int nk, const double *kfrac, const double *kweight,
                       double *alpha_out, double *sum_out, double *ksum_out)
{
    double alpha = cbrt(volume) / (2.0 * PI);
    alpha = alpha * alpha;
    for (int ik = 0; ik < nk; ++ik) {
        double sum = 0.0, ks[3] = {0.0, 0.0, 0.0};
        int match_any = 0;
        for (int iq = 0; iq < nk; ++iq) {
            if (kweight && kweight[iq] < 1e-8) continue;
            double kq[3];
            for (int d = 0; d < 3; ++d) {
                kq[d] = kfrac[ik * 3 + d] - kfrac[iq * 3 + d];
                while (kq[d] < -1e-8) kq[d] += 1.0;
                while (kq[d] > 1.0 - 1e-8) kq[d] -= 1.0;
            }
            const int match = (kq[0] * kq[0] + kq[1] * kq[1] + kq[2] * kq[2]) < 1e-8;
            if (match) match_any = 1;
            if (!match) {
                double c[3];
                for (int d = 0; d < 3; ++d)
                    c[d] = recip[0 * 3 + d] * kq[0] + recip[1 * 3 + d] * kq[1] + recip[2 * 3 + d] * kq[2];
                const double kq2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
                const double w = exp(-alpha * kq2) / kq2;
                sum += w;
                for (int d = 0; d < 3; ++d) ks[d] += w * c[d];
            }
            /* scalar and vector shells stop independently (:135, :245) */
            int done_s = 0, done_v = 0;
            for (int nshell = 1; !(done_s && done_v); ++nshell) {
                double shell = 0.0, vshell[3] = {0.0, 0.0, 0.0};
                for (int gx = -nshell; gx <= nshell; ++gx)
                    for (int gy = -nshell; gy <= nshell; ++gy)
                        for (int gz = -nshell; gz <= nshell; ++gz) {
                            if (gx != nshell && -gx != nshell && gy != nshell && -gy != nshell &&
                                gz != nshell && -gz != nshell) continue;
                            double m[3] = {kq[0] + gx, kq[1] + gy, kq[2] + gz}, c[3];
                            for (int d = 0; d < 3; ++d)
                                c[d] = recip[0 * 3 + d] * m[0] + recip[1 * 3 + d] * m[1] + recip[2 * 3 + d] * m[2];
                            const double g2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
                            const double w = exp(-alpha * g2) / g2;
                            shell += w;
                            for (int d = 0; d < 3; ++d) vshell[d] += w * c[d];
                        }
                if (!done_s) {
                    sum += shell;
                    if (shell < 1e-20) done_s = 1;
                }
                if (!done_v) {
                    for (int d = 0; d < 3; ++d) ks[d] += vshell[d];
                    if (vshell[0] * vshell[0] + vshell[1] * vshell[1] + vshell[2] * vshell[2] < 1e-20) done_v = 1;
                }
            }
        }
        sum *= FOURPI;
        if (match_any) sum -= FOURPI * alpha;
        sum_out[ik] = sum;
        if (ksum_out)
            for (int d = 0; d < 3; ++d) ksum_out[3 * ik + d] = match_any ? 0.0 : FOURPI * ks[d];
    }
    *alpha_out = alpha;
}

/* ------------------------------------------------------------------ */
/* Density, bloch_states_scfloop_filling_density.cpp:256-358            */
/* ------------------------------------------------------------------ */

void orc_density(const orc_problem *p, double *out_d)
{
    cplx *out = (cplx *)out_d;
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const int bitc = p->method == 2;
    cplx *orb = malloc(sizeof(cplx) * N), *orbl = malloc(sizeof(cplx) * N);
    memset(out, 0, sizeof(cplx) * 2 * N);
    const cplx *phi = (const cplx *)p->phi_occ;
    const cplx *chi = p->chi_occ ? (const cplx *)p->chi_occ : phi;
    for (int ik = 0; ik < p->nk; ++ik) {
        long co = occ_coef_offset(p, ik), fo = occ_offset(p, ik);
        for (int ib = 0; ib < p->nocc[ik]; ++ib) {
            scatter(p, ik, phi + co + (long)ib * p->npw[ik], orb, N);
            orc_fft3d((double *)orb, p->n, +1);
            if (bitc) {
                scatter(p, ik, chi + co + (long)ib * p->npw[ik], orbl, N);
                orc_fft3d((double *)orbl, p->n, +1);
            }
            double f = p->filling[fo + ib];
            for (long r = 0; r < N; ++r)
                out[r] += f * (bitc ? conj(orbl[r]) * orb[r] : (cplx)(creal(orb[r]) * creal(orb[r]) + cimag(orb[r]) * cimag(orb[r])));
        }
    }
    if (p->nspin == 1) {                                    /* :343, :356-358 */
        for (long r = 0; r < N; ++r) { out[r] /= 2.0; out[N + r] = out[r]; }
    }
    free(orb); free(orbl);
}

/* ------------------------------------------------------------------ */
/* hf2x, calc_hamiltonian_hf2x.cpp:8-205                                */
/* ------------------------------------------------------------------ */

static double div_corr_value(const orc_problem *p, int ik) /* hf2x.cpp:153-163, tc2x.cpp:412-422 */
{
    double dc = 1.0 / sqrt(PI * p->alpha);
    dc -= p->sum_vaux[ik] / (p->num_kpoints * p->volume);
    dc *= -1.0;
    if (p->nspin == 1) dc /= 2.0;
    return dc;
}

void orc_hf2x(const orc_problem *p, const int *nb, const double *phi_d, double *hphi_d, int nthreads)
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const cplx *phi = (const cplx *)phi_d;
    cplx *hphi = (cplx *)hphi_d;
    const cplx *occ = (const cplx *)p->phi_occ;
    double factor = -FOURPI / (p->num_kpoints * p->volume);  /* :29 */
    if (p->nspin == 1) factor /= 2.0;                        /* :30 */
    double *G = malloc(sizeof(double) * 3 * N);
    grid_gcart(p, G);
    cplx *phij = malloc(sizeof(cplx) * N), *hsub = malloc(sizeof(cplx) * N);
    if (nthreads <= 0) nthreads = 0;

    for (int ik = 0; ik < p->nk; ++ik) {
        const long ao = app_coef_offset(p, nb, ik);
        const int npwk = p->npw[ik];
        const int *gi = p->gindex + pw_offset(p, ik);
        double kc[3];
        kcart(p, p->kfrac + 3 * ik, kc);
        for (int jb = 0; jb < nb[ik]; ++jb) {
            scatter(p, ik, phi + ao + (long)jb * npwk, phij, N);  /* :60-64 */
            orc_fft3d((double *)phij, p->n, +1);
            memset(hsub, 0, sizeof(cplx) * N);
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
            {
                cplx *phiq = malloc(sizeof(cplx) * N), *phiqj = malloc(sizeof(cplx) * N);
                cplx *hloc = calloc(N, sizeof(cplx));
                double *V = malloc(sizeof(double) * N);
#pragma omp for schedule(dynamic)
                for (int iq = 0; iq < p->nk; ++iq) {                  /* :80 */
                    if (p->nocc[iq] == 0 || zero_weight(p, iq)) continue; /* :90 */
                    double qc[3];
                    kcart(p, p->kfrac + 3 * iq, qc);
                    for (long r = 0; r < N; ++r) {                     /* :91-104 */
                        double x = kc[0] - qc[0] + G[3 * r], y = kc[1] - qc[1] + G[3 * r + 1],
                               z = kc[2] - qc[2] + G[3 * r + 2];
                        double t = x * x + y * y + z * z;
                        V[r] = t < 1e-8 ? 0.0 : 1.0 / t;
                    }
                    const long co = occ_coef_offset(p, iq), fo = occ_offset(p, iq);
                    for (int ib = 0; ib < p->nocc[iq]; ++ib) {         /* :107 */
                        scatter(p, iq, occ + co + (long)ib * p->npw[iq], phiq, N);
                        orc_fft3d((double *)phiq, p->n, +1);           /* :123 */
                        for (long r = 0; r < N; ++r) phiqj[r] = conj(phiq[r]) * phij[r];
                        orc_fft3d((double *)phiqj, p->n, -1);          /* :125 */
                        for (long r = 0; r < N; ++r) phiqj[r] *= V[r];
                        orc_fft3d((double *)phiqj, p->n, +1);          /* :127 */
                        const double f = p->filling[fo + ib];
                        for (long r = 0; r < N; ++r) hloc[r] += f * phiqj[r] * phiq[r]; /* :134 */
                    }
                }
#pragma omp critical
                for (long r = 0; r < N; ++r) hsub[r] += hloc[r];
                free(phiq); free(phiqj); free(hloc); free(V);
            }
            orc_fft3d((double *)hsub, p->n, -1);                       /* :143 */
            cplx *out = hphi + ao + (long)jb * npwk;
            for (int l = 0; l < npwk; ++l) out[l] += factor * hsub[gi[l]]; /* :144-147 */

            if (p->div_corr) {                                          /* :150-201 */
                const double dc = div_corr_value(p, ik);
                const div_set D = div_states(p, ik);                    /* :166-191 */
                for (int b = 0; b < D.n; ++b) {
                    const double f = D.fill[b];
                    if (f < 1e-8) continue;
                    const cplx *pk = D.phi + (long)b * npwk;
                    cplx prod = zdotc(pk, phi + ao + (long)jb * npwk, npwk) * f * dc;
                    for (int l = 0; l < npwk; ++l) out[l] += prod * pk[l];
                }
            }
        }
    }
    free(G); free(phij); free(hsub);
}

/* ------------------------------------------------------------------ */
/* tc2x, calc_hamiltonian_tc2x.cpp:11-516                               */
/* (the kVaux branches :117-222, :483-508 act at k off the weighted     */
/*  mesh only: sum_of_kVaux = 0 on it, potentials_coulomb.cpp:187-191)  */
/* ------------------------------------------------------------------ */

void orc_tc2x(const orc_problem *p, const int *nb, const double *phi_d, double *hphi_d, int nthreads)
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const cplx *phi = (const cplx *)phi_d;
    cplx *hphi = (cplx *)hphi_d;
    const cplx *occ = (const cplx *)p->phi_occ;
    const int bitc = p->method == 2;
    const cplx *occl = (bitc && p->chi_occ) ? (const cplx *)p->chi_occ : occ;
    const int s = 0; /* ispin, num_independent_spins == 1 */
    double factor = -1.0 / (p->num_kpoints * p->volume);   /* :32 */
    if (p->nspin == 1) factor /= 2.0;                       /* :33 */
    double *G = malloc(sizeof(double) * 3 * N);
    grid_gcart(p, G);

    for (int ik = 0; ik < p->nk; ++ik) {
        const long ao = app_coef_offset(p, nb, ik);
        const int npwk = p->npw[ik];
        const int *gi = p->gindex + pw_offset(p, ik);
        double kc[3];
        kcart(p, p->kfrac + 3 * ik, kc);                    /* :115 */
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
        {
            double *V2 = malloc(sizeof(double) * N), *duk = malloc(sizeof(double) * 3 * N);
            cplx *phij = malloc(sizeof(cplx) * N), *gphij = malloc(sizeof(cplx) * 3 * N);
            cplx *phiq = malloc(sizeof(cplx) * N), *chiq = malloc(sizeof(cplx) * N);
            cplx *gphiq = malloc(sizeof(cplx) * 3 * N), *phiqj = malloc(sizeof(cplx) * N);
            cplx *pgp = malloc(sizeof(cplx) * 3 * N), *hsub = malloc(sizeof(cplx) * N);
#pragma omp for schedule(dynamic)
            for (int iq = 0; iq < p->nk; ++iq) {                              /* :241 */
                if (p->nocc[iq] == 0 || zero_weight(p, iq)) continue;          /* :251 */
                double qc[3];
                kcart(p, p->kfrac + 3 * iq, qc);                               /* :253-255 */
                for (long r = 0; r < N; ++r) {                                 /* :256-276 */
                    double kq[3], g2 = 0;
                    for (int d = 0; d < 3; ++d) { kq[d] = kc[d] - qc[d] + G[3 * r + d]; g2 += kq[d] * kq[d]; }
                    V2[r] = orc_tc_2body(p, g2, s, s);
                    double u = orc_uk(p, g2, s, s);
                    for (int d = 0; d < 3; ++d) duk[d * N + r] = kq[d] * u;
                }
                const long co = occ_coef_offset(p, iq), fo = occ_offset(p, iq);
                for (int jb = 0; jb < nb[ik]; ++jb) {                          /* :279 */
                    scatter(p, ik, phi + ao + (long)jb * npwk, phij, N);      /* :283-298 */
                    for (int d = 0; d < 3; ++d) {
                        for (long r = 0; r < N; ++r) gphij[d * N + r] = (kc[d] + G[3 * r + d]) * phij[r];
                        orc_fft3d((double *)(gphij + d * N), p->n, +1);
                    }
                    orc_fft3d((double *)phij, p->n, +1);
                    memset(hsub, 0, sizeof(cplx) * N);                         /* :300 */
                    for (int ib = 0; ib < p->nocc[iq]; ++ib) {                 /* :303 */
                        scatter(p, iq, occ + co + (long)ib * p->npw[iq], phiq, N); /* :310-313 */
                        if (bitc) {                                             /* :323-340 */
                            scatter(p, iq, occl + co + (long)ib * p->npw[iq], chiq, N);
                            orc_fft3d((double *)chiq, p->n, +1);
                        }
                        for (int d = 0; d < 3; ++d) {                           /* :342-346 */
                            for (long r = 0; r < N; ++r) gphiq[d * N + r] = (qc[d] + G[3 * r + d]) * phiq[r];
                            orc_fft3d((double *)(gphiq + d * N), p->n, +1);
                        }
                        orc_fft3d((double *)phiq, p->n, +1);                    /* :347 */
                        const cplx *chi = bitc ? chiq : phiq;
                        for (long r = 0; r < N; ++r) phiqj[r] = conj(chi[r]) * phij[r]; /* :350 */
                        orc_fft3d((double *)phiqj, p->n, -1);                   /* :351 */
                        for (int d = 0; d < 3; ++d) {                           /* :354-358 */
                            for (long r = 0; r < N; ++r) pgp[d * N + r] = phiqj[r] * duk[d * N + r];
                            orc_fft3d((double *)(pgp + d * N), p->n, +1);
                        }
                        const double f = p->filling[fo + ib];                   /* :360 */
                        for (long r = 0; r < N; ++r) {                          /* :363-369 */
                            cplx t = 0;
                            for (int d = 0; d < 3; ++d) t += pgp[d * N + r] * gphiq[d * N + r];
                            hsub[r] -= f * t;
                        }
                        for (int d = 0; d < 3; ++d) {                           /* :372-376 */
                            for (long r = 0; r < N; ++r) pgp[d * N + r] = conj(chi[r]) * gphij[d * N + r];
                            orc_fft3d((double *)(pgp + d * N), p->n, -1);
                        }
                        for (long r = 0; r < N; ++r) {                          /* :379-385 */
                            cplx t = phiqj[r] * V2[r];
                            for (int d = 0; d < 3; ++d) t += pgp[d * N + r] * duk[d * N + r];
                            phiqj[r] = t;
                        }
                        orc_fft3d((double *)phiqj, p->n, +1);                   /* :386 */
                        for (long r = 0; r < N; ++r) hsub[r] += f * phiqj[r] * phiq[r]; /* :387-388 */
                    }
                    orc_fft3d((double *)hsub, p->n, -1);                        /* :391 */
#pragma omp critical
                    {
                        cplx *out = hphi + ao + (long)jb * npwk;                /* :397-400 */
                        for (int l = 0; l < npwk; ++l) out[l] += factor * hsub[gi[l]];
                    }
                }
            }
            free(V2); free(duk); free(phij); free(gphij); free(phiq); free(chiq);
            free(gphiq); free(phiqj); free(pgp); free(hsub);
        }

        if (p->div_corr) {                                                      /* :409-513 */
            const double dc = div_corr_value(p, ik);
            const div_set D = div_states(p, ik);                                /* :440-479 */
            /* kVaux = sum_of_kVaux / (Nk V) (/2 for one spin) * A_long[s][s] (:117-126) */
            double kv[3];
            kvaux_of(p, ik, kv);
            for (int d = 0; d < 3; ++d) kv[d] *= p->A[0] / (p->num_kpoints * p->volume) / (p->nspin == 1 ? 2.0 : 1.0);
            const int use_kv = !p->jastrow_A_zero && (kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]) > 1e-8;
            double *kg = malloc(sizeof(double) * 3 * (npwk > 0 ? npwk : 1));
            for (int l = 0; l < npwk; ++l) kg_of(p, ik, l, kg + 3 * l);
            for (int jb = 0; jb < nb[ik]; ++jb) {
                cplx *out = hphi + ao + (long)jb * npwk;
                const cplx *pj = phi + ao + (long)jb * npwk;
                for (int b = 0; b < D.n; ++b) {
                    const double f = D.fill[b];
                    if (f < 1e-8) continue;                                     /* :463 */
                    const cplx *pk = D.phi + (long)b * npwk;
                    const cplx *ck = D.chi + (long)b * npwk;                    /* :476-479 */
                    cplx prod = zdotc(ck, pj, npwk);                            /* :481 */
                    if (use_kv)                                                 /* [2b_x1] :483-492, :169-180 */
                        for (int l = 0; l < npwk; ++l) {
                            const double kd = kv[0] * kg[3 * l] + kv[1] * kg[3 * l + 1] + kv[2] * kg[3 * l + 2];
                            out[l] -= f * kd * pk[l] * prod;
                        }
                    prod *= dc;                                                 /* :496 */
                    if (use_kv)                                                 /* [2b_x2] :498-507, :193-204 */
                        for (int l = 0; l < npwk; ++l) {
                            const double kd = kv[0] * kg[3 * l] + kv[1] * kg[3 * l + 1] + kv[2] * kg[3 * l + 2];
                            prod += conj(ck[l]) * kd * pj[l];
                        }
                    prod *= f;                                                  /* :509 */
                    for (int l = 0; l < npwk; ++l) out[l] += prod * pk[l];      /* :510 */
                }
            }
            free(kg);
        }
    }
    free(G);
}

/* ------------------------------------------------------------------ */
/* kinetic, calc_hamiltonian_kinetic.cpp:6-52: H1 += |k+G|^2/2 phi      */
/* ------------------------------------------------------------------ */

void orc_kinetic(const orc_problem *p, const int *nb, const double *phi_d, double *hphi_d)
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const cplx *phi = (const cplx *)phi_d;
    cplx *hphi = (cplx *)hphi_d;
    double *G = malloc(sizeof(double) * 3 * N);
    grid_gcart(p, G);
    for (int ik = 0; ik < p->nk; ++ik) {
        const long ao = app_coef_offset(p, nb, ik);
        const int *gi = p->gindex + pw_offset(p, ik);
        double kc[3];
        kcart(p, p->kfrac + 3 * ik, kc);
        for (int jb = 0; jb < nb[ik]; ++jb)
            for (int l = 0; l < p->npw[ik]; ++l) {
                double k2 = 0;
                for (int d = 0; d < 3; ++d) {
                    double c = kc[d] + G[3 * (long)gi[l] + d];
                    k2 += c * c;
                }
                long o = ao + (long)jb * p->npw[ik] + l;
                hphi[o] += 0.5 * k2 * phi[o]; /* :44-45 */
            }
    }
    free(G);
}

/* ------------------------------------------------------------------ */
/* local potential applied to phi on the FFT grid                       */
/* ------------------------------------------------------------------ */

static void apply_local_potential(const orc_problem *p, const int *nb, const double *phi_d,
                                  double *hphi_d, const cplx *V, double factor);

/* hf2h, calc_hamiltonian_hf2h.cpp:8-85: V_H = FFT_backward[4 pi n(G)/G^2], G=0 dropped */
void orc_hf2h(const orc_problem *p, const int *nb, const double *phi_d, double *hphi_d)
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const cplx *dens = (const cplx *)p->density;
    double *G = malloc(sizeof(double) * 3 * N);
    grid_gcart(p, G);
    cplx *V = malloc(sizeof(cplx) * N);
    for (long r = 0; r < N; ++r) V[r] = dens[r] + dens[N + r];     /* :28 */
    orc_fft3d((double *)V, p->n, -1);                                /* :29 */
    V[0] = 0.0;                                                      /* :30 */
    for (long r = 1; r < N; ++r) {                                   /* :31-38 */
        double g2 = G[3 * r] * G[3 * r] + G[3 * r + 1] * G[3 * r + 1] + G[3 * r + 2] * G[3 * r + 2];
        V[r] *= FOURPI / g2;
    }
    orc_fft3d((double *)V, p->n, +1);                                /* :39 */
    apply_local_potential(p, nb, phi_d, hphi_d, V, 1.0 / (p->num_kpoints * p->volume)); /* :23, :55-82 */
    free(G); free(V);
}

/* tc2h, calc_hamiltonian_tc2h.cpp:10-236 (one independent spin) */
void orc_tc2h(const orc_problem *p, const int *nb, const double *phi_d, double *hphi_d)
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const cplx *dens = (const cplx *)p->density;
    const cplx *phi = (const cplx *)phi_d;
    cplx *hphi = (cplx *)hphi_d;
    const cplx *occ = (const cplx *)p->phi_occ;
    const int bitc = p->method == 2;
    const cplx *occl = (bitc && p->chi_occ) ? (const cplx *)p->chi_occ : occ;
    const double factor = 1.0 / (p->num_kpoints * p->volume);        /* :32 */
    double *G = malloc(sizeof(double) * 3 * N);
    grid_gcart(p, G);
    cplx *orb = malloc(sizeof(cplx) * N), *gorb = malloc(sizeof(cplx) * 3 * N);
    cplx *pgp = calloc(3 * N, sizeof(cplx)), *V2 = calloc(N, sizeof(cplx)), *dnu = calloc(3 * N, sizeof(cplx));
    /* phiq_gphiq_d = sum_{q,i} f conj(chi_qi) FFT_backward((q+G)_d phi_qi)  (:60-127) */
    for (int iq = 0; iq < p->nk; ++iq) {
        const long co = occ_coef_offset(p, iq), fo = occ_offset(p, iq);
        double qc[3];
        kcart(p, p->kfrac + 3 * iq, qc);
        for (int ib = 0; ib < p->nocc[iq]; ++ib) {
            scatter(p, iq, occ + co + (long)ib * p->npw[iq], orb, N);
            for (int d = 0; d < 3; ++d)
                for (long r = 0; r < N; ++r) gorb[d * N + r] = (qc[d] + G[3 * r + d]) * orb[r];
            scatter(p, iq, occl + co + (long)ib * p->npw[iq], orb, N);
            orc_fft3d((double *)orb, p->n, +1);
            const double f = p->filling[fo + ib];
            for (long r = 0; r < N; ++r) orb[r] = f * conj(orb[r]);
            for (int d = 0; d < 3; ++d) {
                orc_fft3d((double *)(gorb + d * N), p->n, +1);
                for (long r = 0; r < N; ++r) pgp[d * N + r] += orb[r] * gorb[d * N + r];
            }
        }
    }
    for (int d = 0; d < 3; ++d) orc_fft3d((double *)(pgp + d * N), p->n, -1); /* :125 */
    const double spin_factor = p->nspin == 1 ? 2.0 : 1.0;           /* :145 */
    for (int js = 0; js < 2; ++js) {                                 /* :146-171 */
        memcpy(orb, dens + js * N, sizeof(cplx) * N);
        orc_fft3d((double *)orb, p->n, -1);
        for (long r = 0; r < N; ++r) {
            double g2 = G[3 * r] * G[3 * r] + G[3 * r + 1] * G[3 * r + 1] + G[3 * r + 2] * G[3 * r + 2];
            V2[r] += orc_tc_2body(p, g2, 0, js) * orb[r];
            double u = orc_uk(p, g2, 0, js);
            for (int d = 0; d < 3; ++d) {
                V2[r] += u * G[3 * r + d] * pgp[d * N + r] / spin_factor;
                dnu[d * N + r] += u * G[3 * r + d] * orb[r];
            }
        }
    }
    orc_fft3d((double *)V2, p->n, +1);                                /* :172 */
    for (int d = 0; d < 3; ++d) orc_fft3d((double *)(dnu + d * N), p->n, +1);
    cplx *phij = malloc(sizeof(cplx) * N);
    for (int ik = 0; ik < p->nk; ++ik) {                              /* :178-234 */
        const long ao = app_coef_offset(p, nb, ik);
        const int npwk = p->npw[ik];
        const int *gi = p->gindex + pw_offset(p, ik);
        double kc[3];
        kcart(p, p->kfrac + 3 * ik, kc);
        for (int jb = 0; jb < nb[ik]; ++jb) {
            scatter(p, ik, phi + ao + (long)jb * npwk, phij, N);
            for (int d = 0; d < 3; ++d) {
                for (long r = 0; r < N; ++r) gorb[d * N + r] = (kc[d] + G[3 * r + d]) * phij[r];
                orc_fft3d((double *)(gorb + d * N), p->n, +1);
            }
            orc_fft3d((double *)phij, p->n, +1);
            for (long r = 0; r < N; ++r) {
                cplx t = V2[r] * phij[r];
                for (int d = 0; d < 3; ++d) t -= gorb[d * N + r] * dnu[d * N + r];
                orb[r] = t;
            }
            orc_fft3d((double *)orb, p->n, -1);
            cplx *out = hphi + ao + (long)jb * npwk;
            for (int l = 0; l < npwk; ++l) out[l] += factor * orb[gi[l]];
        }
    }
    free(G); free(orb); free(gorb); free(pgp); free(V2); free(dnu); free(phij);
}

/* ------------------------------------------------------------------ */
/* tc3a1, calc_hamiltonian_tc3a1.cpp:8-117                              */
/* ------------------------------------------------------------------ */

static void apply_local_potential(const orc_problem *p, const int *nb, const double *phi_d,
                                  double *hphi_d, const cplx *V, double factor)
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const cplx *phi = (const cplx *)phi_d;
    cplx *hphi = (cplx *)hphi_d;
    cplx *orb = malloc(sizeof(cplx) * N);
    for (int ik = 0; ik < p->nk; ++ik) {
        const long ao = app_coef_offset(p, nb, ik);
        const int npwk = p->npw[ik];
        const int *gi = p->gindex + pw_offset(p, ik);
        for (int jb = 0; jb < nb[ik]; ++jb) {
            scatter(p, ik, phi + ao + (long)jb * npwk, orb, N);
            orc_fft3d((double *)orb, p->n, +1);
            for (long r = 0; r < N; ++r) orb[r] *= V[r];
            orc_fft3d((double *)orb, p->n, -1);
            cplx *out = hphi + ao + (long)jb * npwk;
            for (int l = 0; l < npwk; ++l) out[l] += factor * orb[gi[l]];
        }
    }
    free(orb);
}

void orc_tc3a1(const orc_problem *p, const int *nb, const double *phi_d, double *hphi_d)
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const double NkV = p->num_kpoints * p->volume;
    const double factor = 0.5 / (NkV * NkV);                         /* :26-28 */
    const cplx *dens = (const cplx *)p->density;
    double *G = malloc(sizeof(double) * 3 * N);
    grid_gcart(p, G);
    cplx *dnu = calloc(3 * N, sizeof(cplx)), *nG = malloc(sizeof(cplx) * N);
    cplx *V = calloc(N, sizeof(cplx));
    const int s = 0;
    for (int js = 0; js < 2; ++js) {                                  /* :56-70 */
        memcpy(nG, dens + js * N, sizeof(cplx) * N);
        orc_fft3d((double *)nG, p->n, -1);
        for (long r = 0; r < N; ++r) {
            double g2 = G[3 * r] * G[3 * r] + G[3 * r + 1] * G[3 * r + 1] + G[3 * r + 2] * G[3 * r + 2];
            double u = orc_uk(p, g2, s, js);
            for (int d = 0; d < 3; ++d) dnu[d * N + r] += u * G[3 * r + d] * nG[r];
        }
    }
    for (int d = 0; d < 3; ++d) orc_fft3d((double *)(dnu + d * N), p->n, +1); /* :71-74 */
    for (int d = 0; d < 3; ++d)                                       /* :78-84 */
        for (long r = 0; r < N; ++r) V[r] += dnu[d * N + r] * dnu[d * N + r];
    apply_local_potential(p, nb, phi_d, hphi_d, V, factor);          /* :86-116 */
    free(G); free(dnu); free(nG); free(V);
}

/* ------------------------------------------------------------------ */
/* tc3b1, calc_hamiltonian_tc3b1.cpp:8-135                              */
/* ------------------------------------------------------------------ */

void orc_tc3b1(const orc_problem *p, const int *nb, const double *phi_d, double *hphi_d)
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const double NkV = p->num_kpoints * p->volume;
    const double factor = -1.0 / (NkV * NkV);                        /* :26-28 */
    const cplx *dens = (const cplx *)p->density;
    double *G = malloc(sizeof(double) * 3 * N);
    grid_gcart(p, G);
    cplx *dnu = calloc(3 * N, sizeof(cplx)), *nG = malloc(sizeof(cplx) * N);
    cplx *V = calloc(N, sizeof(cplx));
    const int s2 = 0; /* ispin2 < num_independent_spins == 1 */
    for (int s3 = 0; s3 < 2; ++s3) {                                  /* :55-70 */
        memcpy(nG, dens + s3 * N, sizeof(cplx) * N);
        orc_fft3d((double *)nG, p->n, -1);
        for (long r = 0; r < N; ++r) {
            double g2 = G[3 * r] * G[3 * r] + G[3 * r + 1] * G[3 * r + 1] + G[3 * r + 2] * G[3 * r + 2];
            double u = orc_uk(p, g2, s2, s3);
            for (int d = 0; d < 3; ++d) dnu[d * N + r] += u * G[3 * r + d] * nG[r];
        }
    }
    for (int d = 0; d < 3; ++d) {                                     /* :71-82 */
        orc_fft3d((double *)(dnu + d * N), p->n, +1);
        for (long r = 0; r < N; ++r) dnu[d * N + r] *= dens[s2 * N + r];
        orc_fft3d((double *)(dnu + d * N), p->n, -1);
    }
    const int s1 = 0;
    for (int s2b = 0; s2b < 2; ++s2b) {                               /* :86-100 (ispin2_ref = 0) */
        for (long r = 0; r < N; ++r) {
            double g2 = G[3 * r] * G[3 * r] + G[3 * r + 1] * G[3 * r + 1] + G[3 * r + 2] * G[3 * r + 2];
            double u = orc_uk(p, g2, s1, s2b);
            for (int d = 0; d < 3; ++d) V[r] += u * G[3 * r + d] * dnu[d * N + r];
        }
    }
    orc_fft3d((double *)V, p->n, +1);                                 /* :101 */
    apply_local_potential(p, nb, phi_d, hphi_d, V, factor);          /* :104-134 */
    free(G); free(dnu); free(nG); free(V);
}

/* ------------------------------------------------------------------ */
/* tc3a3b3b4b6, calc_hamiltonian_tc3a3b3b4b6.cpp:11-863                 */
/* (kVaux branches :537-627, :772-858 act at k off the weighted mesh)   */
/* ------------------------------------------------------------------ */

static void dnu_grids(const orc_problem *p, const double *G, cplx *dnu) /* :128-150 */
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const cplx *dens = (const cplx *)p->density;
    cplx *nG = malloc(sizeof(cplx) * N);
    memset(dnu, 0, sizeof(cplx) * 3 * N);
    for (int s3 = 0; s3 < 2; ++s3) {
        memcpy(nG, dens + s3 * N, sizeof(cplx) * N);
        orc_fft3d((double *)nG, p->n, -1);
        for (long r = 0; r < N; ++r) {
            double g2 = G[3 * r] * G[3 * r] + G[3 * r + 1] * G[3 * r + 1] + G[3 * r + 2] * G[3 * r + 2];
            double u = orc_uk(p, g2, 0, s3);
            for (int d = 0; d < 3; ++d) dnu[d * N + r] += u * G[3 * r + d] * nG[r];
        }
    }
    for (int d = 0; d < 3; ++d) orc_fft3d((double *)(dnu + d * N), p->n, +1);
    free(nG);
}

/* du[s][d](r) = (kq + G)_d * uk(kq + G, s, 0) for s = 0, 1 on the grid */
static void du_grids(const orc_problem *p, const double *G, const double kq[3], int nspin2, int swap, double *du)
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    for (long r = 0; r < N; ++r) {
        double x[3], g2 = 0;
        for (int d = 0; d < 3; ++d) {
            x[d] = kq[d] + G[3 * r + d];
            g2 += x[d] * x[d];
        }
        for (int s = 0; s < nspin2; ++s) {
            double u = swap ? orc_uk(p, g2, 0, s) : orc_uk(p, g2, s, 0);
            for (int d = 0; d < 3; ++d) du[(s * 3 + d) * N + r] = x[d] * u;
        }
    }
}

static void orbital_grid(const orc_problem *p, int ik, const cplx *c, cplx *grid, long N)
{
    scatter(p, ik, c, grid, N);
    orc_fft3d((double *)grid, p->n, +1);
}

void orc_tc3a3b3b4b6(const orc_problem *p, const int *nb, const double *phi_d, double *hphi_d, int nthreads)
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const cplx *phi = (const cplx *)phi_d;
    cplx *hphi = (cplx *)hphi_d;
    const cplx *dens = (const cplx *)p->density;
    const cplx *occ = (const cplx *)p->phi_occ;
    const int bitc = p->method == 2;
    const cplx *occl = (bitc && p->chi_occ) ? (const cplx *)p->chi_occ : occ;
    const double NkV = p->num_kpoints * p->volume;
    const double base = 1.0 / (NkV * NkV);                            /* :33-35 */
    double fa3 = -base, fb3 = base, fb4 = -base, fb6 = base;          /* :37-47 */
    if (p->nspin == 1) { fa3 /= 2.0; fb3 /= 2.0; fb4 /= 4.0; fb6 /= 2.0; }
    const int need_div = p->div_corr && !p->jastrow_A_zero;          /* :99 */
    const int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
    double *G = malloc(sizeof(double) * 3 * N);
    grid_gcart(p, G);
    cplx *dnu = malloc(sizeof(cplx) * 3 * N);
    dnu_grids(p, G, dnu);                                             /* [3a3,b3] :126-151 */

    for (int ik = 0; ik < p->nk; ++ik) {
        const long ao = app_coef_offset(p, nb, ik);
        const int npwk = p->npw[ik];
        const int *gi = p->gindex + pw_offset(p, ik);
        double kc[3];
        kcart(p, p->kfrac + 3 * ik, kc);
        cplx *qsum = calloc((size_t)nb[ik] * 3 * N, sizeof(cplx));  /* qsum_temp_3a3b4 :162-172 */

        /* phase 1: [3b3], [3a3,b4,b6] (:179-371) */
#pragma omp parallel num_threads(nth)
        {
            double *du = malloc(sizeof(double) * 6 * N);
            cplx *phij = malloc(sizeof(cplx) * N), *phiq = malloc(sizeof(cplx) * N), *chiq = malloc(sizeof(cplx) * N);
            cplx *pp = malloc(sizeof(cplx) * N), *orb = malloc(sizeof(cplx) * N), *orb2 = malloc(sizeof(cplx) * N);
            cplx *dnuphi = malloc(sizeof(cplx) * 3 * N), *hsub = malloc(sizeof(cplx) * N);
#pragma omp for schedule(dynamic)
            for (int iq = 0; iq < p->nk; ++iq) {
                if (p->nocc[iq] == 0 || zero_weight(p, iq)) continue;  /* :201 */
                double qc[3], kq[3];
                kcart(p, p->kfrac + 3 * iq, qc);
                for (int d = 0; d < 3; ++d) kq[d] = kc[d] - qc[d];
                du_grids(p, G, kq, 2, 0, du);                          /* :207-215 */
                const long co = occ_coef_offset(p, iq), fo = occ_offset(p, iq);
                for (int jb = 0; jb < nb[ik]; ++jb) {
                    orbital_grid(p, ik, phi + ao + (long)jb * npwk, phij, N);  /* :223-228 */
                    for (int d = 0; d < 3; ++d)                                 /* :231-235 */
                        for (long r = 0; r < N; ++r) dnuphi[d * N + r] = dnu[d * N + r] * phij[r];
                    memset(hsub, 0, sizeof(cplx) * N);
                    for (int ib = 0; ib < p->nocc[iq]; ++ib) {
                        orbital_grid(p, iq, occ + co + (long)ib * p->npw[iq], phiq, N);
                        const cplx *chi = phiq;
                        if (bitc) {
                            orbital_grid(p, iq, occl + co + (long)ib * p->npw[iq], chiq, N);
                            chi = chiq;
                        }
                        for (long r = 0; r < N; ++r) pp[r] = conj(chi[r]) * phij[r];   /* :280-281 */
                        orc_fft3d((double *)pp, p->n, -1);
                        /* [3b3] :284-303 */
                        memset(orb2, 0, sizeof(cplx) * N);
                        for (int d = 0; d < 3; ++d) {
                            for (long r = 0; r < N; ++r) orb[r] = dnuphi[d * N + r] * conj(chi[r]);
                            orc_fft3d((double *)orb, p->n, -1);
                            for (long r = 0; r < N; ++r) orb2[r] += du[(0 * 3 + d) * N + r] * orb[r];
                        }
                        orc_fft3d((double *)orb2, p->n, +1);
                        const double f = p->filling[fo + ib];
                        for (long r = 0; r < N; ++r) hsub[r] += (f * fb3) * orb2[r] * phiq[r];
                        /* [3a3,b4,b6] :305-349 */
                        for (long r = 0; r < N; ++r) pp[r] *= f;
                        memset(orb2, 0, sizeof(cplx) * N);
                        for (int s2 = 0; s2 < 2; ++s2)
                            for (int d = 0; d < 3; ++d) {
                                const double *dus = du + (s2 * 3 + d) * N;
                                for (long r = 0; r < N; ++r) orb[r] = pp[r] * dus[r];
                                orc_fft3d((double *)orb, p->n, +1);
                                if (s2 == 0) {
#pragma omp critical
                                    {
                                        cplx *qs = qsum + ((size_t)jb * 3 + d) * N;
                                        for (long r = 0; r < N; ++r) qs[r] += orb[r] * phiq[r];
                                    }
                                }
                                for (long r = 0; r < N; ++r) orb[r] *= dens[s2 * N + r];
                                orc_fft3d((double *)orb, p->n, -1);
                                for (long r = 0; r < N; ++r) orb2[r] += orb[r] * dus[r];
                            }
                        orc_fft3d((double *)orb2, p->n, +1);                  /* :352-358 */
                        for (long r = 0; r < N; ++r) hsub[r] += fb6 * orb2[r] * phiq[r];
                    }
                    orc_fft3d((double *)hsub, p->n, -1);                      /* :363-372 */
#pragma omp critical
                    {
                        cplx *out = hphi + ao + (long)jb * npwk;
                        for (int l = 0; l < npwk; ++l) out[l] += hsub[gi[l]];
                    }
                }
            }
            free(du); free(phij); free(phiq); free(chiq); free(pp); free(orb); free(orb2);
            free(dnuphi); free(hsub);
        }

        /* per band: [3b6] divergence correction (:393-535), the kVaux terms of [3b3] and
         * [3a3,b4] (:537-627), [3a3] (:629-642) */
        const div_set D = div_states(p, ik);
        const int match = if_match(p, ik);                                       /* :396-410 */
        double kv3[3];                                     /* -sum_of_kVaux A_long[s][s] (:539-543) */
        kvaux_of(p, ik, kv3);
        for (int d = 0; d < 3; ++d) kv3[d] *= -p->A[0];
        const int use_kv = need_div && (kv3[0] * kv3[0] + kv3[1] * kv3[1] + kv3[2] * kv3[2]) > 1e-8;
        {
            cplx *phij = malloc(sizeof(cplx) * N), *phib = malloc(sizeof(cplx) * N), *chib = malloc(sizeof(cplx) * N);
            cplx *pp = malloc(sizeof(cplx) * N), *hsub = malloc(sizeof(cplx) * N);
            cplx *tG = malloc(sizeof(cplx) * N), *tR = malloc(sizeof(cplx) * N), *nG = malloc(sizeof(cplx) * N);
            for (int jb = 0; jb < nb[ik]; ++jb) {
                cplx *out = hphi + ao + (long)jb * npwk;
                orbital_grid(p, ik, phi + ao + (long)jb * npwk, phij, N);
                if (need_div) {
                    const double dcp = 1.0 / sqrt(PI * p->alpha) * NkV - p->sum_vaux[ik]; /* :415-419 */
                    double kvd[3];                                                 /* -sum_of_kVaux (:422-425) */
                    kvaux_of(p, ik, kvd);
                    cplx t1 = 0.0;
                    memset(tG, 0, sizeof(cplx) * N);
                    for (int js = 0; js < 2; ++js) {                              /* :427-450 */
                        memcpy(nG, dens + js * N, sizeof(cplx) * N);
                        orc_fft3d((double *)nG, p->n, -1);
                        const double Aj = p->A[0 * 2 + js];
                        double Adiv = dcp * Aj;
                        if (match) Adiv += orc_uk(p, 0.0, 0, js);
                        t1 += FOURPI * Adiv * Aj * nG[0];
                        for (long r = 1; r < N; ++r) {                            /* AuG :99-112, G=0 excluded */
                            double g2 = G[3 * r] * G[3 * r] + G[3 * r + 1] * G[3 * r + 1] + G[3 * r + 2] * G[3 * r + 2];
                            const double AuG = Aj * orc_uk(p, g2, 0, js);
                            if (match)
                                tG[r] += FOURPI * nG[r] * AuG;
                            else                                                  /* :445-448 */
                                tG[r] -= (kvd[0] * G[3 * r] + kvd[1] * G[3 * r + 1] + kvd[2] * G[3 * r + 2]) * nG[r] * AuG;
                        }
                    }
                    memcpy(tR, tG, sizeof(cplx) * N);
                    orc_fft3d((double *)tR, p->n, +1);                             /* :451 */
                    memset(hsub, 0, sizeof(cplx) * N);
                    for (int b = 0; b < D.n; ++b) {                                  /* :456-532 */
                        const double f = D.fill[b];
                        if (f < 1e-8) continue;
                        const cplx *pk = D.phi + (long)b * npwk;
                        orbital_grid(p, ik, pk, phib, N);
                        const cplx *chi = phib;
                        if (bitc) {
                            orbital_grid(p, ik, D.chi + (long)b * npwk, chib, N);
                            chi = chib;
                        }
                        for (long r = 0; r < N; ++r) pp[r] = chi[r] * conj(phij[r]);   /* :511-512 */
                        orc_fft3d((double *)pp, p->n, -1);
                        const cplx Ib = conj(pp[0]);
                        cplx dot = 0.0;                                                /* phiphi.dot(temp_G) */
                        for (long r = 0; r < N; ++r) dot += conj(pp[r]) * tG[r];
                        const double sign = match ? 1.0 : -1.0;                          /* :519 */
                        const cplx c = f * fb6 * t1 * Ib + sign * f * fb6 * dot;         /* contributions 1, 3 or 5 */
                        for (int l = 0; l < npwk; ++l) out[l] += c * pk[l];
                        for (long r = 0; r < N; ++r) hsub[r] += (f * Ib * fb6) * tR[r] * phib[r]; /* 2 or 4 */
                    }
                    orc_fft3d((double *)hsub, p->n, -1);
                    for (int l = 0; l < npwk; ++l) out[l] += hsub[gi[l]];
                }
                if (use_kv) {                                                      /* :545-623 */
                    for (int b = 0; b < D.n; ++b) {
                        const double f = D.fill[b];
                        if (f < 1e-8) continue;
                        const cplx *pk = D.phi + (long)b * npwk;
                        orbital_grid(p, ik, pk, phib, N);
                        const cplx *chi = phib;
                        if (bitc) {
                            orbital_grid(p, ik, D.chi + (long)b * npwk, chib, N);
                            chi = chib;
                        }
                        /* [3b3]: sum_d kVaux_d <chi_b|dnu_d phi_j> f fb3 / N, times phi_b (:596-609) */
                        cplx t3 = 0.0;
                        for (int d = 0; d < 3; ++d) {
                            cplx c = 0.0;
                            for (long r = 0; r < N; ++r) c += conj(chi[r]) * dnu[d * N + r] * phij[r];
                            t3 += c * (f * fb3 / (double)N) * kv3[d];
                        }
                        for (int l = 0; l < npwk; ++l) out[l] += t3 * pk[l];
                        /* [3a3,b4]: qsum_d += kVaux_d f <chi_b|phi_j> phi_b(r) (:612-620) */
                        const cplx pj = zdotc(D.chi + (long)b * npwk, phi + ao + (long)jb * npwk, npwk) * f;
                        for (int d = 0; d < 3; ++d) {
                            cplx *qs = qsum + ((size_t)jb * 3 + d) * N;
                            for (long r = 0; r < N; ++r) qs[r] += pj * kv3[d] * phib[r];
                        }
                    }
                }
                /* [3a3] */
                memset(hsub, 0, sizeof(cplx) * N);
                for (int d = 0; d < 3; ++d) {
                    const cplx *qs = qsum + ((size_t)jb * 3 + d) * N;
                    for (long r = 0; r < N; ++r) hsub[r] += qs[r] * dnu[d * N + r];
                }
                orc_fft3d((double *)hsub, p->n, -1);
                for (int l = 0; l < npwk; ++l) out[l] += fa3 * hsub[gi[l]];
            }
            free(phij); free(phib); free(chib); free(pp); free(hsub); free(tG); free(tR); free(nG);
        }

        /* phase 2: [3b4] (:646-770) */
#pragma omp parallel num_threads(nth)
        {
            double *du = malloc(sizeof(double) * 3 * N);
            cplx *phiq = malloc(sizeof(cplx) * N), *chiq = malloc(sizeof(cplx) * N);
            cplx *pp = malloc(sizeof(cplx) * N), *orb = malloc(sizeof(cplx) * N), *hsub = malloc(sizeof(cplx) * N);
#pragma omp for schedule(dynamic)
            for (int iq = 0; iq < p->nk; ++iq) {
                if (p->nocc[iq] == 0 || zero_weight(p, iq)) continue;
                double qc[3], kq[3];
                kcart(p, p->kfrac + 3 * iq, qc);
                for (int d = 0; d < 3; ++d) kq[d] = kc[d] - qc[d];
                du_grids(p, G, kq, 1, 0, du);                                  /* :680-687 */
                const long co = occ_coef_offset(p, iq), fo = occ_offset(p, iq);
                for (int jb = 0; jb < nb[ik]; ++jb) {
                    memset(hsub, 0, sizeof(cplx) * N);
                    for (int ib = 0; ib < p->nocc[iq]; ++ib) {
                        orbital_grid(p, iq, occ + co + (long)ib * p->npw[iq], phiq, N);
                        const cplx *chi = phiq;
                        if (bitc) {
                            orbital_grid(p, iq, occl + co + (long)ib * p->npw[iq], chiq, N);
                            chi = chiq;
                        }
                        memset(pp, 0, sizeof(cplx) * N);
                        for (int d = 0; d < 3; ++d) {                                   /* :740-748 */
                            const cplx *qs = qsum + ((size_t)jb * 3 + d) * N;
                            for (long r = 0; r < N; ++r) orb[r] = conj(chi[r]) * qs[r];
                            orc_fft3d((double *)orb, p->n, -1);
                            for (long r = 0; r < N; ++r) pp[r] += orb[r] * du[d * N + r];
                        }
                        orc_fft3d((double *)pp, p->n, +1);
                        const double f = p->filling[fo + ib];
                        for (long r = 0; r < N; ++r) hsub[r] += f * pp[r] * phiq[r];    /* :754 */
                    }
                    orc_fft3d((double *)hsub, p->n, -1);                                /* :757-766 */
#pragma omp critical
                    {
                        cplx *out = hphi + ao + (long)jb * npwk;
                        for (int l = 0; l < npwk; ++l) out[l] += fb4 * hsub[gi[l]];
                    }
                }
            }
            free(du); free(phiq); free(chiq); free(pp); free(orb); free(hsub);
        }
        if (use_kv) {                                                        /* [3b4] kVaux (:772-858) */
            cplx *chib = malloc(sizeof(cplx) * N);
            for (int jb = 0; jb < nb[ik]; ++jb) {
                cplx *out = hphi + ao + (long)jb * npwk;
                for (int b = 0; b < D.n; ++b) {
                    const double f = D.fill[b];
                    if (f < 1e-8) continue;
                    orbital_grid(p, ik, D.chi + (long)b * npwk, chib, N);  /* chi_b (phi_b for TC) */
                    cplx t4 = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        const cplx *qs = qsum + ((size_t)jb * 3 + d) * N;
                        cplx c = 0.0;
                        for (long r = 0; r < N; ++r) c += conj(chib[r]) * qs[r];
                        t4 += c * (fb4 * f / (double)N) * kv3[d];
                    }
                    const cplx *pk = D.phi + (long)b * npwk;
                    for (int l = 0; l < npwk; ++l) out[l] += t4 * pk[l];
                }
            }
            free(chib);
        }
        free(qsum);
    }
    free(G); free(dnu);
}

/* ------------------------------------------------------------------ */
/* tc3a2a4b2b5, calc_hamiltonian_tc3a2a4b2b5.cpp:11-1300                */
/* Identity symmetry: the U^+ transforms of :576-643 and :682-765 are   */
/* the identity (phase 1, Udag_cart = 1), so qsum_temp_3a4_trans is     */
/* FFT_backward(FFT_forward(qsum_temp_3a4)).  The kVaux branch           */
/* :1036-1299 acts at k off the weighted mesh only.                      */
/* ------------------------------------------------------------------ */

void orc_tc3a2a4b2b5(const orc_problem *p, const int *nb, const double *phi_d, double *hphi_d, int nthreads)
{
    const long N = (long)p->n[0] * p->n[1] * p->n[2];
    const cplx *phi = (const cplx *)phi_d;
    cplx *hphi = (cplx *)hphi_d;
    const cplx *occ = (const cplx *)p->phi_occ;
    const int bitc = p->method == 2;
    const cplx *occl = (bitc && p->chi_occ) ? (const cplx *)p->chi_occ : occ;
    const double NkV = p->num_kpoints * p->volume;
    const double base = 1.0 / (NkV * NkV);                            /* :34-36 */
    double fa2 = -0.5 * base, fa4 = base, fb2 = base, fb5 = -base;    /* :38-48 */
    if (p->nspin == 1) { fa2 /= 4.0; fa4 /= 4.0; fb2 /= 4.0; fb5 /= 4.0; }
    const int need_div = p->div_corr && !p->jastrow_A_zero;
    const int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
    double *G = malloc(sizeof(double) * 3 * N);
    grid_gcart(p, G);
    double *uG = malloc(sizeof(double) * 2 * N);                       /* uG[0][s] :120-131 */
    for (long r = 0; r < N; ++r) {
        double g2 = G[3 * r] * G[3 * r] + G[3 * r + 1] * G[3 * r + 1] + G[3 * r + 2] * G[3 * r + 2];
        for (int s = 0; s < 2; ++s) uG[s * N + r] = orc_uk(p, g2, 0, s);
    }
    cplx *V3a2 = calloc(N, sizeof(cplx));                              /* V_3a2 :113 */
    cplx *B2 = calloc(3 * N, sizeof(cplx));                            /* qsum_temp_3b2 :160-163 */
    /* flat (k, j) list for the parallel apply loops */
    int napp = 0;
    for (int ik = 0; ik < p->nk; ++ik) napp += nb[ik];
    int *app_k = malloc(sizeof(int) * (napp + 1)), *app_j = malloc(sizeof(int) * (napp + 1));
    for (int ik = 0, a = 0; ik < p->nk; ++ik)
        for (int jb = 0; jb < nb[ik]; ++jb, ++a) { app_k[a] = ik; app_j[a] = jb; }

    for (int iq2 = 0; iq2 < p->nk; ++iq2) {                            /* :165 */
        const int n2 = p->nocc[iq2];
        if (n2 == 0 || zero_weight(p, iq2)) continue;
        double q2c[3];
        kcart(p, p->kfrac + 3 * iq2, q2c);
        const long co2 = occ_coef_offset(p, iq2), fo2 = occ_offset(p, iq2);
        cplx *Q2 = calloc((size_t)n2 * N, sizeof(cplx));                /* qsum_temp_3a2 */
        cplx *Q4 = calloc((size_t)n2 * 3 * N, sizeof(cplx));            /* qsum_temp_3a4 */
        cplx *Q5 = bitc ? calloc((size_t)n2 * 3 * N, sizeof(cplx)) : NULL; /* qsum_temp_3b5 */

        /* (i) q1-q2 phase (:210-443) */
        double *du = malloc(sizeof(double) * 6 * N);
        for (int iq1 = 0; iq1 < p->nk; ++iq1) {
            if (p->nocc[iq1] == 0 || zero_weight(p, iq1)) continue;
            double q1c[3], v[3];
            kcart(p, p->kfrac + 3 * iq1, q1c);
            for (int d = 0; d < 3; ++d) v[d] = q2c[d] - q1c[d];
            du_grids(p, G, v, 2, 1, du);                                 /* :239-251 */
            const long co1 = occ_coef_offset(p, iq1), fo1 = occ_offset(p, iq1);
#pragma omp parallel for num_threads(nth) schedule(dynamic)
            for (int ib2 = 0; ib2 < n2; ++ib2) {
                cplx *phiq2 = malloc(sizeof(cplx) * N), *chiq2 = malloc(sizeof(cplx) * N);
                cplx *phiq1 = malloc(sizeof(cplx) * N), *chiq1 = malloc(sizeof(cplx) * N);
                cplx *pp = malloc(sizeof(cplx) * N), *pj = malloc(sizeof(cplx) * N);
                cplx *orb = malloc(sizeof(cplx) * N), *orb2 = malloc(sizeof(cplx) * N);
                orbital_grid(p, iq2, occ + co2 + (long)ib2 * p->npw[iq2], phiq2, N);
                const cplx *c2 = phiq2;
                if (bitc) {
                    orbital_grid(p, iq2, occl + co2 + (long)ib2 * p->npw[iq2], chiq2, N);
                    c2 = chiq2;
                }
                cplx *q2s = Q2 + (size_t)ib2 * N;
                for (int ib1 = 0; ib1 < p->nocc[iq1]; ++ib1) {
                    orbital_grid(p, iq1, occ + co1 + (long)ib1 * p->npw[iq1], phiq1, N);
                    const cplx *c1 = phiq1;
                    if (bitc) {
                        orbital_grid(p, iq1, occl + co1 + (long)ib1 * p->npw[iq1], chiq1, N);
                        c1 = chiq1;
                    }
                    for (long r = 0; r < N; ++r) pp[r] = conj(c1[r]) * phiq2[r];       /* :356-357 */
                    orc_fft3d((double *)pp, p->n, -1);
                    if (bitc) {                                                        /* :358-363 */
                        for (long r = 0; r < N; ++r) pj[r] = conj(phiq1[r]) * c2[r];
                        orc_fft3d((double *)pj, p->n, -1);
                    }
                    const double f1 = p->filling[fo1 + ib1];
                    for (int s = 0; s < 2; ++s)
                        for (int d = 0; d < 3; ++d) {
                            const double *dus = du + (s * 3 + d) * N;
                            for (long r = 0; r < N; ++r) orb[r] = dus[r] * pp[r];          /* :373-374 */
                            orc_fft3d((double *)orb, p->n, +1);
                            if (bitc) {                                                    /* :376-389 */
                                for (long r = 0; r < N; ++r) orb2[r] = dus[r] * pj[r];
                                orc_fft3d((double *)orb2, p->n, +1);
                                for (long r = 0; r < N; ++r) q2s[r] -= f1 * orb[r] * conj(orb2[r]);
                            } else {                                                       /* :390-401 */
                                for (long r = 0; r < N; ++r) {
                                    const double a = creal(orb[r]), b = cimag(orb[r]);
                                    q2s[r] -= f1 * (a * a + b * b);
                                }
                            }
                            if (s == 0) {                                                  /* :406-427 */
                                cplx *q4 = Q4 + ((size_t)ib2 * 3 + d) * N;
                                for (long r = 0; r < N; ++r) q4[r] += f1 * orb[r] * phiq1[r];
                                if (bitc) {
                                    cplx *q5 = Q5 + ((size_t)ib2 * 3 + d) * N;
                                    for (long r = 0; r < N; ++r) q5[r] -= f1 * conj(orb2[r]) * conj(c1[r]);
                                }
                            }
                        }
                }
                free(phiq2); free(chiq2); free(phiq1); free(chiq1); free(pp); free(pj); free(orb); free(orb2);
            }
        }
        free(du);

        /* (ii) per (q2, i2) (:445-915) */
        for (int ib2 = 0; ib2 < n2; ++ib2) {
            const double f2 = p->filling[fo2 + ib2];
            cplx *phiq2 = malloc(sizeof(cplx) * N), *chiq2 = malloc(sizeof(cplx) * N);
            orbital_grid(p, iq2, occ + co2 + (long)ib2 * p->npw[iq2], phiq2, N);
            const cplx *c2 = phiq2;
            if (bitc) {
                orbital_grid(p, iq2, occl + co2 + (long)ib2 * p->npw[iq2], chiq2, N);
                c2 = chiq2;
            }
            /* [3a2] (:486-573) */
            cplx *T = malloc(sizeof(cplx) * N);
            memcpy(T, Q2 + (size_t)ib2 * N, sizeof(cplx) * N);
            orc_fft3d((double *)T, p->n, -1);
            if (need_div) {
                const double dc2 = 1.0 / sqrt(PI * p->alpha) * NkV - p->sum_vaux[iq2];    /* :501-503 */
                cplx *phiq1 = malloc(sizeof(cplx) * N), *chiq1 = malloc(sizeof(cplx) * N);
                cplx *pa = malloc(sizeof(cplx) * N), *pb = malloc(sizeof(cplx) * N);
                for (int ib1 = 0; ib1 < n2; ++ib1) {                                     /* q1 = q2 */
                    orbital_grid(p, iq2, occ + co2 + (long)ib1 * p->npw[iq2], phiq1, N);
                    const cplx *c1 = phiq1;
                    if (bitc) {
                        orbital_grid(p, iq2, occl + co2 + (long)ib1 * p->npw[iq2], chiq1, N);
                        c1 = chiq1;
                    }
                    for (long r = 0; r < N; ++r) pa[r] = conj(c1[r]) * phiq2[r];          /* :545-547 */
                    orc_fft3d((double *)pa, p->n, -1);
                    for (long r = 0; r < N; ++r) pb[r] = conj(c2[r]) * phiq1[r];          /* :548-550 */
                    orc_fft3d((double *)pb, p->n, -1);
                    const double f1 = p->filling[fo2 + ib1];
                    for (int s = 0; s < 2; ++s) {                                         /* :553-569 */
                        const double As = p->A[0 * 2 + s];
                        T[0] -= FOURPI * As * (As * dc2 + orc_uk(p, 0.0, 0, s)) * f1 * pa[0] * pb[0];
                        const double ct = FOURPI * As * f1;
                        for (long r = 1; r < N; ++r)
                            T[r] -= ct * uG[s * N + r] * (pa[r] * pb[0] + pa[0] * pb[r]);
                    }
                }
                free(phiq1); free(chiq1); free(pa); free(pb);
            }
            for (long r = 0; r < N; ++r) V3a2[r] += f2 * T[r];                             /* :637-641 */
            free(T);

            /* [3a4,b2,b5] (:644-800) */
            cplx *Q4t = malloc(sizeof(cplx) * 3 * N), *Q5t = malloc(sizeof(cplx) * 3 * N);
            for (int d = 0; d < 3; ++d) {
                cplx *a = Q4t + d * N, *b = Q5t + d * N;
                const cplx *q4 = Q4 + ((size_t)ib2 * 3 + d) * N;
                memcpy(a, q4, sizeof(cplx) * N);
                if (bitc) memcpy(b, Q5 + ((size_t)ib2 * 3 + d) * N, sizeof(cplx) * N);
                else for (long r = 0; r < N; ++r) b[r] = -conj(q4[r]);                 /* :662-665 */
                orc_fft3d((double *)a, p->n, -1);
                orc_fft3d((double *)b, p->n, -1);
                orc_fft3d((double *)a, p->n, +1);                                          /* :755-759 */
                orc_fft3d((double *)b, p->n, +1);
            }
            for (int d = 0; d < 3; ++d)                                                   /* [3b2] :790-800 */
                for (long r = 0; r < N; ++r) B2[d * N + r] += f2 * Q5t[d * N + r] * phiq2[r];

#pragma omp parallel num_threads(nth)
            {
                double *du0 = malloc(sizeof(double) * 3 * N);
                cplx *phij = malloc(sizeof(cplx) * N), *pp = malloc(sizeof(cplx) * N);
                cplx *orb = malloc(sizeof(cplx) * N), *orb2 = malloc(sizeof(cplx) * N);
                int last_k = -1;
#pragma omp for schedule(static)
                for (int a = 0; a < napp; ++a) {                                          /* :801-912 */
                    const int ik = app_k[a], jb = app_j[a];
                    if (ik != last_k) {
                        double kc[3], kq[3];
                        kcart(p, p->kfrac + 3 * ik, kc);
                        for (int d = 0; d < 3; ++d) kq[d] = kc[d] - q2c[d];
                        du_grids(p, G, kq, 1, 0, du0);                                     /* uk_save :194-207, :826-832 */
                        last_k = ik;
                    }
                    const long ao = app_coef_offset(p, nb, ik);
                    const int npwk = p->npw[ik];
                    const int *gi = p->gindex + pw_offset(p, ik);
                    orbital_grid(p, ik, phi + ao + (long)jb * npwk, phij, N);
                    memset(orb2, 0, sizeof(cplx) * N);
                    for (long r = 0; r < N; ++r) pp[r] = conj(c2[r]) * phij[r];           /* [3a4] :849-861 */
                    orc_fft3d((double *)pp, p->n, -1);
                    for (long r = 0; r < N; ++r) pp[r] *= f2 * fa4;
                    for (int d = 0; d < 3; ++d) {
                        for (long r = 0; r < N; ++r) orb[r] = pp[r] * du0[d * N + r];
                        orc_fft3d((double *)orb, p->n, +1);
                        for (long r = 0; r < N; ++r) orb2[r] += Q4t[d * N + r] * orb[r];
                    }
                    memset(pp, 0, sizeof(cplx) * N);                                       /* [3b5] :863-875 */
                    for (int d = 0; d < 3; ++d) {
                        for (long r = 0; r < N; ++r) orb[r] = Q5t[d * N + r] * phij[r];
                        orc_fft3d((double *)orb, p->n, -1);
                        for (long r = 0; r < N; ++r) pp[r] += orb[r] * du0[d * N + r];
                    }
                    orc_fft3d((double *)pp, p->n, +1);
                    for (long r = 0; r < N; ++r) orb2[r] += fb5 * f2 * pp[r] * phiq2[r];
                    orc_fft3d((double *)orb2, p->n, -1);                                   /* :877-889 */
                    cplx *out = hphi + ao + (long)jb * npwk;
                    for (int l = 0; l < npwk; ++l) out[l] += orb2[gi[l]];
                }
                free(du0); free(phij); free(pp); free(orb); free(orb2);
            }
            free(Q4t); free(Q5t); free(phiq2); free(chiq2);
        }
        free(Q2); free(Q4); free(Q5);
    }

    /* [3b2] (:922-992) and [3a2] (:995-1034): two local potentials */
    cplx *X = calloc(N, sizeof(cplx));
    for (int d = 0; d < 3; ++d) {
        orc_fft3d((double *)(B2 + d * N), p->n, -1);
        for (long r = 0; r < N; ++r) X[r] += G[3 * r + d] * (uG[r] + uG[N + r]) * B2[d * N + r];
    }
    orc_fft3d((double *)X, p->n, +1);
    orc_fft3d((double *)V3a2, p->n, +1);
    apply_local_potential(p, nb, phi_d, hphi_d, X, fb2);
    apply_local_potential(p, nb, phi_d, hphi_d, V3a2, fa2);

    /* [3a4,b5] kVaux terms at k off the weighted mesh (:1036-1299).  The two |kVaux|^2
     * terms sit inside the (q1) loop, so they count once per q1 with occupied orbitals
     * (one process: the /num_mpi_processes of :1260-1261, :1278-1279 is 1). */
    if (need_div) {
        cplx *phiq2 = malloc(sizeof(cplx) * N), *chiq2 = malloc(sizeof(cplx) * N);
        cplx *phiq1 = malloc(sizeof(cplx) * N), *chiq1 = malloc(sizeof(cplx) * N);
        cplx *pp = malloc(sizeof(cplx) * N), *pj = malloc(sizeof(cplx) * N);
        cplx *orb = malloc(sizeof(cplx) * N), *orb2 = malloc(sizeof(cplx) * N), *phijG = malloc(sizeof(cplx) * N);
        double *duk = malloc(sizeof(double) * N);
        for (int ik = 0; ik < p->nk; ++ik) {
            double kv[3];
            kvaux_of(p, ik, kv);
            for (int d = 0; d < 3; ++d) kv[d] *= -p->A[0];                         /* :1064-1068 */
            const double kv2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
            if (!(kv2 > 1e-8) || nb[ik] == 0) continue;
            const div_set D = div_states(p, ik);
            const long ao = app_coef_offset(p, nb, ik);
            const int npwk = p->npw[ik];
            const int *gi = p->gindex + pw_offset(p, ik);
            double kc[3];
            kcart(p, p->kfrac + 3 * ik, kc);
            for (int iq1 = 0; iq1 < p->nk; ++iq1) {                                 /* :1073-1104 */
                if (p->nocc[iq1] == 0 || zero_weight(p, iq1)) continue;
                double q1c[3];
                kcart(p, p->kfrac + 3 * iq1, q1c);
                for (long r = 0; r < N; ++r) {                                      /* :1106-1110 */
                    double x[3], g2 = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        x[d] = kc[d] - q1c[d] + G[3 * r + d];
                        g2 += x[d] * x[d];
                    }
                    duk[r] = (kv[0] * x[0] + kv[1] * x[1] + kv[2] * x[2]) * orc_uk(p, g2, 0, 0);
                }
                const long co1 = occ_coef_offset(p, iq1), fo1 = occ_offset(p, iq1);
                for (int b2 = 0; b2 < D.n; ++b2) {                                  /* :1114-1152 */
                    const double f2 = D.fill[b2];
                    if (f2 < 1e-8) continue;
                    orbital_grid(p, ik, D.phi + (long)b2 * npwk, phiq2, N);
                    const cplx *c2 = phiq2;
                    if (bitc) {
                        orbital_grid(p, ik, D.chi + (long)b2 * npwk, chiq2, N);
                        c2 = chiq2;
                    }
                    memset(orb, 0, sizeof(cplx) * N);                               /* [3a4] */
                    memset(orb2, 0, sizeof(cplx) * N);                              /* [3b5] */
                    for (int ib1 = 0; ib1 < p->nocc[iq1]; ++ib1) {                  /* :1154-1227 */
                        orbital_grid(p, iq1, occ + co1 + (long)ib1 * p->npw[iq1], phiq1, N);
                        const cplx *c1 = phiq1;
                        if (bitc) {
                            orbital_grid(p, iq1, occl + co1 + (long)ib1 * p->npw[iq1], chiq1, N);
                            c1 = chiq1;
                        }
                        for (long r = 0; r < N; ++r) pp[r] = phiq2[r] * conj(c1[r]);
                        orc_fft3d((double *)pp, p->n, -1);
                        for (long r = 0; r < N; ++r) pj[r] = pp[r] * duk[r];
                        orc_fft3d((double *)pj, p->n, +1);
                        const double f1 = p->filling[fo1 + ib1];
                        for (long r = 0; r < N; ++r) orb[r] += f1 * pj[r] * phiq1[r];
                        if (bitc) {
                            for (long r = 0; r < N; ++r) pp[r] = c2[r] * conj(phiq1[r]);
                            orc_fft3d((double *)pp, p->n, -1);
                            for (long r = 0; r < N; ++r) pj[r] = pp[r] * duk[r];
                            orc_fft3d((double *)pj, p->n, +1);
                        }
                        for (long r = 0; r < N; ++r) orb2[r] -= f1 * pj[r] * c1[r];
                    }
                    orc_fft3d((double *)orb, p->n, -1);
                    orc_fft3d((double *)orb2, p->n, -1);
                    const cplx *pk2 = D.phi + (long)b2 * npwk;
                    const cplx *ck2 = D.chi + (long)b2 * npwk;
                    for (int jb = 0; jb < nb[ik]; ++jb) {                           /* :1231-1290 */
                        const cplx *pjc = phi + ao + (long)jb * npwk;
                        cplx *out = hphi + ao + (long)jb * npwk;
                        scatter(p, ik, pjc, phijG, N);
                        const cplx c2j = zdotc(ck2, pjc, npwk);
                        const cplx a1 = fa4 * f2 * c2j, a2 = fa4 * f2 * f2 * c2j * kv2;
                        for (int l = 0; l < npwk; ++l) out[l] += a1 * orb[gi[l]] + a2 * pk2[l];
                        cplx c5 = 0.0;
                        for (long r = 0; r < N; ++r) c5 += conj(orb2[r]) * phijG[r];
                        c5 = (c5 + f2 * c2j * kv2) * (fb5 * f2);
                        for (int l = 0; l < npwk; ++l) out[l] += c5 * pk2[l];
                    }
                }
            }
        }
        free(phiq2); free(chiq2); free(phiq1); free(chiq1); free(pp); free(pj); free(orb); free(orb2);
        free(phijG); free(duk);
    }
    free(G); free(uG); free(V3a2); free(B2); free(X); free(app_k); free(app_j);
}
