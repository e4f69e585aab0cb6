// kernels_local.hip -- density and the three-body terms that reduce to a local potential.
//
//   density_s(r) = sum_{k,i} f_ki conj(chi_ki(r)) phi_ki(r) (/2 for one independent spin)
//                  (src/bloch_states_scfloop_filling_density.cpp:256-358)
//   tc3a1: dnu_d = FFT_backward[sum_s' u(G,s,s') G_d n_s'(G)], V3 = sum_d dnu_d^2
//          (src/calc_hamiltonian_tc3a1.cpp:53-84)
//   tc3b1: dnu_d = FFT_forward[density_s2 * FFT_backward(sum_s3 u(G,s2,s3) G_d n_s3)],
//          V3 = FFT_backward[sum_{s2,d} u(G,s1,s2) G_d dnu_d]
//          (src/calc_hamiltonian_tc3b1.cpp:53-102)
// and H3 phi_kj += factor * FFT_forward[V3 phi_kj] gathered on the sphere of k.
// These are O(Nk Nb) transforms (no pair loop) and run in FP64.
#include "jastrow_dev.hpp"
#include "kernels_common.hpp"

namespace tcx {

__global__ __launch_bounds__(256) void k_density(const double2* __restrict__ occR, const double2* __restrict__ chiR,
                                                 const double* __restrict__ fill, int nocc, long long N,
                                                 double2* __restrict__ out, int nspin)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    double re = 0.0, im = 0.0;
    for (int b = 0; b < nocc; ++b) {
        const double2 p = occR[(long long)b * N + r];
        const double f = fill[b];
        if (chiR) {
            const double2 c = chiR[(long long)b * N + r];
            re += f * (c.x * p.x + c.y * p.y);
            im += f * (c.x * p.y - c.y * p.x);
        } else {
            re += f * (p.x * p.x + p.y * p.y);
        }
    }
    if (nspin == 2) {  // one spin channel, fillings up to 1
        out[r] = make_double2(re, im);
        return;
    }
    // one independent spin: the spin factor 2 is in the filling (filling_density.cpp:343, :356-358)
    out[r] = make_double2(0.5 * re, 0.5 * im);
    out[N + r] = make_double2(0.5 * re, 0.5 * im);
}

// out_d(G) = sum_s u(G, 0, s) G_d nG_s(G), s = 0, 1   (3 grids)
template <bool kPoly>
__global__ __launch_bounds__(256) void k_dnu(const double2* __restrict__ nG, double2* __restrict__ out, Geom g,
                                             JastrowPair j0, JastrowPair j1, long long N)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= N) return;
    const int i1 = (int)(e % g.n1), i2 = (int)((e / g.n1) % g.n2), i3 = (int)(e / ((long long)g.n1 * g.n2));
    const int m1 = centered(i1, g.n1), m2 = centered(i2, g.n2), m3 = centered(i3, g.n3);
    double G[3], G2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        G[d] = g.b[d] * m1 + g.b[3 + d] * m2 + g.b[6 + d] * m3;
        G2 += G[d] * G[d];
    }
    const double u0 = jastrow_uk<kPoly>(j0, G2), u1 = jastrow_uk<kPoly>(j1, G2);
    const double2 a = nG[e], b = nG[N + e];
    const double sx = u0 * a.x + u1 * b.x, sy = u0 * a.y + u1 * b.y;
    for (int d = 0; d < 3; ++d) out[(long long)d * N + e] = make_double2(G[d] * sx, G[d] * sy);
}

// V(G) = sum_{s2} sum_d u(G, s1, s2) G_d w_s2,d(G)  (tc3b1.cpp:83-100): dnu = w_0, dnu1 = w_1, or
// dnu1 = nullptr for one independent spin (s2_ref = 0: both s2 read w_0)
template <bool kPoly>
__global__ __launch_bounds__(256) void k_tc3b1_V(const double2* __restrict__ dnu, const double2* __restrict__ dnu1,
                                                 double2* __restrict__ V, Geom g, JastrowPair j0, JastrowPair j1,
                                                 long long N)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= N) return;
    const int i1 = (int)(e % g.n1), i2 = (int)((e / g.n1) % g.n2), i3 = (int)(e / ((long long)g.n1 * g.n2));
    const int m1 = centered(i1, g.n1), m2 = centered(i2, g.n2), m3 = centered(i3, g.n3);
    double G[3], G2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        G[d] = g.b[d] * m1 + g.b[3 + d] * m2 + g.b[6 + d] * m3;
        G2 += G[d] * G[d];
    }
    const double u0 = jastrow_uk<kPoly>(j0, G2), u1 = jastrow_uk<kPoly>(j1, G2);
    double re = 0.0, im = 0.0;
    if (dnu1 == nullptr) {  // one independent spin: both s2 read the one sum
        const double u = u0 + u1;
        for (int d = 0; d < 3; ++d) {
            const double2 x = dnu[(long long)d * N + e];
            re += u * G[d] * x.x;
            im += u * G[d] * x.y;
        }
    } else {
        for (int d = 0; d < 3; ++d) {
            const double2 x = dnu[(long long)d * N + e], y = dnu1[(long long)d * N + e];
            re += G[d] * (u0 * x.x + u1 * y.x);
            im += G[d] * (u0 * x.y + u1 * y.y);
        }
    }
    V[e] = make_double2(re, im);
}

// V(r) = sum_d dnu_d(r)^2 (complex square, tc3a1.cpp:78-84)
__global__ __launch_bounds__(256) void k_sum_squares(const double2* __restrict__ dnu, double2* __restrict__ V,
                                                     long long N)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= N) return;
    double re = 0.0, im = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double2 x = dnu[(long long)d * N + e];
        re += x.x * x.x - x.y * x.y;
        im += 2.0 * x.x * x.y;
    }
    V[e] = make_double2(re, im);
}

// grids[b][r] *= V[r] for every grid b (V shared)
__global__ __launch_bounds__(256) void k_mul_shared(double2* __restrict__ grids, const double2* __restrict__ V,
                                                    long long N)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= N) return;
    const double2 v = V[e];
    double2* p = grids + (long long)blockIdx.y * N + e;
    const double2 a = *p;
    *p = make_double2(a.x * v.x - a.y * v.y, a.x * v.y + a.y * v.x);
}

static dim3 grid1(long long N, unsigned y = 1) { return dim3((unsigned)((N + 255) / 256), y); }

void launch_density(const double* occR, const double* chiR, const double* fill, int nocc, long long N, double* out,
                    int nspin, hipStream_t s)
{
    hipLaunchKernelGGL(k_density, grid1(N), dim3(256), 0, s, (const double2*)occR, (const double2*)chiR, fill, nocc,
                       N, (double2*)out, nspin);
    TCX_HIP_CHECK(hipGetLastError());
}

// uk[i] = u(G), v[i] = tc_2body(G) at |G|^2 = G2[i] (Jastrow::uk, tc_2body)
__global__ __launch_bounds__(256) void k_jastrow_eval(const double* __restrict__ G2, double* __restrict__ uk,
                                                      double* __restrict__ v, JastrowPair j, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double g2 = G2[i];
    uk[i] = jastrow_uk(j, g2);
    v[i] = jastrow_tc2body(j, g2);
}

void launch_jastrow_eval(const double* G2, double* uk, double* v, long long n, const JastrowPair& j, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_jastrow_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, G2, uk, v, j, n);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_dnu(const double* nG, double* out, const Geom& g, const JastrowPair& j0, const JastrowPair& j1,
                hipStream_t s)
{
    const long long N = (long long)g.n1 * g.n2 * g.n3;
    TCX_POLY_DISPATCH(j0.pdeg > 0, hipLaunchKernelGGL((k_dnu<kPoly>), grid1(N), dim3(256), 0, s, (const double2*)nG,
                                                      (double2*)out, g, j0, j1, N));
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_tc3b1_V(const double* dnu, const double* dnu1, double* V, const Geom& g, const JastrowPair& j0,
                    const JastrowPair& j1, hipStream_t s)
{
    const long long N = (long long)g.n1 * g.n2 * g.n3;
    TCX_POLY_DISPATCH(j0.pdeg > 0, hipLaunchKernelGGL((k_tc3b1_V<kPoly>), grid1(N), dim3(256), 0, s,
                                                      (const double2*)dnu, (const double2*)dnu1, (double2*)V, g,
                                                      j0, j1, N));
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_sum_squares(const double* dnu, double* V, long long N, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_squares, grid1(N), dim3(256), 0, s, (const double2*)dnu, (double2*)V, N);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_mul_shared(double* grids, int ngrids, const double* V, long long N, hipStream_t s)
{
    for (int b0 = 0; b0 < ngrids; b0 += 65535) {
        const int nb = ngrids - b0 < 65535 ? ngrids - b0 : 65535;
        hipLaunchKernelGGL(k_mul_shared, grid1(N, nb), dim3(256), 0, s, (double2*)grids + (size_t)b0 * N,
                           (const double2*)V, N);
        TCX_HIP_CHECK(hipGetLastError());
    }
}

}  // namespace tcx
