// kernels_tc3x.hip -- building blocks of the three-body exchange terms
//   tc3a3b3b4b6 (src/calc_hamiltonian_tc3a3b3b4b6.cpp:11-863) and
//   tc3a2a4b2b5 (src/calc_hamiltonian_tc3a2a4b2b5.cpp:11-1300)
// and of the local assembly terms kinetic / hf2h / tc2h.
//
// Every (target, source) pair of a launch owns a record of C grids (complex
// FP32 or FP64, consecutive in HBM) that moves through
//   products+z record[c] = A_c(r) * B_c(r) formed in registers from the two
//              orbital grids chosen by the pair and transformed along z in the
//              same pass, then the xy planes: FFT_forward;
//   combine    a per-term G-space map that multiplies by the Jastrow gradient
//              du_d(G) = (kq+G)_d u(kq+G), evaluated in FP64 on the fly;
//   3D FFT     batched FFT_backward (xy planes, then z pencils);
//   accumulate one thread owns one point (or one z pencil, when the z pass of
//              FFT_backward is fused in) of one target and loops over the
//              target's pairs, summing in FP64 registers, then adds once into
//              the FP64 real-space grid of the target (no atomics).
// All point-wise kernels are one thread per grid point, 256 threads, with the
// record / target / band index in blockIdx.y (< 65536, the launchers split).
#include "jastrow_dev.hpp"
#include "kernels_common.hpp"
#include "size_switch.hpp"

namespace tcx {

namespace {

constexpr double kFourPi = 4.0 * 3.14159265358979323846;

__device__ __forceinline__ void miller_of(const Geom& g, long long e, int& m1, int& m2, int& m3)
{
    const int i1 = (int)(e % g.n1), i2 = (int)((e / g.n1) % g.n2), i3 = (int)(e / ((long long)g.n1 * g.n2));
    m1 = centered(i1, g.n1);
    m2 = centered(i2, g.n2);
    m3 = centered(i3, g.n3);
}

// x = kq + G of grid point e (kq may be null: G alone); returns |x|^2
__device__ __forceinline__ double kq_plus_G(const Geom& g, long long e, const double* kq, double x[3])
{
    int m1, m2, m3;
    miller_of(g, e, m1, m2, m3);
    double x2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        x[d] = (kq ? kq[d] : 0.0) + g.b[d] * m1 + g.b[3 + d] * m2 + g.b[6 + d] * m3;
        x2 += x[d] * x[d];
    }
    return x2;
}

__device__ __forceinline__ double2 dmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 dconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 dadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 dscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

template <typename T>
__device__ __forceinline__ double2 up(cpx<T> a)
{
    return make_double2((double)a.re, (double)a.im);
}
template <typename T>
__device__ __forceinline__ cpx<T> down(double2 a)
{
    return cpx<T>{(T)a.x, (T)a.y};
}

inline dim3 grid_pts(long long N, unsigned y) { return dim3((unsigned)((N + 255) / 256), y); }

// Run f(first, count) over [0, n) in slices that fit blockIdx.y.
template <typename F>
void for_slices(long long n, F&& f)
{
    for (long long a = 0; a < n; a += 65535) f((int)a, (unsigned)(n - a < 65535 ? n - a : 65535));
}

}  // namespace

// ----------------------------------------------------------------- products
// Products fused with the forward z pass: one thread owns pencil (x, y) of one record
// channel, forms A_c * B_c along z from the two orbital grids (shared by many records),
// transforms along z, scales, and stores; launch_fft_xy(-1) over the records completes
// FFT_forward.  Same records as k_products + launch_fft3d(-1).
//
// Workgroup order.  Workgroups go to the 8 XCDs in turn (linear id mod 8), each with its
// own L2.  Workgroup L takes XCD r = L mod 8, channel c = (L / 8) mod nchan and the
// (pencil group, pair) item u = 8 ((L / 8) / nchan) + r, pencil-group-major over the
// launch's target-major pairs.  So the channels of one pair and pencil group run back to
// back on one XCD (the source orbital A is read once into its L2), and an XCD takes every
// 8th pair of a pencil group (a target's B grids, and with 8 | sources the source's A
// across targets, are re-read from the same L2).
constexpr int kXcds = 8;

template <typename T, int NZ>
__global__ __launch_bounds__(kPencilBlock) void k_products_z(const PairEntry* __restrict__ pairs, int npairs,
                                                             const ProdChan* __restrict__ chans, int nchan,
                                                             int ngroups, cpx<T>* __restrict__ out, int nxy, T sc)
{
    const unsigned lin = blockIdx.x, r = lin % kXcds, rest = lin / kXcds;
    const int c = (int)(rest % (unsigned)nchan);
    const size_t u = (size_t)(rest / (unsigned)nchan) * kXcds + r;
    if (u >= (size_t)ngroups * npairs) return;
    const int grp = (int)(u / (size_t)npairs), p = (int)(u % (size_t)npairs);
    const int xy = grp * kPencilBlock + threadIdx.x;
    if (xy >= nxy) return;
    const int rec = p * nchan + c;
    const PairEntry e = pairs[p];
    const ProdChan ch = chans[c];
    const size_t ng = (size_t)nxy * NZ;
    const long long ia = (long long)(ch.a_sel ? e.occ : e.app) * ch.a_mul + ch.a_off;
    const long long ib = (long long)(ch.b_sel ? e.occ : e.app) * ch.b_mul + ch.b_off;
    const cpx<T>* a = (const cpx<T>*)ch.a + (size_t)ia * ng + xy;
    const cpx<T>* b = (const cpx<T>*)ch.b + (size_t)ib * ng + xy;
    cpx<T> v[NZ];
    static_for<0, NZ>([&](auto zc) {
        constexpr int z = decltype(zc)::value;
        cpx<T> x = a[(size_t)z * nxy], y = b[(size_t)z * nxy];
        if (ch.a_conj) x = conjg(x);
        if (ch.b_conj) y = conjg(y);
        v[z] = x * y;
    });
    fft<T, NZ, -1>(v);
    static_for<0, NZ>([&](auto zc) { v[decltype(zc)::value] = scale(v[decltype(zc)::value], sc); });
    store_pencil_nt<T, NZ>(v, out + (size_t)rec * ng + xy, nxy);  // streaming: read once by the xy pass
}

template <typename T>
static void products_fft_impl(const PairEntry* pairs, int npairs, const ProdChan* chans, int nchan, T* out,
                              const Geom& g, T sc, hipStream_t s)
{
    const int nxy = g.n1 * g.n2;
    const int ngroups = (nxy + kPencilBlock - 1) / kPencilBlock;
    const size_t items = (size_t)ngroups * npairs;
    const size_t nwg = (items + kXcds - 1) / kXcds * kXcds * nchan;
    if (nwg > 0x7fffffffu) throw Error(TCX_ERR_UNSUPPORTED, "products pass: too many workgroups in one launch");
    const dim3 grid((unsigned)nwg);
#define TCX_PZ(NZ)                                                                                                \
    hipLaunchKernelGGL((k_products_z<T, NZ>), grid, dim3(kPencilBlock), 0, s, pairs, npairs, chans, nchan, ngroups, \
                       (cpx<T>*)out, nxy, sc);
    TCX_SIZE_SWITCH(g.n3, TCX_PZ)
#undef TCX_PZ
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_products_z(int prec, const PairEntry* pairs, int npairs, const ProdChan* chans, int nchan, void* out,
                       const Geom& g, double sc, hipStream_t s)
{
    if (npairs == 0 || nchan == 0) return;
    if (prec == TCX_PREC_FP64)
        products_fft_impl<double>(pairs, npairs, chans, nchan, (double*)out, g, sc, s);
    else
        products_fft_impl<float>(pairs, npairs, chans, nchan, (float*)out, g, (float)sc, s);
}

void launch_products_fft(int prec, const PairEntry* pairs, int npairs, const ProdChan* chans, int nchan, void* out,
                         const Geom& g, double sc, hipStream_t s)
{
    launch_products_z(prec, pairs, npairs, chans, nchan, out, g, sc, s);
    launch_fft_xy(prec, -1, out, npairs * nchan, g, s);
}

// ----------------------------------------------------------------- tc3a3b3b4b6
// out[j][d](r) = dnu_d(r) * phi_j(r) in the working precision (dnu FP64, 3 grids)
template <typename T>
__global__ __launch_bounds__(256) void k_mul_dnu(const cpx<T>* __restrict__ appR, const double2* __restrict__ dnu,
                                                 cpx<T>* __restrict__ out, int j0, long long N)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    const int j = j0 + (int)blockIdx.y;
    const double2 ph = up(appR[(long long)j * N + r]);
    for (int d = 0; d < 3; ++d) out[((long long)j * 3 + d) * N + r] = down<T>(dmul(dnu[(long long)d * N + r], ph));
}

void launch_mul_dnu(int prec, const void* appR, const double* dnu, void* out, int napp, long long N, hipStream_t s)
{
    for_slices(napp, [&](int j0, unsigned ny) {
        if (prec == TCX_PREC_FP64)
            hipLaunchKernelGGL(k_mul_dnu<double>, grid_pts(N, ny), dim3(256), 0, s, (const cpx<double>*)appR,
                               (const double2*)dnu, (cpx<double>*)out, j0, N);
        else
            hipLaunchKernelGGL(k_mul_dnu<float>, grid_pts(N, ny), dim3(256), 0, s, (const cpx<float>*)appR,
                               (const double2*)dnu, (cpx<float>*)out, j0, N);
        TCX_HIP_CHECK(hipGetLastError());
    });
}

// [3a3] (tc3a3b3b4b6.cpp:629-642): W[j] += fa3 sum_d qsum[j][d] dnu_d
__global__ __launch_bounds__(256) void k_a3(double2* __restrict__ W, const double2* __restrict__ qsum,
                                            const double2* __restrict__ dnu, int j0, long long N, double fa3)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    const int j = j0 + (int)blockIdx.y;
    double2 acc = make_double2(0.0, 0.0);
    for (int d = 0; d < 3; ++d) acc = dadd(acc, dmul(qsum[((long long)j * 3 + d) * N + r], dnu[(long long)d * N + r]));
    double2* w = W + (long long)j * N + r;
    *w = dadd(*w, dscale(acc, fa3));
}

void launch_a3(double* W, const double* qsum, const double* dnu, int napp, long long N, double fa3, hipStream_t s)
{
    for_slices(napp, [&](int j0, unsigned ny) {
        hipLaunchKernelGGL(k_a3, grid_pts(N, ny), dim3(256), 0, s, (double2*)W, (const double2*)qsum,
                           (const double2*)dnu, j0, N, fa3);
        TCX_HIP_CHECK(hipGetLastError());
    });
}

// [3b6] divergence correction (tc3a3b3b4b6.cpp:424-451):
//   tG(G) = sum_js 4 pi n_js(G) A_js u(G, 0, js) (G != 0),  tG(0) = 0
template <bool kPoly>
__global__ __launch_bounds__(256) void k_b6_tG(const double2* __restrict__ nG, double2* __restrict__ tG, Geom g,
                                               long long N, JastrowPair j0, JastrowPair j1)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    double x[3];
    const double x2 = kq_plus_G(g, r, nullptr, x);
    if (r == 0) {
        tG[0] = make_double2(0.0, 0.0);
        return;
    }
    const double c0 = kFourPi * j0.A * jastrow_uk<kPoly>(j0, x2), c1 = kFourPi * j1.A * jastrow_uk<kPoly>(j1, x2);
    tG[r] = dadd(dscale(nG[r], c0), dscale(nG[N + r], c1));
}

//   t1[k] = sum_js 4 pi (dcp[k] A_js + u(0, 0, js) [k on the weighted mesh]) A_js n_js(G=0)
// (match null: every k is on the mesh)
template <bool kPoly>
__global__ void k_b6_t1(const double2* __restrict__ nG, const double* __restrict__ dcp, const int* __restrict__ match,
                        int nk, long long N, JastrowPair j0, JastrowPair j1, double2* __restrict__ t1)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const double m = (match == nullptr || match[k]) ? 1.0 : 0.0;
    const double a0 = kFourPi * (dcp[k] * j0.A + m * jastrow_uk<kPoly>(j0, 0.0)) * j0.A;
    const double a1 = kFourPi * (dcp[k] * j1.A + m * jastrow_uk<kPoly>(j1, 0.0)) * j1.A;
    t1[k] = dadd(dscale(nG[0], a0), dscale(nG[N], a1));
}

// One workgroup per applied band j at k (tc3a3b3b4b6.cpp:456-532), for every
// occupied band b at k with f_b >= 1e-8:
//   Ib    = <chi_b|phi_j>  (= conj(FFT_forward(chi_b conj(phi_j))(G=0)))
//   dot_b = sum_G conj(FFT_forward(chi_b conj(phi_j))) tG
//         = (1/N) sum_r conj(chi_b(r)) phi_j(r) tR(r)        (Parseval, tR = FFT_backward(tG))
//   out_j += f b6 (t1[k] Ib + sgn dot_b) phi_b               (contributions 1 and 3, or 5
//                                                             with sgn = -1 off the mesh)
//   cR[j][b] = f Ib b6                                       (contribution 2, real space)
constexpr int kB6MaxOcc = 512;

__global__ __launch_bounds__(256) void k_b6_coef(const double2* __restrict__ phi, const double2* __restrict__ occ_r,
                                                 const double2* __restrict__ occ_l, const DivEntry* __restrict__ jobs,
                                                 const double* __restrict__ fill, const double2* __restrict__ t1,
                                                 const double2* __restrict__ appR, const double2* __restrict__ chiR,
                                                 const double2* __restrict__ tR, double fb6, double sgn,
                                                 long long N, double2* __restrict__ out, double2* __restrict__ cR,
                                                 int maxocc)
{
    const int j = blockIdx.x;
    const DivEntry e = jobs[j];
    __shared__ double2 coef[kB6MaxOcc];
    __shared__ double red[2][256];
    const int t = threadIdx.x;
    const double2* pj = phi + e.app_off;
    const double2* aj = appR + (long long)j * N;
    const double invN = 1.0 / (double)N;
    for (int b = 0; b < e.nocc; ++b) {
        const double f = fill[e.fill_off + b];
        const double2* cb = occ_l + e.occ_off + (long long)b * e.npw;
        const double2* xb = chiR + (long long)(e.fill_off + b) * N;
        double2 ib = make_double2(0.0, 0.0), db = make_double2(0.0, 0.0);
        for (int l = t; l < e.npw; l += 256) ib = dadd(ib, dmul(dconj(cb[l]), pj[l]));
        for (long long r = t; r < N; r += 256) db = dadd(db, dmul(dmul(dconj(xb[r]), aj[r]), tR[r]));
        // block reduction of the two complex sums, one at a time through LDS
        double2 sums[2] = {ib, db};
        for (int q = 0; q < 2; ++q) {
            red[0][t] = sums[q].x;
            red[1][t] = sums[q].y;
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if (t < w) {
                    red[0][t] += red[0][t + w];
                    red[1][t] += red[1][t + w];
                }
                __syncthreads();
            }
            sums[q] = make_double2(red[0][0], red[1][0]);
            __syncthreads();
        }
        if (t == 0) {
            double2 c = make_double2(0.0, 0.0), cr = make_double2(0.0, 0.0);
            if (f >= 1e-8) {
                const double2 Ib = sums[0], dot = dscale(sums[1], sgn * invN);
                c = dscale(dadd(dmul(t1[e.kidx], Ib), dot), f * fb6);
                cr = dscale(Ib, f * fb6);
            }
            coef[b] = c;
            cR[(long long)j * maxocc + b] = cr;
        }
        __syncthreads();
    }
    double2* o = out + e.app_off;
    for (int l = t; l < e.npw; l += 256) {
        double2 acc = o[l];
        for (int b = 0; b < e.nocc; ++b) acc = dadd(acc, dmul(coef[b], occ_r[e.occ_off + (long long)b * e.npw + l]));
        o[l] = acc;
    }
}

// W[j](r) += tR(r) sum_b cR[j][b] phi_b(r)
__global__ __launch_bounds__(256) void k_b6_W(const double2* __restrict__ cR, int maxocc,
                                              const DivEntry* __restrict__ jobs, const double2* __restrict__ occR,
                                              const double2* __restrict__ tR, double2* __restrict__ W, int j0,
                                              long long N)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    const int j = j0 + (int)blockIdx.y;
    const DivEntry e = jobs[j];
    double2 acc = make_double2(0.0, 0.0);
    for (int b = 0; b < e.nocc; ++b)
        acc = dadd(acc, dmul(cR[(long long)j * maxocc + b], occR[(long long)(e.fill_off + b) * N + r]));
    double2* w = W + (long long)j * N + r;
    *w = dadd(*w, dmul(acc, tR[r]));
}

void launch_b6(const double* nG, double* tG, double* t1, const double* dcp, const int* match, int nk, const Geom& g,
               long long N, const JastrowPair& j0, const JastrowPair& j1, hipStream_t s)
{
    TCX_POLY_DISPATCH(j0.pdeg > 0, hipLaunchKernelGGL((k_b6_tG<kPoly>), grid_pts(N, 1), dim3(256), 0, s,
                                                      (const double2*)nG, (double2*)tG, g, N, j0, j1));
    TCX_HIP_CHECK(hipGetLastError());
    TCX_POLY_DISPATCH(j0.pdeg > 0, hipLaunchKernelGGL((k_b6_t1<kPoly>), dim3((nk + 63) / 64), dim3(64), 0, s,
                                                      (const double2*)nG, dcp, match, nk, N, j0, j1, (double2*)t1));
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_b6_coef(const double* phi, const double* occ_r, const double* occ_l, const DivEntry* jobs, int njobs,
                    int maxocc, const double* fill, const double* t1, const double* appR, const double* chiR,
                    const double* occR, const double* tR, double fb6, double sgn, long long N, double* out,
                    double* cR, double* W, hipStream_t s)
{
    if (njobs == 0) return;
    if (maxocc > kB6MaxOcc) throw Error(TCX_ERR_UNSUPPORTED, "tc3a3b3b4b6: more than 512 occupied bands at one k");
    hipLaunchKernelGGL(k_b6_coef, dim3(njobs), dim3(256), 0, s, (const double2*)phi, (const double2*)occ_r,
                       (const double2*)occ_l, jobs, fill, (const double2*)t1, (const double2*)appR,
                       (const double2*)chiR, (const double2*)tR, fb6, sgn, N, (double2*)out, (double2*)cR, maxocc);
    TCX_HIP_CHECK(hipGetLastError());
    for_slices(njobs, [&](int j0, unsigned ny) {
        hipLaunchKernelGGL(k_b6_W, grid_pts(N, ny), dim3(256), 0, s, (const double2*)cR, maxocc, jobs,
                           (const double2*)occR, (const double2*)tR, (double2*)W, j0, N);
        TCX_HIP_CHECK(hipGetLastError());
    });
}

// ----------------------------------------------------------------- tc3a2a4b2b5
// [3a2] divergence correction (tc3a2a4b2b5.cpp:496-573), per target (q2,i2), over
// the occupied i1 at q2 (rec[p] = FFT_forward(conj(chi_q2,i1) phi_q2,i2),
// FFT_forward(conj(chi_q2,i2) phi_q2,i1)):
//   T(0)  -= sum_s 4 pi A_s (A_s dc2[q2] + u_s(0)) f1 pa(0) pb(0)
//   T(G)  -= sum_s 4 pi A_s f1 u_s(G) (pa(G) pb(0) + pa(0) pb(G))
template <bool kPoly>
__global__ __launch_bounds__(256) void k_a8_div(double2* __restrict__ Tg, const double2* __restrict__ rec,
                                                const PairEntry* __restrict__ pairs,
                                                const TargetEntry* __restrict__ targets, int t0,
                                                const double* __restrict__ fill, const double* __restrict__ dc2,
                                                Geom g, long long N, JastrowPair j0, JastrowPair j1)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    const TargetEntry t = targets[t0 + blockIdx.y];
    double x[3];
    const double x2 = kq_plus_G(g, r, nullptr, x);
    const double dc = dc2[t.pad];
    const double w0 = r == 0 ? kFourPi * j0.A * (j0.A * dc + jastrow_uk<kPoly>(j0, 0.0)) : kFourPi * j0.A * jastrow_uk<kPoly>(j0, x2);
    const double w1 = r == 0 ? kFourPi * j1.A * (j1.A * dc + jastrow_uk<kPoly>(j1, 0.0)) : kFourPi * j1.A * jastrow_uk<kPoly>(j1, x2);
    double2 acc = make_double2(0.0, 0.0);
    for (int i = 0; i < t.npair; ++i) {
        const int p = t.pair0 + i;
        const double f1 = fill[pairs[p].occ];
        const double2* pa = rec + (long long)p * 2 * N;
        const double2* pb = pa + N;
        const double2 pa0 = pa[0], pb0 = pb[0];
        const double2 v = r == 0 ? dmul(pa0, pb0) : dadd(dmul(pa[r], pb0), dmul(pa0, pb[r]));
        acc = dadd(acc, dscale(v, f1 * (w0 + w1)));
    }
    double2* o = Tg + (long long)t.app * N + r;
    *o = make_double2(o->x - acc.x, o->y - acc.y);
}

// V3a2(G) = sum_t f_t T[t](G);  B2[d](r) = sum_t f_t Q5'[t][d](r) phi_t(r) with
// Q5' = Q5 (BITC) or -conj(Q4) (TC, tc3a2a4b2b5.cpp:662-665)  ([3a2] :637-641, [3b2] :790-800)
__global__ __launch_bounds__(256) void k_a8_reduce(const double2* __restrict__ Tg, const double2* __restrict__ Q4,
                                                   const double2* __restrict__ Q5, int bitc,
                                                   const double2* __restrict__ occR, const double* __restrict__ fill,
                                                   int nocc, double2* __restrict__ V3a2, double2* __restrict__ B2,
                                                   long long N)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    double2 v = make_double2(0.0, 0.0), b0 = v, b1 = v, b2 = v;
    for (int t = 0; t < nocc; ++t) {
        const double f = fill[t];
        v = dadd(v, dscale(Tg[(long long)t * N + r], f));
        const double2 ph = dscale(occR[(long long)t * N + r], f);
        double2 q[3];
        for (int d = 0; d < 3; ++d) {
            const long long o = ((long long)t * 3 + d) * N + r;
            q[d] = bitc ? Q5[o] : make_double2(-Q4[o].x, Q4[o].y);
        }
        b0 = dadd(b0, dmul(q[0], ph));
        b1 = dadd(b1, dmul(q[1], ph));
        b2 = dadd(b2, dmul(q[2], ph));
    }
    V3a2[r] = v;
    B2[r] = b0;
    B2[N + r] = b1;
    B2[2 * N + r] = b2;
}

// Q4T = Q4, Q5T = Q5' in the working precision (3 grids per occupied orbital)
template <typename T>
__global__ __launch_bounds__(256) void k_a8_q45(const double2* __restrict__ Q4, const double2* __restrict__ Q5,
                                                int bitc, cpx<T>* __restrict__ Q4T, cpx<T>* __restrict__ Q5T, int g0,
                                                long long N)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    const long long o = (long long)(g0 + blockIdx.y) * N + r;
    const double2 a = Q4[o];
    Q4T[o] = down<T>(a);
    Q5T[o] = down<T>(bitc ? Q5[o] : make_double2(-a.x, a.y));
}

// Symmetry rotation of q2-resolved sums from an irreducible q2 to a replica
// (tc3a2a4b2b5.cpp:576-640 scalar [3a2], :667-745 Cartesian vectors [3a4] [3b5]), in G space:
//   G' = s U^+ G wrapped into the grid, s = -1 with time reversal,
//   dst(G') = e^{2 pi i tau.G'} c(src(G))                           (scalar)
//   dst_d(G') = s e^{2 pi i tau.G'} sum_e U^+_cart(d,e) c(src_e(G))  (vector)
// with c = conj under time reversal.  One thread per source point; the map
// G -> G' is a bijection of the grid (checked when the symmetry is set), so the
// scattered stores of one job never collide, and jobs write distinct grids.
template <int NC>
__global__ __launch_bounds__(256) void k_sym_rotate(double2* __restrict__ Q, const SymRotJob* __restrict__ jobs,
                                                    int j0, Geom g, long long N)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    const SymRotJob& J = jobs[j0 + blockIdx.y];
    int G[3];
    miller_of(g, r, G[0], G[1], G[2]);
    const int s = J.trev ? -1 : 1;
    const int n[3] = {g.n1, g.n2, g.n3};
    int w[3];
    for (int i = 0; i < 3; ++i) {
        int v = s * (J.ud[3 * i] * G[0] + J.ud[3 * i + 1] * G[1] + J.ud[3 * i + 2] * G[2]);
        v %= n[i];
        w[i] = v < 0 ? v + n[i] : v;
    }
    double sn, cs;
    sincos(2.0 * 3.14159265358979323846 * (J.tau[0] * w[0] + J.tau[1] * w[1] + J.tau[2] * w[2]), &sn, &cs);
    const double2 ph = make_double2(cs, sn);
    const long long rd = w[0] + (long long)g.n1 * (w[1] + (long long)g.n2 * w[2]);
    if (NC == 1) {
        double2 v = Q[(long long)J.src * N + r];
        if (J.trev) v = dconj(v);
        Q[(long long)J.dst * N + rd] = dmul(ph, v);
    } else {
        double2 v[3];
        for (int e = 0; e < 3; ++e) {
            v[e] = Q[((long long)J.src * 3 + e) * N + r];
            if (J.trev) v[e] = dconj(v[e]);
        }
        for (int d = 0; d < 3; ++d) {
            double2 a = make_double2(0.0, 0.0);
            for (int e = 0; e < 3; ++e) a = dadd(a, dscale(v[e], J.udc[3 * d + e]));
            Q[((long long)J.dst * 3 + d) * N + rd] = dscale(dmul(ph, a), (double)s);
        }
    }
}

void launch_sym_rotate(double* Q, int ncomp, const SymRotJob* jobs, int njobs, const Geom& g, long long N,
                       hipStream_t s)
{
    for_slices(njobs, [&](int j0, unsigned ny) {
        if (ncomp == 1)
            hipLaunchKernelGGL(k_sym_rotate<1>, grid_pts(N, ny), dim3(256), 0, s, (double2*)Q, jobs, j0, g, N);
        else
            hipLaunchKernelGGL(k_sym_rotate<3>, grid_pts(N, ny), dim3(256), 0, s, (double2*)Q, jobs, j0, g, N);
        TCX_HIP_CHECK(hipGetLastError());
    });
}

void launch_a8_div(double* Tg, const double* rec, const PairEntry* pairs, const TargetEntry* targets, int ntargets,
                   const double* fill, const double* dc2, const Geom& g, long long N, const JastrowPair& j0,
                   const JastrowPair& j1, hipStream_t s)
{
    for_slices(ntargets, [&](int t0, unsigned ny) {
        TCX_POLY_DISPATCH(j0.pdeg > 0,
                          hipLaunchKernelGGL((k_a8_div<kPoly>), grid_pts(N, ny), dim3(256), 0, s, (double2*)Tg,
                                             (const double2*)rec, pairs, targets, t0, fill, dc2, g, N, j0, j1));
        TCX_HIP_CHECK(hipGetLastError());
    });
}

void launch_a8_reduce(const double* Tg, const double* Q4, const double* Q5, int bitc, const double* occR,
                      const double* fill, int nocc, double* V3a2, double* B2, long long N, hipStream_t s)
{
    hipLaunchKernelGGL(k_a8_reduce, grid_pts(N, 1), dim3(256), 0, s, (const double2*)Tg, (const double2*)Q4,
                       (const double2*)Q5, bitc, (const double2*)occR, fill, nocc, (double2*)V3a2, (double2*)B2, N);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_a8_q45(int prec, const double* Q4, const double* Q5, int bitc, void* Q4T, void* Q5T, int ngrids,
                   long long N, hipStream_t s)
{
    for_slices(ngrids, [&](int g0, unsigned ny) {
        if (prec == TCX_PREC_FP64)
            hipLaunchKernelGGL(k_a8_q45<double>, grid_pts(N, ny), dim3(256), 0, s, (const double2*)Q4,
                               (const double2*)Q5, bitc, (cpx<double>*)Q4T, (cpx<double>*)Q5T, g0, N);
        else
            hipLaunchKernelGGL(k_a8_q45<float>, grid_pts(N, ny), dim3(256), 0, s, (const double2*)Q4,
                               (const double2*)Q5, bitc, (cpx<float>*)Q4T, (cpx<float>*)Q5T, g0, N);
        TCX_HIP_CHECK(hipGetLastError());
    });
}

// ----------------------------------------------------------------- assembly terms
// kinetic (src/calc_hamiltonian_kinetic.cpp:38-46): out += 0.5 |k+G|^2 phi, one job per band
__global__ __launch_bounds__(256) void k_kinetic(const double2* __restrict__ phi, const int* __restrict__ gindex,
                                                 const OrbEntry* __restrict__ jobs, const double* __restrict__ kcart,
                                                 Geom g, double2* __restrict__ out)
{
    const OrbEntry e = jobs[blockIdx.y];
    const int l = blockIdx.x * 256 + threadIdx.x;
    if (l >= e.npw) return;
    double x[3];
    const double x2 = kq_plus_G(g, gindex[e.gidx_off + l], kcart + 3 * e.kidx, x);
    const long long o = e.coef_off + l;
    out[o] = dadd(out[o], dscale(phi[o], 0.5 * x2));
}

void launch_kinetic(const double* phi, const int* gindex, const OrbEntry* jobs, int njobs, int max_npw,
                    const double* kcart, const Geom& g, double* out, hipStream_t s)
{
    for_slices(njobs, [&](int j0, unsigned ny) {
        hipLaunchKernelGGL(k_kinetic, dim3((max_npw + 255) / 256, ny), dim3(256), 0, s, (const double2*)phi, gindex,
                           jobs + j0, kcart, g, (double2*)out);
        TCX_HIP_CHECK(hipGetLastError());
    });
}

// hf2h (src/calc_hamiltonian_hf2h.cpp:28-38): V(G) = 4 pi (n_0 + n_1)(G) / G^2, V(0) = 0
__global__ __launch_bounds__(256) void k_hartree(const double2* __restrict__ nG, double2* __restrict__ V, Geom g,
                                                 long long N)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    double x[3];
    const double x2 = kq_plus_G(g, r, nullptr, x);
    V[r] = r == 0 ? make_double2(0.0, 0.0) : dscale(dadd(nG[r], nG[N + r]), kFourPi / x2);
}

void launch_hartree(const double* nG, double* V, const Geom& g, long long N, hipStream_t s)
{
    hipLaunchKernelGGL(k_hartree, grid_pts(N, 1), dim3(256), 0, s, (const double2*)nG, (double2*)V, g, N);
    TCX_HIP_CHECK(hipGetLastError());
}

// tc2h (src/calc_hamiltonian_tc2h.cpp:60-124): pgp_d(r) = sum_{q,i} f conj(chi_qi(r)) gphi_qi,d(r)
__global__ __launch_bounds__(256) void k_pgp(const double2* __restrict__ chiR, const double2* __restrict__ gradR,
                                             const double* __restrict__ fill, int nocc, double2* __restrict__ pgp,
                                             long long N)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    double2 a0 = make_double2(0.0, 0.0), a1 = a0, a2 = a0;
    for (int b = 0; b < nocc; ++b) {
        const double2 c = dscale(dconj(chiR[(long long)b * N + r]), fill[b]);
        a0 = dadd(a0, dmul(c, gradR[((long long)b * 3 + 0) * N + r]));
        a1 = dadd(a1, dmul(c, gradR[((long long)b * 3 + 1) * N + r]));
        a2 = dadd(a2, dmul(c, gradR[((long long)b * 3 + 2) * N + r]));
    }
    pgp[r] = a0;
    pgp[N + r] = a1;
    pgp[2 * N + r] = a2;
}

// tc2h.cpp:146-171 for spin s (j0 = u(s, 0), j1 = u(s, 1)):
//   V2(G) = sum_js [tc_2body(G,s,js) n_js(G) + u(G,s,js) sum_d G_d pgp_js,d(G) / spin_factor]
// pgp0/pgp1: the <q|nabla|q> sums of spin 0 / 1 (one independent spin: both the one sum,
// pf = 1 / spin_factor = 1/2; two: pf = 1)
template <bool kPoly>
__global__ __launch_bounds__(256) void k_tc2h_V(const double2* __restrict__ nG, const double2* __restrict__ pgp0,
                                                const double2* __restrict__ pgp1, double pf,
                                                double2* __restrict__ V, Geom g, long long N, JastrowPair j0,
                                                JastrowPair j1)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    double x[3];
    const double x2 = kq_plus_G(g, r, nullptr, x);
    const double2 n0 = nG[r], n1 = nG[N + r];
    double2 v = dadd(dscale(n0, jastrow_tc2body<kPoly>(j0, x2)), dscale(n1, jastrow_tc2body<kPoly>(j1, x2)));
    const double u0 = pf * jastrow_uk<kPoly>(j0, x2), u1 = pf * jastrow_uk<kPoly>(j1, x2);
    for (int d = 0; d < 3; ++d)
        v = dadd(v, dadd(dscale(pgp0[(long long)d * N + r], u0 * x[d]), dscale(pgp1[(long long)d * N + r], u1 * x[d])));
    V[r] = v;
}

void launch_tc2h_V(const double* chiR, const double* gradR, const double* fill, int nocc, double* pgp,
                   const Geom& g, long long N, hipStream_t s)
{
    hipLaunchKernelGGL(k_pgp, grid_pts(N, 1), dim3(256), 0, s, (const double2*)chiR, (const double2*)gradR, fill,
                       nocc, (double2*)pgp, N);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_tc2h_combine(const double* nG, const double* pgp0, const double* pgp1, double pf, double* V,
                         const Geom& g, long long N, const JastrowPair& j0, const JastrowPair& j1, hipStream_t s)
{
    TCX_POLY_DISPATCH(j0.pdeg > 0, hipLaunchKernelGGL((k_tc2h_V<kPoly>), grid_pts(N, 1), dim3(256), 0, s,
                                                      (const double2*)nG, (const double2*)pgp0, (const double2*)pgp1,
                                                      pf, (double2*)V, g, N, j0, j1));
    TCX_HIP_CHECK(hipGetLastError());
}

// tc2h.cpp:212-217: out[b] = V2 phi_b - sum_d gphi_b,d dnu_d (bands of a chunk; grads 3 per band)
__global__ __launch_bounds__(256) void k_tc2h_apply(double2* __restrict__ orb, const double2* __restrict__ grad,
                                                    const double2* __restrict__ V2, const double2* __restrict__ dnu,
                                                    int b0, long long N)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    const int b = b0 + (int)blockIdx.y;
    double2* o = orb + (long long)b * N + r;
    double2 v = dmul(V2[r], *o);
    for (int d = 0; d < 3; ++d) {
        const double2 t = dmul(grad[((long long)b * 3 + d) * N + r], dnu[(long long)d * N + r]);
        v = make_double2(v.x - t.x, v.y - t.y);
    }
    *o = v;
}

void launch_tc2h_apply(double* orb, const double* grad, const double* V2, const double* dnu, int nb, long long N,
                       hipStream_t s)
{
    for_slices(nb, [&](int b0, unsigned ny) {
        hipLaunchKernelGGL(k_tc2h_apply, grid_pts(N, ny), dim3(256), 0, s, (double2*)orb, (const double2*)grad,
                           (const double2*)V2, (const double2*)dnu, b0, N);
        TCX_HIP_CHECK(hipGetLastError());
    });
}

// out = a X + b Y (FP64 grids)
__global__ __launch_bounds__(256) void k_axpby(const double2* __restrict__ X, const double2* __restrict__ Y,
                                               double2* __restrict__ out, double a, double b, long long N)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    out[r] = dadd(dscale(X[r], a), dscale(Y[r], b));
}

void launch_axpby(const double* X, const double* Y, double* out, double a, double b, long long N, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpby, grid_pts(N, 1), dim3(256), 0, s, (const double2*)X, (const double2*)Y,
                       (double2*)out, a, b, N);
    TCX_HIP_CHECK(hipGetLastError());
}

// W += C (W FP64, C FP32 complex grids)
__global__ __launch_bounds__(256) void k_add_into(double2* __restrict__ W, const float2* __restrict__ C, long long N)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    const float2 c = C[r];
    W[r] = dadd(W[r], make_double2((double)c.x, (double)c.y));
}

void launch_add_into(double* W, const float* C, long long N, hipStream_t s)
{
    if (N == 0) return;
    hipLaunchKernelGGL(k_add_into, grid_pts(N, 1), dim3(256), 0, s, (double2*)W, (const float2*)C, N);
    TCX_HIP_CHECK(hipGetLastError());
}

}  // namespace tcx

namespace tcx {

// ----------------------------------------------------------------- kernel-gradient reductions
// Backward pass of the occupied-set energies with respect to the two-body kernels.  For
// every (target kj, source qi) pair of occupied orbitals the forward pair densities are
// recomputed (FFT_forward of the products, in the working precision) and, with
// x = k - q + G and w = c f_kj f_qi / (2 Nk):
//   E_x : dE/dv(x)   = w |rho(G)|^2,                         rho = FFT(conj(phi_qi) phi_kj)
//   E_2x: dE/dV2(x)  = w conj(rho'(G)) rho(G)
//         dE/du(x)   = w sum_d x_d [conj(rho'(G)) sig_d(G) - conj(tau_d(G)) rho(G)]
//         rho = FFT(conj(chi_qi) phi_kj), sig_d = FFT(conj(chi_qi) g_kj,d),
//         rho' = FFT(conj(phi_qi) chi_kj), tau_d = FFT(conj(g_qi,d) chi_kj)
// (the G-space form of <chi_kj|tc2x phi_kj> in src/calc_hamiltonian_tc2x.cpp:253-401).  The
// values are summed into radial bins of |x| (width `width`, last bin open) through an LDS
// histogram per workgroup and FP64 global atomics (so the bin sums are not bit-reproducible
// between runs), together with two checks: sum_x K(x) dE/dK(x) (the kernel-linear part of
// the energy) and sum_x dK/dA(x) dE/dK(x) (dE/dA at fixed F).
constexpr int kMaxKernelBins = 1024;

template <typename T, bool kPoly>
__global__ __launch_bounds__(256) void k_kgrad(const cpx<T>* __restrict__ rec, int tc2, const PairEntry* __restrict__ pairs,
                                               int p0, const double* __restrict__ kqcart, const double* __restrict__ fill,
                                               Geom g, long long N, JastrowPair j0, JastrowPair j1, double w0,
                                               int nbins, double width, double* __restrict__ bins_v,
                                               double* __restrict__ bins_u, double* __restrict__ sums)
{
    __shared__ double hv[2 * kMaxKernelBins], hu[2 * kMaxKernelBins];
    __shared__ double red[4][256];
    const int t = threadIdx.x;
    for (int b = t; b < 2 * nbins; b += 256) {
        hv[b] = 0.0;
        hu[b] = 0.0;
    }
    __syncthreads();
    const long long r = (long long)blockIdx.x * 256 + t;
    const int p = p0 + (int)blockIdx.y;
    double s_lin[2] = {0.0, 0.0}, s_dA[2] = {0.0, 0.0};
    if (r < N) {
        const PairEntry e = pairs[p];
        double x[3];
        const double x2 = kq_plus_G(g, r, kqcart + 3 * e.kq, x);
        const int b = min(nbins - 1, (int)(sqrt(x2) / width));
        const double w = w0 * fill[e.app] * fill[e.occ];
        if (!tc2) {
            const double2 rho = up(rec[(long long)p * N + r]);
            const double sv = w * (rho.x * rho.x + rho.y * rho.y);
            atomicAdd(&hv[2 * b], sv);
            const double v = x2 < 1e-8 ? 0.0 : 1.0 / x2;  // hf2x.cpp:91-104
            s_lin[0] = v * sv;
        } else {
            const cpx<T>* in = rec + (long long)p * 8 * N + r;
            const double2 rho = up(in[0]), rhop = up(in[4 * N]);
            const double2 sV = dscale(dmul(dconj(rhop), rho), w);
            double2 sU = make_double2(0.0, 0.0);
            for (int d = 0; d < 3; ++d) {
                const double2 sig = up(in[(long long)(1 + d) * N]), tau = up(in[(long long)(5 + d) * N]);
                const double2 a = dmul(dconj(rhop), sig), c = dmul(dconj(tau), rho);
                sU = dadd(sU, dscale(make_double2(a.x - c.x, a.y - c.y), x[d]));
            }
            sU = dscale(sU, w);
            atomicAdd(&hv[2 * b], sV.x);
            atomicAdd(&hv[2 * b + 1], sV.y);
            atomicAdd(&hu[2 * b], sU.x);
            atomicAdd(&hu[2 * b + 1], sU.y);
            // spin pair (ispin, ispin) of tc2x.cpp:269-271
            JastrowPair jd = j0;
            jd.dA = 1;
            const double V2 = jastrow_tc2body<kPoly>(j0, x2), u = jastrow_uk<kPoly>(j0, x2);
            const double dV2 = jastrow_tc2body<kPoly>(jd, x2), du = jastrow_uk<kPoly>(jd, x2);
            s_lin[0] = V2 * sV.x + u * sU.x;
            s_lin[1] = V2 * sV.y + u * sU.y;
            s_dA[0] = dV2 * sV.x + du * sU.x;
            s_dA[1] = dV2 * sV.y + du * sU.y;
        }
    }
    red[0][t] = s_lin[0];
    red[1][t] = s_lin[1];
    red[2][t] = s_dA[0];
    red[3][t] = s_dA[1];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st)
            for (int q = 0; q < 4; ++q) red[q][t] += red[q][t + st];
        __syncthreads();
    }
    if (t < 4) atomicAdd(&sums[t], red[t][0]);
    for (int b = t; b < 2 * nbins; b += 256) {
        if (hv[b] != 0.0) atomicAdd(&bins_v[b], hv[b]);
        if (tc2 && hu[b] != 0.0) atomicAdd(&bins_u[b], hu[b]);
    }
}

void launch_kgrad(int prec, const void* rec, int tc2, const PairEntry* pairs, int npairs, const double* kqcart,
                  const double* fill, const Geom& g, long long N, const JastrowPair& j0, const JastrowPair& j1, double w0,
                  int nbins, double width, double* bins_v, double* bins_u, double* sums, hipStream_t s)
{
    if (nbins < 1 || nbins > kMaxKernelBins) throw Error(TCX_ERR_ARG, "kernel gradient: 1 <= nbins <= 1024");
    for_slices(npairs, [&](int p0, unsigned ny) {
        if (prec == TCX_PREC_FP64)
            TCX_POLY_DISPATCH(j0.pdeg > 0, hipLaunchKernelGGL((k_kgrad<double, kPoly>), grid_pts(N, ny), dim3(256), 0,
                                                              s, (const cpx<double>*)rec, tc2, pairs, p0, kqcart, fill,
                                                              g, N, j0, j1, w0, nbins, width, bins_v, bins_u, sums));
        else
            TCX_POLY_DISPATCH(j0.pdeg > 0, hipLaunchKernelGGL((k_kgrad<float, kPoly>), grid_pts(N, ny), dim3(256), 0,
                                                              s, (const cpx<float>*)rec, tc2, pairs, p0, kqcart, fill,
                                                              g, N, j0, j1, w0, nbins, width, bins_v, bins_u, sums));
        TCX_HIP_CHECK(hipGetLastError());
    });
}

}  // namespace tcx
