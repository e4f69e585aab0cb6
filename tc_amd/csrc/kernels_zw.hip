// kernels_zw.hip -- divergence terms at zero-weight k-points (BAND k-points and
// zero-weight SCF k-points): the kVaux branches of tc2x, tc3a3b3b4b6 and
// tc3a2a4b2b5 and the off-mesh [3b6] tables.
//
// At such a k the divergence terms project on the states given by
// tcx_set_zero_weight_kpoints (phik_band / phik_scf there, with their band
// fillings); their G-space coefficients and FP64 real-space grids sit behind the
// occupied orbitals in the context's orbital arrays, so a DivEntry addresses them
// like occupied ones.  All of this is FP64: these are corrections of O(N_states)
// transforms per band, next to the pair pipelines they correct.
#include "kernels_common.hpp"

namespace tcx {

namespace {

constexpr int kZwMaxStates = 512;  // states per zero-weight k (LDS coefficient arrays)
constexpr int kRedZ = 256;

__device__ __forceinline__ double2 dadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 dsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 dscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 dmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// x = G of grid point e (Cartesian); returns |x|^2
__device__ __forceinline__ double g_of(const Geom& g, long long e, double x[3])
{
    const int i1 = (int)(e % g.n1), i2 = (int)((e / g.n1) % g.n2), i3 = (int)(e / ((long long)g.n1 * g.n2));
    const int m1 = centered(i1, g.n1), m2 = centered(i2, g.n2), m3 = centered(i3, g.n3);
    double x2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        x[d] = g.b[d] * m1 + g.b[3 + d] * m2 + g.b[6 + d] * m3;
        x2 += x[d] * x[d];
    }
    return x2;
}

__device__ __forceinline__ double2 zmulc(double2 a, double2 b)  // conj(a) b
{
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// workgroup sum of n complex values (kRedZ lanes), red: 2 * kRedZ doubles
template <int n>
__device__ void block_sums(double2 (&v)[n], double* red)
{
    const int t = threadIdx.x;
    for (int q = 0; q < n; ++q) {
        red[t] = v[q].x;
        red[kRedZ + t] = v[q].y;
        __syncthreads();
        for (int w = kRedZ / 2; w > 0; w >>= 1) {
            if (t < w) {
                red[t] += red[t + w];
                red[kRedZ + t] += red[kRedZ + t + w];
            }
            __syncthreads();
        }
        v[q] = make_double2(red[0], red[kRedZ]);
        __syncthreads();
    }
}

// Cartesian (k + G) of plane wave l of k (gindex entry e, grid index i1 fastest)
__device__ __forceinline__ void k_plus_G(const Geom& g, const double* kc, int e, double out[3])
{
    const int i1 = e % g.n1, i2 = (e / g.n1) % g.n2, i3 = e / (g.n1 * g.n2);
    const int m1 = i1 <= g.n1 / 2 ? i1 : i1 - g.n1, m2 = i2 <= g.n2 / 2 ? i2 : i2 - g.n2,
              m3 = i3 <= g.n3 / 2 ? i3 : i3 - g.n3;
    for (int d = 0; d < 3; ++d) out[d] = kc[d] + g.b[d] * m1 + g.b[3 + d] * m2 + g.b[6 + d] * m3;
}

}  // namespace

// tc2x divergence terms of one applied band j at k (calc_hamiltonian_tc2x.cpp:408-513):
// for every state b at k with f_b >= 1e-8, p = <chi_b|phi_j>, r = <chi_b|(kv.(k+G)) phi_j>,
//   out_j += f_b (dc p + r) phi_b - f_b p (kv.(k+G)) phi_b
// ([2a_x], [2b_x2], [2b_x1]); kv[k] = kVaux of :117-126 (zero on the mesh: [2a_x] only).
__global__ __launch_bounds__(kRedZ) void k_divcorr_kv(const double2* __restrict__ phi, const double2* __restrict__ occ_r,
                                                      const double2* __restrict__ occ_l,
                                                      const DivEntry* __restrict__ jobs,
                                                      const double* __restrict__ fill, const double* __restrict__ dcs,
                                                      const double* __restrict__ kv, const int* __restrict__ gindex,
                                                      const int* __restrict__ gidx_off, const double* __restrict__ kcart,
                                                      Geom g, double2* __restrict__ out)
{
    const DivEntry e = jobs[blockIdx.x];
    __shared__ double2 ca[kZwMaxStates], cb[kZwMaxStates];
    __shared__ double red[2 * kRedZ];
    const int t = threadIdx.x;
    const double dc = dcs[e.kidx];
    const double* kvk = kv + 3 * e.kidx;
    const double* kc = kcart + 3 * e.kidx;
    const int* gi = gindex + gidx_off[e.kidx];
    const double2* pj = phi + e.app_off;
    for (int b = 0; b < e.nocc; ++b) {
        const double2* xb = occ_l + e.occ_off + (long long)b * e.npw;
        double2 s[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
        for (int l = t; l < e.npw; l += kRedZ) {
            double x[3];
            k_plus_G(g, kc, gi[l], x);
            const double kd = kvk[0] * x[0] + kvk[1] * x[1] + kvk[2] * x[2];
            const double2 c = zmulc(xb[l], pj[l]);
            s[0] = dadd(s[0], c);
            s[1] = dadd(s[1], dscale(c, kd));
        }
        block_sums<2>(s, red);
        if (t == 0) {
            const double f = fill[e.fill_off + b];
            const bool on = f >= 1e-8;
            ca[b] = on ? dscale(dadd(dscale(s[0], dc), s[1]), f) : make_double2(0.0, 0.0);
            cb[b] = on ? dscale(s[0], f) : make_double2(0.0, 0.0);
        }
        __syncthreads();
    }
    double2* o = out + e.app_off;
    for (int l = t; l < e.npw; l += kRedZ) {
        double x[3];
        k_plus_G(g, kc, gi[l], x);
        const double kd = kvk[0] * x[0] + kvk[1] * x[1] + kvk[2] * x[2];
        double2 acc = o[l];
        for (int b = 0; b < e.nocc; ++b) {
            const double2 v = occ_r[e.occ_off + (long long)b * e.npw + l];
            acc = dadd(acc, dmul(dsub(ca[b], dscale(cb[b], kd)), v));
        }
        o[l] = acc;
    }
}

void launch_divcorr_kv(const double* phi, const double* occ_right, const double* occ_left, const DivEntry* jobs,
                       int njobs, int max_states, const double* fill, const double* dc, const double* kv,
                       const int* gindex, const int* gidx_off, const double* kcart, const Geom& g, double* out,
                       hipStream_t s)
{
    if (njobs == 0) return;
    if (max_states > kZwMaxStates) throw Error(TCX_ERR_UNSUPPORTED, "divergence terms: more than 512 states at one k");
    hipLaunchKernelGGL(k_divcorr_kv, dim3(njobs), dim3(kRedZ), 0, s, (const double2*)phi, (const double2*)occ_right,
                       (const double2*)occ_left, jobs, fill, dc, kv, gindex, gidx_off, kcart, g, (double2*)out);
    TCX_HIP_CHECK(hipGetLastError());
}

// Off-mesh [3b6] table (tc3a3b3b4b6.cpp:445-448):
//   tG(G) = kVaux . G sum_js n_js(G) A_js u(G, 0, js) (G != 0), kVaux = -sum_of_kVaux, tG(0) = 0
template <bool kPoly>
__global__ __launch_bounds__(256) void k_b6_tG_kv(const double2* __restrict__ nG, double2* __restrict__ tG, Geom g,
                                                  long long N, JastrowPair j0, JastrowPair j1, double s0, double s1,
                                                  double s2)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    double x[3];
    const double x2 = g_of(g, r, x);
    if (r == 0) {
        tG[0] = make_double2(0.0, 0.0);
        return;
    }
    const double kg = -(s0 * x[0] + s1 * x[1] + s2 * x[2]);
    const double c0 = kg * j0.A * jastrow_uk<kPoly>(j0, x2), c1 = kg * j1.A * jastrow_uk<kPoly>(j1, x2);
    tG[r] = dadd(dscale(nG[r], c0), dscale(nG[N + r], c1));
}

void launch_b6_tG_kv(const double* nG, double* tG, const double skv[3], const Geom& g, long long N,
                     const JastrowPair& j0, const JastrowPair& j1, hipStream_t s)
{
    TCX_POLY_DISPATCH(j0.pdeg > 0, hipLaunchKernelGGL((k_b6_tG_kv<kPoly>), dim3((unsigned)((N + 255) / 256)), dim3(256),
                                                      0, s, (const double2*)nG, (double2*)tG, g, N, j0, j1, skv[0],
                                                      skv[1], skv[2]));
    TCX_HIP_CHECK(hipGetLastError());
}

// [3b3] and [3a3,b4] kVaux coefficients (tc3a3b3b4b6.cpp:596-620), one workgroup per
// (band j of the job range, state b):
//   cA[j][b] = f fb3 / N sum_d kv_d sum_r conj(chi_b(r)) dnu_d(r) phi_j(r)
//   cQ[j][b] = f <chi_b|phi_j>
// (kv = -sum_of_kVaux A_long[s][s]; states with f < 1e-8 get zeros).
__global__ __launch_bounds__(kRedZ) void k_zw_b3_coef(const double2* __restrict__ phi, const double2* __restrict__ occ_l,
                                                      const DivEntry* __restrict__ jobs,
                                                      const double* __restrict__ fill,
                                                      const double2* __restrict__ chiR, const double2* __restrict__ dnu,
                                                      const double2* __restrict__ appR, long long N, double fb3,
                                                      double k0, double k1, double k2, double2* __restrict__ cA,
                                                      double2* __restrict__ cQ, int stride)
{
    const int j = blockIdx.x, b = blockIdx.y;
    const DivEntry e = jobs[j];
    if (b >= e.nocc) return;  // uniform per workgroup
    __shared__ double red[2 * kRedZ];
    const int t = threadIdx.x;
    const double2* xb = chiR + (long long)(e.fill_off + b) * N;
    const double2* aj = appR + (long long)j * N;
    double2 s[4] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    for (long long r = t; r < N; r += kRedZ) {
        const double2 c = zmulc(xb[r], aj[r]);
        for (int d = 0; d < 3; ++d) s[d] = dadd(s[d], dmul(c, dnu[(long long)d * N + r]));
    }
    const double2* cb = occ_l + e.occ_off + (long long)b * e.npw;
    const double2* pj = phi + e.app_off;
    for (int l = t; l < e.npw; l += kRedZ) s[3] = dadd(s[3], zmulc(cb[l], pj[l]));
    block_sums<4>(s, red);
    if (t == 0) {
        const double f = fill[e.fill_off + b];
        const bool on = f >= 1e-8;
        const double2 a = dadd(dadd(dscale(s[0], k0), dscale(s[1], k1)), dscale(s[2], k2));
        cA[(long long)j * stride + b] = on ? dscale(a, f * fb3 / (double)N) : make_double2(0.0, 0.0);
        cQ[(long long)j * stride + b] = on ? dscale(s[3], f) : make_double2(0.0, 0.0);
    }
}

// [3b4] kVaux coefficients (tc3a3b3b4b6.cpp:837-853):
//   cA[j][b] = fb4 f / N sum_d kv_d sum_r conj(chi_b(r)) qsum_j,d(r)
__global__ __launch_bounds__(kRedZ) void k_zw_b4_coef(const DivEntry* __restrict__ jobs,
                                                      const double* __restrict__ fill,
                                                      const double2* __restrict__ chiR, const double2* __restrict__ qsum,
                                                      long long N, double fb4, double k0, double k1, double k2,
                                                      double2* __restrict__ cA, int stride)
{
    const int j = blockIdx.x, b = blockIdx.y;
    const DivEntry e = jobs[j];
    if (b >= e.nocc) return;
    __shared__ double red[2 * kRedZ];
    const int t = threadIdx.x;
    const double2* xb = chiR + (long long)(e.fill_off + b) * N;
    const double2* qj = qsum + (long long)j * 3 * N;
    double2 s[1] = {make_double2(0.0, 0.0)};
    for (long long r = t; r < N; r += kRedZ) {
        const double2 q = dadd(dadd(dscale(qj[r], k0), dscale(qj[N + r], k1)), dscale(qj[2 * N + r], k2));
        s[0] = dadd(s[0], zmulc(xb[r], q));
    }
    block_sums<1>(s, red);
    if (t == 0) {
        const double f = fill[e.fill_off + b];
        cA[(long long)j * stride + b] = f >= 1e-8 ? dscale(s[0], fb4 * f / (double)N) : make_double2(0.0, 0.0);
    }
}

// out_j += sum_b c[j][b] phi_b (G space, on the sphere of k), one workgroup per band j
__global__ __launch_bounds__(kRedZ) void k_zw_proj_add(const DivEntry* __restrict__ jobs,
                                                       const double2* __restrict__ occ_r,
                                                       const double2* __restrict__ c, int stride,
                                                       double2* __restrict__ out)
{
    const int j = blockIdx.x;
    const DivEntry e = jobs[j];
    double2* o = out + e.app_off;
    for (int l = threadIdx.x; l < e.npw; l += kRedZ) {
        double2 acc = o[l];
        for (int b = 0; b < e.nocc; ++b)
            acc = dadd(acc, dmul(c[(long long)j * stride + b], occ_r[e.occ_off + (long long)b * e.npw + l]));
        o[l] = acc;
    }
}

// qsum_j,d(r) += kv_d sum_b cQ[j][b] phi_b(r)  (tc3a3b3b4b6.cpp:612-620)
__global__ __launch_bounds__(256) void k_zw_qsum_add(const DivEntry* __restrict__ jobs,
                                                     const double2* __restrict__ occR,
                                                     const double2* __restrict__ cQ, int stride, long long N,
                                                     double k0, double k1, double k2, double2* __restrict__ qsum)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= N) return;
    const int j = blockIdx.y;
    const DivEntry e = jobs[j];
    double2 acc = make_double2(0.0, 0.0);
    for (int b = 0; b < e.nocc; ++b)
        acc = dadd(acc, dmul(cQ[(long long)j * stride + b], occR[(long long)(e.fill_off + b) * N + r]));
    double2* q = qsum + (long long)j * 3 * N + r;
    q[0] = dadd(q[0], dscale(acc, k0));
    q[N] = dadd(q[N], dscale(acc, k1));
    q[2 * N] = dadd(q[2 * N], dscale(acc, k2));
}

void launch_zw_b3(const double* phi, const double* occ_l, const DivEntry* jobs, int njobs, int max_states,
                  const double* fill, const double* chiR, const double* occR, const double* occ_r, const double* dnu,
                  const double* appR, long long N, double fb3, const double kv[3], double* cA, double* cQ,
                  double* qsum, double* out, hipStream_t s)
{
    if (njobs == 0 || max_states == 0) return;
    if (max_states > kZwMaxStates) throw Error(TCX_ERR_UNSUPPORTED, "divergence terms: more than 512 states at one k");
    hipLaunchKernelGGL(k_zw_b3_coef, dim3(njobs, max_states), dim3(kRedZ), 0, s, (const double2*)phi,
                       (const double2*)occ_l, jobs, fill, (const double2*)chiR, (const double2*)dnu,
                       (const double2*)appR, N, fb3, kv[0], kv[1], kv[2], (double2*)cA, (double2*)cQ, max_states);
    TCX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_zw_proj_add, dim3(njobs), dim3(kRedZ), 0, s, jobs, (const double2*)occ_r,
                       (const double2*)cA, max_states, (double2*)out);
    TCX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_zw_qsum_add, dim3((unsigned)((N + 255) / 256), njobs), dim3(256), 0, s, jobs,
                       (const double2*)occR, (const double2*)cQ, max_states, N, kv[0], kv[1], kv[2], (double2*)qsum);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_zw_b4(const DivEntry* jobs, int njobs, int max_states, const double* fill, const double* chiR,
                  const double* occ_r, const double* qsum, long long N, double fb4, const double kv[3], double* cA,
                  double* out, hipStream_t s)
{
    if (njobs == 0 || max_states == 0) return;
    if (max_states > kZwMaxStates) throw Error(TCX_ERR_UNSUPPORTED, "divergence terms: more than 512 states at one k");
    hipLaunchKernelGGL(k_zw_b4_coef, dim3(njobs, max_states), dim3(kRedZ), 0, s, jobs, fill, (const double2*)chiR,
                       (const double2*)qsum, N, fb4, kv[0], kv[1], kv[2], (double2*)cA, max_states);
    TCX_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_zw_proj_add, dim3(njobs), dim3(kRedZ), 0, s, jobs, (const double2*)occ_r,
                       (const double2*)cA, max_states, (double2*)out);
    TCX_HIP_CHECK(hipGetLastError());
}

// Kernel table of the [3a4,b5] kVaux block (tc3a2a4b2b5.cpp:1106-1110): per k-q1 row,
//   du(G) / N = (kv . (k - q1 + G)) u(k - q1 + G, s, s) / N
// in FP64, stored like the Coulomb table of pass B (y fastest within a plane), so the
// hf2x pair passes carry the block's three transforms.
template <bool kPoly>
__global__ __launch_bounds__(256) void k_duk_table(double* __restrict__ table, const double* __restrict__ kqcart,
                                                   int row0, Geom g, long long N, double invN, JastrowPair j, double k0,
                                                   double k1, double k2)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= N) return;
    const int i2 = (int)(e % g.n2), i1 = (int)((e / g.n2) % g.n1), i3 = (int)(e / ((long long)g.n1 * g.n2));
    const int m1 = i1 <= g.n1 / 2 ? i1 : i1 - g.n1, m2 = i2 <= g.n2 / 2 ? i2 : i2 - g.n2,
              m3 = i3 <= g.n3 / 2 ? i3 : i3 - g.n3;
    const double* kq = kqcart + 3 * (row0 + blockIdx.y);
    double x[3], x2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        x[d] = kq[d] + g.b[d] * m1 + g.b[3 + d] * m2 + g.b[6 + d] * m3;
        x2 += x[d] * x[d];
    }
    table[(long long)blockIdx.y * N + e] = invN * (k0 * x[0] + k1 * x[1] + k2 * x[2]) * jastrow_uk<kPoly>(j, x2);
}

void launch_duk_table(double* table, const double* kqcart, int row0, int nrows, const Geom& g, const JastrowPair& j,
                      const double kv[3], hipStream_t s)
{
    if (nrows == 0) return;
    const long long N = (long long)g.n1 * g.n2 * g.n3;
    TCX_POLY_DISPATCH(j.pdeg > 0, hipLaunchKernelGGL((k_duk_table<kPoly>), dim3((unsigned)((N + 255) / 256), nrows),
                                                     dim3(256), 0, s, table, kqcart, row0, g, N, 1.0 / (double)N, j,
                                                     kv[0], kv[1], kv[2]));
    TCX_HIP_CHECK(hipGetLastError());
}

// [3a4,b5] kVaux apply (tc3a2a4b2b5.cpp:1231-1290), one workgroup per applied band j at k:
// Wa_b (G space) = FFT of the [3a4] orbital of state b, Wb_b = FFT of the [3b5] sum
// (orbital2 = -Wb_b), over all q1; for every state b with f_b >= 1e-8, p = <chi_b|phi_j>,
//   out_j += fa4 f_b p Wa_b(G) + [fa4 f_b^2 p kv2 nq + fb5 f_b (c5 + f_b p kv2 nq)] phi_b,
//   c5 = <orbital2_b|phi_j> = -sum_l conj(Wb_b(G_l)) phi_j(l),
// with nq the number of q1 the reference's loop adds the |kVaux|^2 terms for.
__global__ __launch_bounds__(kRedZ) void k_zw_a4b5_apply(const double2* __restrict__ phi,
                                                         const double2* __restrict__ occ_r,
                                                         const double2* __restrict__ occ_l,
                                                         const DivEntry* __restrict__ jobs,
                                                         const double* __restrict__ fill,
                                                         const double2* __restrict__ Wa, const double2* __restrict__ Wb,
                                                         const int* __restrict__ gi, long long N, double fa4, double fb5,
                                                         double kv2nq, double2* __restrict__ out)
{
    const int j = blockIdx.x;
    const DivEntry e = jobs[j];
    __shared__ double2 ca[kZwMaxStates], cb[kZwMaxStates];
    __shared__ double red[2 * kRedZ];
    const int t = threadIdx.x;
    const double2* pj = phi + e.app_off;
    for (int b = 0; b < e.nocc; ++b) {
        const double2* xb = occ_l + e.occ_off + (long long)b * e.npw;
        const double2* wb = Wb + (long long)b * N;
        double2 s[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
        for (int l = t; l < e.npw; l += kRedZ) {
            s[0] = dadd(s[0], zmulc(xb[l], pj[l]));
            s[1] = dadd(s[1], zmulc(wb[gi[l]], pj[l]));
        }
        block_sums<2>(s, red);
        if (t == 0) {
            const double f = fill[e.fill_off + b];
            const bool on = f >= 1e-8;
            const double2 p = s[0], c5 = dscale(s[1], -1.0);
            ca[b] = on ? dscale(p, fa4 * f) : make_double2(0.0, 0.0);
            const double2 tb = dadd(dscale(p, fa4 * f * f * kv2nq), dscale(dadd(c5, dscale(p, f * kv2nq)), fb5 * f));
            cb[b] = on ? tb : make_double2(0.0, 0.0);
        }
        __syncthreads();
    }
    double2* o = out + e.app_off;
    for (int l = t; l < e.npw; l += kRedZ) {
        double2 acc = o[l];
        const int gl = gi[l];
        for (int b = 0; b < e.nocc; ++b) {
            acc = dadd(acc, dmul(ca[b], Wa[(long long)b * N + gl]));
            acc = dadd(acc, dmul(cb[b], occ_r[e.occ_off + (long long)b * e.npw + l]));
        }
        o[l] = acc;
    }
}

void launch_zw_a4b5_apply(const double* phi, const double* occ_r, const double* occ_l, const DivEntry* jobs, int njobs,
                          int max_states, const double* fill, const double* Wa, const double* Wb, const int* gi,
                          long long N, double fa4, double fb5, double kv2nq, double* out, hipStream_t s)
{
    if (njobs == 0 || max_states == 0) return;
    if (max_states > kZwMaxStates) throw Error(TCX_ERR_UNSUPPORTED, "divergence terms: more than 512 states at one k");
    hipLaunchKernelGGL(k_zw_a4b5_apply, dim3(njobs), dim3(kRedZ), 0, s, (const double2*)phi, (const double2*)occ_r,
                       (const double2*)occ_l, jobs, fill, (const double2*)Wa, (const double2*)Wb, gi, N, fa4, fb5,
                       kv2nq, (double2*)out);
    TCX_HIP_CHECK(hipGetLastError());
}

}  // namespace tcx
