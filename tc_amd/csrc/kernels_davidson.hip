// kernels_davidson.hip -- the FP64 algebra of block Davidson on the device, batched over k.
//
// Diagonalization::block_davidson (src/diagonalization_davidson.cpp:7-689).  A vector set
// names vector i of k-point k at base + 2 * (off[k] + i * npw[k]) doubles (complex128,
// off and npw are device arrays), so the same kernels serve the subspace V (max_blocks * nb
// slots per k), the H_n V sets, the candidate vectors and the packed [k][band][npw] arrays
// of the C ABI.  Inner products and norms are one 256-thread workgroup per (k, vector) over
// npw with FP64 sums; the subspace problem of every k (a small non-Hermitian matrix) is one
// 64-lane workgroup per k with its matrices in LDS: Householder reduction to Hessenberg
// form, single-shift QR with Wilkinson shifts to the complex Schur form (Golub & Van Loan,
// Matrix Computations, Alg. 7.4.2 / 7.5), back-substitution for the eigenvectors, the sort
// by real part (or by overlap, bi-orthogonal variant) and, for BITC, X_L = (X^-1)^H by
// Gauss-Jordan.  The host version of the
// same eigensolver is eig_dense (tcx_davidson.cpp).
#include "kernels_common.hpp"

namespace tcx {

namespace {

struct z2 {
    double re, im;
};
__device__ __forceinline__ z2 zadd(z2 a, z2 b) { return z2{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ z2 zsub(z2 a, z2 b) { return z2{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ z2 zmul(z2 a, z2 b) { return z2{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ z2 zconj(z2 a) { return z2{a.re, -a.im}; }
__device__ __forceinline__ z2 zscale(z2 a, double s) { return z2{a.re * s, a.im * s}; }
__device__ __forceinline__ double zabs(z2 a) { return hypot(a.re, a.im); }
__device__ __forceinline__ double znorm(z2 a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ z2 zdiv(z2 a, z2 b)
{
    const double d = b.re * b.re + b.im * b.im;
    return z2{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
__device__ __forceinline__ z2 zsqrt(z2 a)
{
    const double r = zabs(a);
    const double x = sqrt(0.5 * (r + a.re)), y = sqrt(0.5 * fmax(r - a.re, 0.0));
    return z2{x, a.im < 0.0 ? -y : y};
}

__device__ __forceinline__ z2 ld(const double* base, long long i) { return z2{base[2 * i], base[2 * i + 1]}; }
__device__ __forceinline__ void st(double* base, long long i, z2 v)
{
    base[2 * i] = v.re;
    base[2 * i + 1] = v.im;
}

constexpr int kRed = 256;  // threads of the reduction workgroups

// workgroup-wide sum of a complex value (kRed lanes)
__device__ z2 block_sum(z2 v, double* red)
{
    const int t = threadIdx.x;
    red[t] = v.re;
    red[kRed + t] = v.im;
    __syncthreads();
    for (int w = kRed / 2; w > 0; w >>= 1) {
        if (t < w) {
            red[t] += red[t + w];
            red[kRed + t] += red[kRed + t + w];
        }
        __syncthreads();
    }
    const z2 out{red[0], red[kRed]};
    __syncthreads();
    return out;
}

}  // namespace

// Payne preconditioner of the residual, normalised (make_a_new_trial_vector,
// diagonalization_utils.cpp:182-210): candidate b of k (b < cnt[k]) is
//   T_b = -P(x) (sum_n H_n phi - eps phi) / norm,  x = |k+G|^2 / <phi|-nabla^2|phi>,
// phi = V slot first[k] + b, H_n phi the same slot of the H_n V sets, eps = eig[band0[k] + b];
// a candidate whose norm is below 1e-8 is zero.  Grid (nk, max cnt).
__global__ __launch_bounds__(kRed) void k_dv_precond(const double* __restrict__ V, const long long* __restrict__ voff,
                                                     const int* __restrict__ first, const int* __restrict__ cnt,
                                                     const double* __restrict__ H0, const double* __restrict__ H1,
                                                     const double* __restrict__ H2, const double* __restrict__ kG2,
                                                     const long long* __restrict__ goff, const double* __restrict__ eig,
                                                     const int* __restrict__ band0, double* __restrict__ T,
                                                     const long long* __restrict__ toff, const int* __restrict__ npw)
{
    __shared__ double red[2 * kRed];
    const int k = blockIdx.x, b = blockIdx.y;
    if (b >= cnt[k]) return;  // uniform per workgroup
    const int n = npw[k];
    const long long p0 = voff[k] + (long long)(first[k] + b) * n, t0 = toff[k] + (long long)b * n;
    const double* g = kG2 + goff[k];
    z2 s{0.0, 0.0};
    for (int l = threadIdx.x; l < n; l += kRed) {
        const double a = znorm(ld(V, p0 + l));
        s.re += g[l] * a;
        s.im += a;
    }
    s = block_sum(s, red);
    const double kin = s.re / s.im;
    const z2 e = ld(eig, band0[k] + b);
    double tn = 0.0;
    for (int l = threadIdx.x; l < n; l += kRed) {
        const double x = g[l] / kin;
        const double c = (27 + x * (18 + x * (12 + x * 8))) / (27 + x * (18 + x * (12 + x * (8 + x * 16))));
        z2 h = zadd(ld(H0, p0 + l), ld(H1, p0 + l));
        if (H2) h = zadd(h, ld(H2, p0 + l));
        const z2 tv = zscale(zsub(h, zmul(e, ld(V, p0 + l))), -c);
        tn += znorm(tv);
        st(T, t0 + l, tv);
    }
    const double nrm = sqrt(block_sum(z2{tn, 0.0}, red).re);
    const double f = nrm < 1e-8 ? 0.0 : 1.0 / nrm;
    for (int l = threadIdx.x; l < n; l += kRed) st(T, t0 + l, zscale(ld(T, t0 + l), f));  // own elements only
}

// c[k][i] = <A_i | B_bidx[k]> for i < nA[k]; k with bidx[k] < 0 is skipped.  Grid (max nA, nk).
// (A is the dual set D of launch_dv_project: the set itself for Gram_Schmidt, the other
// side's set for Gram_Schmidt_biortho.)
__global__ __launch_bounds__(kRed) void k_dv_dots(const double* __restrict__ A, const long long* __restrict__ aoff,
                                                  const int* __restrict__ nA, const double* __restrict__ B,
                                                  const long long* __restrict__ boff, const int* __restrict__ bidx,
                                                  const int* __restrict__ npw, double* __restrict__ c, int cstride)
{
    __shared__ double red[2 * kRed];
    const int k = blockIdx.y, i = blockIdx.x;
    if (i >= nA[k] || bidx[k] < 0) return;
    const int n = npw[k];
    const long long a0 = aoff[k] + (long long)i * n, b0 = boff[k] + (long long)bidx[k] * n;
    z2 s{0.0, 0.0};
    for (int l = threadIdx.x; l < n; l += kRed) s = zadd(s, zmul(zconj(ld(A, a0 + l)), ld(B, b0 + l)));
    s = block_sum(s, red);
    if (threadIdx.x == 0) st(c, (long long)k * cstride + i, s);
}

// B_bidx[k] -= sum_{i < nA[k]} c[k][i] A_i.  Grid (ceil(max npw / 256), nk).
__global__ __launch_bounds__(kRed) void k_dv_subtract(const double* __restrict__ A, const long long* __restrict__ aoff,
                                                      const int* __restrict__ nA, double* __restrict__ B,
                                                      const long long* __restrict__ boff, const int* __restrict__ bidx,
                                                      const int* __restrict__ npw, const double* __restrict__ c,
                                                      int cstride)
{
    const int k = blockIdx.y;
    const int n = npw[k];
    const int l = blockIdx.x * kRed + threadIdx.x;
    if (l >= n || bidx[k] < 0) return;
    const long long o = boff[k] + (long long)bidx[k] * n + l;
    z2 v = ld(B, o);
    for (int i = 0; i < nA[k]; ++i)
        v = zsub(v, zmul(ld(c, (long long)k * cstride + i), ld(A, aoff[k] + (long long)i * n + l)));
    st(B, o, v);
}

// nrm[k] = || B_bidx[k] ||.  Grid nk.
__global__ __launch_bounds__(kRed) void k_dv_norm(const double* __restrict__ B, const long long* __restrict__ boff,
                                                  const int* __restrict__ bidx, const int* __restrict__ npw,
                                                  double* __restrict__ nrm)
{
    __shared__ double red[2 * kRed];
    const int k = blockIdx.x;
    if (bidx[k] < 0) return;
    const int n = npw[k];
    const long long b0 = boff[k] + (long long)bidx[k] * n;
    double s = 0.0;
    for (int l = threadIdx.x; l < n; l += kRed) s += znorm(ld(B, b0 + l));
    s = block_sum(z2{s, 0.0}, red).re;
    if (threadIdx.x == 0) nrm[k] = sqrt(s);
}

// O_odst[k] = B_bidx[k] / nrm[k] (in place when O, ooff, odst name B's slot).  Grid (ceil(max npw / 256), nk).
__global__ __launch_bounds__(kRed) void k_dv_store(const double* B, const long long* __restrict__ boff,
                                                   const int* __restrict__ bidx, const double* __restrict__ nrm,
                                                   double* O, const long long* __restrict__ ooff,
                                                   const int* __restrict__ odst, const int* __restrict__ npw)
{
    const int k = blockIdx.y;
    const int n = npw[k];
    const int l = blockIdx.x * kRed + threadIdx.x;
    if (l >= n || bidx[k] < 0 || odst[k] < 0) return;
    const z2 v = ld(B, boff[k] + (long long)bidx[k] * n + l);
    st(O, ooff[k] + (long long)odst[k] * n + l, zscale(v, 1.0 / nrm[k]));
}

// D slots dfirst[k] .. + cnt[k] = S slots sfirst[k] .. + cnt[k] (null first = 0).
// Grid (ceil(max npw / 256), max cnt, nk).
__global__ __launch_bounds__(kRed) void k_dv_copy(const double* __restrict__ S, const long long* __restrict__ soff,
                                                  const int* __restrict__ sfirst, double* __restrict__ D,
                                                  const long long* __restrict__ doff, const int* __restrict__ dfirst,
                                                  const int* __restrict__ cnt, const int* __restrict__ npw)
{
    const int k = blockIdx.z, b = blockIdx.y;
    const int n = npw[k];
    const int l = blockIdx.x * kRed + threadIdx.x;
    if (l >= n || b >= cnt[k]) return;
    st(D, doff[k] + (long long)((dfirst ? dfirst[k] : 0) + b) * n + l,
       ld(S, soff[k] + (long long)((sfirst ? sfirst[k] : 0) + b) * n + l));
}

// G_q[k][i][j] = <A_i | B_q,j> for i, j < nA[k], q < nB (B_q are sets with A's offsets;
// A is V, or V_left in the bi-orthogonal variant, :465-466).
// Grid (gs * gs, nk, nB); G_q[k] at G + 2 * (q * nk + k) * gs * gs.
__global__ __launch_bounds__(kRed) void k_dv_gram(const double* __restrict__ A, const long long* __restrict__ aoff,
                                                  const int* __restrict__ nA, const double* __restrict__ B0,
                                                  const double* __restrict__ B1, const double* __restrict__ B2,
                                                  const int* __restrict__ npw, double* __restrict__ G, int gs)
{
    __shared__ double red[2 * kRed];
    const int k = blockIdx.y, q = blockIdx.z;
    const int i = blockIdx.x / gs, j = blockIdx.x % gs;
    const int m = nA[k];
    if (i >= m || j >= m) return;
    const double* B = q == 0 ? B0 : (q == 1 ? B1 : B2);
    const int n = npw[k];
    const long long a0 = aoff[k] + (long long)i * n, b0 = aoff[k] + (long long)j * n;
    z2 s{0.0, 0.0};
    for (int l = threadIdx.x; l < n; l += kRed) s = zadd(s, zmul(zconj(ld(A, a0 + l)), ld(B, b0 + l)));
    s = block_sum(s, red);
    const long long msz = (long long)gs * gs;
    if (threadIdx.x == 0) st(G, ((long long)q * gridDim.y + k) * msz + (long long)i * gs + j, s);
}

// O_b(k) = sum_{i < nA[k]} C[k][i][b] A_i for b < nout[k] (C[k] row-major with stride gs).
// Grid (ceil(max npw / 256), nk).
__global__ __launch_bounds__(kRed) void k_dv_rotate(const double* __restrict__ A, const long long* __restrict__ aoff,
                                                    const int* __restrict__ nA, const double* __restrict__ C, int gs,
                                                    double* __restrict__ O, const long long* __restrict__ ooff,
                                                    const int* __restrict__ nout, const int* __restrict__ npw)
{
    const int k = blockIdx.y;
    const int n = npw[k];
    const int l = blockIdx.x * kRed + threadIdx.x;
    if (l >= n) return;
    const double* Ck = C + 2 * (long long)k * gs * gs;
    const int m = nA[k];
    for (int b = 0; b < nout[k]; ++b) {
        z2 s{0.0, 0.0};
        for (int i = 0; i < m; ++i)
            s = zadd(s, zmul(ld(Ck, (long long)i * gs + b), ld(A, aoff[k] + (long long)i * n + l)));
        st(O, ooff[k] + (long long)b * n + l, s);
    }
}

// c[k] = <X_xidx[k] | Y_yidx[k]> (X at xoff, Y at yoff); k with xidx[k] < 0 is skipped.  Grid nk.
__global__ __launch_bounds__(kRed) void k_dv_dot(const double* __restrict__ X, const long long* __restrict__ xoff,
                                                 const int* __restrict__ xidx, const double* __restrict__ Y,
                                                 const long long* __restrict__ yoff, const int* __restrict__ yidx,
                                                 const int* __restrict__ npw, double* __restrict__ c)
{
    __shared__ double red[2 * kRed];
    const int k = blockIdx.x;
    if (xidx[k] < 0) return;
    const int n = npw[k];
    const long long x0 = xoff[k] + (long long)xidx[k] * n, y0 = yoff[k] + (long long)yidx[k] * n;
    z2 s{0.0, 0.0};
    for (int l = threadIdx.x; l < n; l += kRed) s = zadd(s, zmul(zconj(ld(X, x0 + l)), ld(Y, y0 + l)));
    s = block_sum(s, red);
    if (threadIdx.x == 0) st(c, k, s);
}

// O_odst[k] = B_bidx[k] / conj(c[k]) (Gram_Schmidt_biortho's <t|t> = 1 on the left vector,
// diagonalization_utils.cpp:69-76).  Grid (ceil(max npw / 256), nk).
__global__ __launch_bounds__(kRed) void k_dv_cdiv_store(const double* __restrict__ B, const long long* __restrict__ boff,
                                                        const int* __restrict__ bidx, const double* __restrict__ c,
                                                        double* __restrict__ O, const long long* __restrict__ ooff,
                                                        const int* __restrict__ odst, const int* __restrict__ npw)
{
    const int k = blockIdx.y;
    const int n = npw[k];
    const int l = blockIdx.x * kRed + threadIdx.x;
    if (l >= n || bidx[k] < 0 || odst[k] < 0) return;
    const z2 v = ld(B, boff[k] + (long long)bidx[k] * n + l);
    st(O, ooff[k] + (long long)odst[k] * n + l, zdiv(v, zconj(ld(c, k))));
}

// Bi-orthogonal pair normalisation after the rotation (diagonalization_davidson.cpp:609-615):
// for b < nb[k], nrm = ||R_b||, R_b /= nrm, L_b *= nrm, out[band0[k] + b] = nrm.  Grid (nk, max nb).
__global__ __launch_bounds__(kRed) void k_dv_binorm(double* __restrict__ R, double* __restrict__ L,
                                                    const long long* __restrict__ off, const int* __restrict__ nb,
                                                    const int* __restrict__ band0, const int* __restrict__ npw,
                                                    double* __restrict__ out)
{
    __shared__ double red[2 * kRed];
    const int k = blockIdx.x, b = blockIdx.y;
    if (b >= nb[k]) return;  // uniform per workgroup
    const int n = npw[k];
    const long long o = off[k] + (long long)b * n;
    double s = 0.0;
    for (int l = threadIdx.x; l < n; l += kRed) s += znorm(ld(R, o + l));
    const double nrm = sqrt(block_sum(z2{s, 0.0}, red).re);
    if (threadIdx.x == 0) out[band0[k] + b] = nrm;
    if (!(nrm >= 1e-10)) return;  // the host reports the breakdown
    for (int l = threadIdx.x; l < n; l += kRed) {
        st(R, o + l, zscale(ld(R, o + l), 1.0 / nrm));
        st(L, o + l, zscale(ld(L, o + l), nrm));
    }
}

// ----------------------------------------------------------------- subspace problem, one workgroup per k
constexpr int kEigLanes = 64;

// Work area of one k (LDS when it fits, else global): Hm, Zm, Im (gs x gs complex),
// vec, wv (gs complex), perm (gs ints).
__host__ __device__ inline size_t dv_eig_bytes(int gs) { return (size_t)(3 * gs * gs + 2 * gs) * 16 + (size_t)gs * 4; }
__host__ __device__ inline size_t dv_eig_stride(int gs) { return dv_eig_bytes(gs) / 8 + 1; }  // in doubles

// Hs_q[k] (q < nH) summed and eigen-decomposed: w[k] ascending by real part (sort_eigen,
// diagonalization_utils.cpp:79-118) or, with sort_overlap, ordered by the largest component
// (sort_eigen_overlap, :121-167), X[k] unit-norm right eigenvectors (columns); for
// want_left X_L[k] = (X^-1)^H (diagonalization_davidson.cpp:498); with en non-null,
// en[band0[k] + b][q] = (X_L^H Hs_q X)_bb for b < nb[k] (X_L = X unless want_left,
// :571-585).  status[k] = 0 ok, 1 QR did not converge, 2 singular eigenvector matrix.
// Every lane takes the same decisions from the same shared values.
__global__ __launch_bounds__(kEigLanes) void k_dv_eig(const double* __restrict__ Hs, int nH,
                                                      const int* __restrict__ dim, int gs, int want_left,
                                                      int sort_overlap,
                                                      double* __restrict__ w, double* __restrict__ X,
                                                      double* __restrict__ XL, double* __restrict__ gwork, int use_lds,
                                                      const int* __restrict__ nb, const int* __restrict__ band0,
                                                      double* __restrict__ en, int* __restrict__ status)
{
    extern __shared__ double lds[];
    const int k = blockIdx.x, t = threadIdx.x, nk = gridDim.x;
    const int n = dim[k];
    const long long msz = (long long)gs * gs;
    double* base = use_lds ? lds : gwork + (long long)k * (long long)dv_eig_stride(gs);
    double* Hm = base;
    double* Zm = Hm + 2 * msz;
    double* Im = Zm + 2 * msz;
    double* vec = Im + 2 * msz;
    double* wv = vec + 2 * gs;
    int* perm = reinterpret_cast<int*>(wv + 2 * gs);
    double* Xk = X + 2 * msz * k;
    double* wk = w + 2 * (long long)gs * k;
    auto h = [&](int i, int j) { return (long long)i * gs + j; };
    if (t == 0) status[k] = 0;
    if (n == 0) return;

    for (int e = t; e < n * n; e += kEigLanes) {
        const int i = e / n, j = e % n;
        z2 s{0.0, 0.0};
        for (int q = 0; q < nH; ++q) s = zadd(s, ld(Hs + 2 * ((long long)q * nk + k) * msz, h(i, j)));
        st(Hm, h(i, j), s);
        st(Zm, h(i, j), z2{i == j ? 1.0 : 0.0, 0.0});
    }
    __syncthreads();
    const double eps = 2.220446049250313e-16;

    // Householder reduction to upper Hessenberg form, A = Z H Z^H
    for (int c = 0; c + 2 < n; ++c) {
        double xn = 0.0;
        for (int i = c + 1; i < n; ++i) xn += znorm(ld(Hm, h(i, c)));
        xn = sqrt(xn);
        if (xn == 0.0) continue;  // uniform
        const z2 x0 = ld(Hm, h(c + 1, c));
        const double ax0 = zabs(x0);
        const z2 phase = ax0 > 0.0 ? zscale(x0, 1.0 / ax0) : z2{1.0, 0.0};
        const z2 alpha = zscale(phase, -xn);
        for (int i = t; i < n; i += kEigLanes) {
            z2 v{0.0, 0.0};
            if (i > c) v = ld(Hm, h(i, c));
            if (i == c + 1) v = zsub(v, alpha);
            st(vec, i, v);
        }
        __syncthreads();
        double vn = 0.0;
        for (int i = c + 1; i < n; ++i) vn += znorm(ld(vec, i));
        vn = sqrt(vn);
        __syncthreads();
        if (vn == 0.0) continue;  // uniform
        for (int i = t; i < n; i += kEigLanes)
            if (i > c) st(vec, i, zscale(ld(vec, i), 1.0 / vn));
        __syncthreads();
        for (int j = t; j < n; j += kEigLanes) {  // H <- (I - 2 v v^H) H, lane per column
            z2 s{0.0, 0.0};
            for (int i = c + 1; i < n; ++i) s = zadd(s, zmul(zconj(ld(vec, i)), ld(Hm, h(i, j))));
            for (int i = c + 1; i < n; ++i) st(Hm, h(i, j), zsub(ld(Hm, h(i, j)), zscale(zmul(ld(vec, i), s), 2.0)));
        }
        __syncthreads();
        for (int i = t; i < n; i += kEigLanes) {  // H <- H (I - 2 v v^H), Z <- Z (I - 2 v v^H), lane per row
            z2 s{0.0, 0.0}, u{0.0, 0.0};
            for (int j = c + 1; j < n; ++j) {
                s = zadd(s, zmul(ld(Hm, h(i, j)), ld(vec, j)));
                u = zadd(u, zmul(ld(Zm, h(i, j)), ld(vec, j)));
            }
            for (int j = c + 1; j < n; ++j) {
                st(Hm, h(i, j), zsub(ld(Hm, h(i, j)), zscale(zmul(s, zconj(ld(vec, j))), 2.0)));
                st(Zm, h(i, j), zsub(ld(Zm, h(i, j)), zscale(zmul(u, zconj(ld(vec, j))), 2.0)));
            }
        }
        __syncthreads();
        for (int i = c + 2 + t; i < n; i += kEigLanes) st(Hm, h(i, c), z2{0.0, 0.0});
        __syncthreads();
    }
    double anorm = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) anorm += znorm(ld(Hm, h(i, j)));
    anorm = fmax(sqrt(anorm), 1e-300);

    // single-shift QR on the active block [l, hi] until H is upper triangular
    int hi = n - 1, iter = 0, total = 0;
    while (hi > 0) {
        if (++total > 200 * n) {
            if (t == 0) status[k] = 1;
            return;
        }
        int l = hi;
        while (l > 0) {
            double s = zabs(ld(Hm, h(l - 1, l - 1))) + zabs(ld(Hm, h(l, l)));
            if (s == 0.0) s = anorm;
            if (zabs(ld(Hm, h(l, l - 1))) <= eps * s) break;
            --l;
        }
        __syncthreads();
        if (l > 0 && t == 0) st(Hm, h(l, l - 1), z2{0.0, 0.0});
        __syncthreads();
        if (l == hi) {
            --hi;
            iter = 0;
            continue;
        }
        ++iter;
        const z2 a = ld(Hm, h(hi - 1, hi - 1)), b = ld(Hm, h(hi - 1, hi)), cc = ld(Hm, h(hi, hi - 1)),
                 d = ld(Hm, h(hi, hi));
        z2 mu;
        if (iter % 11 == 10) {
            mu = zadd(d, z2{fabs(cc.re) + fabs(cc.im), 0.0});  // exceptional shift
        } else {
            const z2 half = zscale(zadd(a, d), 0.5);
            const z2 amd = zsub(a, d);
            const z2 disc = zsqrt(zadd(zscale(zmul(amd, amd), 0.25), zmul(b, cc)));
            const z2 m1 = zadd(half, disc), m2 = zsub(half, disc);
            mu = zabs(zsub(m1, d)) < zabs(zsub(m2, d)) ? m1 : m2;
        }
        z2 x = zsub(ld(Hm, h(l, l)), mu), y = ld(Hm, h(l + 1, l));
        for (int c = l; c < hi; ++c) {
            if (c > l) {
                x = ld(Hm, h(c, c - 1));
                y = ld(Hm, h(c + 1, c - 1));
            }
            __syncthreads();
            const double r = sqrt(znorm(x) + znorm(y));
            double cs = 1.0;
            z2 sn{0.0, 0.0};
            if (r > 0.0) {
                const double ax = zabs(x);
                if (ax == 0.0) {
                    cs = 0.0;
                    sn = zscale(zconj(y), 1.0 / zabs(y));
                } else {
                    cs = ax / r;
                    sn = zscale(zmul(zscale(x, 1.0 / ax), zconj(y)), 1.0 / r);
                }
            }
            const int j0 = c > l ? c - 1 : l;
            for (int j = j0 + t; j < n; j += kEigLanes) {  // rows c, c+1 <- G [rows]
                const z2 h1 = ld(Hm, h(c, j)), h2 = ld(Hm, h(c + 1, j));
                st(Hm, h(c, j), zadd(zscale(h1, cs), zmul(sn, h2)));
                st(Hm, h(c + 1, j), (c > l && j == c - 1) ? z2{0.0, 0.0} : zsub(zscale(h2, cs), zmul(zconj(sn), h1)));
            }
            __syncthreads();
            const int imax = min(c + 2, hi);
            for (int i = t; i < n; i += kEigLanes) {  // columns c, c+1 <- [columns] G^H
                if (i <= imax) {
                    const z2 h1 = ld(Hm, h(i, c)), h2 = ld(Hm, h(i, c + 1));
                    st(Hm, h(i, c), zadd(zscale(h1, cs), zmul(zconj(sn), h2)));
                    st(Hm, h(i, c + 1), zsub(zscale(h2, cs), zmul(sn, h1)));
                }
                const z2 z1 = ld(Zm, h(i, c)), zz = ld(Zm, h(i, c + 1));
                st(Zm, h(i, c), zadd(zscale(z1, cs), zmul(zconj(sn), zz)));
                st(Zm, h(i, c + 1), zsub(zscale(zz, cs), zmul(sn, z1)));
            }
            __syncthreads();
        }
    }

    // eigenvectors of the triangular factor (lane per eigenvector; y in column c of Im)
    const double small = eps * anorm;
    for (int c = t; c < n; c += kEigLanes) {
        const z2 lam = ld(Hm, h(c, c));
        st(wv, c, lam);
        for (int i = 0; i < n; ++i) st(Im, h(i, c), z2{i == c ? 1.0 : 0.0, 0.0});
        for (int i = c - 1; i >= 0; --i) {
            z2 s{0.0, 0.0};
            for (int j = i + 1; j <= c; ++j) s = zadd(s, zmul(ld(Hm, h(i, j)), ld(Im, h(j, c))));
            z2 den = zsub(ld(Hm, h(i, i)), lam);
            if (zabs(den) < small) den = z2{small, 0.0};
            st(Im, h(i, c), zscale(zdiv(s, den), -1.0));
        }
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) {
            z2 s{0.0, 0.0};
            for (int j = 0; j <= c; ++j) s = zadd(s, zmul(ld(Zm, h(i, j)), ld(Im, h(j, c))));
            st(Xk, h(i, c), s);
            nrm += znorm(s);
        }
        nrm = sqrt(nrm);
        for (int i = 0; i < n; ++i) st(Xk, h(i, c), zscale(ld(Xk, h(i, c)), 1.0 / nrm));
    }
    __threadfence_block();
    __syncthreads();
    if (t == 0) {
        for (int i = 0; i < n; ++i) perm[i] = i;
        if (!sort_overlap) {
            // sort by ascending real part in the reference's exchange order (sort_eigen)
            for (int i = 0; i < n; ++i)
                for (int j = i + 1; j < n; ++j)
                    if (ld(wv, perm[i]).re > ld(wv, perm[j]).re) {
                        const int p = perm[i];
                        perm[i] = perm[j];
                        perm[j] = p;
                    }
        } else {
            // sort_eigen_overlap (diagonalization_utils.cpp:121-167): position i takes the
            // not yet chosen eigenvector with the largest |X(i, j)|^2 (the first on ties)
            for (int i = 0; i < n; ++i) {
                double best = 0.0;
                int js = -1, free0 = -1;
                for (int j = 0; j < n; ++j) {
                    bool used = false;
                    for (int q = 0; q < i; ++q) used = used || perm[q] == j;
                    if (used) continue;
                    if (free0 < 0) free0 = j;
                    const double a = znorm(ld(Xk, h(i, j)));
                    if (best < a) {
                        best = a;
                        js = j;
                    }
                }
                perm[i] = js >= 0 ? js : free0;  // an all-zero row keeps the first free column
            }
        }
    }
    __syncthreads();
    for (int i = t; i < n; i += kEigLanes) {  // Hm <- X with sorted columns, lane per row
        for (int c = 0; c < n; ++c) st(Hm, h(i, c), ld(Xk, h(i, perm[c])));
        st(wk, i, ld(wv, perm[i]));
    }
    __syncthreads();
    for (int i = t; i < n; i += kEigLanes)
        for (int c = 0; c < n; ++c) st(Xk, h(i, c), ld(Hm, h(i, c)));
    __threadfence_block();
    __syncthreads();

    const double* L = Xk;
    if (want_left) {  // Gauss-Jordan with partial pivoting on Hm = X, Im = I
        double* XLk = XL + 2 * msz * k;
        for (int e = t; e < n * n; e += kEigLanes) st(Im, h(e / n, e % n), z2{e / n == e % n ? 1.0 : 0.0, 0.0});
        __syncthreads();
        for (int c = 0; c < n; ++c) {
            int p = c;
            for (int r = c + 1; r < n; ++r)
                if (zabs(ld(Hm, h(r, c))) > zabs(ld(Hm, h(p, c)))) p = r;
            if (zabs(ld(Hm, h(p, c))) == 0.0) {  // uniform
                if (t == 0) status[k] = 2;
                return;
            }
            __syncthreads();
            if (p != c)
                for (int j = t; j < n; j += kEigLanes) {
                    const z2 a1 = ld(Hm, h(p, j)), b1 = ld(Im, h(p, j));
                    st(Hm, h(p, j), ld(Hm, h(c, j)));
                    st(Im, h(p, j), ld(Im, h(c, j)));
                    st(Hm, h(c, j), a1);
                    st(Im, h(c, j), b1);
                }
            __syncthreads();
            const z2 piv = ld(Hm, h(c, c));
            __syncthreads();
            for (int j = t; j < n; j += kEigLanes) {
                st(Hm, h(c, j), zdiv(ld(Hm, h(c, j)), piv));
                st(Im, h(c, j), zdiv(ld(Im, h(c, j)), piv));
            }
            __syncthreads();
            for (int r = t; r < n; r += kEigLanes) {  // lane per row
                if (r == c) continue;
                const z2 f = ld(Hm, h(r, c));
                if (f.re == 0.0 && f.im == 0.0) continue;
                for (int j = 0; j < n; ++j) {
                    st(Hm, h(r, j), zsub(ld(Hm, h(r, j)), zmul(f, ld(Hm, h(c, j)))));
                    st(Im, h(r, j), zsub(ld(Im, h(r, j)), zmul(f, ld(Im, h(c, j)))));
                }
            }
            __syncthreads();
        }
        for (int e = t; e < n * n; e += kEigLanes) {
            const int i = e / n, j = e % n;
            st(XLk, h(i, j), zconj(ld(Im, h(j, i))));
        }
        __threadfence_block();
        __syncthreads();
        L = XLk;
    }
    if (!en) return;
    for (int e = t; e < 3 * nb[k]; e += kEigLanes) {  // e_q of band b = (X_L^H Hs_q X)_bb
        const int b = e / 3, q = e % 3;
        z2 acc{0.0, 0.0};
        if (q < nH) {
            const double* Hq = Hs + 2 * ((long long)q * nk + k) * msz;
            for (int i = 0; i < n; ++i) {
                z2 s{0.0, 0.0};
                for (int j = 0; j < n; ++j) s = zadd(s, zmul(ld(Hq, h(i, j)), ld(Xk, h(j, b))));
                acc = zadd(acc, zmul(zconj(ld(L, h(i, b))), s));
            }
        }
        st(en, (long long)(band0[k] + b) * 3 + q, acc);
    }
}

// ----------------------------------------------------------------- launchers
namespace {
inline unsigned blocks_of(int n) { return (unsigned)((n + kRed - 1) / kRed); }
}  // namespace

void launch_dv_precond(const double* V, const long long* voff, const int* first, const int* cnt, int max_cnt,
                       const double* H0, const double* H1, const double* H2, const double* kG2, const long long* goff,
                       const double* eig, const int* band0, double* T, const long long* toff, const int* npw, int nk,
                       hipStream_t s)
{
    if (max_cnt <= 0 || nk <= 0) return;
    hipLaunchKernelGGL(k_dv_precond, dim3(nk, max_cnt), dim3(kRed), 0, s, V, voff, first, cnt, H0, H1, H2, kG2, goff,
                       eig, band0, T, toff, npw);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_dv_project(const double* A, const long long* aoff, const int* nA, int max_nA, double* B,
                       const long long* boff, const int* bidx, const int* npw, int max_npw, int nk, double* c,
                       double* nrm, hipStream_t s, const double* D)
{
    if (nk <= 0) return;
    if (max_nA > 0) {
        hipLaunchKernelGGL(k_dv_dots, dim3(max_nA, nk), dim3(kRed), 0, s, D ? D : A, aoff, nA, B, boff, bidx, npw, c,
                           max_nA);
        TCX_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_dv_subtract, dim3(blocks_of(max_npw), nk), dim3(kRed), 0, s, A, aoff, nA, B, boff, bidx,
                           npw, c, max_nA);
        TCX_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_dv_norm, dim3(nk), dim3(kRed), 0, s, B, boff, bidx, npw, nrm);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_dv_store(const double* B, const long long* boff, const int* bidx, const double* nrm, double* O,
                     const long long* ooff, const int* odst, const int* npw, int max_npw, int nk, hipStream_t s)
{
    if (nk <= 0) return;
    hipLaunchKernelGGL(k_dv_store, dim3(blocks_of(max_npw), nk), dim3(kRed), 0, s, B, boff, bidx, nrm, O, ooff, odst,
                       npw);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_dv_copy(const double* S, const long long* soff, const int* sfirst, double* D, const long long* doff,
                    const int* dfirst, const int* cnt, int max_cnt, const int* npw, int max_npw, int nk, hipStream_t s)
{
    if (max_cnt <= 0 || nk <= 0) return;
    hipLaunchKernelGGL(k_dv_copy, dim3(blocks_of(max_npw), max_cnt, nk), dim3(kRed), 0, s, S, soff, sfirst, D, doff,
                       dfirst, cnt, npw);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_dv_gram(const double* A, const long long* aoff, const int* nA, const double* B0, const double* B1,
                    const double* B2, int nB, const int* npw, int nk, double* G, int gs, hipStream_t s)
{
    if (nk <= 0 || gs <= 0) return;
    hipLaunchKernelGGL(k_dv_gram, dim3(gs * gs, nk, nB), dim3(kRed), 0, s, A, aoff, nA, B0, B1, B2, npw, G, gs);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_dv_rotate(const double* A, const long long* aoff, const int* nA, const double* C, int gs, double* O,
                      const long long* ooff, const int* nout, const int* npw, int max_npw, int nk, hipStream_t s)
{
    if (nk <= 0) return;
    hipLaunchKernelGGL(k_dv_rotate, dim3(blocks_of(max_npw), nk), dim3(kRed), 0, s, A, aoff, nA, C, gs, O, ooff, nout,
                       npw);
    TCX_HIP_CHECK(hipGetLastError());
}

size_t dv_eig_work_bytes(int gs, int nk) { return (size_t)nk * dv_eig_stride(gs) * 8; }

bool dv_eig_fits_lds(int gs) { return dv_eig_bytes(gs) <= 64 * 1024; }

void launch_dv_eig(const double* Hs, int nH, const int* dim, int gs, int nk, int want_left, int sort_overlap,
                   double* w, double* X, double* XL, double* gwork, const int* nb, const int* band0, double* en,
                   int* status, hipStream_t s)
{
    if (nk <= 0) return;
    const bool lds = dv_eig_fits_lds(gs);
    hipLaunchKernelGGL(k_dv_eig, dim3(nk), dim3(kEigLanes), lds ? dv_eig_bytes(gs) : 0, s, Hs, nH, dim, gs, want_left,
                       sort_overlap, w, X, XL, gwork, lds ? 1 : 0, nb, band0, en, status);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_dv_dot(const double* X, const long long* xoff, const int* xidx, const double* Y, const long long* yoff,
                   const int* yidx, const int* npw, int nk, double* c, hipStream_t s)
{
    if (nk <= 0) return;
    hipLaunchKernelGGL(k_dv_dot, dim3(nk), dim3(kRed), 0, s, X, xoff, xidx, Y, yoff, yidx, npw, c);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_dv_cdiv_store(const double* B, const long long* boff, const int* bidx, const double* c, double* O,
                          const long long* ooff, const int* odst, const int* npw, int max_npw, int nk, hipStream_t s)
{
    if (nk <= 0) return;
    hipLaunchKernelGGL(k_dv_cdiv_store, dim3(blocks_of(max_npw), nk), dim3(kRed), 0, s, B, boff, bidx, c, O, ooff,
                       odst, npw);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_dv_binorm(double* R, double* L, const long long* off, const int* nb, int max_nb, const int* band0,
                      const int* npw, int nk, double* out, hipStream_t s)
{
    if (nk <= 0 || max_nb <= 0) return;
    hipLaunchKernelGGL(k_dv_binorm, dim3(nk, max_nb), dim3(kRed), 0, s, R, L, off, nb, band0, npw, out);
    TCX_HIP_CHECK(hipGetLastError());
}

}  // namespace tcx
