// kernels_tc2x.hip -- fused pair pipeline of the TC/BITC two-body exchange (tc2x).
//
// Per pair (k,j | q,i) the reference computes (src/calc_hamiltonian_tc2x.cpp:253-401)
//   rho~      = FFT_forward(conj(chi_qi) phi_kj)                         (:350-351)
//   sigma~_d  = FFT_forward(conj(chi_qi) g_kj,d),  g_kj,d = FFT_backward((k+G)_d phi_kj) (:372-376)
//   a_d       = FFT_backward(rho~ du_d),  du_d = (k-q+G)_d u(k-q+G)       (:354-358)
//   b         = FFT_backward(rho~ V2 + sum_d sigma~_d du_d), V2 = tc_2body(k-q+G) (:379-386)
//   W_kj     += f_qi [ b phi_qi - sum_d a_d g_qi,d ],  g_qi,d = FFT_backward((q+G)_d phi_qi) (:363-369, :387)
// (chi = phi for TC).  Eight 3D FFTs per pair, done here as three streaming passes
// over a 4-channel pair record [rho, sigma_0, sigma_1, sigma_2] -> [b, a_0, a_1, a_2]:
//   A (z-pencils): form the 4 channels from cached real-space orbitals, forward FFT along z;
//   B (xy-planes): forward FFT along y, x; combine the channels with the kernel table;
//                  inverse FFT along x, y (4 padded LDS tiles);
//   C (z-pencils): inverse FFT along z, multiply by phi_qi / g_qi,d, accumulate per target in FP64.
#include "jastrow_dev.hpp"
#include "kernels_common.hpp"
#include "size_switch.hpp"

namespace tcx {

// ----------------------------------------------------------------- kernel table (V2, u) per k-q row
template <typename T, bool kPoly>
__global__ __launch_bounds__(256) void k_tc_table(cpx<T>* __restrict__ table, const double* __restrict__ kqcart,
                                                  int row0, Geom g, JastrowPair jp, long long N)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= N) return;
    const int row = blockIdx.y;
    const int i1 = (int)(e % g.n1), i2 = (int)((e / g.n1) % g.n2), i3 = (int)(e / ((long long)g.n1 * g.n2));
    const int m1 = centered(i1, g.n1), m2 = centered(i2, g.n2), m3 = centered(i3, g.n3);
    const double* kq = kqcart + 3 * (row0 + row);
    double G2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double c = kq[d] + g.b[d] * m1 + g.b[3 + d] * m2 + g.b[6 + d] * m3;
        G2 += c * c;
    }
    table[(long long)row * N + e] = cpx<T>{static_cast<T>(jastrow_tc2body<kPoly>(jp, G2)), static_cast<T>(jastrow_uk<kPoly>(jp, G2))};
}

// ----------------------------------------------------------------- pass A
template <typename T, int NZ>
__global__ __launch_bounds__(kPencilBlock) void k_tc2x_A(const cpx<T>* __restrict__ appR,
                                                         const cpx<T>* __restrict__ appGrad,
                                                         const cpx<T>* __restrict__ chiR,
                                                         const PairEntry* __restrict__ pairs,
                                                         cpx<T>* __restrict__ buf, int nxy)
{
    constexpr int S = SplitOf<T, NZ>::value;  // pencil split over S lanes, see PencilSplit
    using Split = PencilSplit<S>;
    constexpr int M = NZ / S;
    const int h = threadIdx.x / Split::pencils;
    const int xy0 = blockIdx.x * Split::pencils + (threadIdx.x % Split::pencils);
    const bool valid = xy0 < nxy;
    const int xy = valid ? xy0 : 0;
    const PairEntry pe = pairs[blockIdx.y];
    const size_t ng = (size_t)nxy * NZ;
    cpx<T> x[M];
    load_pencil<T, M>(x, chiR + (size_t)pe.occ * ng + xy + (size_t)h * nxy, S * nxy);           // z = S m + h
    cpx<T>* out = buf + (size_t)blockIdx.y * 4 * ng + xy + (size_t)Split::out_block(h) * M * nxy;  // z = ob*M + k
    for (int c = 0; c < 4; ++c) {
        const cpx<T>* src = (c == 0 ? appR + (size_t)pe.app * ng : appGrad + ((size_t)pe.app * 3 + (c - 1)) * ng) + xy +
                            (size_t)h * nxy;
        cpx<T> v[M];
        static_for<0, M>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            v[m] = cmulc(x[m], src[(size_t)(S * m) * nxy]);
        });
        Split::template fft<T, NZ, -1, kSwapF64<T>>(v, h);
        if (valid) store_pencil_nt<T, M>(v, out + (size_t)c * ng, nxy);  // streaming: read once by pass B
    }
}

// ----------------------------------------------------------------- pass B
// A workgroup owns P planes z of one pair, all four channels: P*4 plane-channels of NP
// lines each, packed onto consecutive lanes (lane l -> plane-channel l / NP, line l % NP)
// so that almost every lane transforms a line.  The channel combination at every G of
// the P planes is spread over all lanes of the workgroup.
// LDS per workgroup capped so that two or more workgroups share a CU (one 160 KB
// workgroup per CU measured 40% slower at 40^3 than 52 KB ones)
constexpr int kTc2xLdsCap = 80 * 1024;
template <typename T, int NP>
struct Tc2xPlanes {
    static constexpr int tile_bytes = NP * (NP + 1) * (int)sizeof(cpx<T>);
    static constexpr int pick()
    {
        int best = 1;
        double best_util = 0.0;
        for (int p = 1; p <= 8; ++p) {
            if (4 * p * tile_bytes > kTc2xLdsCap || 4 * p * NP > 1024) break;
            const int lanes = ((4 * p * NP + 63) / 64) * 64;
            const double util = double(4 * p * NP) / lanes;
            if (util > best_util + 1e-9) {
                best_util = util;
                best = p;
            }
        }
        return best;
    }
    static constexpr int planes = pick();
    static constexpr int threads = ((4 * planes * NP + 63) / 64) * 64;
};

template <typename T, int NP>
__global__ __launch_bounds__((Tc2xPlanes<T, NP>::threads)) void k_tc2x_B(cpx<T>* __restrict__ buf,
                                                                       const PairEntry* __restrict__ pairs,
                                                                       const cpx<T>* __restrict__ table, int kq0,
                                                                       const double* __restrict__ kqcart, int nz,
                                                                       Geom g, double invN)
{
    constexpr int P = Tc2xPlanes<T, NP>::planes;
    constexpr int NT = Tc2xPlanes<T, NP>::threads;
    __shared__ PlaneTile<T, NP> S[P * 4];
    const int pc = threadIdx.x / NP, t = threadIdx.x % NP;
    const int p = pc >> 2, c = pc & 3;  // plane slot and channel of this lane's line
    const int z = blockIdx.x * P + p;
    const bool on = pc < 4 * P && z < nz;
    const size_t ng = (size_t)NP * NP * nz;
    PlaneTile<T, NP>& tile = S[on ? pc : 0];
    cpx<T>* base = buf + ((size_t)blockIdx.y * 4 + (on ? c : 0)) * ng + (size_t)(on ? z : 0) * NP * NP;
    // (V2, u) at the thread's combine points e = threadIdx.x + i NT, loaded before the
    // forward transform so that the loads overlap it (coalesced over e)
    constexpr int EPT = (P * NP * NP + NT - 1) / NT;
    const PairEntry pe = pairs[blockIdx.y];
    const cpx<T>* tab = table + (size_t)(pe.kq - kq0) * ng + (size_t)blockIdx.x * P * NP * NP;
    const int nvalid = (nz - (int)blockIdx.x * P < P ? nz - (int)blockIdx.x * P : P) * NP * NP;
    cpx<T> vus[EPT];
    static_for<0, EPT>([&](auto ic) {
        const int e = threadIdx.x + decltype(ic)::value * NT;
        if (e < nvalid) vus[decltype(ic)::value] = tab[e];
    });
    cpx<T> v[NP];
    if (on) {
        load_pencil_nt<T, NP>(v, base + t, NP);  // column x = t; planes are streamed once
        fft<T, NP, -1>(v);                     // along y
        tile_put_col<T, NP>(tile, v, t);
    }
    __syncthreads();
    if (on) {
        tile_get_row<T, NP>(tile, v, t);       // row y = t
        fft<T, NP, -1>(v);                     // along x
        tile_put_row<T, NP>(tile, v, t);
    }
    __syncthreads();
    // combine at frequency (x, y, z): b = rho V2 + sum_d sigma_d du_d, a_d = rho du_d (1/N folded in)
    const double* kq = kqcart + 3 * pe.kq;
    static_for<0, EPT>([&](auto ic) {
        const int e = threadIdx.x + decltype(ic)::value * NT;
        if (e < nvalid) {
            const int q = e / (NP * NP), xy = e % (NP * NP);
            const int zz = blockIdx.x * P + q;
            const int x = xy % NP, y = xy / NP;
            const int mx = centered(x, NP), my = centered(y, NP), mz = centered(zz, nz);
            const cpx<T> vu = vus[decltype(ic)::value];
            const double us = (double)vu.im * invN;
            const T du0 = static_cast<T>((kq[0] + g.b[0] * mx + g.b[3] * my + g.b[6] * mz) * us);
            const T du1 = static_cast<T>((kq[1] + g.b[1] * mx + g.b[4] * my + g.b[7] * mz) * us);
            const T du2 = static_cast<T>((kq[2] + g.b[2] * mx + g.b[5] * my + g.b[8] * mz) * us);
            const T v2 = static_cast<T>((double)vu.re * invN);
            PlaneTile<T, NP>* Q = S + 4 * q;
            const cpx<T> r = Q[0].a[y][x];
            const cpx<T> s0 = Q[1].a[y][x], s1 = Q[2].a[y][x], s2 = Q[3].a[y][x];
            Q[0].a[y][x] = scale(r, v2) + scale(s0, du0) + scale(s1, du1) + scale(s2, du2);
            Q[1].a[y][x] = scale(r, du0);
            Q[2].a[y][x] = scale(r, du1);
            Q[3].a[y][x] = scale(r, du2);
        }
    });
    __syncthreads();
    if (on) {
        tile_get_row<T, NP>(tile, v, t);
        fft<T, NP, 1>(v);
        tile_put_row<T, NP>(tile, v, t);
    }
    __syncthreads();
    if (on) {
        tile_get_col<T, NP>(tile, v, t);
        fft<T, NP, 1>(v);
        store_pencil_nt<T, NP>(v, base + t, NP);
    }
}

// ----------------------------------------------------------------- pass C
template <typename T, int NZ>
__global__ __launch_bounds__(kPencilBlock) void k_tc2x_C(const cpx<T>* __restrict__ buf,
                                                         const cpx<T>* __restrict__ occR,
                                                         const cpx<T>* __restrict__ gradR,
                                                         const PairEntry* __restrict__ pairs,
                                                         const TargetEntry* __restrict__ targets,
                                                         const double* __restrict__ fill,
                                                         cpx<double>* __restrict__ W, float2* __restrict__ Wc,
                                                         int nxy)
{
    constexpr int S = SplitOf<T, NZ>::value;  // pencil split over S lanes, see PencilSplit
    using Split = PencilSplit<S>;
    constexpr int M = NZ / S;
    const int h = threadIdx.x / Split::pencils;
    const int xy0 = blockIdx.x * Split::pencils + (threadIdx.x % Split::pencils);
    const bool valid = xy0 < nxy;
    const int xy = valid ? xy0 : 0;
    const int ob = Split::out_block(h);
    const TargetEntry te = targets[blockIdx.y];
    const size_t ng = (size_t)nxy * NZ;
    cpx<double> acc[M];
    static_for<0, M>([&](auto kc) { acc[decltype(kc)::value] = cpx<double>{0.0, 0.0}; });
    // steps it = 4 (pair) + c; the pencil of step it + 1 is loaded before step it is
    // transformed (FP32)
    constexpr bool kPrefetch = sizeof(T) == 4 && M <= 20;  // 48^3 pencils would drop to one wave
    const int nsteps = 4 * te.npair;
    auto grid_of = [&](int it) { return buf + ((size_t)te.pair0 * 4 + it) * ng + xy + (size_t)h * nxy; };
    cpx<T> vn[M];
    if (kPrefetch && nsteps > 0) load_pencil<T, M>(vn, grid_of(0), S * nxy);
#pragma unroll 1
    for (int it = 0; it < nsteps; ++it) {
        const int pp = te.pair0 + it / 4, c = it % 4;
        const int occ = pairs[pp].occ;
        const double f = fill[occ];
        {
            cpx<T> v[M];
            if (kPrefetch) {
                static_for<0, M>([&](auto mc) { v[decltype(mc)::value] = vn[decltype(mc)::value]; });
                if (it + 1 < nsteps) load_pencil<T, M>(vn, grid_of(it + 1), S * nxy);
            } else {
                load_pencil<T, M>(v, grid_of(it), S * nxy);  // z = S m + h
            }
            Split::template fft<T, NZ, 1, kSwapF64<T>>(v, h);
            // channel 0: b * phi_qi (+f); channels 1..3: a_d * g_qi,d (-f)   [tc2x.cpp:363-369, :387-388]
            const cpx<T>* m = (c == 0 ? occR + (size_t)occ * ng : gradR + ((size_t)occ * 3 + (c - 1)) * ng) + xy +
                              (size_t)ob * M * nxy;
            const double s = c == 0 ? f : -f;
            static_for<0, M>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                const cpx<T> prod = v[k] * m[(size_t)k * nxy];
                acc[k].re += s * static_cast<double>(prod.re);
                acc[k].im += s * static_cast<double>(prod.im);
            });
        }
    }
    if (!valid) return;
    const size_t o = (size_t)te.app * ng + xy + (size_t)ob * M * nxy;
    static_for<0, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        add_partial((double2*)(W + o + (size_t)k * nxy), Wc ? Wc + o + (size_t)k * nxy : nullptr, acc[k].re,
                    acc[k].im);
    });
}

// ----------------------------------------------------------------- launchers
void launch_tc_table(int prec, void* table, const double* kqcart, int row0, int nrows, const Geom& g,
                     const JastrowPair& jp, hipStream_t s)
{
    if (nrows == 0) return;
    const long long N = (long long)g.n1 * g.n2 * g.n3;
    const dim3 grid((unsigned)((N + 255) / 256), nrows);
    if (prec == TCX_PREC_FP64)
        TCX_POLY_DISPATCH(jp.pdeg > 0, hipLaunchKernelGGL((k_tc_table<double, kPoly>), grid, dim3(256), 0, s,
                                                          (cpx<double>*)table, kqcart, row0, g, jp, N));
    else
        TCX_POLY_DISPATCH(jp.pdeg > 0, hipLaunchKernelGGL((k_tc_table<float, kPoly>), grid, dim3(256), 0, s,
                                                          (cpx<float>*)table, kqcart, row0, g, jp, N));
    TCX_HIP_CHECK(hipGetLastError());
}

template <typename T>
static void tc2x_A_impl(const Geom& g, const void* appR, const void* appGrad, const void* chiR, const PairEntry* pairs,
                        int npairs, void* buf, hipStream_t s)
{
    const int nxy = g.n1 * g.n2;
#define TCX_A(NZ)                                                                                              \
    {                                                                                                          \
        constexpr int P = 64 / SplitOf<T, NZ>::value;                                                          \
        hipLaunchKernelGGL((k_tc2x_A<T, NZ>), dim3((nxy + P - 1) / P, npairs), dim3(kPencilBlock), 0, s,       \
                           (const cpx<T>*)appR, (const cpx<T>*)appGrad, (const cpx<T>*)chiR, pairs, (cpx<T>*)buf, \
                           nxy);                                                                               \
    }
    TCX_SIZE_SWITCH(g.n3, TCX_A)
#undef TCX_A
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_tc2x_A(int prec, const Geom& g, const void* appR, const void* appGrad, const void* chiR,
                   const PairEntry* pairs, int npairs, void* buf, hipStream_t s)
{
    if (npairs == 0) return;
    if (npairs > 65535) throw Error(TCX_ERR_INTERNAL, "tc2x: too many pairs in one launch");
    if (prec == TCX_PREC_FP64)
        tc2x_A_impl<double>(g, appR, appGrad, chiR, pairs, npairs, buf, s);
    else
        tc2x_A_impl<float>(g, appR, appGrad, chiR, pairs, npairs, buf, s);
}

template <typename T>
static void tc2x_B_impl(const Geom& g, void* buf, const PairEntry* pairs, int npairs, const void* table, int kq0,
                        const double* kqcart, hipStream_t s)
{
    if (g.n1 != g.n2) throw Error(TCX_ERR_UNSUPPORTED, "FFT grid must have n1 == n2");
    const double invN = 1.0 / ((double)g.n1 * g.n2 * g.n3);
#define TCX_B(NP)                                                                                         \
    {                                                                                                     \
        using TP = Tc2xPlanes<T, NP>;                                                                     \
        hipLaunchKernelGGL((k_tc2x_B<T, NP>), dim3((g.n3 + TP::planes - 1) / TP::planes, npairs),         \
                           dim3(TP::threads), 0, s, (cpx<T>*)buf, pairs, (const cpx<T>*)table, kq0, kqcart, \
                           g.n3, g, invN);                                                                \
    }
    TCX_SIZE_SWITCH(g.n1, TCX_B)
#undef TCX_B
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_tc2x_B(int prec, const Geom& g, void* buf, const PairEntry* pairs, int npairs, const void* table, int kq0,
                   const double* kqcart, hipStream_t s)
{
    if (npairs == 0) return;
    if (prec == TCX_PREC_FP64)
        tc2x_B_impl<double>(g, buf, pairs, npairs, table, kq0, kqcart, s);
    else
        tc2x_B_impl<float>(g, buf, pairs, npairs, table, kq0, kqcart, s);
}

template <typename T>
static void tc2x_C_impl(const Geom& g, const void* buf, const void* occR, const void* gradR, const PairEntry* pairs,
                        const TargetEntry* targets, int ntargets, const double* fill, double* W, float* Wc,
                        hipStream_t s)
{
    const int nxy = g.n1 * g.n2;
#define TCX_C(NZ)                                                                                           \
    {                                                                                                       \
        constexpr int P = 64 / SplitOf<T, NZ>::value;                                                       \
        hipLaunchKernelGGL((k_tc2x_C<T, NZ>), dim3((nxy + P - 1) / P, ntargets), dim3(kPencilBlock), 0, s,  \
                           (const cpx<T>*)buf, (const cpx<T>*)occR, (const cpx<T>*)gradR, pairs, targets, fill, \
                           (cpx<double>*)W, (float2*)Wc, nxy);                                              \
    }
    TCX_SIZE_SWITCH(g.n3, TCX_C)
#undef TCX_C
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_tc2x_C(int prec, const Geom& g, const void* buf, const void* occR, const void* gradR,
                   const PairEntry* pairs, const TargetEntry* targets, int ntargets, const double* fill, double* W,
                   float* Wc, hipStream_t s)
{
    if (ntargets == 0) return;
    if (ntargets > 65535) throw Error(TCX_ERR_INTERNAL, "tc2x: too many targets in one launch");
    if (prec == TCX_PREC_FP64)
        tc2x_C_impl<double>(g, buf, occR, gradR, pairs, targets, ntargets, fill, W, Wc, s);
    else
        tc2x_C_impl<float>(g, buf, occR, gradR, pairs, targets, ntargets, fill, W, Wc, s);
}

}  // namespace tcx
