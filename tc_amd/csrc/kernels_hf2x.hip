// kernels_hf2x.hip -- fused pair pipeline of the Fock exchange (hf2x).
//
// For every pair (k,j | q,i) of a launch the reference computes
// (src/calc_hamiltonian_hf2x.cpp:123-135)
//     rho(r)  = conj(phi_qi(r)) phi_kj(r)
//     rho~    = FFT_forward(rho)            (1/N)
//     rho~   *= v(G) = 1/|k-q+G|^2  (0 below 1e-8, :91-104)
//     V(r)    = FFT_backward(rho~)
//     W_kj(r) += f_qi V(r) phi_qi(r)
// Here the two 3D transforms are split into three streaming passes so that
// every pass is one coalesced read and one coalesced write of the pair grid:
//   A  (z-pencils):  form rho from the two cached real-space orbitals,
//                    forward FFT along z, store;
//   B  (xy-planes):  forward FFT along y and x, multiply by v(G)/N,
//                    inverse FFT along x and y, store (LDS transposes);
//   C  (z-pencils):  inverse FFT along z, multiply by f phi_qi, accumulate
//                    over all pairs of the target (k,j) in FP64 registers,
//                    and add once into the FP64 grid W_kj.
// No atomics: every target (k,j) of a launch is owned by one set of threads,
// so the FP64 sums are deterministic.
#include "kernels_common.hpp"
#include "size_switch.hpp"

namespace tcx {

// The pair record is written once and read by the next pass only: streaming
// stores keep it from displacing the orbital grids in L2 / MALL.
template <typename T, int NZ>
__global__ __launch_bounds__(kPencilBlock) void k_hf2x_A(const cpx<T>* __restrict__ phiR_app,
                                                         const cpx<T>* __restrict__ phiR_occ,
                                                         const PairEntry* __restrict__ pairs,
                                                         cpx<T>* __restrict__ buf, int nxy)
{
    const int xy = blockIdx.x * kPencilBlock + threadIdx.x;
    if (xy >= nxy) return;
    const PairEntry pe = pairs[blockIdx.y];
    const size_t ng = (size_t)nxy * NZ;
    const cpx<T>* a = phiR_app + (size_t)pe.app * ng + xy;
    const cpx<T>* b = phiR_occ + (size_t)pe.occ * ng + xy;
    cpx<T> v[NZ];
    static_for<0, NZ>([&](auto zc) {
        constexpr int z = decltype(zc)::value;
        v[z] = cmulc(b[(size_t)z * nxy], a[(size_t)z * nxy]);
    });
    fft<T, NZ, -1>(v);
    store_pencil_nt<T, NZ>(v, buf + (size_t)blockIdx.y * ng + xy, nxy);
}

// Coulomb kernel table of a launch: v(G)/N = 1/(N |k-q+G|^2), 0 below 1e-8
// (hf2x.cpp:91-104; the 1/N is FFT_forward's), one row per k-q of the launch,
// evaluated in FP64 and rounded once to the working precision.  Within a plane
// the table is stored y-fastest, e = i2 + n2 (i1 + n1 i3): pass B holds one
// y line per lane when it multiplies, so its table loads are coalesced.
template <typename T>
__global__ __launch_bounds__(256) void k_coulomb_table(T* __restrict__ table, const double* __restrict__ kqcart,
                                                       int row0, Geom g, long long N, double invN)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= N) return;
    const int i2 = (int)(e % g.n2), i1 = (int)((e / g.n2) % g.n1), i3 = (int)(e / ((long long)g.n1 * g.n2));
    const int m1 = centered(i1, g.n1), m2 = centered(i2, g.n2), m3 = centered(i3, g.n3);
    const double* kq = kqcart + 3 * (row0 + blockIdx.y);
    double g2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double c = kq[d] + g.b[d] * m1 + g.b[3 + d] * m2 + g.b[6 + d] * m3;
        g2 += c * c;
    }
    table[(long long)blockIdx.y * N + e] = static_cast<T>(g2 < 1e-8 ? 0.0 : invN / g2);
}

// LDS tile of one component (real or imaginary part) of a plane, [y][x] with one
// element of padding per row.  The FP64 pass B transposes the two parts one after
// the other: that halves its LDS per plane and doubles its waves per SIMD (1 -> 2 at
// 40^3).  FP32 keeps whole complex tiles: split tiles would double its waves too
// (2 -> 4), but with twice the LDS instructions and three times the barriers the
// pass ran 5% slower.
template <typename T, int NP>
struct PartTile {
    T a[NP][NP + 1];
};
template <int C, typename T>
__device__ __forceinline__ T& part(cpx<T>& z)
{
    if constexpr (C == 0) return z.re; else return z.im;
}

// Transposes of the PL planes of a workgroup: thread t (active) holds column x=t
// (values along y) -> row y=t (values along x), and back.  Every thread of the
// workgroup calls them (they contain the barriers).
template <typename T, int NP>
__device__ __forceinline__ void cols_to_rows(PlaneTile<T, NP>& s, cpx<T> (&v)[NP], int t, bool active)
{
    if (active) tile_put_col<T, NP>(s, v, t);
    __syncthreads();
    if (active) tile_get_row<T, NP>(s, v, t);
}
template <typename T, int NP>
__device__ __forceinline__ void rows_to_cols(PlaneTile<T, NP>& s, cpx<T> (&v)[NP], int t, bool active)
{
    if (active) tile_put_row<T, NP>(s, v, t);  // row t was read by thread t alone
    __syncthreads();
    if (active) tile_get_col<T, NP>(s, v, t);
}
template <typename T, int NP>
__device__ __forceinline__ void cols_to_rows(PartTile<T, NP>& s, cpx<T> (&v)[NP], int t, bool active)
{
    static_for<0, 2>([&](auto cc) {
        constexpr int C = decltype(cc)::value;
        if constexpr (C == 1) __syncthreads();  // every thread has read the real parts
        if (active) static_for<0, NP>([&](auto yc) { s.a[decltype(yc)::value][t] = part<C>(v[decltype(yc)::value]); });
        __syncthreads();
        if (active) static_for<0, NP>([&](auto xc) { part<C>(v[decltype(xc)::value]) = s.a[t][decltype(xc)::value]; });
    });
}
template <typename T, int NP>
__device__ __forceinline__ void rows_to_cols(PartTile<T, NP>& s, cpx<T> (&v)[NP], int t, bool active)
{
    static_for<0, 2>([&](auto cc) {
        constexpr int C = decltype(cc)::value;
        if constexpr (C == 1) __syncthreads();
        if (active) static_for<0, NP>([&](auto xc) { s.a[t][decltype(xc)::value] = part<C>(v[decltype(xc)::value]); });
        __syncthreads();
        if (active) static_for<0, NP>([&](auto yc) { part<C>(v[decltype(yc)::value]) = s.a[decltype(yc)::value][t]; });
    });
}

// One workgroup transforms PL planes of one pair (PlaneGroup): thread t owns
// line t % NP of plane t / NP.  Column load (coalesced) -> FFT_y -> LDS
// transpose -> FFT_x -> x v(G)/N -> inverse FFT_x -> LDS transpose -> inverse
// FFT_y -> column store.  Each plane is read and written once (streaming).
template <typename T, int NP>
__global__ __launch_bounds__((PlaneGroup<T, NP>::threads)) void k_hf2x_B(cpx<T>* __restrict__ buf,
                                                                       const PairEntry* __restrict__ pairs,
                                                                       const T* __restrict__ table, int kq0, int nz)
{
    constexpr int PL = PlaneGroup<T, NP>::planes;
    using Tile = std::conditional_t<sizeof(T) == 8, PartTile<T, NP>, PlaneTile<T, NP>>;
    __shared__ Tile tiles[PL];
    const int p = threadIdx.x / NP, t = threadIdx.x % NP;
    const int z = blockIdx.x * PL + p;
    const bool active = p < PL && z < nz;
    const size_t ng = (size_t)NP * NP * nz;
    cpx<T>* plane = buf + (size_t)blockIdx.y * ng + (size_t)(active ? z : 0) * (NP * NP);
    Tile& tile = tiles[active ? p : 0];
    cpx<T> v[NP];
    if (active) {
        load_pencil_nt<T, NP>(v, plane + t, NP);
        fft<T, NP, -1>(v);
    }
    cols_to_rows<T, NP>(tile, v, t, active);
    if (active) {
        fft<T, NP, -1>(v);
        // frequency (x, y=t, z); the table is y-fastest within the plane
        const T* w = table + (size_t)(pairs[blockIdx.y].kq - kq0) * ng + (size_t)z * NP * NP + t;
        static_for<0, NP>([&](auto xc) {
            constexpr int x = decltype(xc)::value;
            v[x] = scale(v[x], w[x * NP]);
        });
        fft<T, NP, 1>(v);
    }
    rows_to_cols<T, NP>(tile, v, t, active);
    if (active) {
        fft<T, NP, 1>(v);
        store_pencil_nt<T, NP>(v, plane + t, NP);
    }
}

// Each pencil is split over S lanes (PencilSplit), so a 64-lane workgroup
// covers 64/S pencils and each lane keeps NZ/S FP64 accumulators.
template <typename T, int NZ>
__global__ __launch_bounds__(kPencilBlock) void k_hf2x_C(const cpx<T>* __restrict__ buf,
                                                         const cpx<T>* __restrict__ phiR_occ,
                                                         const PairEntry* __restrict__ pairs,
                                                         const TargetEntry* __restrict__ targets,
                                                         const double* __restrict__ fill,
                                                         cpx<double>* __restrict__ W, float2* __restrict__ Wc,
                                                         int nxy)
{
    constexpr int S = SplitOf<T, NZ>::value;
    using Split = PencilSplit<S>;
    constexpr int M = NZ / S;
    const int h = threadIdx.x / Split::pencils;
    const int xy0 = blockIdx.x * Split::pencils + (threadIdx.x % Split::pencils);
    const bool valid = xy0 < nxy;
    const int xy = valid ? xy0 : 0;  // out-of-range lanes still join the shuffles
    const int ob = Split::out_block(h);
    const TargetEntry te = targets[blockIdx.y];
    const size_t ng = (size_t)nxy * NZ;
    cpx<double> acc[M];
    static_for<0, M>([&](auto kc) { acc[decltype(kc)::value] = cpx<double>{0.0, 0.0}; });
    // the pencil of pair pp + 1 is loaded before pair pp is transformed (FP32)
    constexpr bool kPrefetch = sizeof(T) == 4 && M <= 20;  // 48^3 pencils would drop to one wave
    auto grid_of = [&](int pp) { return buf + (size_t)pp * ng + xy + (size_t)h * nxy; };
    cpx<T> vn[M];
    if (kPrefetch && te.npair > 0) load_pencil<T, M>(vn, grid_of(te.pair0), S * nxy);
#pragma unroll 1
    for (int pp = te.pair0; pp < te.pair0 + te.npair; ++pp) {
        const int occ = pairs[pp].occ;
        cpx<T> v[M];
        if (kPrefetch) {
            static_for<0, M>([&](auto mc) { v[decltype(mc)::value] = vn[decltype(mc)::value]; });
            if (pp + 1 < te.pair0 + te.npair) load_pencil<T, M>(vn, grid_of(pp + 1), S * nxy);
        } else {
            load_pencil<T, M>(v, grid_of(pp), S * nxy);  // z = S m + h
        }
        Split::template fft<T, NZ, 1, kSwapF64<T>>(v, h);  // -> z = ob*M + k
        const cpx<T>* q = phiR_occ + (size_t)occ * ng + xy + (size_t)ob * M * nxy;
        const double f = fill ? fill[occ] : 1.0;
        static_for<0, M>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const cpx<T> prod = v[k] * q[(size_t)k * nxy];
            acc[k].re += f * static_cast<double>(prod.re);
            acc[k].im += f * static_cast<double>(prod.im);
        });
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
template <typename T>
static void hf2x_A_impl(const Geom& g, const void* app, const void* occ, const PairEntry* pairs, int npairs,
                        void* buf, hipStream_t s)
{
    const int nxy = g.n1 * g.n2;
    const dim3 grid((nxy + kPencilBlock - 1) / kPencilBlock, npairs);
#define TCX_A(NZ)                                                                                         \
    hipLaunchKernelGGL((k_hf2x_A<T, NZ>), grid, dim3(kPencilBlock), 0, s, (const cpx<T>*)app,             \
                       (const cpx<T>*)occ, pairs, (cpx<T>*)buf, nxy);
    TCX_SIZE_SWITCH(g.n3, TCX_A)
#undef TCX_A
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_hf2x_A(int prec, const Geom& g, const void* app, const void* occ, const PairEntry* pairs, int npairs,
                   void* buf, hipStream_t s)
{
    if (npairs == 0) return;
    if (npairs > 65535) throw Error(TCX_ERR_INTERNAL, "hf2x: too many pairs in one launch");
    if (prec == TCX_PREC_FP64)
        hf2x_A_impl<double>(g, app, occ, pairs, npairs, buf, s);
    else
        hf2x_A_impl<float>(g, app, occ, pairs, npairs, buf, s);
}

template <typename T>
static void hf2x_B_impl(const Geom& g, void* buf, const PairEntry* pairs, int npairs, const void* table, int kq0,
                        hipStream_t s)
{
    if (g.n1 != g.n2) throw Error(TCX_ERR_UNSUPPORTED, "FFT grid must have n1 == n2");
#define TCX_B(NP)                                                                                         \
    {                                                                                                     \
        using PG = PlaneGroup<T, NP>;                                                                     \
        hipLaunchKernelGGL((k_hf2x_B<T, NP>), dim3((g.n3 + PG::planes - 1) / PG::planes, npairs),         \
                           dim3(PG::threads), 0, s, (cpx<T>*)buf, pairs, (const T*)table, kq0, g.n3);     \
    }
    TCX_SIZE_SWITCH(g.n1, TCX_B)
#undef TCX_B
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_coulomb_table(int prec, void* table, const double* kqcart, int row0, int nrows, const Geom& g,
                          hipStream_t s)
{
    if (nrows == 0) return;
    const long long N = (long long)g.n1 * g.n2 * g.n3;
    const dim3 grid((unsigned)((N + 255) / 256), nrows);
    if (prec == TCX_PREC_FP64)
        hipLaunchKernelGGL(k_coulomb_table<double>, grid, dim3(256), 0, s, (double*)table, kqcart, row0, g, N,
                           1.0 / (double)N);
    else
        hipLaunchKernelGGL(k_coulomb_table<float>, grid, dim3(256), 0, s, (float*)table, kqcart, row0, g, N,
                           1.0 / (double)N);
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_hf2x_B(int prec, const Geom& g, void* buf, const PairEntry* pairs, int npairs, const void* table,
                   int kq0, hipStream_t s)
{
    if (npairs == 0) return;
    if (prec == TCX_PREC_FP64)
        hf2x_B_impl<double>(g, buf, pairs, npairs, table, kq0, s);
    else
        hf2x_B_impl<float>(g, buf, pairs, npairs, table, kq0, s);
}

template <typename T>
static void hf2x_C_impl(const Geom& g, const void* buf, const void* occ, const PairEntry* pairs,
                        const TargetEntry* targets, int ntargets, const double* fill, double* W, float* Wc,
                        hipStream_t s)
{
    const int nxy = g.n1 * g.n2;
    // 64/S split pencils per 64-lane workgroup
#define TCX_C(NZ)                                                                                          \
    {                                                                                                      \
        constexpr int P = 64 / SplitOf<T, NZ>::value;                                                      \
        hipLaunchKernelGGL((k_hf2x_C<T, NZ>), dim3((nxy + P - 1) / P, ntargets), dim3(kPencilBlock), 0, s, \
                           (const cpx<T>*)buf, (const cpx<T>*)occ, pairs, targets, fill, (cpx<double>*)W,  \
                           (float2*)Wc, nxy);                                                              \
    }
    TCX_SIZE_SWITCH(g.n3, TCX_C)
#undef TCX_C
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_hf2x_C(int prec, const Geom& g, const void* buf, const void* occ, const PairEntry* pairs,
                   const TargetEntry* targets, int ntargets, const double* fill, double* W, float* Wc, hipStream_t s)
{
    if (ntargets == 0) return;
    if (ntargets > 65535) throw Error(TCX_ERR_INTERNAL, "hf2x: too many targets in one launch");
    if (prec == TCX_PREC_FP64)
        hf2x_C_impl<double>(g, buf, occ, pairs, targets, ntargets, fill, W, Wc, s);
    else
        hf2x_C_impl<float>(g, buf, occ, pairs, targets, ntargets, fill, W, Wc, s);
}

}  // namespace tcx
