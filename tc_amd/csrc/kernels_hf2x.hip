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
    store_pencil<T, NZ>(v, buf + (size_t)blockIdx.y * ng + xy, nxy);
}

template <typename T, int NP>
__global__ __launch_bounds__(kPlaneBlock) void k_hf2x_B(cpx<T>* __restrict__ buf, const PairEntry* __restrict__ pairs,
                                                        const double* __restrict__ kqcart, int nz, Geom g,
                                                        double invN)
{
    __shared__ PlaneTile<T, NP> tile;
    const int t = threadIdx.x;
    const int z = blockIdx.x;
    cpx<T>* plane = buf + ((size_t)blockIdx.y * nz + z) * (NP * NP);
    cpx<T> v[NP];
    if (t < NP) {
        load_pencil<T, NP>(v, plane + t, NP);
        fft<T, NP, -1>(v);
        tile_put_col<T, NP>(tile, v, t);
    }
    __syncthreads();
    if (t < NP) {
        tile_get_row<T, NP>(tile, v, t);
        fft<T, NP, -1>(v);
        // frequency (x, y=t, z): Cartesian k-q+G, Miller wrapped (plane_wave_basis.cpp:205-216);
        // v = 1/|k-q+G|^2, 0 below 1e-8 (hf2x.cpp:91-104), times the 1/N of FFT_forward.
        // Evaluated in the working precision: the row offset is formed in FP64 and
        // rounded once, so the FP32 kernel carries one rounding of |k-q+G|^2.
        const int kq = pairs[blockIdx.y].kq;
        const int my = centered(t, NP), mz = centered(z, nz);
        const T bx = static_cast<T>(kqcart[3 * kq + 0] + g.b[3] * my + g.b[6] * mz);
        const T by = static_cast<T>(kqcart[3 * kq + 1] + g.b[4] * my + g.b[7] * mz);
        const T bz = static_cast<T>(kqcart[3 * kq + 2] + g.b[5] * my + g.b[8] * mz);
        const T b0 = static_cast<T>(g.b[0]), b1 = static_cast<T>(g.b[1]), b2 = static_cast<T>(g.b[2]);
        const T sN = static_cast<T>(invN);
        static_for<0, NP>([&](auto xc) {
            constexpr int x = decltype(xc)::value;
            constexpr T mx = static_cast<T>(x <= NP / 2 ? x : x - NP);
            const T gx = bx + b0 * mx, gy = by + b1 * mx, gz = bz + b2 * mx;
            const T g2 = gx * gx + gy * gy + gz * gz;
            const T w = g2 < T(1e-8) ? T(0) : sN / g2;
            v[x] = scale(v[x], w);
        });
        fft<T, NP, 1>(v);
        tile_put_row<T, NP>(tile, v, t);
    }
    __syncthreads();
    if (t < NP) {
        tile_get_col<T, NP>(tile, v, t);
        fft<T, NP, 1>(v);
        store_pencil<T, NP>(v, plane + t, NP);
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
                                                         cpx<double>* __restrict__ W, int nxy)
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
    for (int pp = te.pair0; pp < te.pair0 + te.npair; ++pp) {
        const int occ = pairs[pp].occ;
        cpx<T> v[M];
        load_pencil<T, M>(v, buf + (size_t)pp * ng + xy + (size_t)h * nxy, S * nxy);  // z = S m + h
        Split::template fft<T, NZ, 1>(v, h);                                         // -> z = ob*M + k
        const cpx<T>* q = phiR_occ + (size_t)occ * ng + xy + (size_t)ob * M * nxy;
        const double f = fill[occ];
        static_for<0, M>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const cpx<T> prod = v[k] * q[(size_t)k * nxy];
            acc[k].re += f * static_cast<double>(prod.re);
            acc[k].im += f * static_cast<double>(prod.im);
        });
    }
    if (!valid) return;
    cpx<double>* w = W + (size_t)te.app * ng + xy + (size_t)ob * M * nxy;
    static_for<0, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        w[(size_t)k * nxy] = w[(size_t)k * nxy] + acc[k];
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
static void hf2x_B_impl(const Geom& g, void* buf, const PairEntry* pairs, int npairs, const double* kqcart,
                        hipStream_t s)
{
    if (g.n1 != g.n2) throw Error(TCX_ERR_UNSUPPORTED, "FFT grid must have n1 == n2");
    const double invN = 1.0 / ((double)g.n1 * g.n2 * g.n3);
#define TCX_B(NP)                                                                                       \
    hipLaunchKernelGGL((k_hf2x_B<T, NP>), dim3(g.n3, npairs), dim3(kPlaneBlock), 0, s, (cpx<T>*)buf,    \
                       pairs, kqcart, g.n3, g, invN);
    TCX_SIZE_SWITCH(g.n1, TCX_B)
#undef TCX_B
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_hf2x_B(int prec, const Geom& g, void* buf, const PairEntry* pairs, int npairs, const double* kqcart,
                   hipStream_t s)
{
    if (npairs == 0) return;
    if (prec == TCX_PREC_FP64)
        hf2x_B_impl<double>(g, buf, pairs, npairs, kqcart, s);
    else
        hf2x_B_impl<float>(g, buf, pairs, npairs, kqcart, s);
}

template <typename T>
static void hf2x_C_impl(const Geom& g, const void* buf, const void* occ, const PairEntry* pairs,
                        const TargetEntry* targets, int ntargets, const double* fill, double* W, hipStream_t s)
{
    const int nxy = g.n1 * g.n2;
    // 64/S split pencils per 64-lane workgroup
#define TCX_C(NZ)                                                                                             \
    {                                                                                                         \
        constexpr int P = 64 / SplitOf<T, NZ>::value;                                                         \
        hipLaunchKernelGGL((k_hf2x_C<T, NZ>), dim3((nxy + P - 1) / P, ntargets), dim3(kPencilBlock), 0, s,    \
                           (const cpx<T>*)buf, (const cpx<T>*)occ, pairs, targets, fill, (cpx<double>*)W, nxy); \
    }
    TCX_SIZE_SWITCH(g.n3, TCX_C)
#undef TCX_C
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_hf2x_C(int prec, const Geom& g, const void* buf, const void* occ, const PairEntry* pairs,
                   const TargetEntry* targets, int ntargets, const double* fill, double* W, hipStream_t s)
{
    if (ntargets == 0) return;
    if (ntargets > 65535) throw Error(TCX_ERR_INTERNAL, "hf2x: too many targets in one launch");
    if (prec == TCX_PREC_FP64)
        hf2x_C_impl<double>(g, buf, occ, pairs, targets, ntargets, fill, W, s);
    else
        hf2x_C_impl<float>(g, buf, occ, pairs, targets, ntargets, fill, W, s);
}

}  // namespace tcx
