// kernels_tc3z.hip -- z passes of FFT_backward fused with the real-space consumers of the
// three-body record pipelines.
//
// A record channel arrives inverse-transformed along x and y (the fused combine pass);
// one thread group owns a z pencil of one target and loops over the target's pairs: it
// finishes FFT_backward along z with the pencil split over S lanes (PencilSplit), then
// applies the consumer at the lane's output points, summing in FP64 registers, and adds
// once into the FP64 grids of the target (no atomics, deterministic).  The consumers are
// the reference's real-space steps:
//   tc3a3b3b4b6 (src/calc_hamiltonian_tc3a3b3b4b6.cpp:316-335): qsum_j,d += P_0,d phi_qi,
//     then P_s,d *= density_s, transformed along z again for the next FFT_forward;
//   tc3a2a4b2b5 (src/calc_hamiltonian_tc3a2a4b2b5.cpp:377-427): the q-sums Q2, Q4, Q5;
//     (:848-913) the apply-phase accumulation into W.
#include "kernels_common.hpp"
#include "size_switch.hpp"

namespace tcx {

namespace {

// lanes per pencil: Want when it divides NZ, else 2
template <int NZ, int Want>
struct SplitFor {
    static constexpr int value = (NZ % Want == 0) ? Want : 2;
};

template <typename T>
__device__ __forceinline__ double2 upd(cpx<T> a)
{
    return make_double2((double)a.re, (double)a.im);
}
__device__ __forceinline__ double2 cmadd(double2 acc, double2 a, double2 b, double s)  // acc + s a b
{
    return make_double2(acc.x + s * (a.x * b.x - a.y * b.y), acc.y + s * (a.x * b.y + a.y * b.x));
}

// Lane geometry of a split pencil.
template <int S>
struct LaneOf {
    int h, xy, ob;
    bool valid;
    __device__ LaneOf(int nxy)
    {
        using Split = PencilSplit<S>;
        h = threadIdx.x / Split::pencils;
        const int xy0 = blockIdx.x * Split::pencils + (threadIdx.x % Split::pencils);
        valid = xy0 < nxy;
        xy = valid ? xy0 : 0;  // out-of-range lanes still join the shuffles
        ob = Split::out_block(h);
    }
};

// v = FFT_backward along z of the pencil of grid g (split layout in, blocked layout out)
template <typename T, int NZ, int S>
__device__ __forceinline__ void zinv(cpx<T> (&v)[NZ / S], const cpx<T>* grid, const LaneOf<S>& L, int nxy)
{
    load_pencil<T, NZ / S>(v, grid + L.xy + (size_t)L.h * nxy, S * nxy);  // z = S m + h
    PencilSplit<S>::template fft<T, NZ, 1>(v, L.h);                      // -> z = ob M + k
}

}  // namespace

// ----------------------------------------------------------------- tc3a3b3b4b6 real-space step
// Per target (k,j) and direction d = blockIdx.z: qsum[j][d] += sum_p P[p][d] phi_qi (the
// s2 = 0 channel, before the density), and P[p][3s+d] <- FFT_z(density_s * P[p][3s+d]) * sc
// for s = 0, 1 (in place): the real-space step between the z pass of FFT_backward and the
// z pass of the next FFT_forward.  The lanes of a pencil move from the blocked output layout
// of one split transform to the strided input layout of the next through LDS.
template <typename T, int NZ, int S>
__global__ __launch_bounds__(kPencilBlock) void k_za7_mid(cpx<T>* __restrict__ P, const cpx<T>* __restrict__ occR,
                                                          const PairEntry* __restrict__ pairs,
                                                          const TargetEntry* __restrict__ targets,
                                                          const double2* __restrict__ dens, size_t dstride,
                                                          double2* __restrict__ qsum, int nxy, T sc)
{
    constexpr int M = NZ / S;
    constexpr int PN = PencilSplit<S>::pencils;
    __shared__ cpx<T> xs[PN][NZ + 1];  // one pencil per row (+1 against bank conflicts)
    const LaneOf<S> L(nxy);
    const int pi = threadIdx.x % PN;
    const TargetEntry te = targets[blockIdx.y];
    const int d = blockIdx.z;
    const size_t ng = (size_t)nxy * NZ;
    const size_t o0 = L.xy + (size_t)L.ob * M * nxy;  // first output point of the lane
    double2 acc[M];
    static_for<0, M>([&](auto kc) { acc[decltype(kc)::value] = make_double2(0.0, 0.0); });
    // steps it = 2 (pair) + s over the target's pairs; the pencil of step it + 1 is
    // loaded before step it is transformed (steps touch distinct grids)
    const int nsteps = 2 * te.npair;
    auto grid_of = [&](int it) { return P + ((size_t)(te.pair0 + (it >> 1)) * 6 + 3 * (it & 1) + d) * ng; };
    constexpr bool kPrefetch = sizeof(T) == 4;  // FP64 instances have no VGPRs to spare
    cpx<T> vn[M];
    if (kPrefetch && nsteps > 0) load_pencil<T, M>(vn, grid_of(0) + L.xy + (size_t)L.h * nxy, S * nxy);
#pragma unroll 1
    for (int it = 0; it < nsteps; ++it) {
        const int pp = te.pair0 + (it >> 1), s = it & 1;
        const cpx<T>* ph = occR + (size_t)pairs[pp].occ * ng + o0;
        cpx<T>* g = grid_of(it);
        cpx<T> v[M];
        if (kPrefetch) {
            static_for<0, M>([&](auto mc) { v[decltype(mc)::value] = vn[decltype(mc)::value]; });
            if (it + 1 < nsteps) load_pencil<T, M>(vn, grid_of(it + 1) + L.xy + (size_t)L.h * nxy, S * nxy);
        } else {
            load_pencil<T, M>(v, g + L.xy + (size_t)L.h * nxy, S * nxy);  // z = S m + h
        }
        const double2* n = dens + s * dstride + o0;  // dstride 0: one independent spin
        // FP32: the orbital and density points of the step are loaded before the transform,
        // so that their latency overlaps it (FP64 loads them where they are used)
        double2 dnv[kPrefetch ? M : 1];
        cpx<T> phv[kPrefetch ? M : 1];
        if constexpr (kPrefetch)
            static_for<0, M>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                dnv[k] = n[(size_t)k * nxy];
                if (s == 0) phv[k] = ph[(size_t)k * nxy];
            });
        // the exchanges of both transforms use lane swaps here (3.5% faster than ds_bpermute)
        PencilSplit<S>::template fft<T, NZ, 1, true>(v, L.h);  // -> z = ob M + k
        {
            static_for<0, M>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                const double2 x = upd(v[k]);
                double2 dn;
                if constexpr (kPrefetch) {
                    if (s == 0) acc[k] = cmadd(acc[k], x, upd(phv[k]), 1.0);
                    dn = dnv[k];
                } else {
                    if (s == 0) acc[k] = cmadd(acc[k], x, upd(ph[(size_t)k * nxy]), 1.0);
                    dn = n[(size_t)k * nxy];
                }
                xs[pi][L.ob * M + k] = cpx<T>{(T)(x.x * dn.x - x.y * dn.y), (T)(x.x * dn.y + x.y * dn.x)};
            });
            __syncthreads();
            static_for<0, M>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                v[m] = xs[pi][S * m + L.h];  // z = S m + h
            });
            __syncthreads();
            PencilSplit<S>::template fft<T, NZ, -1, true>(v, L.h);  // -> z = ob M + k
            if (L.valid)
                static_for<0, M>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    g[o0 + (size_t)k * nxy] = scale(v[k], sc);
                });
        }
    }
    if (!L.valid) return;
    double2* q = qsum + ((size_t)te.app * 3 + d) * ng + o0;
    static_for<0, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const double2 a = q[(size_t)k * nxy];
        q[(size_t)k * nxy] = make_double2(a.x + acc[k].x, a.y + acc[k].y);
    });
}

// ----------------------------------------------------------------- tc3a2a4b2b5 q-sums
// Per target (q2,i2) over pairs (q1,i1), O[p][6h + 3s + d] (h = 1 only for BITC):
//   Q2 -= f1 sum_{s,d} O_0,s,d conj(O_h2,s,d)     (h2 = 1 for BITC, 0 for TC)
template <typename T, int NZ, int S>
__global__ __launch_bounds__(kPencilBlock) void k_za8_q2(const cpx<T>* __restrict__ O, int bitc,
                                                         const PairEntry* __restrict__ pairs,
                                                         const TargetEntry* __restrict__ targets,
                                                         const double* __restrict__ fill, double2* __restrict__ Q2,
                                                         int nxy)
{
    constexpr int M = NZ / S;
    const LaneOf<S> L(nxy);
    const TargetEntry te = targets[blockIdx.y];
    const size_t ng = (size_t)nxy * NZ;
    const int nch = bitc ? 12 : 6;
    double2 acc[M];
    static_for<0, M>([&](auto kc) { acc[decltype(kc)::value] = make_double2(0.0, 0.0); });
    // steps it = 6 (pair) + c; the pencil(s) of step it + 1 are loaded before step it is
    // transformed
    const int nsteps = 6 * te.npair;
    auto grid_of = [&](int it, int half) {
        return O + ((size_t)(te.pair0 + it / 6) * nch + 6 * half + it % 6) * ng + L.xy + (size_t)L.h * nxy;
    };
    constexpr bool kPrefetch = sizeof(T) == 4;  // FP64 instances have no VGPRs to spare
    cpx<T> vn[M], wn[M];
    if (kPrefetch && nsteps > 0) {
        load_pencil<T, M>(vn, grid_of(0, 0), S * nxy);  // z = S m + h
        if (bitc) load_pencil<T, M>(wn, grid_of(0, 1), S * nxy);
    }
#pragma unroll 1
    for (int it = 0; it < nsteps; ++it) {
        const double f1 = fill[pairs[te.pair0 + it / 6].occ];
        {
            cpx<T> v[M];
            if (kPrefetch) {
                static_for<0, M>([&](auto mc) { v[decltype(mc)::value] = vn[decltype(mc)::value]; });
                if (it + 1 < nsteps) load_pencil<T, M>(vn, grid_of(it + 1, 0), S * nxy);
            } else {
                load_pencil<T, M>(v, grid_of(it, 0), S * nxy);
            }
            PencilSplit<S>::template fft<T, NZ, 1>(v, L.h);  // -> z = ob M + k
            if (bitc) {
                cpx<T> w[M];
                if (kPrefetch) {
                    static_for<0, M>([&](auto mc) { w[decltype(mc)::value] = wn[decltype(mc)::value]; });
                    if (it + 1 < nsteps) load_pencil<T, M>(wn, grid_of(it + 1, 1), S * nxy);
                } else {
                    load_pencil<T, M>(w, grid_of(it, 1), S * nxy);
                }
                PencilSplit<S>::template fft<T, NZ, 1>(w, L.h);
                static_for<0, M>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    const double2 a = upd(v[k]), b = upd(w[k]);
                    acc[k] = cmadd(acc[k], a, make_double2(b.x, -b.y), -f1);
                });
            } else {
                static_for<0, M>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    const double2 a = upd(v[k]);
                    acc[k].x -= f1 * (a.x * a.x + a.y * a.y);
                });
            }
        }
    }
    if (!L.valid) return;
    double2* q = Q2 + (size_t)te.app * ng + L.xy + (size_t)L.ob * M * nxy;
    static_for<0, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const double2 a = q[(size_t)k * nxy];
        q[(size_t)k * nxy] = make_double2(a.x + acc[k].x, a.y + acc[k].y);
    });
}

// Q4[d] += f1 O_0,0,d phi_q1     (conj5 = 0, channel d, X = occupied right orbitals)
// Q5[d] -= f1 conj(O_1,0,d chi_q1)  (conj5 = 1, channel 6 + d, X = occupied left orbitals; BITC)
// with d = blockIdx.z.
template <typename T, int NZ, int S>
__global__ __launch_bounds__(kPencilBlock) void k_za8_q45(const cpx<T>* __restrict__ O, int nch, int ch0, int conj5,
                                                          const cpx<T>* __restrict__ X,
                                                          const PairEntry* __restrict__ pairs,
                                                          const TargetEntry* __restrict__ targets,
                                                          const double* __restrict__ fill, double2* __restrict__ Q,
                                                          int nxy)
{
    constexpr int M = NZ / S;
    const LaneOf<S> L(nxy);
    const TargetEntry te = targets[blockIdx.y];
    const int d = blockIdx.z;
    const size_t ng = (size_t)nxy * NZ;
    const size_t o0 = L.xy + (size_t)L.ob * M * nxy;
    const double sy = conj5 ? -1.0 : 1.0;  // conj flips the imaginary part; Q5 also flips the sign
    double2 acc[M];
    static_for<0, M>([&](auto kc) { acc[decltype(kc)::value] = make_double2(0.0, 0.0); });
    // the pencil of pair pp + 1 is loaded before pair pp is transformed (FP32)
    constexpr bool kPrefetch = sizeof(T) == 4;
    auto grid_of = [&](int pp) { return O + ((size_t)pp * nch + ch0 + d) * ng + L.xy + (size_t)L.h * nxy; };
    cpx<T> vn[M];
    if (kPrefetch && te.npair > 0) load_pencil<T, M>(vn, grid_of(te.pair0), S * nxy);
#pragma unroll 1
    for (int pp = te.pair0; pp < te.pair0 + te.npair; ++pp) {
        const PairEntry e = pairs[pp];
        const double f1 = fill[e.occ];
        const cpx<T>* xg = X + (size_t)e.occ * ng + o0;
        cpx<T> v[M];
        if (kPrefetch) {
            static_for<0, M>([&](auto mc) { v[decltype(mc)::value] = vn[decltype(mc)::value]; });
            if (pp + 1 < te.pair0 + te.npair) load_pencil<T, M>(vn, grid_of(pp + 1), S * nxy);
        } else {
            load_pencil<T, M>(v, grid_of(pp), S * nxy);  // z = S m + h
        }
        // FP32 up to 40^3: the orbital points are loaded before the transform, so that their
        // latency overlaps it (at 48^3 the registers would halve the waves per SIMD; FP64
        // loads them where they are used)
        constexpr bool kHoist = kPrefetch && NZ <= 40;
        cpx<T> xv[kHoist ? M : 1];
        if constexpr (kHoist)
            static_for<0, M>([&](auto kc) { xv[decltype(kc)::value] = xg[(size_t)decltype(kc)::value * nxy]; });
        PencilSplit<S>::template fft<T, NZ, 1>(v, L.h);  // -> z = ob M + k
        static_for<0, M>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const double2 a = upd(v[k]);
            double2 b;
            if constexpr (kHoist)
                b = upd(xv[k]);
            else
                b = upd(xg[(size_t)k * nxy]);
            acc[k].x += (sy * f1) * (a.x * b.x - a.y * b.y);
            acc[k].y += f1 * (a.x * b.y + a.y * b.x);
        });
    }
    if (!L.valid) return;
    double2* q = Q + ((size_t)te.app * 3 + d) * ng + o0;
    static_for<0, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const double2 a = q[(size_t)k * nxy];
        q[(size_t)k * nxy] = make_double2(a.x + acc[k].x, a.y + acc[k].y);
    });
}

// Apply phase: W[t.app] += sum_p Z[p][0] phi_occ(p) + sum_d Z[p][1+d] Q4T[occ(p)][d]
template <typename T, int NZ, int S>
__global__ __launch_bounds__(kPencilBlock) void k_za8_apply(const cpx<T>* __restrict__ Z,
                                                            const cpx<T>* __restrict__ occR,
                                                            const cpx<T>* __restrict__ Q4T,
                                                            const PairEntry* __restrict__ pairs,
                                                            const TargetEntry* __restrict__ targets,
                                                            double2* __restrict__ W, float2* __restrict__ Wc,
                                                            int nxy)
{
    constexpr int M = NZ / S;
    const LaneOf<S> L(nxy);
    const TargetEntry te = targets[blockIdx.y];
    const size_t ng = (size_t)nxy * NZ;
    const size_t o0 = L.xy + (size_t)L.ob * M * nxy;
    double2 acc[M];
    static_for<0, M>([&](auto kc) { acc[decltype(kc)::value] = make_double2(0.0, 0.0); });
    // steps it = 4 (pair) + c; the pencil of step it + 1 is loaded before step it is
    // transformed (FP32)
    constexpr bool kPrefetch = sizeof(T) == 4;
    const int nsteps = 4 * te.npair;
    auto grid_of = [&](int it) { return Z + ((size_t)te.pair0 * 4 + it) * ng + L.xy + (size_t)L.h * nxy; };
    cpx<T> vn[M];
    if (kPrefetch && nsteps > 0) load_pencil<T, M>(vn, grid_of(0), S * nxy);
#pragma unroll 1
    for (int it = 0; it < nsteps; ++it) {
        const int pp = te.pair0 + it / 4, c = it % 4;
        const int o = pairs[pp].occ;
        {
            cpx<T> v[M];
            if (kPrefetch) {
                static_for<0, M>([&](auto mc) { v[decltype(mc)::value] = vn[decltype(mc)::value]; });
                if (it + 1 < nsteps) load_pencil<T, M>(vn, grid_of(it + 1), S * nxy);
            } else {
                load_pencil<T, M>(v, grid_of(it), S * nxy);  // z = S m + h
            }
            PencilSplit<S>::template fft<T, NZ, 1>(v, L.h);  // -> z = ob M + k
            const cpx<T>* m = (c == 0 ? occR + (size_t)o * ng : Q4T + ((size_t)o * 3 + (c - 1)) * ng) + o0;
            static_for<0, M>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                acc[k] = cmadd(acc[k], upd(v[k]), upd(m[(size_t)k * nxy]), 1.0);
            });
        }
    }
    if (!L.valid) return;
    const size_t o = (size_t)te.app * ng + o0;
    static_for<0, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        add_partial(W + o + (size_t)k * nxy, Wc ? Wc + o + (size_t)k * nxy : nullptr, acc[k].x, acc[k].y);
    });
}

// ----------------------------------------------------------------- launchers
namespace {

template <typename F>
void for_target_slices(int n, F&& f)
{
    for (int a = 0; a < n; a += 65535) f(a, n - a < 65535 ? n - a : 65535);
}

}  // namespace

#define TCX_Z_DISPATCH(prec, g, LAUNCH)                        \
    do {                                                       \
        if ((prec) == TCX_PREC_FP64) {                         \
            using T = double;                                  \
            TCX_SIZE_SWITCH((g).n3, LAUNCH)                    \
        } else {                                               \
            using T = float;                                   \
            TCX_SIZE_SWITCH((g).n3, LAUNCH)                    \
        }                                                      \
        TCX_HIP_CHECK(hipGetLastError());                      \
    } while (0)

void launch_za7_mid(int prec, const Geom& g, void* P, const void* occR, const PairEntry* pairs,
                    const TargetEntry* targets, int ntargets, const double* dens, long long dens_stride,
                    double* qsum, double sc, hipStream_t s)
{
    const int nxy = g.n1 * g.n2;
    for_target_slices(ntargets, [&](int a, int n) {
#define TCX_L(NZ)                                                                                            \
    {                                                                                                        \
        constexpr int S = SplitFor<NZ, 4>::value;                                                            \
        constexpr int PN = PencilSplit<S>::pencils;                                                          \
        hipLaunchKernelGGL((k_za7_mid<T, NZ, S>), dim3((nxy + PN - 1) / PN, n, 3), dim3(kPencilBlock), 0, s,\
                           (cpx<T>*)P, (const cpx<T>*)occR, pairs, targets + a, (const double2*)dens,        \
                           (size_t)dens_stride, (double2*)qsum, nxy, (T)sc);                                 \
    }
        TCX_Z_DISPATCH(prec, g, TCX_L);
#undef TCX_L
    });
}

void launch_za8_q(int prec, const Geom& g, const void* O, int bitc, const void* occR, const void* chiR,
                  const PairEntry* pairs, const TargetEntry* targets, int ntargets, const double* fill, double* Q2,
                  double* Q4, double* Q5, hipStream_t s)
{
    const int nxy = g.n1 * g.n2;
    const int nch = bitc ? 12 : 6;
    for_target_slices(ntargets, [&](int a, int n) {
#define TCX_L(NZ)                                                                                              \
    {                                                                                                          \
        /* 4 lanes per pencil for Q2 (2 would need 256 VGPRs); Q4/Q5 take 2 in FP32 (longer load runs) */  \
        constexpr int SQ2 = SplitFor<NZ, 4>::value;                                                            \
        constexpr int SQ45 = SplitFor<NZ, sizeof(T) == 8 ? 4 : 2>::value;                                     \
        constexpr int P2 = PencilSplit<SQ2>::pencils, P4 = PencilSplit<SQ45>::pencils;                         \
        hipLaunchKernelGGL((k_za8_q2<T, NZ, SQ2>), dim3((nxy + P2 - 1) / P2, n), dim3(kPencilBlock), 0, s,    \
                           (const cpx<T>*)O, bitc, pairs, targets + a, fill, (double2*)Q2, nxy);               \
        hipLaunchKernelGGL((k_za8_q45<T, NZ, SQ45>), dim3((nxy + P4 - 1) / P4, n, 3), dim3(kPencilBlock), 0, s, \
                           (const cpx<T>*)O, nch, 0, 0, (const cpx<T>*)occR, pairs, targets + a, fill,        \
                           (double2*)Q4, nxy);                                                                 \
        if (bitc)                                                                                              \
            hipLaunchKernelGGL((k_za8_q45<T, NZ, SQ45>), dim3((nxy + P4 - 1) / P4, n, 3), dim3(kPencilBlock), 0, \
                               s,                                                                              \
                               (const cpx<T>*)O, nch, 6, 1, (const cpx<T>*)chiR, pairs, targets + a, fill,    \
                               (double2*)Q5, nxy);                                                             \
    }
        TCX_Z_DISPATCH(prec, g, TCX_L);
#undef TCX_L
    });
}

void launch_za8_apply(int prec, const Geom& g, const void* Z, const void* occR, const void* Q4T,
                      const PairEntry* pairs, const TargetEntry* targets, int ntargets, double* W, float* Wc,
                      hipStream_t s)
{
    const int nxy = g.n1 * g.n2;
    for_target_slices(ntargets, [&](int a, int n) {
#define TCX_L(NZ)                                                                                              \
    {                                                                                                          \
        constexpr int S = SplitFor<NZ, sizeof(T) == 8 ? 4 : 2>::value;                                         \
        constexpr int PN = PencilSplit<S>::pencils;                                                            \
        hipLaunchKernelGGL((k_za8_apply<T, NZ, S>), dim3((nxy + PN - 1) / PN, n), dim3(kPencilBlock), 0, s,   \
                           (const cpx<T>*)Z, (const cpx<T>*)occR, (const cpx<T>*)Q4T, pairs, targets + a,      \
                           (double2*)W, (float2*)Wc, nxy);                                                     \
    }
        TCX_Z_DISPATCH(prec, g, TCX_L);
#undef TCX_L
    });
}

#undef TCX_Z_DISPATCH

}  // namespace tcx
