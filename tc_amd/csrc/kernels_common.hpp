// kernels_common.hpp -- device building blocks shared by the gfx950 kernels.
//
// Data layout in HBM: every real-space or G-space grid is n1*n2*n3 complex
// values with i1 fastest (the reference's ipw = i1 + n1*(i2 + n2*i3)); batches
// of grids are consecutive.  Two access patterns cover a 3D transform:
//   * z-pencil: one thread owns column (i1,i2) and walks i3 with stride n1*n2;
//     a wavefront's 64 lanes are 64 consecutive (i1,i2), so every load and
//     store is a fully coalesced 512 B (float2) or 1 KB (double2) transaction;
//   * xy-plane: one workgroup owns plane i3; thread x loads column x
//     (coalesced across x), transforms along y in registers, and a padded LDS
//     tile transposes it so that thread y transforms row y along x.
#pragma once

#include "fft_reg.hpp"
#include "tcx_internal.hpp"

namespace tcx {

constexpr int kPencilBlock = 64;  // threads per z-pencil workgroup
constexpr int kPlaneBlock = 64;   // threads per xy-plane workgroup (plane side <= 64)

// Launch a kernel template instantiated with kPoly = (polynomial Jastrow present):
// kernels without the polynomial part carry no call and no stack.
#define TCX_POLY_DISPATCH(poly, ...)         \
    do {                                     \
        if (poly) {                          \
            constexpr bool kPoly = true;     \
            __VA_ARGS__;                     \
        } else {                             \
            constexpr bool kPoly = false;    \
            __VA_ARGS__;                     \
        }                                    \
    } while (0)

__device__ __forceinline__ int centered(int i, int n) { return i <= n / 2 ? i : i - n; }

// w += a; returns the rounding error of that addition (Knuth's TwoSum: exact in binary
// floating point, so the code must not be reassociated; it has no products to contract).
__device__ __forceinline__ double two_sum_add(double& w, double a)
{
    const double s = w + a;
    const double ap = s - w;
    const double e = (w - (s - ap)) + (a - ap);
    w = s;
    return e;
}

// Adds the partial band-pair sum (are, aim) of a launch into the FP64 accumulator point w.
// With a compensation grid (wc non-null) the rounding errors of these additions are summed
// into wc, so that w + wc carries the sum with about one rounding of the result instead of
// one per launch.  The errors are below 1.2e-16 |w| each, so their sum is kept in FP32: its
// own rounding is 1e-7 of a correction of order 1e-16 |w|, at half the traffic of an FP64 grid.
__device__ __forceinline__ void add_partial(double2* w, float2* wc, double are, double aim)
{
    double2 a = *w;
    if (wc) {
        const double ex = two_sum_add(a.x, are), ey = two_sum_add(a.y, aim);
        const float2 c = *wc;
        *wc = make_float2((float)((double)c.x + ex), (float)((double)c.y + ey));
    } else {
        a.x += are;
        a.y += aim;
    }
    *w = a;
}


template <typename T, int NZ>
__device__ __forceinline__ void load_pencil(cpx<T> (&v)[NZ], const cpx<T>* __restrict__ p, int stride)
{
    static_for<0, NZ>([&](auto zc) {
        constexpr int z = decltype(zc)::value;
        v[z] = p[(size_t)z * stride];
    });
}

// Streaming (non-temporal) variants for grids that a pass touches once, so that
// they do not displace the orbital grids that many workgroups re-read from cache.
template <typename T, int NZ>
__device__ __forceinline__ void store_pencil_nt(const cpx<T> (&v)[NZ], cpx<T>* __restrict__ p, int stride)
{
    typedef T v2 __attribute__((ext_vector_type(2)));
    static_for<0, NZ>([&](auto zc) {
        constexpr int z = decltype(zc)::value;
        const v2 x = {v[z].re, v[z].im};
        __builtin_nontemporal_store(x, reinterpret_cast<v2*>(p + (size_t)z * stride));
    });
}

template <typename T, int NZ>
__device__ __forceinline__ void load_pencil_nt(cpx<T> (&v)[NZ], const cpx<T>* __restrict__ p, int stride)
{
    typedef T v2 __attribute__((ext_vector_type(2)));
    static_for<0, NZ>([&](auto zc) {
        constexpr int z = decltype(zc)::value;
        const v2 x = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p + (size_t)z * stride));
        v[z] = cpx<T>{x.x, x.y};
    });
}

template <typename T, int NZ>
__device__ __forceinline__ void store_pencil(const cpx<T> (&v)[NZ], cpx<T>* __restrict__ p, int stride)
{
    static_for<0, NZ>([&](auto zc) {
        constexpr int z = decltype(zc)::value;
        p[(size_t)z * stride] = v[z];
    });
}

// Component-wise select: a ?: on the 16-byte cpx<double> struct is lowered
// through private memory, two scalar selects stay in registers.
template <typename T>
__device__ __forceinline__ cpx<T> csel(bool c, cpx<T> a, cpx<T> b)
{
    return cpx<T>{c ? a.re : b.re, c ? a.im : b.im};
}

// Lane exchange of one value with lane ^ mask (ds_bpermute), doubles as two
// 32-bit halves kept in registers.
__device__ __forceinline__ float xshfl(float v, int mask) { return __shfl_xor(v, mask); }
__device__ __forceinline__ double xshfl(double v, int mask)
{
    const int lo = __shfl_xor(__double2loint(v), mask);
    const int hi = __shfl_xor(__double2hiint(v), mask);
    return __hiloint2double(hi, lo);
}

// Radix-2 butterfly between lane and lane ^ MASK (MASK = 32 or 16) through gfx950's
// lane-swap instructions, which exchange values in the VALU rather than through LDS.
// With both operands holding x, v_permlane32_swap returns (own x, partner's x) in
// lanes 0-31 and (partner's x, own x) in lanes 32-63; v_permlane16_swap does the same
// for the even and odd 16-lane rows.  So a lane whose MASK bit is clear gets
// r0 + r1 = own + partner, and a lane whose MASK bit is set gets r0 - r1 =
// partner - own, with no select of the partner's value.  `upper` must be
// (lane & MASK) != 0.  A double is exchanged as its two 32-bit halves.
// The swap is in-order VALU work, where ds_bpermute (xshfl) can have several
// exchanges in flight; which one is faster depends on the kernel around it.
template <int MASK>
__device__ __forceinline__ void lane_swap(unsigned x, unsigned& r0, unsigned& r1)
{
    static_assert(MASK == 32 || MASK == 16, "lane swaps exchange lane ^ 32 or lane ^ 16");
    if constexpr (MASK == 32) {
        const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        r0 = r[0];
        r1 = r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
        r0 = r[0];
        r1 = r[1];
    }
}
template <int MASK>
__device__ __forceinline__ float xbfly(float x, bool upper)
{
    unsigned a, b;
    lane_swap<MASK>(__float_as_uint(x), a, b);
    const float r0 = __uint_as_float(a), r1 = __uint_as_float(b);
    return upper ? r0 - r1 : r0 + r1;
}
template <int MASK>
__device__ __forceinline__ double xbfly(double x, bool upper)
{
    unsigned lo0, lo1, hi0, hi1;
    lane_swap<MASK>((unsigned)__double2loint(x), lo0, lo1);
    lane_swap<MASK>((unsigned)__double2hiint(x), hi0, hi1);
    const double r0 = __hiloint2double((int)hi0, (int)lo0), r1 = __hiloint2double((int)hi1, (int)lo1);
    return upper ? r0 - r1 : r0 + r1;
}
template <int MASK, typename T>
__device__ __forceinline__ cpx<T> xbfly(cpx<T> x, bool upper)
{
    return cpx<T>{xbfly<MASK>(x.re, upper), xbfly<MASK>(x.im, upper)};
}

// ---------------------------------------------------------------------------
// Pencil split over two lanes (lane and lane ^ 32 of a wavefront).
// Lane half h holds the samples z = 2m + h (m < NZ/2) on entry: a radix-2
// decimation in time whose length-NZ/2 sub-transforms run in each lane's
// registers and whose final butterfly crosses the two halves with one
// __shfl_xor(.., 32) per value.  On exit lane half h holds outputs
// z = h*NZ/2 + k (k < NZ/2).  Halves the registers a pencil needs, so the
// FP64 accumulators of the accumulate pass fit at twice the occupancy.
// ---------------------------------------------------------------------------
template <typename T, int NZ, int SIGN, bool kLaneSwap = false>
__device__ __forceinline__ void fft_split2(cpx<T> (&v)[NZ / 2], int h)
{
    constexpr int M = NZ / 2;
    fft<T, M, SIGN>(v);
    static_for<1, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const cpx<T> t = v[k] * twiddle<T, NZ, SIGN, k>();
        v[k] = csel(h != 0, t, v[k]);
    });
    static_for<0, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (kLaneSwap) {
            v[k] = xbfly<32>(v[k], h != 0);  // h = lane >> 5
        } else {
            const cpx<T> o{xshfl(v[k].re, 32), xshfl(v[k].im, 32)};
            v[k] = csel(h != 0, o - v[k], v[k] + o);
        }
    });
}

// The transpose of fft_split2 (decimation in frequency): lane half h holds the samples
// z = h*NZ/2 + k on entry (the layout fft_split2 leaves); the cross-lane butterfly comes
// first, then the twiddle of the upper half and the length-NZ/2 transform in registers.
// On exit lane half h holds outputs z = 2m + h.
template <typename T, int NZ, int SIGN, bool kLaneSwap = false>
__device__ __forceinline__ void fft_split2_dif(cpx<T> (&v)[NZ / 2], int h)
{
    constexpr int M = NZ / 2;
    static_for<0, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (kLaneSwap) {
            v[k] = xbfly<32>(v[k], h != 0);  // h = lane >> 5: lower X[k] + X[k+M], upper X[k] - X[k+M]
        } else {
            const cpx<T> o{xshfl(v[k].re, 32), xshfl(v[k].im, 32)};
            v[k] = csel(h != 0, o - v[k], v[k] + o);
        }
    });
    static_for<1, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const cpx<T> t = v[k] * twiddle<T, NZ, SIGN, k>();
        v[k] = csel(h != 0, t, v[k]);
    });
    fft<T, M, SIGN>(v);
}

// Pencil split over four lanes (lane ^ 16, ^ 32, ^ 48): lane quarter h holds
// z = 4m + h on entry; the radix-4 cross-lane DFT runs as two radix-2 stages
// (shuffle with lane ^ 32, twiddle, shuffle with lane ^ 16).  On exit lane
// quarter h holds outputs z = ob*NZ/4 + k with ob = (h >> 1) + 2*(h & 1).
template <typename T, int NZ, int SIGN, bool kLaneSwap = false>
__device__ __forceinline__ void fft_split4(cpx<T> (&v)[NZ / 4], int h)
{
    constexpr int M = NZ / 4;
    fft<T, M, SIGN>(v);
    // Y'_h[k] = W_NZ^{h k} Y_h[k] = W^{h0 k} W^{2 h1 k} Y_h[k]  (two predicated
    // multiplies by constants; a select among constants would become a table)
    const int h0 = h & 1, h1 = h >> 1;
    static_for<1, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const cpx<T> t1 = v[k] * twiddle<T, NZ, SIGN, k>();
        const cpx<T> a = csel(h0 != 0, t1, v[k]);
        const cpx<T> t2 = a * twiddle<T, NZ, SIGN, (2 * k) % NZ>();
        v[k] = csel(h1 != 0, t2, a);
    });
    static_for<0, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (kLaneSwap) {
            cpx<T> z = xbfly<32>(v[k], h1 != 0);                        // h1 = (lane >> 5) & 1
            z = csel((h0 & h1) != 0, mul_isign<T, SIGN>(z, T(1)), z);
            v[k] = xbfly<16>(z, h0 != 0);                               // h0 = (lane >> 4) & 1
        } else {
            const cpx<T> o{xshfl(v[k].re, 32), xshfl(v[k].im, 32)};
            cpx<T> z = csel(h1 != 0, o - v[k], v[k] + o);               // Z_{h0, s0=h1}
            z = csel((h0 & h1) != 0, mul_isign<T, SIGN>(z, T(1)), z);     // W_4^{h0 s0}
            const cpx<T> o2{xshfl(z.re, 16), xshfl(z.im, 16)};
            v[k] = csel(h0 != 0, o2 - z, z + o2);                       // X_{s0 + 2 s1}, s1 = h0
        }
    });
}

// Exchange choice of the hf2x and tc2x pair passes: lane swaps in FP64 (tc2x passes A
// and C then fit one more wave per SIMD; hf2x and tc2x FP64 applies ran 1% faster),
// ds_bpermute in FP32 (no gain there).
template <typename T>
constexpr bool kSwapF64 = sizeof(T) == 8;

// Lanes per pencil and the output block of a lane for the split transforms.
template <int S>
struct PencilSplit;
template <>
struct PencilSplit<2> {
    static constexpr int pencils = 32;
    __device__ static int out_block(int h) { return h; }
    template <typename T, int NZ, int SIGN, bool kLaneSwap = false>
    __device__ static void fft(cpx<T> (&v)[NZ / 2], int h) { fft_split2<T, NZ, SIGN, kLaneSwap>(v, h); }
};
template <>
struct PencilSplit<4> {
    static constexpr int pencils = 16;
    __device__ static int out_block(int h) { return (h >> 1) + 2 * (h & 1); }
    template <typename T, int NZ, int SIGN, bool kLaneSwap = false>
    __device__ static void fft(cpx<T> (&v)[NZ / 4], int h) { fft_split4<T, NZ, SIGN, kLaneSwap>(v, h); }
};

// FP32 pair records split a pencil over 2 lanes, FP64 ones over 4 (registers
// of the FP64 accumulators plus the pencil stay under 256 VGPRs).
template <typename T, int NZ>
struct SplitOf {
    static constexpr int value = (sizeof(T) == 8 && NZ % 4 == 0) ? 4 : 2;
};

// Planes per workgroup for the xy-plane passes: one thread per line of a
// plane, so a workgroup of PL planes has PL*NP busy lanes; choose PL for the
// best lane utilisation with the padded LDS tiles under kPlaneLdsCap bytes.
constexpr int kPlaneLdsCap = 80 * 1024;
template <typename T, int NP>
struct PlaneGroup {
    static constexpr int tile_bytes = NP * (NP + 1) * (int)sizeof(cpx<T>);
    static constexpr int pick()
    {
        int best = 1;
        double best_util = 0.0;
        for (int p = 1; p <= 64; ++p) {
            if (p * tile_bytes > kPlaneLdsCap || p * NP > 1024) break;
            const int lanes = ((p * NP + 63) / 64) * 64;
            const double util = double(p * NP) / lanes;
            if (util > best_util + 1e-9) {
                best_util = util;
                best = p;
            }
        }
        return best;
    }
    static constexpr int planes = pick();
    static constexpr int threads = ((planes * NP + 63) / 64) * 64;
};

// LDS tile of one plane: [y][x] with one element of padding per row.
template <typename T, int NP>
struct PlaneTile {
    cpx<T> a[NP][NP + 1];
};

// thread t (< NP) holds column x=t (values along y) -> writes it to the tile
template <typename T, int NP>
__device__ __forceinline__ void tile_put_col(PlaneTile<T, NP>& s, const cpx<T> (&v)[NP], int t)
{
    static_for<0, NP>([&](auto yc) {
        constexpr int y = decltype(yc)::value;
        s.a[y][t] = v[y];
    });
}
template <typename T, int NP>
__device__ __forceinline__ void tile_get_col(const PlaneTile<T, NP>& s, cpx<T> (&v)[NP], int t)
{
    static_for<0, NP>([&](auto yc) {
        constexpr int y = decltype(yc)::value;
        v[y] = s.a[y][t];
    });
}
template <typename T, int NP>
__device__ __forceinline__ void tile_put_row(PlaneTile<T, NP>& s, const cpx<T> (&v)[NP], int t)
{
    static_for<0, NP>([&](auto xc) {
        constexpr int x = decltype(xc)::value;
        s.a[t][x] = v[x];
    });
}
template <typename T, int NP>
__device__ __forceinline__ void tile_get_row(const PlaneTile<T, NP>& s, cpx<T> (&v)[NP], int t)
{
    static_for<0, NP>([&](auto xc) {
        constexpr int x = decltype(xc)::value;
        v[x] = s.a[t][x];
    });
}

}  // namespace tcx
