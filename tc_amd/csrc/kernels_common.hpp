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

__device__ __forceinline__ int centered(int i, int n) { return i <= n / 2 ? i : i - n; }

template <typename T, int NZ>
__device__ __forceinline__ void load_pencil(cpx<T> (&v)[NZ], const cpx<T>* __restrict__ p, int stride)
{
    static_for<0, NZ>([&](auto zc) {
        constexpr int z = decltype(zc)::value;
        v[z] = p[(size_t)z * stride];
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
