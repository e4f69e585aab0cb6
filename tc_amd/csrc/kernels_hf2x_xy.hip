// kernels_hf2x_xy.hip -- plane-first pair pipeline of the Fock exchange (hf2x).
//
// The same operator as kernels_hf2x.hip (src/calc_hamiltonian_hf2x.cpp:123-135):
//     rho(r) = conj(phi_qi(r)) phi_kj(r),  rho~ = FFT_forward(rho) v(G)/N,
//     W_kj(r) += f_qi phi_qi(r) FFT_backward(rho~)(r),
// with the two 3D transforms split the other way round:
//   A  (xy planes):  rho on plane z, FFT along y and x, store the plane's
//                    support columns;
//   B  (z pencils):  one per support column (kx, ky): FFT along z, x v(G)/N,
//                    inverse FFT along z, store in place;
//   C  (xy planes):  inverse FFT along x and y, x f phi_qi, accumulate over the
//                    pairs of the target in FP64 registers, add once into W_kj.
// rho has G-space support only inside |G + k - q| < 2 G_w (the sum of the two
// orbitals' cutoff spheres), and so has every quantity between the passes.  A z
// pencil after pass A covers every (kx, ky) column, but between xy and z only the
// columns in the support's projection on the (kx, ky) plane are non-zero: about
// two thirds of them at 40^3 (fcc, 4x4x4 k).  The pair record holds just those,
// [z][c] with c enumerating, row ky by row ky, a contiguous run of kx (the row's
// range [lo, lo + len) in centred kx, taken modulo n1), so every pass moves about a
// third fewer bytes than the z-first order, whose records must hold every column.
// The column maps are per k - q row of a launch:
//   colxy[coloff[r] + c] = i1 | i2 << 16 (grid indices of column c),
//   rowinfo[r * n2 + i2]  = (lo, len, first column) of row ky = i2.
#include "kernels_common.hpp"
#include "size_switch.hpp"

namespace tcx {

namespace {

// Pass C splits every line of its planes over S lanes (PencilSplit: lanes l, l ^ 16,
// l ^ 32, l ^ 48 for S = 4), which keeps its NP/S FP64 accumulators per lane within
// the registers of three waves per SIMD (two at NP = 30, whose lines split over 2 lanes
// only: under three it spilled 380 bytes per lane); a wavefront then holds 64/S lines.  Planes
// per workgroup: the best fill of the wavefronts with at most 56 KiB of LDS tiles (two
// per plane).
template <typename T, int NP>
struct PlaneGroupC {
    static constexpr int split = (NP % 4 == 0 && NP >= 16) ? 4 : 2;
    static constexpr int lines_per_wave = 64 / split;
    static constexpr int tile_bytes = 2 * NP * (NP + 1) * (int)sizeof(cpx<T>);  // two buffers
    static constexpr int pick()
    {
        int best = 1;
        double best_util = 0.0;
        for (int p = 1; p <= 16; ++p) {
            if (p * tile_bytes > 56 * 1024 || p * NP > 16 * lines_per_wave) break;
            const int waves = (p * NP + lines_per_wave - 1) / lines_per_wave;
            const double util = double(p * NP) / (waves * lines_per_wave);
            if (util > best_util + 1e-9) {
                best_util = util;
                best = p;
            }
        }
        return best;
    }
    static constexpr int planes = pick();
    static constexpr int waves = (planes * NP + lines_per_wave - 1) / lines_per_wave;
    static constexpr int threads = waves * 64;
};

__device__ __forceinline__ int posmod(int a, int n)
{
    const int r = a % n;
    return r < 0 ? r + n : r;
}

}  // namespace

// Pass A: one workgroup holds PL planes of one pair (PlaneGroup), thread t owns line
// t % NP of plane t / NP.  Column x = t of rho -> FFT_y -> LDS transpose -> FFT_x ->
// row ky = t back into the tile -> the planes' support columns leave coalesced.  The
// column map of the pair's k - q row is read before the transforms, so its loads are
// done by the time the stores need it.
template <typename T, int NP>
__global__ __launch_bounds__((PlaneGroup<T, NP>::threads)) void k_hf2x_xyA(
    const cpx<T>* __restrict__ phiR_app, const cpx<T>* __restrict__ phiR_occ, const PairEntry* __restrict__ pairs,
    const int* __restrict__ colxy, const int* __restrict__ coloff, cpx<T>* __restrict__ rec, int cstride, int nz)
{
    constexpr int PL = PlaneGroup<T, NP>::planes;
    constexpr int NT = PlaneGroup<T, NP>::threads;
    constexpr int CPT = (NP * NP + NT - 1) / NT;  // columns per thread, at most
    __shared__ PlaneTile<T, NP> tiles[PL];
    const int p = threadIdx.x / NP, t = threadIdx.x % NP;
    const int z0 = blockIdx.x * PL, z = z0 + p;
    const bool active = p < PL && z < nz;
    const PairEntry pe = pairs[blockIdx.y];
    const int c0 = coloff[pe.kq], nc = coloff[pe.kq + 1] - c0;
    int cxy[CPT];
    static_for<0, CPT>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const int c = (int)threadIdx.x + i * NT;
        cxy[i] = c < nc ? colxy[c0 + c] : 0;
    });
    const size_t ng = (size_t)NP * NP * nz;
    PlaneTile<T, NP>& tile = tiles[active ? p : 0];
    cpx<T> v[NP];
    if (active) {
        const cpx<T>* a = phiR_app + (size_t)pe.app * ng + (size_t)z * NP * NP + t;
        const cpx<T>* b = phiR_occ + (size_t)pe.occ * ng + (size_t)z * NP * NP + t;
        static_for<0, NP>([&](auto yc) {
            constexpr int y = decltype(yc)::value;
            v[y] = cmulc(b[y * NP], a[y * NP]);
        });
        fft<T, NP, -1>(v);
        tile_put_col<T, NP>(tile, v, t);
    }
    __syncthreads();
    if (active) {
        tile_get_row<T, NP>(tile, v, t);
        fft<T, NP, -1>(v);
        tile_put_row<T, NP>(tile, v, t);  // row ky = t, frequency kx along the row
    }
    __syncthreads();
    const int npl = min(PL, nz - z0);
    cpx<T>* out = rec + ((size_t)blockIdx.y * nz + z0) * cstride;
    typedef T v2 __attribute__((ext_vector_type(2)));
    static_for<0, CPT>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const int c = (int)threadIdx.x + i * NT;
        if (c < nc) {
            const int xy = cxy[i];
            static_for<0, PL>([&](auto plc) {
                constexpr int pl = decltype(plc)::value;
                if (pl < npl) {
                    const cpx<T> x = tiles[pl].a[xy >> 16][xy & 0xffff];
                    const v2 s2 = {x.re, x.im};
                    __builtin_nontemporal_store(s2, reinterpret_cast<v2*>(out + (size_t)pl * cstride + c));
                }
            });
        }
    });
}

// Pass B: two lanes per support column c of a pair (lanes l and l ^ 32), each holding
// half of the z pencil: FFT_z (fft_split2, samples z = 2m + h in, kz = h NZ/2 + k out)
// -> x v(G)/N -> inverse FFT_z (fft_split2_dif, back to z = 2m + h), in place.  The
// record is [z][c], so the lanes of a load are consecutive columns.  v(G)/N =
// 1/(N |k-q+G|^2) (0 below 1e-8, hf2x.cpp:91-104) is evaluated here in FP64 and rounded
// once to FP32: |k-q+G|^2 is a quadratic in kz along the column, its reciprocal one
// v_rcp_f64 with a Newton step.
template <typename T, int NZ>
__global__ __launch_bounds__(kPencilBlock) void k_hf2x_zB(cpx<T>* __restrict__ rec, const PairEntry* __restrict__ pairs,
                                                          const int* __restrict__ colxy, const int* __restrict__ coloff,
                                                          const double* __restrict__ kqcart, int cstride, Geom g,
                                                          double invN)
{
    constexpr int M = NZ / 2;
    const int h = threadIdx.x >> 5;
    const int c = blockIdx.x * 32 + (threadIdx.x & 31);
    const PairEntry pe = pairs[blockIdx.y];
    const int c0 = coloff[pe.kq], nc = coloff[pe.kq + 1] - c0;
    const bool valid = c < nc;  // both halves of a column agree; invalid lanes still shuffle
    const int cc = valid ? c : 0;
    const int xy = colxy[c0 + cc];
    const int m1 = centered(xy & 0xffff, g.n1), m2 = centered(xy >> 16, g.n2);
    // |k-q+G|^2 = a0 + a1 m3 + a2 m3^2 along the column
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double base = kqcart[3 * pe.kq + d] + g.b[d] * m1 + g.b[3 + d] * m2;
        a0 += base * base;
        a1 += 2.0 * base * g.b[6 + d];
        a2 += g.b[6 + d] * g.b[6 + d];
    }
    cpx<T>* col = rec + (size_t)blockIdx.y * NZ * cstride + cc + (size_t)h * cstride;
    cpx<T> v[M];
    if (valid) {
        load_pencil_nt<T, M>(v, col, 2 * cstride);  // z = 2m + h
    } else {
        static_for<0, M>([&](auto mc) { v[decltype(mc)::value] = cpx<T>{T(0), T(0)}; });
    }
    fft_split2<T, NZ, -1>(v, h);  // -> kz = h M + k
    static_for<0, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const int kz = h * M + k;
        const double m3 = kz <= NZ / 2 ? kz : kz - NZ;
        const double g2 = fma(fma(a2, m3, a1), m3, a0);
        double r = __builtin_amdgcn_rcp(g2);  // one Newton step: 1/g2 to ~1 ulp
        r = fma(fma(-g2, r, 1.0), r, r);
        v[k] = scale(v[k], static_cast<T>(g2 < 1e-8 ? 0.0 : invN * r));
    });
    fft_split2_dif<T, NZ, 1>(v, h);  // -> z = 2m + h
    if (valid) store_pencil_nt<T, M>(v, col, 2 * cstride);
}

// Pass C: one workgroup holds PL planes of one target (k,j) and loops over the pairs
// of that target.  Per pair: row ky of the record (S lanes, kx = S m + h; a row's
// support columns are contiguous in the record) -> inverse FFT_x -> LDS tile; column x
// (S lanes, ky = S m + h) -> inverse FFT_y -> x f phi_qi -> FP64 accumulators (NP/S per
// lane).  The tiles alternate between two buffers, so one barrier per pair suffices,
// and the next pair's row and phi_qi values are loaded before the current pair is
// transformed.  No atomics: each (target, plane) belongs to one workgroup.
template <typename T, int NP>
__global__ __launch_bounds__((PlaneGroupC<T, NP>::threads), (NP == 30 ? 2 : 3)) void k_hf2x_xyC(
    const cpx<T>* __restrict__ rec, const cpx<T>* __restrict__ phiR_occ, const PairEntry* __restrict__ pairs,
    const TargetEntry* __restrict__ targets, const int4* __restrict__ rowinfo, const double* __restrict__ fill,
    cpx<double>* __restrict__ W, float2* __restrict__ Wc, int cstride, int nz)
{
    using PG = PlaneGroupC<T, NP>;
    constexpr int PL = PG::planes, S = PG::split, LW = PG::lines_per_wave, M = NP / S;
    using Split = PencilSplit<S>;
    __shared__ PlaneTile<T, NP> tiles[2][PL];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int L = w * LW + lane % LW, h = lane / LW;  // line L of the workgroup, part h of it
    const int ob = Split::out_block(h);
    const int p = L / NP, line = L - p * NP;
    const int z = blockIdx.x * PL + p;
    const bool active = p < PL && z < nz;
    const TargetEntry te = targets[blockIdx.y];
    const size_t ng = (size_t)NP * NP * nz;
    // row `line` of plane z of pair pp (kx = S m + h), phi_qi at (x = line, y = ob M + k)
    // and its filling.  The pair entries run two pairs ahead and the row's (lo, len,
    // first column) is reloaded only when the k - q row changes, so the loads of a pair
    // wait on nothing loaded in the same iteration.
    int cur_kq = -1;
    int4 ri = make_int4(0, 0, 0, 0);
    auto load = [&](const PairEntry& pe, int pp, cpx<T> (&r)[M], cpx<T> (&q)[M], double& f) {
        if (pe.kq != cur_kq) {
            ri = rowinfo[(size_t)pe.kq * NP + line];
            cur_kq = pe.kq;
        }
        f = fill ? fill[pe.occ] : 1.0;
        const cpx<T>* row = rec + ((size_t)pp * nz + z) * cstride + ri.z;
        static_for<0, M>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const int d = posmod(S * m + h - ri.x, NP);
            r[m] = d < ri.y ? row[d] : cpx<T>{T(0), T(0)};
        });
        const cpx<T>* o = phiR_occ + (size_t)pe.occ * ng + (size_t)z * NP * NP + (size_t)ob * M * NP + line;
        static_for<0, M>([&](auto kc) { q[decltype(kc)::value] = o[(size_t)decltype(kc)::value * NP]; });
    };
    cpx<double> acc[M];
    static_for<0, M>([&](auto kc) { acc[decltype(kc)::value] = cpx<double>{0.0, 0.0}; });
    cpx<T> rn[M], qn[M];
    double fn = 0.0;
    static_for<0, M>([&](auto mc) {
        rn[decltype(mc)::value] = cpx<T>{T(0), T(0)};
        qn[decltype(mc)::value] = cpx<T>{T(0), T(0)};
    });
    const int pend = te.pair0 + te.npair;
    PairEntry pe_next = te.npair > 1 ? pairs[te.pair0 + 1] : PairEntry{0, 0, 0, 0};
    if (active && te.npair > 0) load(pairs[te.pair0], te.pair0, rn, qn, fn);
    int buf = 0;
#pragma unroll 1
    for (int pp = te.pair0; pp < pend; ++pp, buf ^= 1) {
        cpx<T> v[M], q[M];
        static_for<0, M>([&](auto mc) {
            v[decltype(mc)::value] = rn[decltype(mc)::value];
            q[decltype(mc)::value] = qn[decltype(mc)::value];
        });
        const double f = fn;
        if (active && pp + 1 < pend) load(pe_next, pp + 1, rn, qn, fn);
        if (pp + 2 < pend) pe_next = pairs[pp + 2];
        PlaneTile<T, NP>& tile = tiles[buf][active ? p : 0];
        Split::template fft<T, NP, 1>(v, h);  // row ky = line -> x = ob M + k
        if (active) static_for<0, M>([&](auto kc) { tile.a[line][ob * M + decltype(kc)::value] = v[decltype(kc)::value]; });
        __syncthreads();  // buffer `buf` was last read two pairs ago, before the previous barrier
        if (active) static_for<0, M>([&](auto mc) { v[decltype(mc)::value] = tile.a[S * decltype(mc)::value + h][line]; });
        Split::template fft<T, NP, 1>(v, h);  // column x = line -> y = ob M + k
        if (active)
            static_for<0, M>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                const cpx<T> prod = v[k] * q[k];
                acc[k].re += f * static_cast<double>(prod.re);
                acc[k].im += f * static_cast<double>(prod.im);
            });
    }
    if (!active) return;
    const size_t o = (size_t)te.app * ng + (size_t)z * NP * NP + (size_t)ob * M * NP + line;
    static_for<0, M>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        add_partial((double2*)(W + o + (size_t)k * NP), Wc ? Wc + o + (size_t)k * NP : nullptr, acc[k].re,
                    acc[k].im);
    });
}

// ----------------------------------------------------------------- launchers
void launch_hf2x_xyA(const Geom& g, const void* app, const void* occ, const PairEntry* pairs, int npairs,
                     const int* colxy, const int* coloff, void* rec, int cstride, hipStream_t s)
{
    if (npairs == 0) return;
    if (npairs > 65535) throw Error(TCX_ERR_INTERNAL, "hf2x: too many pairs in one launch");
    if (g.n1 != g.n2) throw Error(TCX_ERR_UNSUPPORTED, "FFT grid must have n1 == n2");
#define TCX_XA(NP)                                                                                            \
    {                                                                                                         \
        using PG = PlaneGroup<float, NP>;                                                                     \
        hipLaunchKernelGGL((k_hf2x_xyA<float, NP>), dim3((g.n3 + PG::planes - 1) / PG::planes, npairs),       \
                           dim3(PG::threads), 0, s, (const cpx<float>*)app, (const cpx<float>*)occ, pairs, colxy, \
                           coloff, (cpx<float>*)rec, cstride, g.n3);                                          \
    }
    TCX_SIZE_SWITCH(g.n1, TCX_XA)
#undef TCX_XA
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_hf2x_zB(const Geom& g, void* rec, const PairEntry* pairs, int npairs, const int* colxy, const int* coloff,
                    const double* kqcart, int cstride, hipStream_t s)
{
    if (npairs == 0 || cstride == 0) return;
    const double invN = 1.0 / ((double)g.n1 * g.n2 * g.n3);
#define TCX_ZB(NZ)                                                                                            \
    hipLaunchKernelGGL((k_hf2x_zB<float, NZ>), dim3((cstride + 31) / 32, npairs),                             \
                       dim3(kPencilBlock), 0, s, (cpx<float>*)rec, pairs, colxy, coloff, kqcart, cstride, g, invN);
    TCX_SIZE_SWITCH(g.n3, TCX_ZB)
#undef TCX_ZB
    TCX_HIP_CHECK(hipGetLastError());
}

void launch_hf2x_xyC(const Geom& g, const void* rec, const void* occ, const PairEntry* pairs,
                     const TargetEntry* targets, int ntargets, const int* rowinfo, const double* fill, double* W,
                     float* Wc, int cstride, hipStream_t s)
{
    if (ntargets == 0) return;
    if (ntargets > 65535) throw Error(TCX_ERR_INTERNAL, "hf2x: too many targets in one launch");
#define TCX_XC(NP)                                                                                               \
    {                                                                                                            \
        using PG = PlaneGroupC<float, NP>;                                                                       \
        hipLaunchKernelGGL((k_hf2x_xyC<float, NP>), dim3((g.n3 + PG::planes - 1) / PG::planes, ntargets),        \
                           dim3(PG::threads), 0, s, (const cpx<float>*)rec, (const cpx<float>*)occ, pairs, targets, \
                           (const int4*)rowinfo, fill, (cpx<double>*)W, (float2*)Wc, cstride, g.n3);              \
    }
    TCX_SIZE_SWITCH(g.n1, TCX_XC)
#undef TCX_XC
    TCX_HIP_CHECK(hipGetLastError());
}

}  // namespace tcx
