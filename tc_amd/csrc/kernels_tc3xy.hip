// kernels_tc3xy.hip -- xy-plane passes of the three-body record pipelines with the
// per-term G-space combine fused in.
//
// A record channel arrives transformed along z only (the products pass); one workgroup
// owns plane z of one pair and three to six LDS tiles (channel planes), and runs
//     FFT_y, FFT_x of the input channels        (completing FFT_forward),
//     the term's combine at every G of the plane,
//     inverse FFT_x, FFT_y of the output channels (the xy half of FFT_backward),
// so that every input channel is read once and every output written once; the z pass of
// FFT_backward follows separately or fused with the consumer.  Line t of tile c runs on
// thread c NP + t, so a phase that transforms R tiles occupies ceil(R NP / 64) wavefronts
// and the rest skip it; the du_s,d outputs leave in rounds of their three d components.
// The combines are the reference's G-space products with
// du_s,d(G) = (kq+G)_d u(kq+G, s):
//   tc3a3b3b4b6 (src/calc_hamiltonian_tc3a3b3b4b6.cpp): [3b3] and the [3a3,b4,b6] channels
//     (:280-349), the [3b3]+[3b6] merge (:344-358), [3b4] (:740-754);
//   tc3a2a4b2b5 (src/calc_hamiltonian_tc3a2a4b2b5.cpp): the q1-q2 channels (:371-375) and
//     the apply-phase channels (:849-876).
// u(kq+G, s) comes from a per-launch table (FP64 evaluation rounded once to the working
// precision); (kq+G)_d is formed in FP64 per point; products are FP64, stored in T.
#include "jastrow_dev.hpp"
#include "kernels_common.hpp"
#include "size_switch.hpp"

namespace tcx {

namespace {

// Workgroup of the xy passes at grid NP: NT channel planes (LDS tiles), line t of tile c on
// thread c NP + t.
template <int NP, int NT>
struct XyGroup {
    static constexpr int threads = ((NT * NP + 63) / 64) * 64;
    static constexpr int ept = (NP * NP + threads - 1) / threads;  // combine points per thread
};

template <typename T, int NP, int NT>
struct XyTiles {
    PlaneTile<T, NP> t[NT];
};

// Wavefronts per SIMD that the LDS of the tiles allows (160 KB per CU, 4 SIMDs), as the
// register bound of the four-tile pass (a8c2 in FP32) where it is 3 or fewer, so that registers
// do not lower the occupancy further (above that the bound would make the small grids spill;
// the three-tile passes spill under it at 24^3-36^3 and take no bound).
template <typename T, int NP, int NT>
constexpr int xy_waves_per_simd()
{
    if (NT < 4) return 0;  // no bound
    const int wgs = (160 * 1024) / (int)sizeof(XyTiles<T, NP, NT>);
    const int w = (wgs * XyGroup<NP, NT>::threads / 64 + 3) / 4;
    return w < 1 || w > 3 ? 1 : w;
}

// The calling thread's tile and line: c = tile (NT or more on padding lanes), t = line.
template <int NP, int NT>
struct XyLane {
    int c, t;
    __device__ XyLane() : c((int)threadIdx.x / NP), t((int)threadIdx.x % NP) {}
    __device__ int tile() const { return c < NT ? c : 0; }
};

template <typename T>
__device__ __forceinline__ double2 up2(cpx<T> a)
{
    return make_double2((double)a.re, (double)a.im);
}
template <typename T>
__device__ __forceinline__ cpx<T> down2(double2 a)
{
    return cpx<T>{(T)a.x, (T)a.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> mulr(cpx<T> a, double s)  // FP64 product, rounded once
{
    return cpx<T>{(T)((double)a.re * s), (T)((double)a.im * s)};
}

// Threads with a job: column x = t -> FFT_y -> tile -> row y = t -> FFT_x -> tile.  Every
// thread of the workgroup calls it (two barriers).
template <typename T, int NP>
__device__ __forceinline__ void xy_forward(PlaneTile<T, NP>& tile, const cpx<T>* plane, bool job, int t)
{
    cpx<T> v[NP];
    if (job) {
        load_pencil_nt<T, NP>(v, plane + t, NP);  // record planes are streamed once
        fft<T, NP, -1>(v);
        tile_put_col<T, NP>(tile, v, t);
    }
    __syncthreads();
    if (job) {
        tile_get_row<T, NP>(tile, v, t);
        fft<T, NP, -1>(v);
        tile_put_row<T, NP>(tile, v, t);
    }
    __syncthreads();
}

// The reverse: row y = t -> inverse FFT_x -> tile -> column x = t -> inverse FFT_y -> store.
template <typename T, int NP>
__device__ __forceinline__ void xy_inverse_store(PlaneTile<T, NP>& tile, cpx<T>* plane, bool job, int t)
{
    cpx<T> v[NP];
    if (job) {
        tile_get_row<T, NP>(tile, v, t);
        fft<T, NP, 1>(v);
        tile_put_row<T, NP>(tile, v, t);
    }
    __syncthreads();
    if (job) {
        tile_get_col<T, NP>(tile, v, t);
        fft<T, NP, 1>(v);
        store_pencil_nt<T, NP>(v, plane + t, NP);
    }
    __syncthreads();
}

// src[e] at the thread's plane points e = threadIdx.x + i threads.  Issued before the
// forward transform so that the loads overlap it: a load inside a combine loop waits once
// per iteration, and with it for every store issued before it.
template <typename T, int NP, int NT>
__device__ __forceinline__ void load_points(cpx<T> (&v)[(XyGroup<NP, NT>::ept)], const cpx<T>* __restrict__ src)
{
    static_for<0, XyGroup<NP, NT>::ept>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const int e = threadIdx.x + i * XyGroup<NP, NT>::threads;
        if (e < NP * NP) v[i] = src[e];
    });
}

// f(ic, e, x, y) at the thread's plane points, ic = std::integral_constant<int, i>
template <int NP, int NT, typename F>
__device__ __forceinline__ void for_points(F&& f)
{
    static_for<0, XyGroup<NP, NT>::ept>([&](auto ic) {
        const int e = threadIdx.x + decltype(ic)::value * XyGroup<NP, NT>::threads;
        if (e < NP * NP) f(ic, e, e % NP, e / NP);
    });
}

// (kq + G)_d at the points of plane z: x_d = k0[d] + b_1d m_x + b_2d m_y
struct PlanePoint {
    double k0[3];
    __device__ PlanePoint(const double* kq, const Geom& g, int z, int nz)
    {
        const int mz = centered(z, nz);
        for (int d = 0; d < 3; ++d) k0[d] = kq[d] + g.b[6 + d] * mz;
    }
    __device__ void at(const Geom& g, int x, int y, int np, double xd[3]) const
    {
        const int mx = centered(x, np), my = centered(y, np);
        for (int d = 0; d < 3; ++d) xd[d] = k0[d] + g.b[d] * mx + g.b[3 + d] * my;
    }
};

// Outputs of the form du_s,d src: NRND rounds of the three d components, round k in
// tiles 0..2 (tile d), inverse-transformed into dst(3k + d).  Tile d at the thread's point
// i takes src(ic, k) (kq+G)_d scale(k, uv[i]).  With the d components in fixed tiles the
// per-point work has no index arithmetic; three tiles of NP lines fill two wavefronts at 40^3.
template <typename T, int NP, int NT, int NRND, typename Src, typename Scale, typename Dst>
__device__ __forceinline__ void xy_du_rounds(XyTiles<T, NP, NT>& S, const PlanePoint& pp, const Geom& g,
                                             const cpx<T> (&uv)[(XyGroup<NP, NT>::ept)], Src&& src, Scale&& scale,
                                             Dst&& dst)
{
    static_assert(NT >= 3, "the three d components need three tiles");
    const XyLane<NP, NT> L;
#pragma unroll 1
    for (int k = 0; k < NRND; ++k) {
        // the per-point terms are recomputed each round rather than held across rounds:
        // the empty asm hides that the thread index does not change
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        static_for<0, XyGroup<NP, NT>::ept>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const int e = tid + i * XyGroup<NP, NT>::threads;
            if (e < NP * NP) {
                const int x = e % NP, y = e / NP;
                double xd[3];
                pp.at(g, x, y, NP, xd);
                const cpx<T> r = src(ic, k);
                const double sc = scale(k, uv[i]);
                for (int d = 0; d < 3; ++d) S.t[d].a[y][x] = mulr(r, xd[d] * sc);
            }
        });
        __syncthreads();
        const bool job = L.c < 3;
        xy_inverse_store<T, NP>(S.t[L.tile()], dst(3 * k + (job ? L.c : 0)), job, L.t);
    }
}

// a7c2 transforms its six channels at once where two workgroups of six tiles fit in a CU's
// LDS, else in two groups of three (one six-tile workgroup per CU was slower: FP64 40^3
// tc3a3b3b4b6 13.2 -> 13.8 s; FP32 40^3 gained 6.66 -> 6.61 s from six tiles)
template <typename T, int NP>
constexpr int tiles_a7c2()
{
    return 2 * 6 * (int)sizeof(PlaneTile<T, NP>) <= 160 * 1024 ? 6 : 3;
}

// Tiles per workgroup of each pass: as many as its largest group of channels transformed
// together (three tiles give 4 workgroups of 2 wavefronts per CU at 40^3 in FP32, four
// tiles 3 of 3).  a7c1 transforms its four inputs as three, then one.
constexpr int kTilesA7c1 = 3, kTilesB4 = 3, kTilesA8c1 = 3;

// a8c2 transforms its four inputs at once in FP32 and as three, then one, in FP64 (at 40^3
// two workgroups of 2 wavefronts per CU against one of 3): tc3a2a4b2b5 at 40^3 took 6.59 s
// (four) against 6.69 s (three) in FP32, 13.67 s against 13.10 s in FP64
template <typename T>
constexpr int tiles_a8c2()
{
    return sizeof(T) == 4 ? 4 : 3;
}

}  // namespace

// ----------------------------------------------------------------- u table
// tab[row][e] = (u(kq_row + G_e, j0), u(kq_row + G_e, j1)) in the working precision.
template <typename T, bool kPoly>
__global__ __launch_bounds__(256) void k_u2_table(cpx<T>* __restrict__ tab, const double* __restrict__ kqcart, int row0,
                                                  Geom g, JastrowPair j0, JastrowPair j1, long long N)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= N) return;
    const int i1 = (int)(e % g.n1), i2 = (int)((e / g.n1) % g.n2), i3 = (int)(e / ((long long)g.n1 * g.n2));
    const int m1 = centered(i1, g.n1), m2 = centered(i2, g.n2), m3 = centered(i3, g.n3);
    const double* kq = kqcart + 3 * (row0 + blockIdx.y);
    double x2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double c = kq[d] + g.b[d] * m1 + g.b[3 + d] * m2 + g.b[6 + d] * m3;
        x2 += c * c;
    }
    tab[(long long)blockIdx.y * N + e] = cpx<T>{(T)jastrow_uk<kPoly>(j0, x2), (T)jastrow_uk<kPoly>(j1, x2)};
}

// ----------------------------------------------------------------- tc3a3b3b4b6
// Phase-1 combine.  In: rec[p][0..3] = z-transformed (conj(chi_qi) phi_j, conj(chi_qi) dnu_d phi_j).
// Out: Y[p] = f b3 sum_d du_0,d rec[p][1+d] in G space (completed later by the second combine);
//      P[p][3s+d] = f du_s,d rec[p][0], inverse-transformed along x and y.
template <typename T, int NP>
__global__ __launch_bounds__((XyGroup<NP, kTilesA7c1>::threads))
void k_xy_a7c1(
    const cpx<T>* __restrict__ rec, cpx<T>* __restrict__ P, cpx<T>* __restrict__ Y, const PairEntry* __restrict__ pairs,
    const cpx<T>* __restrict__ tab, int kq0, const double* __restrict__ kqcart, const double* __restrict__ fill, int nz,
    Geom g, double fb3)
{
    constexpr int NT = kTilesA7c1;
    constexpr int EPT = XyGroup<NP, NT>::ept;
    __shared__ XyTiles<T, NP, NT> S;
    const XyLane<NP, NT> L;
    const int z = blockIdx.x, p = blockIdx.y;
    const size_t ng = (size_t)NP * NP * nz, pz = (size_t)z * NP * NP;
    const PairEntry pe = pairs[p];
    const PlanePoint pp(kqcart + 3 * pe.kq, g, z, nz);
    const cpx<T>* tp = tab + (size_t)(pe.kq - kq0) * ng + pz;
    const double f = fill[pe.occ];
    cpx<T> u[EPT];
    load_points<T, NP, NT>(u, tp);
    // rec[p][1+d] in tiles d -> Y; then rec[p][0] in tile 0, kept in registers
    xy_forward<T, NP>(S.t[L.tile()], rec + ((size_t)p * 4 + 1 + L.tile()) * ng + pz, L.c < 3, L.t);
    for_points<NP, NT>([&](auto ic, int e, int x, int y) {
        constexpr int i = decltype(ic)::value;
        double xd[3];
        pp.at(g, x, y, NP, xd);
        const double u0 = (double)u[i].re;
        double2 acc = make_double2(0.0, 0.0);
        for (int d = 0; d < 3; ++d) {
            const double2 s = up2(S.t[d].a[y][x]);
            acc.x += s.x * (xd[d] * u0);
            acc.y += s.y * (xd[d] * u0);
        }
        Y[(size_t)p * ng + pz + e] = down2<T>(make_double2(acc.x * (f * fb3), acc.y * (f * fb3)));
    });
    __syncthreads();  // the tiles have been read before rec[p][0] overwrites tile 0
    xy_forward<T, NP>(S.t[0], rec + (size_t)p * 4 * ng + pz, L.c == 0, L.t);
    cpx<T> r0[EPT];
    for_points<NP, NT>([&](auto ic, int e, int x, int y) { r0[decltype(ic)::value] = S.t[0].a[y][x]; });
    __syncthreads();  // the input has been read before outputs overwrite the tiles
    xy_du_rounds<T, NP, NT, 2>(
        S, pp, g, u, [&](auto ic, int) { return r0[decltype(ic)::value]; },
        [&](int k, cpx<T> uk) { return f * (double)(k == 0 ? uk.re : uk.im); },
        [&](int o) { return P + ((size_t)p * 6 + o) * ng + pz; });
}

// Phase-1 second combine ([3b3] and [3b6] merged by linearity).  In: P[p][0..5] z-transformed
// (after the density step), Y[p] in G space.  Out: Y[p] += b6 sum_{s,d} du_s,d P[p][3s+d],
// inverse-transformed along x and y (in place).
template <typename T, int NP>
__global__ __launch_bounds__((XyGroup<NP, tiles_a7c2<T, NP>()>::threads))
void k_xy_a7c2(
    const cpx<T>* __restrict__ P, cpx<T>* __restrict__ Y, const PairEntry* __restrict__ pairs,
    const cpx<T>* __restrict__ tab, int kq0, const double* __restrict__ kqcart, int nz, Geom g, double fb6)
{
    constexpr int NT = tiles_a7c2<T, NP>();
    constexpr int EPT = XyGroup<NP, NT>::ept;
    __shared__ XyTiles<T, NP, NT> S;
    const XyLane<NP, NT> L;
    const int z = blockIdx.x, p = blockIdx.y;
    const size_t ng = (size_t)NP * NP * nz, pz = (size_t)z * NP * NP;
    const PairEntry pe = pairs[p];
    const PlanePoint pp(kqcart + 3 * pe.kq, g, z, nz);
    const cpx<T>* tp = tab + (size_t)(pe.kq - kq0) * ng + pz;
    cpx<T> u[EPT], yv[EPT];
    load_points<T, NP, NT>(u, tp);
    load_points<T, NP, NT>(yv, Y + (size_t)p * ng + pz);
    if constexpr (NT == 6) {
        // all six channels at once: channel 3s + d in tile 3s + d
        xy_forward<T, NP>(S.t[L.tile()], P + ((size_t)p * 6 + L.tile()) * ng + pz, L.c < 6, L.t);
        for_points<NP, NT>([&](auto ic, int e, int x, int y) {
            constexpr int i = decltype(ic)::value;
            double xd[3];
            pp.at(g, x, y, NP, xd);
            double2 a = up2(yv[i]);
            for (int sp = 0; sp < 2; ++sp) {
                const double us = fb6 * (double)(sp == 0 ? u[i].re : u[i].im);
                for (int d = 0; d < 3; ++d) {
                    const double2 v = up2(S.t[3 * sp + d].a[y][x]);
                    a.x += v.x * (xd[d] * us);
                    a.y += v.y * (xd[d] * us);
                }
            }
            S.t[0].a[y][x] = down2<T>(a);
        });
    } else {
        // s = 0, 1 in turn: channels 3s + d in tiles d, the partial sum in FP64 registers
        double2 acc[EPT];
        static_for<0, 2>([&](auto sc) {
            constexpr int sp = decltype(sc)::value;
            if constexpr (sp == 1) __syncthreads();  // the tiles have been read before the next channels overwrite them
            xy_forward<T, NP>(S.t[L.tile()], P + ((size_t)p * 6 + 3 * sp + L.tile()) * ng + pz, L.c < 3, L.t);
            for_points<NP, NT>([&](auto ic, int e, int x, int y) {
                constexpr int i = decltype(ic)::value;
                double xd[3];
                pp.at(g, x, y, NP, xd);
                const double us = fb6 * (double)(sp == 0 ? u[i].re : u[i].im);
                double2 a = sp == 0 ? up2(yv[i]) : acc[i];
                for (int d = 0; d < 3; ++d) {
                    const double2 v = up2(S.t[d].a[y][x]);
                    a.x += v.x * (xd[d] * us);
                    a.y += v.y * (xd[d] * us);
                }
                if constexpr (sp == 0)
                    acc[i] = a;
                else
                    S.t[0].a[y][x] = down2<T>(a);
            });
        });
    }
    __syncthreads();
    xy_inverse_store<T, NP>(S.t[0], Y + (size_t)p * ng + pz, L.c == 0, L.t);
}

// Phase-2 combine ([3b4]).  In: rec[p][0..2] z-transformed conj(chi_qi) qsum_j,d.
// Out: Z[p] = b4 f sum_d du_0,d rec[p][d], inverse-transformed along x and y.
template <typename T, int NP>
__global__ __launch_bounds__((XyGroup<NP, kTilesB4>::threads))
void k_xy_b4(
    const cpx<T>* __restrict__ rec, cpx<T>* __restrict__ Z, const PairEntry* __restrict__ pairs,
    const cpx<T>* __restrict__ tab, int kq0, const double* __restrict__ kqcart, const double* __restrict__ fill, int nz,
    Geom g, double fb4)
{
    constexpr int NT = kTilesB4;
    constexpr int EPT = XyGroup<NP, NT>::ept;
    __shared__ XyTiles<T, NP, NT> S;
    const XyLane<NP, NT> L;
    const int z = blockIdx.x, p = blockIdx.y;
    const size_t ng = (size_t)NP * NP * nz, pz = (size_t)z * NP * NP;
    const PairEntry pe = pairs[p];
    const PlanePoint pp(kqcart + 3 * pe.kq, g, z, nz);
    const cpx<T>* tp = tab + (size_t)(pe.kq - kq0) * ng + pz;
    const double f = fill[pe.occ];
    cpx<T> u[EPT];
    load_points<T, NP, NT>(u, tp);
    xy_forward<T, NP>(S.t[L.tile()], rec + ((size_t)p * 3 + (L.c < 3 ? L.c : 0)) * ng + pz, L.c < 3, L.t);
    for_points<NP, NT>([&](auto ic, int e, int x, int y) {
        double xd[3];
        pp.at(g, x, y, NP, xd);
        const double u0 = (double)u[decltype(ic)::value].re;
        double2 acc = make_double2(0.0, 0.0);
        for (int d = 0; d < 3; ++d) {
            const double2 s = up2(S.t[d].a[y][x]);
            acc.x += s.x * (xd[d] * u0);
            acc.y += s.y * (xd[d] * u0);
        }
        S.t[0].a[y][x] = down2<T>(make_double2(acc.x * (fb4 * f), acc.y * (fb4 * f)));
    });
    __syncthreads();
    xy_inverse_store<T, NP>(S.t[0], Z + (size_t)p * ng + pz, L.c == 0, L.t);
}

// ----------------------------------------------------------------- tc3a2a4b2b5
// q1-q2 combine.  In: rec[p][h] z-transformed (h = 0: conj(chi_q1) phi_q2; BITC h = 1:
// conj(phi_q1) chi_q2).  Out: O[p][6h + 3s + d] = du12_s,d rec[p][h], inverse-transformed
// along x and y.  The NIN inputs are transformed side by side; each thread then keeps its
// points of them in registers, so every tile can take outputs.
template <typename T, int NP, int NIN>
__global__ __launch_bounds__((XyGroup<NP, kTilesA8c1>::threads))
void k_xy_a8c1(
    const cpx<T>* __restrict__ rec, cpx<T>* __restrict__ O, const PairEntry* __restrict__ pairs,
    const cpx<T>* __restrict__ tab, int kq0, const double* __restrict__ kqcart, int nz, Geom g)
{
    constexpr int NOUT = 6 * NIN;
    constexpr int NT = kTilesA8c1;
    constexpr int EPT = XyGroup<NP, NT>::ept;
    __shared__ XyTiles<T, NP, NT> S;
    const XyLane<NP, NT> L;
    const int z = blockIdx.x, p = blockIdx.y;
    const size_t ng = (size_t)NP * NP * nz, pz = (size_t)z * NP * NP;
    const PairEntry pe = pairs[p];
    const PlanePoint pp(kqcart + 3 * pe.kq, g, z, nz);
    const cpx<T>* tp = tab + (size_t)(pe.kq - kq0) * ng + pz;
    cpx<T> uv[EPT];
    load_points<T, NP, NT>(uv, tp);
    xy_forward<T, NP>(S.t[L.tile()], rec + ((size_t)p * NIN + (L.c < NIN ? L.c : 0)) * ng + pz, L.c < NIN, L.t);
    cpx<T> r[NIN][EPT];
    for_points<NP, NT>([&](auto ic, int e, int x, int y) {
        static_for<0, NIN>([&](auto hc) {
            r[decltype(hc)::value][decltype(ic)::value] = S.t[decltype(hc)::value].a[y][x];
        });
    });
    __syncthreads();  // the inputs have been read before outputs overwrite the tiles
    // round k: input h = k / 2, s = k % 2 (outputs 6h + 3s + d)
    xy_du_rounds<T, NP, NT, 2 * NIN>(
        S, pp, g, uv,
        [&](auto ic, int k) {
            constexpr int i = decltype(ic)::value;
            return (NIN == 1 || k < 2) ? r[0][i] : r[NIN - 1][i];
        },
        [&](int k, cpx<T> uk) { return (double)((k & 1) ? uk.im : uk.re); },
        [&](int o) { return O + ((size_t)p * NOUT + o) * ng + pz; });
}

// Apply-phase combine.  In: rec[p][0] z-transformed conj(chi_t) phi_kj, rec[p][1+d]
// z-transformed Q5'_t,d phi_kj.  Out (a separate record): Z[p][0] = b5 f2 sum_d du0_d rec[p][1+d],
// Z[p][1+d] = f2 a4 du0_d rec[p][0], inverse-transformed along x and y.
template <typename T, int NP>
__global__ __launch_bounds__((XyGroup<NP, tiles_a8c2<T>()>::threads),
                             (xy_waves_per_simd<T, NP, tiles_a8c2<T>()>()))
void k_xy_a8c2(
    const cpx<T>* __restrict__ rec, cpx<T>* __restrict__ Z, const PairEntry* __restrict__ pairs,
    const cpx<T>* __restrict__ tab, int kq0, const double* __restrict__ kqcart, const double* __restrict__ fill, int nz,
    Geom g, double fa4, double fb5)
{
    constexpr int NT = tiles_a8c2<T>();
    constexpr int EPT = XyGroup<NP, NT>::ept;
    __shared__ XyTiles<T, NP, NT> S;
    const XyLane<NP, NT> L;
    const int z = blockIdx.x, p = blockIdx.y;
    const size_t ng = (size_t)NP * NP * nz, pz = (size_t)z * NP * NP;
    const PairEntry pe = pairs[p];
    const PlanePoint pp(kqcart + 3 * pe.kq, g, z, nz);
    const cpx<T>* tp = tab + (size_t)(pe.kq - kq0) * ng + pz;
    const double f2 = fill[pe.occ];
    cpx<T> uv[EPT];
    load_points<T, NP, NT>(uv, tp);
    if constexpr (NT == 4) {
        xy_forward<T, NP>(S.t[L.tile()], rec + ((size_t)p * 4 + L.tile()) * ng + pz, L.c < 4, L.t);
        for_points<NP, NT>([&](auto ic, int e, int x, int y) {
            double xd[3];
            pp.at(g, x, y, NP, xd);
            const double u = (double)uv[decltype(ic)::value].re;
            const cpx<T> r0 = S.t[0].a[y][x];
            double2 b = make_double2(0.0, 0.0);
            for (int d = 0; d < 3; ++d) {
                const double2 s = up2(S.t[1 + d].a[y][x]);
                b.x += s.x * (xd[d] * u);
                b.y += s.y * (xd[d] * u);
            }
            S.t[0].a[y][x] = down2<T>(make_double2(b.x * (fb5 * f2), b.y * (fb5 * f2)));
            for (int d = 0; d < 3; ++d) S.t[1 + d].a[y][x] = mulr(r0, f2 * fa4 * xd[d] * u);
        });
        __syncthreads();
        xy_inverse_store<T, NP>(S.t[L.tile()], Z + ((size_t)p * 4 + L.tile()) * ng + pz, L.c < 4, L.t);
    } else {
        // rec[p][1+d] in tiles d -> Z[p][0] (tile 0); then rec[p][0] in tile 0, kept in registers
        xy_forward<T, NP>(S.t[L.tile()], rec + ((size_t)p * 4 + 1 + L.tile()) * ng + pz, L.c < 3, L.t);
        for_points<NP, NT>([&](auto ic, int e, int x, int y) {
            double xd[3];
            pp.at(g, x, y, NP, xd);
            const double u = (double)uv[decltype(ic)::value].re;
            double2 b = make_double2(0.0, 0.0);
            for (int d = 0; d < 3; ++d) {
                const double2 s = up2(S.t[d].a[y][x]);
                b.x += s.x * (xd[d] * u);
                b.y += s.y * (xd[d] * u);
            }
            S.t[0].a[y][x] = down2<T>(make_double2(b.x * (fb5 * f2), b.y * (fb5 * f2)));
        });
        __syncthreads();
        xy_inverse_store<T, NP>(S.t[0], Z + (size_t)p * 4 * ng + pz, L.c == 0, L.t);
        xy_forward<T, NP>(S.t[0], rec + (size_t)p * 4 * ng + pz, L.c == 0, L.t);
        cpx<T> r0[EPT];
        for_points<NP, NT>([&](auto ic, int e, int x, int y) { r0[decltype(ic)::value] = S.t[0].a[y][x]; });
        __syncthreads();  // the input has been read before outputs overwrite the tiles
        xy_du_rounds<T, NP, NT, 1>(
            S, pp, g, uv, [&](auto ic, int) { return r0[decltype(ic)::value]; },
            [&](int, cpx<T> uk) { return f2 * fa4 * (double)uk.re; },
            [&](int o) { return Z + ((size_t)p * 4 + 1 + o) * ng + pz; });
    }
}

// ----------------------------------------------------------------- launchers
namespace {

template <typename F>
void for_pair_slices(int npairs, F&& f)
{
    for (int a = 0; a < npairs; a += 65535) f(a, npairs - a < 65535 ? npairs - a : 65535);
}

template <int NP, int NT>
dim3 xy_block()
{
    return dim3(XyGroup<NP, NT>::threads);
}

void check_plane(const Geom& g)
{
    if (g.n1 != g.n2) throw Error(TCX_ERR_UNSUPPORTED, "FFT grid must have n1 == n2");
}

}  // namespace

void launch_u2_table(int prec, void* tab, const double* kqcart, int row0, int nrows, const Geom& g,
                     const JastrowPair& j0, const JastrowPair& j1, hipStream_t s)
{
    if (nrows == 0) return;
    const long long N = (long long)g.n1 * g.n2 * g.n3;
    const dim3 grid((unsigned)((N + 255) / 256), nrows);
    if (prec == TCX_PREC_FP64)
        TCX_POLY_DISPATCH(j0.pdeg > 0, hipLaunchKernelGGL((k_u2_table<double, kPoly>), grid, dim3(256), 0, s,
                                                          (cpx<double>*)tab, kqcart, row0, g, j0, j1, N));
    else
        TCX_POLY_DISPATCH(j0.pdeg > 0, hipLaunchKernelGGL((k_u2_table<float, kPoly>), grid, dim3(256), 0, s,
                                                          (cpx<float>*)tab, kqcart, row0, g, j0, j1, N));
    TCX_HIP_CHECK(hipGetLastError());
}

#define TCX_XY_DISPATCH(prec, g, LAUNCH)                      \
    do {                                                      \
        check_plane(g);                                       \
        if ((prec) == TCX_PREC_FP64) {                        \
            using T = double;                                 \
            TCX_SIZE_SWITCH((g).n1, LAUNCH)                   \
        } else {                                              \
            using T = float;                                  \
            TCX_SIZE_SWITCH((g).n1, LAUNCH)                   \
        }                                                     \
        TCX_HIP_CHECK(hipGetLastError());                     \
    } while (0)

void launch_xy_a7c1(int prec, const Geom& g, const void* rec, void* P, void* Y, const PairEntry* pairs, int npairs,
                    const void* tab, int kq0, const double* kqcart, const double* fill, double fb3, hipStream_t s)
{
    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8, gb = (size_t)g.n1 * g.n2 * g.n3 * es;
    for_pair_slices(npairs, [&](int a, int n) {
#define TCX_L(NP)                                                                                               \
    hipLaunchKernelGGL((k_xy_a7c1<T, NP>), dim3(g.n3, n), (xy_block<NP, kTilesA7c1>()), 0, s,                   \
                       (const cpx<T>*)((const char*)rec + (size_t)a * 4 * gb),                                  \
                       (cpx<T>*)((char*)P + (size_t)a * 6 * gb), (cpx<T>*)((char*)Y + (size_t)a * gb), pairs + a, \
                       (const cpx<T>*)tab, kq0, kqcart, fill, g.n3, g, fb3);
        TCX_XY_DISPATCH(prec, g, TCX_L);
#undef TCX_L
    });
}

void launch_xy_a7c2(int prec, const Geom& g, const void* P, void* Y, const PairEntry* pairs, int npairs,
                    const void* tab, int kq0, const double* kqcart, double fb6, hipStream_t s)
{
    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8, gb = (size_t)g.n1 * g.n2 * g.n3 * es;
    for_pair_slices(npairs, [&](int a, int n) {
#define TCX_L(NP)                                                                                              \
    hipLaunchKernelGGL((k_xy_a7c2<T, NP>), dim3(g.n3, n), (xy_block<NP, tiles_a7c2<T, NP>()>()), 0, s,         \
                       (const cpx<T>*)((const char*)P + (size_t)a * 6 * gb), (cpx<T>*)((char*)Y + (size_t)a * gb), \
                       pairs + a, (const cpx<T>*)tab, kq0, kqcart, g.n3, g, fb6);
        TCX_XY_DISPATCH(prec, g, TCX_L);
#undef TCX_L
    });
}

void launch_xy_b4(int prec, const Geom& g, const void* rec, void* Z, const PairEntry* pairs, int npairs,
                  const void* tab, int kq0, const double* kqcart, const double* fill, double fb4, hipStream_t s)
{
    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8, gb = (size_t)g.n1 * g.n2 * g.n3 * es;
    for_pair_slices(npairs, [&](int a, int n) {
#define TCX_L(NP)                                                                                                \
    hipLaunchKernelGGL((k_xy_b4<T, NP>), dim3(g.n3, n), (xy_block<NP, kTilesB4>()), 0, s,                        \
                       (const cpx<T>*)((const char*)rec + (size_t)a * 3 * gb), (cpx<T>*)((char*)Z + (size_t)a * gb), \
                       pairs + a, (const cpx<T>*)tab, kq0, kqcart, fill, g.n3, g, fb4);
        TCX_XY_DISPATCH(prec, g, TCX_L);
#undef TCX_L
    });
}

void launch_xy_a8c1(int prec, const Geom& g, const void* rec, int nin, void* O, const PairEntry* pairs, int npairs,
                    const void* tab, int kq0, const double* kqcart, hipStream_t s)
{
    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8, gb = (size_t)g.n1 * g.n2 * g.n3 * es;
    for_pair_slices(npairs, [&](int a, int n) {
#define TCX_L(NP)                                                                                   \
    {                                                                                               \
        const cpx<T>* in = (const cpx<T>*)((const char*)rec + (size_t)a * nin * gb);                \
        cpx<T>* out = (cpx<T>*)((char*)O + (size_t)a * 6 * nin * gb);                               \
        const dim3 grid(g.n3, n), block = xy_block<NP, kTilesA8c1>();                               \
        if (nin == 2)                                                                               \
            hipLaunchKernelGGL((k_xy_a8c1<T, NP, 2>), grid, block, 0, s, in, out, pairs + a,        \
                               (const cpx<T>*)tab, kq0, kqcart, g.n3, g);                           \
        else                                                                                        \
            hipLaunchKernelGGL((k_xy_a8c1<T, NP, 1>), grid, block, 0, s, in, out, pairs + a,        \
                               (const cpx<T>*)tab, kq0, kqcart, g.n3, g);                           \
    }
        TCX_XY_DISPATCH(prec, g, TCX_L);
#undef TCX_L
    });
}

void launch_xy_a8c2(int prec, const Geom& g, const void* rec, void* Z, const PairEntry* pairs, int npairs,
                    const void* tab, int kq0, const double* kqcart, const double* fill, double fa4, double fb5,
                    hipStream_t s)
{
    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8, gb = (size_t)g.n1 * g.n2 * g.n3 * es;
    for_pair_slices(npairs, [&](int a, int n) {
#define TCX_L(NP)                                                                                                 \
    hipLaunchKernelGGL((k_xy_a8c2<T, NP>), dim3(g.n3, n), (xy_block<NP, tiles_a8c2<T>()>()), 0, s,                     \
                       (const cpx<T>*)((const char*)rec + (size_t)a * 4 * gb),                                    \
                       (cpx<T>*)((char*)Z + (size_t)a * 4 * gb), pairs + a, (const cpx<T>*)tab, kq0, kqcart, fill, \
                       g.n3, g, fa4, fb5);
        TCX_XY_DISPATCH(prec, g, TCX_L);
#undef TCX_L
    });
}

#undef TCX_XY_DISPATCH

}  // namespace tcx
