// kernels_tc3xy.hip -- xy-plane passes of the three-body record pipelines with the
// per-term G-space combine fused in.
//
// A record channel arrives transformed along z only (the products pass); one workgroup
// owns plane z of one pair and runs, in LDS tiles of four wavefronts,
//     FFT_y, FFT_x of the input channels        (completing FFT_forward),
//     the term's combine at every G of the plane,
//     inverse FFT_x, FFT_y of the output channels (the xy half of FFT_backward),
// so that every input channel is read once and every output written once; the z pass of
// FFT_backward follows separately or fused with the consumer.  The combines are the
// reference's G-space products with du_s,d(G) = (kq+G)_d u(kq+G, s):
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

constexpr int kXyWaves = 4;
constexpr int kXyBlock = 64 * kXyWaves;

template <typename T, int NP>
struct XyTiles {
    PlaneTile<T, NP> t[kXyWaves];
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

// Lanes t < NP of a wavefront with a job: column x = t -> FFT_y -> tile -> row y = t ->
// FFT_x -> tile.  Every thread of the workgroup calls it (two barriers).
template <typename T, int NP>
__device__ __forceinline__ void xy_forward(PlaneTile<T, NP>& tile, const cpx<T>* plane, bool job, int t)
{
    cpx<T> v[NP];
    const bool on = job && t < NP;
    if (on) {
        load_pencil_nt<T, NP>(v, plane + t, NP);  // record planes are streamed once
        fft<T, NP, -1>(v);
        tile_put_col<T, NP>(tile, v, t);
    }
    __syncthreads();
    if (on) {
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
    const bool on = job && t < NP;
    if (on) {
        tile_get_row<T, NP>(tile, v, t);
        fft<T, NP, 1>(v);
        tile_put_row<T, NP>(tile, v, t);
    }
    __syncthreads();
    if (on) {
        tile_get_col<T, NP>(tile, v, t);
        fft<T, NP, 1>(v);
        store_pencil_nt<T, NP>(v, plane + t, NP);
    }
    __syncthreads();
}

// Plane points per thread: thread t owns e = t + i kXyBlock, i < xy_ept(NP).
constexpr int xy_ept(int np) { return (np * np + kXyBlock - 1) / kXyBlock; }

// src[e] at the thread's plane points.  Issued before the forward transform so that the
// loads overlap it: a load inside a combine loop waits once per iteration, and with it
// for every store issued before it.
template <typename T, int NP>
__device__ __forceinline__ void load_points(cpx<T> (&v)[xy_ept(NP)], const cpx<T>* __restrict__ src)
{
    static_for<0, xy_ept(NP)>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const int e = threadIdx.x + i * kXyBlock;
        if (e < NP * NP) v[i] = src[e];
    });
}

// f(ic, e, x, y) at the thread's plane points, ic = std::integral_constant<int, i>
template <int NP, typename F>
__device__ __forceinline__ void for_points(F&& f)
{
    static_for<0, xy_ept(NP)>([&](auto ic) {
        const int e = threadIdx.x + decltype(ic)::value * kXyBlock;
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
__global__ __launch_bounds__(kXyBlock) void k_xy_a7c1(const cpx<T>* __restrict__ rec, cpx<T>* __restrict__ P,
                                                      cpx<T>* __restrict__ Y, const PairEntry* __restrict__ pairs,
                                                      const cpx<T>* __restrict__ tab, int kq0,
                                                      const double* __restrict__ kqcart,
                                                      const double* __restrict__ fill, int nz, Geom g, double fb3)
{
    __shared__ XyTiles<T, NP> S;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int z = blockIdx.x, p = blockIdx.y;
    const size_t ng = (size_t)NP * NP * nz, pz = (size_t)z * NP * NP;
    const PairEntry pe = pairs[p];
    const PlanePoint pp(kqcart + 3 * pe.kq, g, z, nz);
    const cpx<T>* tp = tab + (size_t)(pe.kq - kq0) * ng + pz;
    const double f = fill[pe.occ];
    cpx<T> u[xy_ept(NP)];
    load_points<T, NP>(u, tp);
    xy_forward<T, NP>(S.t[w], rec + ((size_t)p * 4 + w) * ng + pz, true, t);
    for_points<NP>([&](auto ic, int e, int x, int y) {
        double xd[3];
        pp.at(g, x, y, NP, xd);
        const double u0 = (double)u[decltype(ic)::value].re;
        const cpx<T> r = S.t[0].a[y][x];
        double2 acc = make_double2(0.0, 0.0);
        for (int d = 0; d < 3; ++d) {
            const double2 s = up2(S.t[1 + d].a[y][x]);
            acc.x += s.x * (xd[d] * u0);
            acc.y += s.y * (xd[d] * u0);
        }
        Y[(size_t)p * ng + pz + e] = down2<T>(make_double2(acc.x * (f * fb3), acc.y * (f * fb3)));
        for (int d = 0; d < 3; ++d) S.t[1 + d].a[y][x] = mulr(r, f * xd[d] * u0);
    });
    __syncthreads();
    const int ow = w >= 1 ? w - 1 : 0;  // output slot of wavefront w (wavefront 0 stores nothing)
    xy_inverse_store<T, NP>(S.t[w], P + ((size_t)p * 6 + ow) * ng + pz, w >= 1, t);
    for_points<NP>([&](auto ic, int e, int x, int y) {
        double xd[3];
        pp.at(g, x, y, NP, xd);
        const double u1 = (double)u[decltype(ic)::value].im;
        const cpx<T> r = S.t[0].a[y][x];
        for (int d = 0; d < 3; ++d) S.t[1 + d].a[y][x] = mulr(r, f * xd[d] * u1);
    });
    __syncthreads();
    xy_inverse_store<T, NP>(S.t[w], P + ((size_t)p * 6 + 3 + ow) * ng + pz, w >= 1, t);
}

// Phase-1 second combine ([3b3] and [3b6] merged by linearity).  In: P[p][0..5] z-transformed
// (after the density step), Y[p] in G space.  Out: Y[p] += b6 sum_{s,d} du_s,d P[p][3s+d],
// inverse-transformed along x and y (in place).
template <typename T, int NP>
__global__ __launch_bounds__(kXyBlock) void k_xy_a7c2(const cpx<T>* __restrict__ P, cpx<T>* __restrict__ Y,
                                                      const PairEntry* __restrict__ pairs,
                                                      const cpx<T>* __restrict__ tab, int kq0,
                                                      const double* __restrict__ kqcart, int nz, Geom g, double fb6)
{
    __shared__ XyTiles<T, NP> S;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int z = blockIdx.x, p = blockIdx.y;
    const size_t ng = (size_t)NP * NP * nz, pz = (size_t)z * NP * NP;
    const PairEntry pe = pairs[p];
    const PlanePoint pp(kqcart + 3 * pe.kq, g, z, nz);
    const cpx<T>* tp = tab + (size_t)(pe.kq - kq0) * ng + pz;
    cpx<T> u[xy_ept(NP)], yv[xy_ept(NP)];
    load_points<T, NP>(u, tp);
    load_points<T, NP>(yv, Y + (size_t)p * ng + pz);
    // channels 0..3 = (s,d) (0,0) (0,1) (0,2) (1,0) in tiles 0..3
    xy_forward<T, NP>(S.t[w], P + ((size_t)p * 6 + w) * ng + pz, true, t);
    for_points<NP>([&](auto ic, int e, int x, int y) {
        constexpr int i = decltype(ic)::value;
        double xd[3];
        pp.at(g, x, y, NP, xd);
        const double u0 = (double)u[i].re, u1 = (double)u[i].im;
        const double c[4] = {xd[0] * u0, xd[1] * u0, xd[2] * u0, xd[0] * u1};
        double2 acc = make_double2(0.0, 0.0);
        for (int q = 0; q < 4; ++q) {
            const double2 s = up2(S.t[q].a[y][x]);
            acc.x += s.x * c[q];
            acc.y += s.y * c[q];
        }
        const double2 y0 = up2(yv[i]);
        S.t[0].a[y][x] = down2<T>(make_double2(y0.x + fb6 * acc.x, y0.y + fb6 * acc.y));
    });
    __syncthreads();
    // channels 4, 5 = (1,1), (1,2) in tiles 1, 2
    xy_forward<T, NP>(S.t[w], P + ((size_t)p * 6 + 3 + (w == 1 || w == 2 ? w : 0)) * ng + pz, w == 1 || w == 2, t);
    for_points<NP>([&](auto ic, int e, int x, int y) {
        double xd[3];
        pp.at(g, x, y, NP, xd);
        const double u1 = (double)u[decltype(ic)::value].im;
        const double2 s1 = up2(S.t[1].a[y][x]), s2 = up2(S.t[2].a[y][x]);
        const double2 y0 = up2(S.t[0].a[y][x]);
        const double c1 = fb6 * xd[1] * u1, c2 = fb6 * xd[2] * u1;
        S.t[0].a[y][x] = down2<T>(make_double2(y0.x + s1.x * c1 + s2.x * c2, y0.y + s1.y * c1 + s2.y * c2));
    });
    __syncthreads();
    xy_inverse_store<T, NP>(S.t[0], Y + (size_t)p * ng + pz, w == 0, t);
}

// Phase-2 combine ([3b4]).  In: rec[p][0..2] z-transformed conj(chi_qi) qsum_j,d.
// Out: Z[p] = b4 f sum_d du_0,d rec[p][d], inverse-transformed along x and y.
template <typename T, int NP>
__global__ __launch_bounds__(kXyBlock) void k_xy_b4(const cpx<T>* __restrict__ rec, cpx<T>* __restrict__ Z,
                                                    const PairEntry* __restrict__ pairs, const cpx<T>* __restrict__ tab,
                                                    int kq0, const double* __restrict__ kqcart,
                                                    const double* __restrict__ fill, int nz, Geom g, double fb4)
{
    __shared__ XyTiles<T, NP> S;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int z = blockIdx.x, p = blockIdx.y;
    const size_t ng = (size_t)NP * NP * nz, pz = (size_t)z * NP * NP;
    const PairEntry pe = pairs[p];
    const PlanePoint pp(kqcart + 3 * pe.kq, g, z, nz);
    const cpx<T>* tp = tab + (size_t)(pe.kq - kq0) * ng + pz;
    const double f = fill[pe.occ];
    cpx<T> u[xy_ept(NP)];
    load_points<T, NP>(u, tp);
    xy_forward<T, NP>(S.t[w], rec + ((size_t)p * 3 + (w < 3 ? w : 0)) * ng + pz, w < 3, t);
    for_points<NP>([&](auto ic, int e, int x, int y) {
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
    xy_inverse_store<T, NP>(S.t[0], Z + (size_t)p * ng + pz, w == 0, t);
}

// ----------------------------------------------------------------- tc3a2a4b2b5
// q1-q2 combine.  In: rec[p][h] z-transformed (h = 0: conj(chi_q1) phi_q2; BITC h = 1:
// conj(phi_q1) chi_q2).  Out: O[p][6h + 3s + d] = du12_s,d rec[p][h], inverse-transformed
// along x and y.  The NIN inputs are transformed side by side (waves 0..NIN-1); each
// thread then keeps its points of them in registers, so all four tiles take outputs and
// the 6 NIN outputs leave in rounds of four (three full rounds for BITC).
template <typename T, int NP, int NIN>
__global__ __launch_bounds__(kXyBlock, sizeof(T) == 4 ? 3 : 1) void k_xy_a8c1(const cpx<T>* __restrict__ rec, cpx<T>* __restrict__ O,
                                                      const PairEntry* __restrict__ pairs,
                                                      const cpx<T>* __restrict__ tab, int kq0,
                                                      const double* __restrict__ kqcart, int nz, Geom g)
{
    constexpr int NOUT = 6 * NIN;
    constexpr int EPT = xy_ept(NP);
    __shared__ XyTiles<T, NP> S;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int z = blockIdx.x, p = blockIdx.y;
    const size_t ng = (size_t)NP * NP * nz, pz = (size_t)z * NP * NP;
    const PairEntry pe = pairs[p];
    const PlanePoint pp(kqcart + 3 * pe.kq, g, z, nz);
    const cpx<T>* tp = tab + (size_t)(pe.kq - kq0) * ng + pz;
    cpx<T> uv[EPT];
    load_points<T, NP>(uv, tp);
    xy_forward<T, NP>(S.t[w], rec + ((size_t)p * NIN + (w < NIN ? w : 0)) * ng + pz, w < NIN, t);
    cpx<T> r[NIN][EPT];
    static_for<0, EPT>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const int e = threadIdx.x + i * kXyBlock;
        if (e < NP * NP) {
            const int x = e % NP, y = e / NP;
            static_for<0, NIN>([&](auto hc) { r[decltype(hc)::value][i] = S.t[decltype(hc)::value].a[y][x]; });
        }
    });
    __syncthreads();  // the inputs have been read before outputs overwrite the tiles
#pragma unroll 1
    for (int o0 = 0; o0 < NOUT; o0 += kXyWaves) {
        // the per-point terms are recomputed each round rather than held across rounds:
        // the empty asm hides that the thread index does not change
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        static_for<0, EPT>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const int e = tid + i * kXyBlock;
            if (e < NP * NP) {
                const int x = e % NP, y = e / NP;
                double xd[3];
                pp.at(g, x, y, NP, xd);
                const cpx<T> u = uv[i];
                static_for<0, kXyWaves>([&](auto wc) {
                    constexpr int wi = decltype(wc)::value;
                    const int o = o0 + wi;
                    if (o < NOUT) {
                        const int sd = o % 6, d = sd % 3;
                        const cpx<T> rh = (NIN == 1 || o < 6) ? r[0][i] : r[NIN - 1][i];
                        const double us = sd < 3 ? (double)u.re : (double)u.im;
                        const double xs = d == 0 ? xd[0] : (d == 1 ? xd[1] : xd[2]);
                        S.t[wi].a[y][x] = mulr(rh, xs * us);
                    }
                });
            }
        });
        __syncthreads();
        const int o = o0 + w;
        xy_inverse_store<T, NP>(S.t[w], O + ((size_t)p * NOUT + (o < NOUT ? o : 0)) * ng + pz, o < NOUT, t);
    }
}

// Apply-phase combine.  In: rec[p][0] z-transformed conj(chi_t) phi_kj, rec[p][1+d]
// z-transformed Q5'_t,d phi_kj.  Out (a separate record): Z[p][0] = b5 f2 sum_d du0_d rec[p][1+d],
// Z[p][1+d] = f2 a4 du0_d rec[p][0], inverse-transformed along x and y.
template <typename T, int NP>
__global__ __launch_bounds__(kXyBlock) void k_xy_a8c2(const cpx<T>* __restrict__ rec, cpx<T>* __restrict__ Z,
                                                      const PairEntry* __restrict__ pairs,
                                                      const cpx<T>* __restrict__ tab, int kq0,
                                                      const double* __restrict__ kqcart,
                                                      const double* __restrict__ fill, int nz, Geom g, double fa4,
                                                      double fb5)
{
    __shared__ XyTiles<T, NP> S;
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int z = blockIdx.x, p = blockIdx.y;
    const size_t ng = (size_t)NP * NP * nz, pz = (size_t)z * NP * NP;
    const PairEntry pe = pairs[p];
    const PlanePoint pp(kqcart + 3 * pe.kq, g, z, nz);
    const cpx<T>* tp = tab + (size_t)(pe.kq - kq0) * ng + pz;
    const double f2 = fill[pe.occ];
    cpx<T> uv[xy_ept(NP)];
    load_points<T, NP>(uv, tp);
    xy_forward<T, NP>(S.t[w], rec + ((size_t)p * 4 + w) * ng + pz, true, t);
    for_points<NP>([&](auto ic, int e, int x, int y) {
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
    xy_inverse_store<T, NP>(S.t[w], Z + ((size_t)p * 4 + w) * ng + pz, true, t);
}

// ----------------------------------------------------------------- launchers
namespace {

template <typename F>
void for_pair_slices(int npairs, F&& f)
{
    for (int a = 0; a < npairs; a += 65535) f(a, npairs - a < 65535 ? npairs - a : 65535);
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
#define TCX_L(NP)                                                                                                    \
    hipLaunchKernelGGL((k_xy_a7c1<T, NP>), dim3(g.n3, n), dim3(kXyBlock), 0, s,                                    \
                       (const cpx<T>*)((const char*)rec + (size_t)a * 4 * gb), (cpx<T>*)((char*)P + (size_t)a * 6 * gb), \
                       (cpx<T>*)((char*)Y + (size_t)a * gb), pairs + a, (const cpx<T>*)tab, kq0, kqcart, fill, g.n3, g,  \
                       fb3);
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
    hipLaunchKernelGGL((k_xy_a7c2<T, NP>), dim3(g.n3, n), dim3(kXyBlock), 0, s,                              \
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
    hipLaunchKernelGGL((k_xy_b4<T, NP>), dim3(g.n3, n), dim3(kXyBlock), 0, s,                                  \
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
#define TCX_L(NP)                                                                                           \
    {                                                                                                       \
        const cpx<T>* in = (const cpx<T>*)((const char*)rec + (size_t)a * nin * gb);                        \
        cpx<T>* out = (cpx<T>*)((char*)O + (size_t)a * 6 * nin * gb);                                       \
        if (nin == 2)                                                                                       \
            hipLaunchKernelGGL((k_xy_a8c1<T, NP, 2>), dim3(g.n3, n), dim3(kXyBlock), 0, s, in, out, pairs + a, \
                               (const cpx<T>*)tab, kq0, kqcart, g.n3, g);                                   \
        else                                                                                                \
            hipLaunchKernelGGL((k_xy_a8c1<T, NP, 1>), dim3(g.n3, n), dim3(kXyBlock), 0, s, in, out, pairs + a, \
                               (const cpx<T>*)tab, kq0, kqcart, g.n3, g);                                   \
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
#define TCX_L(NP)                                                                                                    \
    hipLaunchKernelGGL((k_xy_a8c2<T, NP>), dim3(g.n3, n), dim3(kXyBlock), 0, s,                                    \
                       (const cpx<T>*)((const char*)rec + (size_t)a * 4 * gb), (cpx<T>*)((char*)Z + (size_t)a * 4 * gb), \
                       pairs + a, (const cpx<T>*)tab, kq0, kqcart, fill, g.n3, g, fa4, fb5);
        TCX_XY_DISPATCH(prec, g, TCX_L);
#undef TCX_L
    });
}

#undef TCX_XY_DISPATCH

}  // namespace tcx
