// tcx_internal.hpp -- shared declarations of the HIP engine behind include/tcx.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/tcx.h"
#include "jastrow_dev.hpp"

namespace tcx {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define TCX_HIP_CHECK(call)                                                                   \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::tcx::Error(e_ == hipErrorOutOfMemory ? TCX_ERR_NOMEM : TCX_ERR_HIP,       \
                               std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

// The fixed device array of tcx_hf2x_noalloc (tcx_noalloc.hip): looks up its address on
// `device` (which must be current) once per device; returns a TCX_ status.
int scratch_setup(int device);
size_t scratch_bytes();

// Grid sizes (per axis) with compiled kernels.
constexpr int kSizes[] = {6, 8, 10, 12, 16, 18, 20, 24, 30, 32, 36, 40, 48};
bool size_supported(int n);

// Geometry passed by value to kernels.
struct Geom {
    int n1, n2, n3;
    double b[9];   // reciprocal rows b_i: b[3*i + d]
};

// One (k,j,q,i) pair of a launch: indices into the real-space grid arrays.
struct PairEntry {
    int app;   // applied-orbital grid (k, j)
    int occ;   // occupied-orbital grid (q, i); also indexes fillings
    int kq;    // row of the launch's k-q table
    int pad;
};

// One accumulation target (k, j) of a launch: pairs [pair0, pair0+npair).
struct TargetEntry {
    int app;
    int pair0;
    int npair;
    int pad;
};

// One symmetry rotation of a q2-resolved sum: the grids of occupied orbital
// `src` (an irreducible q2) mapped to those of `dst` (a replica of it).
struct SymRotJob {
    int dst, src;
    int trev;       // time reversal used at the replica
    int ud[9];      // U^+ = U^T (integer, crystal), row-major
    double udc[9];  // U^+ in Cartesian coordinates
    double tau[3];  // fractional translation (crystal)
};

// One orbital scatter/gather job: coefficients <-> grid.  The grid slot of
// job i of a launch is i (grids are consecutive from the base pointer).
struct OrbEntry {
    long long coef_off;   // offset (complex numbers) into the coefficient array
    int npw;
    int gidx_off;         // offset into the concatenated gindex array
    int kidx;             // k-point (for gradients)
    int dir;              // -1: plain scatter; 0..2: multiply by the Cartesian (k+G)_dir
};

// Divergence-correction job for one applied band (k, j).
struct DivEntry {
    long long app_off;    // coefficient offset of phi_kj (and of hphi_kj)
    long long occ_off;    // coefficient offset of the occupied block at k
    int npw;
    int nocc;
    int fill_off;         // offset of the fillings at k
    int kidx;             // k-point: divergence factor dc[kidx]
};

// One channel of a pair record: record = A * B with A, B grids of the working
// precision; grid index of A = (sel ? pair.occ : pair.app) * mul + off.
struct ProdChan {
    const void* a;
    const void* b;
    int a_sel, a_mul, a_off, a_conj;
    int b_sel, b_mul, b_off, b_conj;
};

// ------------------------------------------------------------- launchers
// prec: TCX_PREC_FP64 (double grids) or TCX_PREC_FP32 (float grids).

void launch_zero(void* p, size_t bytes, hipStream_t s);
void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

// scatter complex128 coefficients to double grids (grids pre-zeroed), grid i <- job i;
// a job with dir >= 0 multiplies by the Cartesian component dir of (k+G).
void launch_scatter(const double* coef, const int* gindex, const OrbEntry* jobs, int njobs, int max_npw,
                    double* grids, long long ngrid_pts, const double* kcart, const Geom& g, hipStream_t s);

// hphi[job] += factor * W[job.grid][gindex]   (W: double G-space grids)
void launch_gather_add(const double* W, const int* gindex, const OrbEntry* jobs, int njobs, int max_npw,
                       double* out, long long ngrid_pts, double factor, hipStream_t s);

void launch_convert_d2f(const double* in, float* out, long long ncomplex, hipStream_t s);

// in-place 3D FFT of ngrids consecutive grids; sign -1 forward, +1 backward; result scaled by scale
void launch_fft3d(int prec, int sign, void* grids, int ngrids, const Geom& g, double scale, hipStream_t s);
// the xy-plane half of a 3D transform (no scaling): with a z-pencil pass before it, a
// complete 3D FFT
void launch_fft_xy(int prec, int sign, void* grids, int ngrids, const Geom& g, hipStream_t s);
// the z-pencil half, scaled by scale
void launch_fft_z(int prec, int sign, void* grids, int ngrids, const Geom& g, double scale, hipStream_t s);

// hf2x pair pipeline
// plane-first hf2x (kernels_hf2x_xy.hip), FP32-split
void launch_hf2x_xyA(const Geom& g, const void* app, const void* occ, const PairEntry* pairs, int npairs,
                     const int* colxy, const int* coloff, void* rec, int cstride, hipStream_t s);
void launch_hf2x_zB(const Geom& g, void* rec, const PairEntry* pairs, int npairs, const int* colxy, const int* coloff,
                    const double* kqcart, int cstride, hipStream_t s);
void launch_hf2x_xyC(const Geom& g, const void* rec, const void* occ, const PairEntry* pairs,
                     const TargetEntry* targets, int ntargets, const int* rowinfo, const double* fill, double* W,
                     float* Wc, int cstride, hipStream_t s);
void launch_hf2x_A(int prec, const Geom& g, const void* phiR_app, const void* phiR_occ, const PairEntry* pairs,
                   int npairs, void* buf, hipStream_t s);
void launch_coulomb_table(int prec, void* table, const double* kqcart, int row0, int nrows, const Geom& g,
                          hipStream_t s);
void launch_hf2x_B(int prec, const Geom& g, void* buf, const PairEntry* pairs, int npairs, const void* table,
                   int kq0, hipStream_t s);
// W[target] += the launch's band-pair sums; with Wc (complex FP32, the layout of W) non-null
// the rounding errors of these additions are summed into Wc (W + Wc carries the sum); the
// same for the other pair accumulators below
void launch_hf2x_C(int prec, const Geom& g, const void* buf, const void* phiR_occ, const PairEntry* pairs,
                   const TargetEntry* targets, int ntargets, const double* fill, double* W, float* Wc,
                   hipStream_t s);

// divergence correction: hphi += sum_b dc f_b <chi_b|phi_j> phi_b
void launch_divcorr(const double* phi, const double* occ_right, const double* occ_left, const DivEntry* jobs,
                    int njobs, const double* fill, const double* dc, double* out, hipStream_t s);

// tc2x pair pipeline (4-channel pair records)
void launch_tc_table(int prec, void* table, const double* kqcart, int row0, int nrows, const Geom& g,
                     const JastrowPair& jp, hipStream_t s);
void launch_tc2x_A(int prec, const Geom& g, const void* appR, const void* appGrad, const void* chiR,
                   const PairEntry* pairs, int npairs, void* buf, hipStream_t s);
void launch_tc2x_B(int prec, const Geom& g, void* buf, const PairEntry* pairs, int npairs, const void* table, int kq0,
                   const double* kqcart, hipStream_t s);
void launch_tc2x_C(int prec, const Geom& g, const void* buf, const void* occR, const void* gradR,
                   const PairEntry* pairs, const TargetEntry* targets, int ntargets, const double* fill, double* W,
                   float* Wc, hipStream_t s);

// density and local three-body potentials (FP64 grids)
// nspin 1: out[0] = out[1] = half the filled density of the nocc orbitals (the filling holds
// the spin factor 2); nspin 2: out[0] = the filled density (one spin channel)
void launch_density(const double* occR, const double* chiR, const double* fill, int nocc, long long N, double* out,
                    int nspin, hipStream_t s);
void launch_jastrow_eval(const double* G2, double* uk, double* v, long long n, const JastrowPair& j, hipStream_t s);
void launch_dnu(const double* nG, double* out, const Geom& g, const JastrowPair& j0, const JastrowPair& j1,
                hipStream_t s);
void launch_tc3b1_V(const double* dnu, const double* dnu1, double* V, const Geom& g, const JastrowPair& j0,
                    const JastrowPair& j1, hipStream_t s);
void launch_sum_squares(const double* dnu, double* V, long long N, hipStream_t s);
void launch_mul_shared(double* grids, int ngrids, const double* V, long long N, hipStream_t s);

// three-body exchange building blocks (records of the working precision, FP64 accumulators)
// records = FFT_forward(A * B) scaled by sc (the products fused with the z pass)
void launch_products_fft(int prec, const PairEntry* pairs, int npairs, const ProdChan* chans, int nchan, void* out,
                         const Geom& g, double sc, hipStream_t s);
// the same records transformed along z only (an xy pass with a fused combine completes them)
void launch_products_z(int prec, const PairEntry* pairs, int npairs, const ProdChan* chans, int nchan, void* out,
                       const Geom& g, double sc, hipStream_t s);
// xy-plane passes with the three-body combines fused in (kernels_tc3xy.hip); tab holds
// (u(kq+G, j0), u(kq+G, j1)) per k-q row of the launch (launch_u2_table)
void launch_u2_table(int prec, void* tab, const double* kqcart, int row0, int nrows, const Geom& g,
                     const JastrowPair& j0, const JastrowPair& j1, hipStream_t s);
void launch_xy_a7c1(int prec, const Geom& g, const void* rec, void* P, void* Y, const PairEntry* pairs, int npairs,
                    const void* tab, int kq0, const double* kqcart, const double* fill, double fb3, hipStream_t s);
void launch_xy_a7c2(int prec, const Geom& g, const void* P, void* Y, const PairEntry* pairs, int npairs,
                    const void* tab, int kq0, const double* kqcart, double fb6, hipStream_t s);
void launch_xy_b4(int prec, const Geom& g, const void* rec, void* Z, const PairEntry* pairs, int npairs,
                  const void* tab, int kq0, const double* kqcart, const double* fill, double fb4, hipStream_t s);
void launch_xy_a8c1(int prec, const Geom& g, const void* rec, int nin, void* O, const PairEntry* pairs, int npairs,
                    const void* tab, int kq0, const double* kqcart, hipStream_t s);
void launch_xy_a8c2(int prec, const Geom& g, const void* rec, void* Z, const PairEntry* pairs, int npairs,
                    const void* tab, int kq0, const double* kqcart, const double* fill, double fa4, double fb5,
                    hipStream_t s);
// z passes of FFT_backward fused with the real-space consumers (kernels_tc3z.hip)
// dens_stride: complex elements between the two spin densities (0: one independent spin)
void launch_za7_mid(int prec, const Geom& g, void* P, const void* occR, const PairEntry* pairs,
                    const TargetEntry* targets, int ntargets, const double* dens, long long dens_stride,
                    double* qsum, double sc, hipStream_t s);
void launch_za8_q(int prec, const Geom& g, const void* O, int bitc, const void* occR, const void* chiR,
                  const PairEntry* pairs, const TargetEntry* targets, int ntargets, const double* fill, double* Q2,
                  double* Q4, double* Q5, hipStream_t s);
void launch_za8_apply(int prec, const Geom& g, const void* Z, const void* occR, const void* Q4T,
                      const PairEntry* pairs, const TargetEntry* targets, int ntargets, double* W, float* Wc,
                      hipStream_t s);
void launch_mul_dnu(int prec, const void* appR, const double* dnu, void* out, int napp, long long N, hipStream_t s);
void launch_a3(double* W, const double* qsum, const double* dnu, int napp, long long N, double fa3, hipStream_t s);
void launch_b6(const double* nG, double* tG, double* t1, const double* dcp, const int* match, int nk, const Geom& g,
               long long N,
               const JastrowPair& j0, const JastrowPair& j1, hipStream_t s);
void launch_b6_coef(const double* phi, const double* occ_r, const double* occ_l, const DivEntry* jobs, int njobs,
                    int maxocc, const double* fill, const double* t1, const double* appR, const double* chiR,
                    const double* occR, const double* tR, double fb6, double sgn, long long N, double* out,
                    double* cR, double* W, hipStream_t s);
void launch_a8_div(double* Tg, const double* rec, const PairEntry* pairs, const TargetEntry* targets, int ntargets,
                   const double* fill, const double* dc2, const Geom& g, long long N, const JastrowPair& j0,
                   const JastrowPair& j1, hipStream_t s);
void launch_a8_reduce(const double* Tg, const double* Q4, const double* Q5, int bitc, const double* occR,
                      const double* fill, int nocc, double* V3a2, double* B2, long long N, hipStream_t s);
void launch_a8_q45(int prec, const double* Q4, const double* Q5, int bitc, void* Q4T, void* Q5T, int ngrids,
                   long long N, hipStream_t s);
// G-space rotation of ncomp-component sums (ncomp 1: scalar, 3: Cartesian vector) from src to dst grids
void launch_sym_rotate(double* Q, int ncomp, const SymRotJob* jobs, int njobs, const Geom& g, long long N,
                       hipStream_t s);

// kernel-gradient reduction of the occupied-set energies (rec: 1 channel E_x, 8 channels E_2x)
void launch_kgrad(int prec, const void* rec, int tc2, const PairEntry* pairs, int npairs, const double* kqcart,
                  const double* fill, const Geom& g, long long N, const JastrowPair& j0, const JastrowPair& j1, double w0,
                  int nbins, double width, double* bins_v, double* bins_u, double* sums, hipStream_t s);

// assembly terms (FP64)
void launch_kinetic(const double* phi, const int* gindex, const OrbEntry* jobs, int njobs, int max_npw,
                    const double* kcart, const Geom& g, double* out, hipStream_t s);
void launch_hartree(const double* nG, double* V, const Geom& g, long long N, hipStream_t s);
void launch_tc2h_V(const double* chiR, const double* gradR, const double* fill, int nocc, double* pgp,
                   const Geom& g, long long N, hipStream_t s);
// pgp0 / pgp1: the <q|nabla|q> sums (G space) of spin 0 / 1; pf = 1 / spin_factor
void launch_tc2h_combine(const double* nG, const double* pgp0, const double* pgp1, double pf, double* V,
                         const Geom& g, long long N, const JastrowPair& j0, const JastrowPair& j1, hipStream_t s);
void launch_tc2h_apply(double* orb, const double* grad, const double* V2, const double* dnu, int nb, long long N,
                       hipStream_t s);
void launch_axpby(const double* X, const double* Y, double* out, double a, double b, long long N, hipStream_t s);
// W += C over N complex points (W FP64, C FP32)
void launch_add_into(double* W, const float* C, long long N, hipStream_t s);

// divergence terms at zero-weight k-points (kernels_zw.hip); the states projected on are
// addressed by DivEntry like occupied orbitals
void launch_divcorr_kv(const double* phi, const double* occ_right, const double* occ_left, const DivEntry* jobs,
                       int njobs, int max_states, const double* fill, const double* dc, const double* kv,
                       const int* gindex, const int* gidx_off, const double* kcart, const Geom& g, double* out,
                       hipStream_t s);
void launch_b6_tG_kv(const double* nG, double* tG, const double skv[3], const Geom& g, long long N,
                     const JastrowPair& j0, const JastrowPair& j1, hipStream_t s);
void launch_zw_b3(const double* phi, const double* occ_l, const DivEntry* jobs, int njobs, int max_states,
                  const double* fill, const double* chiR, const double* occR, const double* occ_r, const double* dnu,
                  const double* appR, long long N, double fb3, const double kv[3], double* cA, double* cQ,
                  double* qsum, double* out, hipStream_t s);
void launch_zw_b4(const DivEntry* jobs, int njobs, int max_states, const double* fill, const double* chiR,
                  const double* occ_r, const double* qsum, long long N, double fb4, const double kv[3], double* cA,
                  double* out, hipStream_t s);
void launch_duk_table(double* table, const double* kqcart, int row0, int nrows, const Geom& g, const JastrowPair& j,
                      const double kv[3], hipStream_t s);
void launch_zw_a4b5_apply(const double* phi, const double* occ_r, const double* occ_l, const DivEntry* jobs, int njobs,
                          int max_states, const double* fill, const double* Wa, const double* Wb, const int* gi,
                          long long N, double fa4, double fb5, double kv2nq, double* out, hipStream_t s);

// block Davidson algebra (FP64, batched over k; vector i of k at base + 2 (off[k] + i npw[k]))
void launch_dv_precond(const double* V, const long long* voff, const int* first, const int* cnt, int max_cnt,
                       const double* H0, const double* H1, const double* H2, const double* kG2, const long long* goff,
                       const double* eig, const int* band0, double* T, const long long* toff, const int* npw, int nk,
                       hipStream_t s);
// B_bidx -= sum_i <D_i|B_bidx> A_i over i < nA (D = A when null: Gram_Schmidt; D = the
// other side's set: Gram_Schmidt_biortho), then nrm = ||B_bidx||
void launch_dv_project(const double* A, const long long* aoff, const int* nA, int max_nA, double* B,
                       const long long* boff, const int* bidx, const int* npw, int max_npw, int nk, double* c,
                       double* nrm, hipStream_t s, const double* D = nullptr);
void launch_dv_store(const double* B, const long long* boff, const int* bidx, const double* nrm, double* O,
                     const long long* ooff, const int* odst, const int* npw, int max_npw, int nk, hipStream_t s);
void launch_dv_copy(const double* S, const long long* soff, const int* sfirst, double* D, const long long* doff,
                    const int* dfirst, const int* cnt, int max_cnt, const int* npw, int max_npw, int nk, hipStream_t s);
void launch_dv_gram(const double* A, const long long* aoff, const int* nA, const double* B0, const double* B1,
                    const double* B2, int nB, const int* npw, int nk, double* G, int gs, hipStream_t s);
void launch_dv_rotate(const double* A, const long long* aoff, const int* nA, const double* C, int gs, double* O,
                      const long long* ooff, const int* nout, const int* npw, int max_npw, int nk, hipStream_t s);
size_t dv_eig_work_bytes(int gs, int nk);
bool dv_eig_fits_lds(int gs);
void launch_dv_eig(const double* Hs, int nH, const int* dim, int gs, int nk, int want_left, int sort_overlap,
                   double* w, double* X, double* XL, double* gwork, const int* nb, const int* band0, double* en,
                   int* status, hipStream_t s);
void launch_dv_dot(const double* X, const long long* xoff, const int* xidx, const double* Y, const long long* yoff,
                   const int* yidx, const int* npw, int nk, double* c, hipStream_t s);
void launch_dv_cdiv_store(const double* B, const long long* boff, const int* bidx, const double* c, double* O,
                          const long long* ooff, const int* odst, const int* npw, int max_npw, int nk, hipStream_t s);
void launch_dv_binorm(double* R, double* L, const long long* off, const int* nb, int max_nb, const int* band0,
                      const int* npw, int nk, double* out, hipStream_t s);

}  // namespace tcx
