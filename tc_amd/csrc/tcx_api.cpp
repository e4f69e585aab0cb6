// tcx_api.cpp -- C ABI (include/tcx.h) over the gfx950 kernels.
//
// The context keeps, on one device and one HIP stream:
//   * the cell/grid geometry, k-points and plane-wave index lists;
//   * the occupied orbitals in G space (complex128) and, cached per
//     precision, in real space (one FFT grid per orbital);
//   * workspaces: applied orbitals in real space, the FP64 accumulation grids
//     W_kj, and the pair buffer whose size bounds the pairs in flight.
// Each term call builds (and caches) a launch plan: the (k,j | q,i) pairs are
// grouped into launches of "k-chunk x q-chunk" so that every launch owns a
// set of accumulation targets (k,j) exclusively and W is updated without
// atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "tcx_internal.hpp"
#include "tcx_terms.hpp"

using namespace tcx;

namespace tcx {

thread_local bool g_capturing = false;

namespace {
std::mutex g_graph_mu;
int g_live_graphs = 0;
std::vector<void*> g_retired;
}  // namespace

void graph_created()
{
    std::lock_guard<std::mutex> lk(g_graph_mu);
    ++g_live_graphs;
}

void graph_destroyed()
{
    std::lock_guard<std::mutex> lk(g_graph_mu);
    if (--g_live_graphs > 0) return;
    if (!g_retired.empty()) (void)hipDeviceSynchronize();
    for (void* p : g_retired) (void)hipFree(p);
    g_retired.clear();
}

void DBuf::ensure(size_t bytes)
{
    if (bytes <= cap) return;
    if (g_capturing)
        throw Error(TCX_ERR_STATE, "graph capture: a workspace would grow; run the same call once before capturing");
    release();
    if (bytes == 0) return;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        (void)hipGetLastError();
        throw Error(TCX_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    cap = bytes;
}

void DBuf::release()
{
    if (p) {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        if (g_live_graphs > 0) {
            g_retired.push_back(p);  // a captured graph may still address it
        } else {
            (void)hipDeviceSynchronize();  // kernels queued on any stream may still read it
            (void)hipFree(p);
        }
    }
    p = nullptr;
    cap = 0;
}

void Context::require(bool cond, int code, const char* what)
{
    if (!cond) throw Error(code, what);
}

void Context::timed(int cls, long pairs, const std::function<void()>& launch)
{
    if (!timing_on || g_capturing) {
        launch();
        return;
    }
    hipEvent_t a, b;
    TCX_HIP_CHECK(hipEventCreate(&a));
    TCX_HIP_CHECK(hipEventCreate(&b));
    TCX_HIP_CHECK(hipEventRecord(a, stream));
    launch();
    TCX_HIP_CHECK(hipEventRecord(b, stream));
    pending.push_back({cls, pairs, a, b});
}

void Context::collect_timing()
{
    if (pending.empty()) return;
    TCX_HIP_CHECK(hipStreamSynchronize(stream));
    for (auto& ev : pending) {
        float ms = 0.f;
        TCX_HIP_CHECK(hipEventElapsedTime(&ms, ev.a, ev.b));
        t_ms[ev.cls] += ms;
        t_launches[ev.cls] += 1;
        t_pairs[ev.cls] += ev.pairs;
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    pending.clear();
}

// ------------------------------------------------------------------ state setters
void Context::set_cell(const double recip_rows[9], double vol, const int grid[3], int ns, int meth)
{
    require(recip_rows && grid, TCX_ERR_ARG, "set_cell: null pointer");
    require(vol > 1e-8, TCX_ERR_ARG, "set_cell: non-positive volume");
    require(grid[0] > 0 && grid[1] > 0 && grid[2] > 0, TCX_ERR_ARG, "set_cell: bad grid");
    require(meth >= TCX_METHOD_HF && meth <= TCX_METHOD_BITC, TCX_ERR_ARG, "set_cell: bad method");
    require(ns == 1 || ns == 2, TCX_ERR_ARG, "set_cell: num_independent_spins must be 1 or 2");
    require(grid[0] == grid[1], TCX_ERR_UNSUPPORTED, "set_cell: grid must have n1 == n2");
    require(size_supported(grid[0]) && size_supported(grid[2]), TCX_ERR_UNSUPPORTED,
            "set_cell: grid size not compiled (supported: 6 8 10 12 16 18 20 24 30 32 36 40 48)");
    geo.n1 = grid[0];
    geo.n2 = grid[1];
    geo.n3 = grid[2];
    for (int i = 0; i < 9; ++i) geo.b[i] = recip_rows[i];
    N = (long long)grid[0] * grid[1] * grid[2];
    volume = vol;
    if (ns != nspin) k_set = false;  // the k-blocks repeat per spin: set the k-points again
    nspin = ns;
    method = meth;
    cell_set = true;
    if (poly_deg > 0) poly_dirty = true;  // the J3 mesh follows the grid
    sym_set = false;
    basis_set = false;
    states_set = false;
    invalidate_states();
}

void Context::set_kpoints(int nk_, const double* kf, int numk)
{
    require(cell_set, TCX_ERR_STATE, "set_kpoints: call tcx_set_cell first");
    require(nk_ > 0 && kf, TCX_ERR_ARG, "set_kpoints: bad k list");
    require(numk > 0, TCX_ERR_ARG, "set_kpoints: num_kpoints must be positive");
    nk_phys = nk_;
    nk = nspin * nk_;
    num_kpoints = numk;
    kfrac.clear();
    for (int s = 0; s < nspin; ++s) kfrac.insert(kfrac.end(), kf, kf + 3 * nk_);  // the k list per spin
    kcart.assign(3 * nk, 0.0);
    for (int k = 0; k < nk; ++k)
        for (int d = 0; d < 3; ++d)
            kcart[3 * k + d] = geo.b[d] * kfrac[3 * k] + geo.b[3 + d] * kfrac[3 * k + 1] + geo.b[6 + d] * kfrac[3 * k + 2];
    d_kcart.ensure(kcart.size() * sizeof(double));
    TCX_HIP_CHECK(hipMemcpy(d_kcart.p, kcart.data(), kcart.size() * sizeof(double), hipMemcpyHostToDevice));
    k_set = true;
    basis_set = false;
    states_set = false;
    div_enabled = false;
    sym_set = false;
    // no occupied orbitals until tcx_set_states (the one-body term needs none)
    nocc.assign(nk, 0);
    occ_first.assign(nk + 1, 0);
    occ_coef.assign(nk + 1, 0);
    n_occ = 0;
    fill.clear();
    invalidate_states();
}

// The setters below drop the state they replace on entry (so a rejected call
// leaves the context without it: later applies return TCX_ERR_STATE instead of
// reading buffers sized for the old counts), validate into locals, and commit
// only when every check has passed.
void Context::set_basis(const int* npw_, const int* gindex)
{
    require(k_set, TCX_ERR_STATE, "set_basis: call tcx_set_kpoints first");
    basis_set = false;
    states_set = false;
    invalidate_states();
    require(npw_ && gindex, TCX_ERR_ARG, "set_basis: null pointer");
    // the basis of the nk_phys k-points, repeated for the second spin
    std::vector<int> n_(nk);
    for (int k = 0; k < nk; ++k) n_[k] = npw_[k % nk_phys];
    std::vector<long long> off(nk + 1, 0);
    for (int k = 0; k < nk; ++k) {
        require(n_[k] > 0, TCX_ERR_ARG, "set_basis: npw must be positive");
        off[k + 1] = off[k] + n_[k];
    }
    const long long nphys = off[nk_phys];
    for (long long l = 0; l < nphys; ++l)
        require(gindex[l] >= 0 && gindex[l] < N, TCX_ERR_ARG, "set_basis: gindex outside the FFT grid");
    std::vector<int> gi;
    for (int s = 0; s < nspin; ++s) gi.insert(gi.end(), gindex, gindex + nphys);
    d_gindex.ensure(gi.size() * sizeof(int));
    TCX_HIP_CHECK(hipMemcpy(d_gindex.p, gi.data(), gi.size() * sizeof(int), hipMemcpyHostToDevice));
    npw = std::move(n_);
    pw_off = std::move(off);
    max_npw = *std::max_element(npw.begin(), npw.end());
    gindex_h = std::move(gi);
    basis_set = true;
}

void Context::invalidate_local()
{
    for (LocalCache* c : {&c_a8, &c_hartree, &c_tc2h, &c_tc3a1, &c_tc3b1, &a_a8, &a_hartree, &a_tc2h, &a_tc3a1,
                          &a_tc3b1}) {
        c->valid = false;
        c->prec = -1;
    }
}

void Context::hermitian_conj()
{
    // the density of the unswapped orbitals stays (built first if it was not)
    if (!conj_mode) ensure_density();
    negate_u();
    std::swap(d_phi_occ, d_chi_occ);
    for (int p = 0; p < 2; ++p) {
        std::swap(d_occR[p], d_chiR[p]);
        std::swap(occR_valid[p], chiR_valid[p]);
        std::swap(d_gradR[p], a_gradR[p]);
        std::swap(gradR_valid[p], a_gradR_valid[p]);
    }
    std::swap(c_a8, a_a8);
    std::swap(c_hartree, a_hartree);
    std::swap(c_tc2h, a_tc2h);
    std::swap(c_tc3a1, a_tc3a1);
    std::swap(c_tc3b1, a_tc3b1);
    std::swap(d_Q4T, a_Q4T);
    std::swap(d_Q5T, a_Q5T);
    conj_mode = !conj_mode;
}

void Context::invalidate_states()
{
    clear_zero_weight();  // its states live in the orbital arrays that a setter replaces
    invalidate_local();
    occR_valid[0] = occR_valid[1] = false;
    chiR_valid[0] = chiR_valid[1] = false;
    gradR_valid[0] = gradR_valid[1] = false;
    a_gradR_valid[0] = a_gradR_valid[1] = false;
    density_valid = false;
    sym_jobs_valid = false;
    plans.clear();
    pairsets.clear();  // pair lists index the occupied orbitals and k-points
}

void Context::set_states(const int* nocc_, const double* filling, const double* phi_occ, const double* chi_occ)
{
    require(basis_set, TCX_ERR_STATE, "set_states: call tcx_set_basis first");
    states_set = false;
    invalidate_states();
    require(nocc_ && filling && phi_occ, TCX_ERR_ARG, "set_states: null pointer");
    require(method == TCX_METHOD_BITC || chi_occ == nullptr, TCX_ERR_ARG,
            "set_states: left orbitals are only meaningful for BITC");
    require(method != TCX_METHOD_BITC || chi_occ != nullptr, TCX_ERR_ARG, "set_states: BITC needs left orbitals");
    std::vector<int> no(nocc_, nocc_ + nk);
    std::vector<long long> first(nk + 1, 0), coef(nk + 1, 0);
    for (int k = 0; k < nk; ++k) {
        require(no[k] >= 0 && no[k] <= 512, TCX_ERR_ARG, "set_states: nocc must be in [0, 512]");
        first[k + 1] = first[k] + no[k];
        coef[k + 1] = coef[k] + (long long)no[k] * npw[k];
    }
    const int nall = (int)first[nk];
    std::vector<double> f_(filling, filling + nall);
    for (double f : f_) require(std::isfinite(f) && f >= 0.0, TCX_ERR_ARG, "set_states: filling must be >= 0");
    const size_t cb = (size_t)coef[nk] * 16;
    for (size_t i = 0; i < (size_t)coef[nk] * 2; ++i)
        require(std::isfinite(phi_occ[i]), TCX_ERR_ARG, "set_states: non-finite orbital coefficient");
    if (chi_occ)
        for (size_t i = 0; i < (size_t)coef[nk] * 2; ++i)
            require(std::isfinite(chi_occ[i]), TCX_ERR_ARG, "set_states: non-finite left orbital coefficient");
    nocc = std::move(no);
    occ_first = std::move(first);
    occ_coef = std::move(coef);
    n_occ = nall;
    fill = std::move(f_);
    d_phi_occ.ensure(std::max<size_t>(cb, 16));
    TCX_HIP_CHECK(hipMemcpy(d_phi_occ.p, phi_occ, cb, hipMemcpyHostToDevice));
    if (chi_occ) {
        d_chi_occ.ensure(std::max<size_t>(cb, 16));
        TCX_HIP_CHECK(hipMemcpy(d_chi_occ.p, chi_occ, cb, hipMemcpyHostToDevice));
    }
    upload(d_fill, fill.data(), fill.size() * sizeof(double));
    std::vector<OrbEntry> jobs, gjobs;
    for (int q = 0; q < nk; ++q)
        for (int i = 0; i < nocc[q]; ++i) {
            const OrbEntry e{occ_coef[q] + (long long)i * npw[q], npw[q], (int)pw_off[q], q, -1};
            jobs.push_back(e);
            for (int d = 0; d < 3; ++d) gjobs.push_back({e.coef_off, e.npw, e.gidx_off, q, d});
        }
    upload(d_occjobs, jobs.data(), jobs.size() * sizeof(OrbEntry));
    upload(d_occgradjobs, gjobs.data(), gjobs.size() * sizeof(OrbEntry));
    ++states_version;
    states_set = true;
}

void Context::set_jastrow(const double A_[4], const double F_[4])
{
    jastrow_set = false;
    plans.clear();
    invalidate_local();
    require(A_ && F_, TCX_ERR_ARG, "set_jastrow: null pointer");
    bool zero = true;
    for (int i = 0; i < 4; ++i) {
        require(std::isfinite(A_[i]) && std::isfinite(F_[i]), TCX_ERR_ARG, "set_jastrow: non-finite parameter");
        if (std::fabs(A_[i]) > 1e-8) zero = false;
    }
    if (!zero)
        for (int i = 0; i < 4; ++i)
            require(std::fabs(F_[i]) > 1e-8, TCX_ERR_ARG, "set_jastrow: F must be non-zero");
    for (int i = 0; i < 4; ++i) {
        A[i] = A_[i];
        F[i] = F_[i];
    }
    jastrow_zero = zero;
    jastrow_set = true;
    if (poly_deg > 0) poly_dirty = true;  // the RPA+polynomial parameters follow A and F
}

void Context::set_div(int enabled, double alpha_, const double* svaux)
{
    require(k_set, TCX_ERR_STATE, "set_div_correction: call tcx_set_kpoints first");
    invalidate_local();
    div_enabled = false;
    if (enabled) {
        require(svaux != nullptr, TCX_ERR_ARG, "set_div_correction: sum_of_vaux is null");
        require(std::isfinite(alpha_) && alpha_ > 0.0, TCX_ERR_ARG, "set_div_correction: alpha must be positive");
        for (int k = 0; k < nk_phys; ++k)
            require(std::isfinite(svaux[k]), TCX_ERR_ARG, "set_div_correction: non-finite sum_of_vaux");
        alpha = alpha_;
        sum_vaux.resize(nk);  // per k-block: the value of its k-point
        for (int k = 0; k < nk; ++k) sum_vaux[k] = svaux[k % nk_phys];
        std::vector<double> dc(nk);
        for (int k = 0; k < nk; ++k) dc[k] = div_factor(k);
        upload(d_dc, dc.data(), dc.size() * sizeof(double));
    }
    div_enabled = enabled != 0;
}

// k-mesh symmetry: each k is built from the orbitals of k_src[k] by the TC++
// rotation U (symmetry.cpp:17-33), translation tau and optional time reversal,
// as in Kpoints::set_kvectors_by_symmetry (kpoints.cpp:34-110) and
// PlaneWaveBasis (plane_wave_basis.cpp:115-140).
void Context::set_symmetry(const int* k_src, const int* rot, const double* trans, const int* trev)
{
    require(k_set, TCX_ERR_STATE, "set_kpoint_symmetry: call tcx_set_kpoints first");
    sym_set = false;
    invalidate_states();
    if (k_src == nullptr) return;  // identity symmetry
    require_one_spin("set_kpoint_symmetry");
    require(rot && trans && trev, TCX_ERR_ARG, "set_kpoint_symmetry: null pointer");
    // lattice rows a_i = 2 pi (B^-1)^T, B = reciprocal rows
    const double* B = geo.b;
    const double det = B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                       B[2] * (B[3] * B[7] - B[4] * B[6]);
    double Binv[9];  // B^-1
    Binv[0] = (B[4] * B[8] - B[5] * B[7]) / det;
    Binv[1] = (B[2] * B[7] - B[1] * B[8]) / det;
    Binv[2] = (B[1] * B[5] - B[2] * B[4]) / det;
    Binv[3] = (B[5] * B[6] - B[3] * B[8]) / det;
    Binv[4] = (B[0] * B[8] - B[2] * B[6]) / det;
    Binv[5] = (B[2] * B[3] - B[0] * B[5]) / det;
    Binv[6] = (B[3] * B[7] - B[4] * B[6]) / det;
    Binv[7] = (B[1] * B[6] - B[0] * B[7]) / det;
    Binv[8] = (B[0] * B[4] - B[1] * B[3]) / det;
    double Alat[9];  // rows a_i
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Alat[3 * i + j] = 2.0 * kPiHost * Binv[3 * j + i];
    const int n[3] = {geo.n1, geo.n2, geo.n3};
    std::vector<SymRotJob> ops(nk);
    std::vector<char> seen((size_t)N);
    for (int k = 0; k < nk; ++k) {
        const int s = k_src[k];
        require(s >= 0 && s < nk, TCX_ERR_ARG, "set_kpoint_symmetry: k_src out of range");
        require(k_src[s] == s, TCX_ERR_ARG, "set_kpoint_symmetry: k_src must name an irreducible k (k_src[q] == q)");
        const int* U = rot + 9 * k;
        const double* t = trans + 3 * k;
        require(trev[k] == 0 || trev[k] == 1, TCX_ERR_ARG, "set_kpoint_symmetry: trev must be 0 or 1");
        const long dU = (long)U[0] * (U[4] * U[8] - U[5] * U[7]) - (long)U[1] * (U[3] * U[8] - U[5] * U[6]) +
                        (long)U[2] * (U[3] * U[7] - U[4] * U[6]);
        require(dU == 1 || dU == -1, TCX_ERR_ARG, "set_kpoint_symmetry: rotation is not unimodular");
        if (s == k) {
            bool ident = trev[k] == 0;
            for (int i = 0; i < 9; ++i) ident = ident && U[i] == (i % 4 == 0 ? 1 : 0);
            for (int d = 0; d < 3; ++d) ident = ident && std::fabs(t[d]) < 1e-12;
            require(ident, TCX_ERR_ARG,
                    "set_kpoint_symmetry: an irreducible k needs the identity (no translation, no time reversal)");
        }
        for (int d = 0; d < 3; ++d) {
            const double tn = t[d] * n[d];
            require(std::isfinite(tn) && std::fabs(tn - std::round(tn)) < 1e-6, TCX_ERR_ARG,
                    "set_kpoint_symmetry: fractional translation not commensurate with the FFT grid");
        }
        SymRotJob& J = ops[k];
        J.dst = J.src = 0;
        J.trev = trev[k];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J.ud[3 * i + j] = U[3 * j + i];  // U^+ = U^T
        for (int d = 0; d < 3; ++d) J.tau[d] = t[d];
        // k = s U^+ k_src modulo reciprocal-lattice vectors (kpoints.cpp:10-26)
        const int sg = J.trev ? -1 : 1;
        double d2 = 0.0;
        for (int i = 0; i < 3; ++i) {
            double v = 0.0;
            for (int j = 0; j < 3; ++j) v += J.ud[3 * i + j] * kfrac[3 * s + j];
            double dd = kfrac[3 * k + i] - sg * v;
            dd -= std::round(dd);
            d2 += dd * dd;
        }
        require(d2 < 1e-8, TCX_ERR_ARG, "set_kpoint_symmetry: k is not the rotated k_src");
        // U^+ in Cartesian coordinates: B^T U^+ A / (2 pi) (tc3a2a4b2b5.cpp:677-679)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double a = 0.0;
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) a += B[3 * p + i] * J.ud[3 * p + q] * Alat[3 * q + j];
                J.udc[3 * i + j] = a / (2.0 * kPiHost);
            }
        if (s == k) continue;
        // the wrapped map G -> s U^+ G must be a bijection of the FFT grid
        std::fill(seen.begin(), seen.end(), 0);
        for (long long r = 0; r < N; ++r) {
            const int i1 = (int)(r % n[0]), i2 = (int)((r / n[0]) % n[1]), i3 = (int)(r / ((long long)n[0] * n[1]));
            const int G[3] = {i1 <= n[0] / 2 ? i1 : i1 - n[0], i2 <= n[1] / 2 ? i2 : i2 - n[1],
                              i3 <= n[2] / 2 ? i3 : i3 - n[2]};
            long long rd = 0, mul = 1;
            for (int i = 0; i < 3; ++i) {
                int v = sg * (J.ud[3 * i] * G[0] + J.ud[3 * i + 1] * G[1] + J.ud[3 * i + 2] * G[2]);
                v %= n[i];
                if (v < 0) v += n[i];
                rd += mul * v;
                mul *= n[i];
            }
            require(!seen[rd], TCX_ERR_ARG, "set_kpoint_symmetry: rotation does not map the FFT grid onto itself");
            seen[rd] = 1;
        }
    }
    sym_src.assign(k_src, k_src + nk);
    sym_ops = std::move(ops);
    sym_set = true;
}

// The real-space density the terms read (BlochStates::density_): formed from
// the occupied orbitals on first use, or replaced by the caller (density mixing).
void Context::get_density(double* out)
{
    require(states_set, TCX_ERR_STATE, "get_density: call tcx_set_states first");
    require(out != nullptr, TCX_ERR_ARG, "get_density: null pointer");
    ensure_density();
    TCX_HIP_CHECK(hipStreamSynchronize(stream));
    TCX_HIP_CHECK(hipMemcpy(out, d_density.p, (size_t)2 * N * 16, hipMemcpyDeviceToHost));
}

void Context::set_density(const double* in)
{
    require(states_set, TCX_ERR_STATE, "set_density: call tcx_set_states first");
    require(in != nullptr, TCX_ERR_ARG, "set_density: null pointer");
    for (long long i = 0; i < (nspin == 1 ? 2 : 4) * N; ++i)
        require(std::isfinite(in[i]), TCX_ERR_ARG, "set_density: non-finite value");
    invalidate_local();
    TCX_HIP_CHECK(hipStreamSynchronize(stream));
    d_density.ensure((size_t)2 * N * 16);
    // one independent spin: density_[1] = density_[0] (bloch_states_scfloop_filling_density.cpp:355-358)
    const double* in1 = nspin == 1 ? in : in + 2 * N;
    TCX_HIP_CHECK(hipMemcpy(d_density.p, in, (size_t)N * 16, hipMemcpyHostToDevice));
    TCX_HIP_CHECK(hipMemcpy((char*)d_density.p + (size_t)N * 16, in1, (size_t)N * 16, hipMemcpyHostToDevice));
    density_valid = true;
}

// ------------------------------------------------------------------ real-space grids
// Transforms ncoef-addressed orbitals (complex128, G space) to real space in
// precision prec: FP64 scatter + FP64 inverse FFT, then rounding to float for
// FP32.  Jobs with dir >= 0 multiply by (k+G)_dir first (gradient components).
void Context::orbitals_to_real(const double* coef, const OrbEntry* d_jobs, int n, int prec, void* out)
{
    if (n == 0) return;
    const size_t g64 = (size_t)N * 16;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(n, ((size_t)1 << 30) / g64));
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int nc = std::min(chunk, n - c0);
        double* dst64;
        if (prec == TCX_PREC_FP64) {
            dst64 = (double*)out + (size_t)c0 * N * 2;
        } else {
            d_tmp64.ensure(g64 * chunk);
            dst64 = (double*)d_tmp64.p;
        }
        launch_zero(dst64, g64 * nc, stream);
        launch_scatter(coef, (const int*)d_gindex.p, d_jobs + c0, nc, max_npw, dst64, N, (const double*)d_kcart.p,
                       geo, stream);
        launch_fft3d(TCX_PREC_FP64, +1, dst64, nc, geo, 1.0, stream);
        if (prec == TCX_PREC_FP32)
            launch_convert_d2f(dst64, (float*)out + (size_t)c0 * N * 2, (long long)nc * N, stream);
    }
}

void upload(DBuf& buf, const void* src, size_t bytes)
{
    if (g_capturing)
        throw Error(TCX_ERR_STATE, "graph capture: a launch plan would be built; run the same call once before capturing");
    buf.ensure(std::max<size_t>(bytes, 16));
    if (bytes) TCX_HIP_CHECK(hipMemcpy(buf.p, src, bytes, hipMemcpyHostToDevice));
}

void Context::ensure_occR(int prec)
{
    if (occR_valid[prec]) return;
    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8;
    // occupied orbitals, then the states of zero-weight k-points (divergence terms only)
    const int ng = n_occ + n_zw;
    d_occR[prec].ensure(std::max<size_t>((size_t)ng * N * es, 16));
    orbitals_to_real((const double*)d_phi_occ.p, (const OrbEntry*)d_occjobs.p, ng, prec, d_occR[prec].p);
    occR_valid[prec] = true;
}

void Context::ensure_chiR(int prec)
{
    if (chiR_valid[prec]) return;
    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8;
    const int ng = n_occ + n_zw;
    d_chiR[prec].ensure(std::max<size_t>((size_t)ng * N * es, 16));
    orbitals_to_real((const double*)d_chi_occ.p, (const OrbEntry*)d_occjobs.p, ng, prec, d_chiR[prec].p);
    chiR_valid[prec] = true;
}

void Context::ensure_gradR(int prec)
{
    if (gradR_valid[prec]) return;
    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8;
    d_gradR[prec].ensure(std::max<size_t>((size_t)3 * n_occ * N * es, 16));
    orbitals_to_real((const double*)d_phi_occ.p, (const OrbEntry*)d_occgradjobs.p, 3 * n_occ, prec, d_gradR[prec].p);
    gradR_valid[prec] = true;
}

// ------------------------------------------------------------------ launch plans
// Launch shape: every launch pairs a block of consecutive targets (k,j) with a
// block of consecutive sources (q,i).  Targets own their accumulation grid
// within a launch, so launches of one apply are serialised on the stream and
// W needs no atomics.  Auto shape: enough targets that pass C issues
// kAutoTargetWaves wavefronts, then as many sources as kAutoBufferGrids pair
// grids allow (channels = grids per pair record: 1 for hf2x, 4 for tc2x).
void Context::resolve_shape(int term, long& tc, long& sc, long ntargets, long nsources) const
{
    constexpr long kAutoTargetWaves = 8192, kAutoBufferGrids = 8192;
    const int channels = term == 1 ? 4 : 1;
    const long waves_per_target = ((long)geo.n1 * geo.n2 + 31) / 32;
    tc = shape_targets[term] > 0 ? shape_targets[term] : (kAutoTargetWaves + waves_per_target - 1) / waves_per_target;
    tc = std::max<long>(1, std::min(tc, std::max<long>(1, ntargets)));
    sc = shape_sources[term] > 0 ? shape_sources[term] : std::max<long>(1, kAutoBufferGrids / channels / tc);
    sc = std::max<long>(1, std::min(sc, std::max<long>(1, nsources)));
    while (tc * sc > 65535) {
        if (sc > 1)
            sc = std::max<long>(1, 65535 / tc);
        else
            tc = 65535;
    }
}

void build_pairs(PairSet& S, const std::vector<PairItem>& tg, const std::vector<PairItem>& src, long tc, long sc,
                 const std::vector<double>& kcart, bool same_k)
{
    for (size_t t0 = 0; t0 < tg.size(); t0 += tc) {
        const size_t t1 = std::min(tg.size(), t0 + (size_t)tc);
        const int kf = tg[t0].k, kl = tg[t1 - 1].k;
        for (size_t s0 = 0; s0 < src.size(); s0 += sc) {
            const size_t s1 = std::min(src.size(), s0 + (size_t)sc);
            const int qf = src[s0].k, ql = src[s1 - 1].k;
            if (same_k && (ql < kf || qf > kl)) continue;
            Launch L;
            L.pair0 = (int)S.pairs.size();
            L.target0 = (int)S.targets.size();
            L.kq0 = (int)S.kq.size() / 3;
            L.nkq = (kl - kf + 1) * (ql - qf + 1);
            for (int k = kf; k <= kl; ++k)
                for (int q = qf; q <= ql; ++q) {
                    for (int d = 0; d < 3; ++d) S.kq.push_back(kcart[3 * k + d] - kcart[3 * q + d]);
                    S.kq_kq.push_back(k);
                    S.kq_kq.push_back(q);
                }
            int local = 0;
            for (size_t t = t0; t < t1; ++t) {
                TargetEntry te{tg[t].idx, local, 0, tg[t].k};
                for (size_t s = s0; s < s1; ++s) {
                    if (same_k && src[s].k != tg[t].k) continue;
                    const int kq = L.kq0 + (tg[t].k - kf) * (ql - qf + 1) + (src[s].k - qf);
                    S.pairs.push_back({tg[t].idx, src[s].idx, kq, 0});
                    ++local;
                    ++te.npair;
                }
                S.targets.push_back(te);
            }
            L.npairs = local;
            L.ntargets = (int)(t1 - t0);
            if (local == 0) {
                S.targets.resize(L.target0);
                S.kq.resize(3 * (size_t)L.kq0);
                S.kq_kq.resize(2 * (size_t)L.kq0);
                continue;
            }
            S.launches.push_back(L);
        }
    }
    for (const auto& L : S.launches) S.max_launch_pairs = std::max(S.max_launch_pairs, L.npairs);
    upload(S.d_pairs, S.pairs.data(), S.pairs.size() * sizeof(PairEntry));
    upload(S.d_targets, S.targets.data(), S.targets.size() * sizeof(TargetEntry));
    upload(S.d_kq, S.kq.data(), S.kq.size() * sizeof(double));
}

std::vector<PairItem> Context::applied_items(const std::vector<int>& nb) const
{
    std::vector<PairItem> v;
    int idx = 0;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nb[k]; ++j) v.push_back({k, idx++});
    return v;
}

std::vector<PairItem> Context::occupied_items(bool irreducible_only) const
{
    std::vector<PairItem> v;
    for (int q = 0; q < nk; ++q) {
        if (irreducible_only && sym_set && sym_src[q] != q) continue;
        for (int i = 0; i < nocc[q]; ++i) v.push_back({q, (int)occ_first[q] + i});
    }
    return v;
}

const Plan& Context::get_plan(const std::vector<int>& nb, int term)
{
    long nt = 0;
    for (int k = 0; k < nk; ++k) nt += nb[k];
    long tc, sc;
    resolve_shape(term, tc, sc, nt, n_occ);
    auto key = std::make_tuple(nb, tc, sc);
    auto it = plans.find(key);
    if (it != plans.end()) return it->second;
    Plan P;
    P.app_first.assign(nk + 1, 0);
    P.app_coef.assign(nk + 1, 0);
    for (int k = 0; k < nk; ++k) {
        P.app_first[k + 1] = P.app_first[k] + nb[k];
        P.app_coef[k + 1] = P.app_coef[k] + (long long)nb[k] * npw[k];
    }
    P.n_app = (int)P.app_first[nk];
    if (nspin == 1) {
        build_pairs(P, applied_items(nb), occupied_items(), tc, sc, kcart, false);
    } else {
        // the exchange pairs states of one spin only (calc_hamiltonian_hf2x.cpp:46, :90-134)
        const std::vector<PairItem> app = applied_items(nb), occ = occupied_items();
        for (int s = 0; s < nspin; ++s) {
            std::vector<PairItem> a, o;
            for (const PairItem& x : app)
                if (x.k / nk_phys == s) a.push_back(x);
            for (const PairItem& x : occ)
                if (x.k / nk_phys == s) o.push_back(x);
            const size_t l0 = P.launches.size();
            build_pairs(P, a, o, tc, sc, kcart, false);
            for (size_t l = l0; l < P.launches.size(); ++l) P.launches[l].spin = s;
        }
    }
    // one scatter/gather job and one divergence job per applied band, in band order
    std::vector<OrbEntry> appjobs, appgradjobs;
    std::vector<DivEntry> divjobs;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nb[k]; ++j) {
            const OrbEntry e{P.app_coef[k] + (long long)j * npw[k], npw[k], (int)pw_off[k], k, -1};
            appjobs.push_back(e);
            for (int d = 0; d < 3; ++d) appgradjobs.push_back({e.coef_off, e.npw, e.gidx_off, k, d});
            if (zw_set && zw_k[k])  // the states given for a zero-weight k
                divjobs.push_back({e.coef_off, zw_coef[k], npw[k], zw_n[k], zw_first[k], k});
            else
                divjobs.push_back({e.coef_off, occ_coef[k], npw[k], nocc[k], (int)occ_first[k], k});
        }
    upload(P.d_appjobs, appjobs.data(), appjobs.size() * sizeof(OrbEntry));
    upload(P.d_appgradjobs, appgradjobs.data(), appgradjobs.size() * sizeof(OrbEntry));
    upload(P.d_divjobs, divjobs.data(), divjobs.size() * sizeof(DivEntry));
    auto res = plans.emplace(key, std::move(P));
    return res.first->second;
}

void Context::require_one_spin(const char* what) const
{
    if (nspin == 1) return;
    throw Error(TCX_ERR_UNSUPPORTED, std::string(what) +
                                         ": two independent spins are implemented for kinetic, hf2h, hf2x, tc2x "
                                         "and the HF all / Davidson only");
}

// ------------------------------------------------------------------ term driver
void Context::check_apply_args(const int* nb, const double* phi, const double* out, int prec)
{
    require(states_set, TCX_ERR_STATE, "term: call tcx_set_states first");
    require(nb && phi && out, TCX_ERR_ARG, "term: null pointer");
    require(prec == TCX_PREC_FP64 || prec == TCX_PREC_FP32, TCX_ERR_ARG, "term: bad precision");
    for (int k = 0; k < nk; ++k) require(nb[k] >= 0, TCX_ERR_ARG, "term: negative band count");
}

ApplyIO Context::begin_apply(const int* nb, const double* phi, double* out, unsigned flags, const Plan& P)
{
    ApplyIO io;
    io.host = !(flags & TCX_FLAG_DEVICE_PTRS);
    require(!(io.host && g_capturing), TCX_ERR_STATE, "graph capture: term calls need TCX_FLAG_DEVICE_PTRS");
    io.ncoef = P.app_coef[nk];
    io.out_host = out;
    if (io.host) {
        for (long long i = 0; i < io.ncoef * 2; ++i)
            require(std::isfinite(phi[i]), TCX_ERR_ARG, "term: non-finite input coefficient");
        d_phi_app.ensure(std::max<size_t>(16, (size_t)io.ncoef * 16));
        d_out.ensure(std::max<size_t>(16, (size_t)io.ncoef * 16));
        TCX_HIP_CHECK(hipMemcpyAsync(d_phi_app.p, phi, (size_t)io.ncoef * 16, hipMemcpyHostToDevice, stream));
        launch_zero(d_out.p, (size_t)io.ncoef * 16, stream);
        io.phi = (const double*)d_phi_app.p;
        io.out = (double*)d_out.p;
    } else {
        io.phi = phi;
        io.out = out;
    }
    return io;
}

void Context::end_apply(const ApplyIO& io)
{
    if (g_capturing) return;  // the graph replays asynchronously; tcx_synchronize waits
    if (io.host && io.ncoef > 0) {
        std::vector<double> tmp((size_t)io.ncoef * 2);
        TCX_HIP_CHECK(hipMemcpyAsync(tmp.data(), io.out, (size_t)io.ncoef * 16, hipMemcpyDeviceToHost, stream));
        TCX_HIP_CHECK(hipStreamSynchronize(stream));
        for (size_t i = 0; i < tmp.size(); ++i) io.out_host[i] += tmp[i];
    }
    TCX_HIP_CHECK(hipStreamSynchronize(stream));
    collect_timing();
}

double Context::div_factor(int k) const
{
    // src/calc_hamiltonian_hf2x.cpp:153-163 (identical in tc2x.cpp:412-422)
    double dc = 1.0 / std::sqrt(kPiHost * alpha);
    dc -= sum_vaux[k] / (num_kpoints * volume);
    dc *= -1.0;
    if (nspin == 1) dc /= 2.0;
    return dc;
}

void Context::apply_divcorr(const Plan& P, const ApplyIO& io, const double* occ_left)
{
    if (!div_enabled || P.n_app == 0) return;
    launch_divcorr(io.phi, (const double*)d_phi_occ.p, occ_left, (const DivEntry*)P.d_divjobs.p, P.n_app,
                   (const double*)d_fill.p, (const double*)d_dc.p, io.out, stream);
}

// W (FP64 real-space accumulators, one grid per applied band) and Wc (FP32: the summed
// rounding errors of the band-pair sums added into W across launches, see add_partial)
double* Context::begin_W(int n_app)
{
    const size_t bytes = (size_t)n_app * N * 16;
    d_W.ensure(bytes);
    d_Wc.ensure(bytes / 2);
    launch_zero(d_W.p, bytes, stream);
    launch_zero(d_Wc.p, bytes / 2, stream);
    return (double*)d_W.p;
}

// W + Wc -> FFT_forward -> gather with factor
void Context::finish_W(const Plan& P, const ApplyIO& io, double factor)
{
    launch_add_into((double*)d_W.p, (const float*)d_Wc.p, (long long)P.n_app * N, stream);
    timed(3, 0, [&] { launch_fft3d(TCX_PREC_FP64, -1, d_W.p, P.n_app, geo, 1.0 / (double)N, stream); });
    launch_gather_add((const double*)d_W.p, (const int*)d_gindex.p, (const OrbEntry*)P.d_appjobs.p, P.n_app,
                      max_npw, io.out, N, factor, stream);
}

void Context::hf2x(const int* nb_, const double* phi, double* out, int prec, unsigned flags)
{
    check_apply_args(nb_, phi, out, prec);
    std::vector<int> nb(nb_, nb_ + nk);
    const Plan& P = get_plan(nb, 0);
    ApplyIO io = begin_apply(nb_, phi, out, flags, P);
    if (P.n_app > 0) {
        ensure_occR(prec);
        const size_t es = prec == TCX_PREC_FP64 ? 16 : 8;
        d_appR.ensure((size_t)P.n_app * N * es);
        timed(3, 0, [&] { orbitals_to_real(io.phi, (const OrbEntry*)P.d_appjobs.p, P.n_app, prec, d_appR.p); });
        begin_W(P.n_app);
        if (use_plane_first(prec)) {
            hf2x_plane_first(P);
            double factor = -4.0 * kPiHost / (num_kpoints * volume);  // two_body_factor, hf2x.cpp:29-30
            if (nspin == 1) factor /= 2.0;
            finish_W(P, io, factor);
            apply_divcorr(P, io, (const double*)d_phi_occ.p);
            end_apply(io);
            return;
        }
        hf2x_record_cols = (double)geo.n1 * geo.n2;
        d_pairbuf.ensure((size_t)std::max(1, P.max_launch_pairs) * N * es);
        int max_rows = 1;
        for (const Launch& L : P.launches) max_rows = std::max(max_rows, L.nkq);
        d_ktab.ensure((size_t)max_rows * N * (es / 2));  // one real per k-q row and grid point
        for (const Launch& L : P.launches) {
            const PairEntry* pairs = (const PairEntry*)P.d_pairs.p + L.pair0;
            const TargetEntry* targets = (const TargetEntry*)P.d_targets.p + L.target0;
            launch_coulomb_table(prec, d_ktab.p, (const double*)P.d_kq.p, L.kq0, L.nkq, geo, stream);
            timed(0, L.npairs, [&] {
                launch_hf2x_A(prec, geo, d_appR.p, d_occR[prec].p, pairs, L.npairs, d_pairbuf.p, stream);
            });
            timed(1, L.npairs, [&] {
                launch_hf2x_B(prec, geo, d_pairbuf.p, pairs, L.npairs, d_ktab.p, L.kq0, stream);
            });
            timed(2, L.npairs, [&] {
                launch_hf2x_C(prec, geo, d_pairbuf.p, d_occR[prec].p, pairs, targets, L.ntargets,
                              (const double*)d_fill.p, (double*)d_W.p, (float*)d_Wc.p, stream);
            });
        }
        // two_body_factor, src/calc_hamiltonian_hf2x.cpp:29-30
        double factor = -4.0 * kPiHost / (num_kpoints * volume);
        if (nspin == 1) factor /= 2.0;
        finish_W(P, io, factor);
        apply_divcorr(P, io, (const double*)d_phi_occ.p);
    }
    end_apply(io);
}

}  // namespace tcx

// ======================================================================= C ABI
namespace {

// an error inside a captured region ends the capture and drops the partial graph
void abort_capture(tcx_context* ctx)
{
    if (!g_capturing) return;
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(ctx->impl.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    g_capturing = false;
}

template <typename F>
int guarded(tcx_context* ctx, F&& f)
{
    if (!ctx) return TCX_ERR_ARG;
    try {
        TCX_HIP_CHECK(hipSetDevice(ctx->impl.device));
        f();
        ctx->impl.err.clear();
        return TCX_OK;
    } catch (const Error& e) {
        abort_capture(ctx);
        ctx->impl.err = e.what();
        ctx->impl.pending.clear();
        return e.code;
    } catch (const std::exception& e) {
        abort_capture(ctx);
        ctx->impl.err = e.what();
        return TCX_ERR_INTERNAL;
    } catch (...) {
        abort_capture(ctx);
        ctx->impl.err = "unknown error";
        return TCX_ERR_INTERNAL;
    }
}

}  // namespace

extern "C" {

int tcx_create(tcx_context** out, int device)
{
    if (!out) return TCX_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return TCX_ERR_HIP;
    if (device < 0 || device >= ndev) return TCX_ERR_ARG;
    tcx_context* c = new (std::nothrow) tcx_context();
    if (!c) return TCX_ERR_NOMEM;
    c->impl.device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->impl.stream) != hipSuccess) {
        delete c;
        return TCX_ERR_HIP;
    }
    // address of the no-allocation entry point's device array on this device (once per
    // device; a failure is reported by tcx_hf2x_noalloc, the other entry points do not use it)
    (void)tcx::scratch_setup(device);
    *out = c;
    return TCX_OK;
}

int tcx_destroy(tcx_context* ctx)
{
    if (!ctx) return TCX_ERR_ARG;
    (void)hipSetDevice(ctx->impl.device);
    delete ctx;  // ~Context tears down graphs and events, then destroys the stream
    return TCX_OK;
}

const char* tcx_last_error(const tcx_context* ctx)
{
    if (!ctx) return "null context";
    return ctx->impl.err.c_str();
}

int tcx_synchronize(tcx_context* ctx)
{
    return guarded(ctx, [&] { TCX_HIP_CHECK(hipStreamSynchronize(ctx->impl.stream)); });
}

int tcx_set_cell(tcx_context* ctx, const double recip_rows[9], double volume, const int grid[3], int nspin, int method)
{
    return guarded(ctx, [&] { ctx->impl.set_cell(recip_rows, volume, grid, nspin, method); });
}

int tcx_set_kpoints(tcx_context* ctx, int nk, const double* kfrac, int num_kpoints)
{
    return guarded(ctx, [&] { ctx->impl.set_kpoints(nk, kfrac, num_kpoints); });
}

int tcx_set_basis(tcx_context* ctx, const int* npw, const int* gindex)
{
    return guarded(ctx, [&] { ctx->impl.set_basis(npw, gindex); });
}

int tcx_set_states(tcx_context* ctx, const int* nocc, const double* filling, const double* phi_occ,
                   const double* chi_occ)
{
    return guarded(ctx, [&] { ctx->impl.set_states(nocc, filling, phi_occ, chi_occ); });
}

int tcx_set_jastrow_rpa(tcx_context* ctx, const double A[4], const double F[4])
{
    return guarded(ctx, [&] { ctx->impl.set_jastrow(A, F); });
}

int tcx_set_jastrow_polynomial(tcx_context* ctx, int deg, int C, const double L[4], const double* c_input, int cusp)
{
    return guarded(ctx, [&] { ctx->impl.set_jastrow_poly(deg, C, L, c_input, cusp); });
}

int tcx_get_jastrow_num_rmesh(tcx_context* ctx, int* num_rmesh)
{
    return guarded(ctx, [&] {
        tcx::Context::require(num_rmesh != nullptr, TCX_ERR_ARG, "get_jastrow_num_rmesh: null pointer");
        if (ctx->impl.poly_dirty) ctx->impl.build_jastrow_poly();
        *num_rmesh = ctx->impl.poly_nr;
    });
}

int tcx_jastrow_kernels(tcx_context* ctx, int s1, int s2, long long n, const double* G2, double* uk, double* tc2body)
{
    return guarded(ctx, [&] { ctx->impl.jastrow_kernels(s1, s2, n, G2, uk, tc2body); });
}

int tcx_set_div_correction(tcx_context* ctx, int enabled, double alpha, const double* sum_of_vaux)
{
    return guarded(ctx, [&] { ctx->impl.set_div(enabled, alpha, sum_of_vaux); });
}

int tcx_set_kpoint_symmetry(tcx_context* ctx, const int* k_src, const int* rot, const double* trans, const int* trev)
{
    return guarded(ctx, [&] { ctx->impl.set_symmetry(k_src, rot, trans, trev); });
}

int tcx_get_density(tcx_context* ctx, double* density)
{
    return guarded(ctx, [&] { ctx->impl.get_density(density); });
}

int tcx_set_density(tcx_context* ctx, const double* density)
{
    return guarded(ctx, [&] { ctx->impl.set_density(density); });
}

int tcx_hf2x(tcx_context* ctx, const int* nb, const double* phi, double* h2phi, int precision, unsigned flags)
{
    return guarded(ctx, [&] { ctx->impl.hf2x(nb, phi, h2phi, precision, flags); });
}

int tcx_set_zero_weight_kpoints(tcx_context* ctx, const double* kweight, const double* sum_of_kvaux,
                                const int* nstate, const double* filling, const double* phi, const double* chi)
{
    return guarded(ctx, [&] { ctx->impl.set_zero_weight(kweight, sum_of_kvaux, nstate, filling, phi, chi); });
}

int tcx_hf2x_noalloc(tcx_context* ctx, const int* nb, const double* phi, double* h2phi, int precision)
{
    return guarded(ctx, [&] { ctx->impl.hf2x_noalloc(nb, phi, h2phi, precision); });
}

int tcx_tc2x(tcx_context* ctx, const int* nb, const double* phi, double* h2phi, int precision, unsigned flags)
{
    return guarded(ctx, [&] { ctx->impl.tc2x(nb, phi, h2phi, precision, flags); });
}

int tcx_tc3a1(tcx_context* ctx, const int* nb, const double* phi, double* h3phi, int precision, unsigned flags)
{
    return guarded(ctx, [&] { ctx->impl.tc3_local(0, nb, phi, h3phi, precision, flags); });
}

int tcx_tc3b1(tcx_context* ctx, const int* nb, const double* phi, double* h3phi, int precision, unsigned flags)
{
    return guarded(ctx, [&] { ctx->impl.tc3_local(1, nb, phi, h3phi, precision, flags); });
}

int tcx_tc3a3b3b4b6(tcx_context* ctx, const int* nb, const double* phi, double* h3phi, int precision,
                    unsigned flags)
{
    return guarded(ctx, [&] { ctx->impl.tc3a3b3b4b6(nb, phi, h3phi, precision, flags); });
}

int tcx_tc3a2a4b2b5(tcx_context* ctx, const int* nb, const double* phi, double* h3phi, int precision,
                    unsigned flags)
{
    return guarded(ctx, [&] { ctx->impl.tc3a2a4b2b5(nb, phi, h3phi, precision, flags); });
}

int tcx_kinetic(tcx_context* ctx, const int* nb, const double* phi, double* h1phi, unsigned flags)
{
    return guarded(ctx, [&] { ctx->impl.kinetic(nb, phi, h1phi, flags); });
}

int tcx_hf2h(tcx_context* ctx, const int* nb, const double* phi, double* h2phi, unsigned flags)
{
    return guarded(ctx, [&] { ctx->impl.hf2h(nb, phi, h2phi, flags); });
}

int tcx_tc2h(tcx_context* ctx, const int* nb, const double* phi, double* h2phi, unsigned flags)
{
    return guarded(ctx, [&] { ctx->impl.tc2h(nb, phi, h2phi, flags); });
}

int tcx_apply_all(tcx_context* ctx, const int* nb, const double* phi, double* h1phi, double* h2phi, double* h3phi,
                  int precision, unsigned flags)
{
    return guarded(ctx, [&] { ctx->impl.apply_all(nb, phi, h1phi, h2phi, h3phi, precision, flags); });
}

int tcx_davidson(tcx_context* ctx, const int* nb, double* phi, double* phi_left, double* eig, double* energies,
                 int max_blocks, int num_refresh, int biortho, int precision)
{
    return guarded(ctx, [&] {
        ctx->impl.davidson(nb, phi, phi_left, eig, energies, max_blocks, num_refresh, biortho, precision);
    });
}

int tcx_total_energy(int nk, const int* nb, const double* filling, const double* energies, int num_kpoints,
                     int uses_3body, double* out)
{
    if (nk < 0 || !nb || !filling || !energies || !out || num_kpoints <= 0) return TCX_ERR_ARG;
    // TotalEnergy::calc_total_energy (src/total_energy.cpp:98-163), identity symmetry, no Ewald term;
    // band sums compensated (Neumaier)
    double e[3][2] = {{0, 0}, {0, 0}, {0, 0}}, c[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    auto add = [](double& sum, double& comp, double x) {
        const double t = sum + x;
        comp += std::fabs(sum) >= std::fabs(x) ? (sum - t) + x : (x - t) + sum;
        sum = t;
    };
    long long b = 0;
    for (int k = 0; k < nk; ++k) {
        if (nb[k] < 0) return TCX_ERR_ARG;
        for (int j = 0; j < nb[k]; ++j, ++b) {
            const double f = filling[b];
            for (int n = 0; n < 3; ++n) {
                if (n == 2 && !uses_3body) continue;
                add(e[n][0], c[n][0], f * energies[(b * 3 + n) * 2] / (n + 1));
                add(e[n][1], c[n][1], f * energies[(b * 3 + n) * 2 + 1] / (n + 1));
            }
        }
    }
    for (int n = 0; n < 3; ++n) {
        out[2 * n] = (e[n][0] + c[n][0]) / num_kpoints;
        out[2 * n + 1] = (e[n][1] + c[n][1]) / num_kpoints;
    }
    out[6] = out[0] + out[2] + out[4];
    out[7] = out[1] + out[3] + out[5];
    return TCX_OK;
}

int tcx_set_filling(int nk, const int* nb, const double* eig, const double* kweight, double num_electrons,
                    int smearing, double width, double* filling, double* fermi_energy, int* num_occupied)
{
    // BlochStates::set_filling + Kpoints::return_band_filling
    // (bloch_states_scfloop_filling_density.cpp:9-194, kpoints.cpp:349-387), one independent spin
    if (nk <= 0 || !nb || !eig || !kweight || !filling || !fermi_energy || !num_occupied) return TCX_ERR_ARG;
    if (smearing != 0 && smearing != 1) return TCX_ERR_ARG;
    if (smearing == 1 && !(width > 0.0)) return TCX_ERR_ARG;
    const double spin_factor = 2.0;
    std::vector<long long> off(nk + 1, 0);
    for (int k = 0; k < nk; ++k) {
        if (nb[k] < 0) return TCX_ERR_ARG;
        off[k + 1] = off[k] + nb[k];
    }
    double mu = 0.0;
    if (smearing == 0) {  // fixed: the lowest round(N_e / 2) bands are doubly occupied
        const int nocc = (int)std::lround(num_electrons / 2.0);
        if (std::fabs(nocc - num_electrons / 2.0) > 1e-5) return TCX_ERR_ARG;
        mu = -1e5;
        for (int k = 0; k < nk; ++k) {
            if (nocc > nb[k]) return TCX_ERR_ARG;  // num_bands_tc too small
            for (int b = 0; b < nb[k]; ++b) {
                const bool occ = kweight[k] >= 1e-8 && b < nocc;
                filling[off[k] + b] = occ ? spin_factor : 0.0;
                if (occ) mu = std::max(mu, eig[off[k] + b]);
            }
        }
    } else {  // gaussian: bisection for the Fermi energy, 60 halvings of [-1e5, 1e5]
        double lo = -1e5, hi = 1e5;
        for (int it = 0; it < 60; ++it) {
            mu = 0.5 * (hi + lo);
            double ne = 0.0;
            for (int k = 0; k < nk; ++k)
                for (int b = 0; b < nb[k]; ++b) {
                    double f = 0.0;
                    if (kweight[k] >= 1e-8) f = spin_factor * 0.5 * (1.0 - std::erf((eig[off[k] + b] - mu) / width));
                    filling[off[k] + b] = f;
                    ne += f * (kweight[k] / spin_factor);
                }
            if (ne > num_electrons)
                hi = mu;
            else
                lo = mu;
        }
    }
    double ne = 0.0;
    for (int k = 0; k < nk; ++k) {
        num_occupied[k] = 0;
        for (int b = 0; b < nb[k]; ++b) {
            ne += filling[off[k] + b] * (kweight[k] / spin_factor);
            if (filling[off[k] + b] > 1e-8) ++num_occupied[k];
        }
    }
    *fermi_energy = mu;
    return std::fabs(ne - num_electrons) > 1e-8 ? TCX_ERR_ARG : TCX_OK;  // "set Fermi energy failed"
}

int tcx_eig_dense(int n, const double* A, double* w, double* X)
{
    if (n < 0 || n > 4096 || (n > 0 && (!A || !w || !X))) return TCX_ERR_ARG;
    using cd = std::complex<double>;
    return eig_dense(n, reinterpret_cast<const cd*>(A), reinterpret_cast<cd*>(w), reinterpret_cast<cd*>(X))
               ? TCX_OK
               : TCX_ERR_INTERNAL;
}

int tcx_grad_exchange(tcx_context* ctx, double* grad, double* energy, int precision)
{
    return guarded(ctx, [&] { ctx->impl.grad_x(grad, energy, precision); });
}

int tcx_grad_tc2(tcx_context* ctx, double* grad_left, double* grad_right, double* energy, double* dE_dA,
                 int precision)
{
    return guarded(ctx, [&] { ctx->impl.grad_tc2(grad_left, grad_right, energy, dE_dA, precision); });
}

int tcx_kernel_gradient(tcx_context* ctx, int term, int nbins, double width, double* dE_dK, double* dE_du,
                        double* sums, int precision)
{
    return guarded(ctx, [&] { ctx->impl.kernel_grad(term, nbins, width, dE_dK, dE_du, sums, precision); });
}

int tcx_graph_begin(tcx_context* ctx)
{
    return guarded(ctx, [&] {
        Context::require(!g_capturing, TCX_ERR_STATE, "graph_begin: a capture is already open on this thread");
        Context::require(!ctx->impl.timing_on, TCX_ERR_STATE, "graph_begin: switch timing off first");
        TCX_HIP_CHECK(hipStreamBeginCapture(ctx->impl.stream, hipStreamCaptureModeThreadLocal));
        g_capturing = true;
    });
}

int tcx_graph_end(tcx_context* ctx, int* graph_id)
{
    return guarded(ctx, [&] {
        Context::require(g_capturing, TCX_ERR_STATE, "graph_end: no open capture");
        Context::require(graph_id != nullptr, TCX_ERR_ARG, "graph_end: null id");
        hipGraph_t g = nullptr;
        g_capturing = false;
        TCX_HIP_CHECK(hipStreamEndCapture(ctx->impl.stream, &g));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        TCX_HIP_CHECK(e);
        ctx->impl.graphs.push_back(exec);
        graph_created();
        *graph_id = (int)ctx->impl.graphs.size() - 1;
    });
}

int tcx_graph_launch(tcx_context* ctx, int graph_id)
{
    return guarded(ctx, [&] {
        Context::require(graph_id >= 0 && graph_id < (int)ctx->impl.graphs.size() && ctx->impl.graphs[graph_id],
                         TCX_ERR_ARG, "graph_launch: unknown graph");
        TCX_HIP_CHECK(hipGraphLaunch(ctx->impl.graphs[graph_id], ctx->impl.stream));
    });
}

int tcx_graph_destroy(tcx_context* ctx, int graph_id)
{
    return guarded(ctx, [&] {
        Context::require(graph_id >= 0 && graph_id < (int)ctx->impl.graphs.size() && ctx->impl.graphs[graph_id],
                         TCX_ERR_ARG, "graph_destroy: unknown graph");
        TCX_HIP_CHECK(hipStreamSynchronize(ctx->impl.stream));
        TCX_HIP_CHECK(hipGraphExecDestroy(ctx->impl.graphs[graph_id]));
        ctx->impl.graphs[graph_id] = nullptr;
        graph_destroyed();
    });
}

int tcx_device_alloc(tcx_context* ctx, unsigned long long bytes, void** ptr)
{
    if (!ptr) return TCX_ERR_ARG;
    *ptr = nullptr;
    return guarded(ctx, [&] {
        void* p = nullptr;
        TCX_HIP_CHECK(hipMalloc(&p, bytes ? bytes : 1));
        *ptr = p;
    });
}

int tcx_device_free(tcx_context* ctx, void* ptr)
{
    return guarded(ctx, [&] {
        TCX_HIP_CHECK(hipStreamSynchronize(ctx->impl.stream));
        if (ptr) TCX_HIP_CHECK(hipFree(ptr));
    });
}

int tcx_memcpy(tcx_context* ctx, void* dst, const void* src, unsigned long long bytes, int kind)
{
    return guarded(ctx, [&] {
        if (kind < 0 || kind > 2) throw Error(TCX_ERR_ARG, "memcpy: kind must be 0, 1 or 2");
        if (bytes == 0) return;
        if (!dst || !src) throw Error(TCX_ERR_ARG, "memcpy: null pointer");
        if (kind == 2) {
            launch_copy(dst, src, bytes, ctx->impl.stream);  // streaming kernel, not a DMA engine
        } else {
            const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
            TCX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, k, ctx->impl.stream));
        }
        TCX_HIP_CHECK(hipStreamSynchronize(ctx->impl.stream));
    });
}

int tcx_memset_zero(tcx_context* ctx, void* ptr, unsigned long long bytes)
{
    return guarded(ctx, [&] {
        if (bytes == 0) return;
        if (!ptr) throw Error(TCX_ERR_ARG, "memset_zero: null pointer");
        launch_zero(ptr, bytes, ctx->impl.stream);  // a fill kernel, also inside a captured graph
        if (!g_capturing) TCX_HIP_CHECK(hipStreamSynchronize(ctx->impl.stream));
    });
}

int tcx_set_launch_shape(tcx_context* ctx, int term, long targets, long sources)
{
    return guarded(ctx, [&] {
        if (term < 0 || term > 2)
            throw Error(TCX_ERR_ARG, "launch_shape: term must be 0 (hf2x), 1 (tc2x) or 2 (three-body terms)");
        if (targets < 0 || sources < 0 || targets > 65535 || sources > 65535)
            throw Error(TCX_ERR_ARG, "launch_shape: targets and sources must be in [0, 65535]");
        if (term == 2 && (ctx->impl.shape_targets[2] != targets || ctx->impl.shape_sources[2] != sources))
            ctx->impl.pairsets.clear();  // the three-body pair sets are cached per shape
        ctx->impl.shape_targets[term] = targets;
        ctx->impl.shape_sources[term] = sources;
    });
}

int tcx_set_pipeline_order(tcx_context* ctx, int term, int order)
{
    return guarded(ctx, [&] {
        if (term != 0) throw Error(TCX_ERR_ARG, "pipeline_order: term must be 0 (hf2x)");
        if (order < -1 || order > 1) throw Error(TCX_ERR_ARG, "pipeline_order: order must be -1, 0 or 1");
        ctx->impl.hf2x_order = order;
    });
}

int tcx_get_record_columns(tcx_context* ctx, int term, double* columns)
{
    return guarded(ctx, [&] {
        if (term != 0) throw Error(TCX_ERR_ARG, "record_columns: term must be 0 (hf2x)");
        if (!columns) throw Error(TCX_ERR_ARG, "record_columns: null pointer");
        *columns = ctx->impl.hf2x_record_cols;
    });
}

int tcx_enable_timing(tcx_context* ctx, int on)
{
    return guarded(ctx, [&] {
        ctx->impl.timing_on = on != 0;
        for (int c = 0; c < 4; ++c) {
            ctx->impl.t_ms[c] = 0;
            ctx->impl.t_launches[c] = 0;
            ctx->impl.t_pairs[c] = 0;
        }
    });
}

int tcx_get_timing(tcx_context* ctx, double out_ms[4], long out_launches[4], long out_pairs[4])
{
    return guarded(ctx, [&] {
        ctx->impl.collect_timing();
        for (int c = 0; c < 4; ++c) {
            if (out_ms) out_ms[c] = ctx->impl.t_ms[c];
            if (out_launches) out_launches[c] = ctx->impl.t_launches[c];
            if (out_pairs) out_pairs[c] = ctx->impl.t_pairs[c];
        }
    });
}

int tcx_sum_of_vaux(const double recip[9], double volume, int nk, const double* kfrac, double* alpha_out,
                    double* sum_out)
{
    if (!recip || !kfrac || !alpha_out || !sum_out || nk <= 0 || !(volume > 0)) return TCX_ERR_ARG;
    // Restates Potentials::set_sum_of_Vaux_each (src/potentials_coulomb.cpp:60-161) for an
    // identity-symmetry k set that is also the q set.
    double alpha = std::cbrt(volume) / (2.0 * kPiHost);
    alpha = alpha * alpha;
    auto cart2 = [&](const double m[3]) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double c = recip[d] * m[0] + recip[3 + d] * m[1] + recip[6 + d] * m[2];
            s += c * c;
        }
        return s;
    };
    for (int ik = 0; ik < nk; ++ik) {
        double sum = 0.0;
        bool match_any = false;
        for (int iq = 0; iq < nk; ++iq) {
            double kq[3];
            for (int d = 0; d < 3; ++d) {
                kq[d] = kfrac[3 * ik + d] - kfrac[3 * iq + d];
                while (kq[d] < -1e-8) kq[d] += 1.0;
                while (kq[d] > 1.0 - 1e-8) kq[d] -= 1.0;
            }
            const bool match = kq[0] * kq[0] + kq[1] * kq[1] + kq[2] * kq[2] < 1e-8;
            match_any = match_any || match;
            if (!match) {
                const double k2 = cart2(kq);
                sum += std::exp(-alpha * k2) / k2;
            }
            for (int shell = 1;; ++shell) {
                double on_shell = 0.0;
                for (int gx = -shell; gx <= shell; ++gx)
                    for (int gy = -shell; gy <= shell; ++gy)
                        for (int gz = -shell; gz <= shell; ++gz) {
                            if (std::abs(gx) != shell && std::abs(gy) != shell && std::abs(gz) != shell) continue;
                            const double m[3] = {kq[0] + gx, kq[1] + gy, kq[2] + gz};
                            const double g2 = cart2(m);
                            on_shell += std::exp(-alpha * g2) / g2;
                        }
                sum += on_shell;
                if (on_shell < 1e-20) break;
            }
        }
        sum *= 4.0 * kPiHost;
        if (match_any) sum -= 4.0 * kPiHost * alpha;
        sum_out[ik] = sum;
    }
    *alpha_out = alpha;
    return TCX_OK;
}

}  // extern "C"
