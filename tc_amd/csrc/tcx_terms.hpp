// tcx_terms.hpp -- the context object behind tcx_context.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <functional>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "tcx_internal.hpp"

namespace tcx {

constexpr double kPiHost = 3.14159265358979323846;

// Set while a context's stream is being captured into a HIP graph: nothing may
// allocate, copy synchronously or wait then.
extern thread_local bool g_capturing;

// Live captured graphs hold raw pointers into workspaces: while any exists, a
// released workspace is retired (kept allocated) and freed with the last graph.
void graph_created();
void graph_destroyed();

// Device buffer that only grows.
struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), cap(o.cap) { o.p = nullptr; o.cap = 0; }
    DBuf& operator=(DBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            cap = o.cap;
            o.p = nullptr;
            o.cap = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void ensure(size_t bytes);
    void release();
};

// One launch of the pair pipeline.
struct Launch {
    int pair0 = 0, npairs = 0;      // slice of Plan::pairs; slot p of the pair buffer = pair0 + p
    int target0 = 0, ntargets = 0;  // slice of Plan::targets (pair indices are launch-local)
    int kq0 = 0, nkq = 0;           // rows of Plan::kq used by this launch
    int spin = 0;                   // spin channel of its targets and sources (two independent spins)
};

// (target, source) pairs grouped into launches; TargetEntry::pad holds the
// k-point of the target.
struct PairSet {
    std::vector<PairEntry> pairs;
    std::vector<TargetEntry> targets;
    std::vector<double> kq;  // Cartesian k - q per (k,q) row
    std::vector<int> kq_kq;  // (k, q) per row
    std::vector<Launch> launches;
    int max_launch_pairs = 0;
    DBuf d_pairs, d_targets, d_kq;
};

// A grid of a pair set: k-point and grid index (applied band or occupied orbital).
struct PairItem {
    int k, idx;
};

// Every target paired with every source (same_k: only sources at the target's
// k), in launches of tc consecutive targets x sc consecutive sources.
void build_pairs(PairSet& S, const std::vector<PairItem>& tg, const std::vector<PairItem>& src, long tc, long sc,
                 const std::vector<double>& kcart, bool same_k);

// Support columns of the plane-first hf2x records, per k-q row of a pair set
// (kernels_hf2x_xy.hip): colxy[coloff[r] + c] = i1 | i2 << 16, rowinfo[r][i2] = (lo, len).
struct XyColumns {
    bool built = false;
    int cstride = 0;      // record columns per plane (the largest row)
    double mean_cols = 0; // mean columns per pair
    DBuf d_colxy, d_coloff, d_rowinfo;
};

struct Plan : PairSet {
    std::vector<long long> app_first, app_coef;  // per k: first applied orbital, coefficient offset
    int n_app = 0;
    DBuf d_appjobs;      // OrbEntry per applied band (band order)
    DBuf d_appgradjobs;  // three OrbEntry (dir 0..2) per applied band
    DBuf d_divjobs;      // DivEntry per applied band
    mutable XyColumns xy;
};

// synchronous host -> device copy into a growing buffer
void upload(DBuf& buf, const void* src, size_t bytes);

struct ApplyIO {
    bool host = true;
    long long ncoef = 0;
    const double* phi = nullptr;  // device
    double* out = nullptr;        // device
    double* out_host = nullptr;
};

struct PendingEvent {
    int cls;
    long pairs;
    hipEvent_t a, b;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // geometry and state
    bool cell_set = false, k_set = false, basis_set = false, states_set = false, jastrow_set = false;
    Geom geo{};
    long long N = 0;
    double volume = 0.0;
    int nspin = 1, method = TCX_METHOD_HF;
    // Two independent spins: every per-k array of the states and of the applies (occupied
    // counts, fillings, orbitals, applied band counts) runs over (spin, k), spin-major, and
    // the k-points and the basis are those of the nk_phys given k-points repeated per spin.
    // nk = nspin * nk_phys; spin of the k-block k = k / nk_phys.
    int nk = 0, nk_phys = 0, num_kpoints = 0;
    std::vector<double> kfrac, kcart;
    std::vector<int> npw;
    std::vector<long long> pw_off;
    std::vector<int> gindex_h;  // host copy of the plane-wave grid indices
    int max_npw = 0;
    std::vector<int> nocc;
    std::vector<long long> occ_first, occ_coef;
    int n_occ = 0;
    std::vector<double> fill;
    double A[4] = {0, 0, 0, 0}, F[4] = {1, 1, 1, 1};
    bool jastrow_zero = true;
    bool div_enabled = false;
    double alpha = 0.0;
    std::vector<double> sum_vaux;
    long long states_version = 0;  // counts tcx_set_states calls
    // k-mesh symmetry (tcx_set_kpoint_symmetry): sym_src[k] = irreducible k of k,
    // sym_ops[k] = its rotation (dst/src unused); off = identity symmetry
    bool sym_set = false;
    std::vector<int> sym_src;
    std::vector<SymRotJob> sym_ops;
    bool sym_jobs_valid = false;
    int n_sym_jobs = 0;

    // device data
    DBuf d_gindex, d_kcart, d_phi_occ, d_chi_occ, d_fill;
    DBuf d_occR[2], d_chiR[2], d_gradR[2];  // per precision
    bool occR_valid[2] = {false, false}, chiR_valid[2] = {false, false}, gradR_valid[2] = {false, false};
    DBuf d_density;  // double2 [2][N], real space
    bool density_valid = false;
    DBuf d_occjobs;      // OrbEntry per occupied orbital
    DBuf d_occgradjobs;  // three OrbEntry (dir 0..2) per occupied orbital
    DBuf d_dc;       // divergence factor per k
    DBuf d_symjobs;  // SymRotJob per occupied orbital at a replica k
    DBuf d_phi_app, d_out, d_appR, d_appGrad, d_tmp64, d_W, d_pairbuf, d_ktab, d_work;
    DBuf d_Wc;  // compensation grids of the band-pair sums in d_W

    // launch shape per term (0 hf2x, 1 tc2x): targets x sources per launch, 0 = auto
    long shape_targets[3] = {0, 0, 0};  // hf2x, tc2x, three-body pair sets (0 = automatic)
    long shape_sources[3] = {0, 0, 0};
    int hf2x_order = -1;          // -1 automatic, 0 z-first, 1 plane-first (FP32)
    double hf2x_record_cols = 0;  // (kx, ky) columns per plane of the last hf2x apply's records
    bool use_plane_first(int prec) const;
    void build_xy_columns(const Plan& P);
    void hf2x_plane_first(const Plan& P);
    std::map<std::tuple<std::vector<int>, long, long>, Plan> plans;
    // pair sets of the three-body exchange terms, keyed by (applied bands, kind)
    std::map<std::tuple<std::vector<int>, int>, PairSet> pairsets;
    DBuf d_qsum, d_qsumT, d_rec, d_Q2, d_Q4, d_Q5, d_Q4T, d_Q5T, d_appR64, d_chans, d_aux, d_local;

    // Intermediates that depend on the occupied set (and density, Jastrow, divergence
    // correction, k-mesh symmetry) but not on the applied bands, kept across applies:
    // tc3a2a4b2b5's q-sum products Q4T/Q5T in the working precision `prec` plus its
    // [3b2]+[3a2] local potential (one FP64 grid in d_a8X), and the FP64 real-space local
    // potentials of hf2h (1 grid), tc2h (V2 and dnu: 4 grids), tc3a1 and tc3b1 (1 each).
    struct LocalCache {
        DBuf buf;
        bool valid = false;
        int prec = -1;
    };
    LocalCache c_a8, c_hartree, c_tc2h, c_tc3a1, c_tc3b1;
    void invalidate_local();

    // The Hermitian conjugate H^dagger that the bi-orthogonal Davidson applies to its left
    // vectors: u -> -u and the right and left occupied orbitals swapped, with the density
    // left as it is (Jastrow::take_Hermitian_conj, BlochStates::switch_left_right_orbitals;
    // diagonalization_davidson.cpp:234-251).  hermitian_conj() toggles it.  The caches of
    // the other orientation (local potentials, tc3a2a4b2b5's q-sum products, the gradient
    // grids) are kept beside the current ones and swapped with them, so that alternating
    // H and H^dagger applies do not rebuild them; every invalidation clears both.
    // Zero-weight k-points (tcx_set_zero_weight_kpoints, tcx_zeroweight.cpp): zw_k[k] marks
    // one; zw_match[k]: k lies on the weighted mesh; zw_skv: sum_of_kVaux [nk][3]; the
    // states of k are grids zw_first[k] .. + zw_n[k] (behind the n_occ occupied ones) and
    // coefficients from zw_coef[k] in the orbital arrays.
    bool zw_set = false;
    std::vector<char> zw_k;
    std::vector<int> zw_match, zw_n, zw_first;
    std::vector<long long> zw_coef;
    std::vector<double> zw_skv;
    int n_zw = 0;
    DBuf d_zw_match, d_zw_kv, d_pwoff, d_zwtab;
    void set_zero_weight(const double* kweight, const double* skv, const int* nstate, const double* filling,
                         const double* phi, const double* chi);
    void clear_zero_weight();
    void tc2x_kvaux(std::vector<double>& kv) const;
    bool tc3_kvaux(int k, double kv[3]) const;
    // states a divergence job projects on at most (occupied orbitals, or zero-weight states)
    int max_div_states() const
    {
        int m = 0;
        for (int k = 0; k < nk; ++k) m = std::max(m, zw_set && zw_k[k] ? zw_n[k] : nocc[k]);
        return m;
    }

    bool conj_mode = false;
    LocalCache a_a8, a_hartree, a_tc2h, a_tc3a1, a_tc3b1;
    DBuf a_Q4T, a_Q5T, a_gradR[2];
    bool a_gradR_valid[2] = {false, false};
    void hermitian_conj();

    // HIP graphs captured from this context's stream (tcx_graph_*)
    std::vector<hipGraphExec_t> graphs;

    // timing
    bool timing_on = false;
    double t_ms[4] = {0, 0, 0, 0};
    long t_launches[4] = {0, 0, 0, 0};
    long t_pairs[4] = {0, 0, 0, 0};
    std::vector<PendingEvent> pending;

    // Graphs and events go first, then the stream they were recorded on; the
    // workspaces (DBuf members) are released after this body, with a device-wide
    // synchronisation each.
    ~Context()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& g : graphs)
            if (g) {
                (void)hipGraphExecDestroy(g);
                graph_destroyed();
            }
        graphs.clear();
        for (auto& ev : pending) {
            (void)hipEventDestroy(ev.a);
            (void)hipEventDestroy(ev.b);
        }
        pending.clear();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    static void require(bool cond, int code, const char* what);
    void timed(int cls, long pairs, const std::function<void()>& launch);
    void collect_timing();

    void set_cell(const double recip_rows[9], double vol, const int grid[3], int ns, int meth);
    void set_kpoints(int nk, const double* kf, int numk);
    void set_basis(const int* npw, const int* gindex);
    void set_states(const int* nocc, const double* filling, const double* phi_occ, const double* chi_occ);
    void set_jastrow(const double A[4], const double F[4]);
    void set_div(int enabled, double alpha, const double* svaux);
    void set_symmetry(const int* k_src, const int* rot, const double* trans, const int* trev);
    void get_density(double* out);
    void set_density(const double* in);
    void rotate_qsums(double* Q2, double* Q4, double* Q5);
    void invalidate_states();

    void orbitals_to_real(const double* coef, const OrbEntry* d_jobs, int n, int prec, void* out);
    void ensure_occR(int prec);
    void ensure_chiR(int prec);
    void ensure_gradR(int prec);
    bool jastrow_dA = false;  // evaluate d/dA of the Jastrow kernels (adjoint of the A-dependence)
    JastrowPair jastrow_pair(int s1, int s2);  // builds the polynomial blocks if they are stale
    // polynomial Jastrow (tcx_set_jastrow_polynomial): deg 0 = none
    int poly_deg = 0, poly_C = 3;
    bool poly_cusp = false;
    double poly_L[4] = {1, 1, 1, 1};
    std::vector<double> poly_c;  // [4][deg] as given (entry 0 of each pair replaced at build)
    bool poly_dirty = false;     // derived parameters to rebuild (RPA F or the grid changed)
    int poly_nr = 0;             // num_rmesh_J3_integ
    size_t poly_stride = 0;      // doubles per pair block
    DBuf d_jpoly[2];             // [4] pair blocks: u as given, u negated
    bool u_negated = false;      // A and the polynomial negated (Hermitian conjugate)
    void set_jastrow_poly(int deg, int C, const double* L, const double* c_in, int cusp);
    void build_jastrow_poly();
    void negate_u();
    bool u_zero() const { return jastrow_zero && poly_deg == 0; }
    void jastrow_kernels(int s1, int s2, long long n, const double* G2, double* uk, double* v);
    void ensure_density();

    void resolve_shape(int term, long& tc, long& sc, long ntargets, long nsources) const;
    const Plan& get_plan(const std::vector<int>& nb, int term);
    void check_apply_args(const int* nb, const double* phi, const double* out, int prec);
    ApplyIO begin_apply(const int* nb, const double* phi, double* out, unsigned flags, const Plan& P);
    void end_apply(const ApplyIO& io);
    double div_factor(int k) const;
    void apply_divcorr(const Plan& P, const ApplyIO& io, const double* occ_left);
    // clears W and its compensation grids for n_app applied bands; returns W
    double* begin_W(int n_app);
    // the terms whose spin structure is not the same-spin pairing of the two-body exchange
    void require_one_spin(const char* what) const;
    void finish_W(const Plan& P, const ApplyIO& io, double factor);

    void hf2x(const int* nb, const double* phi, double* out, int prec, unsigned flags);
    // hf2x on device pointers with all working memory in the library's fixed device array
    // (tcx_noalloc.hip); no device allocation
    void hf2x_noalloc(const int* nb, const double* phi, double* out, int prec);
    void tc2x(const int* nb, const double* phi, double* out, int prec, unsigned flags);
    void tc3_local(int which, const int* nb, const double* phi, double* out, int prec, unsigned flags);

    // three-body exchange terms and the assembly terms
    const PairSet& pair_set(const std::vector<int>& nb, int kind, int channels, int prec);
    std::vector<PairItem> applied_items(const std::vector<int>& nb) const;
    std::vector<PairItem> occupied_items(bool irreducible_only = false) const;
    void apply_local(const Plan& P, const ApplyIO& io, const double* V, double factor);
    void density_G(double* nG);
    void tc3a3b3b4b6(const int* nb, const double* phi, double* out, int prec, unsigned flags);
    void tc3a2a4b2b5(const int* nb, const double* phi, double* out, int prec, unsigned flags);
    void a8_occupied_phase(const std::vector<int>& nb, int prec);
    void a8_kvaux(const Plan& P, const ApplyIO& io, const std::vector<int>& nb, double fa4, double fb5);
    void kinetic(const int* nb, const double* phi, double* out, unsigned flags);
    void hf2h(const int* nb, const double* phi, double* out, unsigned flags);
    void tc2h(const int* nb, const double* phi, double* out, unsigned flags);
    void apply_all(const int* nb, const double* phi, double* h1, double* h2, double* h3, int prec, unsigned flags);
    void davidson(const int* nb, double* phi, double* phi_left, double* eig, double* energies, int max_blocks,
                  int num_refresh, int biortho, int prec);

    // gradients of the occupied-set energies (tcx_adjoint.cpp)
    void swap_left_right();
    void grad_x(double* grad, double* energy, int prec);
    void grad_tc2(double* grad_left, double* grad_right, double* energy, double* dE_dA, int prec);
    void kernel_grad(int term, int nbins, double width, double* dE_dK, double* dE_du, double* sums, int prec);
};

// Dense complex eigen-decomposition A X = X diag(w), A and X n x n row-major,
// unit-norm eigenvector columns.  Returns false if the QR iteration does not converge.
bool eig_dense(int n, const std::complex<double>* A, std::complex<double>* w, std::complex<double>* X);

}  // namespace tcx

struct tcx_context {
    tcx::Context impl;
};
