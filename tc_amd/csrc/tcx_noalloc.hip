// tcx_noalloc.hip -- the Fock-exchange apply without device allocation (tcx_hf2x_noalloc).
//
// For callers that cannot allocate device memory while they run (serving loops, code
// that must not call hipMalloc).  The per-call working memory of hf2x -- launch tables,
// the applied bands in real space, the FP64 accumulators W, the Coulomb table and the
// pair records -- lives in `scratch`, a file-scope device array of kScratchBytes that
// every GPU the library runs on holds until the process ends.  Its device address is
// looked up once per device (scratch_setup, called by tcx_create with the device
// current) and kept in g_scratch[device]; host code passes that address, never the
// array's name, to kernels and copies.
//
// The apply runs the same kernels as Context::hf2x over consecutive chunks of applied
// bands sized to fit the array: per chunk, the launch tables are copied in (after the
// previous chunk has finished with them), the chunk's bands go to real space, every
// occupied source block runs passes A, B and C into the chunk's W, and W is transformed
// and gathered into the output, followed by the divergence correction.  All contexts on
// one GPU share its array: calls that overlap on one GPU race.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "tcx_terms.hpp"

namespace tcx {

namespace {

constexpr size_t kScratchBytes = (size_t)256 << 20;  // 256 MiB per GPU
constexpr size_t kTableBytes = (size_t)4 << 20;      // launch tables at the start of the array
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};
int g_scratch_status[kMaxDevices] = {};  // 0 not looked up, 1 ok, -1 lookup failed

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return TCX_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_status[device] != 0) return g_scratch_status[device] > 0 ? TCX_OK : TCX_ERR_HIP;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        g_scratch_status[device] = -1;
        return TCX_ERR_HIP;
    }
    g_scratch[device] = p;
    g_scratch_status[device] = 1;
    return TCX_OK;
}

size_t scratch_bytes() { return kScratchBytes; }

void Context::hf2x_noalloc(const int* nb_, const double* phi, double* out, int prec)
{
    require(!g_capturing, TCX_ERR_STATE, "hf2x_noalloc: not capturable (it waits between band chunks)");
    check_apply_args(nb_, phi, out, prec);
    require_one_spin("hf2x_noalloc");
    for (const void* p : {(const void*)phi, (const void*)out}) {
        hipPointerAttribute_t at{};
        const hipError_t e = hipPointerGetAttributes(&at, p);
        if (e != hipSuccess) (void)hipGetLastError();
        require(e == hipSuccess && at.type == hipMemoryTypeDevice && at.device == device, TCX_ERR_ARG,
                "hf2x_noalloc: phi and h2phi must be device memory of the context's GPU");
    }
    void* base_p = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (device >= 0 && device < kMaxDevices && g_scratch_status[device] > 0) base_p = g_scratch[device];
    }
    require(base_p != nullptr, TCX_ERR_HIP, "hf2x_noalloc: the scratch array of this GPU was not found at set-up");
    require(!zw_set, TCX_ERR_UNSUPPORTED, "hf2x_noalloc: zero-weight k-points need tcx_hf2x");
    require(occR_valid[prec], TCX_ERR_STATE,
            "hf2x_noalloc: the occupied orbitals of this precision are not in real space yet; one tcx_hf2x call "
            "in this precision after tcx_set_states builds them");
    char* base = (char*)base_p;

    // applied bands in band order: k and coefficient offset
    std::vector<int> tk;
    std::vector<long long> tcoef;
    long long coef = 0;
    for (int k = 0; k < nk; ++k) {
        for (int j = 0; j < nb_[k]; ++j) {
            tk.push_back(k);
            tcoef.push_back(coef + (long long)j * npw[k]);
        }
        coef += (long long)nb_[k] * npw[k];
    }
    const int ntarg = (int)tk.size();
    if (ntarg == 0 || n_occ == 0) {
        TCX_HIP_CHECK(hipStreamSynchronize(stream));
        return;
    }
    std::vector<PairItem> src = occupied_items();
    const int nsrc = (int)src.size();

    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8;
    const size_t g64 = (size_t)N * 16, gT = (size_t)N * es;

    // chunk shape: T targets with (FP32) an FP64 staging grid that later holds W, the
    // working-precision grids, and per source a pair record plus at most one Coulomb row
    // per target k-point; tables: every pair of the chunk, one target entry per (target,
    // source block), the k-q rows, one scatter job and one divergence job per target
    auto chunk_fits = [&](int T, int S) {
        const long long blocks = (nsrc + S - 1) / S;
        const size_t tb = align256((size_t)T * nsrc * sizeof(PairEntry)) +
                          align256((size_t)T * blocks * sizeof(TargetEntry)) +
                          align256((size_t)T * nsrc * 3 * sizeof(double)) + align256((size_t)T * sizeof(OrbEntry)) +
                          align256((size_t)T * sizeof(DivEntry));
        const size_t grids = align256((size_t)T * g64) + align256((size_t)T * gT);
        const size_t per_launch = align256((size_t)T * S * gT) + align256((size_t)T * S * (gT / 2));
        return tb <= kTableBytes && kTableBytes + grids + per_launch <= kScratchBytes && (long)T * S <= 65535;
    };
    int T = std::min(ntarg, 16), S = 0;
    for (; T >= 1; T /= 2) {
        int lo = 0, hi = nsrc;  // largest S in [1, nsrc] that fits
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (chunk_fits(T, mid))
                lo = mid;
            else
                hi = mid - 1;
        }
        if (lo >= 1) {
            S = lo;
            break;
        }
    }
    require(T >= 1 && S >= 1, TCX_ERR_ARG,
            "hf2x_noalloc: one band pair of this grid and occupied set exceeds the 256 MiB scratch array");

    // scratch layout
    char* p = base;
    PairEntry* d_pairs = (PairEntry*)p;
    p += align256((size_t)T * nsrc * sizeof(PairEntry));
    const long long blocks = (nsrc + S - 1) / S;
    TargetEntry* d_targets = (TargetEntry*)p;
    p += align256((size_t)T * blocks * sizeof(TargetEntry));
    double* d_kq = (double*)p;
    p += align256((size_t)T * nsrc * 3 * sizeof(double));
    OrbEntry* d_jobs = (OrbEntry*)p;
    p += align256((size_t)T * sizeof(OrbEntry));
    DivEntry* d_div = (DivEntry*)p;
    char* grids = base + kTableBytes;
    double* W = (double*)grids;  // FP32: first the FP64 staging grids of the chunk's bands
    void* appR = grids + align256((size_t)T * g64);
    char* launch_base = (char*)appR + align256((size_t)T * gT);
    void* pairbuf = launch_base;
    void* ktab = launch_base + align256((size_t)T * S * gT);  // at most T * S rows of N reals

    double factor = -4.0 * kPiHost / (num_kpoints * volume);  // two_body_factor, hf2x.cpp:29-30
    if (nspin == 1) factor /= 2.0;

    for (int t0 = 0; t0 < ntarg; t0 += T) {
        const int nt = std::min(T, ntarg - t0);
        // launch tables of this chunk (targets chunk-local: app and W slot t - t0)
        std::vector<PairEntry> pairs;
        std::vector<TargetEntry> targets;
        std::vector<double> kq;
        struct L {
            int pair0, npairs, target0, ntargets, kq0, nkq;
        };
        std::vector<L> launches;
        // the distinct k-points of the chunk's targets (band order keeps them sorted)
        std::vector<int> ks;
        for (int t = 0; t < nt; ++t)
            if (ks.empty() || ks.back() != tk[t0 + t]) ks.push_back(tk[t0 + t]);
        for (int s0 = 0; s0 < nsrc; s0 += S) {
            const int s1 = std::min(nsrc, s0 + S);
            std::vector<int> qs;  // the distinct q of the source block, sorted likewise
            for (int s = s0; s < s1; ++s)
                if (qs.empty() || qs.back() != src[s].k) qs.push_back(src[s].k);
            const int nq = (int)qs.size();
            L l{(int)pairs.size(), 0, (int)targets.size(), nt, (int)(kq.size() / 3), (int)ks.size() * nq};
            for (int k : ks)
                for (int q : qs)
                    for (int d = 0; d < 3; ++d) kq.push_back(kcart[3 * k + d] - kcart[3 * q + d]);
            int local = 0, ik = 0;
            for (int t = 0; t < nt; ++t) {
                while (ks[ik] != tk[t0 + t]) ++ik;
                TargetEntry te{t, local, 0, tk[t0 + t]};
                int iq = 0;
                for (int s = s0; s < s1; ++s) {
                    while (qs[iq] != src[s].k) ++iq;
                    pairs.push_back({t, src[s].idx, l.kq0 + ik * nq + iq, 0});
                    ++local;
                    ++te.npair;
                }
                targets.push_back(te);
            }
            l.npairs = local;
            launches.push_back(l);
        }
        std::vector<OrbEntry> jobs;
        std::vector<DivEntry> div;
        for (int t = 0; t < nt; ++t) {
            const int k = tk[t0 + t];
            jobs.push_back({tcoef[t0 + t], npw[k], (int)pw_off[k], k, -1});
            div.push_back({tcoef[t0 + t], occ_coef[k], npw[k], nocc[k], (int)occ_first[k], k});
        }
        int max_rows = 1;
        for (const L& l : launches) max_rows = std::max(max_rows, l.nkq);
        require((size_t)max_rows * (gT / 2) <= align256((size_t)T * S * (gT / 2)), TCX_ERR_INTERNAL,
                "hf2x_noalloc: Coulomb rows exceed their scratch slice");
        require(pairs.size() <= (size_t)T * nsrc && targets.size() <= (size_t)T * blocks && kq.size() <= (size_t)T * nsrc * 3,
                TCX_ERR_INTERNAL, "hf2x_noalloc: launch tables exceed their scratch slice");

        // the previous chunk's kernels read the tables: wait, then copy this chunk's in
        TCX_HIP_CHECK(hipStreamSynchronize(stream));
        TCX_HIP_CHECK(hipMemcpy(d_pairs, pairs.data(), pairs.size() * sizeof(PairEntry), hipMemcpyHostToDevice));
        TCX_HIP_CHECK(hipMemcpy(d_targets, targets.data(), targets.size() * sizeof(TargetEntry), hipMemcpyHostToDevice));
        TCX_HIP_CHECK(hipMemcpy(d_kq, kq.data(), kq.size() * sizeof(double), hipMemcpyHostToDevice));
        TCX_HIP_CHECK(hipMemcpy(d_jobs, jobs.data(), jobs.size() * sizeof(OrbEntry), hipMemcpyHostToDevice));
        TCX_HIP_CHECK(hipMemcpy(d_div, div.data(), div.size() * sizeof(DivEntry), hipMemcpyHostToDevice));

        // applied bands of the chunk to real space (FP64 transform, rounded once for FP32)
        double* dst64 = prec == TCX_PREC_FP64 ? (double*)appR : W;
        launch_zero(dst64, (size_t)nt * g64, stream);
        launch_scatter(phi, (const int*)d_gindex.p, d_jobs, nt, max_npw, dst64, N, (const double*)d_kcart.p, geo,
                       stream);
        launch_fft3d(TCX_PREC_FP64, +1, dst64, nt, geo, 1.0, stream);
        if (prec == TCX_PREC_FP32) launch_convert_d2f(dst64, (float*)appR, (long long)nt * N, stream);
        launch_zero(W, (size_t)nt * g64, stream);
        for (const L& l : launches) {
            const PairEntry* lp = d_pairs + l.pair0;
            const TargetEntry* lt = d_targets + l.target0;
            launch_coulomb_table(prec, ktab, d_kq, l.kq0, l.nkq, geo, stream);
            launch_hf2x_A(prec, geo, appR, d_occR[prec].p, lp, l.npairs, pairbuf, stream);
            launch_hf2x_B(prec, geo, pairbuf, lp, l.npairs, ktab, l.kq0, stream);
            // plain FP64 band-pair sums: the working array holds no compensation grids
            launch_hf2x_C(prec, geo, pairbuf, d_occR[prec].p, lp, lt, l.ntargets, (const double*)d_fill.p, W, nullptr,
                          stream);
        }
        launch_fft3d(TCX_PREC_FP64, -1, W, nt, geo, 1.0 / (double)N, stream);
        launch_gather_add(W, (const int*)d_gindex.p, d_jobs, nt, max_npw, out, N, factor, stream);
        if (div_enabled)
            launch_divcorr(phi, (const double*)d_phi_occ.p, (const double*)d_phi_occ.p, d_div, nt,
                           (const double*)d_fill.p, (const double*)d_dc.p, out, stream);
    }
    TCX_HIP_CHECK(hipStreamSynchronize(stream));
}

}  // namespace tcx
