// tcx_tc3x.cpp -- three-body exchange terms tc3a3b3b4b6 / tc3a2a4b2b5 and the
// assembly terms kinetic / hf2h / tc2h / all on the context.
//
// The pair work of the three-body terms runs in the working precision
// (FP32-split or FP64) through batched record pipelines
// (kernels_tc3x.hip); all accumulators, the divergence corrections and the
// local potentials are FP64.  Without a k-mesh symmetry the q2-resolved sums of
// tc3a2a4b2b5 are formed at every q2 directly; with one (tcx_set_kpoint_symmetry)
// they are formed at the irreducible q2 only and rotated to the other q2 by the
// U^+ transforms of tc3a2a4b2b5.cpp:576-765, as the reference does.  The kVaux
// branches act at zero-weight k-points off the mesh (tcx_set_zero_weight_kpoints):
// the off-mesh [3b6] table and the [3b3], [3a3,b4], [3b4] terms of tc3a3b3b4b6
// (kernels_zw.hip) and the [3a4,b5] block of tc3a2a4b2b5 (a8_kvaux).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "tcx_terms.hpp"

namespace tcx {

namespace {

// Pair records in flight per launch are bounded by this many bytes.
constexpr size_t kRecordBudget = (size_t)3 << 30;
constexpr long kTc3Targets = 32;

// kPairsIrr*: as kPairsOcc*, with targets at the irreducible k of the k-mesh symmetry only
enum PairKind { kPairsAppOcc = 0, kPairsOccOcc = 1, kPairsOccSameK = 2, kPairsIrrOcc = 3, kPairsIrrSameK = 4 };

ProdChan chan(const void* a, int a_sel, int a_mul, int a_off, int a_conj, const void* b, int b_sel, int b_mul,
              int b_off, int b_conj)
{
    return ProdChan{a, b, a_sel, a_mul, a_off, a_conj, b_sel, b_mul, b_off, b_conj};
}

}  // namespace

// Pair set of one kind, with launches sized so that `channels` record grids
// per pair stay under kRecordBudget.
const PairSet& Context::pair_set(const std::vector<int>& nb, int kind, int channels, int prec)
{
    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8;
    // only the (k,j) x (q,i) kind depends on the applied bands
    auto key = std::make_tuple(kind == kPairsAppOcc ? nb : std::vector<int>(), kind * 1000 + channels * 2 + prec);
    auto it = pairsets.find(key);
    if (it != pairsets.end()) return it->second;
    const bool irr = kind == kPairsIrrOcc || kind == kPairsIrrSameK;
    const std::vector<PairItem> tg = kind == kPairsAppOcc ? applied_items(nb) : occupied_items(irr);
    const std::vector<PairItem> src = occupied_items();
    // the launch shape (tcx_set_launch_shape term 2) within the record budget, else automatic
    const long maxpairs = std::max<long>(1, (long)(kRecordBudget / ((size_t)channels * N * es)));
    const long want_t = shape_targets[2] > 0 ? shape_targets[2] : kTc3Targets;
    const long tc = std::max<long>(1, std::min<long>({want_t, maxpairs, (long)std::max<size_t>(1, tg.size())}));
    long sc = std::max<long>(1, maxpairs / tc);
    if (shape_sources[2] > 0) sc = std::min<long>(sc, shape_sources[2]);
    sc = std::min<long>(sc, (long)std::max<size_t>(1, src.size()));
    PairSet S;
    build_pairs(S, tg, src, tc, sc, kcart, kind == kPairsOccSameK || kind == kPairsIrrSameK);
    auto res = pairsets.emplace(key, std::move(S));
    return res.first->second;
}

// nG[2][N] = FFT_forward(density_s), FP64
void Context::density_G(double* nG)
{
    ensure_density();
    const size_t g = (size_t)N * 16;
    TCX_HIP_CHECK(hipMemcpyAsync(nG, d_density.p, 2 * g, hipMemcpyDeviceToDevice, stream));
    launch_fft3d(TCX_PREC_FP64, -1, nG, 2, geo, 1.0 / (double)N, stream);
}

// H phi_kj += factor * FFT_forward[V phi_kj] gathered on the sphere of k (V real space, FP64),
// the apply loop of tc3a1.cpp:86-116, hf2h.cpp:55-82 and tc3a2a4b2b5.cpp:922-1034.
void Context::apply_local(const Plan& P, const ApplyIO& io, const double* V, double factor)
{
    const size_t g = (size_t)N * 16;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(P.n_app, ((size_t)1 << 30) / g));
    d_W.ensure((size_t)chunk * g);
    const OrbEntry* jobs = (const OrbEntry*)P.d_appjobs.p;
    for (int c0 = 0; c0 < P.n_app; c0 += chunk) {
        const int nc = std::min(chunk, P.n_app - c0);
        double* Wc = (double*)d_W.p;
        launch_zero(Wc, (size_t)nc * g, stream);
        launch_scatter(io.phi, (const int*)d_gindex.p, jobs + c0, nc, max_npw, Wc, N, (const double*)d_kcart.p, geo,
                       stream);
        launch_fft3d(TCX_PREC_FP64, +1, Wc, nc, geo, 1.0, stream);
        launch_mul_shared(Wc, nc, V, N, stream);
        launch_fft3d(TCX_PREC_FP64, -1, Wc, nc, geo, 1.0 / (double)N, stream);
        launch_gather_add(Wc, (const int*)d_gindex.p, jobs + c0, nc, max_npw, io.out, N, factor, stream);
    }
}

// ------------------------------------------------------------------ tc3a3b3b4b6
void Context::tc3a3b3b4b6(const int* nb_, const double* phi, double* out, int prec, unsigned flags)
{
    check_apply_args(nb_, phi, out, prec);
    require_one_spin("tc3a3b3b4b6");
    require(method != TCX_METHOD_HF, TCX_ERR_STATE, "tc3a3b3b4b6: the cell was set up for HF");
    require(jastrow_set, TCX_ERR_STATE, "tc3a3b3b4b6: call tcx_set_jastrow_rpa first");
    std::vector<int> nb(nb_, nb_ + nk);
    const Plan& P = get_plan(nb, 0);
    ApplyIO io = begin_apply(nb_, phi, out, flags, P);
    // u = 0 makes every contribution vanish, and the G=0 correction is skipped with it (:99)
    if (P.n_app > 0 && n_occ > 0 && !u_zero()) {
        const bool bitc = method == TCX_METHOD_BITC;
        const size_t es = prec == TCX_PREC_FP64 ? 16 : 8;
        const size_t g64 = (size_t)N * 16, gT = (size_t)N * es;
        const double NkV = num_kpoints * volume;
        const double base = 1.0 / (NkV * NkV);  // :33-47
        double fa3 = -base, fb3 = base, fb4 = -base, fb6 = base;
        if (nspin == 1) {
            fa3 /= 2.0;
            fb3 /= 2.0;
            fb4 /= 4.0;
            fb6 /= 2.0;
        }
        const JastrowPair j00 = jastrow_pair(0, 0), j10 = jastrow_pair(1, 0), j01 = jastrow_pair(0, 1);
        ensure_occR(prec);
        if (bitc) ensure_chiR(prec);
        const void* occR = d_occR[prec].p;
        const void* chiR = bitc ? d_chiR[prec].p : occR;  // bra orbital (:245)
        ensure_density();

        // n_s(G), dnu_d(r) = FFT_backward[sum_s3 u(G,0,s3) G_d n_s3(G)] (:126-151), tG, tR
        d_local.ensure(7 * g64);
        double* nG = (double*)d_local.p;
        double* dnu = nG + 4 * N;
        double* tG = nG + 10 * N;
        double* tR = nG + 12 * N;
        density_G(nG);
        launch_dnu(nG, dnu, geo, j00, j01, stream);
        launch_fft3d(TCX_PREC_FP64, +1, dnu, 3, geo, 1.0, stream);

        // applied orbitals and dnu_d phi_j in real space
        d_appR.ensure((size_t)P.n_app * gT);
        d_appGrad.ensure((size_t)3 * P.n_app * gT);
        orbitals_to_real(io.phi, (const OrbEntry*)P.d_appjobs.p, P.n_app, prec, d_appR.p);
        launch_mul_dnu(prec, d_appR.p, dnu, d_appGrad.p, P.n_app, N, stream);

        begin_W(P.n_app);
        d_qsum.ensure((size_t)3 * P.n_app * g64);
        launch_zero(d_qsum.p, (size_t)3 * P.n_app * g64, stream);
        double* W = (double*)d_W.p;
        double* qsum = (double*)d_qsum.p;

        const PairSet& S = pair_set(nb, kPairsAppOcc, 11, prec);
        d_rec.ensure((size_t)std::max(1, S.max_launch_pairs) * 11 * gT);
        const double* kq = (const double*)S.d_kq.p;
        int max_rows = 1;
        for (const Launch& L : S.launches) max_rows = std::max(max_rows, L.nkq);
        d_ktab.ensure((size_t)max_rows * gT);  // (u(kq+G, 00), u(kq+G, 10)) per k-q row
        const double* fill = (const double*)d_fill.p;

        // phase 1: [3b3], [3a3,b4,b6] (:179-371)
        {
            ProdChan ch[4];
            ch[0] = chan(chiR, 1, 1, 0, 1, d_appR.p, 0, 1, 0, 0);
            for (int d = 0; d < 3; ++d) ch[1 + d] = chan(chiR, 1, 1, 0, 1, d_appGrad.p, 0, 3, d, 0);
            upload(d_chans, ch, sizeof(ch));
        }
        for (const Launch& L : S.launches) {
            const PairEntry* pairs = (const PairEntry*)S.d_pairs.p + L.pair0;
            const TargetEntry* targets = (const TargetEntry*)S.d_targets.p + L.target0;
            const int np = L.npairs;
            char* rec4 = (char*)d_rec.p;
            char* Pb = rec4 + (size_t)4 * np * gT;
            char* Y = Pb + (size_t)6 * np * gT;
            launch_u2_table(prec, d_ktab.p, kq, L.kq0, L.nkq, geo, j00, j10, stream);
            timed(0, np, [&] {
                launch_products_z(prec, pairs, np, (const ProdChan*)d_chans.p, 4, rec4, geo, 1.0 / (double)N,
                                  stream);
            });
            timed(1, np, [&] {
                launch_xy_a7c1(prec, geo, rec4, Pb, Y, pairs, np, d_ktab.p, L.kq0, kq, fill, fb3, stream);
                // one independent spin: both spin channels read the one density grid
                launch_za7_mid(prec, geo, Pb, occR, pairs, targets, L.ntargets, (const double*)d_density.p,
                               nspin == 1 ? 0 : N, qsum, 1.0 / (double)N, stream);
                launch_xy_a7c2(prec, geo, Pb, Y, pairs, np, d_ktab.p, L.kq0, kq, fb6, stream);
            });
            timed(2, np, [&] {  // z pass of FFT_backward fused with the accumulation
                launch_hf2x_C(prec, geo, Y, occR, pairs, targets, L.ntargets, nullptr, W, (float*)d_Wc.p, stream);
            });
        }

        // [3b6] divergence correction (:393-535): k on the weighted mesh (if_match) takes the
        // shared tables; a zero-weight k off it its own kVaux . G table with the opposite sign
        // of contribution 3 (:445-448, :519), and then the [3b3] and [3a3,b4] kVaux terms
        // (:537-627).  The states projected on are the occupied orbitals at k, or those of
        // tcx_set_zero_weight_kpoints at a zero-weight k.
        const double* div_r = (const double*)d_phi_occ.p;
        const double* div_l = (const double*)(bitc ? d_chi_occ.p : d_phi_occ.p);
        const DivEntry* djobs = (const DivEntry*)P.d_divjobs.p;
        const int maxst = std::max(1, max_div_states());
        const double* occR64 = nullptr;
        const double* chiR64 = nullptr;
        if (div_enabled) {
            std::vector<double> dcp(nk);
            for (int k = 0; k < nk; ++k) dcp[k] = 1.0 / std::sqrt(kPiHost * alpha) * NkV - sum_vaux[k];
            d_aux.ensure((size_t)nk * 8 + (size_t)nk * 16 + (size_t)P.n_app * maxst * 16 * 2);
            double* d_dcp = (double*)d_aux.p;
            double* t1 = d_dcp + nk;
            double* cR = t1 + 2 * nk;
            double* cQ = cR + 2 * (size_t)P.n_app * maxst;
            TCX_HIP_CHECK(hipMemcpy(d_dcp, dcp.data(), nk * sizeof(double), hipMemcpyHostToDevice));
            launch_b6(nG, tG, t1, d_dcp, zw_set ? (const int*)d_zw_match.p : nullptr, nk, geo, N, j00, j01, stream);
            launch_copy(tR, tG, g64, stream);
            launch_fft3d(TCX_PREC_FP64, +1, tR, 1, geo, 1.0, stream);
            ensure_occR(TCX_PREC_FP64);
            if (bitc) ensure_chiR(TCX_PREC_FP64);
            occR64 = (const double*)d_occR[TCX_PREC_FP64].p;
            chiR64 = bitc ? (const double*)d_chiR[TCX_PREC_FP64].p : occR64;
            const double* appR64;
            if (prec == TCX_PREC_FP64) {
                appR64 = (const double*)d_appR.p;
            } else {
                d_appR64.ensure((size_t)P.n_app * g64);
                orbitals_to_real(io.phi, (const OrbEntry*)P.d_appjobs.p, P.n_app, TCX_PREC_FP64, d_appR64.p);
                appR64 = (const double*)d_appR64.p;
            }
            if (!zw_set) {
                launch_b6_coef(io.phi, div_r, div_l, djobs, P.n_app, maxst, fill, t1, appR64, chiR64, occR64, tR, fb6,
                               1.0, N, io.out, cR, W, stream);
            } else {
                d_zwtab.ensure(2 * g64);
                double* tGk = (double*)d_zwtab.p;
                double* tRk = tGk + 2 * N;
                for (int k = 0; k < nk; ++k) {
                    const int j0 = (int)P.app_first[k], nj = nb[k];
                    if (nj == 0) continue;
                    const DivEntry* jk = djobs + j0;
                    double* Wk = W + 2 * (size_t)j0 * N;
                    const double* aRk = appR64 + 2 * (size_t)j0 * N;
                    if (zw_match[k]) {
                        launch_b6_coef(io.phi, div_r, div_l, jk, nj, maxst, fill, t1, aRk, chiR64, occR64, tR, fb6, 1.0,
                                       N, io.out, cR, Wk, stream);
                    } else {
                        launch_b6_tG_kv(nG, tGk, &zw_skv[3 * (size_t)k], geo, N, j00, j01, stream);
                        launch_copy(tRk, tGk, g64, stream);
                        launch_fft3d(TCX_PREC_FP64, +1, tRk, 1, geo, 1.0, stream);
                        launch_b6_coef(io.phi, div_r, div_l, jk, nj, maxst, fill, t1, aRk, chiR64, occR64, tRk, fb6,
                                       -1.0, N, io.out, cR, Wk, stream);
                    }
                    double kv3[3];
                    if (!jastrow_zero && tc3_kvaux(k, kv3) && zw_n[k] > 0)
                        launch_zw_b3(io.phi, div_l, jk, nj, zw_n[k], fill, chiR64, occR64, div_r, dnu, aRk, N, fb3, kv3,
                                     cR, cQ, qsum + 2 * (size_t)j0 * 3 * N, io.out, stream);
                }
            }
        }
        // [3a3] (:629-642)
        launch_a3(W, qsum, dnu, P.n_app, N, fa3, stream);

        // phase 2: [3b4] (:646-770), products conj(chi_qi) qsum_j,d in the working precision
        const void* qsumT = qsum;
        if (prec == TCX_PREC_FP32) {
            d_qsumT.ensure((size_t)3 * P.n_app * gT);
            launch_convert_d2f(qsum, (float*)d_qsumT.p, (long long)3 * P.n_app * N, stream);
            qsumT = d_qsumT.p;
        }
        {
            ProdChan ch[3];
            for (int d = 0; d < 3; ++d) ch[d] = chan(chiR, 1, 1, 0, 1, qsumT, 0, 3, d, 0);
            upload(d_chans, ch, sizeof(ch));
        }
        for (const Launch& L : S.launches) {
            const PairEntry* pairs = (const PairEntry*)S.d_pairs.p + L.pair0;
            const TargetEntry* targets = (const TargetEntry*)S.d_targets.p + L.target0;
            const int np = L.npairs;
            char* rec3 = (char*)d_rec.p;
            char* Z = rec3 + (size_t)3 * np * gT;
            launch_u2_table(prec, d_ktab.p, kq, L.kq0, L.nkq, geo, j00, j10, stream);
            timed(0, np, [&] {
                launch_products_z(prec, pairs, np, (const ProdChan*)d_chans.p, 3, rec3, geo, 1.0 / (double)N,
                                  stream);
            });
            timed(1, np, [&] {
                launch_xy_b4(prec, geo, rec3, Z, pairs, np, d_ktab.p, L.kq0, kq, fill, fb4, stream);
            });
            timed(2, np, [&] {
                launch_hf2x_C(prec, geo, Z, occR, pairs, targets, L.ntargets, nullptr, W, (float*)d_Wc.p, stream);
            });
        }
        // [3b4] kVaux term at a zero-weight k (:772-858), on the complete qsum
        if (div_enabled && zw_set && !jastrow_zero) {
            double* cA = (double*)d_aux.p + nk + 2 * (size_t)nk;
            for (int k = 0; k < nk; ++k) {
                double kv3[3];
                const int j0 = (int)P.app_first[k];
                if (nb[k] == 0 || zw_match[k] || zw_n[k] == 0 || !tc3_kvaux(k, kv3)) continue;
                launch_zw_b4(djobs + j0, nb[k], zw_n[k], fill, chiR64, div_r, qsum + 2 * (size_t)j0 * 3 * N, N, fb4,
                             kv3, cA, io.out, stream);
            }
        }
        finish_W(P, io, 1.0);
    }
    end_apply(io);
}

// ------------------------------------------------------------------ tc3a2a4b2b5
// The occupied-set phase of tc3a2a4b2b5 (:210-800): the q1-q2 sums Q2, Q4, Q5 in FP64,
// the [3a2] G=0 correction, the symmetry rotation, the working-precision copies Q4T/Q5T
// read by the apply phase, and the [3b2]+[3a2] local potential (real space, FP64) into
// c_a8.buf.
void Context::a8_occupied_phase(const std::vector<int>& nb, int prec)
{
    const bool bitc = method == TCX_METHOD_BITC;
    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8;
    const size_t g64 = (size_t)N * 16, gT = (size_t)N * es;
    const double NkV = num_kpoints * volume;
    const double base = 1.0 / (NkV * NkV);  // :34-48
    double fa2 = -0.5 * base, fb2 = base;
    if (nspin == 1) {
        fa2 /= 4.0;
        fb2 /= 4.0;
    }
    const JastrowPair j00 = jastrow_pair(0, 0), j01 = jastrow_pair(0, 1);
    const void* occR = d_occR[prec].p;
    const void* chiR = bitc ? d_chiR[prec].p : occR;
    ensure_occR(TCX_PREC_FP64);
    const double* occR64 = (const double*)d_occR[TCX_PREC_FP64].p;
    const double* fill = (const double*)d_fill.p;

    // (i) q1-q2 phase (:210-443): targets (q2,i2) x sources (q1,i1)
    const int nin = bitc ? 2 : 1;
    d_Q2.ensure((size_t)n_occ * g64);
    d_Q4.ensure((size_t)3 * n_occ * g64);
    launch_zero(d_Q2.p, (size_t)n_occ * g64, stream);
    launch_zero(d_Q4.p, (size_t)3 * n_occ * g64, stream);
    double* Q2 = (double*)d_Q2.p;
    double* Q4 = (double*)d_Q4.p;
    double* Q5 = nullptr;
    if (bitc) {
        d_Q5.ensure((size_t)3 * n_occ * g64);
        launch_zero(d_Q5.p, (size_t)3 * n_occ * g64, stream);
        Q5 = (double*)d_Q5.p;
    }
    {
        const PairSet& S = pair_set(nb, sym_set ? kPairsIrrOcc : kPairsOccOcc, 7 * nin, prec);
        d_rec.ensure((size_t)std::max(1, S.max_launch_pairs) * 7 * nin * gT);
        int max_rows = 1;
        for (const Launch& L : S.launches) max_rows = std::max(max_rows, L.nkq);
        d_ktab.ensure((size_t)max_rows * gT);
        ProdChan ch[2];
        ch[0] = chan(chiR, 1, 1, 0, 1, occR, 0, 1, 0, 0);  // conj(chi_q1) phi_q2 (:356-357)
        ch[1] = chan(occR, 1, 1, 0, 1, chiR, 0, 1, 0, 0);  // conj(phi_q1) chi_q2 (:358-363)
        upload(d_chans, ch, sizeof(ch));
        for (const Launch& L : S.launches) {
            const PairEntry* pairs = (const PairEntry*)S.d_pairs.p + L.pair0;
            const TargetEntry* targets = (const TargetEntry*)S.d_targets.p + L.target0;
            const int np = L.npairs;
            char* rec = (char*)d_rec.p;
            char* O = rec + (size_t)nin * np * gT;
            launch_u2_table(prec, d_ktab.p, (const double*)S.d_kq.p, L.kq0, L.nkq, geo, j00, j01, stream);
            timed(0, np, [&] {
                launch_products_z(prec, pairs, np, (const ProdChan*)d_chans.p, nin, rec, geo, 1.0 / (double)N,
                                  stream);
            });
            timed(1, np, [&] {
                launch_xy_a8c1(prec, geo, rec, nin, O, pairs, np, d_ktab.p, L.kq0, (const double*)S.d_kq.p,
                               stream);
            });
            timed(2, np, [&] {  // z pass of FFT_backward fused with the q-sums
                launch_za8_q(prec, geo, O, bitc ? 1 : 0, occR, chiR, pairs, targets, L.ntargets, fill, Q2, Q4,
                             bitc ? Q5 : Q4, stream);
            });
        }
    }

    // (ii) [3a2]: T = FFT_forward(qsum_3a2) (:486-494), G=0 correction (:496-573)
    launch_fft3d(TCX_PREC_FP64, -1, Q2, n_occ, geo, 1.0 / (double)N, stream);
    if (div_enabled) {
        if (bitc) ensure_chiR(TCX_PREC_FP64);
        const double* chiR64 = bitc ? (const double*)d_chiR[TCX_PREC_FP64].p : occR64;
        std::vector<double> dc2(nk);
        for (int k = 0; k < nk; ++k) dc2[k] = 1.0 / std::sqrt(kPiHost * alpha) * NkV - sum_vaux[k];
        d_aux.ensure((size_t)nk * 8);
        TCX_HIP_CHECK(hipMemcpy(d_aux.p, dc2.data(), nk * sizeof(double), hipMemcpyHostToDevice));
        const PairSet& S = pair_set(nb, sym_set ? kPairsIrrSameK : kPairsOccSameK, 2, TCX_PREC_FP64);
        d_rec.ensure((size_t)std::max(1, S.max_launch_pairs) * 2 * g64);
        ProdChan ch[2];
        ch[0] = chan(chiR64, 1, 1, 0, 1, occR64, 0, 1, 0, 0);  // conj(chi_q2,i1) phi_q2,i2 (:545-547)
        ch[1] = chan(chiR64, 0, 1, 0, 1, occR64, 1, 1, 0, 0);  // conj(chi_q2,i2) phi_q2,i1 (:548-550)
        upload(d_chans, ch, sizeof(ch));
        for (const Launch& L : S.launches) {
            const PairEntry* pairs = (const PairEntry*)S.d_pairs.p + L.pair0;
            const TargetEntry* targets = (const TargetEntry*)S.d_targets.p + L.target0;
            launch_products_fft(TCX_PREC_FP64, pairs, L.npairs, (const ProdChan*)d_chans.p, 2, d_rec.p, geo,
                                1.0 / (double)N, stream);
            launch_a8_div(Q2, (const double*)d_rec.p, pairs, targets, L.ntargets, fill, (const double*)d_aux.p,
                          geo, N, j00, j01, stream);
        }
    }
    // replicas of the irreducible q2: rotated sums (:576-765)
    if (sym_set) rotate_qsums(Q2, Q4, bitc ? Q5 : nullptr);
    // V_3a2 (G) and qsum_3b2 (r) (:637-641, :790-800)
    d_local.ensure(5 * g64);
    double* V3a2 = (double*)d_local.p;
    double* B2 = V3a2 + 2 * N;
    launch_a8_reduce(Q2, Q4, bitc ? Q5 : Q4, bitc ? 1 : 0, occR64, fill, n_occ, V3a2, B2, N, stream);
    d_Q4T.ensure((size_t)3 * n_occ * gT);
    d_Q5T.ensure((size_t)3 * n_occ * gT);
    launch_a8_q45(prec, Q4, bitc ? Q5 : Q4, bitc ? 1 : 0, d_Q4T.p, d_Q5T.p, 3 * n_occ, N, stream);
    // [3b2] (:922-992) and [3a2] (:995-1034) as one local potential X (real space)
    c_a8.buf.ensure(g64);
    double* X = (double*)c_a8.buf.p;
    launch_fft3d(TCX_PREC_FP64, -1, B2, 3, geo, 1.0 / (double)N, stream);
    launch_tc3b1_V(B2, nullptr, X, geo, j00, j01, stream);  // sum_d G_d (u(G,0,0) + u(G,0,1)) B2_d(G)
    launch_fft3d(TCX_PREC_FP64, +1, X, 1, geo, 1.0, stream);
    launch_fft3d(TCX_PREC_FP64, +1, V3a2, 1, geo, 1.0, stream);
    launch_axpby(X, V3a2, X, fb2, fa2, N, stream);
}

void Context::tc3a2a4b2b5(const int* nb_, const double* phi, double* out, int prec, unsigned flags)
{
    check_apply_args(nb_, phi, out, prec);
    require_one_spin("tc3a2a4b2b5");
    require(method != TCX_METHOD_HF, TCX_ERR_STATE, "tc3a2a4b2b5: the cell was set up for HF");
    require(jastrow_set, TCX_ERR_STATE, "tc3a2a4b2b5: call tcx_set_jastrow_rpa first");
    std::vector<int> nb(nb_, nb_ + nk);
    const Plan& P = get_plan(nb, 0);
    ApplyIO io = begin_apply(nb_, phi, out, flags, P);
    if (P.n_app > 0 && n_occ > 0 && !u_zero()) {
        const bool bitc = method == TCX_METHOD_BITC;
        const size_t es = prec == TCX_PREC_FP64 ? 16 : 8;
        const size_t g64 = (size_t)N * 16, gT = (size_t)N * es;
        const double NkV = num_kpoints * volume;
        const double base = 1.0 / (NkV * NkV);  // :34-48
        double fa4 = base, fb5 = -base;
        if (nspin == 1) {
            fa4 /= 4.0;
            fb5 /= 4.0;
        }
        const JastrowPair j00 = jastrow_pair(0, 0), j01 = jastrow_pair(0, 1);
        ensure_occR(prec);
        if (bitc) ensure_chiR(prec);
        const void* occR = d_occR[prec].p;
        const void* chiR = bitc ? d_chiR[prec].p : occR;
        const double* fill = (const double*)d_fill.p;
        // Everything up to the (k,j) x (q2,i2) apply depends on the occupied set only: the
        // q-sums in the working precision (Q4T, Q5T) and the [3b2]+[3a2] local potential are
        // kept for later applies until the states, Jastrow, divergence correction or
        // symmetry change.
        if (!(c_a8.valid && c_a8.prec == prec)) {
            c_a8.valid = false;
            a8_occupied_phase(nb, prec);
            c_a8.valid = true;
            c_a8.prec = prec;
        }
        // [3a4], [3b5] applied to (k,j) (:801-913): targets (k,j) x sources (q2,i2)
        d_appR.ensure((size_t)P.n_app * gT);
        orbitals_to_real(io.phi, (const OrbEntry*)P.d_appjobs.p, P.n_app, prec, d_appR.p);
        begin_W(P.n_app);
        {
            const PairSet& S = pair_set(nb, kPairsAppOcc, 8, prec);  // 4 input + 4 output channels
            d_rec.ensure((size_t)std::max(1, S.max_launch_pairs) * 8 * gT);
            int max_rows = 1;
            for (const Launch& L : S.launches) max_rows = std::max(max_rows, L.nkq);
            d_ktab.ensure((size_t)max_rows * gT);
            ProdChan ch[4];
            ch[0] = chan(chiR, 1, 1, 0, 1, d_appR.p, 0, 1, 0, 0);  // conj(chi_q2) phi_j (:849-850)
            for (int d = 0; d < 3; ++d) ch[1 + d] = chan(d_Q5T.p, 1, 3, d, 0, d_appR.p, 0, 1, 0, 0);  // (:867-868)
            upload(d_chans, ch, sizeof(ch));
            for (const Launch& L : S.launches) {
                const PairEntry* pairs = (const PairEntry*)S.d_pairs.p + L.pair0;
                const TargetEntry* targets = (const TargetEntry*)S.d_targets.p + L.target0;
                const int np = L.npairs;
                char* recA = (char*)d_rec.p;
                char* Zo = recA + (size_t)4 * np * gT;
                launch_u2_table(prec, d_ktab.p, (const double*)S.d_kq.p, L.kq0, L.nkq, geo, j00, j01, stream);
                timed(0, np, [&] {
                    launch_products_z(prec, pairs, np, (const ProdChan*)d_chans.p, 4, recA, geo, 1.0 / (double)N,
                                      stream);
                });
                timed(1, np, [&] {
                    launch_xy_a8c2(prec, geo, recA, Zo, pairs, np, d_ktab.p, L.kq0, (const double*)S.d_kq.p, fill, fa4,
                                   fb5, stream);
                });
                timed(2, np, [&] {
                    launch_za8_apply(prec, geo, Zo, occR, d_Q4T.p, pairs, targets, L.ntargets, (double*)d_W.p,
                                     (float*)d_Wc.p, stream);
                });
            }
        }
        finish_W(P, io, 1.0);
        // [3b2] (:922-992) and [3a2] (:995-1034) as one local potential
        apply_local(P, io, (const double*)c_a8.buf.p, 1.0);
        // [3a4,b5] kVaux block at zero-weight k (:1036-1299)
        if (div_enabled && zw_set) a8_kvaux(P, io, nb, fa4, fb5);
    }
    end_apply(io);
}

// Q2 (G space, one grid per occupied orbital) and Q4, Q5 (real space, three
// Cartesian grids each) at every replica q2 from those of its irreducible q2:
// FFT_forward, the U^+ map with the translation phase, FFT_backward for the
// vector sums (tc3a2a4b2b5.cpp:576-640, :644-745).  For TC, Q5 is not stored:
// the consumers form -conj(Q4) of the rotated Q4, which equals the rotated
// -conj(Q4) because the translations are commensurate with the grid.
void Context::rotate_qsums(double* Q2, double* Q4, double* Q5)
{
    if (!sym_jobs_valid) {
        std::vector<SymRotJob> jobs;
        for (int k = 0; k < nk; ++k) {
            const int s = sym_src[k];
            if (s == k) continue;
            require(nocc[k] == nocc[s], TCX_ERR_STATE,
                    "tc3a2a4b2b5: a symmetry replica holds a different number of occupied orbitals than its "
                    "irreducible k");
            for (int i = 0; i < nocc[k]; ++i) {
                SymRotJob J = sym_ops[k];
                J.dst = (int)occ_first[k] + i;
                J.src = (int)occ_first[s] + i;
                require(std::fabs(fill[J.dst] - fill[J.src]) <= 1e-12 * std::max(1.0, std::fabs(fill[J.src])),
                        TCX_ERR_STATE, "tc3a2a4b2b5: a symmetry replica has other fillings than its irreducible k");
                jobs.push_back(J);
            }
        }
        upload(d_symjobs, jobs.data(), jobs.size() * sizeof(SymRotJob));
        n_sym_jobs = (int)jobs.size();
        sym_jobs_valid = true;
    }
    if (n_sym_jobs == 0) return;
    const SymRotJob* jobs = (const SymRotJob*)d_symjobs.p;
    launch_sym_rotate(Q2, 1, jobs, n_sym_jobs, geo, N, stream);
    for (double* Q : {Q4, Q5}) {
        if (!Q) continue;
        launch_fft3d(TCX_PREC_FP64, -1, Q, 3 * n_occ, geo, 1.0 / (double)N, stream);
        launch_sym_rotate(Q, 3, jobs, n_sym_jobs, geo, N, stream);
        launch_fft3d(TCX_PREC_FP64, +1, Q, 3 * n_occ, geo, 1.0, stream);
    }
}

// ------------------------------------------------------------------ assembly terms
void Context::kinetic(const int* nb_, const double* phi, double* out, unsigned flags)
{
    // reads no orbital state: the cell, k-points and basis suffice
    require(basis_set, TCX_ERR_STATE, "kinetic: call tcx_set_basis first");
    require(nb_ && phi && out, TCX_ERR_ARG, "kinetic: null pointer");
    for (int k = 0; k < nk; ++k) require(nb_[k] >= 0, TCX_ERR_ARG, "kinetic: negative band count");
    std::vector<int> nb(nb_, nb_ + nk);
    const Plan& P = get_plan(nb, 0);
    ApplyIO io = begin_apply(nb_, phi, out, flags, P);
    if (P.n_app > 0)
        launch_kinetic(io.phi, (const int*)d_gindex.p, (const OrbEntry*)P.d_appjobs.p, P.n_app, max_npw,
                       (const double*)d_kcart.p, geo, io.out, stream);
    end_apply(io);
}

void Context::hf2h(const int* nb_, const double* phi, double* out, unsigned flags)
{
    check_apply_args(nb_, phi, out, TCX_PREC_FP64);
    std::vector<int> nb(nb_, nb_ + nk);
    const Plan& P = get_plan(nb, 0);
    ApplyIO io = begin_apply(nb_, phi, out, flags, P);
    if (P.n_app > 0 && n_occ > 0) {
        const size_t g = (size_t)N * 16;
        if (!c_hartree.valid) {
            d_local.ensure(2 * g);
            double* nG = (double*)d_local.p;
            c_hartree.buf.ensure(g);
            density_G(nG);
            launch_hartree(nG, (double*)c_hartree.buf.p, geo, N, stream);  // hf2h.cpp:28-38
            launch_fft3d(TCX_PREC_FP64, +1, c_hartree.buf.p, 1, geo, 1.0, stream);
            c_hartree.valid = true;
        }
        const double* V = (const double*)c_hartree.buf.p;
        apply_local(P, io, V, 1.0 / (num_kpoints * volume));  // two_body_factor, hf2h.cpp:23
    }
    end_apply(io);
}

void Context::tc2h(const int* nb_, const double* phi, double* out, unsigned flags)
{
    check_apply_args(nb_, phi, out, TCX_PREC_FP64);
    require(method != TCX_METHOD_HF, TCX_ERR_STATE, "tc2h: the cell was set up for HF");
    require(jastrow_set, TCX_ERR_STATE, "tc2h: call tcx_set_jastrow_rpa first");
    std::vector<int> nb(nb_, nb_ + nk);
    const Plan& P = get_plan(nb, 0);
    ApplyIO io = begin_apply(nb_, phi, out, flags, P);
    if (P.n_app > 0 && n_occ > 0) {
        const bool bitc = method == TCX_METHOD_BITC;
        const size_t g = (size_t)N * 16;
        ensure_occR(TCX_PREC_FP64);
        ensure_gradR(TCX_PREC_FP64);
        if (bitc) ensure_chiR(TCX_PREC_FP64);
        const double* chiR = (const double*)(bitc ? d_chiR[TCX_PREC_FP64].p : d_occR[TCX_PREC_FP64].p);
        const double* gradR = (const double*)d_gradR[TCX_PREC_FP64].p;
        // per spin s: V2_s (1 grid) and dnu_s (3 grids), 4 consecutive grids
        c_tc2h.buf.ensure((size_t)4 * nspin * g);
        auto V2_of = [&](int s) { return (double*)c_tc2h.buf.p + (size_t)s * 8 * N; };
        if (!c_tc2h.valid) {
            d_local.ensure((size_t)(2 + 3 * nspin) * g);
            double* nG = (double*)d_local.p;  // 2 grids: n_0, n_1
            double* pgp = nG + 4 * N;         // 3 grids per spin
            // phiq_gphiq of each spin from its occupied states (tc2h.cpp:60-127)
            for (int s = 0; s < nspin; ++s) {
                const long long b0 = occ_first[(size_t)s * nk_phys], b1 = occ_first[(size_t)(s + 1) * nk_phys];
                launch_tc2h_V(chiR + 2 * b0 * N, gradR + 6 * b0 * N, (const double*)d_fill.p + b0, (int)(b1 - b0),
                              pgp + (size_t)s * 6 * N, geo, N, stream);
            }
            launch_fft3d(TCX_PREC_FP64, -1, pgp, 3 * nspin, geo, 1.0 / (double)N, stream);
            density_G(nG);
            const double* pgp1 = nspin == 1 ? pgp : pgp + 6 * N;  // one spin: jspin_ref = 0
            const double pf = nspin == 1 ? 0.5 : 1.0;             // 1 / spin_factor (tc2h.cpp:144)
            for (int s = 0; s < nspin; ++s) {
                const JastrowPair js0 = jastrow_pair(s, 0), js1 = jastrow_pair(s, 1);
                launch_tc2h_combine(nG, pgp, pgp1, pf, V2_of(s), geo, N, js0, js1, stream);  // tc2h.cpp:146-171
                launch_dnu(nG, V2_of(s) + 2 * N, geo, js0, js1, stream);
            }
            launch_fft3d(TCX_PREC_FP64, +1, V2_of(0), 4 * nspin, geo, 1.0, stream);
            c_tc2h.valid = true;
        }
        // per band: FFT_forward[V2_s phi_j - sum_d gphi_j,d dnu_s,d] (tc2h.cpp:178-234), the
        // bands of spin s being the jobs of its k-blocks
        const double factor = 1.0 / (num_kpoints * volume);  // tc2h.cpp:32
        const int chunk = (int)std::max<size_t>(1, std::min<size_t>(P.n_app, ((size_t)1 << 30) / (4 * g)));
        d_W.ensure((size_t)chunk * g);
        d_appGrad.ensure((size_t)3 * chunk * g);
        const OrbEntry* jobs = (const OrbEntry*)P.d_appjobs.p;
        const OrbEntry* gjobs = (const OrbEntry*)P.d_appgradjobs.p;
        for (int s = 0; s < nspin; ++s) {
            const int a0 = (int)P.app_first[(size_t)s * nk_phys], a1 = (int)P.app_first[(size_t)(s + 1) * nk_phys];
            const double* V2 = V2_of(s);
            for (int c0 = a0; c0 < a1; c0 += chunk) {
                const int nc = std::min(chunk, a1 - c0);
                double* Wc = (double*)d_W.p;
                orbitals_to_real(io.phi, jobs + c0, nc, TCX_PREC_FP64, Wc);
                orbitals_to_real(io.phi, gjobs + 3 * c0, 3 * nc, TCX_PREC_FP64, d_appGrad.p);
                launch_tc2h_apply(Wc, (const double*)d_appGrad.p, V2, V2 + 2 * N, nc, N, stream);
                launch_fft3d(TCX_PREC_FP64, -1, Wc, nc, geo, 1.0 / (double)N, stream);
                launch_gather_add(Wc, (const int*)d_gindex.p, jobs + c0, nc, max_npw, io.out, N, factor, stream);
            }
        }
    }
    end_apply(io);
}

// calc_hamiltonian_all.cpp:28-159: zero H1/H2/H3 phi, then kinetic into H1, the
// two-body terms into H2 and (TC/BITC) the three-body terms into H3.  Pseudopotentials
// are not part of this engine (homogeneous-electron-gas cells, :58-66).
void Context::apply_all(const int* nb_, const double* phi, double* h1, double* h2, double* h3, int prec,
                        unsigned flags)
{
    check_apply_args(nb_, phi, h1, prec);
    require(h2 != nullptr, TCX_ERR_ARG, "all: null h2phi");
    const bool tc = method != TCX_METHOD_HF;
    require(!tc || h3 != nullptr, TCX_ERR_ARG, "all: null h3phi for TC/BITC");
    long long ncoef = 0;
    for (int k = 0; k < nk; ++k) ncoef += (long long)nb_[k] * npw[k];
    const size_t bytes = (size_t)ncoef * 16;
    double* outs[3] = {h1, h2, tc ? h3 : nullptr};
    for (double* o : outs) {
        if (!o || bytes == 0) continue;
        if (flags & TCX_FLAG_DEVICE_PTRS)
            launch_zero(o, bytes, stream);
        else
            std::memset(o, 0, bytes);
    }
    kinetic(nb_, phi, h1, flags);
    if (!tc) {
        hf2h(nb_, phi, h2, flags);
        hf2x(nb_, phi, h2, prec, flags);
    } else {
        tc2h(nb_, phi, h2, flags);
        tc2x(nb_, phi, h2, prec, flags);
        tc3_local(0, nb_, phi, h3, prec, flags);
        tc3_local(1, nb_, phi, h3, prec, flags);
        tc3a2a4b2b5(nb_, phi, h3, prec, flags);
        tc3a3b3b4b6(nb_, phi, h3, prec, flags);
    }
}

}  // namespace tcx
