// tcx_zeroweight.cpp -- zero-weight k-points: BAND k-points and zero-weight SCF
// k-points (Kpoints::kweight_scf < 1e-8).
//
// The reference runs BAND mode as a second k list (kvectors_band) whose bands
// see the SCF occupied orbitals as sources; zero-weight SCF k-points ("fake
// SCF") are k-points of the SCF list that are left out of every q-sum.  Both
// are one k list here: tcx_set_zero_weight_kpoints marks the zero-weight
// entries, which hold no occupied orbitals (nocc = 0), and gives for each of
// them the states its divergence terms project on (phik_band or phik_scf there,
// with their Kpoints::return_band_filling values) and the vector sums
// sum_of_kVaux of every k (potentials_coulomb.cpp:163-254).  Those states are
// appended behind the occupied orbitals in the orbital arrays, so divergence
// jobs address them like occupied ones while the pair sources stay the
// occupied orbitals only.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tcx_terms.hpp"

namespace tcx {

void Context::set_zero_weight(const double* kweight, const double* skv, const int* nstate, const double* filling,
                              const double* phi, const double* chi)
{
    require(!g_capturing, TCX_ERR_STATE, "set_zero_weight_kpoints: not capturable");
    require(states_set, TCX_ERR_STATE, "set_zero_weight_kpoints: call tcx_set_states first");
    require_one_spin("set_zero_weight_kpoints");
    require(!sym_set, TCX_ERR_UNSUPPORTED, "set_zero_weight_kpoints: not combined with a k-mesh symmetry");
    require(kweight && skv && nstate, TCX_ERR_ARG, "set_zero_weight_kpoints: null pointer");
    const bool bitc = method == TCX_METHOD_BITC;
    // validate into locals
    std::vector<char> zk(nk, 0);
    std::vector<int> n(nstate, nstate + nk);
    int nweighted = 0, total = 0;
    long long tcoef = 0;
    for (int k = 0; k < nk; ++k) {
        require(std::isfinite(kweight[k]) && kweight[k] >= 0.0, TCX_ERR_ARG,
                "set_zero_weight_kpoints: k weights must be finite and non-negative");
        zk[k] = kweight[k] < 1e-8;
        if (!zk[k]) ++nweighted;
        require(!zk[k] || nocc[k] == 0, TCX_ERR_ARG,
                "set_zero_weight_kpoints: a zero-weight k-point holds no occupied orbitals (nocc = 0)");
        require(n[k] >= 0 && n[k] <= 512, TCX_ERR_ARG, "set_zero_weight_kpoints: 0 <= states per k <= 512");
        require(zk[k] || n[k] == 0, TCX_ERR_ARG, "set_zero_weight_kpoints: states are given at zero-weight k only");
        for (int d = 0; d < 3; ++d)
            require(std::isfinite(skv[3 * k + d]), TCX_ERR_ARG, "set_zero_weight_kpoints: non-finite sum_of_kvaux");
        total += n[k];
        tcoef += (long long)n[k] * npw[k];
    }
    require(nweighted > 0, TCX_ERR_ARG, "set_zero_weight_kpoints: no weighted k-point");
    require(total == 0 || (filling && phi), TCX_ERR_ARG, "set_zero_weight_kpoints: null states");
    require(total == 0 || !bitc || chi, TCX_ERR_ARG, "set_zero_weight_kpoints: BITC needs the left states");
    for (int i = 0; i < total; ++i)
        require(std::isfinite(filling[i]) && filling[i] >= 0.0, TCX_ERR_ARG,
                "set_zero_weight_kpoints: fillings must be finite and non-negative");
    for (long long i = 0; i < 2 * tcoef; ++i) {
        require(std::isfinite(phi[i]), TCX_ERR_ARG, "set_zero_weight_kpoints: non-finite state coefficient");
        if (bitc) require(std::isfinite(chi[i]), TCX_ERR_ARG, "set_zero_weight_kpoints: non-finite state coefficient");
    }
    // k on the weighted mesh (modulo reciprocal lattice vectors): tc3a3b3b4b6.cpp:396-410
    std::vector<int> match(nk, 0);
    for (int k = 0; k < nk; ++k)
        for (int q = 0; q < nk && !match[k]; ++q) {
            if (zk[q]) continue;
            double s = 0.0;
            for (int d = 0; d < 3; ++d) {
                double x = kfrac[3 * k + d] - kfrac[3 * q + d];
                while (x < -1e-8) x += 1.0;
                while (x > 1.0 - 1e-8) x -= 1.0;
                s += x * x;
            }
            match[k] = s < 1e-8;
        }

    // the orbital arrays become [occupied | states of the zero-weight k], grids likewise
    const long long ocoef = occ_coef[nk];
    const size_t cbytes = (size_t)(ocoef + tcoef) * 16;
    DBuf nphi, nchi, nfill;
    nphi.ensure(std::max<size_t>(cbytes, 16));
    if (ocoef) TCX_HIP_CHECK(hipMemcpy(nphi.p, d_phi_occ.p, (size_t)ocoef * 16, hipMemcpyDeviceToDevice));
    if (tcoef) TCX_HIP_CHECK(hipMemcpy((char*)nphi.p + ocoef * 16, phi, (size_t)tcoef * 16, hipMemcpyHostToDevice));
    if (bitc) {
        nchi.ensure(std::max<size_t>(cbytes, 16));
        if (ocoef) TCX_HIP_CHECK(hipMemcpy(nchi.p, d_chi_occ.p, (size_t)ocoef * 16, hipMemcpyDeviceToDevice));
        if (tcoef) TCX_HIP_CHECK(hipMemcpy((char*)nchi.p + ocoef * 16, chi, (size_t)tcoef * 16, hipMemcpyHostToDevice));
    }
    nfill.ensure(std::max<size_t>((size_t)(n_occ + total) * 8, 16));
    if (n_occ) TCX_HIP_CHECK(hipMemcpy(nfill.p, d_fill.p, (size_t)n_occ * 8, hipMemcpyDeviceToDevice));
    if (total)
        TCX_HIP_CHECK(hipMemcpy((char*)nfill.p + (size_t)n_occ * 8, filling, (size_t)total * 8, hipMemcpyHostToDevice));
    // one scatter job per orbital and state (occupied first, as tcx_set_states lays them out)
    std::vector<OrbEntry> jobs;
    std::vector<int> first(nk, 0);
    std::vector<long long> coef(nk, 0);
    for (int k = 0; k < nk; ++k)
        for (int i = 0; i < nocc[k]; ++i)
            jobs.push_back({occ_coef[k] + (long long)i * npw[k], npw[k], (int)pw_off[k], k, -1});
    long long c = ocoef;
    int g = n_occ;
    for (int k = 0; k < nk; ++k) {
        first[k] = g;
        coef[k] = c;
        for (int b = 0; b < n[k]; ++b) jobs.push_back({c + (long long)b * npw[k], npw[k], (int)pw_off[k], k, -1});
        g += n[k];
        c += (long long)n[k] * npw[k];
    }
    std::vector<double> kv(skv, skv + 3 * nk);

    // commit
    invalidate_local();
    occR_valid[0] = occR_valid[1] = false;
    chiR_valid[0] = chiR_valid[1] = false;
    plans.clear();  // divergence jobs address the new states
    d_phi_occ = std::move(nphi);
    if (bitc) d_chi_occ = std::move(nchi);
    d_fill = std::move(nfill);
    upload(d_occjobs, jobs.data(), jobs.size() * sizeof(OrbEntry));
    std::vector<int> pwo(nk);
    for (int k = 0; k < nk; ++k) pwo[k] = (int)pw_off[k];
    upload(d_pwoff, pwo.data(), pwo.size() * sizeof(int));
    upload(d_zw_match, match.data(), match.size() * sizeof(int));
    zw_k = std::move(zk);
    zw_match = std::move(match);
    zw_skv = std::move(kv);
    zw_n = std::move(n);
    zw_first = std::move(first);
    zw_coef = std::move(coef);
    n_zw = total;
    zw_set = true;
}

void Context::clear_zero_weight()
{
    zw_set = false;
    zw_k.clear();
    zw_match.clear();
    zw_skv.clear();
    zw_n.clear();
    zw_first.clear();
    zw_coef.clear();
    n_zw = 0;
}

// kVaux of tc2x at k (tc2x.cpp:117-126): sum_of_kVaux / (Nk V) (/2 for one spin) A_long[s][s],
// zero where its squared norm is at most 1e-8 (the reference's branch)
void Context::tc2x_kvaux(std::vector<double>& kv) const
{
    kv.assign(3 * (size_t)nk, 0.0);
    if (!zw_set || jastrow_zero) return;
    const double s = A[0] / (num_kpoints * volume) / (nspin == 1 ? 2.0 : 1.0);
    for (int k = 0; k < nk; ++k) {
        double n2 = 0.0, v[3];
        for (int d = 0; d < 3; ++d) {
            v[d] = zw_skv[3 * k + d] * s;
            n2 += v[d] * v[d];
        }
        if (n2 > 1e-8)
            for (int d = 0; d < 3; ++d) kv[3 * k + d] = v[d];
    }
}

// kVaux of the three-body terms at k (tc3a3b3b4b6.cpp:539-543, tc3a2a4b2b5.cpp:1064-1068):
// -sum_of_kVaux A_long[s][s]; false where its squared norm is at most 1e-8
bool Context::tc3_kvaux(int k, double kv[3]) const
{
    if (!zw_set) return false;
    double n2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        kv[d] = -zw_skv[3 * k + d] * A[0];
        n2 += kv[d] * kv[d];
    }
    return n2 > 1e-8;
}

// [3a4,b5] kVaux block of tc3a2a4b2b5 at every zero-weight k off the mesh with applied
// bands (:1036-1299), FP64.  For each state b at k and occupied (q1, i1) the orbitals of
// :1154-1227 are pair transforms with the kernel du = (kv . (k - q1 + G)) u(k - q1 + G):
//   [3a4] Wa_b = sum f1 FFT_backward(FFT_forward(phi_b conj(chi_q1i1)) du) phi_q1i1,
//   [3b5] Wb_b = sum f1 FFT_backward(FFT_forward(chi_b conj(phi_q1i1)) du) chi_q1i1
// (orbital2 = -Wb; TC: Wb = Wa), which the hf2x passes A, B, C carry with du/N as their
// kernel table.  Both go to G space, and k_zw_a4b5_apply adds them and the |kVaux|^2
// terms to every applied band at k.  The reference adds those |kVaux|^2 terms inside its
// (q1) loop, so once per q1 with occupied orbitals (num_mpi_processes = 1 here).
void Context::a8_kvaux(const Plan& P, const ApplyIO& io, const std::vector<int>& nb, double fa4, double fb5)
{
    const bool bitc = method == TCX_METHOD_BITC;
    const size_t g64 = (size_t)N * 16;
    ensure_occR(TCX_PREC_FP64);
    if (bitc) ensure_chiR(TCX_PREC_FP64);
    const double* occR64 = (const double*)d_occR[TCX_PREC_FP64].p;
    const double* chiR64 = bitc ? (const double*)d_chiR[TCX_PREC_FP64].p : occR64;
    const double* div_r = (const double*)d_phi_occ.p;
    const double* div_l = (const double*)(bitc ? d_chi_occ.p : d_phi_occ.p);
    const JastrowPair j00 = jastrow_pair(0, 0);  // u(k - q1 + G, s, s) (:1109)
    int nq = 0;
    for (int q = 0; q < nk; ++q)
        if (!zw_k[q] && nocc[q] > 0) ++nq;
    const std::vector<PairItem> src = occupied_items();
    for (int k = 0; k < nk; ++k) {
        double kv[3];
        if (nb[k] == 0 || zw_n[k] == 0 || zw_match[k] || !tc3_kvaux(k, kv) || src.empty()) continue;
        const int ns = zw_n[k];
        const double kv2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        // targets: the states of k (grids zw_first[k] .. + ns), all occupied sources
        std::vector<PairItem> tg;
        for (int b = 0; b < ns; ++b) tg.push_back({k, b});
        const long sc = std::max<long>(1, std::min<long>({(long)src.size(), 65535L / ns,
                                                          (long)(((size_t)2 << 30) / ((size_t)ns * g64))}));
        PairSet S;
        build_pairs(S, tg, src, ns, sc, kcart, false);
        d_zwtab.ensure((size_t)(bitc ? 2 : 1) * ns * g64);
        double* Wa = (double*)d_zwtab.p;
        double* Wb = bitc ? Wa + 2 * (size_t)ns * N : Wa;
        launch_zero(Wa, (size_t)(bitc ? 2 : 1) * ns * g64, stream);
        d_rec.ensure((size_t)std::max(1, S.max_launch_pairs) * g64);
        int max_rows = 1;
        for (const Launch& L : S.launches) max_rows = std::max(max_rows, L.nkq);
        d_ktab.ensure((size_t)max_rows * N * 8);
        const double* phib = occR64 + 2 * (size_t)zw_first[k] * N;  // phi_b of k
        const double* chib = chiR64 + 2 * (size_t)zw_first[k] * N;  // chi_b of k
        for (const Launch& L : S.launches) {
            const PairEntry* pairs = (const PairEntry*)S.d_pairs.p + L.pair0;
            const TargetEntry* targets = (const TargetEntry*)S.d_targets.p + L.target0;
            launch_duk_table((double*)d_ktab.p, (const double*)S.d_kq.p, L.kq0, L.nkq, geo, j00, kv, stream);
            // [3a4]: phi_b conj(chi_q1i1) -> x du -> x f1 phi_q1i1, summed into Wa_b
            launch_hf2x_A(TCX_PREC_FP64, geo, phib, chiR64, pairs, L.npairs, d_rec.p, stream);
            launch_hf2x_B(TCX_PREC_FP64, geo, d_rec.p, pairs, L.npairs, d_ktab.p, L.kq0, stream);
            launch_hf2x_C(TCX_PREC_FP64, geo, d_rec.p, occR64, pairs, targets, L.ntargets, (const double*)d_fill.p, Wa,
                          nullptr, stream);
            if (bitc) {  // [3b5]: chi_b conj(phi_q1i1) -> x du -> x f1 chi_q1i1, summed into Wb_b
                launch_hf2x_A(TCX_PREC_FP64, geo, chib, occR64, pairs, L.npairs, d_rec.p, stream);
                launch_hf2x_B(TCX_PREC_FP64, geo, d_rec.p, pairs, L.npairs, d_ktab.p, L.kq0, stream);
                launch_hf2x_C(TCX_PREC_FP64, geo, d_rec.p, chiR64, pairs, targets, L.ntargets,
                              (const double*)d_fill.p, Wb, nullptr, stream);
            }
        }
        launch_fft3d(TCX_PREC_FP64, -1, Wa, ns, geo, 1.0 / (double)N, stream);
        if (bitc) launch_fft3d(TCX_PREC_FP64, -1, Wb, ns, geo, 1.0 / (double)N, stream);
        const int j0 = (int)P.app_first[k];
        launch_zw_a4b5_apply(io.phi, div_r, div_l, (const DivEntry*)P.d_divjobs.p + j0, nb[k], ns,
                             (const double*)d_fill.p, Wa, Wb, (const int*)d_gindex.p + pw_off[k], N, fa4, fb5,
                             kv2 * nq, io.out, stream);
        TCX_HIP_CHECK(hipStreamSynchronize(stream));  // S (its tables) leaves scope
    }
}

}  // namespace tcx

// ======================================================================= C ABI
extern "C" int tcx_sum_of_vaux_weighted(const double recip[9], double volume, int nk, const double* kfrac,
                                        const double* kweight, double* alpha_out, double* sum_out, double* ksum_out)
{
    using tcx::kPiHost;
    if (!recip || !kfrac || !kweight || !alpha_out || !sum_out || nk <= 0 || !(volume > 0)) return TCX_ERR_ARG;
    // Potentials::set_sum_of_Vaux_each and set_sum_of_kVaux_each (src/potentials_coulomb.cpp:60-254)
    // with the q-sums over the weighted k only (:86, :198); each shell sum runs until that
    // shell's value (squared norm for the vector sums) drops below 1e-20.
    double alpha = std::cbrt(volume) / (2.0 * kPiHost);
    alpha = alpha * alpha;
    auto cart = [&](const double m[3], double c[3]) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) {
            c[d] = recip[d] * m[0] + recip[3 + d] * m[1] + recip[6 + d] * m[2];
            s += c[d] * c[d];
        }
        return s;
    };
    bool any = false;
    for (int q = 0; q < nk; ++q) {
        if (!std::isfinite(kweight[q]) || kweight[q] < 0.0) return TCX_ERR_ARG;
        any = any || kweight[q] >= 1e-8;
    }
    if (!any) return TCX_ERR_ARG;
    for (int ik = 0; ik < nk; ++ik) {
        double sum = 0.0, vs[3] = {0.0, 0.0, 0.0};
        bool match_any = false;
        for (int iq = 0; iq < nk; ++iq) {
            if (kweight[iq] < 1e-8) continue;
            double kq[3];
            for (int d = 0; d < 3; ++d) {
                kq[d] = kfrac[3 * ik + d] - kfrac[3 * iq + d];
                while (kq[d] < -1e-8) kq[d] += 1.0;
                while (kq[d] > 1.0 - 1e-8) kq[d] -= 1.0;
            }
            const bool match = kq[0] * kq[0] + kq[1] * kq[1] + kq[2] * kq[2] < 1e-8;
            match_any = match_any || match;
            if (!match) {
                double c[3];
                const double k2 = cart(kq, c);
                const double w = std::exp(-alpha * k2) / k2;
                sum += w;
                for (int d = 0; d < 3; ++d) vs[d] += w * c[d];
            }
            bool done_s = false, done_v = false;
            for (int shell = 1; !(done_s && done_v); ++shell) {
                double on = 0.0, von[3] = {0.0, 0.0, 0.0};
                for (int gx = -shell; gx <= shell; ++gx)
                    for (int gy = -shell; gy <= shell; ++gy)
                        for (int gz = -shell; gz <= shell; ++gz) {
                            if (std::abs(gx) != shell && std::abs(gy) != shell && std::abs(gz) != shell) continue;
                            const double m[3] = {kq[0] + gx, kq[1] + gy, kq[2] + gz};
                            double c[3];
                            const double g2 = cart(m, c);
                            const double w = std::exp(-alpha * g2) / g2;
                            on += w;
                            for (int d = 0; d < 3; ++d) von[d] += w * c[d];
                        }
                if (!done_s) {
                    sum += on;
                    done_s = on < 1e-20;
                }
                if (!done_v) {
                    for (int d = 0; d < 3; ++d) vs[d] += von[d];
                    done_v = von[0] * von[0] + von[1] * von[1] + von[2] * von[2] < 1e-20;
                }
            }
        }
        sum *= 4.0 * kPiHost;
        if (match_any) sum -= 4.0 * kPiHost * alpha;
        sum_out[ik] = sum;
        if (ksum_out)
            for (int d = 0; d < 3; ++d) ksum_out[3 * ik + d] = match_any ? 0.0 : 4.0 * kPiHost * vs[d];
    }
    *alpha_out = alpha;
    return TCX_OK;
}
