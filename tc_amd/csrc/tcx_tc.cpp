// tcx_tc.cpp -- TC/BITC terms on the context: tc2x (pair pipeline) and the
// local three-body terms tc3a1 / tc3b1.
#include <algorithm>
#include <cmath>

#include "tcx_terms.hpp"

namespace tcx {

void Context::tc2x(const int* nb_, const double* phi, double* out, int prec, unsigned flags)
{
    check_apply_args(nb_, phi, out, prec);
    require(method != TCX_METHOD_HF, TCX_ERR_STATE, "tc2x: the cell was set up for HF");
    require(jastrow_set, TCX_ERR_STATE, "tc2x: call tcx_set_jastrow_rpa first");
    std::vector<int> nb(nb_, nb_ + nk);
    const Plan& P = get_plan(nb, 1);
    ApplyIO io = begin_apply(nb_, phi, out, flags, P);
    if (P.n_app > 0) {
        const bool bitc = method == TCX_METHOD_BITC;
        ensure_occR(prec);
        ensure_gradR(prec);
        if (bitc) ensure_chiR(prec);
        const void* chiR = bitc ? d_chiR[prec].p : d_occR[prec].p;  // bra orbital (tc2x.cpp:306)
        const size_t es = prec == TCX_PREC_FP64 ? 16 : 8;
        d_appR.ensure((size_t)P.n_app * N * es);
        d_appGrad.ensure((size_t)3 * P.n_app * N * es);
        timed(3, 0, [&] {
            orbitals_to_real(io.phi, (const OrbEntry*)P.d_appjobs.p, P.n_app, prec, d_appR.p);
            orbitals_to_real(io.phi, (const OrbEntry*)P.d_appgradjobs.p, 3 * P.n_app, prec, d_appGrad.p);
        });
        begin_W(P.n_app);
        d_pairbuf.ensure((size_t)std::max(1, P.max_launch_pairs) * 4 * N * es);
        int max_rows = 1;
        for (const Launch& L : P.launches) max_rows = std::max(max_rows, L.nkq);
        d_ktab.ensure((size_t)max_rows * N * es);
        for (const Launch& L : P.launches) {
            const PairEntry* pairs = (const PairEntry*)P.d_pairs.p + L.pair0;
            const TargetEntry* targets = (const TargetEntry*)P.d_targets.p + L.target0;
            const JastrowPair jp = jastrow_pair(L.spin, L.spin);  // tc2x.cpp:269-271 (ispin, ispin)
            launch_tc_table(prec, d_ktab.p, (const double*)P.d_kq.p, L.kq0, L.nkq, geo, jp, stream);
            timed(0, L.npairs, [&] {
                launch_tc2x_A(prec, geo, d_appR.p, d_appGrad.p, chiR, pairs, L.npairs, d_pairbuf.p, stream);
            });
            timed(1, L.npairs, [&] {
                launch_tc2x_B(prec, geo, d_pairbuf.p, pairs, L.npairs, d_ktab.p, L.kq0, (const double*)P.d_kq.p,
                              stream);
            });
            timed(2, L.npairs, [&] {
                launch_tc2x_C(prec, geo, d_pairbuf.p, d_occR[prec].p, d_gradR[prec].p, pairs, targets, L.ntargets,
                              (const double*)d_fill.p, (double*)d_W.p, (float*)d_Wc.p, stream);
            });
        }
        // two_body_factor, src/calc_hamiltonian_tc2x.cpp:32-33
        double factor = -1.0 / (num_kpoints * volume);
        if (nspin == 1) factor /= 2.0;
        finish_W(P, io, factor);
        std::vector<double> kv;
        tc2x_kvaux(kv);
        if (div_enabled && std::any_of(kv.begin(), kv.end(), [](double x) { return x != 0.0; })) {
            // zero-weight k: [2a_x] with the [2b_x1], [2b_x2] kVaux terms (tc2x.cpp:117-222, :483-508)
            upload(d_zw_kv, kv.data(), kv.size() * sizeof(double));
            launch_divcorr_kv(io.phi, (const double*)d_phi_occ.p, (const double*)(bitc ? d_chi_occ.p : d_phi_occ.p),
                              (const DivEntry*)P.d_divjobs.p, P.n_app, max_div_states(), (const double*)d_fill.p,
                              (const double*)d_dc.p, (const double*)d_zw_kv.p, (const int*)d_gindex.p,
                              (const int*)d_pwoff.p, (const double*)d_kcart.p, geo, io.out, stream);
        } else {
            apply_divcorr(P, io, (const double*)(bitc ? d_chi_occ.p : d_phi_occ.p));
        }
    }
    end_apply(io);
}

void Context::ensure_density()
{
    if (density_valid) return;
    const bool bitc = method == TCX_METHOD_BITC;
    ensure_occR(TCX_PREC_FP64);
    if (bitc) ensure_chiR(TCX_PREC_FP64);
    d_density.ensure((size_t)2 * N * 16);
    if (nspin == 1) {
        launch_density((const double*)d_occR[TCX_PREC_FP64].p,
                       bitc ? (const double*)d_chiR[TCX_PREC_FP64].p : nullptr, (const double*)d_fill.p, n_occ, N,
                       (double*)d_density.p, 1, stream);
    } else {
        // density_[s] from the occupied states of spin s (the k-blocks s nk_phys .. (s+1) nk_phys - 1)
        for (int s = 0; s < 2; ++s) {
            const long long b0 = occ_first[(size_t)s * nk_phys], b1 = occ_first[(size_t)(s + 1) * nk_phys];
            launch_density((const double*)d_occR[TCX_PREC_FP64].p + 2 * b0 * N,
                           bitc ? (const double*)d_chiR[TCX_PREC_FP64].p + 2 * b0 * N : nullptr,
                           (const double*)d_fill.p + b0, (int)(b1 - b0), N, (double*)d_density.p + 2 * (size_t)s * N,
                           2, stream);
        }
    }
    density_valid = true;
}

// which = 0: tc3a1, 1: tc3b1.  Always FP64 (O(Nk Nb) transforms, no pair loop).
void Context::tc3_local(int which, const int* nb_, const double* phi, double* out, int prec, unsigned flags)
{
    check_apply_args(nb_, phi, out, prec);
    require(method != TCX_METHOD_HF, TCX_ERR_STATE, "tc3: the cell was set up for HF");
    require(jastrow_set, TCX_ERR_STATE, "tc3: call tcx_set_jastrow_rpa first");
    std::vector<int> nb(nb_, nb_ + nk);
    const Plan& P = get_plan(nb, 0);  // only the per-band job tables are used
    ApplyIO io = begin_apply(nb_, phi, out, flags, P);
    if (P.n_app > 0) {
        const size_t g = (size_t)N * 16;
        const double NkV = num_kpoints * volume;
        // tc3a1.cpp:26-28, tc3b1.cpp:26-28
        const double factor = which == 0 ? 0.5 / (NkV * NkV) : -1.0 / (NkV * NkV);
        LocalCache& C = which == 0 ? c_tc3a1 : c_tc3b1;
        C.buf.ensure((size_t)nspin * g);  // the local potential of each spin s1, real space
        if (!C.valid) {
            ensure_density();
            d_work.ensure((size_t)(2 + 3 * nspin) * g);
            double* nG = (double*)d_work.p;                    // 2 grids: n_0, n_1
            double* dnu = (double*)((char*)d_work.p + 2 * g);  // 3 grids per spin
            TCX_HIP_CHECK(hipMemcpyAsync(nG, d_density.p, 2 * g, hipMemcpyDeviceToDevice, stream));
            launch_fft3d(TCX_PREC_FP64, -1, nG, 2, geo, 1.0 / (double)N, stream);
            // dnu_s = sum_s' u(s, s') G n_s' (tc3a1.cpp:54-75, tc3b1.cpp:52-71)
            for (int s = 0; s < nspin; ++s)
                launch_dnu(nG, dnu + (size_t)s * 6 * N, geo, jastrow_pair(s, 0), jastrow_pair(s, 1), stream);
            launch_fft3d(TCX_PREC_FP64, +1, dnu, 3 * nspin, geo, 1.0, stream);
            for (int s = 0; s < nspin; ++s) {
                double* Vc = (double*)C.buf.p + (size_t)s * 2 * N;
                double* ds = dnu + (size_t)s * 6 * N;
                if (which == 0) {
                    launch_sum_squares(ds, Vc, N, stream);  // tc3a1.cpp:78-84
                } else {
                    // w_s = FFT_forward[n_s dnu_s] (tc3b1.cpp:73-81)
                    launch_mul_shared(ds, 3, (const double*)d_density.p + (size_t)s * 2 * N, N, stream);
                }
            }
            if (which == 1) {
                launch_fft3d(TCX_PREC_FP64, -1, dnu, 3 * nspin, geo, 1.0 / (double)N, stream);
                // V_s1 = sum_s2 u(s1, s2) G.w_s2 (tc3b1.cpp:83-100; one spin: s2_ref = 0)
                for (int s = 0; s < nspin; ++s) {
                    double* Vc = (double*)C.buf.p + (size_t)s * 2 * N;
                    launch_tc3b1_V(dnu, nspin == 1 ? nullptr : dnu + 6 * N, Vc, geo, jastrow_pair(s, 0),
                                   jastrow_pair(s, 1), stream);
                }
                launch_fft3d(TCX_PREC_FP64, +1, (double*)C.buf.p, nspin, geo, 1.0, stream);
            }
            C.valid = true;
        }
        // H3 phi_kj += factor * FFT_forward[V_s phi_kj] on the sphere of k (tc3a1.cpp:86-116), the
        // bands of spin s being the jobs of its k-blocks
        const int chunk = (int)std::max<size_t>(1, std::min<size_t>(P.n_app, ((size_t)1 << 30) / g));
        d_W.ensure((size_t)chunk * g);
        const OrbEntry* jobs = (const OrbEntry*)P.d_appjobs.p;
        timed(3, 0, [&] {
            for (int s = 0; s < nspin; ++s) {
                const double* V = (const double*)C.buf.p + (size_t)s * 2 * N;
                const int a0 = (int)P.app_first[(size_t)s * nk_phys], a1 = (int)P.app_first[(size_t)(s + 1) * nk_phys];
                for (int c0 = a0; c0 < a1; c0 += chunk) {
                    const int nc = std::min(chunk, a1 - c0);
                    double* Wc = (double*)d_W.p;
                    launch_zero(Wc, (size_t)nc * g, stream);
                    launch_scatter(io.phi, (const int*)d_gindex.p, jobs + c0, nc, max_npw, Wc, N,
                                   (const double*)d_kcart.p, geo, stream);
                    launch_fft3d(TCX_PREC_FP64, +1, Wc, nc, geo, 1.0, stream);
                    launch_mul_shared(Wc, nc, V, N, stream);
                    launch_fft3d(TCX_PREC_FP64, -1, Wc, nc, geo, 1.0 / (double)N, stream);
                    launch_gather_add(Wc, (const int*)d_gindex.p, jobs + c0, nc, max_npw, io.out, N, factor,
                                      stream);
                }
            }
        });
    }
    end_apply(io);
}

}  // namespace tcx
