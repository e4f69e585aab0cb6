// tcx_adjoint.cpp -- gradients of the exchange and TC two-body energies of the
// occupied set with respect to the orbital coefficients and the Jastrow strength.
//
// With the applied bands equal to the occupied orbitals,
//   E_x = (1/(2 Nk)) sum_kj f_kj <phi_kj| V_x phi_kj>          (TotalEnergy, total_energy.cpp:120-123)
//   E_2 = (1/(2 Nk)) sum_kj f_kj <chi_kj| (tc2h + tc2x) phi_kj>   (left orbitals for BITC)
// are quadratic forms of a two-body operator that is symmetric under exchange
// of its two particles, so the derivative through the orbitals inside the
// operator equals the derivative through the bra/ket, and
//   dE_x / d conj(phi_kj) = (f_kj / Nk) V_x phi_kj
//   dE_2 / d conj(chi_kj) = (f_kj / Nk) H_2 phi_kj
//   dE_2 / d phi_kj       = (f_kj / Nk) conj(H_2^dagger chi_kj),
// where H_2^dagger is H_2 with u -> -u and the right and left orbitals swapped
// (Jastrow::take_Hermitian_conj, BlochStates::switch_left_right_orbitals;
// diagonalization_davidson.cpp:234-251); the density is rebuilt from the
// swapped orbitals, which makes it the exact adjoint.  u and tc_2body enter
// H_2 linearly, so dE_2/dA at fixed F is E_2 evaluated with the d/dA kernels
// (JastrowPair::dA); the divergence correction does not depend on A.
// The backward pass recomputes everything from the orbitals (no pair densities
// are stored between the forward and the backward calls).
#include <algorithm>
#include <complex>
#include <vector>

#include "tcx_terms.hpp"

namespace tcx {

using cd = std::complex<double>;

namespace {

// (1/(2 Nk)) sum_b f_b <bra_b|H_b> with compensated (Neumaier) sums of both parts
cd half_energy(const std::vector<cd>& bra, const std::vector<cd>& Hphi, const std::vector<long long>& coef,
               const std::vector<int>& nocc, const std::vector<int>& npw, const std::vector<double>& fill,
               int num_kpoints)
{
    double s[2] = {0, 0}, c[2] = {0, 0};
    auto add = [&](int i, double v) {
        const double t = s[i] + v;
        c[i] += std::fabs(s[i]) >= std::fabs(v) ? (s[i] - t) + v : (v - t) + s[i];
        s[i] = t;
    };
    int b = 0;
    for (size_t k = 0; k < nocc.size(); ++k)
        for (int i = 0; i < nocc[k]; ++i, ++b) {
            cd e = 0.0;
            const long long o = coef[k] + (long long)i * npw[k];
            for (int l = 0; l < npw[k]; ++l) e += std::conj(bra[o + l]) * Hphi[o + l];
            add(0, fill[b] * e.real());
            add(1, fill[b] * e.imag());
        }
    return cd(s[0] + c[0], s[1] + c[1]) / (2.0 * num_kpoints);
}

}  // namespace

void Context::swap_left_right()
{
    std::swap(d_phi_occ, d_chi_occ);
    for (int p = 0; p < 2; ++p) {
        std::swap(d_occR[p], d_chiR[p]);
        std::swap(occR_valid[p], chiR_valid[p]);
        gradR_valid[p] = false;  // gradients are of the right orbitals
        a_gradR_valid[p] = false;
    }
    density_valid = false;       // rebuilt from the swapped pair: conj of the density
    invalidate_local();
}

void Context::grad_x(double* grad, double* energy, int prec)
{
    require(states_set, TCX_ERR_STATE, "grad_x: call tcx_set_states first");
    require_one_spin("grad_exchange");
    require(grad && energy, TCX_ERR_ARG, "grad_x: null pointer");
    const long long ncoef = occ_coef[nk];
    std::vector<cd> phi(ncoef), h(ncoef, cd(0.0));
    if (ncoef) {
        TCX_HIP_CHECK(hipMemcpy(phi.data(), d_phi_occ.p, ncoef * 16, hipMemcpyDeviceToHost));
        hf2x(nocc.data(), reinterpret_cast<const double*>(phi.data()), reinterpret_cast<double*>(h.data()), prec, 0);
    }
    const cd e = half_energy(phi, h, occ_coef, nocc, npw, fill, num_kpoints);
    energy[0] = e.real();
    energy[1] = e.imag();
    cd* g = reinterpret_cast<cd*>(grad);
    int b = 0;
    for (int k = 0; k < nk; ++k)
        for (int i = 0; i < nocc[k]; ++i, ++b) {
            const long long o = occ_coef[k] + (long long)i * npw[k];
            for (int l = 0; l < npw[k]; ++l) g[o + l] = (fill[b] / num_kpoints) * h[o + l];
        }
}

void Context::grad_tc2(double* grad_left, double* grad_right, double* energy, double* dE_dA, int prec)
{
    require(states_set, TCX_ERR_STATE, "grad_tc2: call tcx_set_states first");
    require_one_spin("grad_tc2");
    require(method != TCX_METHOD_HF && jastrow_set, TCX_ERR_STATE, "grad_tc2: needs a TC/BITC cell and a Jastrow");
    require(grad_left && grad_right && energy && dE_dA, TCX_ERR_ARG, "grad_tc2: null pointer");
    const bool bitc = method == TCX_METHOD_BITC;
    const long long ncoef = occ_coef[nk];
    std::vector<cd> phi(ncoef), chi(ncoef);
    if (ncoef) {
        TCX_HIP_CHECK(hipMemcpy(phi.data(), d_phi_occ.p, ncoef * 16, hipMemcpyDeviceToHost));
        TCX_HIP_CHECK(hipMemcpy(chi.data(), (bitc ? d_chi_occ : d_phi_occ).p, ncoef * 16, hipMemcpyDeviceToHost));
    }
    auto H2 = [&](const std::vector<cd>& in) {
        std::vector<cd> out(ncoef, cd(0.0));
        if (ncoef) {
            tc2h(nocc.data(), reinterpret_cast<const double*>(in.data()), reinterpret_cast<double*>(out.data()), 0);
            tc2x(nocc.data(), reinterpret_cast<const double*>(in.data()), reinterpret_cast<double*>(out.data()), prec,
                 0);
        }
        return out;
    };
    const std::vector<cd> Hphi = H2(phi);
    const cd e = half_energy(chi, Hphi, occ_coef, nocc, npw, fill, num_kpoints);

    // H_2^dagger chi: u -> -u, right <-> left (TC: the orbitals are their own left set)
    // (every change of the kernels or orbitals drops the cached local potentials)
    negate_u();
    invalidate_local();
    if (bitc) swap_left_right();  // TC: phi = chi, nothing to swap
    std::vector<cd> Hdchi;
    try {
        Hdchi = H2(chi);
    } catch (...) {
        negate_u();
        invalidate_local();
        if (bitc) swap_left_right();
        throw;
    }
    negate_u();
    invalidate_local();
    if (bitc) swap_left_right();

    // dE_2/dA: the same energy with the d/dA kernels and no divergence term
    const bool div_saved = div_enabled;
    jastrow_dA = true;
    div_enabled = false;
    invalidate_local();
    std::vector<cd> Hd;
    try {
        Hd = H2(phi);
    } catch (...) {
        jastrow_dA = false;
        div_enabled = div_saved;
        invalidate_local();
        throw;
    }
    jastrow_dA = false;
    div_enabled = div_saved;
    invalidate_local();
    const cd dA = half_energy(chi, Hd, occ_coef, nocc, npw, fill, num_kpoints);

    cd* gl = reinterpret_cast<cd*>(grad_left);
    cd* gr = reinterpret_cast<cd*>(grad_right);
    int b = 0;
    for (int k = 0; k < nk; ++k)
        for (int i = 0; i < nocc[k]; ++i, ++b) {
            const long long o = occ_coef[k] + (long long)i * npw[k];
            const double w = fill[b] / num_kpoints;
            for (int l = 0; l < npw[k]; ++l) {
                gl[o + l] = w * Hphi[o + l];
                gr[o + l] = w * std::conj(Hdchi[o + l]);
            }
        }
    energy[0] = e.real();
    energy[1] = e.imag();
    dE_dA[0] = dA.real();
    dE_dA[1] = dA.imag();
}

// Kernel gradients (the backward pair reduction of kernels_tc3x.hip, k_kgrad): pair densities
// of every occupied (target, source) pair are recomputed and reduced into radial bins of
// |k - q + G|.  term 0: E_x and its Coulomb kernel v; term 1: the exchange part of E_2
// (tc2x without the divergence term) and its kernels tc_2body (dE_dK) and u (dE_du).
void Context::kernel_grad(int term, int nbins, double width, double* dE_dK, double* dE_du, double* sums, int prec)
{
    require(states_set, TCX_ERR_STATE, "kernel_grad: call tcx_set_states first");
    require_one_spin("kernel_gradient");
    require(term == 0 || term == 1, TCX_ERR_ARG, "kernel_grad: term is 0 (exchange) or 1 (TC two-body)");
    require(dE_dK && sums && (term == 0 || dE_du), TCX_ERR_ARG, "kernel_grad: null pointer");
    require(nbins >= 1 && nbins <= 1024 && width > 0.0, TCX_ERR_ARG, "kernel_grad: 1 <= nbins <= 1024, width > 0");
    const bool tc2 = term == 1;
    if (tc2) require(method != TCX_METHOD_HF && jastrow_set, TCX_ERR_STATE, "kernel_grad: needs a TC/BITC cell");
    const bool bitc = method == TCX_METHOD_BITC;
    const size_t es = prec == TCX_PREC_FP64 ? 16 : 8, gT = (size_t)N * es;
    const int C = tc2 ? 8 : 1;
    d_aux.ensure((4 * (size_t)nbins + 4) * sizeof(double));
    launch_zero(d_aux.p, (4 * (size_t)nbins + 4) * sizeof(double), stream);
    double* bins_v = (double*)d_aux.p;
    double* bins_u = bins_v + 2 * nbins;
    double* d_sums = bins_u + 2 * nbins;
    if (n_occ > 0) {
        ensure_occR(prec);
        if (bitc) ensure_chiR(prec);
        if (tc2) ensure_gradR(prec);
        const void* occR = d_occR[prec].p;
        const void* chiR = bitc ? d_chiR[prec].p : occR;
        const void* gradR = tc2 ? d_gradR[prec].p : nullptr;
        std::vector<ProdChan> ch;
        if (!tc2) {
            ch.push_back(ProdChan{occR, occR, 1, 1, 0, 1, 0, 1, 0, 0});  // conj(phi_qi) phi_kj
        } else {
            ch.push_back(ProdChan{chiR, occR, 1, 1, 0, 1, 0, 1, 0, 0});  // rho = conj(chi_qi) phi_kj
            for (int d = 0; d < 3; ++d) ch.push_back(ProdChan{chiR, gradR, 1, 1, 0, 1, 0, 3, d, 0});  // sig_d
            ch.push_back(ProdChan{occR, chiR, 1, 1, 0, 1, 0, 1, 0, 0});  // rho' = conj(phi_qi) chi_kj
            for (int d = 0; d < 3; ++d) ch.push_back(ProdChan{gradR, chiR, 1, 3, d, 1, 0, 1, 0, 0});  // tau_d
        }
        upload(d_chans, ch.data(), ch.size() * sizeof(ProdChan));
        const PairSet& S = pair_set(nocc, 1, C, prec);  // occupied x occupied
        d_rec.ensure((size_t)std::max(1, S.max_launch_pairs) * C * gT);
        const double NkV = num_kpoints * volume;
        double c = tc2 ? -1.0 / NkV : -4.0 * kPiHost / NkV;  // two_body_factor of tc2x / hf2x
        if (nspin == 1) c /= 2.0;
        const double w0 = c / (2.0 * num_kpoints);
        const JastrowPair j0 = jastrow_pair(0, 0), j1 = jastrow_pair(0, 1);
        for (const Launch& L : S.launches) {
            const PairEntry* pairs = (const PairEntry*)S.d_pairs.p + L.pair0;
            launch_products_fft(prec, pairs, L.npairs, (const ProdChan*)d_chans.p, C, d_rec.p, geo, 1.0 / (double)N,
                                stream);
            launch_kgrad(prec, d_rec.p, tc2 ? 1 : 0, pairs, L.npairs, (const double*)S.d_kq.p, (const double*)d_fill.p,
                         geo, N, j0, j1, w0, nbins, width, bins_v, bins_u, d_sums, stream);
        }
    }
    TCX_HIP_CHECK(hipStreamSynchronize(stream));
    TCX_HIP_CHECK(hipMemcpy(dE_dK, bins_v, 2 * (size_t)nbins * sizeof(double), hipMemcpyDeviceToHost));
    if (tc2) TCX_HIP_CHECK(hipMemcpy(dE_du, bins_u, 2 * (size_t)nbins * sizeof(double), hipMemcpyDeviceToHost));
    TCX_HIP_CHECK(hipMemcpy(sums, d_sums, 4 * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace tcx
