// scf.cpp -- C++ host of TC++'s SCF loop over the C ABI (see scf.hpp).
#include "scf.hpp"

#include <cmath>
#include <stdexcept>

namespace tcx_host {

namespace {

constexpr double kPi = 3.14159265358979323846;

void check(tcx_context* ctx, int st, const char* what)
{
    if (st != TCX_OK)
        throw StatusError(st, std::string(what) + ": " + (ctx ? tcx_last_error(ctx) : "status " + std::to_string(st)));
}

// Equal modulo reciprocal-lattice vectors (kpoints.cpp:10-26).
bool matched(const double v[3], const std::vector<std::array<double, 3>>& lst)
{
    for (const auto& w : lst) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) {
            double x = v[d] - w[d];
            x -= std::nearbyint(x);
            s += x * x;
        }
        if (s < 1e-8) return true;
    }
    return false;
}

void inverse3(const double a[3][3], double out[3][3])
{
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                       a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                       a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (std::fabs(det) < 1e-300) throw std::runtime_error("singular 3x3 matrix");
    out[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    out[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    out[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    out[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    out[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    out[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    out[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    out[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    out[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
}

double det3(const double a[3][3])
{
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

using Bands = std::vector<Complex>;  // [band * npw + ipw] at one irreducible k

}  // namespace

std::array<int, 9> tc_rotation(const std::array<int, 9>& r)
{
    const long det = (long)r[0] * (r[4] * r[8] - r[5] * r[7]) - (long)r[1] * (r[3] * r[8] - r[5] * r[6]) +
                     (long)r[2] * (r[3] * r[7] - r[4] * r[6]);
    if (det != 1 && det != -1) throw std::runtime_error("symmetry: QE rotation is not unimodular");
    const int d = (int)det;
    return {(r[4] * r[8] - r[5] * r[7]) / d, (r[2] * r[7] - r[1] * r[8]) / d, (r[1] * r[5] - r[2] * r[4]) / d,
            (r[5] * r[6] - r[3] * r[8]) / d, (r[0] * r[8] - r[2] * r[6]) / d, (r[2] * r[3] - r[0] * r[5]) / d,
            (r[3] * r[7] - r[4] * r[6]) / d, (r[1] * r[6] - r[0] * r[7]) / d, (r[0] * r[4] - r[1] * r[3]) / d};
}

std::vector<KReplica> replicate_kpoints(const QEInputs& qe)
{
    const int nirr = qe.nirr();
    double binv[3][3];
    inverse3(qe.b_norm, binv);
    std::vector<std::array<double, 3>> kcrys(nirr);
    for (int ik = 0; ik < nirr; ++ik)  // kcart = (b1 b2 b3) kcrys
        for (int i = 0; i < 3; ++i)
            kcrys[ik][i] = binv[0][i] * qe.kcart[ik][0] + binv[1][i] * qe.kcart[ik][1] + binv[2][i] * qe.kcart[ik][2];
    std::vector<std::array<int, 9>> U;
    for (const auto& r : qe.rot_qe) U.push_back(tc_rotation(r));
    std::vector<std::vector<KReplica>> per_k(nirr);
    std::vector<std::array<double, 3>> vec_list;
    for (int ik = 0; ik < nirr; ++ik) {
        per_k[ik].push_back({ik, 0, false, {kcrys[ik][0], kcrys[ik][1], kcrys[ik][2]}});
        if (qe.kweight[ik] >= 1e-8) vec_list.push_back(kcrys[ik]);
    }
    for (int ik = 0; ik < nirr; ++ik) {
        if (qe.kweight[ik] < 1e-8) continue;
        for (int isym = 0; isym < (int)U.size(); ++isym) {
            double v[3], mv[3];
            for (int i = 0; i < 3; ++i) {  // U^T k
                v[i] = U[isym][i] * kcrys[ik][0] + U[isym][3 + i] * kcrys[ik][1] + U[isym][6 + i] * kcrys[ik][2];
                if (std::fabs(v[i]) < 1e-4) v[i] = 0.0;
                mv[i] = -v[i];
            }
            if (!matched(v, vec_list)) {
                vec_list.push_back({v[0], v[1], v[2]});
                per_k[ik].push_back({ik, isym, false, {v[0], v[1], v[2]}});
            }
            if (qe.time_reversal && !matched(mv, vec_list)) {
                vec_list.push_back({mv[0], mv[1], mv[2]});
                per_k[ik].push_back({ik, isym, true, {mv[0], mv[1], mv[2]}});
            }
        }
    }
    std::vector<KReplica> out;
    for (const auto& lst : per_k) out.insert(out.end(), lst.begin(), lst.end());
    return out;
}

void replica_basis(const QEInputs& qe, const KReplica& r, std::vector<int>& gidx, std::vector<Complex>& phase)
{
    const std::array<int, 9> U = tc_rotation(qe.rot_qe[r.isym]);
    const auto& mill = qe.mill[r.ik];
    const auto& tau = qe.trans[r.isym];
    const int* n = qe.grid;
    gidx.resize(mill.size());
    phase.resize(mill.size());
    for (size_t ipw = 0; ipw < mill.size(); ++ipw) {
        const auto& G = mill[ipw];
        int Gs[3];
        for (int j = 0; j < 3; ++j) {  // U^T G
            Gs[j] = G[0] * U[j] + G[1] * U[3 + j] + G[2] * U[6 + j];
            if (r.trev) Gs[j] = -Gs[j];
        }
        const double arg = 2.0 * kPi * (Gs[0] * tau[0] + Gs[1] * tau[1] + Gs[2] * tau[2]);
        phase[ipw] = Complex(std::cos(arg), std::sin(arg));
        int w[3];
        for (int j = 0; j < 3; ++j) w[j] = ((Gs[j] % n[j]) + n[j]) % n[j];
        gidx[ipw] = w[0] + n[0] * (w[1] + n[1] * w[2]);
    }
}

SCFResult Diagonalization::scf(tcx_context* ctx, const QEInputs& qe) const
{
    const SCFSettings& S = set_;
    const bool bitc = S.method == TCX_METHOD_BITC;
    const bool dm = S.mixes_density_matrix;
    const double beta = S.mixing_beta;
    const int nirr = qe.nirr(), nb = qe.nb;
    if (nirr == 0 || nb <= 0) throw std::runtime_error("scf: no k-points or bands");

    // ---- cell, full k-mesh, basis, Jastrow, divergence correction, symmetry
    const std::vector<KReplica> reps = replicate_kpoints(qe);
    const int nk = (int)reps.size();
    double ainv[3][3];
    inverse3(qe.cell, ainv);
    double recip[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) recip[3 * i + j] = 2.0 * kPi * ainv[j][i];  // rows b_i = 2 pi (A^-1)^T
    const double volume = std::fabs(det3(qe.cell));
    check(ctx, tcx_set_cell(ctx, recip, volume, qe.grid, 1, S.method), "set_cell");
    std::vector<double> kfrac(3 * nk);
    for (int k = 0; k < nk; ++k)
        for (int d = 0; d < 3; ++d) kfrac[3 * k + d] = reps[k].kcrys[d];
    check(ctx, tcx_set_kpoints(ctx, nk, kfrac.data(), nk), "set_kpoints");
    std::vector<std::vector<int>> gidx(nk);
    std::vector<std::vector<Complex>> phase(nk);
    std::vector<int> npw(nk), gcat;
    for (int k = 0; k < nk; ++k) {
        replica_basis(qe, reps[k], gidx[k], phase[k]);
        npw[k] = (int)gidx[k].size();
        gcat.insert(gcat.end(), gidx[k].begin(), gidx[k].end());
    }
    check(ctx, tcx_set_basis(ctx, npw.data(), gcat.data()), "set_basis");
    double A[4] = {0, 0, 0, 0}, F[4] = {1, 1, 1, 1};
    if (S.method != TCX_METHOD_HF) {  // jastrow_initialize.cpp:28-75
        const double a = S.A_input * std::sqrt(volume / (4.0 * kPi * qe.nelec));
        for (int i = 0; i < 4; ++i) A[i] = a;
        F[0] = F[3] = std::sqrt(2.0 * a);
        F[1] = F[2] = std::sqrt(a);
    }
    check(ctx, tcx_set_jastrow_rpa(ctx, A, F), "set_jastrow_rpa");
    // the run's Jastrow is the RPA form alone: clear a polynomial part left on the context
    check(ctx, tcx_set_jastrow_polynomial(ctx, 0, 0, nullptr, nullptr, 0), "set_jastrow_polynomial");
    double alpha = 0.0;
    std::vector<double> sv(nk);
    check(nullptr, tcx_sum_of_vaux(recip, volume, nk, kfrac.data(), &alpha, sv.data()), "sum_of_vaux");
    check(ctx, tcx_set_div_correction(ctx, 1, alpha, sv.data()), "set_div_correction");
    std::vector<int> first(nirr, -1), k_src(nk), rot(9 * nk), trev(nk);
    std::vector<double> trans(3 * nk);
    for (int k = 0; k < nk; ++k)
        if (first[reps[k].ik] < 0) first[reps[k].ik] = k;
    for (int k = 0; k < nk; ++k) {
        const std::array<int, 9> U = tc_rotation(qe.rot_qe[reps[k].isym]);
        k_src[k] = first[reps[k].ik];
        for (int i = 0; i < 9; ++i) rot[9 * k + i] = U[i];
        for (int d = 0; d < 3; ++d) trans[3 * k + d] = qe.trans[reps[k].isym][d];
        trev[k] = reps[k].trev ? 1 : 0;
    }
    check(ctx, tcx_set_kpoint_symmetry(ctx, k_src.data(), rot.data(), trans.data(), trev.data()),
          "set_kpoint_symmetry");  // kept across tcx_set_states
    const long long N = (long long)qe.grid[0] * qe.grid[1] * qe.grid[2];

    // ---- state of the loop (BlochStates at the irreducible k)
    std::vector<Bands> phi(nirr), chi, old_phi, old_chi;
    for (int ik = 0; ik < nirr; ++ik) {
        const size_t np = qe.mill[ik].size();
        if (qe.coef[ik].size() != (size_t)nb * np) throw std::runtime_error("scf: coefficient array size");
        phi[ik] = qe.coef[ik];
    }
    if (bitc) chi = phi;  // phik_left_scf = phik_scf at the start (bloch_states_initialize.cpp:370)
    std::vector<std::vector<Complex>> eig(nirr, std::vector<Complex>(nb));
    for (int ik = 0; ik < nirr; ++ik)
        for (int b = 0; b < nb; ++b) eig[ik][b] = qe.eig[ik][b];
    std::vector<Complex> en;  // [ik][band][3] from the last Davidson
    std::vector<std::vector<double>> fill(nirr), fill_old;
    std::vector<int> nocc(nirr, 0);
    bool have_old = false;
    std::vector<Complex> rho((size_t)2 * N);
    SCFResult out;

    auto filling_step = [&](bool keep_old) {  // set_filling (:9-199)
        if (keep_old) {
            fill_old.assign(nirr, {});
            for (int ik = 0; ik < nirr; ++ik) fill_old[ik].assign(fill[ik].begin(), fill[ik].begin() + nocc[ik]);
        }
        std::vector<int> nbk(nirr, nb);
        std::vector<double> e((size_t)nirr * nb), f((size_t)nirr * nb);
        for (int ik = 0; ik < nirr; ++ik)
            for (int b = 0; b < nb; ++b) e[(size_t)ik * nb + b] = eig[ik][b].real();
        double mu = 0.0;
        check(nullptr,
              tcx_set_filling(nirr, nbk.data(), e.data(), qe.kweight.data(), qe.nelec, S.smearing, S.smearing_width,
                              f.data(), &mu, nocc.data()),
              "set_filling");
        for (int ik = 0; ik < nirr; ++ik) fill[ik].assign(f.begin() + (size_t)ik * nb, f.begin() + (size_t)(ik + 1) * nb);
        out.fillings.push_back({mu, eig, fill});
    };

    auto total_energy = [&]() {  // calc_total_energy (total_energy.cpp:101-170)
        Complex e[3] = {0.0, 0.0, 0.0};
        for (const KReplica& r : reps)
            for (int b = 0; b < nb; ++b)
                for (int n = 0; n < 3; ++n) {
                    Complex w = en[((size_t)r.ik * nb + b) * 3 + n];
                    if (r.trev) w = std::conj(w);
                    e[n] += fill[r.ik][b] * w / double(n + 1);
                }
        if (S.method == TCX_METHOD_HF) e[2] = 0.0;
        std::array<Complex, 4> E;
        for (int n = 0; n < 3; ++n) E[n] = e[n] / double(nk);
        E[3] = E[0] + E[1] + E[2];
        return E;
    };

    auto density_step = [&](bool first_iter) -> double {  // occupied set, set_density (:256-358)
        std::vector<int> nocc_k(nk);
        std::vector<double> f_cat;
        std::vector<Complex> occ_cat, left_cat;
        for (int k = 0; k < nk; ++k) {
            const KReplica& r = reps[k];
            const int ik = r.ik;
            const size_t np = gidx[k].size();
            auto add = [&](const Bands& src, int b, std::vector<Complex>& dst) {
                for (size_t i = 0; i < np; ++i) {
                    Complex c = src[(size_t)b * np + i];
                    if (r.trev) c = std::conj(c);
                    dst.push_back(c * phase[k][i]);
                }
            };
            int n = 0;
            for (int b = 0; b < nocc[ik]; ++b, ++n) {
                add(phi[ik], b, occ_cat);
                if (bitc) add(chi[ik], b, left_cat);
                f_cat.push_back(fill[ik][b]);
            }
            if (have_old)
                for (int b = 0; b < (int)fill_old[ik].size(); ++b, ++n) {
                    add(old_phi[ik], b, occ_cat);
                    if (bitc) add(old_chi[ik], b, left_cat);
                    f_cat.push_back(fill_old[ik][b]);
                }
            nocc_k[k] = n;
        }
        check(ctx,
              tcx_set_states(ctx, nocc_k.data(), f_cat.data(), (const double*)occ_cat.data(),
                             bitc ? (const double*)left_cat.data() : nullptr),
              "set_states");
        std::vector<Complex> rnew((size_t)2 * N);
        check(ctx, tcx_get_density(ctx, (double*)rnew.data()), "get_density");
        if (first_iter) {
            rho = rnew;
            return 0.0;
        }
        if (!dm) {  // density mixing (:345-349)
            for (size_t i = 0; i < rnew.size(); ++i) rnew[i] = (1.0 - beta) * rho[i] + beta * rnew[i];
            check(ctx, tcx_set_density(ctx, (const double*)rnew.data()), "set_density");
        }
        double diff = 0.0;  // L1 over spin up, x2 for the unpolarised pair, / (N Nk)
        for (long long i = 0; i < N; ++i) diff += std::abs(rnew[i] - rho[i]);
        diff = 2.0 * diff / ((double)N * nk);
        rho = rnew;
        out.density_difference.push_back(diff);
        return diff;
    };

    Complex e_prev = 0.0;
    bool e_conv = false, e_conv_prev = false, c_conv = false, c_conv_prev = false, conv = false;
    for (int it = 0; it < S.max_num_iterations; ++it) {
        const bool first_iter = it == 0;
        filling_step(!first_iter && dm);
        if (!first_iter) {
            const std::array<Complex, 4> E = total_energy();
            out.energies.push_back(E);
            e_conv = std::abs(E[3] - e_prev) < S.energy_tolerance;
            e_prev = E[3];
            if (dm) {  // mix_density_matrix (:204-223)
                for (auto& f : fill)
                    for (double& x : f) x *= beta;
                for (auto& f : fill_old)
                    for (double& x : f) x *= 1.0 - beta;
            }
        }
        const double diff = density_step(first_iter);
        if (!first_iter) {
            c_conv = diff < S.charge_tolerance;
            conv = e_conv && e_conv_prev && c_conv && c_conv_prev;
            if (conv) break;
        }
        e_conv_prev = e_conv;
        c_conv_prev = c_conv;

        // block Davidson on the irreducible k (the first replica of each)
        std::vector<int> nb_full(nk, 0);
        std::vector<Complex> phi_cat, left_cat, eig_in((size_t)nirr * nb);
        for (int ik = 0; ik < nirr; ++ik) {
            nb_full[first[ik]] = nb;
            phi_cat.insert(phi_cat.end(), phi[ik].begin(), phi[ik].end());
            for (int b = 0; b < nb; ++b) eig_in[(size_t)ik * nb + b] = eig[ik][b];
        }
        if (bitc && S.biortho_david)  // the left trial bands are the current left orbitals (:86-103)
            for (int ik = 0; ik < nirr; ++ik) left_cat.insert(left_cat.end(), chi[ik].begin(), chi[ik].end());
        else if (bitc)
            left_cat.assign(phi_cat.size(), 0.0);
        std::vector<Complex> en_out((size_t)nirr * nb * 3);
        check(ctx,
              tcx_davidson(ctx, nb_full.data(), (double*)phi_cat.data(), bitc ? (double*)left_cat.data() : nullptr,
                           (double*)eig_in.data(), (double*)en_out.data(), S.max_num_blocks_david, S.num_refresh_david,
                           bitc && S.biortho_david ? 1 : 0, S.precision),
              "davidson");
        ++out.iterations;
        if (dm) {  // reset_phik keeps the occupied bands before the update (bloch_states_scfloop_phik.cpp:86-99)
            old_phi.assign(nirr, {});
            if (bitc) old_chi.assign(nirr, {});
            for (int ik = 0; ik < nirr; ++ik) {
                const size_t np = qe.mill[ik].size();
                old_phi[ik].assign(phi[ik].begin(), phi[ik].begin() + (size_t)nocc[ik] * np);
                if (bitc) old_chi[ik].assign(chi[ik].begin(), chi[ik].begin() + (size_t)nocc[ik] * np);
            }
        }
        size_t off = 0;
        for (int ik = 0; ik < nirr; ++ik) {
            const size_t sz = (size_t)nb * qe.mill[ik].size();
            phi[ik].assign(phi_cat.begin() + off, phi_cat.begin() + off + sz);
            if (bitc) chi[ik].assign(left_cat.begin() + off, left_cat.begin() + off + sz);
            off += sz;
            for (int b = 0; b < nb; ++b) eig[ik][b] = eig_in[(size_t)ik * nb + b];
        }
        en = en_out;
        if (dm) {
            have_old = true;
            if (!first_iter) {  // recover_filling_for_density_matrix (:225-242)
                for (auto& f : fill)
                    for (double& x : f) x /= beta;
                for (auto& f : fill_old)
                    for (double& x : f) x = 0.0;
            }
        }
    }
    if (!conv) {  // no convergence: final set_filling, calc_total_energy, set_density (:247-258)
        filling_step(dm);
        out.energies.push_back(total_energy());
        if (dm) {
            for (auto& f : fill)
                for (double& x : f) x *= beta;
            for (auto& f : fill_old)
                for (double& x : f) x *= 1.0 - beta;
        }
        density_step(false);
    }
    out.converged = conv;
    return out;
}

}  // namespace tcx_host
