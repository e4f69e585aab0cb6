// tcx_davidson.cpp -- block Davidson consumer of H*phi and the dense subspace
// eigensolver.
//
// Diagonalization::block_davidson (src/diagonalization_davidson.cpp:7-689)
// at fixed occupied orbitals.  Everything stays in HBM: H*phi comes from
// calc_hamiltonian::all (Context::apply_all on device pointers), and the FP64
// subspace algebra -- the Payne preconditioner, Gram-Schmidt, V^H H_n V, the
// batched non-Hermitian eigensolve of every k, the rotations V X and V X_L --
// runs in the kernels of kernels_davidson.hip, batched over k.  The host only
// sequences the launches and reads back one norm per k per Gram-Schmidt step
// (the reference's 1e-8 drop decides which candidates enter the subspace).
//
// eig_dense is the host version of the subspace eigensolver behind
// tcx_eig_dense: Householder reduction to Hessenberg form, single-shift QR with
// Wilkinson shifts to the complex Schur form (Golub & Van Loan, Matrix
// Computations, Alg. 7.4.2 / 7.5), and back-substitution for the eigenvectors
// of the triangular factor; the device kernel follows the same steps.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "tcx_terms.hpp"

namespace tcx {

using cd = std::complex<double>;

bool eig_dense(int n, const cd* A_in, cd* w, cd* X)
{
    if (n <= 0) return true;
    std::vector<cd> H(A_in, A_in + (size_t)n * n), Z((size_t)n * n, cd(0.0));
    auto h = [&](int i, int j) -> cd& { return H[(size_t)i * n + j]; };
    auto z = [&](int i, int j) -> cd& { return Z[(size_t)i * n + j]; };
    for (int i = 0; i < n; ++i) z(i, i) = 1.0;
    const double eps = 2.220446049250313e-16;

    // Householder reduction to upper Hessenberg form, A = Z H Z^H
    std::vector<cd> v(n);
    for (int k = 0; k + 2 < n; ++k) {
        double xn = 0.0;
        for (int i = k + 1; i < n; ++i) xn += std::norm(h(i, k));
        xn = std::sqrt(xn);
        if (xn == 0.0) continue;
        const cd x0 = h(k + 1, k);
        const cd phase = std::abs(x0) > 0.0 ? x0 / std::abs(x0) : cd(1.0);
        const cd alpha = -phase * xn;
        std::fill(v.begin(), v.end(), cd(0.0));
        for (int i = k + 1; i < n; ++i) v[i] = h(i, k);
        v[k + 1] -= alpha;
        double vn = 0.0;
        for (int i = k + 1; i < n; ++i) vn += std::norm(v[i]);
        vn = std::sqrt(vn);
        if (vn == 0.0) continue;
        for (int i = k + 1; i < n; ++i) v[i] /= vn;
        for (int j = 0; j < n; ++j) {  // H <- (I - 2 v v^H) H
            cd s = 0.0;
            for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * h(i, j);
            for (int i = k + 1; i < n; ++i) h(i, j) -= 2.0 * v[i] * s;
        }
        for (int i = 0; i < n; ++i) {  // H <- H (I - 2 v v^H), Z <- Z (I - 2 v v^H)
            cd s = 0.0, t = 0.0;
            for (int j = k + 1; j < n; ++j) {
                s += h(i, j) * v[j];
                t += z(i, j) * v[j];
            }
            for (int j = k + 1; j < n; ++j) {
                h(i, j) -= 2.0 * s * std::conj(v[j]);
                z(i, j) -= 2.0 * t * std::conj(v[j]);
            }
        }
        for (int i = k + 2; i < n; ++i) h(i, k) = 0.0;
    }
    double anorm = 0.0;
    for (const cd& a : H) anorm += std::norm(a);
    anorm = std::max(std::sqrt(anorm), 1e-300);

    // single-shift QR on the active block [l, hi] until H is upper triangular
    int hi = n - 1, iter = 0, total = 0;
    while (hi > 0) {
        if (++total > 200 * n) return false;
        int l = hi;
        while (l > 0) {
            double s = std::abs(h(l - 1, l - 1)) + std::abs(h(l, l));
            if (s == 0.0) s = anorm;
            if (std::abs(h(l, l - 1)) <= eps * s) {
                h(l, l - 1) = 0.0;
                break;
            }
            --l;
        }
        if (l == hi) {
            --hi;
            iter = 0;
            continue;
        }
        ++iter;
        const cd a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1), d = h(hi, hi);
        cd mu;
        if (iter % 11 == 10) {
            mu = d + cd(std::abs(c.real()) + std::abs(c.imag()), 0.0);  // exceptional shift
        } else {
            const cd half = 0.5 * (a + d);
            const cd disc = std::sqrt(0.25 * (a - d) * (a - d) + b * c);
            const cd m1 = half + disc, m2 = half - disc;
            mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
        }
        cd x = h(l, l) - mu, y = h(l + 1, l);
        for (int k = l; k < hi; ++k) {
            if (k > l) {
                x = h(k, k - 1);
                y = h(k + 1, k - 1);
            }
            const double r = std::sqrt(std::norm(x) + std::norm(y));
            double cc = 1.0;
            cd ss = 0.0;
            if (r > 0.0) {
                if (std::abs(x) == 0.0) {
                    cc = 0.0;
                    ss = std::conj(y) / std::abs(y);
                } else {
                    cc = std::abs(x) / r;
                    ss = (x / std::abs(x)) * std::conj(y) / r;
                }
            }
            for (int j = (k > l ? k - 1 : l); j < n; ++j) {  // rows k, k+1 <- G [rows]
                const cd h1 = h(k, j), h2 = h(k + 1, j);
                h(k, j) = cc * h1 + ss * h2;
                h(k + 1, j) = -std::conj(ss) * h1 + cc * h2;
            }
            if (k > l) h(k + 1, k - 1) = 0.0;
            const int imax = std::min(k + 2, hi);
            for (int i = 0; i <= imax; ++i) {  // columns k, k+1 <- [columns] G^H
                const cd h1 = h(i, k), h2 = h(i, k + 1);
                h(i, k) = cc * h1 + std::conj(ss) * h2;
                h(i, k + 1) = -ss * h1 + cc * h2;
            }
            for (int i = 0; i < n; ++i) {
                const cd z1 = z(i, k), z2 = z(i, k + 1);
                z(i, k) = cc * z1 + std::conj(ss) * z2;
                z(i, k + 1) = -ss * z1 + cc * z2;
            }
        }
    }

    // eigenvectors of the triangular factor, back to the original basis, unit norm
    const double small = eps * anorm;
    std::vector<cd> y(n);
    for (int k = 0; k < n; ++k) w[k] = h(k, k);
    for (int k = 0; k < n; ++k) {
        std::fill(y.begin(), y.end(), cd(0.0));
        y[k] = 1.0;
        for (int i = k - 1; i >= 0; --i) {
            cd s = 0.0;
            for (int j = i + 1; j <= k; ++j) s += h(i, j) * y[j];
            cd den = h(i, i) - w[k];
            if (std::abs(den) < small) den = small;
            y[i] = -s / den;
        }
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) {
            cd s = 0.0;
            for (int j = 0; j <= k; ++j) s += z(i, j) * y[j];
            X[(size_t)i * n + k] = s;
            nrm += std::norm(s);
        }
        nrm = std::sqrt(nrm);
        for (int i = 0; i < n; ++i) X[(size_t)i * n + k] /= nrm;
    }
    return true;
}


namespace {

// Device layout of one Davidson call: int and long long tables of nk entries each.
enum IntTab { kNpw, kBand0, kNb, kFirst, kCnt, kSize, kNA, kBidx, kOdst, kStatus, kNumIntTabs };
enum OffTab { kVoff, kPoff, kGoff, kCoff, kNumOffTabs };

}  // namespace

void Context::davidson(const int* nb_, double* phi_d, double* phi_left_d, double* eig_d, double* energies_d,
                       int max_blocks, int num_refresh, int biortho, int prec)
{
    require(!g_capturing, TCX_ERR_STATE, "graph capture: davidson is not capturable");
    require(states_set, TCX_ERR_STATE, "davidson: call tcx_set_states first");
    require(nb_ && phi_d && eig_d && energies_d, TCX_ERR_ARG, "davidson: null pointer");
    require(max_blocks >= 2 && num_refresh >= 1, TCX_ERR_ARG, "davidson: need max_blocks >= 2, num_refresh >= 1");
    const bool bitc = method == TCX_METHOD_BITC, uses3 = method != TCX_METHOD_HF;
    const bool bi = biortho != 0;  // biortho_david_ (diagonalization.cpp:117-121: BITC only)
    require(!bi || bitc, TCX_ERR_ARG, "davidson: the bi-orthogonal variant needs a BITC cell");
    require(!bitc || phi_left_d, TCX_ERR_ARG, "davidson: BITC needs phi_left");
    if (uses3) require(jastrow_set, TCX_ERR_STATE, "davidson: call tcx_set_jastrow_rpa first");
    std::vector<int> nb(nb_, nb_ + nk);
    int max_nb = 0, max_npw = 0;
    for (int k = 0; k < nk; ++k) {
        require(nb[k] >= 0 && nb[k] <= npw[k], TCX_ERR_ARG, "davidson: band count outside [0, npw]");
        max_nb = std::max(max_nb, nb[k]);
        max_npw = std::max(max_npw, npw[k]);
    }
    if (max_nb == 0) return;
    const int gs = max_blocks * max_nb;  // subspace capacity per k (rows of the small matrices)
    const int nH = uses3 ? 3 : 2;

    // tables
    std::vector<int> I((size_t)kNumIntTabs * nk, 0);
    std::vector<long long> L((size_t)kNumOffTabs * nk, 0);
    auto it = [&](IntTab t) { return I.data() + (size_t)t * nk; };
    long long vtot = 0, ptot = 0;
    for (int k = 0; k < nk; ++k) {
        it(kNpw)[k] = npw[k];
        it(kBand0)[k] = k ? it(kBand0)[k - 1] + nb[k - 1] : 0;
        it(kNb)[k] = nb[k];
        L[kVoff * nk + k] = vtot;
        L[kPoff * nk + k] = ptot;
        L[kGoff * nk + k] = pw_off[k];
        vtot += (long long)max_blocks * nb[k] * npw[k];
        ptot += (long long)nb[k] * npw[k];
    }
    const int nbands = it(kBand0)[nk - 1] + nb[nk - 1];

    // |k+G|^2 on the sphere of every k
    std::vector<double> kG2(std::max<long long>(1, pw_off[nk]));
    for (int k = 0; k < nk; ++k)
        for (int l = 0; l < npw[k]; ++l) {
            const int e = gindex_h[pw_off[k] + l];
            const int i1 = e % geo.n1, i2 = (e / geo.n1) % geo.n2, i3 = e / (geo.n1 * geo.n2);
            const int m[3] = {i1 <= geo.n1 / 2 ? i1 : i1 - geo.n1, i2 <= geo.n2 / 2 ? i2 : i2 - geo.n2,
                              i3 <= geo.n3 / 2 ? i3 : i3 - geo.n3};
            double s = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double x = kcart[3 * k + d] + geo.b[d] * m[0] + geo.b[3 + d] * m[1] + geo.b[6 + d] * m[2];
                s += x * x;
            }
            kG2[pw_off[k] + l] = s;
        }

    // device memory of this call (vtot, ptot coefficients of 16 bytes)
    DBuf dV[2], dHV[3], dP, dT, dIn, dO[3], dkG2, deig, dHs, dX, dXL, dw, den, dc, dnrm, dI, dL, dwork;
    DBuf dVL[2], dHLV[3], dPL, dTL, deigc, dcc, dbn;  // bi-orthogonal variant: left side
    const size_t vb = (size_t)std::max<long long>(1, vtot) * 16, pb = (size_t)std::max<long long>(1, ptot) * 16;
    const size_t mb = (size_t)nk * gs * gs * 16;
    for (auto& b : dV) b.ensure(vb);
    for (int n = 0; n < nH; ++n) dHV[n].ensure(vb);
    if (bi) {
        for (auto& b : dVL) b.ensure(vb);
        for (int n = 0; n < nH; ++n) dHLV[n].ensure(vb);
        dPL.ensure(pb);
        dTL.ensure(pb);
        deigc.ensure((size_t)nbands * 16);
        dcc.ensure((size_t)nk * 16);
        dbn.ensure((size_t)nbands * 8);
    }
    for (DBuf* b : {&dP, &dT, &dIn, &dO[0], &dO[1]}) b->ensure(pb);
    if (uses3) dO[2].ensure(pb);
    dkG2.ensure(kG2.size() * 8);
    deig.ensure((size_t)nbands * 16);
    dHs.ensure(mb * nH);
    dX.ensure(mb);
    if (bitc) dXL.ensure(mb);
    dw.ensure((size_t)nk * gs * 16);
    den.ensure((size_t)nbands * 3 * 16);
    dc.ensure((size_t)nk * gs * 16);
    dnrm.ensure((size_t)nk * 8);
    dI.ensure(I.size() * 4);
    dL.ensure(L.size() * 8);
    if (!dv_eig_fits_lds(gs)) dwork.ensure(dv_eig_work_bytes(gs, nk));

    // copies on the context stream, waited for (host buffers are reused right after)
    auto h2d = [&](void* dst, const void* src, size_t bytes) {
        TCX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
        TCX_HIP_CHECK(hipStreamSynchronize(stream));
    };
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
        TCX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
        TCX_HIP_CHECK(hipStreamSynchronize(stream));
    };
    int* dIp = (int*)dI.p;
    const long long* dLp = (const long long*)dL.p;
    auto di = [&](IntTab t) { return dIp + (size_t)t * nk; };
    auto dl = [&](OffTab t) { return dLp + (size_t)t * nk; };
    auto push_ints = [&](IntTab t0, int ntabs) { h2d(di(t0), it(t0), (size_t)ntabs * nk * 4); };
    auto d64 = [](DBuf& b) { return (double*)b.p; };

    h2d(dI.p, I.data(), I.size() * 4);
    h2d(dL.p, L.data(), L.size() * 8);
    h2d(dkG2.p, kG2.data(), kG2.size() * 8);
    h2d(deig.p, eig_d, (size_t)nbands * 16);  // bloch_states.eigenvalues shift the residuals
    h2d(dP.p, phi_d, (size_t)ptot * 16);
    if (bi) {
        std::vector<double> ec(eig_d, eig_d + 2 * (size_t)nbands);  // the left residuals take conj(eps) (:322-326)
        for (int b = 0; b < nbands; ++b) ec[2 * b + 1] = -ec[2 * b + 1];
        h2d(deigc.p, ec.data(), ec.size() * 8);
        h2d(dPL.p, phi_left_d, (size_t)ptot * 16);
    }

    // Gram-Schmidt of slots 0..nb[k] of B in place, in band order; a vector whose first
    // projection leaves a norm below 1e-8 is an error (diagonalization_davidson.cpp:527-545)
    auto orthonormalise = [&](double* B) {
        for (int b = 0; b < max_nb; ++b) {
            for (int k = 0; k < nk; ++k) {
                const bool on = b < nb[k];
                it(kNA)[k] = on ? b : 0;
                it(kBidx)[k] = on ? b : -1;
                it(kOdst)[k] = on ? b : -1;
            }
            push_ints(kNA, 3);
            launch_dv_project(B, dl(kVoff), di(kNA), b, B, dl(kVoff), di(kBidx), di(kNpw), max_npw, nk, d64(dc),
                              d64(dnrm), stream);
            std::vector<double> nrm(nk);
            d2h(nrm.data(), dnrm.p, (size_t)nk * 8);
            for (int k = 0; k < nk; ++k)
                require(it(kBidx)[k] < 0 || nrm[k] >= 1e-8, TCX_ERR_INTERNAL, "davidson: linear independence breaks");
            launch_dv_project(B, dl(kVoff), di(kNA), b, B, dl(kVoff), di(kBidx), di(kNpw), max_npw, nk, d64(dc),
                              d64(dnrm), stream);
            launch_dv_store(B, dl(kVoff), di(kBidx), d64(dnrm), B, dl(kVoff), di(kOdst), di(kNpw), max_npw, nk,
                            stream);
        }
    };

    int cur = 0;
    std::vector<int> size(nk);
    // trial vectors V (initialised by phi, :48-65; orthonormalised for BITC unless
    // bi-orthogonal, where V_left starts from phi_left as it is, :86-103)
    launch_dv_copy(d64(dP), dl(kPoff), nullptr, d64(dV[cur]), dl(kVoff), nullptr, di(kNb), max_nb, di(kNpw), max_npw,
                   nk, stream);
    if (bi)
        launch_dv_copy(d64(dPL), dl(kPoff), nullptr, d64(dVL[cur]), dl(kVoff), nullptr, di(kNb), max_nb, di(kNpw),
                       max_npw, nk, stream);
    else if (bitc)
        orthonormalise(d64(dV[cur]));
    size = nb;

    // H_n (or, left, H_n^dagger: :234-251) applied to slots first[k] .. + cnt[k] of V (V_left),
    // into the same slots of the H_n V (H_n^dagger V_left) sets
    auto apply_block = [&](bool left) {
        long long c = 0;
        int max_cnt = 0;
        for (int k = 0; k < nk; ++k) {
            L[kCoff * nk + k] = c;
            c += (long long)it(kCnt)[k] * npw[k];
            max_cnt = std::max(max_cnt, it(kCnt)[k]);
        }
        h2d((void*)dl(kCoff), L.data() + (size_t)kCoff * nk, (size_t)nk * 8);
        push_ints(kFirst, 2);
        launch_dv_copy(d64(left ? dVL[cur] : dV[cur]), dl(kVoff), di(kFirst), d64(dIn), dl(kCoff), nullptr, di(kCnt),
                       max_cnt, di(kNpw), max_npw, nk, stream);
        if (left) hermitian_conj();
        try {
            apply_all(it(kCnt), d64(dIn), d64(dO[0]), d64(dO[1]), uses3 ? d64(dO[2]) : nullptr, prec,
                      TCX_FLAG_DEVICE_PTRS);
        } catch (...) {
            if (left) hermitian_conj();
            throw;
        }
        if (left) hermitian_conj();
        for (int n = 0; n < nH; ++n)
            launch_dv_copy(d64(dO[n]), dl(kCoff), nullptr, d64(left ? dHLV[n] : dHV[n]), dl(kVoff), di(kFirst),
                           di(kCnt), max_cnt, di(kNpw), max_npw, nk, stream);
    };

    // Gram_Schmidt_biortho (diagonalization_utils.cpp:31-77) of candidate b (right in dT, left
    // in dTL) against the pairs in V, V_left, batched over k; a kept pair enters slot size[k]
    auto biortho_candidate = [&](int b) {
        int max_nA = 0;
        for (int k = 0; k < nk; ++k) {
            it(kNA)[k] = size[k];
            it(kBidx)[k] = b < it(kCnt)[k] ? b : -1;
            it(kOdst)[k] = size[k];
            max_nA = std::max(max_nA, size[k]);
        }
        push_ints(kNA, 3);
        // right: t -= sum_j <V_left_j|t> V_j, then |t| >= 1e-10 and t /= |t| (:41-55)
        launch_dv_project(d64(dV[cur]), dl(kVoff), di(kNA), max_nA, d64(dT), dl(kPoff), di(kBidx), di(kNpw), max_npw,
                          nk, d64(dc), d64(dnrm), stream, d64(dVL[cur]));
        std::vector<double> nrm(nk);
        d2h(nrm.data(), dnrm.p, (size_t)nk * 8);
        bool any = false;
        for (int k = 0; k < nk; ++k) {
            if (it(kBidx)[k] >= 0 && !(nrm[k] >= 1e-10)) it(kBidx)[k] = -1;
            any = any || it(kBidx)[k] >= 0;
        }
        if (!any) return;
        push_ints(kBidx, 1);
        // a second pass of the same projection, as for Gram_Schmidt (exact arithmetic: no change)
        launch_dv_project(d64(dV[cur]), dl(kVoff), di(kNA), max_nA, d64(dT), dl(kPoff), di(kBidx), di(kNpw), max_npw,
                          nk, d64(dc), d64(dnrm), stream, d64(dVL[cur]));
        launch_dv_store(d64(dT), dl(kPoff), di(kBidx), d64(dnrm), d64(dV[cur]), dl(kVoff), di(kOdst), di(kNpw),
                        max_npw, nk, stream);  // tentative: slot size[k] counts only if the left passes
        // left: s -= sum_j <V_j|s> V_left_j (:58-63), then s /= conj(<s|t>) with |<s|t>| >= 1e-10 (:65-76)
        for (int pass = 0; pass < 2; ++pass)
            launch_dv_project(d64(dVL[cur]), dl(kVoff), di(kNA), max_nA, d64(dTL), dl(kPoff), di(kBidx), di(kNpw),
                              max_npw, nk, d64(dc), d64(dnrm), stream, d64(dV[cur]));
        launch_dv_dot(d64(dTL), dl(kPoff), di(kBidx), d64(dV[cur]), dl(kVoff), di(kOdst), di(kNpw), nk, d64(dcc),
                      stream);
        std::vector<double> cc(2 * (size_t)nk);
        d2h(cc.data(), dcc.p, cc.size() * 8);
        for (int k = 0; k < nk; ++k)
            if (it(kBidx)[k] >= 0 && !(std::hypot(cc[2 * k], cc[2 * k + 1]) >= 1e-10)) it(kBidx)[k] = -1;
        push_ints(kBidx, 1);
        launch_dv_cdiv_store(d64(dTL), dl(kPoff), di(kBidx), d64(dcc), d64(dVL[cur]), dl(kVoff), di(kOdst), di(kNpw),
                             max_npw, nk, stream);
        for (int k = 0; k < nk; ++k)
            if (it(kBidx)[k] >= 0) ++size[k];
    };

    for (int iter = 0; iter < num_refresh; ++iter) {
        const bool final_loop = iter == num_refresh - 1;
        for (int k = 0; k < nk; ++k) it(kCnt)[k] = nb[k];
        for (int blk = 0; blk < max_blocks - 1; ++blk) {  // grow the subspace (:190-365)
            int max_cnt = 0;
            for (int k = 0; k < nk; ++k) {
                it(kFirst)[k] = size[k] - it(kCnt)[k];
                max_cnt = std::max(max_cnt, it(kCnt)[k]);
            }
            apply_block(false);
            if (bi) apply_block(true);
            launch_dv_precond(d64(dV[cur]), dl(kVoff), di(kFirst), di(kCnt), max_cnt, d64(dHV[0]), d64(dHV[1]),
                              uses3 ? d64(dHV[2]) : nullptr, d64(dkG2), dl(kGoff), d64(deig), di(kBand0), d64(dT),
                              dl(kPoff), di(kNpw), nk, stream);
            const std::vector<int> nold = size;
            if (bi) {
                launch_dv_precond(d64(dVL[cur]), dl(kVoff), di(kFirst), di(kCnt), max_cnt, d64(dHLV[0]),
                                  d64(dHLV[1]), uses3 ? d64(dHLV[2]) : nullptr, d64(dkG2), dl(kGoff), d64(deigc),
                                  di(kBand0), d64(dTL), dl(kPoff), di(kNpw), nk, stream);
                for (int b = 0; b < max_cnt; ++b) biortho_candidate(b);
                for (int k = 0; k < nk; ++k) it(kCnt)[k] = size[k] - nold[k];
                continue;
            }
            for (int b = 0; b < max_cnt; ++b) {  // candidates in band order, each against the grown V
                int max_nA = 0;
                for (int k = 0; k < nk; ++k) {
                    it(kNA)[k] = size[k];
                    it(kBidx)[k] = b < it(kCnt)[k] ? b : -1;
                    it(kOdst)[k] = size[k];
                    max_nA = std::max(max_nA, size[k]);
                }
                push_ints(kNA, 3);
                launch_dv_project(d64(dV[cur]), dl(kVoff), di(kNA), max_nA, d64(dT), dl(kPoff), di(kBidx), di(kNpw),
                                  max_npw, nk, d64(dc), d64(dnrm), stream);
                std::vector<double> nrm(nk);
                d2h(nrm.data(), dnrm.p, (size_t)nk * 8);
                bool any = false;
                for (int k = 0; k < nk; ++k) {  // Gram_Schmidt's 1e-8 drop (diagonalization_utils.cpp:7-29)
                    if (it(kBidx)[k] >= 0 && !(nrm[k] >= 1e-8)) it(kBidx)[k] = -1;
                    any = any || it(kBidx)[k] >= 0;
                }
                if (!any) continue;
                push_ints(kBidx, 1);
                launch_dv_project(d64(dV[cur]), dl(kVoff), di(kNA), max_nA, d64(dT), dl(kPoff), di(kBidx), di(kNpw),
                                  max_npw, nk, d64(dc), d64(dnrm), stream);
                launch_dv_store(d64(dT), dl(kPoff), di(kBidx), d64(dnrm), d64(dV[cur]), dl(kVoff), di(kOdst),
                                di(kNpw), max_npw, nk, stream);
                for (int k = 0; k < nk; ++k)
                    if (it(kBidx)[k] >= 0) ++size[k];
            }
            for (int k = 0; k < nk; ++k) it(kCnt)[k] = size[k] - nold[k];
        }
        for (int k = 0; k < nk; ++k) it(kFirst)[k] = size[k] - it(kCnt)[k];
        apply_block(false);  // remaining matrix elements of H V (:408-437)

        // subspace problem (:438-640): H_ij = <V_i|H V_j>, <V_left_i|H V_j> bi-orthogonal (:465-466)
        for (int k = 0; k < nk; ++k) it(kSize)[k] = size[k];
        push_ints(kSize, 1);
        launch_dv_gram(d64(bi ? dVL[cur] : dV[cur]), dl(kVoff), di(kSize), d64(dHV[0]), d64(dHV[1]),
                       uses3 ? d64(dHV[2]) : nullptr, nH, di(kNpw), nk, d64(dHs), gs, stream);
        launch_dv_eig(d64(dHs), nH, di(kSize), gs, nk, (bitc && final_loop) || bi, bi ? 1 : 0, d64(dw), d64(dX),
                      bitc ? d64(dXL) : nullptr, dwork.p ? d64(dwork) : nullptr, di(kNb), di(kBand0),
                      final_loop ? d64(den) : nullptr, di(kStatus), stream);
        d2h(it(kStatus), di(kStatus), (size_t)nk * 4);
        for (int k = 0; k < nk; ++k) {
            require(it(kStatus)[k] != 1, TCX_ERR_INTERNAL, "davidson: diagonalization of H_subspace failed");
            require(it(kStatus)[k] != 2, TCX_ERR_INTERNAL, "davidson: singular eigenvector matrix");
        }
        const int alt = 1 - cur;
        launch_dv_rotate(d64(dV[cur]), dl(kVoff), di(kSize), d64(dX), gs, d64(dV[alt]), dl(kVoff), di(kNb), di(kNpw),
                         max_npw, nk, stream);
        if (bi) {
            // V X and V_left X_L, rescaled so that |V X| = 1 with <left|right> kept (:596-615)
            launch_dv_rotate(d64(dVL[cur]), dl(kVoff), di(kSize), d64(dXL), gs, d64(dVL[alt]), dl(kVoff), di(kNb),
                             di(kNpw), max_npw, nk, stream);
            launch_dv_binorm(d64(dV[alt]), d64(dVL[alt]), dl(kVoff), di(kNb), max_nb, di(kBand0), di(kNpw), nk,
                             d64(dbn), stream);
            std::vector<double> bn(nbands);
            d2h(bn.data(), dbn.p, (size_t)nbands * 8);
            for (int b = 0; b < nbands; ++b)
                require(bn[b] >= 1e-10, TCX_ERR_INTERNAL, "davidson: linear independence breaks");
        } else if (!(bitc && final_loop)) {
            orthonormalise(d64(dV[alt]));
        }
        if (final_loop) {
            launch_dv_copy(d64(dV[alt]), dl(kVoff), nullptr, d64(dP), dl(kPoff), nullptr, di(kNb), max_nb, di(kNpw),
                           max_npw, nk, stream);
            d2h(phi_d, dP.p, (size_t)ptot * 16);
            if (bi) {
                launch_dv_copy(d64(dVL[alt]), dl(kVoff), nullptr, d64(dPL), dl(kPoff), nullptr, di(kNb), max_nb,
                               di(kNpw), max_npw, nk, stream);
                d2h(phi_left_d, dPL.p, (size_t)ptot * 16);
            } else if (bitc) {
                launch_dv_rotate(d64(dV[cur]), dl(kVoff), di(kSize), d64(dXL), gs, d64(dT), dl(kPoff), di(kNb),
                                 di(kNpw), max_npw, nk, stream);
                d2h(phi_left_d, dT.p, (size_t)ptot * 16);
            }
            std::vector<double> w((size_t)nk * gs * 2);
            d2h(w.data(), dw.p, w.size() * 8);
            for (int k = 0; k < nk; ++k)
                for (int b = 0; b < nb[k]; ++b) {
                    eig_d[2 * (it(kBand0)[k] + b)] = w[2 * ((size_t)k * gs + b)];
                    eig_d[2 * (it(kBand0)[k] + b) + 1] = w[2 * ((size_t)k * gs + b) + 1];
                }
            d2h(energies_d, den.p, (size_t)nbands * 3 * 16);
        } else {
            cur = alt;
            size = nb;
        }
    }
}

}  // namespace tcx
