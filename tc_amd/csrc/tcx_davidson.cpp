// tcx_davidson.cpp -- block Davidson consumer of H*phi and the dense subspace
// eigensolver.
//
// Diagonalization::block_davidson (src/diagonalization_davidson.cpp:7-689)
// at fixed occupied orbitals: H*phi comes from calc_hamiltonian::all on the
// GPU (Context::apply_all); the subspace algebra (Gram-Schmidt, the Payne
// preconditioner, V^H H V, the rotation V X) is O(dim^2 npw) per k and runs on
// the host, as the reference runs it on MPI rank 0.  The complex non-Hermitian
// subspace problem (Eigen::ComplexEigenSolver in the reference, :480-482) is
// solved by eig_dense: Householder reduction to Hessenberg form, single-shift
// QR with Wilkinson shifts to the complex Schur form (Golub & Van Loan,
// Matrix Computations, Alg. 7.4.2 / 7.5), and back-substitution for the
// eigenvectors of the triangular factor.
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

using Vec = std::vector<cd>;

cd dotc(const Vec& a, const Vec& b)
{
    cd s = 0.0;
    for (size_t l = 0; l < a.size(); ++l) s += std::conj(a[l]) * b[l];
    return s;
}

double norm2(const Vec& a)
{
    double s = 0.0;
    for (const cd& x : a) s += std::norm(x);
    return std::sqrt(s);
}

// Diagonalization::Gram_Schmidt (diagonalization_utils.cpp:7-29): orthonormalise
// t against vs[0..target) by classical Gram-Schmidt; false if |t| < 1e-8 after the
// first projection (the reference's criterion).  A second projection pass follows
// ("twice is enough"): identical in exact arithmetic, it keeps the basis orthonormal
// when a residual is nearly inside the span (a subspace that fills the basis).
bool gram_schmidt(const std::vector<Vec>& vs, int target, Vec& t)
{
    for (int pass = 0; pass < 2; ++pass) {
        for (int b = 0; b < target; ++b) {
            const cd c = dotc(vs[b], t);
            for (size_t l = 0; l < t.size(); ++l) t[l] -= c * vs[b][l];
        }
        if (pass == 0 && norm2(t) < 1e-8) return false;
    }
    const double nrm = norm2(t);
    for (cd& x : t) x /= nrm;
    return true;
}

// make_a_new_trial_vector (diagonalization_utils.cpp:182-210): Payne preconditioner on
// the residual H phi - eps phi, normalised (zero if its norm is below 1e-8).
Vec new_trial_vector(const std::vector<double>& kG2, const Vec& phi, const cd* Hphi, cd eps)
{
    double kin = 0.0, pn = 0.0;
    for (size_t l = 0; l < phi.size(); ++l) {
        kin += kG2[l] * std::norm(phi[l]);
        pn += std::norm(phi[l]);
    }
    kin /= pn;
    Vec t(phi.size());
    for (size_t l = 0; l < phi.size(); ++l) {
        const double x = kG2[l] / kin;
        const double c = (27 + x * (18 + x * (12 + x * 8))) / (27 + x * (18 + x * (12 + x * (8 + x * 16))));
        t[l] = -c * (Hphi[l] - eps * phi[l]);
    }
    const double nrm = norm2(t);
    if (nrm < 1e-8)
        std::fill(t.begin(), t.end(), cd(0.0));
    else
        for (cd& x : t) x /= nrm;
    return t;
}

// sort_eigen (diagonalization_utils.cpp:79-118): ascending real part, same swap order
void sort_eigen(int n, std::vector<cd>& w, std::vector<cd>& X)
{
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (w[i].real() > w[j].real()) {
                std::swap(w[i], w[j]);
                for (int r = 0; r < n; ++r) std::swap(X[(size_t)r * n + i], X[(size_t)r * n + j]);
            }
}

// inverse of a small complex matrix (Gauss-Jordan, partial pivoting), row-major
bool invert(int n, std::vector<cd>& A)
{
    std::vector<cd> I((size_t)n * n, cd(0.0));
    for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::abs(A[(size_t)r * n + c]) > std::abs(A[(size_t)p * n + c])) p = r;
        if (std::abs(A[(size_t)p * n + c]) == 0.0) return false;
        if (p != c)
            for (int j = 0; j < n; ++j) {
                std::swap(A[(size_t)p * n + j], A[(size_t)c * n + j]);
                std::swap(I[(size_t)p * n + j], I[(size_t)c * n + j]);
            }
        const cd piv = A[(size_t)c * n + c];
        for (int j = 0; j < n; ++j) {
            A[(size_t)c * n + j] /= piv;
            I[(size_t)c * n + j] /= piv;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const cd f = A[(size_t)r * n + c];
            if (f == cd(0.0)) continue;
            for (int j = 0; j < n; ++j) {
                A[(size_t)r * n + j] -= f * A[(size_t)c * n + j];
                I[(size_t)r * n + j] -= f * I[(size_t)c * n + j];
            }
        }
    }
    A.swap(I);
    return true;
}

}  // namespace

void Context::davidson(const int* nb_, double* phi_d, double* phi_left_d, double* eig_d, double* energies_d,
                       int max_blocks, int num_refresh, int biortho, int prec)
{
    require(states_set, TCX_ERR_STATE, "davidson: call tcx_set_states first");
    require(nb_ && phi_d && eig_d && energies_d, TCX_ERR_ARG, "davidson: null pointer");
    require(max_blocks >= 2 && num_refresh >= 1, TCX_ERR_ARG, "davidson: need max_blocks >= 2, num_refresh >= 1");
    require(biortho == 0, TCX_ERR_UNSUPPORTED, "davidson: the bi-orthogonal variant is not implemented");
    const bool bitc = method == TCX_METHOD_BITC, uses3 = method != TCX_METHOD_HF;
    require(!bitc || phi_left_d, TCX_ERR_ARG, "davidson: BITC needs phi_left");
    if (uses3) require(jastrow_set, TCX_ERR_STATE, "davidson: call tcx_set_jastrow_rpa first");
    std::vector<int> nb(nb_, nb_ + nk);
    for (int k = 0; k < nk; ++k) require(nb[k] >= 0, TCX_ERR_ARG, "davidson: negative band count");
    cd* phi = reinterpret_cast<cd*>(phi_d);
    cd* phi_left = reinterpret_cast<cd*>(phi_left_d);
    cd* eig = reinterpret_cast<cd*>(eig_d);
    cd* energies = reinterpret_cast<cd*>(energies_d);
    std::vector<long long> coef(nk + 1, 0), band(nk + 1, 0);
    for (int k = 0; k < nk; ++k) {
        coef[k + 1] = coef[k] + (long long)nb[k] * npw[k];
        band[k + 1] = band[k] + nb[k];
    }
    // |k+G|^2 on the sphere of every k
    std::vector<std::vector<double>> kG2(nk);
    for (int k = 0; k < nk; ++k) {
        kG2[k].resize(npw[k]);
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
            kG2[k][l] = s;
        }
    }
    const std::vector<cd> eig_ref(eig, eig + band[nk]);  // bloch_states.eigenvalues (residual shifts)

    // trial vectors V (initialised by phi, :48-65; orthonormalised for BITC)
    std::vector<std::vector<Vec>> V(nk);
    for (int k = 0; k < nk; ++k)
        for (int b = 0; b < nb[k]; ++b) {
            Vec v(phi + coef[k] + (long long)b * npw[k], phi + coef[k] + (long long)(b + 1) * npw[k]);
            if (bitc) require(gram_schmidt(V[k], b, v), TCX_ERR_INTERNAL, "davidson: linear independence breaks");
            V[k].push_back(std::move(v));
        }

    // H_n applied to the last `cnt[k]` vectors of V[k]; appends to HnV
    std::vector<std::vector<Vec>> HV[3];
    auto apply_block = [&](const std::vector<int>& cnt) {
        std::vector<long long> off(nk + 1, 0);
        for (int k = 0; k < nk; ++k) off[k + 1] = off[k] + (long long)cnt[k] * npw[k];
        std::vector<cd> in(std::max<long long>(1, off[nk])), h[3];
        for (int n = 0; n < 3; ++n) h[n].assign(std::max<long long>(1, off[nk]), cd(0.0));
        for (int k = 0; k < nk; ++k) {
            const int first = (int)V[k].size() - cnt[k];
            for (int b = 0; b < cnt[k]; ++b)
                std::copy(V[k][first + b].begin(), V[k][first + b].end(), in.begin() + off[k] + (long long)b * npw[k]);
        }
        apply_all(cnt.data(), reinterpret_cast<const double*>(in.data()), reinterpret_cast<double*>(h[0].data()),
                  reinterpret_cast<double*>(h[1].data()), uses3 ? reinterpret_cast<double*>(h[2].data()) : nullptr,
                  prec, 0);
        for (int n = 0; n < 3; ++n)
            for (int k = 0; k < nk; ++k)
                for (int b = 0; b < cnt[k]; ++b) {
                    const cd* src = h[n].data() + off[k] + (long long)b * npw[k];
                    HV[n][k].emplace_back(src, src + npw[k]);
                }
    };

    for (int it = 0; it < num_refresh; ++it) {
        const bool final_loop = it == num_refresh - 1;
        for (int n = 0; n < 3; ++n) HV[n].assign(nk, {});
        std::vector<int> cnt = nb;
        for (int blk = 0; blk < max_blocks - 1; ++blk) {  // grow the subspace (:190-365)
            apply_block(cnt);
            for (int k = 0; k < nk; ++k) {
                const int nold = (int)V[k].size();
                const int first = nold - cnt[k];
                for (int b = 0; b < cnt[k]; ++b) {
                    std::vector<cd> hsum(npw[k]);
                    for (int l = 0; l < npw[k]; ++l)
                        hsum[l] = HV[0][k][first + b][l] + HV[1][k][first + b][l] + HV[2][k][first + b][l];
                    Vec t = new_trial_vector(kG2[k], V[k][first + b], hsum.data(), eig_ref[band[k] + b]);
                    if (gram_schmidt(V[k], (int)V[k].size(), t)) V[k].push_back(std::move(t));
                }
                cnt[k] = (int)V[k].size() - nold;
            }
        }
        apply_block(cnt);  // remaining matrix elements of H V (:408-437)

        for (int k = 0; k < nk; ++k) {  // subspace problem (:438-640)
            const int dim = (int)V[k].size(), nbk = nb[k];
            if (nbk == 0) continue;
            std::vector<cd> Hs[3], H((size_t)dim * dim, cd(0.0));
            for (int n = 0; n < 3; ++n) {
                Hs[n].assign((size_t)dim * dim, cd(0.0));
                if (n == 2 && !uses3) continue;
                for (int i = 0; i < dim; ++i)
                    for (int j = 0; j < dim; ++j) Hs[n][(size_t)i * dim + j] = dotc(V[k][i], HV[n][k][j]);
            }
            for (size_t e = 0; e < H.size(); ++e) H[e] = Hs[0][e] + Hs[1][e] + Hs[2][e];
            std::vector<cd> w(dim), X((size_t)dim * dim), XL;
            require(eig_dense(dim, H.data(), w.data(), X.data()), TCX_ERR_INTERNAL,
                    "davidson: diagonalization of H_subspace failed");
            sort_eigen(dim, w, X);
            if (bitc) {  // X_L = (X^-1)^H (:498)
                std::vector<cd> Xi = X;
                require(invert(dim, Xi), TCX_ERR_INTERNAL, "davidson: singular eigenvector matrix");
                XL.assign((size_t)dim * dim, cd(0.0));
                for (int i = 0; i < dim; ++i)
                    for (int j = 0; j < dim; ++j) XL[(size_t)i * dim + j] = std::conj(Xi[(size_t)j * dim + i]);
            }
            auto rotate = [&](const std::vector<cd>& C) {
                std::vector<Vec> out(nbk, Vec(npw[k], cd(0.0)));
                for (int b = 0; b < nbk; ++b)
                    for (int i = 0; i < dim; ++i) {
                        const cd c = C[(size_t)i * dim + b];
                        for (int l = 0; l < npw[k]; ++l) out[b][l] += c * V[k][i][l];
                    }
                return out;
            };
            std::vector<Vec> Vt = rotate(X);
            if (!(bitc && final_loop)) {
                for (int b = 0; b < nbk; ++b)
                    require(gram_schmidt(Vt, b, Vt[b]), TCX_ERR_INTERNAL, "davidson: linear independence breaks");
            }
            if (final_loop) {
                for (int b = 0; b < nbk; ++b) std::copy(Vt[b].begin(), Vt[b].end(), phi + coef[k] + (long long)b * npw[k]);
                if (bitc) {
                    std::vector<Vec> Vl = rotate(XL);
                    for (int b = 0; b < nbk; ++b)
                        std::copy(Vl[b].begin(), Vl[b].end(), phi_left + coef[k] + (long long)b * npw[k]);
                }
                // eigenvalues and e_n = diag(X_L^H H_n X) (X_L = X unless BITC, :571-585)
                const std::vector<cd>& L = bitc ? XL : X;
                for (int b = 0; b < nbk; ++b) {
                    eig[band[k] + b] = w[b];
                    for (int n = 0; n < 3; ++n) {
                        cd e = 0.0;
                        for (int i = 0; i < dim; ++i) {
                            cd s = 0.0;
                            for (int j = 0; j < dim; ++j) s += Hs[n][(size_t)i * dim + j] * X[(size_t)j * dim + b];
                            e += std::conj(L[(size_t)i * dim + b]) * s;
                        }
                        energies[(band[k] + b) * 3 + n] = e;
                    }
                }
            } else {
                V[k].resize(nbk);
                for (int b = 0; b < nbk; ++b) V[k][b] = Vt[b];
            }
        }
    }
}

}  // namespace tcx
