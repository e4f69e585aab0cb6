// tcx_jastrow.cpp -- polynomial and RPA+polynomial Jastrow set-up: the per-spin-pair
// parameter blocks that jastrow_poly_dev.hpp evaluates on the GPU.
//
// Follows reference src/jastrow_initialize_polynomial.cpp:
//   set_polynomial_parameters (:6-101)          -> cusp condition, c_poly_internal_
//   set_derived_polynomial_parameters (:103-118) -> 4 pi L^i
//   set_derived_RPA_polynomial_parameters (:121-155) -> exp(-L/F), M_integ_ (RPA terms present)
//   prepare_for_J3_integ / test_J3_integ / set_arrays_J3_integ (:157-354)
//                                               -> the Simpson mesh of J3_integ, doubled from
//                                                  101 points until two meshes agree to 1e-6
//                                                  at every |G| of the FFT grid
// with the reference's double arithmetic in its order, so the blocks hold the same values.
#include <algorithm>
#include <cmath>
#include <vector>

#include "jastrow_poly_dev.hpp"
#include "tcx_terms.hpp"

namespace tcx {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kFourPi = 4.0 * kPi;

// set_arrays_J3_integ (:295-354) into the mesh part of one block
void fill_rmesh(double* blk, int nr)
{
    const double L = blk[jpoly::kL], F = blk[jpoly::kF];
    const double dr = L / F / (double)(nr - 1);
    const double scale = (dr / 3.0) * kFourPi;
    double* r = blk + jpoly::kMesh;
    double* w = r + nr;
    blk[jpoly::kNr] = nr;
    for (int i = 0; i < nr; ++i) r[i] = i * dr;
    w[0] = 0.0;
    for (int i = 1; i < nr; ++i) {
        const double ex = std::exp(-r[i]);
        const double simpson = (i == nr - 1) ? 1.0 : (i % 2 == 0 ? 2.0 : 4.0);
        w[i] = ((ex - 1.0) / r[i] + ex) * (simpson * scale);
    }
}

}  // namespace

void Context::set_jastrow_poly(int deg, int C, const double* L, const double* c_in, int cusp)
{
    // validate everything before any state changes (:12-32, :43-46)
    require(deg >= 0 && deg <= jpoly::kMaxDeg, TCX_ERR_ARG, "set_jastrow_polynomial: deg must be in 0..16");
    if (deg > 0) {
        require(C == 2 || C == 3, TCX_ERR_ARG, "set_jastrow_polynomial: C must be 2 or 3");
        require(L != nullptr && c_in != nullptr, TCX_ERR_ARG, "set_jastrow_polynomial: null pointer");
        for (int p = 0; p < 4; ++p) {
            require(std::isfinite(L[p]) && L[p] >= 1e-8, TCX_ERR_ARG, "set_jastrow_polynomial: L must be positive");
            for (int i = 0; i < deg; ++i)
                require(std::isfinite(c_in[p * deg + i]), TCX_ERR_ARG,
                        "set_jastrow_polynomial: non-finite coefficient");
        }
    } else {
        require(!cusp, TCX_ERR_ARG, "set_jastrow_polynomial: the cusp condition needs deg >= 1");
    }
    plans.clear();
    invalidate_local();
    poly_deg = deg;
    poly_C = deg > 0 ? C : 3;
    poly_cusp = cusp != 0;
    poly_c.assign(c_in ? c_in : nullptr, c_in ? c_in + 4 * (size_t)deg : nullptr);
    for (int p = 0; p < 4; ++p) poly_L[p] = deg > 0 ? L[p] : 1.0;
    if (!jastrow_set) {  // polynomial only: no RPA part
        for (int p = 0; p < 4; ++p) {
            A[p] = 0.0;
            F[p] = 1.0;
        }
        jastrow_zero = true;
        jastrow_set = true;
    }
    poly_dirty = true;
    if (cell_set) build_jastrow_poly();
}

void Context::build_jastrow_poly()
{
    if (poly_deg == 0) {
        poly_dirty = false;
        poly_nr = 0;
        poly_stride = 0;
        return;
    }
    require(cell_set, TCX_ERR_STATE, "polynomial Jastrow: call tcx_set_cell first (its FFT grid sets the J3 mesh)");
    const int deg = poly_deg, C = poly_C, nc = deg + C;
    // the sign of u at set-up: blocks are built for u as given, negated copies below
    const double sgn = u_negated ? -1.0 : 1.0;
    const bool rpa = !jastrow_zero;
    double base[4][jpoly::kMesh];
    for (int p = 0; p < 4; ++p) {
        double* b = base[p];
        std::fill(b, b + jpoly::kMesh, 0.0);
        const double Lp = poly_L[p];
        b[jpoly::kDeg] = deg;
        b[jpoly::kC] = C;
        b[jpoly::kL] = Lp;
        b[jpoly::kA] = sgn * A[p];
        b[jpoly::kF] = F[p];
        // cusp condition on c_1 (:47-60)
        std::vector<double> c(poly_c.begin() + (size_t)p * deg, poly_c.begin() + (size_t)(p + 1) * deg);
        const int s1 = p / 2, s2 = p % 2;
        const double cusp_const = poly_cusp ? (s1 == s2 ? 0.25 : 0.5) : 0.0;
        const double higher = deg == 1 ? 0.0 : c[1];
        c[0] = (Lp / (double)C) * (higher - cusp_const);
        // c_poly_internal_ with the minus sign (:62-100)
        double* ci = b + jpoly::kCoef;
        for (int i = 0; i < deg; ++i) {
            if (C == 3) {
                ci[i] -= c[i];
                ci[i + 1] += 3.0 * c[i] / Lp;
                ci[i + 2] -= 3.0 * c[i] / (Lp * Lp);
                ci[i + 3] += c[i] / (Lp * Lp * Lp);
            } else {
                ci[i] -= c[i];
                ci[i + 1] += 2.0 * c[i] / Lp;
                ci[i + 2] -= c[i] / (Lp * Lp);
            }
        }
        // 4 pi L^i (:103-118)
        double* pw = b + jpoly::kPowL;
        pw[0] = kFourPi;
        for (int i = 1; i < 2 * deg + 6; ++i) pw[i] = Lp * pw[i - 1];
        if (rpa) {  // :121-150 (M_integ_ above index 29 stays zero, as sized there)
            require(std::fabs(F[p]) > 1e-8, TCX_ERR_ARG, "polynomial Jastrow: F must be non-zero");
            b[jpoly::kFinv] = 1.0 / F[p];
            b[jpoly::kExpLF] = std::exp(-Lp / F[p]);
            double* M = b + jpoly::kM;
            M[0] = F[p] * (1.0 - b[jpoly::kExpLF]);
            for (int i = 1; i < jpoly::kMaxDegTaylor2; ++i)
                M[i] = F[p] * (-b[jpoly::kExpLF] + i * M[i - 1] / Lp);
        }
        (void)nc;
    }
    // J3 mesh (:157-204), only for the RPA+polynomial form of degree > 1
    int nr = 0;
    if (rpa && deg > 1) {
        std::vector<double> glist;
        glist.reserve((size_t)N);
        for (int i3 = 0; i3 < geo.n3; ++i3)
            for (int i2 = 0; i2 < geo.n2; ++i2)
                for (int i1 = 0; i1 < geo.n1; ++i1) {
                    const int m1 = i1 <= geo.n1 / 2 ? i1 : i1 - geo.n1;
                    const int m2 = i2 <= geo.n2 / 2 ? i2 : i2 - geo.n2;
                    const int m3 = i3 <= geo.n3 / 2 ? i3 : i3 - geo.n3;
                    double g2 = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        const double g = geo.b[d] * m1 + geo.b[3 + d] * m2 + geo.b[6 + d] * m3;
                        g2 += g * g;
                    }
                    glist.push_back(std::sqrt(g2));
                }
        std::sort(glist.begin(), glist.end());
        glist.erase(std::unique(glist.begin(), glist.end()), glist.end());
        // J3 depends on (L, F) only: test each distinct pair once
        std::vector<int> pairs;
        for (int p = 0; p < 4; ++p) {
            bool seen = false;
            for (int q : pairs) seen = seen || (poly_L[q] == poly_L[p] && F[q] == F[p]);
            if (!seen) pairs.push_back(p);
        }
        auto j3_all = [&](int mesh) {
            std::vector<double> out;
            std::vector<double> blk(jpoly::block_doubles(mesh));
            for (int p : pairs) {
                std::copy(base[p], base[p] + jpoly::kMesh, blk.begin());
                fill_rmesh(blk.data(), mesh);
                for (double g : glist) out.push_back(jpoly::J3_integ(blk.data(), g));
            }
            return out;
        };
        nr = 101;
        std::vector<double> prev = j3_all(nr);
        bool small = false;
        for (int it = 0; it < 15; ++it) {
            const int nr_new = 2 * nr - 1;
            const std::vector<double> next = j3_all(nr_new);
            double err = 0.0;
            for (size_t i = 0; i < prev.size(); ++i) err = std::max(err, std::fabs(prev[i] - next[i]));
            if (err < 1e-6) {
                small = true;
                break;
            }
            nr = nr_new;
            prev = next;
        }
        require(small, TCX_ERR_ARG, "polynomial Jastrow: the J3_integ mesh did not converge (reference :200)");
    }
    const size_t stride = jpoly::block_doubles(nr);
    std::vector<double> h(4 * stride, 0.0), hn(4 * stride, 0.0);
    for (int p = 0; p < 4; ++p) {
        double* b = h.data() + p * stride;
        std::copy(base[p], base[p] + jpoly::kMesh, b);
        if (nr > 0) fill_rmesh(b, nr);
        double* bn = hn.data() + p * stride;
        std::copy(b, b + stride, bn);
        // Jastrow::take_Hermitian_conj (jastrow_functions.cpp:417-440): A and c_poly_internal_ flip
        bn[jpoly::kA] = -b[jpoly::kA];
        for (int i = 0; i < jpoly::kMaxCoef; ++i) bn[jpoly::kCoef + i] = -b[jpoly::kCoef + i];
    }
    // slot 0 holds u as given, slot 1 its negative, whatever the sign at build time
    upload(d_jpoly[0], h.data(), h.size() * sizeof(double));
    upload(d_jpoly[1], hn.data(), hn.size() * sizeof(double));
    poly_nr = nr;
    poly_stride = stride;
    poly_dirty = false;
}

void Context::negate_u()
{
    for (double& a : A) a = -a;
    u_negated = !u_negated;
}

JastrowPair Context::jastrow_pair(int s1, int s2)
{
    if (poly_dirty) build_jastrow_poly();
    const int p = s1 * 2 + s2;
    JastrowPair j{A[p], F[p], jastrow_zero ? 1 : 0, jastrow_dA ? 1 : 0};
    if (poly_deg > 0) {
        j.pdeg = poly_deg;
        j.poly = (const double*)d_jpoly[u_negated ? 1 : 0].p + (size_t)p * poly_stride;
    }
    return j;
}

void Context::jastrow_kernels(int s1, int s2, long long n, const double* G2, double* uk, double* v)
{
    require(jastrow_set, TCX_ERR_STATE, "jastrow_kernels: call tcx_set_jastrow_rpa or tcx_set_jastrow_polynomial first");
    require(s1 >= 0 && s1 < 2 && s2 >= 0 && s2 < 2 && n >= 0, TCX_ERR_ARG, "jastrow_kernels: bad spin pair or count");
    if (n == 0) return;
    require(G2 && uk && v, TCX_ERR_ARG, "jastrow_kernels: null pointer");
    for (const void* ptr : {(const void*)G2, (const void*)uk, (const void*)v}) {
        hipPointerAttribute_t at{};
        const hipError_t e = hipPointerGetAttributes(&at, ptr);
        if (e != hipSuccess) (void)hipGetLastError();
        require(e == hipSuccess && at.type == hipMemoryTypeDevice && at.device == device, TCX_ERR_ARG,
                "jastrow_kernels: G2, uk and tc2body must be device memory of the context's GPU");
    }
    launch_jastrow_eval(G2, uk, v, n, jastrow_pair(s1, s2), stream);
}

}  // namespace tcx
