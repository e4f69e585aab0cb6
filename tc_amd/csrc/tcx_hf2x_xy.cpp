// tcx_hf2x_xy.cpp -- host side of the plane-first hf2x pipeline (kernels_hf2x_xy.hip):
// the support columns of every k - q row and the launch sequence.
//
// The pair density of (k,j | q,i) has G-space components G = G_k - G_q with G_k in the
// basis of k and G_q in the basis of q (src/calc_hamiltonian_hf2x.cpp:123-135 forms it
// from the two orbitals on the FFT grid).  Its (kx, ky) columns are therefore the
// differences of the two bases' columns.  Per basis and per row ky the columns span
// [min kx, max kx]; the difference of two such rows spans [lo_k - hi_q, hi_k - lo_q],
// and a row of the difference set is the hull of every such span that lands on it
// (modulo the grid).  Hulls only add columns, so the record holds every non-zero one.
#include <algorithm>
#include <vector>

#include "tcx_terms.hpp"

namespace tcx {

namespace {

int posmod(int a, int n)
{
    const int r = a % n;
    return r < 0 ? r + n : r;
}
int centered_host(int i, int n) { return i <= n / 2 ? i : i - n; }

}  // namespace

// Automatic order: z-first.  Plane-first moves a third fewer bytes per pair at 40^3 but
// ran slower there (its xy passes wait on LDS barriers and dependent loads per pair).
bool Context::use_plane_first(int prec) const
{
    if (prec != TCX_PREC_FP32 || geo.n1 != geo.n2) return false;
    return hf2x_order == 1;
}

void Context::build_xy_columns(const Plan& P)
{
    XyColumns& X = P.xy;
    if (X.built) return;
    const int n1 = geo.n1, n2 = geo.n2;
    // per k and row i2: [lo, hi] of the basis' centred kx (lo > hi: empty row)
    std::vector<int> blo((size_t)nk * n2, 1 << 20), bhi((size_t)nk * n2, -(1 << 20));
    for (int k = 0; k < nk; ++k)
        for (int l = 0; l < npw[k]; ++l) {
            const int e = gindex_h[pw_off[k] + l];
            const int i1 = e % n1, i2 = (e / n1) % n2;
            const int m1 = centered_host(i1, n1);
            int& lo = blo[(size_t)k * n2 + i2];
            int& hi = bhi[(size_t)k * n2 + i2];
            lo = std::min(lo, m1);
            hi = std::max(hi, m1);
        }
    // per (k, q): the rows of the difference set, computed once
    std::vector<int> dlo((size_t)nk * nk * n2), dhi((size_t)nk * nk * n2);
    std::vector<char> done((size_t)nk * nk, 0);
    auto rows_of = [&](int k, int q) {
        int* lo = &dlo[((size_t)k * nk + q) * n2];
        int* hi = &dhi[((size_t)k * nk + q) * n2];
        if (done[(size_t)k * nk + q]) return;
        std::fill(lo, lo + n2, 1 << 20);
        std::fill(hi, hi + n2, -(1 << 20));
        for (int a = 0; a < n2; ++a) {
            const int alo = blo[(size_t)k * n2 + a], ahi = bhi[(size_t)k * n2 + a];
            if (alo > ahi) continue;
            for (int b = 0; b < n2; ++b) {
                const int qlo = blo[(size_t)q * n2 + b], qhi = bhi[(size_t)q * n2 + b];
                if (qlo > qhi) continue;
                const int r = posmod(centered_host(a, n2) - centered_host(b, n2), n2);
                lo[r] = std::min(lo[r], alo - qhi);
                hi[r] = std::max(hi[r], ahi - qlo);
            }
        }
        done[(size_t)k * nk + q] = 1;
    };
    const int nrows = (int)P.kq.size() / 3;
    std::vector<int> colxy, coloff(1, 0), rowinfo((size_t)nrows * n2 * 4, 0);
    int cstride = 0;
    for (int r = 0; r < nrows; ++r) {
        const int k = P.kq_kq[2 * r], q = P.kq_kq[2 * r + 1];
        rows_of(k, q);
        const int* lo = &dlo[((size_t)k * nk + q) * n2];
        const int* hi = &dhi[((size_t)k * nk + q) * n2];
        for (int i2 = 0; i2 < n2; ++i2) {
            int len = 0, l0 = 0;
            if (lo[i2] <= hi[i2]) {
                len = std::min(hi[i2] - lo[i2] + 1, n1);
                l0 = lo[i2];
            }
            int* ri = &rowinfo[((size_t)r * n2 + i2) * 4];
            ri[0] = l0;
            ri[1] = len;
            ri[2] = (int)colxy.size() - coloff[r];  // first column of the row
            for (int j = 0; j < len; ++j) colxy.push_back(posmod(l0 + j, n1) | (i2 << 16));
        }
        coloff.push_back((int)colxy.size());
        cstride = std::max(cstride, coloff[r + 1] - coloff[r]);
    }
    // mean columns per pair record
    double sum = 0.0;
    for (const PairEntry& pe : P.pairs) sum += coloff[pe.kq + 1] - coloff[pe.kq];
    X.mean_cols = P.pairs.empty() ? 0.0 : sum / (double)P.pairs.size();
    X.cstride = (cstride + 15) / 16 * 16;  // record rows start on 128-byte lines
    upload(X.d_colxy, colxy.data(), colxy.size() * sizeof(int));
    upload(X.d_coloff, coloff.data(), coloff.size() * sizeof(int));
    upload(X.d_rowinfo, rowinfo.data(), rowinfo.size() * sizeof(int));
    X.built = true;
}

// The three plane-first passes of every launch; W (FP64) was cleared by the caller.
void Context::hf2x_plane_first(const Plan& P)
{
    build_xy_columns(P);
    const XyColumns& X = P.xy;
    hf2x_record_cols = X.mean_cols;
    const int cs = X.cstride;
    d_pairbuf.ensure((size_t)std::max(1, P.max_launch_pairs) * geo.n3 * std::max(1, cs) * 8);
    const int* colxy = (const int*)X.d_colxy.p;
    const int* coloff = (const int*)X.d_coloff.p;
    const int* rowinfo = (const int*)X.d_rowinfo.p;
    for (const Launch& L : P.launches) {
        const PairEntry* pairs = (const PairEntry*)P.d_pairs.p + L.pair0;
        const TargetEntry* targets = (const TargetEntry*)P.d_targets.p + L.target0;
        timed(0, L.npairs, [&] {
            launch_hf2x_xyA(geo, d_appR.p, d_occR[TCX_PREC_FP32].p, pairs, L.npairs, colxy, coloff, d_pairbuf.p, cs,
                            stream);
        });
        timed(1, L.npairs, [&] {
            launch_hf2x_zB(geo, d_pairbuf.p, pairs, L.npairs, colxy, coloff, (const double*)P.d_kq.p, cs, stream);
        });
        timed(2, L.npairs, [&] {
            launch_hf2x_xyC(geo, d_pairbuf.p, d_occR[TCX_PREC_FP32].p, pairs, targets, L.ntargets, rowinfo,
                            (const double*)d_fill.p, (double*)d_W.p, (float*)d_Wc.p, cs, stream);
        });
    }
}

}  // namespace tcx
