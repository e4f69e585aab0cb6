// C++ host SCF (tc_amd/host/scf.hpp) on one of the reference's QE examples.
//   scf_host_test <inputs.tcxqe> <precision 0|1>
// The input file is the flat QE input of include/tcx_host.h (tc_amd/qe.py write_flat).
// Prints one line per printed quantity of the run:
//   FILL <fermi> <eigenvalue real parts [ik][band]>   (each set_filling)
//   ENERGY <E1 re im> <E2 re im> <E3 re im> <Etot re im> (each calc_total_energy)
//   DDIFF <density difference>                         (each set_density after the first)
//   DONE <iterations> <converged>
// With "mirror <out>" it applies calc_hamiltonian::tc3a2a4b2b5 of the C++ host mirror
// (tc_amd/host/calc_hamiltonian.hpp) in FP64 to the occupied bands of every k of the
// expanded mesh, fillings from the QE eigenvalues, once with the mesh's Symmetry and
// Kpoints relation and once without, and writes both H3*phi (complex128, concatenated)
// to <out>.  Exit status 0 when the run completes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../tc_amd/host/calc_hamiltonian.hpp"
#include "../tc_amd/host/scf.hpp"

using tcx_host::Complex;

// calc_hamiltonian::tc3a2a4b2b5 through the host mirror on the expanded mesh of a QE example.
static int run_mirror(const tcx_host::QEInputs& qe, const tcx_host::SCFSettings& s, const char* out_path)
{
    const double pi = 3.14159265358979323846;
    const auto reps = tcx_host::replicate_kpoints(qe);
    const int nk = (int)reps.size(), nirr = qe.nirr(), nb = qe.nb;
    // cell
    const double(*a)[3] = qe.cell;
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                       a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    CrystalStructure cs;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {  // rows b_i = 2 pi (A^-1)^T: cofactor / det
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            cs.reciprocal_vectors[i][j] = 2.0 * pi * (a[i1][j1] * a[i2][j2] - a[i1][j2] * a[i2][j1]) / det;
        }
    cs.unit_cell_volume = std::fabs(det);
    // fillings from the QE eigenvalues
    std::vector<int> nbk(nirr, nb), nocc(nirr);
    std::vector<double> e((size_t)nirr * nb), f((size_t)nirr * nb);
    for (int ik = 0; ik < nirr; ++ik)
        for (int b = 0; b < nb; ++b) e[(size_t)ik * nb + b] = qe.eig[ik][b];
    double mu = 0.0;
    if (tcx_set_filling(nirr, nbk.data(), e.data(), qe.kweight.data(), qe.nelec, s.smearing, s.smearing_width,
                        f.data(), &mu, nocc.data()) != TCX_OK)
        throw std::runtime_error("set_filling failed");
    // flattened mesh, basis, orbitals, symmetry relation
    Kpoints kp;
    PlaneWaveBasis pw;
    BlochStates bs;
    Symmetry sym;
    for (int d = 0; d < 3; ++d) pw.size_FFT_grid_vec[d] = qe.grid[d];
    for (size_t i = 0; i < qe.rot_qe.size(); ++i) {
        sym.rotation.push_back(tcx_host::tc_rotation(qe.rot_qe[i]));
        sym.translation.push_back(qe.trans[i]);
    }
    bs.phik_scf.assign(1, {});
    bs.filling.assign(1, {});
    bs.num_occupied_bands.assign(1, {});
    Phi4 phi(1);
    std::vector<int> first(nirr, -1);
    for (int k = 0; k < nk; ++k)
        if (first[reps[k].ik] < 0) first[reps[k].ik] = k;
    for (int k = 0; k < nk; ++k) {
        const auto& r = reps[k];
        std::vector<int> gidx;
        std::vector<Complex> ph;
        tcx_host::replica_basis(qe, r, gidx, ph);
        kp.kvectors_scf.push_back({r.kcrys[0], r.kcrys[1], r.kcrys[2]});
        kp.irreducible_of.push_back(first[r.ik]);
        kp.index_of_rotation.push_back(r.isym);
        kp.time_reversal_used.push_back(r.trev ? 1 : 0);
        pw.Gindex_at_k_scf.push_back(gidx);
        const size_t np = gidx.size();
        std::vector<std::vector<VectorXcd>> bands(nb, std::vector<VectorXcd>(1, VectorXcd(np)));
        for (int b = 0; b < nb; ++b)
            for (size_t i = 0; i < np; ++i) {
                Complex c = qe.coef[r.ik][(size_t)b * np + i];
                if (r.trev) c = std::conj(c);
                bands[b][0][i] = c * ph[i];
            }
        bs.phik_scf[0].push_back(bands);
        bs.filling[0].emplace_back(f.begin() + (size_t)r.ik * nb, f.begin() + (size_t)(r.ik + 1) * nb);
        bs.num_occupied_bands[0].push_back(nocc[r.ik]);
        phi[0].emplace_back(bands.begin(), bands.begin() + nocc[r.ik]);
    }
    kp.num_kpoints = nk;
    bs.phik_left_scf = bs.phik_scf;
    // Jastrow (jastrow_initialize.cpp:28-75) and divergence correction
    Potentials pot;
    const double aj = s.A_input * std::sqrt(cs.unit_cell_volume / (4.0 * pi * qe.nelec));
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            pot.jastrow.A_long[i][j] = aj;
            pot.jastrow.F_long[i][j] = std::sqrt((i == j ? 2.0 : 1.0) * aj);
        }
    std::vector<double> kf;
    for (const auto& k : kp.kvectors_scf) kf.insert(kf.end(), k.begin(), k.end());
    pot.sum_of_Vaux_scf.resize(nk);
    double recip[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) recip[3 * i + j] = cs.reciprocal_vectors[i][j];
    if (tcx_sum_of_vaux(recip, cs.unit_cell_volume, nk, kf.data(), &pot.alpha_Vaux, pot.sum_of_Vaux_scf.data()) !=
        TCX_OK)
        throw std::runtime_error("sum_of_vaux failed");
    Method m;
    m.calc_method = s.method == TCX_METHOD_BITC ? "BITC" : s.method == TCX_METHOD_TC ? "TC" : "HF";
    m.precision = TCX_PREC_FP64;
    Parallelization par;
    Spin spin;
    std::ofstream out(out_path, std::ios::binary);
    for (int with_sym = 1; with_sym >= 0; --with_sym) {
        Kpoints kq = kp;
        if (!with_sym) {
            kq.irreducible_of.clear();
            kq.index_of_rotation.clear();
            kq.time_reversal_used.clear();
        }
        Phi4 H3 = phi;
        for (auto& k : H3[0])
            for (auto& b : k) std::fill(b[0].begin(), b[0].end(), Complex(0.0));
        calc_hamiltonian::tc3a2a4b2b5(par, m, cs, sym, pot, spin, kq, pw, bs, phi, H3, nullptr);
        for (const auto& k : H3[0])
            for (const auto& b : k) out.write((const char*)b[0].data(), (std::streamsize)(b[0].size() * sizeof(Complex)));
    }
    std::printf("MIRROR %d k-points written\n", nk);
    return out ? 0 : 1;
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s inputs.tcxqe precision\n", argv[0]);
        return 2;
    }
    tcx_context* ctx = nullptr;
    try {
        tcx_host::QEInputs qe;
        tcx_host::SCFSettings s;
        tcx_host::read_qe_flat(argv[1], qe, s);
        if (std::string(argv[2]) == "mirror") {
            if (argc < 4) throw std::runtime_error("mirror mode needs an output path");
            return run_mirror(qe, s, argv[3]);
        }
        s.precision = std::atoi(argv[2]);

        if (tcx_create(&ctx, 0) != TCX_OK) throw std::runtime_error("tcx_create failed");
        const tcx_host::SCFResult r = tcx_host::Diagonalization(s).scf(ctx, qe);
        for (const auto& f : r.fillings) {
            std::printf("FILL %.17g", f.fermi);
            for (const auto& row : f.eig)
                for (const Complex& e : row) std::printf(" %.17g", e.real());
            std::printf("\n");
        }
        for (const auto& e : r.energies) {
            std::printf("ENERGY");
            for (const Complex& x : e) std::printf(" %.17g %.17g", x.real(), x.imag());
            std::printf("\n");
        }
        for (double d : r.density_difference) std::printf("DDIFF %.17g\n", d);
        std::printf("DONE %d %d\n", r.iterations, r.converged ? 1 : 0);
        tcx_destroy(ctx);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "scf_host_test: %s\n", e.what());
        if (ctx) tcx_destroy(ctx);
        return 1;
    }
}
