// calc_hamiltonian.cpp -- host mirror of TC++'s term functions over the C ABI (tcx.h).
//
// Each call packs the reference-shaped arguments into the ABI's ragged
// complex128 arrays, brings the per-device context up to date with the state
// the term reads (cell, k-points, basis, occupied orbitals, Jastrow,
// divergence correction, k-mesh symmetry), runs the term and adds the result
// into HNphi.  The context keeps what was pushed: each piece of state is
// fingerprinted (a 64-bit hash of its bytes) and a setter runs only when its
// fingerprint changed, so the repeated calls of one Davidson block reuse the
// context's launch plans, real-space occupied orbitals and cached q-sums.
#include "calc_hamiltonian.hpp"

#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>

namespace {

// 64-bit fingerprint of byte ranges: 8-byte words folded by multiply-rotate
// (two independent lanes), the byte count mixed in at the end.
struct Fingerprint {
    uint64_t a = 0x9E3779B97F4A7C15ull, b = 0xC2B2AE3D27D4EB4Full, n = 0;
    static uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
    void add(const void* p, size_t bytes)
    {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        size_t i = 0;
        for (; i + 16 <= bytes; i += 16) {
            uint64_t w0, w1;
            std::memcpy(&w0, c + i, 8);
            std::memcpy(&w1, c + i + 8, 8);
            a = rotl(a ^ (w0 * 0x87C37B91114253D5ull), 31) * 0x4CF5AD432745937Full;
            b = rotl(b ^ (w1 * 0x4CF5AD432745937Full), 29) * 0x87C37B91114253D5ull;
        }
        for (; i < bytes; ++i) a = rotl(a ^ (c[i] * 0xFF51AFD7ED558CCDull), 23) * 0xC4CEB9FE1A85EC53ull;
        n += bytes;
    }
    template <typename T>
    void add_vec(const std::vector<T>& v)
    {
        const uint64_t sz = v.size();
        add(&sz, sizeof(sz));
        if (!v.empty()) add(v.data(), v.size() * sizeof(T));
    }
    uint64_t value() const
    {
        uint64_t h = a ^ rotl(b, 17) ^ (n * 0x9E3779B97F4A7C15ull);
        h ^= h >> 33;
        h *= 0xFF51AFD7ED558CCDull;
        h ^= h >> 33;
        return h;
    }
};

// What the context currently holds, by fingerprint (0: nothing pushed).
struct Pushed {
    uint64_t geometry = 0, states = 0, density = 0, jastrow = 0, div = 0, symmetry = 0;
};
std::map<const tcx_context*, Pushed>& pushed_state()
{
    static std::map<const tcx_context*, Pushed> m;
    return m;
}

tcx_context* context_for(int device)
{
    static std::map<int, tcx_context*> contexts;  // one per device for the process lifetime
    auto it = contexts.find(device);
    if (it != contexts.end()) return it->second;
    tcx_context* ctx = nullptr;
    if (tcx_create(&ctx, device) != TCX_OK) throw std::runtime_error("calc_hamiltonian: tcx_create failed");
    contexts[device] = ctx;
    return ctx;
}

void check(tcx_context* ctx, int status, const char* what)
{
    if (status != TCX_OK) throw std::runtime_error(std::string("calc_hamiltonian::") + what + ": " + tcx_last_error(ctx));
}

int method_id(const std::string& m)
{
    if (m == "HF") return TCX_METHOD_HF;
    if (m == "TC") return TCX_METHOD_TC;
    if (m == "BITC") return TCX_METHOD_BITC;
    throw std::runtime_error("calc_hamiltonian: unknown calc_method " + m);
}

// append [k][band < count[k]][spinor=0] coefficient vectors of one spin to a ragged array
void pack_bands(const std::vector<std::vector<std::vector<VectorXcd>>>& bands_of_k, const std::vector<int>& count,
                std::vector<double>& out)
{
    for (size_t k = 0; k < bands_of_k.size(); ++k)
        for (int b = 0; b < count[k]; ++b) {
            const VectorXcd& v = bands_of_k[k][b][0];
            const double* d = reinterpret_cast<const double*>(v.data());
            out.insert(out.end(), d, d + 2 * v.size());
        }
}

// The ABI's (spin, k)-blocked arrays of phi[spin][k][band][0] for nspin spins, spin-major:
// band counts nb[s nk + k] and the packed coefficients.
void pack_spins(const Phi4& phi, int nspin, int nk, std::vector<int>& nb, std::vector<double>& out)
{
    if ((int)phi.size() < nspin) throw std::runtime_error("calc_hamiltonian: phi does not cover every spin");
    nb.assign((size_t)nspin * nk, 0);
    out.clear();
    for (int s = 0; s < nspin; ++s) {
        if ((int)phi[s].size() != nk) throw std::runtime_error("calc_hamiltonian: phi does not cover every k");
        std::vector<int> c(nk);
        for (int k = 0; k < nk; ++k) c[k] = nb[(size_t)s * nk + k] = (int)phi[s][k].size();
        pack_bands(phi[s], c, out);
    }
}

void push_geometry(tcx_context* ctx, const Method& method, const CrystalStructure& cs, const Spin& spin,
                   const Kpoints& kp, const PlaneWaveBasis& pw)
{
    const int nspin = spin.num_independent_spins;
    if (nspin != 1 && nspin != 2) throw std::runtime_error("calc_hamiltonian: num_independent_spins must be 1 or 2");
    if (method.calc_mode != "SCF") throw std::runtime_error("calc_hamiltonian: only SCF mode");
    const int nk = (int)kp.kvectors_scf.size();
    double recip[9];
    for (int i = 0; i < 3; ++i)
        for (int d = 0; d < 3; ++d) recip[3 * i + d] = cs.reciprocal_vectors[i][d];
    const int meth = method_id(method.calc_method);
    std::vector<double> kf;
    for (const auto& k : kp.kvectors_scf) {
        if (k.size() < 3) throw std::runtime_error("calc_hamiltonian: a k-vector needs 3 components");
        kf.insert(kf.end(), k.begin(), k.begin() + 3);
    }
    std::vector<int> npw(nk), gidx;
    for (int k = 0; k < nk; ++k) {
        npw[k] = (int)pw.Gindex_at_k_scf.at(k).size();
        gidx.insert(gidx.end(), pw.Gindex_at_k_scf[k].begin(), pw.Gindex_at_k_scf[k].end());
    }
    Fingerprint f;
    f.add(recip, sizeof(recip));
    f.add(&cs.unit_cell_volume, sizeof(double));
    f.add(pw.size_FFT_grid_vec, sizeof(pw.size_FFT_grid_vec));
    f.add(&meth, sizeof(meth));
    f.add(&nspin, sizeof(nspin));
    f.add(&kp.num_kpoints, sizeof(int));
    f.add_vec(kf);
    f.add_vec(npw);
    f.add_vec(gidx);
    Pushed& P = pushed_state()[ctx];
    if (P.geometry != 0 && P.geometry == f.value()) return;
    // set_cell and set_kpoints drop everything that depends on them
    P = Pushed();
    check(ctx, tcx_set_cell(ctx, recip, cs.unit_cell_volume, pw.size_FFT_grid_vec, nspin, meth), "set_cell");
    check(ctx, tcx_set_kpoints(ctx, nk, kf.data(), kp.num_kpoints), "set_kpoints");
    check(ctx, tcx_set_basis(ctx, npw.data(), gidx.data()), "set_basis");
    P.geometry = f.value();
}

// fingerprint of bands [k][b < count[k]][spinor 0] without packing them
void add_bands(Fingerprint& f, const std::vector<std::vector<std::vector<VectorXcd>>>& bands_of_k,
               const std::vector<int>& count)
{
    for (size_t k = 0; k < bands_of_k.size(); ++k)
        for (int b = 0; b < count[k]; ++b) {
            const VectorXcd& v = bands_of_k[k].at(b).at(0);
            const uint64_t n = v.size();
            f.add(&n, sizeof(n));
            f.add(v.data(), v.size() * sizeof(Complex));
        }
}

// The k-mesh symmetry relation for tc3a2a4b2b5 (identity when Kpoints carries none).
void push_symmetry(tcx_context* ctx, const Symmetry& sym, const Kpoints& kp)
{
    Pushed& P = pushed_state()[ctx];
    const size_t nk = kp.kvectors_scf.size();
    std::vector<int> rot;
    std::vector<double> trans;
    if (!kp.irreducible_of.empty()) {
        if (kp.irreducible_of.size() != nk || kp.index_of_rotation.size() != nk || kp.time_reversal_used.size() != nk)
            throw std::runtime_error("calc_hamiltonian: Kpoints symmetry relation does not cover the mesh");
        rot.resize(9 * nk);
        trans.resize(3 * nk);
        for (size_t k = 0; k < nk; ++k) {
            const int s = kp.index_of_rotation[k];
            if (s < 0 || s >= (int)sym.rotation.size() || s >= (int)sym.translation.size())
                throw std::runtime_error("calc_hamiltonian: rotation index outside Symmetry");
            for (int i = 0; i < 9; ++i) rot[9 * k + i] = sym.rotation[s][i];
            for (int d = 0; d < 3; ++d) trans[3 * k + d] = sym.translation[s][d];
        }
    }
    Fingerprint f;
    f.add_vec(kp.irreducible_of);
    f.add_vec(rot);
    f.add_vec(trans);
    f.add_vec(kp.time_reversal_used);
    if (P.symmetry != 0 && P.symmetry == f.value()) return;
    // the symmetry setter drops the occupied-set caches and a caller-set density, not the states
    P.symmetry = 0;
    P.density = 0;
    if (kp.irreducible_of.empty())
        check(ctx, tcx_set_kpoint_symmetry(ctx, nullptr, nullptr, nullptr, nullptr), "set_kpoint_symmetry");
    else
        check(ctx,
              tcx_set_kpoint_symmetry(ctx, kp.irreducible_of.data(), rot.data(), trans.data(),
                                      kp.time_reversal_used.data()),
              "set_kpoint_symmetry");
    P.symmetry = f.value();
}

// sym: the k-mesh symmetry to push (tc3a2a4b2b5 and all); NULL keeps the context's
void push_state(tcx_context* ctx, const Method& method, const CrystalStructure& cs, const Potentials& pot,
                const Spin& spin, const Kpoints& kp, const PlaneWaveBasis& pw, const BlochStates& bs,
                const Symmetry* sym = nullptr)
{
    push_geometry(ctx, method, cs, spin, kp, pw);
    if (sym) push_symmetry(ctx, *sym, kp);
    Pushed& P = pushed_state()[ctx];
    const int nk = (int)kp.kvectors_scf.size();
    const int nspin = spin.num_independent_spins;
    // occupied orbitals: bands < num_occupied_bands (BlochStates::set_filling, f > 1e-8),
    // over (spin, k), spin-major
    std::vector<int> nocc;
    std::vector<double> fill;
    const bool bitc = method.calc_method == "BITC";
    Fingerprint fs;
    for (int s = 0; s < nspin; ++s) {
        const std::vector<int>& ns = bs.num_occupied_bands.at(s);
        if ((int)ns.size() != nk) throw std::runtime_error("calc_hamiltonian: num_occupied_bands does not cover every k");
        nocc.insert(nocc.end(), ns.begin(), ns.end());
        for (int k = 0; k < nk; ++k)
            for (int b = 0; b < ns[k]; ++b) fill.push_back(bs.filling.at(s).at(k).at(b));
        add_bands(fs, bs.phik_scf.at(s), ns);
        if (bitc) add_bands(fs, bs.phik_left_scf.at(s), ns);
    }
    fs.add_vec(nocc);
    fs.add_vec(fill);
    auto push_orbitals = [&] {
        std::vector<double> phi_occ, chi_occ;
        for (int s = 0; s < nspin; ++s) {
            pack_bands(bs.phik_scf[s], bs.num_occupied_bands[s], phi_occ);
            if (bitc) pack_bands(bs.phik_left_scf[s], bs.num_occupied_bands[s], chi_occ);
        }
        check(ctx, tcx_set_states(ctx, nocc.data(), fill.data(), phi_occ.data(), bitc ? chi_occ.data() : nullptr),
              "set_states");
    };
    if (P.states == 0 || P.states != fs.value()) {
        P.states = 0;
        push_orbitals();
        P.states = fs.value();
        P.density = 0;  // the context now holds the density of these orbitals
    }
    // a caller-supplied density (BlochStates::density_) replaces the orbital density
    Fingerprint fn;
    fn.add(&P.states, sizeof(P.states));
    for (const VectorXcd& d : bs.density) {
        const uint64_t n = d.size();
        fn.add(&n, sizeof(n));
        fn.add(d.data(), d.size() * sizeof(Complex));
    }
    if (!bs.density.empty() && P.density != fn.value()) {
        const int* g = pw.size_FFT_grid_vec;
        const size_t N = (size_t)g[0] * g[1] * g[2];
        if (bs.density[0].size() != N || (bs.density.size() > 1 && bs.density[1].size() != N))
            throw std::runtime_error("calc_hamiltonian: BlochStates::density must hold N points per spin");
        std::vector<Complex> both(2 * N);
        std::memcpy(both.data(), bs.density[0].data(), N * sizeof(Complex));
        std::memcpy(both.data() + N, (bs.density.size() > 1 ? bs.density[1] : bs.density[0]).data(),
                    N * sizeof(Complex));
        P.density = 0;
        check(ctx, tcx_set_density(ctx, reinterpret_cast<const double*>(both.data())), "set_density");
        P.density = fn.value();
    } else if (bs.density.empty() && P.density != 0) {
        // back to the orbital density: push the states again
        P.states = 0;
        P.density = 0;
        push_orbitals();
        P.states = fs.value();
    }
    double A[4], F[4];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            A[2 * i + j] = pot.jastrow.A_long[i][j];
            F[2 * i + j] = pot.jastrow.F_long[i][j];
        }
    Fingerprint fj;
    fj.add(A, sizeof(A));
    fj.add(F, sizeof(F));
    if (P.jastrow == 0 || P.jastrow != fj.value()) {
        P.jastrow = 0;
        check(ctx, tcx_set_jastrow_rpa(ctx, A, F), "set_jastrow_rpa");
        P.jastrow = fj.value();
    }
    const int div_on = pot.includes_div_correction ? 1 : 0;
    Fingerprint fd;
    fd.add(&div_on, sizeof(div_on));
    fd.add(&pot.alpha_Vaux, sizeof(double));
    fd.add_vec(pot.sum_of_Vaux_scf);
    if (P.div == 0 || P.div != fd.value()) {
        P.div = 0;
        check(ctx, tcx_set_div_correction(ctx, div_on, pot.alpha_Vaux,
                                          pot.sum_of_Vaux_scf.empty() ? nullptr : pot.sum_of_Vaux_scf.data()),
              "set_div_correction");
        P.div = fd.value();
    }
}

// the (spin, k)-blocked ragged array back into H[spin][k][band][0], spins < nspin
void unpack_bands(const std::vector<double>& buf, Phi4& H, int nspin)
{
    size_t off = 0;
    for (int s = 0; s < nspin; ++s)
        for (auto& k : H[s])
            for (auto& b : k) {
                VectorXcd& v = b[0];
                std::memcpy(reinterpret_cast<double*>(v.data()), buf.data() + off, sizeof(double) * 2 * v.size());
                off += 2 * v.size();
            }
}

using TermFn = int (*)(tcx_context*, const int*, const double*, double*, int, unsigned);

void run_term(TermFn fn, const char* name, const Parallelization& par, const Method& method,
              const CrystalStructure& cs, const Potentials& pot, const Spin& spin, const Kpoints& kp,
              const PlaneWaveBasis& pw, const BlochStates& bs, const Phi4& phi, Phi4& HNphi,
              const Symmetry* sym = nullptr)
{
    tcx_context* ctx = context_for(par.device);
    push_state(ctx, method, cs, pot, spin, kp, pw, bs, sym);
    const int nk = (int)kp.kvectors_scf.size(), nspin = spin.num_independent_spins;
    std::vector<int> nb, nbo;
    std::vector<double> in, out;
    pack_spins(phi, nspin, nk, nb, in);
    pack_spins(HNphi, nspin, nk, nbo, out);  // the term accumulates into the current H*phi
    if (nbo != nb) throw std::runtime_error(std::string("calc_hamiltonian::") + name + ": H*phi does not match phi");
    check(ctx, fn(ctx, nb.data(), in.data(), out.data(), method.precision, 0u), name);
    unpack_bands(out, HNphi, nspin);
}

int hf2h_term(tcx_context* c, const int* nb, const double* phi, double* out, int, unsigned f)
{
    return tcx_hf2h(c, nb, phi, out, f);
}
int tc2h_term(tcx_context* c, const int* nb, const double* phi, double* out, int, unsigned f)
{
    return tcx_tc2h(c, nb, phi, out, f);
}


}  // namespace

namespace calc_hamiltonian {

void all(const Parallelization& par, MyClock&, const Method& m, const CrystalStructure& cs, const Symmetry& sym,
         const Potentials& pot, const Spin& spin, const Kpoints& kp, PlaneWaveBasis& pw, const BlochStates& bs,
         const Phi4& phi, Phi4& H1phi, Phi4& H2phi, Phi4& H3phi, std::ostream*)
{
    tcx_context* ctx = context_for(par.device);
    push_state(ctx, m, cs, pot, spin, kp, pw, bs, &sym);
    const int nk = (int)kp.kvectors_scf.size(), nspin = spin.num_independent_spins;
    std::vector<int> nb;
    const bool tc = m.calc_method != "HF";
    std::vector<double> in, h1, h2, h3;
    pack_spins(phi, nspin, nk, nb, in);
    h1.assign(in.size(), 0.0);
    h2.assign(in.size(), 0.0);
    if (tc) h3.assign(in.size(), 0.0);
    check(ctx,
          tcx_apply_all(ctx, nb.data(), in.data(), h1.data(), h2.data(), tc ? h3.data() : nullptr, m.precision, 0u),
          "all");
    // H*phi arrays take the shape of phi (calc_hamiltonian_all.cpp:28-49)
    for (Phi4* H : {&H1phi, &H2phi, &H3phi}) {
        if (H == &H3phi && !tc) continue;
        H->assign(nspin, std::vector<std::vector<std::vector<VectorXcd>>>(nk));
        for (int s = 0; s < nspin; ++s)
            for (int k = 0; k < nk; ++k) {
                const int n = nb[(size_t)s * nk + k];
                (*H)[s][k].resize(n);
                for (int b = 0; b < n; ++b) (*H)[s][k][b].assign(1, VectorXcd(phi[s][k][b][0].size()));
            }
    }
    unpack_bands(h1, H1phi, nspin);
    unpack_bands(h2, H2phi, nspin);
    if (tc) unpack_bands(h3, H3phi, nspin);
}

void kinetic(const Parallelization& par, const Method& m, const CrystalStructure& cs, const Spin& spin,
             const Kpoints& kp, PlaneWaveBasis& pw, const Phi4& phi, Phi4& H1phi, std::ostream*)
{
    // the one-body term reads no orbital state: only the cell, k-points and basis are pushed
    tcx_context* ctx = context_for(par.device);
    push_geometry(ctx, m, cs, spin, kp, pw);
    const int nk = (int)kp.kvectors_scf.size(), nspin = spin.num_independent_spins;
    std::vector<int> nb, nbo;
    std::vector<double> in, out;
    pack_spins(phi, nspin, nk, nb, in);
    pack_spins(H1phi, nspin, nk, nbo, out);
    if (nbo != nb) throw std::runtime_error("calc_hamiltonian::kinetic: H*phi does not match phi");
    check(ctx, tcx_kinetic(ctx, nb.data(), in.data(), out.data(), 0u), "kinetic");
    unpack_bands(out, H1phi, nspin);
}

void hf2h(const Parallelization& par, const Method& m, const CrystalStructure& cs, const Potentials& pot,
          const Spin& spin, const Kpoints& kp, PlaneWaveBasis& pw, const BlochStates& bs, const Phi4& phi,
          Phi4& H2phi, std::ostream*)
{
    run_term(hf2h_term, "hf2h", par, m, cs, pot, spin, kp, pw, bs, phi, H2phi);
}

void tc2h(const Parallelization& par, const Method& m, const CrystalStructure& cs, const Potentials& pot,
          const Spin& spin, const Kpoints& kp, PlaneWaveBasis& pw, const BlochStates& bs, const Phi4& phi,
          Phi4& H2phi, std::ostream*)
{
    run_term(tc2h_term, "tc2h", par, m, cs, pot, spin, kp, pw, bs, phi, H2phi);
}

void tc3a3b3b4b6(const Parallelization& par, const Method& m, const CrystalStructure& cs, const Potentials& pot,
                 const Spin& spin, const Kpoints& kp, PlaneWaveBasis& pw, const BlochStates& bs, const Phi4& phi,
                 Phi4& H3phi, std::ostream*)
{
    run_term(tcx_tc3a3b3b4b6, "tc3a3b3b4b6", par, m, cs, pot, spin, kp, pw, bs, phi, H3phi);
}

void tc3a2a4b2b5(const Parallelization& par, const Method& m, const CrystalStructure& cs, const Symmetry& sym,
                 const Potentials& pot, const Spin& spin, const Kpoints& kp, PlaneWaveBasis& pw,
                 const BlochStates& bs, const Phi4& phi, Phi4& H3phi, std::ostream*)
{
    run_term(tcx_tc3a2a4b2b5, "tc3a2a4b2b5", par, m, cs, pot, spin, kp, pw, bs, phi, H3phi, &sym);
}

void hf2x(const Parallelization& par, const Method& m, const CrystalStructure& cs, const Potentials& pot,
          const Spin& spin, const Kpoints& kp, PlaneWaveBasis& pw, const BlochStates& bs, const Phi4& phi,
          Phi4& H2phi, std::ostream*)
{
    run_term(tcx_hf2x, "hf2x", par, m, cs, pot, spin, kp, pw, bs, phi, H2phi);
}

void tc2x(const Parallelization& par, const Method& m, const CrystalStructure& cs, const Potentials& pot,
          const Spin& spin, const Kpoints& kp, PlaneWaveBasis& pw, const BlochStates& bs, const Phi4& phi,
          Phi4& H2phi, std::ostream*)
{
    run_term(tcx_tc2x, "tc2x", par, m, cs, pot, spin, kp, pw, bs, phi, H2phi);
}

void tc3a1(const Parallelization& par, const Method& m, const CrystalStructure& cs, const Potentials& pot,
           const Spin& spin, const Kpoints& kp, PlaneWaveBasis& pw, const BlochStates& bs, const Phi4& phi,
           Phi4& H3phi, std::ostream*)
{
    run_term(tcx_tc3a1, "tc3a1", par, m, cs, pot, spin, kp, pw, bs, phi, H3phi);
}

void tc3b1(const Parallelization& par, const Method& m, const CrystalStructure& cs, const Potentials& pot,
           const Spin& spin, const Kpoints& kp, PlaneWaveBasis& pw, const BlochStates& bs, const Phi4& phi,
           Phi4& H3phi, std::ostream*)
{
    run_term(tcx_tc3b1, "tc3b1", par, m, cs, pot, spin, kp, pw, bs, phi, H3phi);
}

}  // namespace calc_hamiltonian
