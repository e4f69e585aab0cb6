// C++ host SCF (tc_amd/host/scf.hpp) on one of the reference's QE examples.
//   scf_host_test <inputs.tcxqe> <precision 0|1>
// The input file holds the example's QE data and SCF settings as named records
// (tc_amd/qe.py write_flat).  Prints one line per printed quantity of the run:
//   FILL <fermi> <eigenvalue real parts [ik][band]>   (each set_filling)
//   ENERGY <E1 re im> <E2 re im> <E3 re im> <Etot re im> (each calc_total_energy)
//   DDIFF <density difference>                         (each set_density after the first)
//   DONE <iterations> <converged>
// With "replicas" in place of the precision it prints the symmetry-expanded k-mesh
// without touching the GPU: per replica "REP ik isym trev k1 k2 k3 npw sum(gidx)
// sum(Re phase) sum(Im phase)".  Exit status 0 when the run completes.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../tc_amd/host/scf.hpp"

using tcx_host::Complex;

struct Record {
    std::string type;
    std::vector<char> data;
    int64_t count = 0;
};

static std::map<std::string, Record> read_records(const char* path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "TCXQE001", 8) != 0) throw std::runtime_error("bad input file");
    std::map<std::string, Record> out;
    for (;;) {
        char name[32], type[8];
        int64_t count;
        if (!f.read(name, 32)) break;
        f.read(type, 8);
        f.read((char*)&count, 8);
        Record r;
        r.type.assign(type, strnlen(type, 8));
        r.count = count;
        const size_t es = r.type == "<f8" ? 8 : r.type == "<i4" ? 4 : r.type == "<c16" ? 16 : 0;
        if (es == 0 || count < 0) throw std::runtime_error("bad record type");
        r.data.resize((size_t)count * es);
        f.read(r.data.data(), (std::streamsize)r.data.size());
        if (!f) throw std::runtime_error("truncated input file");
        out[std::string(name, strnlen(name, 32))] = std::move(r);
    }
    return out;
}

template <typename T>
static std::vector<T> get(const std::map<std::string, Record>& m, const std::string& key, const char* type)
{
    auto it = m.find(key);
    if (it == m.end() || it->second.type != type) throw std::runtime_error("missing record " + key);
    std::vector<T> v((size_t)it->second.count);
    std::memcpy(v.data(), it->second.data.data(), it->second.data.size());
    return v;
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s inputs.tcxqe precision\n", argv[0]);
        return 2;
    }
    tcx_context* ctx = nullptr;
    try {
        const auto m = read_records(argv[1]);
        tcx_host::QEInputs qe;
        const auto cell = get<double>(m, "cell", "<f8"), bn = get<double>(m, "b_norm", "<f8");
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                qe.cell[i][j] = cell[3 * i + j];
                qe.b_norm[i][j] = bn[3 * i + j];
            }
        const auto grid = get<int32_t>(m, "grid", "<i4");
        for (int d = 0; d < 3; ++d) qe.grid[d] = grid[d];
        const auto rot = get<int32_t>(m, "rot_qe", "<i4");
        const auto tr = get<double>(m, "trans", "<f8");
        for (size_t s = 0; s < rot.size() / 9; ++s) {
            std::array<int, 9> r;
            for (int i = 0; i < 9; ++i) r[i] = rot[9 * s + i];
            qe.rot_qe.push_back(r);
            qe.trans.push_back({tr[3 * s], tr[3 * s + 1], tr[3 * s + 2]});
        }
        qe.time_reversal = get<int32_t>(m, "time_reversal", "<i4")[0] != 0;
        qe.nelec = get<double>(m, "nelec", "<f8")[0];
        const auto kc = get<double>(m, "kcart", "<f8");
        qe.kweight = get<double>(m, "kweight", "<f8");
        const int nirr = (int)qe.kweight.size();
        qe.nb = get<int32_t>(m, "nb", "<i4")[0];
        const auto eig = get<double>(m, "eig", "<f8");
        const auto npw = get<int32_t>(m, "npw", "<i4");
        const auto mill = get<int32_t>(m, "mill", "<i4");
        const auto coef = get<Complex>(m, "coef", "<c16");
        size_t mo = 0, co = 0;
        for (int ik = 0; ik < nirr; ++ik) {
            qe.kcart.push_back({kc[3 * ik], kc[3 * ik + 1], kc[3 * ik + 2]});
            qe.eig.emplace_back(eig.begin() + (size_t)ik * qe.nb, eig.begin() + (size_t)(ik + 1) * qe.nb);
            std::vector<std::array<int, 3>> mk;
            for (int i = 0; i < npw[ik]; ++i, mo += 3) mk.push_back({mill[mo], mill[mo + 1], mill[mo + 2]});
            qe.mill.push_back(mk);
            const size_t n = (size_t)qe.nb * npw[ik];
            qe.coef.emplace_back(coef.begin() + co, coef.begin() + co + n);
            co += n;
        }
        tcx_host::SCFSettings s;
        s.method = get<int32_t>(m, "method", "<i4")[0];
        s.smearing = get<int32_t>(m, "smearing", "<i4")[0];
        s.smearing_width = get<double>(m, "smearing_width", "<f8")[0];
        s.mixing_beta = get<double>(m, "mixing_beta", "<f8")[0];
        s.mixes_density_matrix = get<int32_t>(m, "mixes_density_matrix", "<i4")[0] != 0;
        s.max_num_iterations = get<int32_t>(m, "max_num_iterations", "<i4")[0];
        s.energy_tolerance = get<double>(m, "energy_tolerance", "<f8")[0];
        s.charge_tolerance = get<double>(m, "charge_tolerance", "<f8")[0];
        s.max_num_blocks_david = get<int32_t>(m, "max_num_blocks_david", "<i4")[0];
        s.num_refresh_david = get<int32_t>(m, "num_refresh_david", "<i4")[0];
        s.A_input = get<double>(m, "A_input", "<f8")[0];
        if (std::string(argv[2]) == "replicas") {
            for (const auto& r : tcx_host::replicate_kpoints(qe)) {
                std::vector<int> gidx;
                std::vector<Complex> ph;
                tcx_host::replica_basis(qe, r, gidx, ph);
                long long gs = 0;
                Complex ps = 0.0;
                for (size_t i = 0; i < gidx.size(); ++i) {
                    gs += gidx[i];
                    ps += ph[i];
                }
                std::printf("REP %d %d %d %.17g %.17g %.17g %zu %lld %.17g %.17g\n", r.ik, r.isym, r.trev ? 1 : 0,
                            r.kcrys[0], r.kcrys[1], r.kcrys[2], gidx.size(), gs, ps.real(), ps.imag());
            }
            return 0;
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
