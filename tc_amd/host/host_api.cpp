// host_api.cpp -- the flat QE input reader and the C ABI of include/tcx_host.h over
// tcx_host's k-mesh expansion and SCF loop (scf.hpp).
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tcx_host.h"
#include "scf.hpp"

namespace tcx_host {

namespace {

struct Record {
    std::string type;
    std::vector<char> data;
    int64_t count = 0;
};

std::map<std::string, Record> read_records(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::invalid_argument("cannot open " + path);
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "TCXQE001", 8) != 0) throw std::invalid_argument(path + ": not a TCXQE001 file");
    std::map<std::string, Record> out;
    for (;;) {
        char name[32], type[8];
        int64_t count;
        if (!f.read(name, 32)) break;
        f.read(type, 8);
        f.read((char*)&count, 8);
        if (!f) throw std::invalid_argument(path + ": truncated record header");
        Record r;
        r.type.assign(type, strnlen(type, 8));
        r.count = count;
        const size_t es = r.type == "<f8" ? 8 : r.type == "<i4" ? 4 : r.type == "<c16" ? 16 : 0;
        if (es == 0 || count < 0 || count > (int64_t{1} << 34)) throw std::invalid_argument(path + ": bad record");
        r.data.resize((size_t)count * es);
        f.read(r.data.data(), (std::streamsize)r.data.size());
        if (!f) throw std::invalid_argument(path + ": truncated record");
        out[std::string(name, strnlen(name, 32))] = std::move(r);
    }
    return out;
}

template <typename T>
std::vector<T> get(const std::map<std::string, Record>& m, const std::string& key, const char* type, size_t min_n = 1)
{
    auto it = m.find(key);
    if (it == m.end() || it->second.type != type) throw std::invalid_argument("missing record " + key);
    if ((size_t)it->second.count < min_n) throw std::invalid_argument("record " + key + " too short");
    std::vector<T> v((size_t)it->second.count);
    std::memcpy(v.data(), it->second.data.data(), it->second.data.size());
    return v;
}

}  // namespace

void read_qe_flat(const std::string& path, QEInputs& qe, SCFSettings& s)
{
    const auto m = read_records(path);
    qe = QEInputs();
    const auto cell = get<double>(m, "cell", "<f8", 9), bn = get<double>(m, "b_norm", "<f8", 9);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            qe.cell[i][j] = cell[3 * i + j];
            qe.b_norm[i][j] = bn[3 * i + j];
        }
    const auto grid = get<int32_t>(m, "grid", "<i4", 3);
    for (int d = 0; d < 3; ++d) qe.grid[d] = grid[d];
    const auto rot = get<int32_t>(m, "rot_qe", "<i4", 9);
    const auto tr = get<double>(m, "trans", "<f8", 3);
    if (rot.size() % 9 != 0 || tr.size() != rot.size() / 3) throw std::invalid_argument("rot_qe / trans sizes");
    for (size_t i = 0; i < rot.size() / 9; ++i) {
        std::array<int, 9> r;
        for (int e = 0; e < 9; ++e) r[e] = rot[9 * i + e];
        qe.rot_qe.push_back(r);
        qe.trans.push_back({tr[3 * i], tr[3 * i + 1], tr[3 * i + 2]});
    }
    qe.time_reversal = get<int32_t>(m, "time_reversal", "<i4")[0] != 0;
    qe.nelec = get<double>(m, "nelec", "<f8")[0];
    qe.kweight = get<double>(m, "kweight", "<f8");
    const int nirr = (int)qe.kweight.size();
    qe.nb = get<int32_t>(m, "nb", "<i4")[0];
    if (nirr < 1 || qe.nb < 1) throw std::invalid_argument("no k-points or bands");
    const auto kc = get<double>(m, "kcart", "<f8", (size_t)3 * nirr);
    const auto eig = get<double>(m, "eig", "<f8", (size_t)nirr * qe.nb);
    const auto npw = get<int32_t>(m, "npw", "<i4", nirr);
    const auto mill = get<int32_t>(m, "mill", "<i4", 0);
    const auto coef = get<Complex>(m, "coef", "<c16", 0);
    size_t mo = 0, co = 0;
    for (int ik = 0; ik < nirr; ++ik) {
        const size_t n = (size_t)qe.nb * npw[ik];
        if (npw[ik] < 0 || mo + 3 * (size_t)npw[ik] > mill.size() || co + n > coef.size())
            throw std::invalid_argument("mill / coef shorter than npw says");
        qe.kcart.push_back({kc[3 * ik], kc[3 * ik + 1], kc[3 * ik + 2]});
        qe.eig.emplace_back(eig.begin() + (size_t)ik * qe.nb, eig.begin() + (size_t)(ik + 1) * qe.nb);
        std::vector<std::array<int, 3>> mk;
        for (int i = 0; i < npw[ik]; ++i, mo += 3) mk.push_back({mill[mo], mill[mo + 1], mill[mo + 2]});
        qe.mill.push_back(mk);
        qe.coef.emplace_back(coef.begin() + co, coef.begin() + co + n);
        co += n;
    }
    s = SCFSettings();
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
    s.biortho_david = get<int32_t>(m, "biortho_david", "<i4")[0] != 0;
    s.A_input = get<double>(m, "A_input", "<f8")[0];
}

}  // namespace tcx_host

struct tcxh_qe {
    tcx_host::QEInputs qe;
    tcx_host::SCFSettings settings;
    std::vector<tcx_host::KReplica> reps;
};

struct tcxh_scf_result {
    tcx_host::SCFResult r;
    int nirr = 0, nb = 0;
};

namespace {

thread_local std::string g_error;

// Runs f, mapping exceptions to a status and the thread's error message.
template <typename F>
int guarded(const char* what, F&& f)
{
    try {
        const int st = f();
        if (st == TCX_OK) g_error.clear();
        return st;
    } catch (const tcx_host::StatusError& e) {  // an engine call failed: keep its status
        g_error = std::string(what) + ": " + e.what();
        return e.status;
    } catch (const std::invalid_argument& e) {
        g_error = std::string(what) + ": " + e.what();
        return TCX_ERR_ARG;
    } catch (const std::bad_alloc&) {
        g_error = std::string(what) + ": out of host memory";
        return TCX_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_error = std::string(what) + ": " + e.what();
        return TCX_ERR_INTERNAL;
    }
}

int arg_error(const char* what, const char* msg)
{
    g_error = std::string(what) + ": " + msg;
    return TCX_ERR_ARG;
}

}  // namespace

extern "C" {

const char* tcxh_last_error(void) { return g_error.c_str(); }

int tcxh_qe_load(const char* path, tcxh_qe** out)
{
    if (!path || !out) return arg_error("tcxh_qe_load", "null argument");
    *out = nullptr;
    return guarded("tcxh_qe_load", [&] {
        auto q = std::make_unique<tcxh_qe>();
        tcx_host::read_qe_flat(path, q->qe, q->settings);
        try {
            q->reps = tcx_host::replicate_kpoints(q->qe);
        } catch (const std::runtime_error& e) {  // singular cell or rotation: bad input
            throw std::invalid_argument(e.what());
        }
        *out = q.release();
        return (int)TCX_OK;
    });
}

void tcxh_qe_free(tcxh_qe* q) { delete q; }

int tcxh_qe_num_replicas(const tcxh_qe* q, int* nrep)
{
    if (!q || !nrep) return arg_error("tcxh_qe_num_replicas", "null argument");
    *nrep = (int)q->reps.size();
    return TCX_OK;
}

int tcxh_qe_replicas(const tcxh_qe* q, int* ik, int* isym, int* trev, double* kcrys)
{
    if (!q || !ik || !isym || !trev || !kcrys) return arg_error("tcxh_qe_replicas", "null argument");
    for (size_t r = 0; r < q->reps.size(); ++r) {
        const auto& k = q->reps[r];
        ik[r] = k.ik;
        isym[r] = k.isym;
        trev[r] = k.trev ? 1 : 0;
        for (int d = 0; d < 3; ++d) kcrys[3 * r + d] = k.kcrys[d];
    }
    return TCX_OK;
}

int tcxh_qe_rotation(const tcxh_qe* q, int isym, int* U)
{
    if (!q || !U) return arg_error("tcxh_qe_rotation", "null argument");
    if (isym < 0 || isym >= (int)q->qe.rot_qe.size()) return arg_error("tcxh_qe_rotation", "isym out of range");
    return guarded("tcxh_qe_rotation", [&] {
        const auto u = tcx_host::tc_rotation(q->qe.rot_qe[isym]);
        for (int e = 0; e < 9; ++e) U[e] = u[e];
        return (int)TCX_OK;
    });
}

int tcxh_qe_replica_npw(const tcxh_qe* q, int r, int* npw)
{
    if (!q || !npw) return arg_error("tcxh_qe_replica_npw", "null argument");
    if (r < 0 || r >= (int)q->reps.size()) return arg_error("tcxh_qe_replica_npw", "replica out of range");
    *npw = (int)q->qe.mill[q->reps[r].ik].size();
    return TCX_OK;
}

int tcxh_qe_replica_basis(const tcxh_qe* q, int r, int* gidx, double* phase)
{
    if (!q || !gidx || !phase) return arg_error("tcxh_qe_replica_basis", "null argument");
    if (r < 0 || r >= (int)q->reps.size()) return arg_error("tcxh_qe_replica_basis", "replica out of range");
    return guarded("tcxh_qe_replica_basis", [&] {
        std::vector<int> g;
        std::vector<tcx_host::Complex> ph;
        tcx_host::replica_basis(q->qe, q->reps[r], g, ph);
        std::memcpy(gidx, g.data(), g.size() * sizeof(int));
        std::memcpy(phase, ph.data(), ph.size() * sizeof(tcx_host::Complex));
        return (int)TCX_OK;
    });
}

int tcxh_scf_run(tcx_context* ctx, const tcxh_qe* q, int precision, int max_iterations, tcxh_scf_result** out)
{
    if (!ctx || !q || !out) return arg_error("tcxh_scf_run", "null argument");
    if (precision != TCX_PREC_FP64 && precision != TCX_PREC_FP32) return arg_error("tcxh_scf_run", "bad precision");
    *out = nullptr;
    return guarded("tcxh_scf_run", [&] {
        tcx_host::SCFSettings s = q->settings;
        s.precision = precision;
        if (max_iterations > 0) s.max_num_iterations = max_iterations;
        auto res = std::make_unique<tcxh_scf_result>();
        res->r = tcx_host::Diagonalization(s).scf(ctx, q->qe);
        res->nirr = q->qe.nirr();
        res->nb = q->qe.nb;
        *out = res.release();
        return (int)TCX_OK;
    });
}

int tcxh_scf_result_sizes(const tcxh_scf_result* r, int* nfill, int* nenergy, int* nddiff, int* iterations,
                          int* converged, int* nirr, int* nb)
{
    if (!r || !nfill || !nenergy || !nddiff || !iterations || !converged || !nirr || !nb)
        return arg_error("tcxh_scf_result_sizes", "null argument");
    *nfill = (int)r->r.fillings.size();
    *nenergy = (int)r->r.energies.size();
    *nddiff = (int)r->r.density_difference.size();
    *iterations = r->r.iterations;
    *converged = r->r.converged ? 1 : 0;
    *nirr = r->nirr;
    *nb = r->nb;
    return TCX_OK;
}

int tcxh_scf_result_fill(const tcxh_scf_result* r, int i, double* fermi, double* eig, double* filling)
{
    if (!r || !fermi || !eig || !filling) return arg_error("tcxh_scf_result_fill", "null argument");
    if (i < 0 || i >= (int)r->r.fillings.size()) return arg_error("tcxh_scf_result_fill", "index out of range");
    const auto& f = r->r.fillings[i];
    *fermi = f.fermi;
    for (int k = 0; k < r->nirr; ++k)
        for (int b = 0; b < r->nb; ++b) {
            const size_t o = (size_t)k * r->nb + b;
            eig[2 * o] = f.eig[k][b].real();
            eig[2 * o + 1] = f.eig[k][b].imag();
            filling[o] = f.filling[k][b];
        }
    return TCX_OK;
}

int tcxh_scf_result_energy(const tcxh_scf_result* r, int i, double* e)
{
    if (!r || !e) return arg_error("tcxh_scf_result_energy", "null argument");
    if (i < 0 || i >= (int)r->r.energies.size()) return arg_error("tcxh_scf_result_energy", "index out of range");
    for (int c = 0; c < 4; ++c) {
        e[2 * c] = r->r.energies[i][c].real();
        e[2 * c + 1] = r->r.energies[i][c].imag();
    }
    return TCX_OK;
}

int tcxh_scf_result_ddiff(const tcxh_scf_result* r, double* d)
{
    if (!r || !d) return arg_error("tcxh_scf_result_ddiff", "null argument");
    for (size_t i = 0; i < r->r.density_difference.size(); ++i) d[i] = r->r.density_difference[i];
    return TCX_OK;
}

void tcxh_scf_result_free(tcxh_scf_result* r) { delete r; }

}  // extern "C"
