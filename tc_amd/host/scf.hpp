// scf.hpp -- C++ host of TC++'s SCF loop over the C ABI.
//
// Mirrors, for the homogeneous-electron-gas runs the engine covers:
//   * Symmetry::set (src/symmetry.cpp:17-33): TC++ rotation U = (QE rotation)^-1;
//   * Kpoints::set_kvectors_by_symmetry (src/kpoints.cpp:34-110): the full k-mesh as
//     replicas k' = +-U^T k of the irreducible k, in index_all_kscf order;
//   * PlaneWaveBasis::set_Gindex_at_k / get_orbital_FFTgrid
//     (src/plane_wave_basis.cpp:115-140, 221-271): c(G) of an irreducible orbital placed at
//     G' = +-U^T G (conjugated under time reversal) times exp(2 pi i G'.tau);
//   * Diagonalization::scf (src/diagonalization.cpp:154-303) with BlochStates::set_filling,
//     mix_density_matrix, recover_filling_for_density_matrix, set_density
//     (src/bloch_states_scfloop_filling_density.cpp:9-358), BlochStates::reset_phik's old
//     bands (src/bloch_states_scfloop_phik.cpp:86-99) and TotalEnergy::calc_total_energy
//     (src/total_energy.cpp:101-170).
// The operator work (H.psi, the Davidson, the density) runs on the GPU through
// include/tcx.h; this file only sequences it.  Scope: SCF mode, one independent spin,
// no pseudopotentials (is_heg), no forces, biortho_david = false.
//
// Error behaviour: the reference stops the run (error_messages::stop); here a failing
// engine call throws tcx_host::StatusError (a std::runtime_error) carrying its tcx_status
// and tcx_last_error(), and invalid inputs throw std::runtime_error.
#pragma once

#include <array>
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tcx.h"

namespace tcx_host {

using Complex = std::complex<double>;

// A failed engine call: its status and message.
struct StatusError : std::runtime_error {
    int status;
    StatusError(int st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

// What TC++ reads from the QE save directory (io_qe_files_read_xml.cpp:55-273,
// io_qe_files_read_wfc.f90:69-140), for one independent spin.
struct QEInputs {
    double cell[3][3] = {{0}};    // rows a_i (bohr)
    double b_norm[3][3] = {{0}};  // rows b_i in 2 pi / alat
    int grid[3] = {0, 0, 0};
    std::vector<std::array<int, 9>> rot_qe;      // QE rotations, crystal, row-major
    std::vector<std::array<double, 3>> trans;    // fractional translations (crystal)
    bool time_reversal = true;                   // noinv != "true"
    double nelec = 0.0;
    std::vector<std::array<double, 3>> kcart;    // irreducible k in 2 pi / alat
    std::vector<double> kweight;                 // QE weights
    int nb = 0;                                  // num_bands_tc
    std::vector<std::vector<double>> eig;        // [ik][nb] Hartree (QE eigenvalues)
    std::vector<std::vector<std::array<int, 3>>> mill;  // [ik][ipw] Miller indices
    std::vector<std::vector<Complex>> coef;      // [ik][band * npw + ipw]
    int nirr() const { return (int)kcart.size(); }
};

// Diagonalization's SCF parameters (include/diagonalization.hpp:75-89) plus the run's method.
struct SCFSettings {
    int method = TCX_METHOD_HF;
    int smearing = 1;             // 0 fixed, 1 gaussian (Kpoints::smearing_mode)
    double smearing_width = 0.01;
    double mixing_beta = 0.7;
    bool mixes_density_matrix = false;
    int max_num_iterations = 30;
    double energy_tolerance = 1e-6;
    double charge_tolerance = 1e-5;
    int max_num_blocks_david = 2;
    int num_refresh_david = 1;
    bool biortho_david = false;   // BITC: bi-orthogonal block Davidson (diagonalization.cpp:117-122)
    double A_input = 1.0;         // RPA Jastrow scale (jastrow_initialize.cpp:28-50)
    int precision = TCX_PREC_FP64;
};

struct KReplica {
    int ik, isym;
    bool trev;
    double kcrys[3];
};

struct FillingStep {
    double fermi;                      // Hartree
    std::vector<std::vector<Complex>> eig;  // [ik][nb] eigenvalues the fillings came from
    std::vector<std::vector<double>> filling;
};

struct SCFResult {
    std::vector<FillingStep> fillings;             // one per set_filling
    std::vector<std::array<Complex, 4>> energies;  // E_1, E_2, E_3, E_total per calc_total_energy
    std::vector<double> density_difference;        // per set_density after the first
    bool converged = false;
    int iterations = 0;                             // Davidson calls
};

// U = (QE rotation)^-1, exact integer inverse (symmetry.cpp:17-33)
std::array<int, 9> tc_rotation(const std::array<int, 9>& rot_qe);

// The full k-mesh, irreducible k first in each group (kpoints.cpp:34-110)
std::vector<KReplica> replicate_kpoints(const QEInputs& qe);

// FFT-grid indices and phases of the irreducible plane waves at a replica
void replica_basis(const QEInputs& qe, const KReplica& r, std::vector<int>& gidx, std::vector<Complex>& phase);

// Reads the flat input file of include/tcx_host.h (QE inputs and SCF settings); throws
// std::invalid_argument for a missing or malformed file.
void read_qe_flat(const std::string& path, QEInputs& qe, SCFSettings& s);

class Diagonalization {
public:
    explicit Diagonalization(const SCFSettings& s) : set_(s) {}
    // Runs the SCF loop from the QE orbitals on the context's GPU.
    SCFResult scf(tcx_context* ctx, const QEInputs& qe) const;

private:
    SCFSettings set_;
};

}  // namespace tcx_host
