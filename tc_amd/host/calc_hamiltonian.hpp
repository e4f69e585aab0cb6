// calc_hamiltonian.hpp -- C++ host mirror of TC++'s operator interface over the C ABI.
//
// Same namespace, function names and argument roles as the reference's
// src/include/calc_hamiltonian.hpp:43-108:
//   void T(const Parallelization&, const Method&, const CrystalStructure&,
//          const Potentials&, const Spin&, const Kpoints&, PlaneWaveBasis&,
//          const BlochStates&, const phi[spin][k][band][spinor],
//          HNphi[spin][k][band][spinor] (accumulated, +=), std::ostream*)
// The state classes below are minimal, builder-owned stand-ins carrying only
// what these terms read from the reference's classes; a maintainer wiring the
// engine into TC++ maps the reference objects onto them.
// Eigen::VectorXcd becomes std::vector<std::complex<double>>.
//
// Error behaviour: the reference stops the run (error_messages::stop ->
// MPI_Abort, src/error_messages.cpp:60-64); here a failing call throws
// std::runtime_error carrying tcx_last_error().  Scope: SCF mode, no spinors;
// num_independent_spins 1, or 2 for kinetic, hf2h, hf2x, tc2x and the HF all (the
// engine refuses the spin-coupling terms with two spins); the k-mesh is passed flattened
// (every symmetry replica a k of its own), with its symmetry relation for tc3a2a4b2b5.
#pragma once

#include <array>
#include <complex>
#include <memory>
#include <ostream>
#include <string>
#include <vector>

#include "../../include/tcx.h"

using Complex = std::complex<double>;
using VectorXcd = std::vector<Complex>;
using Phi4 = std::vector<std::vector<std::vector<std::vector<VectorXcd>>>>;  // [spin][k][band][spinor]

struct Parallelization {
    int device = 0;  // one GPU per process; no MPI rank split (every band "assigned")
};

struct Method {
    std::string calc_method = "HF";  // "HF", "TC" or "BITC"
    std::string calc_mode = "SCF";
    // TCX_PREC_FP32: FP32 FFTs/products with FP64 pair sums; TCX_PREC_FP64: all FP64
    int precision = TCX_PREC_FP32;
};

struct CrystalStructure {
    double reciprocal_vectors[3][3] = {{0}};  // rows b_i (bohr^-1)
    double unit_cell_volume = 0.0;
};

struct Spin {
    int num_independent_spins = 1;
};

struct Kpoints {
    std::vector<std::vector<double>> kvectors_scf;  // [k][3] crystal coordinates, every k of the mesh
    int num_kpoints = 0;                             // Nk of the 1/(Nk V) prefactors
    // Symmetry relation of the flattened mesh (Kpoints::index_of_rotation_at_k,
    // is_time_reversal_used_at_k; src/kpoints.cpp:34-110); empty: every k irreducible.
    std::vector<int> irreducible_of;      // [k] mesh index of the irreducible k whose orbitals k's are
    std::vector<int> index_of_rotation;   // [k] index into Symmetry::rotation
    std::vector<int> time_reversal_used;  // [k] 0 or 1
};

struct PlaneWaveBasis {
    int size_FFT_grid_vec[3] = {0, 0, 0};
    std::vector<std::vector<int>> Gindex_at_k_scf;  // [k][ipw_at_k] -> FFT-grid index
};

struct Jastrow {
    double A_long[2][2] = {{0, 0}, {0, 0}};  // atomic units (after unnormalize_A_long)
    double F_long[2][2] = {{1, 1}, {1, 1}};  // after impose_cusp
};

struct Potentials {
    bool includes_div_correction = true;
    double alpha_Vaux = 0.0;
    std::vector<double> sum_of_Vaux_scf;  // [k]
    Jastrow jastrow;
};

struct BlochStates {
    Phi4 phik_scf;       // [spin][k][band][spinor], occupied + unoccupied bands
    Phi4 phik_left_scf;  // BITC left orbitals (same shape) or empty
    std::vector<std::vector<std::vector<double>>> filling;  // [spin][k][band]
    std::vector<std::vector<int>> num_occupied_bands;        // [spin][k]
    // BlochStates::density_ (src/bloch_states_scfloop_filling_density.cpp:256-358): the
    // real-space density [spin][N] the terms read, e.g. after density mixing.  Empty: the
    // density of the occupied orbitals above.
    std::vector<VectorXcd> density;
};

// Symmetry::rotation / translation (src/symmetry.cpp:17-33): TC++ rotations U
// (integer, crystal, row-major) and fractional translations; read by tc3a2a4b2b5
// (and by all) together with the Kpoints relation.
struct Symmetry {
    std::vector<std::array<int, 9>> rotation;
    std::vector<std::array<double, 3>> translation;
};
struct MyClock {};   // timing printouts of the reference are not mirrored

namespace calc_hamiltonian {

// src/include/calc_hamiltonian.hpp:10-21: zeroes H1/H2/H3 phi, then every term
void all(const Parallelization&, MyClock&, const Method&, const CrystalStructure&, const Symmetry&,
         const Potentials&, const Spin&, const Kpoints&, PlaneWaveBasis&, const BlochStates&, const Phi4& phi,
         Phi4& H1phi, Phi4& H2phi, Phi4& H3phi, std::ostream* ost);
void kinetic(const Parallelization&, const Method&, const CrystalStructure&, const Spin&, const Kpoints&,
             PlaneWaveBasis&, const Phi4& phi, Phi4& H1phi, std::ostream* ost);
void hf2h(const Parallelization&, const Method&, const CrystalStructure&, const Potentials&, const Spin&,
          const Kpoints&, PlaneWaveBasis&, const BlochStates&, const Phi4& phi, Phi4& H2phi, std::ostream* ost);
void tc2h(const Parallelization&, const Method&, const CrystalStructure&, const Potentials&, const Spin&,
          const Kpoints&, PlaneWaveBasis&, const BlochStates&, const Phi4& phi, Phi4& H2phi, std::ostream* ost);
void tc3a3b3b4b6(const Parallelization&, const Method&, const CrystalStructure&, const Potentials&, const Spin&,
                 const Kpoints&, PlaneWaveBasis&, const BlochStates&, const Phi4& phi, Phi4& H3phi,
                 std::ostream* ost);
void tc3a2a4b2b5(const Parallelization&, const Method&, const CrystalStructure&, const Symmetry&,
                 const Potentials&, const Spin&, const Kpoints&, PlaneWaveBasis&, const BlochStates&,
                 const Phi4& phi, Phi4& H3phi, std::ostream* ost);

void hf2x(const Parallelization&, const Method&, const CrystalStructure&, const Potentials&, const Spin&,
          const Kpoints&, PlaneWaveBasis&, const BlochStates&, const Phi4& phi, Phi4& H2phi, std::ostream* ost);
void tc2x(const Parallelization&, const Method&, const CrystalStructure&, const Potentials&, const Spin&,
          const Kpoints&, PlaneWaveBasis&, const BlochStates&, const Phi4& phi, Phi4& H2phi, std::ostream* ost);
void tc3a1(const Parallelization&, const Method&, const CrystalStructure&, const Potentials&, const Spin&,
           const Kpoints&, PlaneWaveBasis&, const BlochStates&, const Phi4& phi, Phi4& H3phi, std::ostream* ost);
void tc3b1(const Parallelization&, const Method&, const CrystalStructure&, const Potentials&, const Spin&,
           const Kpoints&, PlaneWaveBasis&, const BlochStates&, const Phi4& phi, Phi4& H3phi, std::ostream* ost);

}  // namespace calc_hamiltonian
