/*
 * tcx_host.h -- C ABI of the C++ host around the engine: the QE inputs of a run, the
 * symmetry-expanded k-mesh built from them, and the SCF loop.
 *
 * These entry points wrap tcx_host (tc_amd/host/scf.hpp), the one implementation of
 * TC++'s k-mesh expansion and SCF loop in this library, so that callers in other
 * languages drive the same code:
 *   - Symmetry::set (src/symmetry.cpp:17-33) and Kpoints::set_kvectors_by_symmetry
 *     (src/kpoints.cpp:34-110): tcxh_qe_rotation, tcxh_qe_replicas;
 *   - PlaneWaveBasis::set_Gindex_at_k / get_orbital_FFTgrid
 *     (src/plane_wave_basis.cpp:115-140, 221-271): tcxh_qe_replica_basis;
 *   - Diagonalization::scf (src/diagonalization.cpp:154-303): tcxh_scf_run.
 * The inputs are what TC++ reads from the QE save directory
 * (src/io_qe_files_read_xml.cpp:55-273, io_qe_files_read_wfc.f90:69-140) plus the SCF
 * settings, as one flat file: after the 8-byte magic "TCXQE001", named records of a
 * 32-byte name, an 8-byte type tag ("<f8", "<i4", "<c16"), an int64 element count and
 * the little-endian data.
 *
 * Errors: every function returns a tcx_status from tcx.h: TCX_ERR_ARG for bad input
 * (including a malformed file or a rotation that is not unimodular), and from
 * tcxh_scf_run the status of an engine call that failed.  tcxh_last_error() gives the
 * message of the calling thread's last failure.
 */
#ifndef TCX_HOST_H
#define TCX_HOST_H

#include "tcx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tcxh_qe tcxh_qe;                  /* QE inputs + SCF settings */
typedef struct tcxh_scf_result tcxh_scf_result;  /* what one SCF run printed */

/* Message of the calling thread's last failed tcxh_ call ("" if none). */
const char *tcxh_last_error(void);

/* Reads a flat input file.  TCX_ERR_ARG for a missing, truncated or malformed file. */
int tcxh_qe_load(const char *path, tcxh_qe **out);
void tcxh_qe_free(tcxh_qe *q);

/* Number of k-points of the expanded mesh. */
int tcxh_qe_num_replicas(const tcxh_qe *q, int *nrep);

/* The expanded mesh in the reference's index_all_kscf order: per replica the irreducible
 * k (ik), the symmetry operation (isym, TC++ index), the time-reversal flag and the
 * crystal coordinates (kcrys[3 r + d]); each array holds nrep entries. */
int tcxh_qe_replicas(const tcxh_qe *q, int *ik, int *isym, int *trev, double *kcrys);

/* TC++'s rotation U = (QE rotation isym)^-1, row-major integers (U[9]). */
int tcxh_qe_rotation(const tcxh_qe *q, int isym, int *U);

/* Plane waves of replica r (the irreducible k's count) and, per plane wave of the
 * irreducible basis, the grid index of G' = +-U^T G and the phase exp(2 pi i G'.tau)
 * (phase: npw complex128). */
int tcxh_qe_replica_npw(const tcxh_qe *q, int r, int *npw);
int tcxh_qe_replica_basis(const tcxh_qe *q, int r, int *gidx, double *phase);

/* Diagonalization::scf from the QE orbitals on the context's GPU, in the given
 * precision (TCX_PREC_*); max_iterations > 0 replaces the input's max_num_iterations.
 * The context's state (cell, k-points, basis, states, density) is left as the run's
 * last set_density made it. */
int tcxh_scf_run(tcx_context *ctx, const tcxh_qe *q, int precision, int max_iterations, tcxh_scf_result **out);

/* Counts: set_filling calls, calc_total_energy calls, density differences (one per
 * set_density after the first), Davidson calls, whether the run converged, and the
 * irreducible k and bands of the eigenvalue and filling tables. */
int tcxh_scf_result_sizes(const tcxh_scf_result *r, int *nfill, int *nenergy, int *nddiff, int *iterations,
                          int *converged, int *nirr, int *nb);

/* set_filling call i: Fermi energy (Hartree), the eigenvalues the fillings came from
 * (nirr*nb complex128) and the fillings (nirr*nb). */
int tcxh_scf_result_fill(const tcxh_scf_result *r, int i, double *fermi, double *eig, double *filling);

/* calc_total_energy call i: E_1, E_2, E_3, E_total (4 complex128). */
int tcxh_scf_result_energy(const tcxh_scf_result *r, int i, double *e);

/* The density differences (nddiff doubles). */
int tcxh_scf_result_ddiff(const tcxh_scf_result *r, double *d);

void tcxh_scf_result_free(tcxh_scf_result *r);

#ifdef __cplusplus
}
#endif

#endif /* TCX_HOST_H */
