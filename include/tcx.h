/*
 * tcx.h -- C ABI of the MI355X pair-FFT operator-apply engine.
 *
 * Drop-in boundary for the H*phi terms of TC++ (masaochi/TC).  Each term
 * entry point replaces one function of namespace calc_hamiltonian declared in
 * the reference's src/include/calc_hamiltonian.hpp; the state setters carry
 * the pieces of the reference's CrystalStructure / Kpoints / PlaneWaveBasis /
 * BlochStates / Potentials objects that those functions read.
 *
 * Conventions (the reference's):
 *   - complex numbers are complex128, interleaved (re, im) doubles;
 *   - orbital arrays are "ragged contiguous": k-major, then band, then plane
 *     wave, i.e. phi[k][band][0..npw_k) with the k blocks concatenated
 *     (reference layout phi[spin][k][band][spinor] of Eigen::VectorXcd,
 *     spin and spinor of extent 1 here);
 *   - grid index ipw = i1 + n1*(i2 + n2*i3), i1 fastest
 *     (src/plane_wave_basis.cpp:137-140);
 *   - reciprocal vectors are the rows b_i, Cartesian G = recip^T * m
 *     (src/calc_hamiltonian_hf2x.cpp:35-36);
 *   - term entry points ACCUMULATE into their output (+=), as the reference
 *     terms do (e.g. src/calc_hamiltonian_hf2x.cpp:146).
 *
 * Errors: every function returns a tcx_status; nothing aborts or throws
 * across this boundary (the reference calls MPI_Abort instead,
 * src/error_messages.cpp:60-64).  tcx_last_error() gives the message.
 * Threading: one context is used by one host thread at a time; work runs on
 * the context's own HIP stream and entry points return after it completes.
 */
#ifndef TCX_H
#define TCX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tcx_context tcx_context;

typedef enum {
    TCX_OK = 0,
    TCX_ERR_ARG = 1,          /* invalid argument or size mismatch */
    TCX_ERR_STATE = 2,        /* a required setter has not been called */
    TCX_ERR_HIP = 3,          /* HIP runtime error */
    TCX_ERR_UNSUPPORTED = 4,  /* valid input outside what this build implements */
    TCX_ERR_NOMEM = 5,        /* device allocation failed */
    TCX_ERR_INTERNAL = 6
} tcx_status;

typedef enum { TCX_METHOD_HF = 0, TCX_METHOD_TC = 1, TCX_METHOD_BITC = 2 } tcx_method;

/* Arithmetic of the pair pipeline.  FP64: every FFT and product in double.
 * FP32: FFTs and pointwise products in float, band-pair sums accumulated in
 * double, orbital transforms, final transforms, divergence corrections and
 * all outputs in double.  Both: the band-pair sums of a launch are added into
 * the accumulators with compensation (TwoSum) across launches. */
typedef enum { TCX_PREC_FP64 = 0, TCX_PREC_FP32 = 1 } tcx_precision;

/* flags of the term entry points */
#define TCX_FLAG_DEVICE_PTRS 1u   /* phi / hphi are device pointers on the context's device */

/* ---------------------------------------------------------------- context */
int tcx_create(tcx_context **ctx, int device);
int tcx_destroy(tcx_context *ctx);
const char *tcx_last_error(const tcx_context *ctx);
int tcx_synchronize(tcx_context *ctx);

/* ---------------------------------------------------------------- state */
/* Cell, FFT grid, spin and method.  Replaces what the terms read from
 * CrystalStructure (reciprocal_vectors, unit_cell_volume:
 * src/include/crystal_structure.hpp), PlaneWaveBasis::setup_FFT
 * (src/plane_wave_basis.cpp:175-198), Spin::num_independent_spins and
 * Method::calc_method.
 *
 * grid: n1 = n2 and n1, n3 in {6, 8, 10, 12, 16, 18, 20, 24, 30, 32, 36, 40, 48}
 * (the compiled FFT lengths), else TCX_ERR_UNSUPPORTED.
 *
 * nspin = 1 or 2 (num_independent_spins).  With 2, every per-k array of the
 * states and of the term calls (nocc, filling, phi_occ, chi_occ, the applied
 * band counts nb and the coefficient arrays) runs over (spin, k), spin-major:
 * 2 nk entries, spin 0's k-blocks first.  The k-points, the basis and
 * sum_of_vaux are given once for the nk k-points.  kinetic, hf2h (the sum of
 * the two spin densities), hf2x and tc2x (states of one spin paired, u(s, s),
 * no spin-sum halving of the prefactors), tc2h, tc3a1 and tc3b1 (u(s, s')
 * with each spin's density and <q|nabla|q> sum) and the HF all / Davidson run
 * with two spins; the three-body exchange terms, the k-mesh symmetry,
 * zero-weight k-points, the gradients and tcx_hf2x_noalloc return
 * TCX_ERR_UNSUPPORTED.
 * Changing nspin clears the k-points: call tcx_set_kpoints again. */
int tcx_set_cell(tcx_context *ctx, const double recip_rows[9], double volume,
                 const int grid[3], int nspin, int method);

/* SCF k-points in crystal coordinates, identity symmetry (each k irreducible
 * with one isym, no time reversal).  num_kpoints is Kpoints::num_kpoints(),
 * the Nk of the 1/(Nk V) prefactors.  Replaces Kpoints::kvectors_scf /
 * index_all_kscf (src/kpoints.cpp:34-110). */
int tcx_set_kpoints(tcx_context *ctx, int nk, const double *kfrac, int num_kpoints);

/* Plane-wave basis: npw[k] and the concatenated FFT-grid indices of every
 * G at every k.  Replaces PlaneWaveBasis::num_G_at_k_scf / Gindex_at_k_scf
 * (src/plane_wave_basis.cpp:81-164). */
int tcx_set_basis(tcx_context *ctx, const int *npw, const int *gindex);

/* Occupied orbitals (BlochStates::phik_scf, filling, num_occupied_bands;
 * src/include/bloch_states.hpp:31-40).  nocc[k] orbitals at k with their
 * fillings (spin factor included, as in the reference).  Density-matrix
 * mixing's "old" bands (phik_scf_old, filling_old) are simply extra entries.
 * chi_occ: BITC left orbitals (phik_left_scf), NULL otherwise.  For TC/BITC
 * the real-space density (BlochStates::set_density,
 * src/bloch_states_scfloop_filling_density.cpp:256-358) is formed here. */
int tcx_set_states(tcx_context *ctx, const int *nocc, const double *filling,
                   const double *phi_occ, const double *chi_occ);

/* RPA Jastrow parameters in atomic units, [s1*2+s2] (Jastrow::A_long_,
 * F_long_ after unnormalize_A_long and impose_cusp;
 * src/jastrow_initialize.cpp:28-110).  All |A| <= 1e-8 means "no Jastrow"
 * (Jastrow::is_A_zero_). */
int tcx_set_jastrow_rpa(tcx_context *ctx, const double A[4], const double F[4]);

/* Polynomial Jastrow u_poly(r) = -sum_{i<deg} c_i r^i (1 - r/L)^C for r < L, per
 * spin pair [s1*2+s2] (Jastrow::set_polynomial_parameters,
 * src/jastrow_initialize_polynomial.cpp:6-101, with set_derived_polynomial_parameters
 * :103-118 and, when the RPA part is non-zero, set_derived_RPA_polynomial_parameters
 * :121-155).  deg in 0..16 (0 removes the polynomial part), C in {2, 3}, L[4] >= 1e-8,
 * c_input[p*deg + i] = c_poly_input_ (entry i = 0 of each pair is replaced by the cusp
 * condition: c_0 = (L/C)(c_1 - cusp), cusp = 1/4 parallel, 1/2 antiparallel, or 0 when
 * cusp == 0).  u(G) becomes uk_RPA + uk_POLY and tc_2body becomes tc_2body_POLY (no RPA
 * part) or tc_2body_RPAPOLY (deg > 1) (src/jastrow_functions.cpp:6-51, :99-413).
 * Needs tcx_set_cell: its FFT grid chooses num_rmesh_J3_integ (:157-204).  Without a
 * tcx_set_jastrow_rpa call the RPA part is zero; a later one (or tcx_set_cell) rebuilds
 * the derived parameters.  Errors where the reference stops: TCX_ERR_ARG. */
int tcx_set_jastrow_polynomial(tcx_context *ctx, int deg, int C, const double L[4], const double *c_input,
                               int cusp);

/* num_rmesh_J3_integ of the RPA+polynomial Jastrow (Jastrow::print_RPA_polynomial_parameters,
 * src/jastrow_initialize_polynomial.cpp:426-431); 0 without that form. */
int tcx_get_jastrow_num_rmesh(tcx_context *ctx, int *num_rmesh);

/* The Jastrow kernels at n points (device pointers of the context's GPU): uk[i] = u(G) and
 * tc2body[i] = tc_2body(G) of spin pair (s1, s2) at |G|^2 = G2[i] (Jastrow::uk,
 * Jastrow::tc_2body; src/jastrow_functions.cpp:6-51).  Asynchronous on the context's stream. */
int tcx_jastrow_kernels(tcx_context *ctx, int s1, int s2, long long n, const double *G2, double *uk,
                        double *tc2body);

/* Divergence (G=0) correction: Potentials::includes_div_correction,
 * alpha_Vaux and sum_of_Vaux_scf[nk] (src/potentials_coulomb.cpp:9-161). */
int tcx_set_div_correction(tcx_context *ctx, int enabled, double alpha, const double *sum_of_vaux);

/* Symmetry of the SCF k-mesh: k is built from the orbitals of the
 * irreducible k_src[k] (k_src[q] == q for an irreducible q, which needs the
 * identity below).  rot[9k..9k+8]: the TC++ rotation U of k (integer, crystal
 * coordinates, row-major; the inverse of QE's rotation), trans[3k..3k+2]: its
 * fractional translation, trev[k]: time reversal used.  k = +-U^T k_src in
 * crystal coordinates.  Replaces Kpoints::index_of_rotation_at_k /
 * is_time_reversal_used_at_k (src/kpoints.cpp:34-110) and Symmetry::rotation /
 * translation (src/symmetry.cpp:17-33).  The occupied orbitals of
 * tcx_set_states at k must be those of k_src[k] mapped as in
 * src/plane_wave_basis.cpp:115-140 (c(G) placed at G' = +-U^T G, conjugated
 * under time reversal, times exp(2 pi i G'.tau)), with the same fillings.
 * With a symmetry set, tcx_tc3a2a4b2b5 forms its q2-resolved sums at the
 * irreducible q2 only and rotates them to the other q2, as the reference does
 * (src/calc_hamiltonian_tc3a2a4b2b5.cpp:143-765); the other terms read the
 * full mesh as given.  k_src == NULL restores the identity symmetry.  Cleared
 * by tcx_set_cell / tcx_set_kpoints. */
int tcx_set_kpoint_symmetry(tcx_context *ctx, const int *k_src, const int *rot, const double *trans,
                            const int *trev);

/* Real-space density the terms read (BlochStates::density_;
 * src/bloch_states_scfloop_filling_density.cpp:256-358): complex128 [2][N]
 * (spin up, spin down), normalised so that sum_r rho_s / (N * num_kpoints) is
 * the electron count of spin s.  It is formed from the occupied orbitals and
 * fillings of tcx_set_states on first use.  With one independent spin both
 * spins take the first grid, as the reference sets density_[1] = density_[0]
 * (:355-358); tcx_set_density then ignores the second.  tcx_set_density replaces it, e.g.
 * by the mixed density (1-beta) rho_old + beta rho_new of the reference's
 * density mixing (:345-349), until the next tcx_set_states or
 * tcx_set_kpoint_symmetry.  The gradient entry points always use the density
 * of the orbitals. */
int tcx_get_density(tcx_context *ctx, double *density);
int tcx_set_density(tcx_context *ctx, const double *density);

/* Host utility computing alpha and sum_of_Vaux for a k-set, restating
 * Potentials::set_sum_of_Vaux_each (src/potentials_coulomb.cpp:60-161). */
int tcx_sum_of_vaux(const double recip_rows[9], double volume, int nk, const double *kfrac,
                    double *alpha, double *sum_of_vaux);

/* The same with zero-weight k-points in the list: replaces Potentials::set_sum_of_Vaux_each
 * and set_sum_of_kVaux_each (src/potentials_coulomb.cpp:60-254), whose q-sums run over the
 * weighted k only (kweight[k] >= 1e-8).  Out: alpha, sum_of_vaux[nk] and, when not NULL,
 * sum_of_kvaux[nk][3] (zero at k on the weighted mesh).  No GPU. */
int tcx_sum_of_vaux_weighted(const double recip_rows[9], double volume, int nk, const double *kfrac,
                             const double *kweight, double *alpha, double *sum_of_vaux, double *sum_of_kvaux);

/* Zero-weight k-points: the BAND k-points of a band calculation (kvectors_band) and the
 * zero-weight SCF k-points (Kpoints::kweight_scf < 1e-8) of the k list.  Replaces the
 * BAND / zero-weight branches of the terms: calc_hamiltonian_hf2x.cpp:150-201,
 * tc2x.cpp:117-222 and :408-513, tc3a3b3b4b6.cpp:393-627 and :772-858,
 * tc3a2a4b2b5.cpp:1036-1299 (BlochStates::phik_band, Kpoints::return_band_filling).
 * After tcx_set_states (which holds nocc = 0 at these k) and the divergence correction
 * (sum_of_vaux from tcx_sum_of_vaux_weighted):
 *   kweight      [nk]: < 1e-8 marks a zero-weight k (no source of the q-sums);
 *   sum_of_kvaux [nk][3]: the vector shell sums of tcx_sum_of_vaux_weighted;
 *   nstate       [nk]: states the divergence terms at a zero-weight k project on (0 at
 *                weighted k, at most 512), with their band fillings `filling` and
 *                coefficients phi (and BITC chi), complex128 ragged [k][b][npw_k].
 * Bands applied at a zero-weight k see the occupied orbitals as sources and get the
 * reference's divergence terms at that k, including the kVaux branches off the mesh.
 * Cleared by tcx_set_states, tcx_set_basis and tcx_set_kpoints; not combined with
 * tcx_set_kpoint_symmetry. */
int tcx_set_zero_weight_kpoints(tcx_context *ctx, const double *kweight, const double *sum_of_kvaux,
                                const int *nstate, const double *filling, const double *phi, const double *chi);

/* ---------------------------------------------------------------- terms */
/* nb[k] bands at every k in phi (ragged, see above); hphi has the same
 * layout and is accumulated into. */

/* Fock exchange Vx*phi: replaces calc_hamiltonian::hf2x
 * (src/include/calc_hamiltonian.hpp:54-63; src/calc_hamiltonian_hf2x.cpp:8-205). */
int tcx_hf2x(tcx_context *ctx, const int *nb, const double *phi, double *h2phi,
             int precision, unsigned flags);

/* tcx_hf2x without device allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory while they run (serving loops, code that must not
 * call hipMalloc).  phi and h2phi are device pointers of the context's GPU (h2phi is
 * accumulated into, as for tcx_hf2x); the occupied orbitals must already be in real
 * space in `precision`, which the first tcx_hf2x call of that precision after
 * tcx_set_states does (else TCX_ERR_STATE).  The working memory -- launch tables, the
 * applied bands' grids, the accumulators, the pair records -- is one fixed device array
 * of 256 MiB per GPU, held until the process ends; the bands are processed in chunks
 * that fit it, and a grid whose single band pair does not fit is refused with
 * TCX_ERR_ARG.  Every context on a GPU shares that GPU's array, so calls that overlap
 * on one GPU race.  Synchronous; not capturable into a graph. */
int tcx_hf2x_noalloc(tcx_context *ctx, const int *nb, const double *phi, double *h2phi,
                     int precision);

/* TC/BITC two-body exchange [2a_x, 2b_x1, 2b_x2]: replaces
 * calc_hamiltonian::tc2x (src/include/calc_hamiltonian.hpp:77-86;
 * src/calc_hamiltonian_tc2x.cpp:11-516). */
int tcx_tc2x(tcx_context *ctx, const int *nb, const double *phi, double *h2phi,
             int precision, unsigned flags);

/* Three-body terms that reduce to a local potential: replace
 * calc_hamiltonian::tc3a1 and tc3b1 (src/include/calc_hamiltonian.hpp:89-108;
 * src/calc_hamiltonian_tc3a1.cpp:8-117, src/calc_hamiltonian_tc3b1.cpp:8-135). */
int tcx_tc3a1(tcx_context *ctx, const int *nb, const double *phi, double *h3phi,
              int precision, unsigned flags);
int tcx_tc3b1(tcx_context *ctx, const int *nb, const double *phi, double *h3phi,
              int precision, unsigned flags);

/* Three-body exchange terms [3a3(+3a6), 3b3, 3b4, 3b6] and [3a2, 3a4(+3a5),
 * 3b2, 3b5]: replace calc_hamiltonian::tc3a3b3b4b6 and tc3a2a4b2b5
 * (src/include/calc_hamiltonian.hpp:111-130; src/calc_hamiltonian_tc3a3b3b4b6.cpp:11-863,
 * src/calc_hamiltonian_tc3a2a4b2b5.cpp:11-1300).  Identity symmetry only (the
 * reference's Symmetry argument is the identity); k-points on the SCF mesh. */
int tcx_tc3a3b3b4b6(tcx_context *ctx, const int *nb, const double *phi, double *h3phi,
                    int precision, unsigned flags);
int tcx_tc3a2a4b2b5(tcx_context *ctx, const int *nb, const double *phi, double *h3phi,
                    int precision, unsigned flags);

/* One-body kinetic term 0.5|k+G|^2 phi: replaces calc_hamiltonian::kinetic
 * (src/include/calc_hamiltonian.hpp:23-31; src/calc_hamiltonian_kinetic.cpp:6-52).  FP64. */
int tcx_kinetic(tcx_context *ctx, const int *nb, const double *phi, double *h1phi, unsigned flags);

/* Hartree terms: replace calc_hamiltonian::hf2h (HF; src/calc_hamiltonian_hf2h.cpp:8-85)
 * and tc2h (TC/BITC [2a_h, 2b_h1, 2b_h2]; src/calc_hamiltonian_tc2h.cpp:10-236)
 * (src/include/calc_hamiltonian.hpp:43-52, 66-75).  FP64. */
int tcx_hf2h(tcx_context *ctx, const int *nb, const double *phi, double *h2phi, unsigned flags);
int tcx_tc2h(tcx_context *ctx, const int *nb, const double *phi, double *h2phi, unsigned flags);

/* The whole H*phi: replaces calc_hamiltonian::all (src/include/calc_hamiltonian.hpp:10-21;
 * src/calc_hamiltonian_all.cpp:6-198).  Zeroes h1phi, h2phi and (TC/BITC) h3phi, then
 * kinetic -> h1phi; HF: hf2h + hf2x -> h2phi; TC/BITC: tc2h + tc2x -> h2phi and
 * tc3a1 + tc3b1 + tc3a2a4b2b5 + tc3a3b3b4b6 -> h3phi.  h3phi may be NULL for HF.
 * No pseudopotential term (homogeneous-electron-gas cells, all.cpp:58-66). */
int tcx_apply_all(tcx_context *ctx, const int *nb, const double *phi, double *h1phi,
                  double *h2phi, double *h3phi, int precision, unsigned flags);

/* ---------------------------------------------------------------- gradients */
/* No single counterpart in the reference (which has no derivatives); the energies
 * are those of TotalEnergy::calc_total_energy (src/total_energy.cpp:118-163) for the
 * occupied set, and H_2^dagger is the reference's own Hermitian conjugate
 * (Jastrow::take_Hermitian_conj + BlochStates::switch_left_right_orbitals,
 * src/diagonalization_davidson.cpp:234-251).  Host pointers, complex128 in the
 * occupied layout of tcx_set_states:
 *   tcx_grad_exchange: energy[2] = E_x = (1/(2Nk)) sum f <phi|V_x phi>,
 *                      grad = dE_x/d conj(phi) = (f/Nk) V_x phi;
 *   tcx_grad_tc2:      energy[2] = E_2 = (1/(2Nk)) sum f <chi|(tc2h + tc2x) phi>,
 *                      grad_left = dE_2/d conj(chi), grad_right = dE_2/d phi,
 *                      dE_dA[2] = dE_2/dA at fixed F (all four spin-pair A scaled together). */
int tcx_grad_exchange(tcx_context *ctx, double *grad, double *energy, int precision);
int tcx_grad_tc2(tcx_context *ctx, double *grad_left, double *grad_right, double *energy,
                 double *dE_dA, int precision);

/* Backward pass with respect to the two-body kernels, radial bins of |k-q+G| of width
 * `width` (1 <= nbins <= 1024, last bin open).  Pair densities are recomputed, not stored.
 * term 0: E_x and its Coulomb kernel v = 1/|k-q+G|^2 (src/calc_hamiltonian_hf2x.cpp:91-104):
 *         dE_dK complex128[nbins] = dE_x/dv per bin; dE_du unused (may be NULL);
 * term 1: the exchange part of E_2 (tc2x without the divergence term,
 *         src/calc_hamiltonian_tc2x.cpp:253-401) and its kernels:
 *         dE_dK complex128[nbins] = dE/d tc_2body, dE_du complex128[nbins] = dE/du.
 * sums[4]: sum over all points of K dE/dK (+ u dE/du): the kernel-linear energy (re, im),
 *          and of dK/dA dE/dK + du/dA dE/du: dE/dA at fixed F (re, im; term 1). */
int tcx_kernel_gradient(tcx_context *ctx, int term, int nbins, double width, double *dE_dK,
                        double *dE_du, double *sums, int precision);

/* ---------------------------------------------------------------- HIP graphs */
/* No counterpart in the reference (which launches no device work).  For callers that
 * replay the same applies many times (a Davidson block, batched per-k workloads):
 * tcx_graph_begin starts capturing this context's stream; the term calls that follow
 * must pass TCX_FLAG_DEVICE_PTRS and repeat a call already made once (so that its
 * launch plan and workspaces exist; otherwise TCX_ERR_STATE and the capture is
 * dropped); tcx_graph_end instantiates the graph and returns its id; tcx_graph_launch
 * replays it asynchronously (tcx_synchronize waits).  Terms accumulate, so clear the
 * outputs (tcx_memset_zero, a fill kernel) between replays when needed.  The
 * three-body exchange terms upload small tables and cannot be captured.  A graph
 * addresses the caller's buffers and the context's workspaces as they were at
 * capture: workspaces that later calls outgrow stay allocated until the last graph is
 * destroyed; the caller keeps its buffers alive; capture again after tcx_set_states. */
int tcx_graph_begin(tcx_context *ctx);
int tcx_graph_end(tcx_context *ctx, int *graph_id);
int tcx_graph_launch(tcx_context *ctx, int graph_id);
int tcx_graph_destroy(tcx_context *ctx, int graph_id);

/* ---------------------------------------------------------------- consumer */
/* Block Davidson at fixed occupied orbitals (one Fock operator): replaces
 * Diagonalization::block_davidson (src/include/diagonalization.hpp; src/diagonalization_davidson.cpp:7-689)
 * with make_a_new_trial_vector, Gram_Schmidt, Gram_Schmidt_biortho, sort_eigen and
 * sort_eigen_overlap (src/diagonalization_utils.cpp:7-211).  H*phi is tcx_apply_all on
 * the GPU in `precision`; the subspace algebra (Gram-Schmidt, preconditioner, V^H H V,
 * the per-k eigensolve, rotations) runs in FP64 on the GPU too.  Host pointers:
 *   phi      in/out complex128 ragged [k][nb[k]][npw_k]: trial bands in (orthonormal per k,
 *            as the reference's bands are; BITC bands are orthonormalised on entry unless
 *            biortho), eigenvectors out;
 *   phi_left BITC only, else may be NULL: left eigenvectors out, V (X^-1)^H; with biortho
 *            also in: the left trial bands (bi-orthonormal to phi, as phik_left is);
 *   eig      in/out complex128 [sum nb]: residual shifts in (the previous eigenvalues),
 *            eigenvalues out, ascending real part (biortho: ordered by overlap with the
 *            previous bands);
 *   energies out complex128 [sum nb][3]: e_n = diag(X_L^H H_n X), n = 1, 2, 3.
 * max_blocks >= 2 and num_refresh >= 1 are the reference's max_num_blocks_david and
 * num_refresh_david (defaults 2 and 1); biortho = 1 is biortho_david = true (BITC only):
 * left trial vectors grown from H^dagger (u -> -u, right and left occupied orbitals
 * swapped, the density kept; :234-251), bi-orthogonal Gram-Schmidt, subspace
 * <V_left|H|V>. */
int tcx_davidson(tcx_context *ctx, const int *nb, double *phi, double *phi_left, double *eig,
                 double *energies, int max_blocks, int num_refresh, int biortho, int precision);

/* Total energies from per-band e_n: replaces TotalEnergy::calc_total_energy
 * (src/total_energy.cpp:98-163; identity symmetry, no Ewald term):
 * out complex128 [4] = E_1 = sum f e_1 / Nk, E_2 = sum f e_2 / (2 Nk),
 * E_3 = sum f e_3 / (3 Nk) (0 unless uses_3body), and E_1 + E_2 + E_3.  No GPU. */
int tcx_total_energy(int nk, const int *nb, const double *filling, const double *energies,
                     int num_kpoints, int uses_3body, double *out);

/* Band fillings and Fermi energy: replaces BlochStates::set_filling and
 * Kpoints::return_band_filling (src/bloch_states_scfloop_filling_density.cpp:9-194,
 * src/kpoints.cpp:349-387) for one independent spin.  eig: real parts of the
 * eigenvalues [sum nb]; kweight[nk] as Kpoints::kweight_scf (2/Nk per k with identity
 * symmetry, 0 for zero-weight k); smearing 0 = fixed (lowest round(N_e/2) bands get 2),
 * 1 = gaussian of the given width (f = 1 - erf((e - mu)/width), mu by 60 bisections
 * of [-1e5, 1e5] Ha).  Out: filling [sum nb], *fermi_energy, num_occupied[nk] (f > 1e-8).
 * Returns TCX_ERR_ARG when the electron count cannot be met.  No GPU. */
int tcx_set_filling(int nk, const int *nb, const double *eig, const double *kweight, double num_electrons,
                    int smearing, double width, double *filling, double *fermi_energy, int *num_occupied);

/* Dense complex eigen-decomposition A X = X diag(w), n <= 4096: replaces
 * Eigen::ComplexEigenSolver in the Davidson subspace step
 * (src/diagonalization_davidson.cpp:480-482).  A, X complex128 n x n row-major;
 * unit-norm eigenvector columns; eigenvalues unsorted.  No GPU. */
int tcx_eig_dense(int n, const double *A, double *w, double *X);

/* ---------------------------------------------------------------- device buffers */
/* No counterpart in the reference (which keeps everything in host memory).
 * For callers that hold no HIP runtime of their own and want inputs and
 * outputs resident in HBM for TCX_FLAG_DEVICE_PTRS: allocate on the
 * context's device, copy (kind 0 host->device, 1 device->host,
 * 2 device->device, ordered on the context's stream, returning after the
 * copy completes), and free. */
int tcx_device_alloc(tcx_context *ctx, unsigned long long bytes, void **ptr);
int tcx_device_free(tcx_context *ctx, void *ptr);
int tcx_memcpy(tcx_context *ctx, void *dst, const void *src, unsigned long long bytes, int kind);
int tcx_memset_zero(tcx_context *ctx, void *ptr, unsigned long long bytes);

/* ---------------------------------------------------------------- tuning / timing */
/* Pass order of the exchange pair pipeline (term 0: hf2x): 0 z-first (transform along z,
 * then the xy planes, then z back), 1 plane-first (xy planes, then z pencils, then xy back;
 * FP32-split only: the pair records then hold only the (kx, ky) columns inside the pair
 * density's G-space support, about two thirds of them at 40^3), -1 automatic (z-first;
 * the default: plane-first moves fewer bytes but ran slower at 40^3).  No counterpart in
 * the reference, whose hf2x runs full 3D FFTs (src/calc_hamiltonian_hf2x.cpp:123-135);
 * the two orders agree to rounding. */
int tcx_set_pipeline_order(tcx_context *ctx, int term, int order);

/* Mean number of (kx, ky) columns per plane of the pair records of the last apply of
 * term 0 (hf2x): n1*n2 for the z-first order, the support columns for plane-first. */
int tcx_get_record_columns(tcx_context *ctx, int term, double *columns);

/* Launch shape of a pair pipeline (term 0: hf2x, 1: tc2x, 2: the three-body terms'
 * pair sets): each launch pairs `targets` consecutive applied bands (k,j) (or
 * occupied orbitals, for the three-body occupied-set phases) with `sources`
 * consecutive occupied orbitals (q,i); targets*sources pair records are in flight
 * (the pair-buffer size; for term 2 capped by a 3 GiB record budget).  0 selects the
 * automatic shape.  Set by the autotuner (tc_amd/autotune.py).  No counterpart in
 * the reference (its OpenMP loop has no launch shape). */
int tcx_set_launch_shape(tcx_context *ctx, int term, long targets, long sources);

/* Per-kernel-class timing with HIP events on the context's stream.
 * Classes: 0 pair pass A (z-pencils), 1 pair pass B (xy-planes),
 * 2 pair pass C (z-pencils + accumulate), 3 orbital/final transforms.
 * out_ms[c] = summed duration (ms), out_launches[c] = number of launches,
 * out_pairs[c] = pairs processed by that class. */
int tcx_enable_timing(tcx_context *ctx, int on);
int tcx_get_timing(tcx_context *ctx, double out_ms[4], long out_launches[4], long out_pairs[4]);

#ifdef __cplusplus
}
#endif
#endif /* TCX_H */
