/*
 * oracle.h -- FP64 CPU restatement of TC++'s pair-FFT operator apply.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (tc_amd/, include/,
 * libtcx.so) may include, link or call this code.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py use it, and
 * only as the checker / the timed CPU baseline.
 *
 * The reference (masaochi/TC) needs FFTW3, Eigen3, Boost and MPI, none of
 * which exist in this image, so it cannot be compiled here; this file is a
 * restatement written from reading the reference sources.  Each function
 * cites the reference file:line it follows.  Conventions (reference
 * src/plane_wave_basis.cpp:175-289, src/include/header.hpp:43-44):
 *   - grid index ipw = i1 + n1*(i2 + n2*i3), i1 fastest;
 *   - FFT_forward (R->G) = exp(-iGr) with 1/N; FFT_backward (G->R) = exp(+iGr), unscaled;
 *   - Cartesian G = recip^T * m with recip rows b_i; Miller m_d = i_d if i_d <= n_d/2 else i_d-n_d.
 *
 * Scope of this restatement: identity-only symmetry (every k irreducible,
 * one isym, no time reversal, phase factor 1), collinear spins with
 * num_independent_spins = 1 (the spin-unpolarised configurations); SCF k-points
 * on the mesh and, through kweight, zero-weight SCF k-points and BAND k-points
 * (one k list whose zero-weight entries are the band k-points).
 */
#ifndef TCX_ORACLE_H
#define TCX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* FFT grid and cell */
    int n[3];
    double recip[9];        /* reciprocal vectors as rows b_i (Cartesian, bohr^-1) */
    double volume;          /* unit-cell volume (bohr^3) */
    int nspin;              /* num_independent_spins (only 1 supported) */
    int method;             /* 0 = HF, 1 = TC, 2 = BITC */
    /* k-points (identity symmetry) */
    int nk;
    const double *kfrac;    /* nk*3 crystal coordinates */
    int num_kpoints;        /* Nk in the 1/(Nk V) prefactors */
    /* plane-wave basis */
    const int *npw;         /* nk */
    const int *gindex;      /* concatenated over k, FFT-grid linear indices */
    /* occupied (right) and left orbitals at every k; "old" density-matrix
     * bands are simply extra entries of this list */
    const int *nocc;        /* nk */
    const double *filling;  /* concatenated over k */
    const double *phi_occ;  /* complex128 interleaved, [k][i][npw_k] */
    const double *chi_occ;  /* BITC left orbitals (same layout) or NULL (= phi_occ) */
    /* RPA Jastrow, [s1*2+s2], atomic units (already "unnormalised") */
    double A[4];
    double F[4];
    int jastrow_A_zero;     /* Jastrow::is_A_zero_ */
    /* divergence (G=0) correction */
    int div_corr;
    double alpha;           /* Potentials::alpha_Vaux_ */
    const double *sum_vaux; /* nk, Potentials::sum_of_Vaux_scf_ */
    /* real-space density [2][N] complex128 (BlochStates::density_); may be NULL
     * for terms that do not use it */
    const double *density;
    /* k-points off the weighted mesh: BAND k-points and zero-weight SCF k-points
     * (Kpoints::kweight_scf < 1e-8).  kweight NULL: every k is a weighted mesh point.
     * A zero-weight k is no source of the q-sums (it holds no occupied orbitals,
     * nocc = 0), is matched against the weighted k only (if_match), and its divergence
     * terms project on the states zw_* (phik_band / phik_scf there) with their
     * Kpoints::return_band_filling values; at a weighted k they project on the
     * occupied orbitals at k. */
    const double *kweight;  /* nk or NULL */
    const double *sum_kvaux;/* nk*3, Potentials::sum_of_kVaux_(scf|band); NULL = 0 */
    const int *zw_nstate;   /* nk: states per zero-weight k (0 at weighted k); NULL = none */
    const double *zw_fill;  /* concatenated over k */
    const double *zw_phi;   /* complex128 [k][b][npw_k] */
    const double *zw_chi;   /* BITC left states (same layout) or NULL (= zw_phi) */
    /* polynomial part of the Jastrow (NULL or deg == 0: none); with A != 0 it is the
     * RPA+polynomial form */
    const struct orc_jastrow_poly *poly;
} orc_problem;

/* Polynomial Jastrow u_poly(r) = -sum_i c_i r^(i-1) (1 - r/L)^C theta(L - r) per spin
 * pair [s1*2+s2] with its derived parameters (reference Jastrow members,
 * src/include/jastrow.hpp:41-69). */
typedef struct orc_jastrow_poly {
    int deg;                 /* deg_poly_ (0..16) */
    int C;                   /* C_damp_poly_ (2 or 3) */
    int rpa;                 /* RPA terms present: RPA+polynomial derived parameters set */
    int nr;                  /* num_rmesh_J3_integ_ */
    double L[4], A[4], F[4];
    double c_input[4][16];   /* c_poly_input_ after the cusp condition */
    double c_int[4][19];     /* c_poly_internal_ */
    double fourpi_L[4][38];  /* FourPI_power_L_ */
    double exp_LF[4];        /* exp_LF_ */
    double M[4][47];         /* M_integ_ */
    double *rmesh;           /* [4][nr] rmesh_J3_integ_ */
    double *integrand;       /* [4][nr] integrand_stored_J3_integ_ */
} orc_jastrow_poly;

/* 3D FFT in place on a complex128 grid (interleaved doubles).
 * sign = -1: FFT_forward (with 1/N); sign = +1: FFT_backward (unscaled).
 * Reference: src/plane_wave_basis.cpp:274-289 (FFTW plan dims (n3,n2,n1), :189-196). */
void orc_fft3d(double *grid, const int n[3], int sign);

/* Naive O(N^2) 3D DFT with the same convention (for pinning orc_fft3d). */
void orc_dft3d_naive(const double *in, double *out, const int n[3], int sign);

/* Jastrow kernels u(G) and tc_2body(G), reference src/jastrow_functions.cpp:6-51
 * (RPA :53-97, polynomial and RPA+polynomial :99-413 through p->poly). */
double orc_uk(const orc_problem *p, double G2, int s1, int s2);
double orc_tc_2body(const orc_problem *p, double G2, int s1, int s2);

/* Polynomial Jastrow set-up (jastrow_initialize_polynomial.cpp:6-354): c_input[p*deg + i]
 * (entry 0 of each pair is replaced by the cusp condition); rpa: RPA terms present (A, F
 * used for the RPA+polynomial cross terms); recip, n: the FFT grid whose |G| choose
 * num_rmesh_J3_integ.  Returns 0, -1 (bad argument) or -2 (J3 mesh did not converge). */
int orc_jastrow_poly_init(orc_jastrow_poly *jp, int deg, int C, const double L[4], const double *c_input,
                          int cusp, int rpa, const double A[4], const double F[4], const double recip[9],
                          const int n[3]);
void orc_jastrow_poly_free(orc_jastrow_poly *jp);
void orc_jastrow_poly_set_rmesh(orc_jastrow_poly *jp, int nr);
void orc_J_integ(const orc_jastrow_poly *jp, double G2, int p, double *J, int nJ, int zeroth);
void orc_J2_integ(const orc_jastrow_poly *jp, double G1, int p, double *J2, int nJ2);
double orc_J3_integ(const orc_jastrow_poly *jp, double G1, int p);
double orc_uk_poly(const orc_jastrow_poly *jp, double G2, int p);
double orc_tc2body_poly(const orc_jastrow_poly *jp, double G2, int p);
double orc_tc2body_rpapoly(const orc_jastrow_poly *jp, double G2, int p, double rpa);

/* Divergence-correction shell sums, reference src/potentials_coulomb.cpp:60-161.
 * Writes alpha and sum_of_Vaux[nk] (single process: ibegin..iend covers everything). */
void orc_sum_of_vaux(const double recip[9], double volume, int nk, const double *kfrac,
                     double *alpha_out, double *sum_out);

/* The same with the mesh restricted to the weighted k (kweight >= 1e-8) and the vector
 * sums of Potentials::set_sum_of_kVaux_each (potentials_coulomb.cpp:163-254):
 * ksum_out[nk*3] = 4 pi sum_{q,G} exp(-alpha |k-q+G|^2) (k-q+G) / |k-q+G|^2, zero at k
 * that match a weighted k.  ksum_out may be NULL. */
void orc_sum_of_vaux_w(const double recip[9], double volume, int nk, const double *kfrac, const double *kweight,
                       double *alpha_out, double *sum_out, double *ksum_out);

/* density_s(r), reference src/bloch_states_scfloop_filling_density.cpp:256-358.
 * out: [2][N] complex128. */
void orc_density(const orc_problem *p, double *out);

/* H*phi terms.  phi and hphi are complex128 ragged [k][j][npw_k] with nb[k]
 * applied bands at k; results are ACCUMULATED into hphi (+=), like the
 * reference.  nthreads <= 0 uses the OpenMP default. */
void orc_hf2x(const orc_problem *p, const int *nb, const double *phi, double *hphi, int nthreads);
void orc_tc2x(const orc_problem *p, const int *nb, const double *phi, double *hphi, int nthreads);
void orc_tc3a1(const orc_problem *p, const int *nb, const double *phi, double *hphi);
void orc_tc3b1(const orc_problem *p, const int *nb, const double *phi, double *hphi);
/* Assembly terms: kinetic (src/calc_hamiltonian_kinetic.cpp:6-52), Hartree
 * (src/calc_hamiltonian_hf2h.cpp:8-85), TC Hartree (src/calc_hamiltonian_tc2h.cpp:10-236). */
void orc_kinetic(const orc_problem *p, const int *nb, const double *phi, double *hphi);
void orc_hf2h(const orc_problem *p, const int *nb, const double *phi, double *hphi);
void orc_tc2h(const orc_problem *p, const int *nb, const double *phi, double *hphi);
/* Three-body exchange terms, src/calc_hamiltonian_tc3a3b3b4b6.cpp:11-863 and
 * src/calc_hamiltonian_tc3a2a4b2b5.cpp:11-1300 (identity symmetry, k on the mesh). */
void orc_tc3a3b3b4b6(const orc_problem *p, const int *nb, const double *phi, double *hphi, int nthreads);
void orc_tc3a2a4b2b5(const orc_problem *p, const int *nb, const double *phi, double *hphi, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
