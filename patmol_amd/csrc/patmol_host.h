// patmol_host.h -- libpatmol_host.so: the caller side of the hot path
// (basis model with the reference's basis.h accessors, CPU one-electron
// integrals, RHF/UHF SCF drivers).
#pragma once
#include "patmol_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* basis.h accessors (basis.c:21-195) */
int number_of_orbitals(void);
int number_of_shells(void);
int number_of_centers(void);
int total_number_of_orbitals(void);
int total_number_of_shells(void);
int total_number_of_centers(void);
int number_of_electrons(void);
int spin_multiplicity(void);
int max_scf_iterations(void);
int skip_center(void);
int max_lvalue(void);
int max_lstates(void);
int do_dpm_calculation(void);
int do_finite_field_calculation(void);
int scf_acceleration(void);
int scf_guess(void);
double scf_convergence(void);
const Shell *shells(void);
const Center *centers(void);
int ***lvector_components(void);
double **normalizationFactors(void);
void set_center_position(int icen, const double *r);
void set_center_charge(int icen, double q);
double nuclear_repulsion(void);
void init_Basis(void);
void dpm_augment(void);

/* array loader: shells normalised as shell.c:47-77; returns 0 on success */
int patmol_basis_load(int ncen, const double *q, const double *xyz, int nshell, const int *npr,
                      const int *lsh, const int *cen, const double *al, const double *co, int nelec,
                      int mult, int maxits, double eps, int scf_a);
int patmol_basis_add_dpm_center(void);

/* CPU one-electron integrals, packed (oneints.c:139-234); the device
 * one_electron_ints is declared in patmol_hip.h */
void patmol_host_one_electron_ints(double *S, double *T, double *H);

/* SCF drivers (rhfscf.c:21-174,432-471; uhfscf.c:22-168,485-527) */
typedef struct {
    double energy, ediff, pdiff, enuc, eelec;
    int iterations, converged;
    double gbuild_seconds;   /* wall time inside the G builds */
    double diag_seconds;     /* wall time inside the diagonalisations */
} PatmolScfResult;

int patmol_rhf_scf(double *P_out, double *evals_out, PatmolScfResult *res);
int patmol_uhf_scf(double *PA_out, double *PB_out, PatmolScfResult *res);

/* multipole analysis (scfaux.c:140-340) from the device moment integrals:
 * electronic, nuclear and total (nuclear - electronic) moments, 9 values
 * each (x, y, z, xx, xy, xz, yy, yz, zz) */
int patmol_rhf_moments(const double *P, double *elec, double *nuc, double *total);
int patmol_uhf_moments(const double *PA, const double *PB, double *elec, double *nuc, double *total);
/* finite-field polarisability (rhfscf.c:176-215) through the device SCF:
 * alfa[5] = xx, yy, zz, mean, and the reference's alfa2 */
int patmol_rhf_finite_field(double efx, double *alfa, PatmolScfResult *gs);

/* RHF nuclear gradient (no counterpart in the reference): grad[3 c + x] =
 * dE/dR_{c,x} for every centre at the converged half density P (packed),
 * E = 2 sum P H + sum P G(P) + E_nuc; two-electron part on the device
 * (patmol_hip_eri_gradient over mints, or over a handle built here when
 * mints is null), one-electron and nuclear-repulsion parts on the host with
 * W = P F P.  Returns 0 or an error code. */
int patmol_rhf_gradient(const ERIFS *mints, const double *P, double *grad);
/* The same for the DPM-augmented energy of one grid point (the basis after
 * dpm_augment(), the positron at skip_center()): H carries -V of dpm_ints,
 * so F and W do too, and the DPM term's derivative (patmol_hip_dpm_gradient)
 * is added; the positron centre's entry is dE/dQ.  mints must cover the
 * whole augmented basis (start 0), or null. */
int patmol_rhf_dpm_gradient(const ERIFS *mints, const double *P, double *grad);
/* its host parts: grad += 2 sum P dH/dR - 2 sum W dS/dR (P, W packed), and
 * grad += dE_nuc/dR */
int patmol_host_onee_gradient(const double *P, const double *W, double *grad);
void patmol_nuclear_repulsion_gradient(double *grad);

/* DPM grid loop, rhf_dpm_calculation / uhf_dpm_calculation (rhfscf.c:217-263,
 * uhfscf.c:213-262), on the device SCF.  The DPM centre and shells must have
 * been added (patmol_basis_add_dpm_center, or init_Basis with the dpm flag)
 * and not yet made visible: the ground state is computed in the electron
 * basis, then dpm_augment() and, per grid point pts[3i..3i+2] (bohr): the
 * centre moved (set_center_position), S, T, H of the augmented basis with
 * the DPM potential subtracted (one_electron_ints, dpm_ints), the ERIs with
 * a DPM shell (two_electron_ints(xints, number of electron shells)), the SCF
 * over mints + xints from the ground-state density (zero over the DPM
 * functions), and vpol = E_conv - E_first with E_first the energy after the
 * first update (rhfscf.c:245-253).  alfa = -2 vpol |r|^4 as rhfscf.c:256.
 * uhf != 0 runs the UHF form (uhfscf.c) with the basis's multiplicity.
 * The reference's vstat (rhfscf.c:254) reads an uninitialised variable and
 * is not reproduced. */
typedef struct {
    double vpol, alfa, e_first, e_conv;
    int iterations, converged;
} PatmolDpmPoint;
int patmol_dpm_calculation(int uhf, int npts, const double *pts, PatmolDpmPoint *out, PatmolScfResult *gs);

#ifdef __cplusplus
}
#endif
