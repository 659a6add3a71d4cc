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

#ifdef __cplusplus
}
#endif
