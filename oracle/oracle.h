/*
 * oracle.h -- CPU restatement of Patmol's two-electron hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (patmol_amd/) never links, loads or calls anything in oracle/.
 *
 * Every function restates the reference's algorithm (file:line cited in
 * oracle.c).  Parity status of each function is listed in oracle.c's header.
 */
#ifndef PATMOL_ORACLE_H
#define PATMOL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXL 4          /* s..g shells */
#define OR_MAXLST 15       /* (MAXL+1)(MAXL+2)/2 */
#define OR_MAXROOTS 16

/* Same field order as the reference's Shell (shell.h:9-13) and Center (center.h:5-7). */
typedef struct {
    int npr, lsh, cen, off, nst;
    double *al;
    double *co;
} OShell;

typedef struct {
    double q, r[3];
} OCenter;

typedef struct {
    int nshell, norb, ncen, maxl, maxlst, skipcen;
    OShell *sh;
    OCenter *cen;
    int lxyz[OR_MAXL + 1][OR_MAXLST][3];
    double nf[OR_MAXL + 1][OR_MAXLST];
} OBasis;

/* 16-byte ERI record, same layout as the reference's ERINT (erifs.h:19-22). */
typedef struct {
    double val;
    unsigned short i, j, k, l;
} __attribute__((aligned(16))) OERINT;

/* ---- basis model (shell.c:47-77, basis.c:243-321) ---- */
OBasis *or_basis_new(int ncen, const double *q, const double *xyz,
                     int nshell, const int *npr, const int *lsh, const int *cen,
                     const double *al, const double *co, int normalize);
void or_basis_free(OBasis *b);
int or_basis_norb(const OBasis *b);
void or_basis_coefs(const OBasis *b, double *co_out);
void or_basis_set_center(OBasis *b, int icen, const double *r);
void or_basis_set_skipcen(OBasis *b, int skipcen);
double or_nuclear_repulsion(const OBasis *b);

/* ---- primitives ---- */
void or_boys(int m, double T, double *F);                       /* mdr.c:6-34 */
int or_rys(int n, double T, double *x, double *w);              /* rys.c:2217-2243 (own method) */
void or_set_root_provider(void *fn);
void or_set_verbose(int v);
void or_hermite_e(double *d, int l1, int l2, double abi, double ax, double bx); /* mdd.c:3-43 */
void or_hermite_r(double *r, int ltot, double sr, double T, double w, const double *pq); /* mdr.c:37 */

/* ---- ERIs ---- */
void or_eri_block(const OBasis *b, int I, int J, int K, int L, double *out);
void or_eri_block_noscreen(const OBasis *b, int I, int J, int K, int L, double *out);
int64_t or_two_electron_ints(const OBasis *b, int start, OERINT **out);
void or_free(void *p);
void or_eri_count(const OBasis *b, int start, int64_t *counts);

/* ---- Fock (G) build ---- */
void or_rhf_form_gmatrix(const OERINT *ints, int64_t n, const double *P, double *G);
void or_uhf_form_gmatrix(const OERINT *ints, int64_t n, const double *PA, const double *PB,
                         double *GA, double *GB);
void or_rhf_form_gmatrix_mt(const OERINT *ints, int64_t n, int norb, const double *P,
                            double *G, int nthreads);
void or_rhf_direct_gmatrix(const OBasis *b, int start, const double *P, double *G, int nthreads,
                           int64_t max_quartets, int64_t *done_quartets);
void or_rhf_direct_gmatrix_sampled(const OBasis *b, int start, const double *P, double *G, int nthreads,
                                   int64_t max_quartets, int64_t *done_quartets, const double *q,
                                   double tau, int sample_mod);
void or_pair_schwarz(const OBasis *b, double *q, int nthreads);
void or_g_rows(const OBasis *b, const double *Pfull, int nrows, const int *rows,
               double *Grows, int nthreads);

/* ---- one-electron / DPM ---- */
void or_one_electron_ints(const OBasis *b, double *S, double *T, double *H);
void or_moment_ints(const OBasis *b, double *dip, double *quad);
void or_dpm_ints(const OBasis *b, const double *q, double *hmat);

/* ---- SCF (rhfscf.c:21-174,432-471; uhfscf.c:22-168,485-527) ---- */
typedef struct {
    double energy, ediff, pdiff, enuc, eelec;
    int iterations, converged;
} OScfResult;

int or_rhf_scf_ex(const OBasis *b, int nelec, int maxits, double eps, int scf_a, const OERINT *ints,
                  int64_t nints, const double *Hin, const double *Pguess, double *P_out, double *evals_out,
                  OScfResult *res, int nthreads, double *energy0);
int or_rhf_scf(const OBasis *b, int nelec, int maxits, double eps, int scf_a,
               const OERINT *ints, int64_t nints, double *P_out, double *evals_out,
               OScfResult *res, int nthreads);
int or_uhf_scf(const OBasis *b, int nelec, int mult, int maxits, double eps, int scf_a,
               const OERINT *ints, int64_t nints, double *PA_out, double *PB_out,
               OScfResult *res, int nthreads);
int or_eigh(int n, const double *A, double *evals, double *evecs);
void or_check_phase(int n, double *z);
int or_diag_fock(int n, const double *S, const double *F, int nocc, double *ev, double *C, double *P);

#ifdef __cplusplus
}
#endif
#endif
