/*
 * patmol_hip.h -- C ABI of libpatmol_hip.so, the MI355X (gfx950) two-electron
 * hot path of Patmol.
 *
 * Drop-in entry points keep the reference's names and signatures:
 *   init_ERIFS         replaces erifs.h:27      (erifs.c:32-60)
 *   clean_ERIFS        replaces erifs.h:28      (erifs.c:97-110)
 *   two_electron_ints  replaces twoints.h:14    (twoints.c:188-418)
 *   rhf_form_gmatrix   replaces scfaux.h:22-23  (scfaux.c:428-552)
 *   uhf_form_gmatrix   replaces scfaux.h:24-26  (scfaux.c:554-714)
 *   dpm_ints           replaces dpmints.h:12    (dpmints.c:120-199)
 *   one_electron_ints  replaces oneints.h      (oneints.c:137-234)
 * Instead of writing ERINT files, two_electron_ints builds the device-side
 * shell-pair and screened work tables for the basis current at the call;
 * rhf_form_gmatrix/uhf_form_gmatrix then run a direct (recompute) J/K build
 * on the GPU and ADD the result into the caller's packed lower-triangle G,
 * synchronously, exactly as the reference's file-driven scatter does.
 *
 * The library reads the basis through the reference's basis.h accessors
 * (shells(), centers(), number_of_shells(), number_of_orbitals(),
 * number_of_centers(), skip_center()); a reference driver supplies them from
 * its own basis.c, a standalone caller from libpatmol_host.so.
 *
 * The reference reports failures by exiting (util.c:4-26).  Here every entry
 * point records a status readable with patmol_hip_last_error(); with
 * patmol_hip_set_strict(1) (or PATMOL_HIP_STRICT=1 in the environment) a
 * failure prints the message and aborts instead, as the reference does.
 */
#ifndef PATMOL_HIP_H
#define PATMOL_HIP_H
#include <stdint.h>
#include "patmol_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference-compatible entry points ---- */
ERIFS *init_ERIFS(int nst, int nend, const char *sname);
void clean_ERIFS(ERIFS *e);
void two_electron_ints(ERIFS *tfiles, int start);
void rhf_form_gmatrix(const ERIFS *tfiles, const double *Pmat, double *Gmat);
void uhf_form_gmatrix(const ERIFS *tfiles, const double *PmatA, const double *PmatB,
                      double *GmatA, double *GmatB);
void dpm_ints(double *hmat);
/* S, T and H = T + V of the current basis into caller-owned packed no2
 * arrays (assigned, not accumulated); V leaves out skip_center(). */
void one_electron_ints(double *Smat, double *Tmat, double *Hmat);

/* ---- status ---- */
#define PATMOL_HIP_OK 0
#define PATMOL_HIP_EBADARG 1     /* invalid argument or handle */
#define PATMOL_HIP_EDEVICE 2     /* HIP runtime failure */
#define PATMOL_HIP_EUNSUPPORTED 3 /* angular momentum above d, > 65535 functions */
#define PATMOL_HIP_ESTATE 4      /* two_electron_ints not run for this handle */
int patmol_hip_last_error(void);
const char *patmol_hip_error_string(void);
void patmol_hip_set_strict(int strict);

/* ---- options (apply to subsequent two_electron_ints / G builds) ---- */
/* Schwarz threshold on Q_IJ*Q_KL (default 1e-12; 0 keeps every quartet). */
void patmol_hip_set_schwarz_threshold(double tau);
/* mode 0: all quartets in FP64.  mode 1: work items whose bound
 * Q_IJ*Q_KL*max|P| is below `threshold` run in FP32 (FP64 accumulation). */
void patmol_hip_set_precision(int mode, double threshold);

/* ---- device-resident G build (inputs already in HBM) ---- */
/* d_P: packed half-density (no2 doubles, device); d_G: packed G (device),
 * incremented by 2J[P]-K[P].  `stream` is a hipStream_t (0 = default).
 * Asynchronous with respect to the host. */
int patmol_hip_rhf_gmatrix_device(const ERIFS *tfiles, const double *d_P, double *d_G, void *stream);

/* ---- introspection ---- */
typedef struct {
    int nshell, norb, start;
    int64_t pairs;                 /* shell pairs */
    int64_t shell_quartets_total;  /* unique shell quartets before screening */
    int64_t shell_quartets_kept;   /* surviving the Schwarz (and start) screen */
    int64_t prim_quartets_kept;    /* primitive quartets inside the kept shell quartets */
    double flops_per_build;        /* algorithmic FLOPs of one G build over kept quartets */
    int64_t work_items;            /* tile quartets launched */
    int64_t work_items_fp32;       /* of which run in FP32 at the last G build */
} PatmolHipStats;
int patmol_hip_stats(const ERIFS *tfiles, PatmolHipStats *out);
/* Kept shell quartets and algorithmic FLOPs per class (36 entries each,
 * indexed bra_pc*6 + ket_pc, pair class pc = l(l+1)/2 + l' with l >= l'). */
int patmol_hip_class_stats(const ERIFS *tfiles, int64_t *quartets36, double *flops36);

/* Schwarz bound of every shell pair, indexed by original shell indices
 * (I >= J) at I*(I+1)/2 + J. */
int patmol_hip_pair_schwarz(const ERIFS *tfiles, double *q_out);

/* Work items as built by two_electron_ints: per item the four tile indices
 * (bra tile pair first), its class pair and its Schwarz bound; plus the
 * sorted shell order and tile table they refer to.  Any pointer may be null;
 * returns the item count. */
int64_t patmol_hip_work_items(const ERIFS *tfiles, int *tiles4, int *classes2, double *bounds,
                              int *sorted_orig, int *tile_first, int *tile_size, int *ntiles);

/* Contracted, normalised integrals of shell quartets given by original shell
 * indices (4 per quartet); out receives, per quartet, the nst(I)*nst(J)*
 * nst(K)*nst(L) values in (i,j,k,l) row-major order. */
int patmol_hip_eri_blocks(const ERIFS *tfiles, int nq, const int *quartets, double *out);

/* Wall time of the last G build measured with HIP events around the kernel
 * launches (milliseconds), and of each class kernel (array of 36, indexed
 * bra_pc*6 + ket_pc). */
int patmol_hip_last_timing(const ERIFS *tfiles, double *total_ms, double *class_ms);
void patmol_hip_set_timing(int on);

#ifdef __cplusplus
}
#endif
#endif
