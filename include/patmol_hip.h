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
 *   momints            replaces momints.h      (momints.c:94-190)
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
/* MOMINTS.DAT in the working directory: the reference's records
 * (int ir, int jr, double dip[3], double quad[6]) in its order. */
void momints(void);

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
 * Q_IJ*Q_KL*max|P| is below `threshold` run in FP32 (FP32 integrals and
 * products, FP64 block sums, fixed-point accumulation), max|P| taken over
 * the six density blocks each item contracts with (RHF builds) or over the
 * whole density (UHF, batches); only in the classes whose FP32 kernel the
 * autotuner measured faster than their FP64 kernel (jk3_tuned.inc), and not
 * in (dd|dd) while the McMurchie-Davidson engine runs it (its FP64 items
 * there are faster).  The device SCF (patmol_hip_rhf_scf / _uhf_scf) builds
 * its last iterations in FP64: from the iteration whose density change is
 * within 20x the stop bound, or stops falling below 1e-6. */
void patmol_hip_set_precision(int mode, double threshold);
/* Autotuning aid (no counterpart in the reference): 1 sends the items below
 * the threshold to FP32 in every class, 0 (default) follows jk3_tuned.inc. */
void patmol_hip_set_fp32_policy(int every_class);
/* No counterpart in the reference (the reference has one integral algorithm,
 * Rys quadrature).  Which classes run their FP64 work items on the
 * McMurchie-Davidson kernel (Hermite expansion, the two contractions on the
 * FP64 matrix cores) instead of their Rys kernels, a mask: bit 0 (dd|dd),
 * bit 1 (dd|dp), bit 2 (dp|dp), bit 3 (dd|pp); -1 restores the default,
 * (dd|dd) and (dd|dp), where (dd|dp) stays on its Rys kernel for a basis
 * whose (dd|dp) work items are mostly kept (patmol_hip_jk_engine_mask);
 * bit 4 keeps every class of the mask on the McMurchie-Davidson kernel
 * regardless.  For comparisons and timing: both engines give the
 * reference's integrals. */
void patmol_hip_set_jk_engine(int mask);

/* ---- device-resident G build (inputs already in HBM) ---- */
/* d_P: packed half-density (no2 doubles, device); d_G: packed G (device),
 * incremented by 2J[P]-K[P].  `stream` is a hipStream_t (0 = default).
 * Asynchronous with respect to the host. */
int patmol_hip_rhf_gmatrix_device(const ERIFS *tfiles, const double *d_P, double *d_G, void *stream);

/* ---- allocation-free device G build ---- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: the same G += 2J[P] - K[P] as
 * patmol_hip_rhf_gmatrix_device (bitwise identical in FP64), with every
 * work buffer in a fixed device array of the library (17,301,504 bytes per
 * device, held until the process ends) instead of allocated memory.  No
 * hipMalloc/hipFree and no allocating library call; one stream (the
 * caller's), no events.  Always FP64 (patmol_hip_set_precision does not
 * apply: its split needs max|P| on the host).  Bases above
 * PATMOL_HIP_NOALLOC_MAX_NORB functions fail with PATMOL_HIP_EUNSUPPORTED.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other streams or threads) race and must be serialised by the caller.
 * The handle's tables come from two_electron_ints, which does allocate. */
#define PATMOL_HIP_NOALLOC_MAX_NORB 256
int patmol_hip_rhf_gmatrix_device_noalloc(const ERIFS *tfiles, const double *d_P, double *d_G, void *stream);

/* ---- ERINT record interop (erifs.h:19-22) ---- */
/* Writes every canonical integral of the basis with |v| > 1e-14 as ERINT
 * records, in the order of two_electron_ints (twoints.c:228-379), shells
 * from `start` on, into one file at `path` (no 1 GiB roll-over, no file-name
 * scheme).  The handle must have run two_electron_ints.  Returns the record
 * count, or -1. */
int64_t patmol_hip_write_erints(const ERIFS *tfiles, int start, const char *path);
/* G += 2J[P] - K[P] (scfaux.c:428-552) from ERINT records, as the reference
 * computes it from its files, on the GPU: records from host memory or
 * streamed from a file of records.  Indices beyond the current basis are
 * rejected (PATMOL_HIP_EBADARG) before any device work on them. */
int patmol_hip_rhf_gmatrix_erints(const ERINT *recs, int64_t nrec, const double *Pmat, double *Gmat);
int patmol_hip_rhf_gmatrix_erint_file(const char *path, const double *Pmat, double *Gmat);
/* GA += J[PA+PB] - K[PA], GB likewise (scfaux.c:554-714) from records. */
int patmol_hip_uhf_gmatrix_erints(const ERINT *recs, int64_t nrec, const double *PmatA, const double *PmatB,
                                  double *GmatA, double *GmatB);
int patmol_hip_uhf_gmatrix_erint_file(const char *path, const double *PmatA, const double *PmatB,
                                      double *GmatA, double *GmatB);

/* ---- multipole integrals (momints.c:14-92) ---- */
/* Dipole (3 x no2: x, y, z) and quadrupole (6 x no2: xx, xy, xz, yy, yz, zz)
 * integrals about the coordinate origin, packed lower triangles,
 * component-major, host arrays. */
int patmol_hip_moment_ints(double *dip, double *quad);
/* momints() with an explicit output path. */
int patmol_hip_write_momints(const char *path);

/* ---- device SCF (the whole iteration in HBM) ---- */
/* The reference's RHF / UHF iteration (rhfscf.c:21-95, uhfscf.c:22-100)
 * with every matrix on the GPU: G builds of `mints` (and `xints`, may be
 * NULL) as rhf_form_gmatrix / uhf_form_gmatrix, F' = X^T (H + G) X, the
 * eigenproblem (rocSOLVER dsyevd with the reference's check_phase sign
 * convention, spack_math.c:274-319), C = X V, P = C_occ C_occ^T
 * (scfaux.c:5-41), calc_energy / pdiffNorm (scfaux.c:97-136) and the
 * damping of rhfscf.c:60-69 / scfAccelerator (scfaux.c:64-95).  X = U s^-1/2
 * as rhf_form_xmatrix (rhfscf.c:432-452) fails with PATMOL_HIP_EDEVICE on an
 * overlap eigenvalue <= DBL_EPSILON.  The guess is the core Hamiltonian
 * (rhfscf.c:454-471, uhfscf.c:505-527) unless one is passed.  scf_eps gets
 * the floor 10 DBL_EPSILON N^2 (rhfscf.c:161-164).  scf_accel 0, 1, 2 are the
 * reference's (none, averaging, scfAccelerator on dE > 0); scf_accel 3 (RHF
 * only, no counterpart in the reference) extrapolates the Fock matrix by
 * Pulay DIIS over the last 8 iterations (commutator error X^T (FDS - SDF) X)
 * instead, for systems the reference's damping does not converge.  S, H, guesses and
 * outputs are host packed arrays (Cmat row-major N x N, MO j in column j);
 * any output may be NULL.  Returns 0 or an error code.  Set
 * PATMOL_HIP_SCF_LOG=1 for the reference's per-iteration log lines. */
typedef struct {
    double energy, eelec, enuc, ediff, pdiff;
    int iterations, converged;
    double gbuild_ms;   /* device time in the G builds */
    double other_ms;    /* the rest of the loop (eigensolver, density, sums, host logic) */
    double total_ms;
    double energy0;     /* energy after the first update (rhfscf.c:86, 191) */
    int purify_fallbacks; /* iterations whose purification did not converge */
    int full_builds;    /* G builds of the whole density */
    int delta_builds;   /* incremental G builds (patmol_hip_set_scf_incremental) */
    double items;       /* work items over all G builds of the run */
    double items_fp32;  /* of those, launched in FP32 */
    double items_skipped; /* of those, skipped by density-weighted screening */
} PatmolHipScfResult;
int patmol_hip_rhf_scf(const ERIFS *mints, const ERIFS *xints, const double *Smat, const double *Hmat,
                       int nocc, int maxits, double scf_eps, int scf_accel, double enuc,
                       const double *Pguess, double *Pout, double *evals, double *Cmat,
                       PatmolHipScfResult *res);
int patmol_hip_uhf_scf(const ERIFS *mints, const ERIFS *xints, const double *Smat, const double *Hmat,
                       int nocca, int noccb, int maxits, double scf_eps, int scf_accel, double enuc,
                       const double *PAguess, const double *PBguess, double *PAout, double *PBout,
                       double *evalsA, double *evalsB, PatmolHipScfResult *res);
/* Density inside the device SCF loop: mode 1 (default) by trace-correcting
 * purification of X^T F X for bases of 128 functions or more (GEMMs only;
 * an iteration whose purification does not converge, e.g. for a degenerate
 * Fermi level, is diagonalised instead), mode 0 by the eigensolver every
 * iteration as the reference does.  Mode 2 is mode 1 with the traces of
 * each purification step summed by a kernel of their own (the form mode 1
 * takes by itself above 1024 tiles of 32 x 32, N > 1024).  The guess and the
 * returned orbitals always come from the eigensolver. */
void patmol_hip_set_scf_density(int mode);
/* Incremental G in patmol_hip_rhf_scf (no counterpart in the reference;
 * period 0, the default, builds G of the whole density every iteration as
 * the reference does).  With period k > 0 the whole density is built every
 * k-th iteration (the first included) and in between only G(P - Pref) is
 * built and added to G(Pref), Pref the previous iteration's density: G is
 * linear in P.  The difference build skips quartets whose Schwarz bound
 * times max|P - Pref| is below the Schwarz threshold, and in mixed precision
 * (patmol_hip_set_precision) splits FP32 from FP64 items by that product.
 * The result's items / items_fp32 / items_skipped count the work. */
void patmol_hip_set_scf_incremental(int period);
/* One diagonalisation (spack_trans + spack_diag + form_cmatrix +
 * form_pmatrix) of host packed F in the basis orthogonalised by S: ascending
 * eigenvalues, C (row-major, phase-fixed) and P for nocc occupied orbitals. */
int patmol_hip_diag_fock(const double *Smat, const double *Fmat, int nocc, double *evals, double *Cmat,
                         double *Pmat);
/* UHF G build on device arrays: dGA += J[PA+PB] - K[PA], dGB likewise
 * (scfaux.c:554-714), asynchronous on `stream`. */
int patmol_hip_uhf_gmatrix_device(const ERIFS *tfiles, const double *d_PA, const double *d_PB, double *d_GA,
                                  double *d_GB, void *stream);

/* ---- molecule batches ---- */
/* No counterpart in the reference (it treats one molecule per run).  Many
 * independent molecules of uneven size in one G build: each
 * patmol_hip_batch_add snapshots the basis current at the call (the
 * basis.h accessors, as two_electron_ints reads them, with the current
 * Schwarz threshold) and returns its molecule index; patmol_hip_batch_finalize
 * merges every molecule's screened work list into one queue per
 * angular-momentum class (work items never mix molecules), splits the items
 * whose bound x dmax is below `threshold` to the FP32 kernels when
 * prec_mode is 1, and captures the whole G build (unpack, zeroing, one launch
 * per class over all molecules, finalisation) as a HIP graph.  Densities and
 * G of the batch are the molecules' packed triangles concatenated in the
 * order they were added.  patmol_hip_batch_gmatrix adds 2J[P] - K[P] of
 * every molecule to host G (use_graph 0 issues the same launches without the
 * graph); patmol_hip_batch_buffers / patmol_hip_batch_launch give device
 * callers the batch's own P and G buffers and one graph replay on their
 * stream.  The J/K sums are fixed point: results are bitwise reproducible
 * and independent of the other molecules of the batch. */
typedef struct PatmolHipBatch PatmolHipBatch;
PatmolHipBatch *patmol_hip_batch_new(void);
int patmol_hip_batch_add(PatmolHipBatch *b);
int patmol_hip_batch_finalize(PatmolHipBatch *b, int prec_mode, double threshold, double dmax);
int patmol_hip_batch_gmatrix(PatmolHipBatch *b, const double *P, double *G, int use_graph);
int patmol_hip_batch_buffers(PatmolHipBatch *b, double **d_P, double **d_G);
int patmol_hip_batch_launch(PatmolHipBatch *b, void *stream);
int patmol_hip_batch_info(PatmolHipBatch *b, int *nmol, int64_t *no2tot, int64_t *items, int64_t *items_fp32);
void patmol_hip_batch_free(PatmolHipBatch *b);
/* RHF of every molecule of a finalised batch, all iterating together: per
 * iteration one replay of the batch's G graph, then per molecule not yet
 * converged the density step of patmol_hip_rhf_scf (DIIS for scf_accel 3,
 * else the plain Roothaan step; scf_accel 1 and 2 are not supported here)
 * and the reference's convergence test (rhfscf.c:52-59, per molecule).
 * S, H: host packed one-electron matrices of the molecules concatenated
 * (batch order); nocc, enuc per molecule; outputs per molecule (energy,
 * iterations, converged) and optionally the densities (concatenated);
 * total_ms / gbuild_ms: wall time of the loop and of its G builds. */
int patmol_hip_batch_rhf_scf(PatmolHipBatch *b, const double *S, const double *H, const int *nocc,
                             const double *enuc, int maxits, double scf_eps, int scf_accel, double *energy,
                             int *iterations, int *converged, double *Pout, double *total_ms, double *gbuild_ms);

/* ---- nuclear gradient ---- */
/* No counterpart in the reference (it has no gradients).  The two-electron
 * part of the RHF energy gradient for the packed half density P (host,
 * no2): grad[3 c + x] += sum_{abcd} Gamma_abcd d(ab|cd)/dR_{c,x} with
 * Gamma = 2 P_ab P_cd - (P_ac P_bd + P_ad P_bc)/2, over the handle's
 * Schwarz-screened quartets (the energy's own screen), for every centre c of
 * the basis (grad has 3 * number_of_centers() doubles, accumulated).  FP64,
 * reproducible bit for bit. */
int patmol_hip_eri_gradient(const ERIFS *tfiles, const double *Pmat, double *grad);
/* The one-electron part of the same gradient (no counterpart in the
 * reference): grad[3 c + x] += 2 sum P dH/dR - 2 sum W dS/dR for the packed
 * half density P and energy-weighted density W of the current basis, H = T
 * + V without skip_center()'s attraction (V's derivative includes the
 * nuclei's own Hellmann-Feynman term).  One GPU thread per shell pair;
 * fixed-point sums, reproducible. */
int patmol_hip_onee_gradient(const double *Pmat, const double *Wmat, double *grad);
/* The DPM potential's part of the gradient (no counterpart in the
 * reference): for the packed half density P of the current basis,
 * grad[3 c + x] += d/dR_{c,x} of -2 sum_ab P_ab V_ab, V_ab the dpm_ints
 * integrals (dpmints.c:120-199) of the positron at skip_center(), for every
 * centre, the positron centre included (-(d/dA + d/dB) per pair).  One GPU
 * thread per shell pair; per-pair sums added on the host in pair order. */
int patmol_hip_dpm_gradient(const double *Pmat, double *grad);

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
/* No counterpart in the reference.  The classes whose FP64 items run on
 * the McMurchie-Davidson kernel for this handle (patmol_hip_set_jk_engine's
 * bits): the engine mask in force, less (dd|dp) when more than 40 % of the
 * shell-quartet slots of its work items survive the Schwarz screen (there
 * the Rys kernel measured faster), 0 when no Hermite tables were built;
 * negative on error. */
int patmol_hip_jk_engine_mask(const ERIFS *tfiles);
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
/* The same with prec 1: the integrals computed in the FP32 arithmetic of the
 * mixed-precision J/K kernels (FP32 roots, recursions and products; the sum
 * over primitive quartets in FP64), for the per-class FP32 error. */
int patmol_hip_eri_blocks_prec(const ERIFS *tfiles, int nq, const int *quartets, double *out, int prec);

/* Wall time of the last G build measured with HIP events around the kernel
 * launches (milliseconds), and of each class kernel (array of 36, indexed
 * bra_pc*6 + ket_pc). */
int patmol_hip_last_timing(const ERIFS *tfiles, double *total_ms, double *class_ms);
void patmol_hip_set_timing(int on);

#ifdef __cplusplus
}
#endif
#endif
