/*
 * ref_harness.c -- thin C entry points over the reference's own functions
 * (compiled from /root/reference into _ref/libpatref.so by Makefile.ref), so
 * tests can check the oracle against the reference itself.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing here is reference source: the ERI entry
 * drives the reference's Rys primitives (Rys_findRoots, Rys_set*, RysRecur,
 * RysShift; rys.c:24-64,2217-2335) through the primitive-quartet arithmetic
 * of calc_pteints (twoints.c:23-186), which is static in the reference and
 * so cannot be called directly.  two_electron_ints itself cannot run here:
 * it writes ERINT files under a hard-coded absolute scratch directory
 * (pconfig.h:5, erifs.c:62-95) that does not exist in this environment.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "basis.h"
#include "md.h"
#include "rys.h"
#include "util.h"
#include "dpmints.h"
#include "oneints.h"
#include "spack_math.h"
#include "scfaux.h"

void ref_init_basis(void) { init_Basis(); }
int ref_norb(void) { return number_of_orbitals(); }
int ref_nshell(void) { return number_of_shells(); }

void ref_shell_coefs(double *out)
{
    const Shell *sh = shells();
    int p = 0;
    for (int s = 0; s < number_of_shells(); ++s)
        for (int k = 0; k < sh[s].npr; ++k) out[p++] = sh[s].co[k];
}

void ref_one_electron_ints(double *S, double *T, double *H) { one_electron_ints(S, T, H); }

void ref_rys_roots(int n, double T, double *u, double *w)
{
    Rys_findRoots(n, T);
    memcpy(u, RysRoots(), sizeof(double) * n);
    memcpy(w, RysWeights(), sizeof(double) * n);
}

/* Same, converted to the oracle's convention x = t^2 = u/(1+u). */
void ref_rys_x(int n, double T, double *x, double *w)
{
    Rys_findRoots(n, T);
    const double *u = RysRoots(), *wt = RysWeights();
    for (int i = 0; i < n; ++i) {
        x[i] = u[i] / (1.0 + u[i]);
        w[i] = wt[i];
    }
}

void ref_dpm_ints(double *hmat) { dpm_ints(hmat); }

void ref_set_center(int icen, const double *r) { set_center_position(icen, r); }

double ref_nuclear_repulsion(void) { return nuclear_repulsion(); }

/* formMDD for one direction: out[(i*(l2+1)+j)*(l1+l2+1)+n] */
void ref_formMDD(double *out, int l1, int l2, double abi, double ax, double bx)
{
    int ds = l1 + l2 + 2;
    double ***d = dtensor(ds, ds, 2 * ds + 1);
    formMDD(d, abi, ax, bx, l1, l2);
    for (int i = 0; i <= l1; ++i)
        for (int j = 0; j <= l2; ++j)
            for (int n = 0; n <= i + j; ++n) out[(i * (l2 + 1) + j) * (l1 + l2 + 1) + n] = d[i][j][n];
    free_dtensor(d, ds, ds);
}

/* formMDR: out[(t*(L+1)+u)*(L+1)+v] for t+u+v <= L (L <= 10) */
void ref_formMDR(double *out, int L, double sr, double t, double w, const double *pq)
{
    int rs = 13;
    double ***r = dtensor(rs, rs, rs);
    formMDR(r, sr, t, w, pq, L);
    for (int a = 0; a <= L; ++a)
        for (int b = 0; b <= L - a; ++b)
            for (int c = 0; c <= L - a - b; ++c) out[(a * (L + 1) + b) * (L + 1) + c] = r[a][b][c];
    free_dtensor(r, rs, rs);
}

/* Contracted, normalised block (IJ|KL) in the caller's shell order. */
void ref_eri_block(int I, int J, int K, int L, double *out)
{
    static int init = 0;
    if (!init) {
        RysInit();
        init = 1;
    }
    const double SRterm = 34.9868366552497250;
    const Shell *sh = shells();
    const Center *ce = centers();
    int ***lxyz = lvector_components();
    double **nf = normalizationFactors();
    const Shell *S[4] = {&sh[I], &sh[J], &sh[K], &sh[L]};
    int sw12 = S[0]->lsh < S[1]->lsh, sw34 = S[2]->lsh < S[3]->lsh;
    const Shell *s1 = sw12 ? S[1] : S[0], *s2 = sw12 ? S[0] : S[1];
    const Shell *s3 = sw34 ? S[3] : S[2], *s4 = sw34 ? S[2] : S[3];
    const double *a = ce[s1->cen].r, *b = ce[s2->cen].r, *c = ce[s3->cen].r, *d = ce[s4->cen].r;
    int l1 = s1->lsh, l2 = s2->lsh, l3 = s3->lsh, l4 = s4->lsh;
    int n1 = (l1 + 1) * (l1 + 2) / 2, n2 = (l2 + 1) * (l2 + 2) / 2, n3 = (l3 + 1) * (l3 + 2) / 2,
        n4 = (l4 + 1) * (l4 + 2) / 2;
    int lvt12 = l1 + l2, lvt34 = l3 + l4, nroots = (lvt12 + lvt34) / 2 + 1;
    double abx[3], cdx[3], absq = 0, cdsq = 0;
    for (int x = 0; x < 3; ++x) {
        abx[x] = a[x] - b[x];
        cdx[x] = c[x] - d[x];
        absq += abx[x] * abx[x];
        cdsq += cdx[x] * cdx[x];
    }
    int ntot = n1 * n2 * n3 * n4;
    double *acc = (double *)calloc(ntot, sizeof(double));
    static double Gx[MAXROOTS][MAXROOTS][MAXROOTS], Gy[MAXROOTS][MAXROOTS][MAXROOTS],
        Gz[MAXROOTS][MAXROOTS][MAXROOTS];
    double p[3], q[3], pa[3], qc[3], s12 = 1.0, s34 = 1.0;
    if (a == b) {
        memcpy(p, a, sizeof p);
        memset(pa, 0, sizeof pa);
    }
    if (c == d) {
        memcpy(q, c, sizeof q);
        memset(qc, 0, sizeof qc);
    }
    for (int i = 0; i < s1->npr; ++i) {
        double axp = s1->al[i], c1 = s1->co[i], f12 = 1.0;
        int jend = s2->npr;
        if (s1->al == s2->al) {
            f12 = 2.0;
            jend = i + 1;
        }
        for (int j = 0; j < jend; ++j) {
            if (i == j) f12 = 1.0;
            double c12 = c1 * f12 * s2->co[j], bxp = s2->al[j], pxp = axp + bxp, abi = 1.0 / pxp;
            if (a != b) {
                s12 = exp(-axp * bxp * absq * abi);
                for (int x = 0; x < 3; ++x) {
                    p[x] = (axp * a[x] + bxp * b[x]) * abi;
                    pa[x] = p[x] - a[x];
                }
            }
            for (int k = 0; k < s3->npr; ++k) {
                double cxp = s3->al[k], c3 = s3->co[k], f34 = 1.0;
                int lend = s4->npr;
                if (s3->al == s4->al) {
                    f34 = 2.0;
                    lend = k + 1;
                }
                for (int l = 0; l < lend; ++l) {
                    if (k == l) f34 = 1.0;
                    double c34 = c3 * f34 * s4->co[l], dxp = s4->al[l], qxp = cxp + dxp, cdi = 1.0 / qxp;
                    if (c != d) {
                        s34 = exp(-cxp * dxp * cdsq * cdi);
                        for (int x = 0; x < 3; ++x) {
                            q[x] = (cxp * c[x] + dxp * d[x]) * cdi;
                            qc[x] = q[x] - c[x];
                        }
                    }
                    double txp = pxp + qxp, sr = SRterm * s12 * s34 * abi * cdi / sqrt(txp);
                    if (sr < 1e-15) continue;
                    sr *= c12 * c34;
                    double pq[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
                    double t = (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]) * pxp * qxp / txp;
                    Rys_findRoots(nroots, t);
                    const double *roots = RysRoots();
                    for (int ir = 0; ir < nroots; ++ir) {
                        double rr = roots[ir], drt = rr / (1.0 + rr), fff = drt / txp, b00 = 0.5 * fff;
                        Rys_setB00(b00);
                        Rys_setB1((0.5 - b00 * qxp) / pxp);
                        Rys_setB1p((0.5 - b00 * pxp) / qxp);
                        Rys_setC(pa[0] - qxp * pq[0] * fff);
                        Rys_setCp(qc[0] + pxp * pq[0] * fff);
                        RysRecur(Gx[ir], lvt12, lvt34);
                        Rys_setC(pa[1] - qxp * pq[1] * fff);
                        Rys_setCp(qc[1] + pxp * pq[1] * fff);
                        RysRecur(Gy[ir], lvt12, lvt34);
                        Rys_setC(pa[2] - qxp * pq[2] * fff);
                        Rys_setCp(qc[2] + pxp * pq[2] * fff);
                        RysRecur(Gz[ir], lvt12, lvt34);
                    }
                    const double *wg = RysWeights();
                    int idx = 0;
                    for (int ia = 0; ia < n1; ++ia)
                        for (int ib = 0; ib < n2; ++ib)
                            for (int ic = 0; ic < n3; ++ic)
                                for (int id = 0; id < n4; ++id, ++idx) {
                                    const int *A = lxyz[l1][ia], *B = lxyz[l2][ib], *Cc = lxyz[l3][ic],
                                              *D = lxyz[l4][id];
                                    double sum = 0.0;
                                    for (int ir = 0; ir < nroots; ++ir)
                                        sum += RysShift(abx[0], cdx[0], Gx[ir], A[0] + B[0], B[0], Cc[0] + D[0], D[0]) *
                                               RysShift(abx[1], cdx[1], Gy[ir], A[1] + B[1], B[1], Cc[1] + D[1], D[1]) *
                                               RysShift(abx[2], cdx[2], Gz[ir], A[2] + B[2], B[2], Cc[2] + D[2], D[2]) *
                                               wg[ir];
                                    acc[idx] += sum * sr;
                                }
                }
            }
        }
    }
    int nI = S[0]->nst, nJ = S[1]->nst, nK = S[2]->nst, nL = S[3]->nst;
    for (int i = 0; i < nI; ++i)
        for (int j = 0; j < nJ; ++j)
            for (int k = 0; k < nK; ++k)
                for (int l = 0; l < nL; ++l) {
                    int ia = sw12 ? j : i, ib = sw12 ? i : j, ic = sw34 ? l : k, id = sw34 ? k : l;
                    out[((i * nJ + j) * nK + k) * nL + l] =
                        acc[((ia * n2 + ib) * n3 + ic) * n4 + id] * nf[l1][ia] * nf[l2][ib] * nf[l3][ic] *
                        nf[l4][id];
                }
    free(acc);
}

/* Host eigensolver (spack_math.c:299-319) on a packed symmetric matrix. */
void ref_spack_diag(int n, double *packed, double *z, double *evals)
{
    double *tmp = (double *)malloc(sizeof(double) * n * n);
    spack_diag(n, packed, z, evals, tmp);
    free(tmp);
}

/* One diagonalisation step of the reference's RHF driver from its own
 * functions: X as rhf_form_xmatrix (rhfscf.c:432-452, restated here because
 * it works on rhfscf.c's static globals), then spack_trans, spack_diag,
 * form_cmatrix and form_pmatrix as rhf_form_guess does (rhfscf.c:465-470). */
void ref_diag_fock(int n, const double *S, const double *F, int nocc, double *ev, double *C, double *P)
{
    const size_t no2 = (size_t)n * (n + 1) / 2;
    double *a = (double *)malloc(sizeof(double) * no2), *w = (double *)malloc(sizeof(double) * n * n);
    double *x = (double *)malloc(sizeof(double) * n * n), *tmp = (double *)malloc(sizeof(double) * n * n);
    memcpy(a, S, sizeof(double) * no2);
    spack_diag(n, a, w, ev, tmp);
    for (int j = 0; j < n; ++j) ev[j] = 1. / sqrt(ev[j]);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) x[i * n + j] = w[i * n + j] * ev[j];
    memcpy(a, F, sizeof(double) * no2);
    spack_trans(n, a, x, tmp);
    spack_diag(n, a, w, ev, tmp);
    form_cmatrix(n, C, x, w);
    form_pmatrix(n, nocc, C, P);
    free(a);
    free(w);
    free(x);
    free(tmp);
}

/* The reference's check_phase (spack_math.c:274-297) on the rows of z (row
 * i = vector i), declared here because spack_math.h does not export it. */
extern void check_phase(int n, double *z);
void ref_check_phase(int n, double *z) { check_phase(n, z); }
