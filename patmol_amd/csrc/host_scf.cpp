// host_scf.cpp -- RHF/UHF SCF drivers over the GPU G build.
//
// Restates the reference's iteration exactly (rhfscf.c:21-118 and
// uhfscf.c:22-118): core-Hamiltonian guess, canonical orthogonalisation
// X = U s^-1/2 (fatal when an overlap eigenvalue is <= DBL_EPSILON), energy
// from the density that entered the iteration, convergence on |dE| < 10 eps
// or on the density change below max(eps, 10 eps N^2), and damping only when
// the energy rose (simple averaging, or the two-point Anderson step of
// scfaux.c:63-95).  The first update of a run never tests convergence and
// resets the iteration counter (rhfscf.c:85-88).  The one-electron
// integrals and the G builds go through the library's own device
// one_electron_ints and rhf_form_gmatrix / uhf_form_gmatrix.
#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "patmol_hip.h"
#include "patmol_host.h"

namespace {

using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// Symmetric eigenproblem by Householder reduction to tridiagonal form and
// the implicit QL method (Golub & Van Loan, Matrix Computations, 8.3).
// a (n x n row-major) is overwritten by eigenvectors in its columns; d gets
// the eigenvalues in ascending order.
bool symmetric_eigen(int n, std::vector<double> &a, std::vector<double> &d)
{
    std::vector<double> e(n, 0.0);
    d.assign(n, 0.0);
    auto A = [&](int i, int j) -> double & { return a[(size_t)i * n + j]; };
    for (int i = n - 1; i > 0; --i) {
        int l = i - 1;
        double h = 0.0;
        if (l > 0) {
            double scale = 0.0;
            for (int k = 0; k <= l; ++k) scale += fabs(A(i, k));
            if (scale == 0.0) {
                e[i] = A(i, l);
            } else {
                for (int k = 0; k <= l; ++k) {
                    A(i, k) /= scale;
                    h += A(i, k) * A(i, k);
                }
                double f = A(i, l), g = f >= 0 ? -sqrt(h) : sqrt(h);
                e[i] = scale * g;
                h -= f * g;
                A(i, l) = f - g;
                f = 0.0;
                for (int j = 0; j <= l; ++j) {
                    A(j, i) = A(i, j) / h;
                    g = 0.0;
                    for (int k = 0; k <= j; ++k) g += A(j, k) * A(i, k);
                    for (int k = j + 1; k <= l; ++k) g += A(k, j) * A(i, k);
                    e[j] = g / h;
                    f += e[j] * A(i, j);
                }
                double hh = f / (h + h);
                for (int j = 0; j <= l; ++j) {
                    f = A(i, j);
                    e[j] = g = e[j] - hh * f;
                    for (int k = 0; k <= j; ++k) A(j, k) -= f * e[k] + g * A(i, k);
                }
            }
        } else {
            e[i] = A(i, l);
        }
        d[i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < n; ++i) {
        if (d[i] != 0.0)
            for (int j = 0; j < i; ++j) {
                double g = 0.0;
                for (int k = 0; k < i; ++k) g += A(i, k) * A(k, j);
                for (int k = 0; k < i; ++k) A(k, j) -= g * A(k, i);
            }
        d[i] = A(i, i);
        A(i, i) = 1.0;
        for (int j = 0; j < i; ++j) A(j, i) = A(i, j) = 0.0;
    }
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                volatile double dd = fabs(d[m]) + fabs(d[m + 1]);
                volatile double sum = fabs(e[m]) + dd;
                if (sum == dd) break;
            }
            if (m != l) {
                if (iter++ == 100) return false;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]), r = hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i], b = c * e[i];
                    e[i + 1] = r = hypot(f, g);
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    d[i + 1] = g + (p = s * r);
                    g = c * r - b;
                    for (int k = 0; k < n; ++k) {
                        f = A(k, i + 1);
                        A(k, i + 1) = s * A(k, i) + c * f;
                        A(k, i) = c * A(k, i) - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        for (int j = i + 1; j < n; ++j)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            std::swap(d[i], d[k]);
            for (int j = 0; j < n; ++j) std::swap(A(j, i), A(j, k));
        }
    }
    return true;
}

void unpack(int n, const double *p, std::vector<double> &f)
{
    f.assign((size_t)n * n, 0.0);
    for (int i = 0, ij = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j, ++ij) f[(size_t)i * n + j] = f[(size_t)j * n + i] = p[ij];
}

// F (packed) -> C = X V, V the eigenvectors of X^T F X
void diagonalise(int n, const double *Fp, const std::vector<double> &X, std::vector<double> &C, std::vector<double> &ev)
{
    std::vector<double> F, T((size_t)n * n, 0.0), M((size_t)n * n, 0.0);
    unpack(n, Fp, F);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            double f = F[(size_t)i * n + k];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) T[(size_t)i * n + j] += f * X[(size_t)k * n + j];
        }
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < n; ++i) {
            double x = X[(size_t)k * n + i];
            if (x == 0.0) continue;
            for (int j = 0; j < n; ++j) M[(size_t)i * n + j] += x * T[(size_t)k * n + j];
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) M[(size_t)i * n + j] = M[(size_t)j * n + i] = 0.5 * (M[(size_t)i * n + j] + M[(size_t)j * n + i]);
    symmetric_eigen(n, M, ev);
    C.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            double x = X[(size_t)i * n + k];
            if (x == 0.0) continue;
            for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += x * M[(size_t)k * n + j];
        }
}

bool form_x(int n, const double *S, std::vector<double> &X)
{
    std::vector<double> U, ev;
    unpack(n, S, U);
    symmetric_eigen(n, U, ev);
    X.assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        if (ev[j] <= DBL_EPSILON) return false;
        double s = 1.0 / sqrt(ev[j]);
        for (int i = 0; i < n; ++i) X[(size_t)i * n + j] = U[(size_t)i * n + j] * s;
    }
    return true;
}

void form_p(int n, int nocc, const std::vector<double> &C, double *P)
{
    for (int i = 0, ij = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j, ++ij) {
            double s = 0.0;
            for (int k = 0; k < nocc; ++k) s += C[(size_t)i * n + k] * C[(size_t)j * n + k];
            P[ij] = s;
        }
}

double calc_energy(int n, const double *pm, const double *hm, const double *gm)
{
    double s1 = 0.0, s2 = 0.0;
    for (int i = 0, ij = 0; i < n; ++i) {
        for (int j = 0; j < i; ++j, ++ij) {
            s1 += hm[ij] * pm[ij];
            s2 += gm[ij] * pm[ij];
        }
        s1 += 0.5 * pm[ij] * hm[ij];
        s2 += 0.5 * pm[ij] * gm[ij];
        ++ij;
    }
    return s1 + s1 + s2;
}

double pdiff_norm(int n, const double *p0, const double *p1)
{
    double sum = 0.0;
    for (int i = 0, ij = 0; i < n; ++i, ++ij) {
        for (int j = 0; j < i; ++j, ++ij) {
            double t = p0[ij] - p1[ij];
            sum += t * t;
        }
        double t = p0[ij] - p1[ij];
        sum += 0.5 * t * t;
    }
    return sqrt(2.0 * sum) / n;
}

void accelerate(int iter, size_t no2, double *p0, double *p1, double *p2)
{
    if (iter > 2) {
        double p00 = 0, p01 = 0, p11 = 0;
        for (size_t i = 0; i < no2; ++i) {
            double s0 = p0[i] - p1[i], s1 = p1[i] - p2[i];
            p00 += s0 * s0;
            p01 += s0 * s1;
            p11 += s1 * s1;
        }
        double beta = (p00 - p01) / (p00 - 2.0 * p01 + p11);
        beta = beta < 1e-3 ? 1e-3 : (beta > 1.5 ? 1.5 : beta);
        for (size_t i = 0; i < no2; ++i) {
            p2[i] = p1[i];
            p1[i] = p0[i];
            p0[i] = p1[i] * beta + p2[i] * (1.0 - beta);
        }
        return;
    }
    for (size_t i = 0; i < no2; ++i) {
        p2[i] = p1[i];
        p1[i] = p0[i];
        p0[i] = 0.5 * (p1[i] + p2[i]);
    }
}

}  // namespace

extern "C" int patmol_rhf_scf(double *P_out, double *evals_out, PatmolScfResult *res)
{
    const double EPS10 = 10 * DBL_EPSILON;
    const int n = number_of_orbitals(), ne = number_of_electrons();
    if (ne % 2 || spin_multiplicity() != 1) return 1;
    const int nocc = ne / 2;
    const size_t no2 = (size_t)n * (n + 1) / 2;
    std::vector<double> S(no2), T(no2), H(no2), G(no2), F(no2), P(no2), P1(no2, 0.0), P2(no2, 0.0);
    std::vector<double> X, C, ev;
    one_electron_ints(S.data(), T.data(), H.data());
    if (patmol_hip_last_error()) return 2;
    ERIFS *mints = init_ERIFS(0, n, "MINTS.DAT");
    two_electron_ints(mints, 0);
    if (patmol_hip_last_error()) {
        clean_ERIFS(mints);
        return 2;
    }
    double eps = scf_convergence();
    if (eps < EPS10 * n * n) eps = EPS10 * n * n;
    if (!form_x(n, S.data(), X)) {
        clean_ERIFS(mints);
        return 3;
    }
    diagonalise(n, H.data(), X, C, ev);
    form_p(n, nocc, C, P.data());
    const int maxits = max_scf_iterations(), accel = scf_acceleration();
    double enuc = nuclear_repulsion(), eold = 0.0, enrg = 0.0, ediff = 0.0, pdiff = 0.0, tg = 0.0, td = 0.0;
    int iteration = 0, done = 0;
    bool first = true;
    for (;;) {
        ++iteration;
        std::fill(G.begin(), G.end(), 0.0);
        auto t0 = Clock::now();
        rhf_form_gmatrix(mints, P.data(), G.data());
        auto t1 = Clock::now();
        tg += secs(t0, t1);
        if (patmol_hip_last_error()) {
            clean_ERIFS(mints);
            return 2;
        }
        for (size_t i = 0; i < no2; ++i) F[i] = H[i] + G[i];
        diagonalise(n, F.data(), X, C, ev);
        td += secs(t1, Clock::now());
        P2 = P1;
        P1 = P;
        form_p(n, nocc, C, P.data());
        pdiff = pdiff_norm(n, P.data(), P1.data());
        enrg = 2.0 * calc_energy(n, P1.data(), H.data(), G.data());
        ediff = enrg - eold;
        eold = enrg;
        int ret = 0;
        if (iteration == 1) {
        } else if (fabs(ediff) < EPS10 || pdiff < eps) {
            ret = 1;
        } else if (ediff > 0.0) {
            if (accel == 1)
                for (size_t i = 0; i < no2; ++i) P[i] = 0.5 * (P1[i] + P2[i]);
            if (accel == 2) accelerate(iteration, no2, P.data(), P1.data(), P2.data());
        }
        if (first) {
            first = false;
            iteration = 0;
        } else {
            done = ret;
        }
        // iteration log as rhfscf.c:89-95
        fprintf(stderr, "ITERATION = %12d EDIFF= %lg\n  energy = %.12f\n", iteration, ediff, enrg + enuc);
        if (iteration == 0) continue;
        if (done || iteration >= maxits) break;
    }
    clean_ERIFS(mints);
    res->energy = enrg + enuc;
    res->eelec = enrg;
    res->enuc = enuc;
    res->ediff = ediff;
    res->pdiff = pdiff;
    res->iterations = iteration;
    res->converged = done;
    res->gbuild_seconds = tg;
    res->diag_seconds = td;
    if (P_out) memcpy(P_out, P.data(), no2 * sizeof(double));
    if (evals_out) memcpy(evals_out, ev.data(), n * sizeof(double));
    return 0;
}

extern "C" int patmol_uhf_scf(double *PA_out, double *PB_out, PatmolScfResult *res)
{
    const double EPS10 = 10 * DBL_EPSILON;
    const int n = number_of_orbitals(), ne = number_of_electrons(), sm = spin_multiplicity();
    const int noccB = (ne - (sm - 1)) / 2, noccA = ne - noccB;
    const size_t no2 = (size_t)n * (n + 1) / 2;
    std::vector<double> S(no2), T(no2), H(no2), GA(no2), GB(no2), F(no2);
    std::vector<double> PA(no2), PA1(no2, 0.0), PA2(no2, 0.0), PB(no2), PB1(no2, 0.0), PB2(no2, 0.0);
    std::vector<double> X, C, ev;
    one_electron_ints(S.data(), T.data(), H.data());
    if (patmol_hip_last_error()) return 2;
    ERIFS *mints = init_ERIFS(0, n, "MINTS.DAT");
    two_electron_ints(mints, 0);
    if (patmol_hip_last_error()) {
        clean_ERIFS(mints);
        return 2;
    }
    double eps = scf_convergence();
    if (eps < EPS10 * n * n) eps = EPS10 * n * n;
    if (!form_x(n, S.data(), X)) {
        clean_ERIFS(mints);
        return 3;
    }
    diagonalise(n, H.data(), X, C, ev);
    form_p(n, noccA, C, PA.data());
    form_p(n, noccB, C, PB.data());
    const int maxits = max_scf_iterations(), accel = scf_acceleration();
    double enuc = nuclear_repulsion(), eold = 0.0, enrg = 0.0, ediff = 0.0, pdiff = 0.0, tg = 0.0, td = 0.0;
    int iteration = 0, done = 0;
    bool first = true;
    for (;;) {
        ++iteration;
        std::fill(GA.begin(), GA.end(), 0.0);
        std::fill(GB.begin(), GB.end(), 0.0);
        auto t0 = Clock::now();
        uhf_form_gmatrix(mints, PA.data(), PB.data(), GA.data(), GB.data());
        auto t1 = Clock::now();
        tg += secs(t0, t1);
        if (patmol_hip_last_error()) {
            clean_ERIFS(mints);
            return 2;
        }
        for (size_t i = 0; i < no2; ++i) F[i] = H[i] + GA[i];
        diagonalise(n, F.data(), X, C, ev);
        PA2 = PA1;
        PA1 = PA;
        form_p(n, noccA, C, PA.data());
        pdiff = pdiff_norm(n, PA.data(), PA1.data());
        for (size_t i = 0; i < no2; ++i) F[i] = H[i] + GB[i];
        diagonalise(n, F.data(), X, C, ev);
        td += secs(t1, Clock::now());
        PB2 = PB1;
        PB1 = PB;
        form_p(n, noccB, C, PB.data());
        pdiff += pdiff_norm(n, PB.data(), PB1.data());
        enrg = calc_energy(n, PA1.data(), H.data(), GA.data()) + calc_energy(n, PB1.data(), H.data(), GB.data());
        ediff = enrg - eold;
        eold = enrg;
        int ret = 0;
        if (iteration == 1) {
        } else if (fabs(ediff) < EPS10 || pdiff < eps) {
            ret = 1;
        } else if (ediff > 0.0) {
            if (accel == 1)
                for (size_t i = 0; i < no2; ++i) {
                    PA[i] = 0.5 * (PA1[i] + PA2[i]);
                    PB[i] = 0.5 * (PB1[i] + PB2[i]);
                }
            if (accel == 2) {
                accelerate(iteration, no2, PA.data(), PA1.data(), PA2.data());
                accelerate(iteration, no2, PB.data(), PB1.data(), PB2.data());
            }
        }
        if (first) {
            first = false;
            iteration = 0;
        } else {
            done = ret;
        }
        fprintf(stderr, "ITERATION = %12d EDIFF= %lg\n  energy = %.12f\n", iteration, ediff, enrg + enuc);
        if (iteration == 0) continue;
        if (done || iteration >= maxits) break;
    }
    clean_ERIFS(mints);
    res->energy = enrg + enuc;
    res->eelec = enrg;
    res->enuc = enuc;
    res->ediff = ediff;
    res->pdiff = pdiff;
    res->iterations = iteration;
    res->converged = done;
    res->gbuild_seconds = tg;
    res->diag_seconds = td;
    if (PA_out) memcpy(PA_out, PA.data(), no2 * sizeof(double));
    if (PB_out) memcpy(PB_out, PB.data(), no2 * sizeof(double));
    return 0;
}

// ---------------------------------------------------------------------------
// Multipole analysis and finite field
// ---------------------------------------------------------------------------
namespace {

// nuclear moments sum_c q_c r_c and q_c r_c r_c (scfaux.c:194-215)
void nuclear_moments(double *nuc)
{
    const Center *ce = centers();
    const int nc = number_of_centers();
    for (int m = 0; m < 9; ++m) nuc[m] = 0.0;
    for (int c = 0; c < nc; ++c) {
        const double q = ce[c].q, *r = ce[c].r;
        nuc[0] += q * r[0];
        nuc[1] += q * r[1];
        nuc[2] += q * r[2];
        nuc[3] += q * r[0] * r[0];
        nuc[4] += q * r[0] * r[1];
        nuc[5] += q * r[0] * r[2];
        nuc[6] += q * r[1] * r[1];
        nuc[7] += q * r[1] * r[2];
        nuc[8] += q * r[2] * r[2];
    }
}

// electronic moments sum_ij f_ij D_ij M_ij over the packed triangle
int moments_common(const double *D, double off, double diag, double *elec, double *nuc, double *total)
{
    const int n = number_of_orbitals();
    const size_t no2 = (size_t)n * (n + 1) / 2;
    std::vector<double> dip(3 * no2), quad(6 * no2);
    if (patmol_hip_moment_ints(dip.data(), quad.data())) return 2;
    for (int m = 0; m < 9; ++m) elec[m] = 0.0;
    for (int i = 0, ij = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j, ++ij) {
            const double p = (i == j ? diag : off) * D[ij];
            for (int m = 0; m < 3; ++m) elec[m] += p * dip[m * no2 + ij];
            for (int m = 0; m < 6; ++m) elec[3 + m] += p * quad[m * no2 + ij];
        }
    nuclear_moments(nuc);
    for (int m = 0; m < 9; ++m) total[m] = nuc[m] - elec[m];
    return 0;
}

}  // namespace

// rhf_analyzeMoments (scfaux.c:140-214) without the MOMINTS.DAT round trip:
// P is the half density, electronic moments weigh off-diagonal elements 4 and
// diagonal ones 2; 9 values (x, y, z, xx, xy, xz, yy, yz, zz) each.
extern "C" int patmol_rhf_moments(const double *P, double *elec, double *nuc, double *total)
{
    if (!P || !elec || !nuc || !total) return 1;
    return moments_common(P, 4.0, 2.0, elec, nuc, total);
}

// uhf_analyzeMoments (scfaux.c:267-340): weights 2 and 1 on PA + PB.
extern "C" int patmol_uhf_moments(const double *PA, const double *PB, double *elec, double *nuc, double *total)
{
    if (!PA || !PB || !elec || !nuc || !total) return 1;
    const int n = number_of_orbitals();
    const size_t no2 = (size_t)n * (n + 1) / 2;
    std::vector<double> D(no2);
    for (size_t i = 0; i < no2; ++i) D[i] = PA[i] + PB[i];
    return moments_common(D.data(), 2.0, 1.0, elec, nuc, total);
}

// rhf_finite_field_calculation (rhfscf.c:176-215) through the device SCF:
// ground state, then per direction H = H_gs - efx d_dir (rhf_getDipoleHam,
// rhfscf.c:474-489) from the ground-state density, and
// alpha_dir = -2 (E - E_first) / efx^2 with E_first the energy after the first
// update (enrg0, rhfscf.c:191).  alfa[0..2] = xx, yy, zz, alfa[3] = their
// mean, alfa[4] = (2 a_zz - a_yy - a_zz) / 3 exactly as rhfscf.c:211 writes it.
extern "C" int patmol_rhf_finite_field(double efx, double *alfa, PatmolScfResult *gs)
{
    const int n = number_of_orbitals(), ne = number_of_electrons();
    if (!alfa || efx == 0.0 || ne % 2 || spin_multiplicity() != 1) return 1;
    const size_t no2 = (size_t)n * (n + 1) / 2;
    std::vector<double> S(no2), T(no2), H(no2), Hf(no2), Pgs(no2), dip(3 * no2), quad(6 * no2);
    one_electron_ints(S.data(), T.data(), H.data());
    if (patmol_hip_last_error()) return 2;
    if (patmol_hip_moment_ints(dip.data(), quad.data())) return 2;
    ERIFS *mints = init_ERIFS(0, n, "MINTS.DAT");
    two_electron_ints(mints, 0);
    if (patmol_hip_last_error()) {
        clean_ERIFS(mints);
        return 2;
    }
    const int maxits = max_scf_iterations(), accel = scf_acceleration();
    const double eps = scf_convergence(), enuc = nuclear_repulsion();
    PatmolHipScfResult r{};
    if (patmol_hip_rhf_scf(mints, nullptr, S.data(), H.data(), ne / 2, maxits, eps, accel, enuc, nullptr,
                           Pgs.data(), nullptr, nullptr, &r)) {
        clean_ERIFS(mints);
        return 2;
    }
    if (gs) {
        gs->energy = r.energy;
        gs->eelec = r.eelec;
        gs->enuc = r.enuc;
        gs->ediff = r.ediff;
        gs->pdiff = r.pdiff;
        gs->iterations = r.iterations;
        gs->converged = r.converged;
        gs->gbuild_seconds = r.gbuild_ms * 1e-3;
        gs->diag_seconds = r.other_ms * 1e-3;
    }
    for (int d = 0; d < 3; ++d) {
        for (size_t i = 0; i < no2; ++i) Hf[i] = H[i] - efx * dip[d * no2 + i];
        PatmolHipScfResult rf{};
        if (patmol_hip_rhf_scf(mints, nullptr, S.data(), Hf.data(), ne / 2, maxits, eps, accel, enuc, Pgs.data(),
                               nullptr, nullptr, nullptr, &rf)) {
            clean_ERIFS(mints);
            return 2;
        }
        alfa[d] = -2.0 * (rf.energy - rf.energy0) / efx / efx;
    }
    clean_ERIFS(mints);
    alfa[3] = (alfa[0] + alfa[1] + alfa[2]) / 3.0;
    alfa[4] = (2.0 * alfa[2] - alfa[1] - alfa[2]) / 3.0;
    return 0;
}

// DPM grid loop (rhfscf.c:217-263, uhfscf.c:213-262) on the device SCF.
extern "C" int patmol_dpm_calculation(int uhf, int npts, const double *pts, PatmolDpmPoint *out, PatmolScfResult *gs)
{
    const int ono = number_of_orbitals(), onsh = number_of_shells(), ne = number_of_electrons();
    const int mult = spin_multiplicity();
    if (npts < 0 || (npts > 0 && (!pts || !out)) || ono <= 0) return 1;
    if (!uhf && (ne % 2 || mult != 1)) return 1;
    const int noccb = (ne - (mult - 1)) / 2, nocca = ne - noccb;
    const int maxits = max_scf_iterations(), accel = scf_acceleration();
    const double eps = scf_convergence();
    // ground state in the electron basis (rhf_energy / uhf_energy)
    const size_t ono2 = (size_t)ono * (ono + 1) / 2;
    std::vector<double> S(ono2), T(ono2), H(ono2), PgA(ono2), PgB(ono2);
    one_electron_ints(S.data(), T.data(), H.data());
    if (patmol_hip_last_error()) return 2;
    ERIFS *mints = init_ERIFS(0, ono, "MINTS.DAT");
    two_electron_ints(mints, 0);
    if (patmol_hip_last_error()) {
        clean_ERIFS(mints);
        return 2;
    }
    PatmolHipScfResult g{};
    int rc = uhf ? patmol_hip_uhf_scf(mints, nullptr, S.data(), H.data(), nocca, noccb, maxits, eps, accel,
                                      nuclear_repulsion(), nullptr, nullptr, PgA.data(), PgB.data(), nullptr, nullptr, &g)
                 : patmol_hip_rhf_scf(mints, nullptr, S.data(), H.data(), nocca, maxits, eps, accel, nuclear_repulsion(),
                                      nullptr, PgA.data(), nullptr, nullptr, &g);
    if (rc) {
        clean_ERIFS(mints);
        return 2;
    }
    if (gs) {
        gs->energy = g.energy;
        gs->eelec = g.eelec;
        gs->enuc = g.enuc;
        gs->ediff = g.ediff;
        gs->pdiff = g.pdiff;
        gs->iterations = g.iterations;
        gs->converged = g.converged;
        gs->gbuild_seconds = g.gbuild_ms * 1e-3;
        gs->diag_seconds = g.other_ms * 1e-3;
    }
    // the DPM shells become part of the basis (rhfscf.c:226-229)
    const int poscen = skip_center();
    dpm_augment();
    const int tnorb = number_of_orbitals();
    if (tnorb <= ono || poscen < 0 || poscen >= number_of_centers()) {
        clean_ERIFS(mints);
        return 3;   // no DPM centre and shells were added
    }
    const size_t tno2 = (size_t)tnorb * (tnorb + 1) / 2;
    ERIFS *xints = init_ERIFS(ono, tnorb, "XINTS.DAT");
    std::vector<double> St(tno2), Tt(tno2), Ht(tno2), QA(tno2, 0.0), QB(tno2, 0.0);
    // ground-state densities on the packed prefix of the augmented basis
    // (the DPM functions are numbered after the electron functions)
    std::copy(PgA.begin(), PgA.end(), QA.begin());
    std::copy(PgB.begin(), PgB.end(), QB.begin());
    for (int ip = 0; ip < npts; ++ip) {
        const double *rp = pts + 3 * ip;
        set_center_position(poscen, rp);
        one_electron_ints(St.data(), Tt.data(), Ht.data());
        if (!patmol_hip_last_error()) dpm_ints(Ht.data());
        if (!patmol_hip_last_error()) two_electron_ints(xints, onsh);
        if (patmol_hip_last_error()) {
            rc = 2;
            break;
        }
        PatmolHipScfResult r{};
        rc = uhf ? patmol_hip_uhf_scf(mints, xints, St.data(), Ht.data(), nocca, noccb, maxits, eps, accel,
                                      nuclear_repulsion(), QA.data(), QB.data(), nullptr, nullptr, nullptr, nullptr, &r)
                 : patmol_hip_rhf_scf(mints, xints, St.data(), Ht.data(), nocca, maxits, eps, accel, nuclear_repulsion(),
                                      QA.data(), nullptr, nullptr, nullptr, &r);
        if (rc) {
            rc = 2;
            break;
        }
        const double r2 = rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2];
        PatmolDpmPoint &o = out[ip];
        o.e_first = r.energy0;
        o.e_conv = r.energy;
        o.vpol = r.energy - r.energy0;
        o.alfa = -2.0 * o.vpol * r2 * r2;
        o.iterations = r.iterations;
        o.converged = r.converged;
    }
    clean_ERIFS(xints);
    clean_ERIFS(mints);
    return rc;
}
