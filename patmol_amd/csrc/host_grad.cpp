// host_grad.cpp -- RHF nuclear gradient: the one-electron and nuclear
// repulsion terms on the host, the two-electron term from the device
// (patmol_hip_eri_gradient).  The reference has no gradients; the terms are
// the textbook RHF ones for its energy E = 2 sum P H + sum P G(P) + E_nuc
// (rhfscf.c:49, scfaux.c:97-118) with the half density P = C_occ C_occ^T:
//   dE/dR = 2 sum P dH/dR + sum Gamma d(ab|cd)/dR - 2 sum W dS/dR + dE_nuc/dR
// with W = P F P = C_occ diag(e) C_occ^T, F = H + G(P), at a converged P.
// One-electron derivative integrals by McMurchie-Davidson Hermite
// expansion (the expansion host_oneints.cpp uses for S, T, V): the derivative
// of a primitive Cartesian Gaussian with respect to its centre is
// 2 alpha G(a + 1_x) - a_x G(a - 1_x); two-centre integrals have
// d/dB = -d/dA; for the attraction of nucleus C, d/dC = -(d/dA + d/dB)
// (Hellmann-Feynman term).  Centre skip_center() (the positron centre) is
// left out of the attraction as oneints.c:51-53 leaves it out.
#include <cmath>
#include <cstring>
#include <vector>

#include "patmol_hip.h"
#include "patmol_host.h"

namespace {

constexpr int EM = 8;          // max power per centre + 1
constexpr int EN = 2 * EM + 1;
constexpr int RM = 13;

struct Herm1D {                // E^{ij}_t for one Cartesian direction
    double e[EM][EM][EN];
    void build(int l1, int l2, double half_ip, double pa, double pb)
    {
        for (int i = 0; i <= l1; ++i)
            for (int j = 0; j <= l2; ++j)
                for (int t = 0; t <= l1 + l2 + 1; ++t) e[i][j][t] = 0.0;
        e[0][0][0] = 1.0;
        for (int j = 1; j <= l2; ++j)
            for (int t = 0; t <= j; ++t)
                e[0][j][t] = (t ? half_ip * e[0][j - 1][t - 1] : 0.0) + pb * e[0][j - 1][t] + (t + 1) * e[0][j - 1][t + 1];
        for (int i = 1; i <= l1; ++i)
            for (int j = 0; j <= l2; ++j)
                for (int t = 0; t <= i + j; ++t)
                    e[i][j][t] = (t ? half_ip * e[i - 1][j][t - 1] : 0.0) + pa * e[i - 1][j][t] + (t + 1) * e[i - 1][j][t + 1];
    }
};

void boys(int m, double t, double *f)
{
    if (t > 36.0 + 2 * m) {
        f[0] = 0.88622692545275801365 / sqrt(t);
        for (int i = 1; i <= m; ++i) f[i] = f[i - 1] * (i - 0.5) / t;
        return;
    }
    double mh = m + 0.5, term = 0.5 / mh, sum = term;
    for (int i = 1; i <= 500; ++i) {
        term *= t / (mh + i);
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    double et = exp(-t);
    f[m] = sum * et;
    for (int i = m; i--;) f[i] = (2.0 * t * f[i + 1] + et) / (2.0 * i + 1.0);
}

// R_{tuv} (t + u + v <= L) of the Hermite Coulomb integrals, scaled
void hermite_r(int L, double scale, double p, const double *pc, std::vector<double> &out)
{
    double f[RM + 1];
    const double T = p * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
    boys(L, T, f);
    std::vector<double> R((size_t)(L + 1) * RM * RM * RM, 0.0);
    auto at = [&](int n, int t, int u, int v) -> double & { return R[(((size_t)n * RM + t) * RM + u) * RM + v]; };
    double m2p = 1.0;
    for (int n = 0; n <= L; ++n) {
        at(n, 0, 0, 0) = scale * m2p * f[n];
        m2p *= -2.0 * p;
    }
    for (int n = L - 1; n >= 0; --n) {
        const int lim = L - n;
        for (int t = 0; t <= lim; ++t)
            for (int u = 0; u <= lim - t; ++u)
                for (int v = 0; v <= lim - t - u; ++v) {
                    if (t + u + v == 0) continue;
                    double val;
                    if (t)
                        val = pc[0] * at(n + 1, t - 1, u, v) + (t > 1 ? (t - 1) * at(n + 1, t - 2, u, v) : 0.0);
                    else if (u)
                        val = pc[1] * at(n + 1, t, u - 1, v) + (u > 1 ? (u - 1) * at(n + 1, t, u - 2, v) : 0.0);
                    else
                        val = pc[2] * at(n + 1, t, u, v - 1) + (v > 1 ? (v - 1) * at(n + 1, t, u, v - 2) : 0.0);
                    at(n, t, u, v) = val;
                }
    }
    out.assign((size_t)RM * RM * RM, 0.0);
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) out[((size_t)t * RM + u) * RM + v] = at(0, t, u, v);
}

// kinetic factor of one direction for exponents (l1, l2) (oneints.c:89-99)
double kin1d(const Herm1D &d, int l1, int l2, double a, double b)
{
    double t = 2.0 * a * b * d.e[l1 + 1][l2 + 1][0];
    if (l1) t -= l1 * b * d.e[l1 - 1][l2 + 1][0];
    if (l2) {
        t -= l2 * a * d.e[l1 + 1][l2 - 1][0];
        if (l1) t += 0.5 * l1 * l2 * d.e[l1 - 1][l2 - 1][0];
    }
    return t;
}

// sum_tuv E^{ix jx}_t E^{iy jy}_u E^{iz jz}_v R_tuv
double hermite_sum(const Herm1D *d, const int *i, const int *j, const std::vector<double> &R)
{
    double s = 0.0;
    for (int t = 0; t <= i[0] + j[0]; ++t)
        for (int u = 0; u <= i[1] + j[1]; ++u) {
            const double eu = d[0].e[i[0]][j[0]][t] * d[1].e[i[1]][j[1]][u];
            for (int v = 0; v <= i[2] + j[2]; ++v) s += eu * d[2].e[i[2]][j[2]][v] * R[((size_t)t * RM + u) * RM + v];
        }
    return s;
}

inline double packed_at(const double *p, int i, int j)
{
    return i >= j ? p[(size_t)i * (i + 1) / 2 + j] : p[(size_t)j * (j + 1) / 2 + i];
}

}  // namespace

// grad[3 c + x] += 2 sum P dH/dR - 2 sum W dS/dR over full matrices (P, W
// packed); H = T + V without skip_center()'s attraction.
extern "C" int patmol_host_onee_gradient(const double *P, const double *W, double *grad)
{
    const Shell *sh = shells();
    const Center *ce = centers();
    const int ns = number_of_shells(), nc = number_of_centers(), skip = skip_center();
    int ***lxyz = lvector_components();
    double **nf = normalizationFactors();
    if (!sh || !ce || !P || !W || !grad) return 1;
    const double pi32 = 5.5683279968317079;
    // one row of sums per first shell, filled in parallel and added in
    // shell order: the result does not depend on the number of threads
    std::vector<double> rows((size_t)ns * 3 * nc, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int is = ns - 1; is >= 0; --is) {
        Herm1D d[3];
        std::vector<double> R;
        double *row = &rows[(size_t)is * 3 * nc];
        const Shell &a = sh[is];
        const double *A = ce[a.cen].r;
        for (int js = 0; js <= is; ++js) {
            const Shell &b = sh[js];
            const double *B = ce[b.cen].r;
            const int L = a.lsh + b.lsh;
            const double pairw = is == js ? 1.0 : 2.0;   // both orders of an off-diagonal pair
            double ab2 = 0;
            for (int x = 0; x < 3; ++x) ab2 += (A[x] - B[x]) * (A[x] - B[x]);
            double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
            std::vector<double> gC((size_t)3 * nc, 0.0);
            for (int ip = 0; ip < a.npr; ++ip)
                for (int jp = 0; jp < b.npr; ++jp) {
                    const double al = a.al[ip], be = b.al[jp], p = al + be, c12 = a.co[ip] * b.co[jp];
                    const double s12 = pi32 * exp(-al * be * ab2 / p) / (p * sqrt(p));
                    double Pc[3];
                    for (int x = 0; x < 3; ++x) Pc[x] = (al * A[x] + be * B[x]) / p;
                    for (int x = 0; x < 3; ++x) d[x].build(a.lsh + 2, b.lsh + 2, 0.5 / p, Pc[x] - A[x], Pc[x] - B[x]);
                    const double sr = 2.0 * s12 * sqrt(p / M_PI);
                    for (int ia = 0; ia < a.nst; ++ia)
                        for (int ib = 0; ib < b.nst; ++ib) {
                            const int *la = lxyz[a.lsh][ia], *lb = lxyz[b.lsh][ib];
                            const int ir = a.off + ia, jr = b.off + ib;
                            const double f = pairw * c12 * nf[a.lsh][ia] * nf[b.lsh][ib];
                            const double pw = 2.0 * packed_at(P, ir, jr), ww = -2.0 * packed_at(W, ir, jr);
                            double E[3], K[3], dE[3], dK[3];
                            for (int x = 0; x < 3; ++x) {
                                const int i = la[x], j = lb[x];
                                E[x] = d[x].e[i][j][0];
                                K[x] = kin1d(d[x], i, j, al, be);
                                dE[x] = 2.0 * al * d[x].e[i + 1][j][0] - (i ? i * d[x].e[i - 1][j][0] : 0.0);
                                dK[x] = 2.0 * al * kin1d(d[x], i + 1, j, al, be) - (i ? i * kin1d(d[x], i - 1, j, al, be) : 0.0);
                            }
                            for (int x = 0; x < 3; ++x) {
                                const int y = (x + 1) % 3, z = (x + 2) % 3;
                                const double dS = s12 * dE[x] * E[y] * E[z];
                                const double dT = s12 * (dK[x] * E[y] * E[z] + dE[x] * (K[y] * E[z] + E[y] * K[z]));
                                const double g = f * (ww * dS + pw * dT);   // d/dA; d/dB = -d/dA
                                gA[x] += g;
                                gB[x] -= g;
                            }
                        }
                    // attraction derivatives, nucleus by nucleus
                    for (int cc = 0; cc < nc; ++cc) {
                        if (cc == skip) continue;
                        const double pc[3] = {Pc[0] - ce[cc].r[0], Pc[1] - ce[cc].r[1], Pc[2] - ce[cc].r[2]};
                        hermite_r(L + 1, -ce[cc].q * sr, p, pc, R);
                        for (int ia = 0; ia < a.nst; ++ia)
                            for (int ib = 0; ib < b.nst; ++ib) {
                                const int *la = lxyz[a.lsh][ia], *lb = lxyz[b.lsh][ib];
                                const int ir = a.off + ia, jr = b.off + ib;
                                const double f = pairw * c12 * nf[a.lsh][ia] * nf[b.lsh][ib] * 2.0 * packed_at(P, ir, jr);
                                for (int x = 0; x < 3; ++x) {
                                    int ai[3] = {la[0], la[1], la[2]}, bi[3] = {lb[0], lb[1], lb[2]};
                                    ai[x] += 1;
                                    double dA = 2.0 * al * hermite_sum(d, ai, bi, R);
                                    if (la[x]) {
                                        ai[x] -= 2;
                                        dA -= la[x] * hermite_sum(d, ai, bi, R);
                                    }
                                    ai[x] = la[x];
                                    bi[x] += 1;
                                    double dB = 2.0 * be * hermite_sum(d, ai, bi, R);
                                    if (lb[x]) {
                                        bi[x] -= 2;
                                        dB -= lb[x] * hermite_sum(d, ai, bi, R);
                                    }
                                    gA[x] += f * dA;
                                    gB[x] += f * dB;
                                    gC[3 * cc + x] -= f * (dA + dB);
                                }
                            }
                    }
                }
            for (int x = 0; x < 3; ++x) {
                row[3 * a.cen + x] += gA[x];
                row[3 * b.cen + x] += gB[x];
            }
            for (int cc = 0; cc < nc; ++cc)
                for (int x = 0; x < 3; ++x) row[3 * cc + x] += gC[3 * cc + x];
        }
    }
    for (int is = 0; is < ns; ++is)
        for (int k = 0; k < 3 * nc; ++k) grad[k] += rows[(size_t)is * 3 * nc + k];
    return 0;
}

// d E_nuc / dR: sum over pairs q_i q_j / |R_i - R_j| (basis.c:175-195)
extern "C" void patmol_nuclear_repulsion_gradient(double *grad)
{
    const Center *ce = centers();
    const int nc = number_of_centers();
    for (int i = 0; i < nc; ++i)
        for (int j = 0; j < nc; ++j) {
            if (i == j) continue;
            double r[3], r2 = 0.0;
            for (int x = 0; x < 3; ++x) {
                r[x] = ce[i].r[x] - ce[j].r[x];
                r2 += r[x] * r[x];
            }
            const double f = -ce[i].q * ce[j].q / (r2 * sqrt(r2));
            for (int x = 0; x < 3; ++x) grad[3 * i + x] += f * r[x];
        }
}

// The RHF energy gradient at the converged half density P (packed) of the
// current basis: grad[3 c + x] = dE/dR_{c,x} (assigned), every centre.
// mints: an ERIFS handle of the current basis (two_electron_ints run), or
// null to build one here.  W = P F P with F = H + G(P) from the device.
// With dpm, H carries -V of dpm_ints and the DPM term's derivative is added.
static int rhf_gradient(const ERIFS *mints, const double *P, double *grad, bool dpm)
{
    const int n = number_of_orbitals(), nc = number_of_centers();
    if (!P || !grad || n <= 0) return 1;
    const size_t no2 = (size_t)n * (n + 1) / 2;
    std::vector<double> S(no2), T(no2), H(no2), G(no2, 0.0);
    one_electron_ints(S.data(), T.data(), H.data());
    if (!patmol_hip_last_error() && dpm) dpm_ints(H.data());
    if (patmol_hip_last_error()) return 2;
    ERIFS *own = nullptr;
    if (!mints) {
        own = init_ERIFS(0, n, "MINTS.DAT");
        two_electron_ints(own, 0);
        if (patmol_hip_last_error()) {
            clean_ERIFS(own);
            return 2;
        }
        mints = own;
    }
    rhf_form_gmatrix(mints, P, G.data());
    int rc = patmol_hip_last_error() ? 2 : 0;
    // W = P F P (full, symmetric)
    std::vector<double> Pf((size_t)n * n), Ff((size_t)n * n), PF((size_t)n * n), W(no2);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            const size_t ij = (size_t)i * (i + 1) / 2 + j;
            Pf[(size_t)i * n + j] = Pf[(size_t)j * n + i] = P[ij];
            Ff[(size_t)i * n + j] = Ff[(size_t)j * n + i] = H[ij] + G[ij];
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += Pf[(size_t)i * n + k] * Ff[(size_t)j * n + k];   // F symmetric
            PF[(size_t)i * n + j] = s;
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += PF[(size_t)i * n + k] * Pf[(size_t)j * n + k];   // P symmetric
            W[(size_t)i * (i + 1) / 2 + j] = s;
        }
    for (int k = 0; k < 3 * nc; ++k) grad[k] = 0.0;
    if (!rc && patmol_hip_eri_gradient(mints, P, grad)) rc = 2;
    if (!rc && patmol_hip_onee_gradient(P, W.data(), grad)) rc = 2;
    if (!rc) patmol_nuclear_repulsion_gradient(grad);
    if (!rc && dpm && patmol_hip_dpm_gradient(P, grad)) rc = 2;
    if (own) clean_ERIFS(own);
    return rc;
}

extern "C" int patmol_rhf_gradient(const ERIFS *mints, const double *P, double *grad)
{
    return rhf_gradient(mints, P, grad, false);
}

extern "C" int patmol_rhf_dpm_gradient(const ERIFS *mints, const double *P, double *grad)
{
    return rhf_gradient(mints, P, grad, true);
}
