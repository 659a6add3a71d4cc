// host_oneints.cpp -- CPU one-electron integrals (oneints.c:12-234):
// overlap, kinetic and nuclear attraction over contracted Cartesian shells
// by McMurchie-Davidson Hermite expansion.  The SCF drivers use the device
// one_electron_ints of libpatmol_hip.so; this CPU version serves callers
// without a GPU.
// Centre skip_center() (the positron centre) is left out of the nuclear
// attraction, as oneints.c:51-53 does.  O(N^2) host work.
#include <cmath>
#include <cstring>
#include <vector>

#include "patmol_host.h"

namespace {

constexpr int EM = 8;          // max power per centre + 1 (kinetic needs l+1)
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

// Boys F_0..F_m by series and downward recursion, asymptotic for large T
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

// R_{tuv} = scale * R^0_{tuv}(p, PC) for t+u+v <= L (Hermite Coulomb)
void hermite_r(int L, double scale, double p, const double *pc, std::vector<double> &out)
{
    double f[RM + 1];
    double T = p * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
    boys(L, T, f);
    std::vector<double> R((size_t)(L + 1) * RM * RM * RM, 0.0);
    auto at = [&](int n, int t, int u, int v) -> double & { return R[(((size_t)n * RM + t) * RM + u) * RM + v]; };
    double m2p = 1.0;
    for (int n = 0; n <= L; ++n) {
        at(n, 0, 0, 0) = scale * m2p * f[n];
        m2p *= -2.0 * p;
    }
    for (int n = L - 1; n >= 0; --n) {
        int lim = L - n;
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

// kinetic factor of one direction, oneints.c:89-99
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

}  // namespace

extern "C" void patmol_host_one_electron_ints(double *S, double *T, double *H)
{
    const Shell *sh = shells();
    const Center *ce = centers();
    const int ns = number_of_shells(), nc = number_of_centers(), skip = skip_center();
    int ***lxyz = lvector_components();
    double **nf = normalizationFactors();
    const double pi32 = 5.5683279968317079;
    Herm1D dx, dy, dz;
    std::vector<double> rsum((size_t)RM * RM * RM), rterm;
    for (int is = 0; is < ns; ++is) {
        const Shell &a = sh[is];
        const double *A = ce[a.cen].r;
        for (int js = 0; js <= is; ++js) {
            const Shell &b = sh[js];
            const double *Bc = ce[b.cen].r;
            const int L = a.lsh + b.lsh;
            double ab2 = 0;
            for (int x = 0; x < 3; ++x) ab2 += (A[x] - Bc[x]) * (A[x] - Bc[x]);
            std::vector<double> sv(a.nst * b.nst, 0.0), tv(a.nst * b.nst, 0.0), vv(a.nst * b.nst, 0.0);
            for (int ip = 0; ip < a.npr; ++ip)
                for (int jp = 0; jp < b.npr; ++jp) {
                    double al = a.al[ip], be = b.al[jp], p = al + be, c12 = a.co[ip] * b.co[jp];
                    double s12 = pi32 * exp(-al * be * ab2 / p) / (p * sqrt(p));
                    double P[3];
                    for (int x = 0; x < 3; ++x) P[x] = (al * A[x] + be * Bc[x]) / p;
                    dx.build(a.lsh + 1, b.lsh + 1, 0.5 / p, P[0] - A[0], P[0] - Bc[0]);
                    dy.build(a.lsh + 1, b.lsh + 1, 0.5 / p, P[1] - A[1], P[1] - Bc[1]);
                    dz.build(a.lsh + 1, b.lsh + 1, 0.5 / p, P[2] - A[2], P[2] - Bc[2]);
                    std::fill(rsum.begin(), rsum.end(), 0.0);
                    double sr = 2.0 * s12 * sqrt(p / M_PI);
                    for (int c = 0; c < nc; ++c) {
                        if (c == skip) continue;
                        double pc[3] = {P[0] - ce[c].r[0], P[1] - ce[c].r[1], P[2] - ce[c].r[2]};
                        hermite_r(L, sr, p, pc, rterm);
                        for (size_t k = 0; k < rsum.size(); ++k) rsum[k] -= ce[c].q * rterm[k];
                    }
                    for (int ia = 0; ia < a.nst; ++ia)
                        for (int ib = 0; ib < b.nst; ++ib) {
                            const int *la = lxyz[a.lsh][ia], *lb = lxyz[b.lsh][ib];
                            double f = c12 * nf[a.lsh][ia] * nf[b.lsh][ib];
                            double ex = dx.e[la[0]][lb[0]][0], ey = dy.e[la[1]][lb[1]][0], ez = dz.e[la[2]][lb[2]][0];
                            int k = ia * b.nst + ib;
                            sv[k] += f * s12 * ex * ey * ez;
                            tv[k] += f * s12 *
                                     (kin1d(dx, la[0], lb[0], al, be) * ey * ez + kin1d(dy, la[1], lb[1], al, be) * ex * ez +
                                      kin1d(dz, la[2], lb[2], al, be) * ex * ey);
                            double sum = 0.0;
                            for (int t = 0; t <= la[0] + lb[0]; ++t)
                                for (int u = 0; u <= la[1] + lb[1]; ++u)
                                    for (int v = 0; v <= la[2] + lb[2]; ++v)
                                        sum += dx.e[la[0]][lb[0]][t] * dy.e[la[1]][lb[1]][u] * dz.e[la[2]][lb[2]][v] *
                                               rsum[((size_t)t * RM + u) * RM + v];
                            vv[k] += f * sum;
                        }
                }
            for (int ia = 0; ia < a.nst; ++ia) {
                int ir = a.off + ia;
                for (int ib = 0; ib < b.nst; ++ib) {
                    int jr = b.off + ib;
                    if (jr > ir) break;
                    size_t ij = (size_t)ir * (ir + 1) / 2 + jr;
                    int k = ia * b.nst + ib;
                    S[ij] = sv[k];
                    T[ij] = tv[k];
                    H[ij] = tv[k] + vv[k];
                }
            }
        }
    }
}
