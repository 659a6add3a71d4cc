// dpm.hip -- distributed-positron-model potential, the drop-in dpm_ints
// (dpmints.c:120-199).
//
// The positron is a unit positive charge smeared as a fixed six-Gaussian fit
// centred on skip_center() (dpmints.c:31-36).  Every shell pair (a,b) of the
// current basis gets V_ab = sum_k c_k (ab | exp(-alpha_k |r-Q|^2)), and
// hmat[ij] -= V_ab for the packed lower triangle (dpmints.c:185-188).  Each
// fit Gaussian is a one-primitive s "pair" of exponent alpha_k at Q, so the
// three-centre integral is the (LA LB | 0 0) case of the Rys engine with
// unit overlap factor, which reproduces the prefactor of dpmints.c:82.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "eri_engine.h"
#include "patmol_hip.h"

extern "C" {
const Shell *shells(void);
const Center *centers(void);
int number_of_shells(void);
int number_of_orbitals(void);
int skip_center(void);
}

namespace pm {

// dpmints.c:31-36
__constant__ double kDpmAlpha[6] = {6.8505018000, 4.0491646000, 3.5941062989, 1.2478274000, 0.7927690989, 0.3377107978};
__constant__ double kDpmCoef[6] = {0.0766926784, 0.1483475912, 0.0462334641, 0.0717376427, 0.0447149792, 0.0069678529};

struct DpmArgs {
    const int *pI, *pJ;          // shells of each pair (first has the higher l)
    int npairs;
    const int *shell_pp0;        // unused placeholder for symmetry with ERI tables
    const double *xyz;           // per shell centre
    const PairRef *pairs;        // per listed pair
    const PrimPair *pp;
    const double *Q;             // positron position
    double *out;                 // per pair: NA*NB values
    RysTables rys;
};

template <int LA, int LB>
__global__ __launch_bounds__(64) void dpm_kernel(DpmArgs a)
{
    constexpr int NAB = ncart(LA) * ncart(LB);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.npairs) return;
    const int I = a.pI[t], J = a.pJ[t];
    PairGeom<double> bg, kg;
    for (int x = 0; x < 3; ++x) {
        bg.A[x] = a.xyz[3 * I + x];
        bg.AB[x] = a.xyz[3 * I + x] - a.xyz[3 * J + x];
        kg.A[x] = a.Q[x];
        kg.AB[x] = 0.0;
    }
    double acc[NAB];
#pragma unroll
    for (int i = 0; i < NAB; ++i) acc[i] = 0.0;
    const PairRef ref = a.pairs[t];
    for (int ib = 0; ib < ref.npp; ++ib) {
        const PrimPair bp = a.pp[ref.pp0 + ib];
        for (int k = 0; k < 6; ++k) {
            PrimPair kp;
            kp.p = kDpmAlpha[k];
            kp.P[0] = a.Q[0];
            kp.P[1] = a.Q[1];
            kp.P[2] = a.Q[2];
            kp.s = 1.0;
            kp.c = kDpmCoef[k];
            kp.ip = 1.0 / kp.p;
            primitive_quartet<double, LA, LB, 0, 0>(bp, kp, bg, kg, a.rys, 1.0,
                                                    [&](auto ia, auto ib_, auto, auto, double v) {
                                                        acc[decltype(ia)::value * ncart(LB) + decltype(ib_)::value] += v;
                                                    });
        }
    }
#pragma unroll
    for (int i = 0; i < NAB; ++i) a.out[(size_t)t * 36 + i] = acc[i];
}

// Derivative of the DPM energy term with respect to the shell centres.
// The energy E = 2 sum_ab P_ab H_ab (full sums, half density P; H carries
// -V_ab) gets -2 sum_ab P_ab V_ab from the potential.  For a shell pair
// (I, J) the block enters the full sum twice when I != J, once when I == J,
// so its part of E is w sum_{a in I, b in J} P_ab V_ab with w = -4 or -2.
// d/dA_x of a primitive Cartesian Gaussian is 2 alpha G(a + 1_x) -
// a_x G(a - 1_x): the three-centre integrals (LA+1 LB+1 | 0) of the Rys
// engine's 2-D recursion (rys.c:2269-2292) and transfer (rys.c:2295-2335),
// one root at a time.  Output per pair: w sum P dV/dA (3), w sum P dV/dB (3);
// the positron centre's share is -(dA + dB) (translational invariance).
struct DpmGradArgs {
    const int *pI, *pJ;          // shells of each pair (first has the higher l)
    const int *offI, *offJ;      // first function of each (reference order)
    int npairs;
    const double *xyz;           // per shell centre
    const PairRef *pairs;        // per listed pair
    const PrimPair *pp;
    const double2 *ppab;         // (alpha, beta) of each primitive pair
    const double *Q;             // positron position
    const double *P;             // packed half density
    double *out;                 // per pair: 6 values
    RysTables rys;
};

template <int LA, int LB>
__global__ __launch_bounds__(64) void dpm_grad_kernel(DpmGradArgs a)
{
    constexpr int NA = ncart(LA), NB = ncart(LB);
    constexpr int L12 = LA + LB + 1, NR = L12 / 2 + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.npairs) return;
    const int I = a.pI[t], J = a.pJ[t];
    double A[3], AB[3];
    for (int x = 0; x < 3; ++x) {
        A[x] = a.xyz[3 * I + x];
        AB[x] = a.xyz[3 * I + x] - a.xyz[3 * J + x];
    }
    // the density block, component normalisation folded in
    double Pab[NA * NB];
    for (int ia = 0; ia < NA; ++ia)
        for (int ib = 0; ib < NB; ++ib) {
            const int r = a.offI[t] + ia, c = a.offJ[t] + ib;
            const int hi = r > c ? r : c, lo = r > c ? c : r;
            Pab[ia * NB + ib] = a.P[(size_t)hi * (hi + 1) / 2 + lo] * cart_norm_rt<LA>(ia) * cart_norm_rt<LB>(ib);
        }
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
    const PairRef ref = a.pairs[t];
    for (int ip = 0; ip < ref.npp; ++ip) {
        const PrimPair bp = a.pp[ref.pp0 + ip];
        const double2 ab = a.ppab[ref.pp0 + ip];
        const double al2 = 2.0 * ab.x, be2 = 2.0 * ab.y, p = bp.p;
        for (int k = 0; k < 6; ++k) {
            const double q = kDpmAlpha[k], pq = p + q, iq = 1.0 / q;
            const double rpq = rsqrt_pos(pq), ipq = rpq * rpq;
            const double sr = SRTERM * (bp.s * bp.ip) * iq * rpq * bp.c * kDpmCoef[k];
            double PQ[3], PA[3];
            for (int x = 0; x < 3; ++x) {
                PQ[x] = bp.P[x] - a.Q[x];
                PA[x] = bp.P[x] - A[x];
            }
            const double T = (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]) * p * q * ipq;
#pragma unroll 1
            for (int r = 0; r < NR; ++r) {
                double xr, wr;
                rys_eval_one<NR>(T, r, xr, wr, a.rys);
                const double fff = xr * ipq, B00 = 0.5 * fff;
                const double B1 = (0.5 - B00 * q) * bp.ip, B1p = (0.5 - B00 * p) * iq;
                double tb[3][LA + 2][LB + 2];
                for (int x = 0; x < 3; ++x) {
                    double G[L12 + 1][1];
                    rys_vrr<double, L12, 0>(G, B00, B1, B1p, PA[x] - q * PQ[x] * fff, p * PQ[x] * fff,
                                            x == 2 ? wr * sr : 1.0);
                    double abp[LB + 2];
                    abp[0] = 1.0;
#pragma unroll
                    for (int i = 1; i <= LB + 1; ++i) abp[i] = abp[i - 1] * AB[x];
                    static_for<LA + 2>([&](auto ia) {
                        static_for<LB + 2>([&](auto ib_) {
                            constexpr int A_ = decltype(ia)::value, B_ = decltype(ib_)::value;
                            if constexpr (A_ + B_ <= L12) {
                                double s = 0.0;
#pragma unroll
                                for (int i = 0; i <= B_; ++i) s += binom(B_, i) * abp[i] * G[A_ + B_ - i][0];
                                tb[x][A_][B_] = s;
                            } else {
                                tb[x][A_][B_] = 0.0;
                            }
                        });
                    });
                }
                static_for<NA>([&](auto ia) {
                    static_for<NB>([&](auto ib_) {
                        constexpr int IA = decltype(ia)::value, IB = decltype(ib_)::value;
                        constexpr int e[3][2] = {{cart_x(LA, IA), cart_x(LB, IB)},
                                                 {cart_y(LA, IA), cart_y(LB, IB)},
                                                 {cart_z(LA, IA), cart_z(LB, IB)}};
                        const double gam = Pab[IA * NB + IB];
                        const double v[3] = {tb[0][e[0][0]][e[0][1]], tb[1][e[1][0]][e[1][1]], tb[2][e[2][0]][e[2][1]]};
                        static_for<3>([&](auto xx) {
                            constexpr int X = decltype(xx)::value, Y = X == 0 ? 1 : 0, Z = X == 2 ? 1 : 2;
                            constexpr int EA = e[X][0], EB = e[X][1];
                            const double o = gam * v[Y] * v[Z];
                            double da = al2 * tb[X][EA + 1][EB];
                            if constexpr (EA > 0) da -= EA * tb[X][EA - 1][EB];
                            double db = be2 * tb[X][EA][EB + 1];
                            if constexpr (EB > 0) db -= EB * tb[X][EA][EB - 1];
                            gA[X] += o * da;
                            gB[X] += o * db;
                        });
                    });
                });
            }
        }
    }
    const double w = I == J ? -2.0 : -4.0;
    for (int x = 0; x < 3; ++x) {
        a.out[(size_t)t * 6 + x] = w * gA[x];
        a.out[(size_t)t * 6 + 3 + x] = w * gB[x];
    }
}

int launch_dpm_grad(int la, int lb, const DpmGradArgs &args, hipStream_t s)
{
    dim3 g((args.npairs + 63) / 64), b(64);
    switch (la * 3 + lb) {
    case 0: hipLaunchKernelGGL((dpm_grad_kernel<0, 0>), g, b, 0, s, args); return 0;
    case 3: hipLaunchKernelGGL((dpm_grad_kernel<1, 0>), g, b, 0, s, args); return 0;
    case 4: hipLaunchKernelGGL((dpm_grad_kernel<1, 1>), g, b, 0, s, args); return 0;
    case 6: hipLaunchKernelGGL((dpm_grad_kernel<2, 0>), g, b, 0, s, args); return 0;
    case 7: hipLaunchKernelGGL((dpm_grad_kernel<2, 1>), g, b, 0, s, args); return 0;
    case 8: hipLaunchKernelGGL((dpm_grad_kernel<2, 2>), g, b, 0, s, args); return 0;
    default: return -1;
    }
}

int launch_dpm(int la, int lb, const DpmArgs &args, hipStream_t s)
{
    dim3 g((args.npairs + 63) / 64), b(64);
    switch (la * 3 + lb) {
    case 0: hipLaunchKernelGGL((dpm_kernel<0, 0>), g, b, 0, s, args); return 0;
    case 3: hipLaunchKernelGGL((dpm_kernel<1, 0>), g, b, 0, s, args); return 0;
    case 4: hipLaunchKernelGGL((dpm_kernel<1, 1>), g, b, 0, s, args); return 0;
    case 6: hipLaunchKernelGGL((dpm_kernel<2, 0>), g, b, 0, s, args); return 0;
    case 7: hipLaunchKernelGGL((dpm_kernel<2, 1>), g, b, 0, s, args); return 0;
    case 8: hipLaunchKernelGGL((dpm_kernel<2, 2>), g, b, 0, s, args); return 0;
    default: return -1;
    }
}

// defined in hip_context.cpp
int aux_fail(int code, const char *msg);
RysTables aux_tables();

}  // namespace pm

namespace pm {
namespace {

// The shell pairs of class (la, lb) of the current basis, first shell of
// each pair the one with the higher l, with their primitive pairs.
struct DpmPairs {
    std::vector<int> pI, pJ, offI, offJ;
    std::vector<PairRef> refs;
    std::vector<PrimPair> pp;
    std::vector<double2> ab;   // (alpha, beta) of each primitive pair
};

DpmPairs dpm_pairs(int la, int lb, const Shell *sh, const Center *ce, int ns)
{
    DpmPairs d;
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j <= i; ++j) {
            int a = i, b = j;
            if (sh[a].lsh < sh[b].lsh) std::swap(a, b);
            if (sh[a].lsh != la || sh[b].lsh != lb) continue;
            d.pI.push_back(a);
            d.pJ.push_back(b);
            d.offI.push_back(sh[a].off);
            d.offJ.push_back(sh[b].off);
            PairRef r{(int)d.pp.size(), sh[a].npr * sh[b].npr};
            const double *A = ce[sh[a].cen].r, *B = ce[sh[b].cen].r;
            double ab2 = 0;
            for (int x = 0; x < 3; ++x) ab2 += (A[x] - B[x]) * (A[x] - B[x]);
            for (int u = 0; u < sh[a].npr; ++u)
                for (int v = 0; v < sh[b].npr; ++v) {
                    PrimPair q;
                    double al = sh[a].al[u], be = sh[b].al[v];
                    q.p = al + be;
                    for (int x = 0; x < 3; ++x) q.P[x] = (al * A[x] + be * B[x]) / q.p;
                    q.s = exp(-al * be * ab2 / q.p);
                    q.c = sh[a].co[u] * sh[b].co[v];
                    q.ip = 1.0 / q.p;
                    d.pp.push_back(q);
                    d.ab.push_back(make_double2(al, be));
                }
            d.refs.push_back(r);
        }
    return d;
}

// device copies released at scope exit
template <typename T>
struct Dev {
    T *p = nullptr;
    ~Dev() { if (p) (void)hipFree(p); }
    bool put(const T *h, size_t n)
    {
        return hipMalloc(&p, sizeof(T) * (n ? n : 1)) == hipSuccess &&
               (n == 0 || hipMemcpy(p, h, sizeof(T) * n, hipMemcpyHostToDevice) == hipSuccess);
    }
    bool alloc(size_t n) { return hipMalloc(&p, sizeof(T) * (n ? n : 1)) == hipSuccess; }
};

bool basis_ok(const Shell *sh, const Center *ce, int ns, const char *who)
{
    if (!sh || !ce || ns <= 0) {
        aux_fail(PATMOL_HIP_EBADARG, who);
        return false;
    }
    for (int s = 0; s < ns; ++s)
        if (sh[s].lsh > 2) {
            aux_fail(PATMOL_HIP_EUNSUPPORTED, "DPM integrals: l > 2");
            return false;
        }
    return true;
}

}  // namespace
}  // namespace pm

extern "C" void dpm_ints(double *hmat)
{
    using namespace pm;
    const Shell *sh = shells();
    const Center *ce = centers();
    const int ns = number_of_shells(), sc = skip_center();
    if (!hmat || !basis_ok(sh, ce, ns, "dpm_ints: no basis or null hmat")) return;
    RysTables rys = aux_tables();
    if (!rys.cheb) return;
    const double *Q = ce[sc].r;
    std::vector<double> xyz(3 * ns);
    for (int s = 0; s < ns; ++s)
        for (int x = 0; x < 3; ++x) xyz[3 * s + x] = ce[sh[s].cen].r[x];
    for (int la = 0; la <= 2; ++la)
        for (int lb = 0; lb <= la; ++lb) {
            const DpmPairs d = dpm_pairs(la, lb, sh, ce, ns);
            if (d.pI.empty()) continue;
            const int np = (int)d.pI.size();
            Dev<int> dI, dJ;
            Dev<double> dxyz, dQ, dout;
            Dev<PairRef> dref;
            Dev<PrimPair> dpp;
            std::vector<double> out(36 * (size_t)np);
            bool ok = dI.put(d.pI.data(), np) && dJ.put(d.pJ.data(), np) && dxyz.put(xyz.data(), 3 * ns) &&
                      dQ.put(Q, 3) && dout.alloc(36 * (size_t)np) && dref.put(d.refs.data(), np) &&
                      dpp.put(d.pp.data(), d.pp.size());
            if (ok) {
                DpmArgs args{dI.p, dJ.p, np, nullptr, dxyz.p, dref.p, dpp.p, dQ.p, dout.p, rys};
                ok = launch_dpm(la, lb, args, 0) == 0 && hipGetLastError() == hipSuccess &&
                     hipMemcpy(out.data(), dout.p, sizeof(double) * out.size(), hipMemcpyDeviceToHost) == hipSuccess;
            }
            if (!ok) {
                aux_fail(PATMOL_HIP_EDEVICE, "dpm_ints: device failure");
                return;
            }
            const int nb = (lb + 1) * (lb + 2) / 2, na = (la + 1) * (la + 2) / 2;
            for (int t = 0; t < np; ++t) {
                const Shell &A = sh[d.pI[t]], &B = sh[d.pJ[t]];
                for (int ia = 0; ia < na; ++ia)
                    for (int ib = 0; ib < nb; ++ib) {
                        int r = A.off + ia, c = B.off + ib;
                        int hi = r > c ? r : c, lo = r > c ? c : r;
                        if (d.pI[t] == d.pJ[t] && c > r) continue;  // same shell: lower triangle once
                        hmat[(size_t)hi * (hi + 1) / 2 + lo] -= out[(size_t)t * 36 + ia * nb + ib];
                    }
            }
        }
}

// d/dR of the DPM term -2 sum_ab P_ab V_ab of the energy (half density P,
// packed, current basis) added to grad[3 c + x] for every centre c, the
// positron centre skip_center() included; per-pair sums added in pair order.
extern "C" int patmol_hip_dpm_gradient(const double *P, double *grad)
{
    using namespace pm;
    const Shell *sh = shells();
    const Center *ce = centers();
    const int ns = number_of_shells(), sc = skip_center(), n = number_of_orbitals();
    if (!P || !grad) return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_dpm_gradient: null pointer");
    if (!basis_ok(sh, ce, ns, "patmol_hip_dpm_gradient: no basis")) return patmol_hip_last_error();
    RysTables rys = aux_tables();
    if (!rys.cheb) return patmol_hip_last_error();
    const double *Q = ce[sc].r;
    std::vector<double> xyz(3 * ns);
    for (int s = 0; s < ns; ++s)
        for (int x = 0; x < 3; ++x) xyz[3 * s + x] = ce[sh[s].cen].r[x];
    Dev<double> dP;
    if (!dP.put(P, (size_t)n * (n + 1) / 2)) return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_dpm_gradient: upload");
    for (int la = 0; la <= 2; ++la)
        for (int lb = 0; lb <= la; ++lb) {
            const DpmPairs d = dpm_pairs(la, lb, sh, ce, ns);
            if (d.pI.empty()) continue;
            const int np = (int)d.pI.size();
            Dev<int> dI, dJ, doI, doJ;
            Dev<double> dxyz, dQ, dout;
            Dev<PairRef> dref;
            Dev<PrimPair> dpp;
            Dev<double2> dab;
            std::vector<double> out(6 * (size_t)np);
            bool ok = dI.put(d.pI.data(), np) && dJ.put(d.pJ.data(), np) && doI.put(d.offI.data(), np) &&
                      doJ.put(d.offJ.data(), np) && dxyz.put(xyz.data(), 3 * ns) && dQ.put(Q, 3) &&
                      dout.alloc(6 * (size_t)np) && dref.put(d.refs.data(), np) && dpp.put(d.pp.data(), d.pp.size()) &&
                      dab.put(d.ab.data(), d.ab.size());
            if (ok) {
                DpmGradArgs args{dI.p, dJ.p, doI.p, doJ.p, np, dxyz.p, dref.p, dpp.p, dab.p, dQ.p, dP.p, dout.p, rys};
                ok = launch_dpm_grad(la, lb, args, 0) == 0 && hipGetLastError() == hipSuccess &&
                     hipMemcpy(out.data(), dout.p, sizeof(double) * out.size(), hipMemcpyDeviceToHost) == hipSuccess;
            }
            if (!ok) return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_dpm_gradient: device failure");
            for (int t = 0; t < np; ++t) {
                const int ca = sh[d.pI[t]].cen, cb = sh[d.pJ[t]].cen;
                for (int x = 0; x < 3; ++x) {
                    const double ga = out[6 * (size_t)t + x], gb = out[6 * (size_t)t + 3 + x];
                    grad[3 * ca + x] += ga;
                    grad[3 * cb + x] += gb;
                    grad[3 * sc + x] -= ga + gb;
                }
            }
        }
    return 0;
}
