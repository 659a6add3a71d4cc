// oneints.hip -- one-electron integrals on the GPU, the drop-in
// one_electron_ints (oneints.h; oneints.c:137-234).
//
// For every shell pair I >= J of the current basis this computes the overlap
// S, the kinetic energy T and the core Hamiltonian H = T + V.  V is the
// attraction of the point nuclei of every centre except skip_center()
// (oneints.c:51-53).  The results go into caller-owned packed lower
// triangles, element (r, c) with r >= c at r(r+1)/2 + c (oneints.c:219-229).
//
// The reference expands each primitive pair in Hermite Gaussians
// (formMDD / formMDR, oneints.c:60-125).  Here:
//   * overlap and kinetic use the 1-D Obara-Saika overlaps
//     S(i,0) = PA S(i-1,0) + (i-1)/(2p) S(i-2,0), i <= la+lb+2, moved to the
//     second centre by the binomial transfer S(i,j) = sum_k C(j,k) AB^k
//     S(i+j-k,0).  These equal the reference's E^{ij}_0, so S and the kinetic
//     combination 2ab S(i+1,j+1) - i b S(i-1,j+1) - j a S(i+1,j-1)
//     + i j/2 S(i-1,j-1) of oneints.c:89-99 follow unchanged;
//   * V uses Rys quadrature with floor((la+lb)/2)+1 roots at T = p|P-C|^2.
//     This is the limit of the electron-repulsion recursion (rys.c:2269-2292)
//     for a point charge (ket exponent to infinity): B00 -> 0,
//     B1 -> (1 - t^2)/(2p), C -> PA - t^2 (P-C), prefactor 2 pi/p K_ab.  The
//     reference's sr = 2 s12 sqrt(p/pi) (oneints.c:103) is the same
//     prefactor.
// One thread per shell pair; every packed element has exactly one writer, so
// the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "eri_engine.h"
#include "eri_kernels.h"
#include "patmol_hip.h"

extern "C" {
const Shell *shells(void);
const Center *centers(void);
int number_of_shells(void);
int number_of_orbitals(void);
int number_of_centers(void);
int skip_center(void);
}

namespace pm {

constexpr double PI32 = 5.5683279968317078453;  // pi^(3/2)
constexpr double TWOPI = 6.2831853071795864769;

struct OneArgs {
    const int2 *pairs;        // shells (I, J) of each pair; I has the higher l
    const PairRef *refs;      // primitive pairs of each shell pair
    const PrimPair *pp;
    const double2 *ab;        // (alpha, beta) of each primitive pair
    const double *xyz;        // centre of each shell (3 doubles)
    const int *off;           // first orbital of each shell
    const double4 *nuc;       // point nuclei (q, x, y, z), skip_center() left out
    int npairs, nnuc;
    double *S, *T, *H;        // packed lower triangles
    double *M;                // moment integrals: 9 packed triangles
    size_t no2;
    RysTables rys;
};

// 1-D Obara-Saika overlaps of one primitive pair along one axis, moved to
// both centres: o[i][j] for i <= NI, j <= NJ.
template <int NI, int NJ>
__device__ __forceinline__ void overlap_1d(double (&o)[NI + 1][NJ + 1], double pa, double ab, double hip)
{
    constexpr int LS = NI + NJ;
    double s[LS + 1];
    s[0] = 1.0;
    if constexpr (LS >= 1) s[1] = pa;
#pragma unroll
    for (int i = 2; i <= LS; ++i) s[i] = pa * s[i - 1] + double(i - 1) * hip * s[i - 2];
    double abp[NJ + 1];
    abp[0] = 1.0;
#pragma unroll
    for (int k = 1; k <= NJ; ++k) abp[k] = abp[k - 1] * ab;
#pragma unroll
    for (int i = 0; i <= NI; ++i)
#pragma unroll
        for (int j = 0; j <= NJ; ++j) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k <= j; ++k) v += binom(j, k) * abp[k] * s[i + j - k];
            o[i][j] = v;
        }
}

template <int LA, int LB>
__global__ __launch_bounds__(64) void onee_kernel(OneArgs a)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NAB = NA * NB, L = LA + LB, NR = L / 2 + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.npairs) return;
    const int2 sh = a.pairs[t];
    double A[3], AB[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        A[x] = a.xyz[3 * sh.x + x];
        AB[x] = A[x] - a.xyz[3 * sh.y + x];
    }
    double sacc[NAB], tacc[NAB], vacc[NAB];
#pragma unroll
    for (int i = 0; i < NAB; ++i) sacc[i] = tacc[i] = vacc[i] = 0.0;

    const PairRef ref = a.refs[t];
    for (int k = 0; k < ref.npp; ++k) {
        const PrimPair q = a.pp[ref.pp0 + k];
        const double2 ex = a.ab[ref.pp0 + k];
        const double hip = 0.5 * q.ip;
        double PA[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) PA[x] = q.P[x] - A[x];

        // ---- overlap and kinetic (oneints.c:80-116) ----
        double o[3][LA + 2][LB + 2];
        overlap_1d<LA + 1, LB + 1>(o[0], PA[0], AB[0], hip);
        overlap_1d<LA + 1, LB + 1>(o[1], PA[1], AB[1], hip);
        overlap_1d<LA + 1, LB + 1>(o[2], PA[2], AB[2], hip);
        const double s12 = PI32 * q.s * q.ip * sqrt(q.ip) * q.c;
        const double al = ex.x, be = ex.y;
        auto kin = [&](const double(&ox)[LA + 2][LB + 2], int i, int j) {
            double v = 2.0 * al * be * ox[i + 1][j + 1];
            if (i) v -= double(i) * be * ox[i - 1][j + 1];
            if (j) v -= double(j) * al * ox[i + 1][j - 1];
            if (i && j) v += 0.5 * double(i * j) * ox[i - 1][j - 1];
            return v;
        };
        static_for<NA>([&](auto ia) {
            static_for<NB>([&](auto ib) {
                constexpr int IA = decltype(ia)::value, IB = decltype(ib)::value;
                constexpr int ax = cart_x(LA, IA), ay = cart_y(LA, IA), az = cart_z(LA, IA);
                constexpr int bx = cart_x(LB, IB), by = cart_y(LB, IB), bz = cart_z(LB, IB);
                const double sx = o[0][ax][bx], sy = o[1][ay][by], sz = o[2][az][bz];
                sacc[IA * NB + IB] += s12 * sx * sy * sz;
                tacc[IA * NB + IB] += s12 * (kin(o[0], ax, bx) * sy * sz + sx * kin(o[1], ay, by) * sz +
                                             sx * sy * kin(o[2], az, bz));
            });
        });

        // ---- nuclear attraction (oneints.c:100-125) ----
        const double vpre = TWOPI * q.ip * q.s * q.c;
        double abp[3][LB + 1];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            abp[x][0] = 1.0;
#pragma unroll
            for (int i = 1; i <= LB; ++i) abp[x][i] = abp[x][i - 1] * AB[x];
        }
        for (int c = 0; c < a.nnuc; ++c) {
            const double4 n = a.nuc[c];
            const double PC[3] = {q.P[0] - n.y, q.P[1] - n.z, q.P[2] - n.w};
            const double Tv = q.p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
            double rx[NR], rw[NR];
            rys_eval<NR>(Tv, rx, rw, a.rys);
            const double qv = -n.x * vpre;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double b1 = (1.0 - rx[r]) * hip;
                double G[3][L + 1];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const double C = PA[x] - rx[r] * PC[x];
                    G[x][0] = x == 2 ? rw[r] * qv : 1.0;
                    if constexpr (L >= 1) G[x][1] = C * G[x][0];
#pragma unroll
                    for (int i = 2; i <= L; ++i) G[x][i] = C * G[x][i - 1] + double(i - 1) * b1 * G[x][i - 2];
                }
                static_for<NA>([&](auto ia) {
                    static_for<NB>([&](auto ib) {
                        constexpr int IA = decltype(ia)::value, IB = decltype(ib)::value;
                        constexpr int ax = cart_x(LA, IA), ay = cart_y(LA, IA), az = cart_z(LA, IA);
                        constexpr int bx = cart_x(LB, IB), by = cart_y(LB, IB), bz = cart_z(LB, IB);
                        auto hrr = [&](const double(&g)[L + 1], const double(&p)[LB + 1], auto ai, auto bi) {
                            constexpr int AI = decltype(ai)::value, BI = decltype(bi)::value;
                            double v = 0.0;
#pragma unroll
                            for (int kk = 0; kk <= BI; ++kk) v += binom(BI, kk) * p[kk] * g[AI + BI - kk];
                            return v;
                        };
                        const double ix = hrr(G[0], abp[0], std::integral_constant<int, ax>{}, std::integral_constant<int, bx>{});
                        const double iy = hrr(G[1], abp[1], std::integral_constant<int, ay>{}, std::integral_constant<int, by>{});
                        const double iz = hrr(G[2], abp[2], std::integral_constant<int, az>{}, std::integral_constant<int, bz>{});
                        vacc[IA * NB + IB] += ix * iy * iz;
                    });
                });
            }
        }
    }

    // ---- packed store (oneints.c:219-229); normalisation basis.c:309-319 ----
    const int oI = a.off[sh.x], oJ = a.off[sh.y];
    const bool same = sh.x == sh.y;
    static_for<NA>([&](auto ia) {
        static_for<NB>([&](auto ib) {
            constexpr int IA = decltype(ia)::value, IB = decltype(ib)::value;
            constexpr double nrm = 1.0 / csqrt(cart_dfprod(LA, IA) * cart_dfprod(LB, IB));
            const int r = oI + IA, c = oJ + IB;
            if (!(same && c > r)) {
                const int hi = r > c ? r : c, lo = r > c ? c : r;
                const size_t idx = (size_t)hi * (hi + 1) / 2 + lo;
                const double tv = nrm * tacc[IA * NB + IB];
                a.S[idx] = nrm * sacc[IA * NB + IB];
                a.T[idx] = tv;
                a.H[idx] = tv + nrm * vacc[IA * NB + IB];
            }
        });
    });
}

// Dipole and quadrupole integrals about the coordinate origin (momints.c:14-92):
// <a|x|b>, <a|y|b>, <a|z|b>, then <a|xx|b>, xy, xz, yy, yz, zz.  Per axis the
// 1-D factors are the overlap S(i,j), the first moment
// M1(i,j) = S(i+1,j) + A S(i,j) and the second M2(i,j) = S(i+2,j)
// + 2A S(i+1,j) + A^2 S(i,j) (x = (x - A) + A), from the same Obara-Saika
// overlaps; the reference forms the same quantities from Hermite E
// coefficients (E_0 P + E_1 and E_0 (P^2 + 1/2p) + 2P E_1 + 2 E_2).  One
// moment component per pass over the primitive pairs keeps NA*NB
// accumulators live instead of 9 NA*NB.
// exponent of axis `ax` in moment component m (x, y, z, xx, xy, xz, yy, yz, zz)
constexpr int mom_pow(int m, int ax)
{
    constexpr int t[9][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {1, 1, 0},
                             {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}};
    return t[m][ax];
}

template <int LA, int LB>
__global__ __launch_bounds__(64) void moment_kernel(OneArgs a)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NAB = NA * NB;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.npairs) return;
    const int2 sh = a.pairs[t];
    double A[3], AB[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        A[x] = a.xyz[3 * sh.x + x];
        AB[x] = A[x] - a.xyz[3 * sh.y + x];
    }
    const PairRef ref = a.refs[t];
    const int oI = a.off[sh.x], oJ = a.off[sh.y];
    const bool same = sh.x == sh.y;
    static_for<9>([&](auto mc) {
        constexpr int M = decltype(mc)::value;
        double acc[NAB];
#pragma unroll
        for (int i = 0; i < NAB; ++i) acc[i] = 0.0;
        for (int k = 0; k < ref.npp; ++k) {
            const PrimPair q = a.pp[ref.pp0 + k];
            const double hip = 0.5 * q.ip;
            double o[3][LA + 3][LB + 1];
#pragma unroll
            for (int x = 0; x < 3; ++x) overlap_1d<LA + 2, LB>(o[x], q.P[x] - A[x], AB[x], hip);
            const double s12 = PI32 * q.s * q.ip * sqrt(q.ip) * q.c;
            auto f1 = [&](int x, int i, int j, int pwr) {
                const double *r = &o[x][0][0];
                const double s0 = r[i * (LB + 1) + j];
                if (pwr == 0) return s0;
                const double s1 = r[(i + 1) * (LB + 1) + j];
                if (pwr == 1) return s1 + A[x] * s0;
                return r[(i + 2) * (LB + 1) + j] + 2.0 * A[x] * s1 + A[x] * A[x] * s0;
            };
            static_for<NA>([&](auto ia) {
                static_for<NB>([&](auto ib) {
                    constexpr int IA = decltype(ia)::value, IB = decltype(ib)::value;
                    constexpr int ax = cart_x(LA, IA), ay = cart_y(LA, IA), az = cart_z(LA, IA);
                    constexpr int bx = cart_x(LB, IB), by = cart_y(LB, IB), bz = cart_z(LB, IB);
                    acc[IA * NB + IB] += s12 * f1(0, ax, bx, mom_pow(M, 0)) * f1(1, ay, by, mom_pow(M, 1)) *
                                         f1(2, az, bz, mom_pow(M, 2));
                });
            });
        }
        double *out = a.M + (size_t)M * a.no2;
        static_for<NA>([&](auto ia) {
            static_for<NB>([&](auto ib) {
                constexpr int IA = decltype(ia)::value, IB = decltype(ib)::value;
                constexpr double nrm = 1.0 / csqrt(cart_dfprod(LA, IA) * cart_dfprod(LB, IB));
                const int r = oI + IA, c = oJ + IB;
                if (!(same && c > r)) {
                    const int hi = r > c ? r : c, lo = r > c ? c : r;
                    out[(size_t)hi * (hi + 1) / 2 + lo] = nrm * acc[IA * NB + IB];
                }
            });
        });
    });
}

int launch_moment(int la, int lb, const OneArgs &args, hipStream_t s)
{
    dim3 g((args.npairs + 63) / 64), b(64);
    switch (la * 3 + lb) {
    case 0: hipLaunchKernelGGL((moment_kernel<0, 0>), g, b, 0, s, args); return 0;
    case 3: hipLaunchKernelGGL((moment_kernel<1, 0>), g, b, 0, s, args); return 0;
    case 4: hipLaunchKernelGGL((moment_kernel<1, 1>), g, b, 0, s, args); return 0;
    case 6: hipLaunchKernelGGL((moment_kernel<2, 0>), g, b, 0, s, args); return 0;
    case 7: hipLaunchKernelGGL((moment_kernel<2, 1>), g, b, 0, s, args); return 0;
    case 8: hipLaunchKernelGGL((moment_kernel<2, 2>), g, b, 0, s, args); return 0;
    default: return -1;
    }
}

int launch_onee(int la, int lb, const OneArgs &args, hipStream_t s)
{
    dim3 g((args.npairs + 63) / 64), b(64);
    switch (la * 3 + lb) {
    case 0: hipLaunchKernelGGL((onee_kernel<0, 0>), g, b, 0, s, args); return 0;
    case 3: hipLaunchKernelGGL((onee_kernel<1, 0>), g, b, 0, s, args); return 0;
    case 4: hipLaunchKernelGGL((onee_kernel<1, 1>), g, b, 0, s, args); return 0;
    case 6: hipLaunchKernelGGL((onee_kernel<2, 0>), g, b, 0, s, args); return 0;
    case 7: hipLaunchKernelGGL((onee_kernel<2, 1>), g, b, 0, s, args); return 0;
    case 8: hipLaunchKernelGGL((onee_kernel<2, 2>), g, b, 0, s, args); return 0;
    default: return -1;
    }
}

// The one-electron part of the RHF nuclear gradient (no counterpart in the
// reference): for the packed half density P and energy-weighted density W,
// grad += 2 sum P dH/dR - 2 sum W dS/dR with H = T + V (V without
// skip_center()'s attraction).  One thread per shell pair (both orders of an
// off-diagonal pair: weight 2).  Per axis, with the Obara-Saika overlaps
// S(i,j) and kinetic T(i,j) of the forward kernel, d/dA_x of a primitive
// Cartesian Gaussian is 2 alpha G(a+1_x) - a_x G(a-1_x), d/dB = -d/dA for S
// and T; V's derivatives come from the Rys point-charge tables one order
// higher (d/dA and d/dB from the shifted exponents, the nucleus C taking
// -(d/dA + d/dB), Hellmann-Feynman).  The per-nucleus sums are summed over
// the wave in a fixed butterfly; every sum lands in fixed point (FxAcc), so
// the gradient is reproducible.
struct OneGradArgs {
    OneArgs o;
    const double *P, *W;       // packed lower triangles
    const int *cen;            // centre of each shell
    const int *nuc_cen;        // centre of each point nucleus
    FxAcc *g;                  // 3 per centre
};

__device__ __forceinline__ void fx_add_dev(FxAcc *p, double v)
{
    if (v == 0.0) return;
    const FxAcc s = fx_split(v);
    if (s.hi != 0) __hip_atomic_fetch_add(&p->hi, s.hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (s.lo != 0) __hip_atomic_fetch_add(&p->lo, s.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int LA, int LB>
__global__ __launch_bounds__(64) void onee_grad_kernel(OneGradArgs g)
{
    const OneArgs &a = g.o;
    constexpr int NA = ncart(LA), NB = ncart(LB), L1 = LA + LB + 1, NR = L1 / 2 + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool here = t < a.npairs;
    const int2 sh = here ? a.pairs[t] : make_int2(0, 0);
    const PairRef ref = here ? a.refs[t] : PairRef{0, 0};
    double A[3], AB[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        A[x] = a.xyz[3 * sh.x + x];
        AB[x] = A[x] - a.xyz[3 * sh.y + x];
    }
    const double pairw = sh.x == sh.y ? 1.0 : 2.0;
    const int oI = a.off[sh.x], oJ = a.off[sh.y];
    // 2 P and -2 W of this pair's components, with the pair weight and the
    // component normalisation
    double pw[NA * NB], ww[NA * NB];
    static_for<NA>([&](auto ia) {
        static_for<NB>([&](auto ib) {
            constexpr int IA = decltype(ia)::value, IB = decltype(ib)::value;
            constexpr double nrm = 1.0 / csqrt(cart_dfprod(LA, IA) * cart_dfprod(LB, IB));
            const int r = oI + IA, c = oJ + IB, hi = r > c ? r : c, lo = r > c ? c : r;
            const size_t idx = (size_t)hi * (hi + 1) / 2 + lo;
            pw[IA * NB + IB] = here ? 2.0 * pairw * nrm * g.P[idx] : 0.0;
            ww[IA * NB + IB] = here ? -2.0 * pairw * nrm * g.W[idx] : 0.0;
        });
    });
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};

    // ---- overlap and kinetic ----
    for (int k = 0; k < ref.npp; ++k) {
        const PrimPair q = a.pp[ref.pp0 + k];
        const double2 ex = a.ab[ref.pp0 + k];
        const double hip = 0.5 * q.ip, al = ex.x, be = ex.y;
        double o[3][LA + 3][LB + 2];
#pragma unroll
        for (int x = 0; x < 3; ++x) overlap_1d<LA + 2, LB + 1>(o[x], q.P[x] - A[x], AB[x], hip);
        const double s12 = PI32 * q.s * q.ip * sqrt(q.ip) * q.c;
        auto kin = [&](const double(&ox)[LA + 3][LB + 2], int i, int j) {
            double v = 2.0 * al * be * ox[i + 1][j + 1];
            if (i) v -= double(i) * be * ox[i - 1][j + 1];
            if (j) v -= double(j) * al * ox[i + 1][j - 1];
            if (i && j) v += 0.5 * double(i * j) * ox[i - 1][j - 1];
            return v;
        };
        static_for<NA>([&](auto ia) {
            static_for<NB>([&](auto ib) {
                constexpr int IA = decltype(ia)::value, IB = decltype(ib)::value;
                constexpr int ea[3] = {cart_x(LA, IA), cart_y(LA, IA), cart_z(LA, IA)};
                constexpr int eb[3] = {cart_x(LB, IB), cart_y(LB, IB), cart_z(LB, IB)};
                double E[3], K[3], dE[3], dK[3];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const int i = ea[x], j = eb[x];
                    E[x] = o[x][i][j];
                    K[x] = kin(o[x], i, j);
                    dE[x] = 2.0 * al * o[x][i + 1][j] - (i ? double(i) * o[x][i - 1][j] : 0.0);
                    dK[x] = 2.0 * al * kin(o[x], i + 1, j) - (i ? double(i) * kin(o[x], i - 1, j) : 0.0);
                }
                const double wS = ww[IA * NB + IB] * s12, wT = pw[IA * NB + IB] * s12;
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const int y = (x + 1) % 3, z = (x + 2) % 3;
                    const double dS = dE[x] * E[y] * E[z];
                    const double dT = dK[x] * E[y] * E[z] + dE[x] * (K[y] * E[z] + E[y] * K[z]);
                    const double v = wS * dS + wT * dT;   // d/dA; d/dB = -d/dA
                    gA[x] += v;
                    gB[x] -= v;
                }
            });
        });
    }

    // ---- nuclear attraction, one nucleus at a time ----
    for (int c = 0; c < a.nnuc; ++c) {
        const double4 n = a.nuc[c];
        double gC[3] = {0, 0, 0};
        for (int k = 0; k < ref.npp; ++k) {
            const PrimPair q = a.pp[ref.pp0 + k];
            const double2 ex = a.ab[ref.pp0 + k];
            const double hip = 0.5 * q.ip, al = ex.x, be = ex.y;
            const double PC[3] = {q.P[0] - n.y, q.P[1] - n.z, q.P[2] - n.w};
            const double Tv = q.p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
            double rx[NR], rw[NR];
            rys_eval<NR>(Tv, rx, rw, a.rys);
            const double qv = -n.x * TWOPI * q.ip * q.s * q.c;
            double abp[3][LB + 2];
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                abp[x][0] = 1.0;
#pragma unroll
                for (int i = 1; i <= LB + 1; ++i) abp[x][i] = abp[x][i - 1] * AB[x];
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double b1 = (1.0 - rx[r]) * hip;
                double G[3][L1 + 1];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const double C = q.P[x] - A[x] - rx[r] * PC[x];
                    G[x][0] = x == 2 ? rw[r] * qv : 1.0;
                    G[x][1] = C * G[x][0];
#pragma unroll
                    for (int i = 2; i <= L1; ++i) G[x][i] = C * G[x][i - 1] + double(i - 1) * b1 * G[x][i - 2];
                }
                // transfer: I_x(i, j) = sum_k C(j, k) AB^k G[i + j - k]
                auto tr = [&](int x, int i, int j) {
                    double v = 0.0;
                    for (int kk = 0; kk <= j; ++kk) v += binom(j, kk) * abp[x][kk] * G[x][i + j - kk];
                    return v;
                };
                static_for<NA>([&](auto ia) {
                    static_for<NB>([&](auto ib) {
                        constexpr int IA = decltype(ia)::value, IB = decltype(ib)::value;
                        constexpr int ea[3] = {cart_x(LA, IA), cart_y(LA, IA), cart_z(LA, IA)};
                        constexpr int eb[3] = {cart_x(LB, IB), cart_y(LB, IB), cart_z(LB, IB)};
                        double I0[3];
#pragma unroll
                        for (int x = 0; x < 3; ++x) I0[x] = tr(x, ea[x], eb[x]);
                        const double f = pw[IA * NB + IB];
#pragma unroll
                        for (int x = 0; x < 3; ++x) {
                            const int y = (x + 1) % 3, z = (x + 2) % 3;
                            const int i = ea[x], j = eb[x];
                            double dA = 2.0 * al * tr(x, i + 1, j);
                            if (i) dA -= double(i) * tr(x, i - 1, j);
                            double dB = 2.0 * be * tr(x, i, j + 1);
                            if (j) dB -= double(j) * tr(x, i, j - 1);
                            const double o = f * I0[y] * I0[z];
                            gA[x] += o * dA;
                            gB[x] += o * dB;
                            gC[x] -= o * (dA + dB);
                        }
                    });
                });
            }
        }
        // this nucleus's share: the wave's sum, then one lane into its centre
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) gC[x] += __shfl_xor(gC[x], m, 64);
        if ((threadIdx.x & 63) == 0)
            for (int x = 0; x < 3; ++x) fx_add_dev(&g.g[3 * g.nuc_cen[c] + x], gC[x]);
    }
    if (here) {
        const int cA = g.cen[sh.x], cB = g.cen[sh.y];
        for (int x = 0; x < 3; ++x) {
            fx_add_dev(&g.g[3 * cA + x], gA[x]);
            fx_add_dev(&g.g[3 * cB + x], gB[x]);
        }
    }
}

int launch_onee_grad(int la, int lb, const OneGradArgs &args, hipStream_t s)
{
    dim3 gr((args.o.npairs + 63) / 64), b(64);
    switch (la * 3 + lb) {
    case 0: hipLaunchKernelGGL((onee_grad_kernel<0, 0>), gr, b, 0, s, args); return 0;
    case 3: hipLaunchKernelGGL((onee_grad_kernel<1, 0>), gr, b, 0, s, args); return 0;
    case 4: hipLaunchKernelGGL((onee_grad_kernel<1, 1>), gr, b, 0, s, args); return 0;
    case 6: hipLaunchKernelGGL((onee_grad_kernel<2, 0>), gr, b, 0, s, args); return 0;
    case 7: hipLaunchKernelGGL((onee_grad_kernel<2, 1>), gr, b, 0, s, args); return 0;
    case 8: hipLaunchKernelGGL((onee_grad_kernel<2, 2>), gr, b, 0, s, args); return 0;
    default: return -1;
    }
}

// defined in hip_context.cpp
int aux_fail(int code, const char *msg);
RysTables aux_tables();

namespace {

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

constexpr int kClassL[6][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};

// Class-grouped shell-pair tables of the current basis in one device block,
// followed by `nout` doubles of output space.
struct PairStage {
    char *d = nullptr;
    size_t o_pairs = 0, o_refs = 0, o_pp = 0, o_ab = 0, o_xyz = 0, o_off = 0, o_nuc = 0, o_out = 0, no2 = 0;
    int cls_first[7] = {0};
    int nnuc = 0;
    ~PairStage()
    {
        if (d) (void)hipFree(d);
    }
    double *out() const { return (double *)(d + o_out); }
    OneArgs args(int c, const RysTables &rys) const
    {
        OneArgs a{};
        a.pairs = (const int2 *)(d + o_pairs) + cls_first[c];
        a.refs = (const PairRef *)(d + o_refs) + cls_first[c];
        a.pp = (const PrimPair *)(d + o_pp);
        a.ab = (const double2 *)(d + o_ab);
        a.xyz = (const double *)(d + o_xyz);
        a.off = (const int *)(d + o_off);
        a.nuc = (const double4 *)(d + o_nuc);
        a.npairs = cls_first[c + 1] - cls_first[c];
        a.nnuc = nnuc;
        a.no2 = no2;
        a.rys = rys;
        return a;
    }
};

// Every shell pair I >= J (oneints.c:147-155), first shell of the higher l,
// grouped by class; their primitive pairs; shell centres and offsets; the
// point nuclei except skip_center() (oneints.c:51-53).
int stage_pairs(PairStage &st, size_t nout_per_no2, const char *who)
{
    const Shell *sh = shells();
    const Center *ce = centers();
    const int ns = number_of_shells(), nc = number_of_centers(), skip = skip_center(), N = number_of_orbitals();
    char msg[160];
    if (!sh || !ce || ns <= 0 || N <= 0) {
        snprintf(msg, sizeof msg, "%s: no basis", who);
        return aux_fail(PATMOL_HIP_EBADARG, msg);
    }
    for (int s = 0; s < ns; ++s)
        if (sh[s].lsh < 0 || sh[s].lsh > 2) {
            snprintf(msg, sizeof msg, "%s: l > 2", who);
            return aux_fail(PATMOL_HIP_EUNSUPPORTED, msg);
        }
    std::vector<int2> pairs;
    std::vector<PairRef> refs;
    std::vector<PrimPair> pp;
    std::vector<double2> ab;
    for (int c = 0; c < 6; ++c) {
        st.cls_first[c] = (int)pairs.size();
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j <= i; ++j) {
                int I = i, J = j;
                if (sh[I].lsh < sh[J].lsh) std::swap(I, J);
                if (sh[I].lsh != kClassL[c][0] || sh[J].lsh != kClassL[c][1]) continue;
                const double *Ar = ce[sh[I].cen].r, *Br = ce[sh[J].cen].r;
                double ab2 = 0.0;
                for (int x = 0; x < 3; ++x) ab2 += (Ar[x] - Br[x]) * (Ar[x] - Br[x]);
                refs.push_back(PairRef{(int)pp.size(), sh[I].npr * sh[J].npr});
                pairs.push_back(make_int2(I, J));
                for (int u = 0; u < sh[I].npr; ++u)
                    for (int v = 0; v < sh[J].npr; ++v) {
                        const double al = sh[I].al[u], be = sh[J].al[v];
                        PrimPair q;
                        q.p = al + be;
                        q.ip = 1.0 / q.p;
                        for (int x = 0; x < 3; ++x) q.P[x] = (al * Ar[x] + be * Br[x]) / q.p;
                        q.s = exp(-al * be * ab2 / q.p);
                        q.c = sh[I].co[u] * sh[J].co[v];
                        pp.push_back(q);
                        ab.push_back(make_double2(al, be));
                    }
            }
    }
    st.cls_first[6] = (int)pairs.size();
    std::vector<double> xyz(3 * (size_t)ns);
    std::vector<int> off(ns);
    for (int s = 0; s < ns; ++s) {
        off[s] = sh[s].off;
        for (int x = 0; x < 3; ++x) xyz[3 * s + x] = ce[sh[s].cen].r[x];
    }
    std::vector<double4> nuc;
    for (int c = 0; c < nc; ++c)
        if (c != skip) nuc.push_back(make_double4(ce[c].q, ce[c].r[0], ce[c].r[1], ce[c].r[2]));
    st.nnuc = (int)nuc.size();
    st.no2 = (size_t)N * (N + 1) / 2;
    st.o_pairs = 0;
    st.o_refs = st.o_pairs + align256(sizeof(int2) * pairs.size());
    st.o_pp = st.o_refs + align256(sizeof(PairRef) * refs.size());
    st.o_ab = st.o_pp + align256(sizeof(PrimPair) * pp.size());
    st.o_xyz = st.o_ab + align256(sizeof(double2) * ab.size());
    st.o_off = st.o_xyz + align256(sizeof(double) * xyz.size());
    st.o_nuc = st.o_off + align256(sizeof(int) * off.size());
    st.o_out = st.o_nuc + align256(sizeof(double4) * (nuc.size() + 1));
    std::vector<char> stage(st.o_out);
    memcpy(stage.data() + st.o_pairs, pairs.data(), sizeof(int2) * pairs.size());
    memcpy(stage.data() + st.o_refs, refs.data(), sizeof(PairRef) * refs.size());
    memcpy(stage.data() + st.o_pp, pp.data(), sizeof(PrimPair) * pp.size());
    memcpy(stage.data() + st.o_ab, ab.data(), sizeof(double2) * ab.size());
    memcpy(stage.data() + st.o_xyz, xyz.data(), sizeof(double) * xyz.size());
    memcpy(stage.data() + st.o_off, off.data(), sizeof(int) * off.size());
    if (!nuc.empty()) memcpy(stage.data() + st.o_nuc, nuc.data(), sizeof(double4) * nuc.size());
    const size_t bytes = st.o_out + nout_per_no2 * st.no2 * sizeof(double);
    if (hipMalloc(&st.d, bytes) != hipSuccess ||
        hipMemcpy(st.d, stage.data(), st.o_out, hipMemcpyHostToDevice) != hipSuccess) {
        snprintf(msg, sizeof msg, "%s: device allocation or upload failed", who);
        return aux_fail(PATMOL_HIP_EDEVICE, msg);
    }
    return 0;
}

}  // namespace

// Packed moment integrals of the current basis into host arrays (9 * no2).
int moment_ints_host(double *M, const char *who)
{
    RysTables rys{};
    PairStage st;
    if (stage_pairs(st, 9, who)) return PATMOL_HIP_EDEVICE;
    bool ok = true;
    for (int c = 0; ok && c < 6; ++c) {
        OneArgs a = st.args(c, rys);
        if (!a.npairs) continue;
        a.M = st.out();
        ok = launch_moment(kClassL[c][0], kClassL[c][1], a, 0) == 0 && hipGetLastError() == hipSuccess;
    }
    ok = ok && hipMemcpy(M, st.out(), sizeof(double) * 9 * st.no2, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) {
        char msg[160];
        snprintf(msg, sizeof msg, "%s: device failure", who);
        return aux_fail(PATMOL_HIP_EDEVICE, msg);
    }
    return 0;
}

}  // namespace pm

// oneints.h: void one_electron_ints(double *Smat, double *Tmat, double *Hmat)
extern "C" void one_electron_ints(double *Smat, double *Tmat, double *Hmat)
{
    using namespace pm;
    if (!Smat || !Tmat || !Hmat) {
        aux_fail(PATMOL_HIP_EBADARG, "one_electron_ints: null output");
        return;
    }
    RysTables rys = aux_tables();
    if (!rys.cheb) return;
    PairStage st;
    if (stage_pairs(st, 3, "one_electron_ints")) return;
    const size_t no2 = st.no2;
    double *dS = st.out();
    bool ok = true;
    for (int c = 0; ok && c < 6; ++c) {
        OneArgs a = st.args(c, rys);
        if (!a.npairs) continue;
        a.S = dS;
        a.T = dS + no2;
        a.H = dS + 2 * no2;
        ok = launch_onee(kClassL[c][0], kClassL[c][1], a, 0) == 0 && hipGetLastError() == hipSuccess;
    }
    ok = ok && hipMemcpy(Smat, dS, sizeof(double) * no2, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(Tmat, dS + no2, sizeof(double) * no2, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(Hmat, dS + 2 * no2, sizeof(double) * no2, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) aux_fail(PATMOL_HIP_EDEVICE, "one_electron_ints: device failure");
}

// grad[3 c + x] += 2 sum P dH/dR - 2 sum W dS/dR (packed half density P,
// packed energy-weighted density W, H = T + V without skip_center()'s
// attraction) on the GPU: onee_grad_kernel per pair class.
extern "C" int patmol_hip_onee_gradient(const double *P, const double *W, double *grad)
{
    using namespace pm;
    if (!P || !W || !grad) return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_onee_gradient: null argument");
    RysTables rys = aux_tables();
    if (!rys.cheb) return patmol_hip_last_error();
    PairStage st;
    if (stage_pairs(st, 0, "patmol_hip_onee_gradient")) return PATMOL_HIP_EDEVICE;
    const Shell *sh = shells();
    const int ns = number_of_shells(), nc = number_of_centers(), skip = skip_center();
    std::vector<int> cen(ns), ncen;
    for (int i = 0; i < ns; ++i) cen[i] = sh[i].cen;
    for (int c = 0; c < nc; ++c)
        if (c != skip) ncen.push_back(c);
    const size_t no2 = st.no2;
    char *buf = nullptr;
    const size_t oP = 0, oW = align256(sizeof(double) * no2), oC = oW + align256(sizeof(double) * no2),
                 oN = oC + align256(sizeof(int) * ns), oG = oN + align256(sizeof(int) * (ncen.size() + 1)),
                 bytes = oG + sizeof(FxAcc) * 3 * (size_t)nc;
    bool ok = hipMalloc(&buf, bytes) == hipSuccess;
    ok = ok && hipMemcpy(buf + oP, P, sizeof(double) * no2, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(buf + oW, W, sizeof(double) * no2, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(buf + oC, cen.data(), sizeof(int) * ns, hipMemcpyHostToDevice) == hipSuccess &&
         (ncen.empty() || hipMemcpy(buf + oN, ncen.data(), sizeof(int) * ncen.size(), hipMemcpyHostToDevice) == hipSuccess) &&
         hipMemset(buf + oG, 0, sizeof(FxAcc) * 3 * (size_t)nc) == hipSuccess;
    for (int c = 0; ok && c < 6; ++c) {
        OneGradArgs ga;
        ga.o = st.args(c, rys);
        if (!ga.o.npairs) continue;
        ga.P = (const double *)(buf + oP);
        ga.W = (const double *)(buf + oW);
        ga.cen = (const int *)(buf + oC);
        ga.nuc_cen = (const int *)(buf + oN);
        ga.g = (FxAcc *)(buf + oG);
        ok = launch_onee_grad(kClassL[c][0], kClassL[c][1], ga, 0) == 0 && hipGetLastError() == hipSuccess;
    }
    std::vector<FxAcc> g(3 * (size_t)nc);
    ok = ok && hipMemcpy(g.data(), buf + oG, sizeof(FxAcc) * g.size(), hipMemcpyDeviceToHost) == hipSuccess;
    if (buf) (void)hipFree(buf);
    if (!ok) return aux_fail(PATMOL_HIP_EDEVICE, "patmol_hip_onee_gradient: device failure");
    for (size_t k = 0; k < g.size(); ++k) grad[k] += fx_value(g[k].hi, g[k].lo);
    return 0;
}

// Dipole (3 * no2) and quadrupole (6 * no2: xx, xy, xz, yy, yz, zz) integrals
// of the current basis, packed lower triangles, component-major.
extern "C" int patmol_hip_moment_ints(double *dip, double *quad)
{
    using namespace pm;
    if (!dip || !quad) return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_moment_ints: null output");
    const int N = number_of_orbitals();
    if (N <= 0) return aux_fail(PATMOL_HIP_EBADARG, "patmol_hip_moment_ints: no basis");
    const size_t no2 = (size_t)N * (N + 1) / 2;
    std::vector<double> M(9 * no2);
    if (moment_ints_host(M.data(), "patmol_hip_moment_ints")) return PATMOL_HIP_EDEVICE;
    memcpy(dip, M.data(), sizeof(double) * 3 * no2);
    memcpy(quad, M.data() + 3 * no2, sizeof(double) * 6 * no2);
    return 0;
}

// momints.h: void momints() -- MOMINTS.DAT in the working directory, the
// reference's records and order (momints.c:113-180): for ish, jsh <= ish, the
// components ir = off_I + a, jr = off_J + b with jr <= ir, each as
// int ir, int jr, double dip[3], double quad[6].
extern "C" int patmol_hip_write_momints(const char *path)
{
    using namespace pm;
    const Shell *sh = shells();
    const int ns = number_of_shells(), N = number_of_orbitals();
    if (!path || !sh || ns <= 0 || N <= 0) return aux_fail(PATMOL_HIP_EBADARG, "momints: no basis or path");
    const size_t no2 = (size_t)N * (N + 1) / 2;
    std::vector<double> M(9 * no2);
    if (moment_ints_host(M.data(), "momints")) return PATMOL_HIP_EDEVICE;
    FILE *f = fopen(path, "wb");
    if (!f) return aux_fail(PATMOL_HIP_EBADARG, "momints: cannot open the output file");
    std::vector<char> rec;
    rec.reserve(no2 * 80);
    for (int is = 0; is < ns; ++is)
        for (int js = 0; js <= is; ++js)
            for (int a = 0; a < sh[is].nst; ++a) {
                const int ir = sh[is].off + a;
                for (int b = 0; b < sh[js].nst; ++b) {
                    const int jr = sh[js].off + b;
                    if (jr > ir) break;
                    const size_t ij = (size_t)ir * (ir + 1) / 2 + jr;
                    const size_t at = rec.size();
                    rec.resize(at + 2 * sizeof(int) + 9 * sizeof(double));
                    memcpy(&rec[at], &ir, sizeof(int));
                    memcpy(&rec[at + sizeof(int)], &jr, sizeof(int));
                    for (int m = 0; m < 9; ++m)
                        memcpy(&rec[at + 2 * sizeof(int) + m * sizeof(double)], &M[m * no2 + ij], sizeof(double));
                }
            }
    const bool ok = fwrite(rec.data(), 1, rec.size(), f) == rec.size();
    if (fclose(f) != 0 || !ok) return aux_fail(PATMOL_HIP_EDEVICE, "momints: write failed");
    return 0;
}

extern "C" void momints(void) { (void)patmol_hip_write_momints("MOMINTS.DAT"); }
