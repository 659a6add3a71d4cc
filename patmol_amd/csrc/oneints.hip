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
#include <cstring>
#include <vector>

#include "eri_engine.h"
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

// defined in hip_context.cpp
int aux_fail(int code, const char *msg);
RysTables aux_tables();

namespace {

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

}  // namespace pm

// oneints.h: void one_electron_ints(double *Smat, double *Tmat, double *Hmat)
extern "C" void one_electron_ints(double *Smat, double *Tmat, double *Hmat)
{
    using namespace pm;
    const Shell *sh = shells();
    const Center *ce = centers();
    const int ns = number_of_shells(), nc = number_of_centers(), skip = skip_center(), N = number_of_orbitals();
    if (!sh || !ce || ns <= 0 || N <= 0 || !Smat || !Tmat || !Hmat) {
        aux_fail(PATMOL_HIP_EBADARG, "one_electron_ints: no basis or null output");
        return;
    }
    RysTables rys = aux_tables();
    if (!rys.cheb) return;
    for (int s = 0; s < ns; ++s)
        if (sh[s].lsh < 0 || sh[s].lsh > 2) {
            aux_fail(PATMOL_HIP_EUNSUPPORTED, "one_electron_ints: l > 2");
            return;
        }

    // Host staging: pairs grouped by class (la >= lb), their primitive pairs,
    // shell centres and offsets, and the nuclei.
    std::vector<int2> pairs;
    std::vector<PairRef> refs;
    std::vector<PrimPair> pp;
    std::vector<double2> ab;
    int cls_first[7] = {0};
    const int cls_l[6][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
    for (int c = 0; c < 6; ++c) {
        cls_first[c] = (int)pairs.size();
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j <= i; ++j) {
                int I = i, J = j;
                if (sh[I].lsh < sh[J].lsh) std::swap(I, J);
                if (sh[I].lsh != cls_l[c][0] || sh[J].lsh != cls_l[c][1]) continue;
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
    cls_first[6] = (int)pairs.size();
    std::vector<double> xyz(3 * (size_t)ns);
    std::vector<int> off(ns);
    for (int s = 0; s < ns; ++s) {
        off[s] = sh[s].off;
        for (int x = 0; x < 3; ++x) xyz[3 * s + x] = ce[sh[s].cen].r[x];
    }
    std::vector<double4> nuc;
    for (int c = 0; c < nc; ++c)
        if (c != skip) nuc.push_back(make_double4(ce[c].q, ce[c].r[0], ce[c].r[1], ce[c].r[2]));

    // One device block: inputs, then S | T | H.
    const size_t no2 = (size_t)N * (N + 1) / 2;
    const size_t o_pairs = 0, o_refs = o_pairs + align256(sizeof(int2) * pairs.size());
    const size_t o_pp = o_refs + align256(sizeof(PairRef) * refs.size());
    const size_t o_ab = o_pp + align256(sizeof(PrimPair) * pp.size());
    const size_t o_xyz = o_ab + align256(sizeof(double2) * ab.size());
    const size_t o_off = o_xyz + align256(sizeof(double) * xyz.size());
    const size_t o_nuc = o_off + align256(sizeof(int) * off.size());
    const size_t o_out = o_nuc + align256(sizeof(double4) * (nuc.size() + 1));
    const size_t bytes = o_out + 3 * no2 * sizeof(double);
    std::vector<char> stage(o_out);
    memcpy(stage.data() + o_pairs, pairs.data(), sizeof(int2) * pairs.size());
    memcpy(stage.data() + o_refs, refs.data(), sizeof(PairRef) * refs.size());
    memcpy(stage.data() + o_pp, pp.data(), sizeof(PrimPair) * pp.size());
    memcpy(stage.data() + o_ab, ab.data(), sizeof(double2) * ab.size());
    memcpy(stage.data() + o_xyz, xyz.data(), sizeof(double) * xyz.size());
    memcpy(stage.data() + o_off, off.data(), sizeof(int) * off.size());
    if (!nuc.empty()) memcpy(stage.data() + o_nuc, nuc.data(), sizeof(double4) * nuc.size());

    char *d = nullptr;
    bool ok = hipMalloc(&d, bytes) == hipSuccess;
    ok = ok && hipMemcpy(d, stage.data(), o_out, hipMemcpyHostToDevice) == hipSuccess;
    double *dS = ok ? (double *)(d + o_out) : nullptr;
    for (int c = 0; ok && c < 6; ++c) {
        const int n = cls_first[c + 1] - cls_first[c];
        if (!n) continue;
        OneArgs args;
        args.pairs = (const int2 *)(d + o_pairs) + cls_first[c];
        args.refs = (const PairRef *)(d + o_refs) + cls_first[c];
        args.pp = (const PrimPair *)(d + o_pp);
        args.ab = (const double2 *)(d + o_ab);
        args.xyz = (const double *)(d + o_xyz);
        args.off = (const int *)(d + o_off);
        args.nuc = (const double4 *)(d + o_nuc);
        args.npairs = n;
        args.nnuc = (int)nuc.size();
        args.S = dS;
        args.T = dS + no2;
        args.H = dS + 2 * no2;
        args.rys = rys;
        ok = launch_onee(cls_l[c][0], cls_l[c][1], args, 0) == 0 && hipGetLastError() == hipSuccess;
    }
    ok = ok && hipMemcpy(Smat, dS, sizeof(double) * no2, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(Tmat, dS + no2, sizeof(double) * no2, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(Hmat, dS + 2 * no2, sizeof(double) * no2, hipMemcpyDeviceToHost) == hipSuccess;
    if (d) (void)hipFree(d);
    if (!ok) aux_fail(PATMOL_HIP_EDEVICE, "one_electron_ints: device failure");
}
