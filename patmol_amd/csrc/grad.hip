// grad.hip -- the two-electron part of the RHF nuclear gradient on the GPU.
//
// dE_2e/dR = sum_{abcd} Gamma_abcd d(ab|cd)/dR with, for the half density P
// of the reference (E = 2 sum P H + sum P G(P), rhfscf.c:49, scfaux.c:97-118),
//   Gamma_abcd = 2 P_ab P_cd - (P_ac P_bd + P_ad P_bc) / 2,
// which has the 8-fold symmetry of the integrals, so the sum runs over the
// canonical shell quartets of the screened work list (the same tile
// quartets as the J/K build) with the factor 8 w, w = 2^-([I=J]+[K=L]+[IJ=KL]).
//
// Derivative integrals: d/dA_x of a primitive Cartesian Gaussian is
// 2 alpha G(a + 1_x) - a_x G(a - 1_x), so the derivatives with respect to
// the centres A, B, C come from the Rys 2-D recursion (rys.c:2269-2292) one
// order higher on each side and the transfer (rys.c:2295-2335) with the
// shifted exponents; the centre D follows from translational invariance,
// d/dD = -(d/dA + d/dB + d/dC).  Roots are digested one at a time (the
// contraction with Gamma is linear), the per-root 1-D integrals of all
// shifted exponents held in a per-lane table.
//
// One workgroup per tile quartet; its threads take the item's shell quartets
// in turn.  Per-shell gradient sums go through LDS and global memory as
// fixed-point integers (FxAcc), so the gradient is reproducible bit for bit.
#include <hip/hip_runtime.h>

#include "eri_kernels.h"
#include "grad.h"

namespace pm {

namespace {

__device__ __forceinline__ void fx_lds_add(FxAcc *p, double v)
{
    if (v == 0.0) return;
    const FxAcc s = fx_split(v);
    __hip_atomic_fetch_add(&p->hi, s.hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&p->lo, s.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void fx_global_add(FxAcc *p, FxAcc v)
{
    // the low word brought into [-2^39, 2^39) first, as the J/K flush does
    const long long carry = (v.lo + (1LL << 39)) >> 40;
    const long long lo = v.lo - carry * (1LL << 40), hi = v.hi + carry;
    if (hi != 0) __hip_atomic_fetch_add(&p->hi, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lo != 0) __hip_atomic_fetch_add(&p->lo, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Cartesian exponents of component i of shell l (reference order,
// basis.c:294-308), packed (lx | ly << 2 | lz << 4)
template <int L>
__device__ __forceinline__ int cart_exps(int i)
{
    int v = 0;
    static_for<ncart(L)>([&](auto k) {
        constexpr int K = decltype(k)::value;
        if (i == K) v = cart_x(L, K) | (cart_y(L, K) << 2) | (cart_z(L, K) << 4);
    });
    return v;
}

}  // namespace

// 1-D integrals of one root and one axis with every exponent shift the
// derivatives need: t[ax][bx][cx][dx], ax <= LA+1, bx <= LB+1, cx <= LC+1,
// dx <= LD (entries with ax + bx > LA + LB + 1 or cx + dx > LC + LD + 1 are
// never read)
template <typename R, int LA, int LB, int LC, int LD>
struct GradTab {
    static constexpr int SA = (LB + 2) * (LC + 2) * (LD + 1), SB = (LC + 2) * (LD + 1), SC = LD + 1;
    static constexpr int NX = (LA + 2) * SA;
    R t[3][NX];
    __device__ __forceinline__ R at(int axis, int a, int b, int c, int d) const
    {
        return t[axis][a * SA + b * SB + c * SC + d];
    }
};

template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(256) void grad2e_kernel(GradArgs a)
{
    constexpr int TI = tile_of(LA), TJ = tile_of(LB), TK = tile_of(LC), TL = tile_of(LD);
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int L12 = LA + LB + 1, L34 = LC + LD + 1;     // recursion orders with the shift
    constexpr int NR = (LA + LB + LC + LD + 1) / 2 + 1;       // roots of the derivative integrals
    using Tab = GradTab<R, LA, LB, LC, LD>;
    __shared__ FxAcc sg[(TI + TJ + TK + TL) * 3];

    const int4 it = a.items[blockIdx.x];
    const int tid = threadIdx.x;
    const int fI = a.tile_first[it.x], nI = a.tile_size[it.x];
    const int fJ = a.tile_first[it.y], nJ = a.tile_size[it.y];
    const int fK = a.tile_first[it.z], nK = a.tile_size[it.z];
    const int fL = a.tile_first[it.w], nL = a.tile_size[it.w];
    const bool sameBra = it.x == it.y, sameKet = it.z == it.w, samePair = it.x == it.z && it.y == it.w;
    const int N = a.N;
    for (int e = tid; e < (TI + TJ + TK + TL) * 3; e += 256) sg[e] = FxAcc{0, 0};
    __syncthreads();

    for (int q = tid; q < TI * TJ * TK * TL; q += 256) {
        const int lq = q % TL, kq = (q / TL) % TK, jq = (q / (TL * TK)) % TJ, iq = q / (TL * TK * TJ);
        if (iq >= nI || jq >= nJ || kq >= nK || lq >= nL) continue;
        const int I = fI + iq, J = fJ + jq, K = fK + kq, L = fL + lq;
        if (sameBra && J > I) continue;
        if (sameKet && L > K) continue;
        if (samePair && !(I > K || (I == K && J >= L))) continue;
        const int pb = I * (I + 1) / 2 + J, pk = K * (K + 1) / 2 + L;
        if (a.pair_q[pb] * a.pair_q[pk] < a.tau) continue;
        double w = 8.0;
        if (I == J) w *= 0.5;
        if (K == L) w *= 0.5;
        if (I == K && J == L) w *= 0.5;
        const PairRef bref = a.pairs[pb], kref = a.pairs[pk];
        double xyz[4][3];
        for (int x = 0; x < 3; ++x) {
            xyz[0][x] = a.shell_xyz[3 * I + x];
            xyz[1][x] = a.shell_xyz[3 * J + x];
            xyz[2][x] = a.shell_xyz[3 * K + x];
            xyz[3][x] = a.shell_xyz[3 * L + x];
        }
        const int oI = a.shell_off[I], oJ = a.shell_off[J], oK = a.shell_off[K], oL = a.shell_off[L];
        double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
        for (int ib = 0; ib < bref.npp; ++ib) {
            const PrimPair bp = a.pp[bref.pp0 + ib];
            const double2 bab = a.ppab[bref.pp0 + ib];
            for (int ik = 0; ik < kref.npp; ++ik) {
                const PrimPair kp = a.pp[kref.pp0 + ik];
                const double2 kab = a.ppab[kref.pp0 + ik];
                const double p = bp.p, qq = kp.p, pq = p + qq;
                const double rpq = rsqrt_pos(pq), ipq = rpq * rpq;
                const double sr0 = SRTERM * (bp.s * bp.ip) * (kp.s * kp.ip) * rpq;
                if (sr0 < PRIM_SCREEN) continue;   // twoints.c:110-111
                const double sr = sr0 * bp.c * kp.c * w;
                double PQ[3], PA[3], QC[3], AB[3], CD[3];
                for (int x = 0; x < 3; ++x) {
                    PQ[x] = bp.P[x] - kp.P[x];
                    PA[x] = bp.P[x] - xyz[0][x];
                    QC[x] = kp.P[x] - xyz[2][x];
                    AB[x] = xyz[0][x] - xyz[1][x];
                    CD[x] = xyz[2][x] - xyz[3][x];
                }
                const double T = (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]) * p * qq * ipq;
                const double al2 = 2.0 * bab.x, be2 = 2.0 * bab.y, ga2 = 2.0 * kab.x;
#pragma unroll 1
                for (int r = 0; r < NR; ++r) {
                    double xr, wr;
                    rys_eval_one<NR>(T, r, xr, wr, a.rys);
                    const double fff = xr * ipq, B00 = 0.5 * fff;
                    const double B1 = (0.5 - B00 * qq) * bp.ip, B1p = (0.5 - B00 * p) * kp.ip;
                    Tab tb;
                    for (int x = 0; x < 3; ++x) {
                        double G[L12 + 1][L34 + 1];
                        rys_vrr<double, L12, L34>(G, B00, B1, B1p, PA[x] - qq * PQ[x] * fff, QC[x] + p * PQ[x] * fff,
                                                  x == 2 ? wr * sr : 1.0);
                        double abp[LB + 2], cdp[LD + 1];
                        abp[0] = 1.0;
#pragma unroll
                        for (int i = 1; i <= LB + 1; ++i) abp[i] = abp[i - 1] * AB[x];
                        cdp[0] = 1.0;
#pragma unroll
                        for (int i = 1; i <= LD; ++i) cdp[i] = cdp[i - 1] * CD[x];
                        static_for<LA + 2>([&](auto ia) {
                            static_for<LB + 2>([&](auto ib_) {
                                static_for<LC + 2>([&](auto ic) {
                                    static_for<LD + 1>([&](auto id) {
                                        constexpr int A_ = decltype(ia)::value, B_ = decltype(ib_)::value;
                                        constexpr int C_ = decltype(ic)::value, D_ = decltype(id)::value;
                                        if constexpr (A_ + B_ <= L12 && C_ + D_ <= L34) {
                                            double s = 0.0;
#pragma unroll
                                            for (int i = 0; i <= B_; ++i) {
                                                double t = 0.0;
#pragma unroll
                                                for (int j = 0; j <= D_; ++j)
                                                    t += binom(D_, j) * cdp[j] * G[A_ + B_ - i][C_ + D_ - j];
                                                s += binom(B_, i) * abp[i] * t;
                                            }
                                            tb.t[x][A_ * Tab::SA + B_ * Tab::SB + C_ * Tab::SC + D_] = s;
                                        }
                                    });
                                });
                            });
                        });
                    }
                    // contraction of this root's derivative integrals with Gamma
                    for (int ia = 0; ia < NA; ++ia) {
                        const int ea = cart_exps<LA>(ia);
                        const int ax = ea & 3, ay = (ea >> 2) & 3, az = ea >> 4;
                        const double na = cart_norm_rt<LA>(ia);
                        for (int ib2 = 0; ib2 < NB; ++ib2) {
                            const int eb = cart_exps<LB>(ib2);
                            const int bx = eb & 3, by = (eb >> 2) & 3, bz = eb >> 4;
                            const double nab = na * cart_norm_rt<LB>(ib2);
                            const double Pab = a.D[(size_t)(oI + ia) * N + oJ + ib2];
                            for (int ic = 0; ic < NC; ++ic) {
                                const int ec = cart_exps<LC>(ic);
                                const int cx = ec & 3, cy = (ec >> 2) & 3, cz = ec >> 4;
                                const double nabc = nab * cart_norm_rt<LC>(ic);
                                const double Pac = a.D[(size_t)(oI + ia) * N + oK + ic];
                                const double Pbc = a.D[(size_t)(oJ + ib2) * N + oK + ic];
                                for (int id = 0; id < ND; ++id) {
                                    const int ed = cart_exps<LD>(id);
                                    const int dx = ed & 3, dy = (ed >> 2) & 3, dz = ed >> 4;
                                    const double Pcd = a.D[(size_t)(oK + ic) * N + oL + id];
                                    const double Pad = a.D[(size_t)(oI + ia) * N + oL + id];
                                    const double Pbd = a.D[(size_t)(oJ + ib2) * N + oL + id];
                                    const double gam = (2.0 * Pab * Pcd - 0.5 * (Pac * Pbd + Pad * Pbc)) * nabc *
                                                       cart_norm_rt<LD>(id);
                                    const int e[3][4] = {{ax, bx, cx, dx}, {ay, by, cy, dy}, {az, bz, cz, dz}};
                                    double v[3];
                                    for (int x = 0; x < 3; ++x) v[x] = tb.at(x, e[x][0], e[x][1], e[x][2], e[x][3]);
                                    for (int x = 0; x < 3; ++x) {
                                        const int y = x == 0 ? 1 : 0, z = x == 2 ? 1 : 2;
                                        const double o = gam * v[y] * v[z];
                                        const int ea_ = e[x][0], eb_ = e[x][1], ec_ = e[x][2], ed_ = e[x][3];
                                        double da = al2 * tb.at(x, ea_ + 1, eb_, ec_, ed_);
                                        if (ea_) da -= ea_ * tb.at(x, ea_ - 1, eb_, ec_, ed_);
                                        double db = be2 * tb.at(x, ea_, eb_ + 1, ec_, ed_);
                                        if (eb_) db -= eb_ * tb.at(x, ea_, eb_ - 1, ec_, ed_);
                                        double dc = ga2 * tb.at(x, ea_, eb_, ec_ + 1, ed_);
                                        if (ec_) dc -= ec_ * tb.at(x, ea_, eb_, ec_ - 1, ed_);
                                        gA[x] += o * da;
                                        gB[x] += o * db;
                                        gC[x] += o * dc;
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
        // per-shell sums: I, J, K in tile order, L by translational invariance
        for (int x = 0; x < 3; ++x) {
            fx_lds_add(&sg[3 * iq + x], gA[x]);
            fx_lds_add(&sg[3 * (TI + jq) + x], gB[x]);
            fx_lds_add(&sg[3 * (TI + TJ + kq) + x], gC[x]);
            fx_lds_add(&sg[3 * (TI + TJ + TK + lq) + x], -(gA[x] + gB[x] + gC[x]));
        }
    }
    __syncthreads();
    // flush: one thread per (tile slot, axis) into the per-shell array
    for (int e = tid; e < (TI + TJ + TK + TL) * 3; e += 256) {
        const int slot = e / 3, x = e % 3;
        int sh;
        if (slot < TI) sh = slot < nI ? fI + slot : -1;
        else if (slot < TI + TJ) sh = slot - TI < nJ ? fJ + slot - TI : -1;
        else if (slot < TI + TJ + TK) sh = slot - TI - TJ < nK ? fK + slot - TI - TJ : -1;
        else sh = slot - TI - TJ - TK < nL ? fL + slot - TI - TJ - TK : -1;
        const FxAcc v = sg[e];
        if (sh >= 0 && (v.hi != 0 || v.lo != 0)) fx_global_add(&a.gshell[3 * sh + x], v);
    }
}

// The same sum with one group of GL lanes (64, 32 or 16: one, two or four
// groups per wave) per shell quartet, for the classes with at least
// GRAD_GROUP_MIN Cartesian products, (ds|pp) and (pp|pp) (the others run the
// per-thread kernel):
// lane g of a group owns the products c = g + GL k (a-major), whose Gamma_c
// it writes into the group's LDS row once per quartet (six density values
// each); per primitive quartet and root, three lanes run the 2-D recursion
// of one axis each into LDS, the group's lanes fill the shifted 1-D table
// (transfer) in LDS, and each lane adds its products' derivative terms; at
// the end of the quartet the group sums its lanes in a fixed butterfly and
// its first lane adds the centre sums to the item's fixed-point shell sums.
constexpr int GRAD_GROUP_MIN = 108;
// lanes per group, measured per class at (H2O)16: 32 (two quartets per
// wave) for (dp|ds) (102 -> 94 ms) and (dp|pp) (43 -> 38 ms), a whole wave
// for the others ((dd|ps) 59 ms with 64 lanes, 65 with 32); 16-lane groups
// for the classes of 27-81 products measured slower than the per-thread
// kernel ((ds|ds) 71 against 16 ms, (dd|ss) 35, (pp|ps) 49; (dp|ps) 72 with
// 16 lanes and 87 with 32 against 70) except for (ds|pp) and (pp|pp)
template <int LA, int LB, int LC, int LD>
constexpr int grad_group_lanes()
{
    if (LA == 2 && LB == 0 && LC == 1 && LD == 1) return 16;   // (ds|pp): 62 against 70 ms per-thread
    if (LA == 1 && LB == 1 && LC == 1 && LD == 1) return 32;   // (pp|pp): 10.3 against 13.1 ms per-thread, 15.4 with 64
    return (LA == 2 && LB == 1 && ((LC == 2 && LD == 0) || (LC == 1 && LD == 1))) ? 32 : 64;
}
// the classes that run the group kernel
template <int LA, int LB, int LC, int LD>
constexpr bool grad_grouped()
{
    return ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD) >= GRAD_GROUP_MIN || (LA == 2 && LB == 0 && LC == 1 && LD == 1) ||
           (LA == 1 && LB == 1 && LC == 1 && LD == 1);
}

template <int LA, int LB, int LC, int LD>
struct GradWave {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NCOMP = NA * NB * NC * ND, CPL = (NCOMP + 63) / 64;
    static constexpr int L12 = LA + LB + 1, L34 = LC + LD + 1;
    static constexpr int NR = (LA + LB + LC + LD + 1) / 2 + 1;
    static constexpr int NG = (L12 + 1) * (L34 + 1);
    static constexpr int SA = (LB + 2) * (LC + 2) * (LD + 1), SB = (LC + 2) * (LD + 1), SC = LD + 1;
    static constexpr int NX = (LA + 2) * SA;
    static_assert(NX < 256, "8-bit table offsets");
};

// C(n, k) for n <= 3
__device__ __forceinline__ double binom3(int n, int k)
{
    return (n == 3 && (k == 1 || k == 2)) ? 3.0 : (n == 2 && k == 1) ? 2.0 : 1.0;
}

__device__ __forceinline__ double pick3(const double (&v)[3], int i) { return i == 0 ? v[0] : i == 1 ? v[1] : v[2]; }

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void grad2e_wave_kernel(GradArgs a)
{
    using W = GradWave<LA, LB, LC, LD>;
    constexpr int TI = tile_of(LA), TJ = tile_of(LB), TK = tile_of(LC), TL = tile_of(LD);
    constexpr int NB = W::NB, NC = W::NC, ND = W::ND, NCOMP = W::NCOMP;
    constexpr int GL = grad_group_lanes<LA, LB, LC, LD>(), NGRP = 4 * (64 / GL);
    constexpr int L12 = W::L12, L34 = W::L34, NR = W::NR, NG = W::NG, NX = W::NX;
    constexpr int SA = W::SA, SB = W::SB, SC = W::SC;
    __shared__ FxAcc sg[(TI + TJ + TK + TL) * 3];
    __shared__ double sG[NGRP][3 * NG];
    __shared__ double sT[NGRP][3 * NX];
    // per product: the three axes' table offsets (8 bits each) and the
    // exponents a, b, c of each axis (2 bits each, 6 bits per axis); the
    // normalisation product
    __shared__ int sCo[NCOMP], sCe[NCOMP];
    __shared__ double sNrm[NCOMP];
    // Gamma of the current quartet's products, per group
    __shared__ double sGam[NGRP][NCOMP];

    const int4 it = a.items[blockIdx.x];
    const int tid = threadIdx.x, grp = tid / GL, lane = tid % GL;
    const int fI = a.tile_first[it.x], nI = a.tile_size[it.x];
    const int fJ = a.tile_first[it.y], nJ = a.tile_size[it.y];
    const int fK = a.tile_first[it.z], nK = a.tile_size[it.z];
    const int fL = a.tile_first[it.w], nL = a.tile_size[it.w];
    const bool sameBra = it.x == it.y, sameKet = it.z == it.w, samePair = it.x == it.z && it.y == it.w;
    const int N = a.N;
    for (int e = tid; e < (TI + TJ + TK + TL) * 3; e += 256) sg[e] = FxAcc{0, 0};
    for (int c = tid; c < NCOMP; c += 256) {
        const int ia = c / (NB * NC * ND), ib = (c / (NC * ND)) % NB, ic = (c / ND) % NC, id = c % ND;
        const int ea = cart_exps<LA>(ia), eb = cart_exps<LB>(ib), ec = cart_exps<LC>(ic), ed = cart_exps<LD>(id);
        int co = 0, ce = 0;
        for (int x = 0; x < 3; ++x) {
            const int ax = (ea >> (2 * x)) & 3, bx = (eb >> (2 * x)) & 3, cx = (ec >> (2 * x)) & 3, dx = (ed >> (2 * x)) & 3;
            co |= (ax * SA + bx * SB + cx * SC + dx) << (8 * x);
            ce |= (ax | (bx << 2) | (cx << 4)) << (6 * x);
        }
        sCo[c] = co;
        sCe[c] = ce;
        sNrm[c] = cart_norm_rt<LA>(ia) * cart_norm_rt<LB>(ib) * cart_norm_rt<LC>(ic) * cart_norm_rt<LD>(id);
    }
    __syncthreads();
    double *const G0 = sG[grp];
    double *const T0 = sT[grp];

    for (int q = grp; q < TI * TJ * TK * TL; q += NGRP) {
        const int lq = q % TL, kq = (q / TL) % TK, jq = (q / (TL * TK)) % TJ, iq = q / (TL * TK * TJ);
        if (iq >= nI || jq >= nJ || kq >= nK || lq >= nL) continue;
        const int I = fI + iq, J = fJ + jq, K = fK + kq, L = fL + lq;
        if (sameBra && J > I) continue;
        if (sameKet && L > K) continue;
        if (samePair && !(I > K || (I == K && J >= L))) continue;
        const int pb = I * (I + 1) / 2 + J, pk = K * (K + 1) / 2 + L;
        if (a.pair_q[pb] * a.pair_q[pk] < a.tau) continue;
        double w = 8.0;
        if (I == J) w *= 0.5;
        if (K == L) w *= 0.5;
        if (I == K && J == L) w *= 0.5;
        const PairRef bref = a.pairs[pb], kref = a.pairs[pk];
        double A[3], B[3], Cc[3], D[3];
        for (int x = 0; x < 3; ++x) {
            A[x] = a.shell_xyz[3 * I + x];
            B[x] = a.shell_xyz[3 * J + x];
            Cc[x] = a.shell_xyz[3 * K + x];
            D[x] = a.shell_xyz[3 * L + x];
        }
        const int oI = a.shell_off[I], oJ = a.shell_off[J], oK = a.shell_off[K], oL = a.shell_off[L];
        // Gamma of this lane's products (this group's LDS row)
        double *const gam = sGam[grp];
        for (int c = lane; c < NCOMP; c += GL) {
            const int ia = c / (NB * NC * ND), ib = (c / (NC * ND)) % NB, ic = (c / ND) % NC, id = c % ND;
            const int ra = oI + ia, rb = oJ + ib, rc = oK + ic, rd = oL + id;
            const double Pab = a.D[(size_t)ra * N + rb], Pcd = a.D[(size_t)rc * N + rd];
            const double Pac = a.D[(size_t)ra * N + rc], Pbd = a.D[(size_t)rb * N + rd];
            const double Pad = a.D[(size_t)ra * N + rd], Pbc = a.D[(size_t)rb * N + rc];
            gam[c] = (2.0 * Pab * Pcd - 0.5 * (Pac * Pbd + Pad * Pbc)) * sNrm[c];
        }
        double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
        double AB[3], CD[3];
        for (int x = 0; x < 3; ++x) {
            AB[x] = A[x] - B[x];
            CD[x] = Cc[x] - D[x];
        }
        for (int ib = 0; ib < bref.npp; ++ib) {
            const PrimPair bp = a.pp[bref.pp0 + ib];
            const double2 bab = a.ppab[bref.pp0 + ib];
            for (int ik = 0; ik < kref.npp; ++ik) {
                const PrimPair kp = a.pp[kref.pp0 + ik];
                const double2 kab = a.ppab[kref.pp0 + ik];
                const double p = bp.p, qq = kp.p, pq = p + qq;
                const double rpq = rsqrt_pos(pq), ipq = rpq * rpq;
                const double sr0 = SRTERM * (bp.s * bp.ip) * (kp.s * kp.ip) * rpq;
                if (sr0 < PRIM_SCREEN) continue;   // twoints.c:110-111
                const double sr = sr0 * bp.c * kp.c * w;
                double PQ[3], PA[3], QC[3];
                for (int x = 0; x < 3; ++x) {
                    PQ[x] = bp.P[x] - kp.P[x];
                    PA[x] = bp.P[x] - A[x];
                    QC[x] = kp.P[x] - Cc[x];
                }
                const double T = (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]) * p * qq * ipq;
                const double al2 = 2.0 * bab.x, be2 = 2.0 * bab.y, ga2 = 2.0 * kab.x;
#pragma unroll 1
                for (int r = 0; r < NR; ++r) {
                    double xr, wr;
                    rys_eval_one<NR>(T, r, xr, wr, a.rys);
                    const double fff = xr * ipq, B00 = 0.5 * fff;
                    const double B1 = (0.5 - B00 * qq) * bp.ip, B1p = (0.5 - B00 * p) * kp.ip;
                    if (lane < 3) {
                        const int x = lane;
                        const double pqx = pick3(PQ, x);
                        double G[L12 + 1][L34 + 1];
                        rys_vrr<double, L12, L34>(G, B00, B1, B1p, pick3(PA, x) - qq * pqx * fff, pick3(QC, x) + p * pqx * fff,
                                                  x == 2 ? wr * sr : 1.0);
#pragma unroll
                        for (int i = 0; i <= L12; ++i)
#pragma unroll
                            for (int j = 0; j <= L34; ++j) G0[x * NG + i * (L34 + 1) + j] = G[i][j];
                    }
                    __builtin_amdgcn_wave_barrier();
                    // transfer: entry (axis, a, b, c, d) of the shifted 1-D table
                    for (int e = lane; e < 3 * NX; e += GL) {
                        const int x = e / NX, rem = e % NX;
                        const int A_ = rem / SA, B_ = (rem / SB) % (LB + 2), C_ = (rem / SC) % (LC + 2), D_ = rem % SC;
                        if (A_ + B_ > L12 || C_ + D_ > L34) continue;   // never read
                        const double abx = pick3(AB, x), cdx = pick3(CD, x);
                        const double *g = G0 + x * NG;
                        double sum = 0.0, abp = 1.0;
                        for (int i = 0; i <= B_; ++i) {
                            double t = 0.0, cdp = 1.0;
                            for (int j = 0; j <= D_; ++j) {
                                t += binom3(D_, j) * cdp * g[(A_ + B_ - i) * (L34 + 1) + C_ + D_ - j];
                                cdp *= cdx;
                            }
                            sum += binom3(B_, i) * abp * t;
                            abp *= abx;
                        }
                        T0[e] = sum;
                    }
                    __builtin_amdgcn_wave_barrier();
                    // this lane's products
#pragma unroll 1
                    for (int c = lane; c < NCOMP; c += GL) {
                        const int co = sCo[c], ce = sCe[c];
                        const double gc = gam[c];
                        const int o[3] = {co & 255, (co >> 8) & 255, (co >> 16) & 255};
                        const double v[3] = {T0[o[0]], T0[NX + o[1]], T0[2 * NX + o[2]]};
#pragma unroll
                        for (int x = 0; x < 3; ++x) {
                            const double ox = gc * v[x == 0 ? 1 : 0] * v[x == 2 ? 1 : 2];
                            const double *tx = T0 + x * NX;
                            const int ex = (ce >> (6 * x)) & 63;
                            const int ax = ex & 3, bx = (ex >> 2) & 3, cx = ex >> 4;
                            const int ox_ = o[x];
                            const double da = al2 * tx[ox_ + SA] - ax * tx[max(ox_ - SA, 0)];
                            const double db = be2 * tx[ox_ + SB] - bx * tx[max(ox_ - SB, 0)];
                            const double dc = ga2 * tx[ox_ + SC] - cx * tx[max(ox_ - SC, 0)];
                            gA[x] += ox * da;
                            gB[x] += ox * db;
                            gC[x] += ox * dc;
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
        // the group's sums (fixed butterfly inside the group), then its
        // first lane into the shell sums
        double red[9] = {gA[0], gA[1], gA[2], gB[0], gB[1], gB[2], gC[0], gC[1], gC[2]};
#pragma unroll
        for (int v = 0; v < 9; ++v)
#pragma unroll
            for (int m = GL / 2; m >= 1; m >>= 1) red[v] += __shfl_xor(red[v], m, 64);
        if (lane == 0) {
            for (int x = 0; x < 3; ++x) {
                fx_lds_add(&sg[3 * iq + x], red[x]);
                fx_lds_add(&sg[3 * (TI + jq) + x], red[3 + x]);
                fx_lds_add(&sg[3 * (TI + TJ + kq) + x], red[6 + x]);
                fx_lds_add(&sg[3 * (TI + TJ + TK + lq) + x], -(red[x] + red[3 + x] + red[6 + x]));
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < (TI + TJ + TK + TL) * 3; e += 256) {
        const int slot = e / 3, x = e % 3;
        int sh;
        if (slot < TI) sh = slot < nI ? fI + slot : -1;
        else if (slot < TI + TJ) sh = slot - TI < nJ ? fJ + slot - TI : -1;
        else if (slot < TI + TJ + TK) sh = slot - TI - TJ < nK ? fK + slot - TI - TJ : -1;
        else sh = slot - TI - TJ - TK < nL ? fL + slot - TI - TJ - TK : -1;
        const FxAcc v = sg[e];
        if (sh >= 0 && (v.hi != 0 || v.lo != 0)) fx_global_add(&a.gshell[3 * sh + x], v);
    }
}

template <int LA, int LB, int LC, int LD>
static void launch_grad_t(const GradArgs &args, hipStream_t s)
{
    if (args.nitems <= 0) return;
    if constexpr (grad_grouped<LA, LB, LC, LD>())
        hipLaunchKernelGGL((grad2e_wave_kernel<LA, LB, LC, LD>), dim3(args.nitems), dim3(256), 0, s, args);
    else
        hipLaunchKernelGGL((grad2e_kernel<LA, LB, LC, LD, double>), dim3(args.nitems), dim3(256), 0, s, args);
}

int launch_grad(int bra_pc, int ket_pc, const GradArgs &args, hipStream_t s)
{
    switch (bra_pc * PM_NPC + ket_pc) {
    case 0: launch_grad_t<0, 0, 0, 0>(args, s); return 0;
    case 1: launch_grad_t<0, 0, 1, 0>(args, s); return 0;
    case 3: launch_grad_t<0, 0, 2, 0>(args, s); return 0;
    case 6: launch_grad_t<1, 0, 0, 0>(args, s); return 0;
    case 7: launch_grad_t<1, 0, 1, 0>(args, s); return 0;
    case 9: launch_grad_t<1, 0, 2, 0>(args, s); return 0;
    case 12: launch_grad_t<1, 1, 0, 0>(args, s); return 0;
    case 13: launch_grad_t<1, 1, 1, 0>(args, s); return 0;
    case 14: launch_grad_t<1, 1, 1, 1>(args, s); return 0;
    case 18: launch_grad_t<2, 0, 0, 0>(args, s); return 0;
    case 19: launch_grad_t<2, 0, 1, 0>(args, s); return 0;
    case 20: launch_grad_t<2, 0, 1, 1>(args, s); return 0;
    case 21: launch_grad_t<2, 0, 2, 0>(args, s); return 0;
    case 24: launch_grad_t<2, 1, 0, 0>(args, s); return 0;
    case 25: launch_grad_t<2, 1, 1, 0>(args, s); return 0;
    case 26: launch_grad_t<2, 1, 1, 1>(args, s); return 0;
    case 27: launch_grad_t<2, 1, 2, 0>(args, s); return 0;
    case 28: launch_grad_t<2, 1, 2, 1>(args, s); return 0;
    case 30: launch_grad_t<2, 2, 0, 0>(args, s); return 0;
    case 31: launch_grad_t<2, 2, 1, 0>(args, s); return 0;
    case 32: launch_grad_t<2, 2, 1, 1>(args, s); return 0;
    case 33: launch_grad_t<2, 2, 2, 0>(args, s); return 0;
    case 34: launch_grad_t<2, 2, 2, 1>(args, s); return 0;
    case 35: launch_grad_t<2, 2, 2, 2>(args, s); return 0;
    default: return -1;
    }
}

}  // namespace pm
