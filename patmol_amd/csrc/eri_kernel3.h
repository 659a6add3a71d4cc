// eri_kernel3.h -- J/K kernel with the tile quartet staged in LDS.
//
// Same work decomposition and lane layout as jk_kernel2 (one workgroup per
// tile quartet; lane = (ket sub-slot, bra shell I, bra shell J), bra pair
// fixed per lane; wave w owns ket shells K = w, w+4, ...; each wave steps
// over L in chunks of KSUB), but everything a step reads or writes lives in
// LDS for the duration of the item:
//   * the six density blocks D(IJ), D(KL), D(IK), D(IL), D(JK), D(JL) of the
//     tile quartet (gathered once, coalesced),
//   * the ket pair table (pair reference, Schwarz bound, validity, centres)
//     and, when they fit, the ket primitive pairs,
//   * the six output blocks J'(IJ), J'(KL), K'(IK), K'(IL), K'(JK), K'(JL):
//     cross-lane reductions land there (plain adds where one wave owns the
//     element, LDS atomics where waves share it), and one pass of global
//     atomics at the end of the item moves them to the per-XCD J'/K' copies.
// Classes whose 2-D tables for all roots would not fit the register file
// digest one Rys root at a time (primitive_quartet_rootwise); the choice and
// the register budget per class are measured (jk3_shape).
#pragma once
#include "eri_kernels.h"

namespace pm {

template <int LA, int LB, int LC, int LD>
struct K3 {
    static constexpr int TI = tile_of(LA), TJ = tile_of(LB), TK = tile_of(LC), TL = tile_of(LD);
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int FI = TI * NA, FJ = TJ * NB, FK = TK * NC, FL = TL * ND;
    static constexpr int KLP = TK * TL;
    static constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
    static constexpr int PPCAP = 128;  // staged ket primitive pairs
};

// Per-class launch shape, from timing every class on (H2O)8 with both
// digestion orders at 1 and 2 waves per SIMD: code = waves * 2 + rootwise.
constexpr int jk3_shape(int la, int lb, int lc, int ld)
{
    const int c = ((la * 3 + lb) * 3 + lc) * 3 + ld;
    switch (c) {
    case 0: return 1 * 2 + 1;    // (ss|ss)
    case 27: return 2 * 2 + 1;   // (ps|ss)
    case 30: return 2 * 2 + 0;   // (ps|ps)
    case 36: return 2 * 2 + 0;   // (pp|ss)
    case 39: return 2 * 2 + 0;   // (pp|ps)
    case 40: return 1 * 2 + 1;   // (pp|pp)
    case 54: return 2 * 2 + 0;   // (ds|ss)
    case 57: return 2 * 2 + 1;   // (ds|ps)
    case 58: return 1 * 2 + 0;   // (ds|pp)
    case 60: return 1 * 2 + 0;   // (ds|ds)
    case 63: return 2 * 2 + 1;   // (dp|ss)
    case 66: return 1 * 2 + 0;   // (dp|ps)
    case 67: return 1 * 2 + 1;   // (dp|pp)
    case 69: return 1 * 2 + 1;   // (dp|ds)
    case 70: return 1 * 2 + 1;   // (dp|dp)
    case 72: return 1 * 2 + 0;   // (dd|ss)
    case 75: return 1 * 2 + 1;   // (dd|ps)
    case 78: return 1 * 2 + 1;   // (dd|ds)
    default: return 1 * 2 + 1;
    }
}

template <int LA, int LB, int LC, int LD>
constexpr bool jk3_rootwise()
{
#ifdef PM_K3_ROOTWISE
    return PM_K3_ROOTWISE;
#else
    return jk3_shape(LA, LB, LC, LD) & 1;
#endif
}

template <int LA, int LB, int LC, int LD>
constexpr int jk3_waves_per_eu()
{
#ifdef PM_JK_WAVES
    return PM_JK_WAVES;
#else
    return jk3_shape(LA, LB, LC, LD) >> 1;
#endif
}

template <int LA, int LB, int LC, int LD, typename R>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(jk3_waves_per_eu<LA, LB, LC, LD>()))) void jk_kernel3(JKArgs a)
{
    using S = K3<LA, LB, LC, LD>;
    constexpr int TI = S::TI, TJ = S::TJ, TK = S::TK, TL = S::TL;
    constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND;
    constexpr int FI = S::FI, FJ = S::FJ, FK = S::FK, FL = S::FL, KLP = S::KLP;
    constexpr int IJN = TI * TJ, KSUB = 64 / IJN;
    constexpr int JBITS = ilog2(TJ), IBITS = ilog2(TI), SBITS = ilog2(KSUB);
    constexpr int MASK_J = (1 << JBITS) - 1;
    constexpr int MASK_I = ((1 << IBITS) - 1) << JBITS;
    constexpr int MASK_S = ((1 << SBITS) - 1) << (JBITS + IBITS);
    constexpr int NLSTEP = (TL + KSUB - 1) / KSUB;
    static_assert(IJN * KSUB == 64, "bra tile pair must divide a wavefront");

    __shared__ double sD[FI * FJ + FK * FL + FI * FK + FI * FL + FJ * FK + FJ * FL];
    __shared__ double sO[FI * FJ + FK * FL + FI * FK + FI * FL + FJ * FK + FJ * FL];
    __shared__ int sOff[TI + TJ + TK + TL];
    __shared__ double sXK[TK][3], sXL[TL][3];
    __shared__ PairRef sKref[KLP];
    __shared__ double sKq[KLP];
    __shared__ int sKok[KLP];
    __shared__ int sPPoff[KLP];
    __shared__ PrimPair sPP[S::PPCAP];
    __shared__ int sStaged, sKmax;
    // block offsets inside sD / sO
    constexpr int oIJ = 0, oKL = oIJ + FI * FJ, oIK = oKL + FK * FL, oIL = oIK + FI * FK, oJK = oIL + FI * FL,
                  oJL = oJK + FJ * FK, oEND = oJL + FJ * FL;

    const int item = blockIdx.x;
    if (item >= a.nitems) return;
    const int4 it = a.items[item];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int jj = lane & (TJ - 1), ii = (lane >> JBITS) & (TI - 1), ks = lane >> (JBITS + IBITS);
    const int fI = a.tile_first[it.x], nI = a.tile_size[it.x];
    const int fJ = a.tile_first[it.y], nJ = a.tile_size[it.y];
    const int fK = a.tile_first[it.z], nK = a.tile_size[it.z];
    const int fL = a.tile_first[it.w], nL = a.tile_size[it.w];
    const bool sameBra = it.x == it.y, sameKet = it.z == it.w, samePair = it.x == it.z && it.y == it.w;
    const int N = a.N;

    // ---- stage: function offsets, ket pairs, density blocks, zero outputs ----
    if (tid < TI) sOff[tid] = a.shell_off[fI + (tid < nI ? tid : 0)];
    else if (tid < TI + TJ) sOff[tid] = a.shell_off[fJ + (tid - TI < nJ ? tid - TI : 0)];
    else if (tid < TI + TJ + TK) sOff[tid] = a.shell_off[fK + (tid - TI - TJ < nK ? tid - TI - TJ : 0)];
    else if (tid < TI + TJ + TK + TL) sOff[tid] = a.shell_off[fL + (tid - TI - TJ - TK < nL ? tid - TI - TJ - TK : 0)];
    if (tid < 3 * TK) {
        const int k = tid / 3, x = tid % 3;
        sXK[k][x] = a.shell_xyz[3 * (fK + (k < nK ? k : 0)) + x];
    } else if (tid >= 64 && tid < 64 + 3 * TL) {
        const int l = (tid - 64) / 3, x = (tid - 64) % 3;
        sXL[l][x] = a.shell_xyz[3 * (fL + (l < nL ? l : 0)) + x];
    }
    if (tid >= 128 && tid < 128 + KLP) {
        const int t = tid - 128, kq = t / TL, lq = t % TL;
        const bool kex = kq < nK, lex = lq < nL;
        const int K = fK + (kex ? kq : 0), L = fL + (lex ? lq : 0);
        const bool ok = kex && lex && (!sameKet || K >= L);
        // a structurally valid stand-in pair for lanes whose (K, L) does not exist
        const int Ls = ok ? L : (sameKet ? K : fL);
        const int Kp = K >= Ls ? K : Ls, Lp = K >= Ls ? Ls : K;
        const int pk = Kp * (Kp + 1) / 2 + Lp;
        sKref[t] = a.pairs[pk];
        sKq[t] = a.pair_q[pk];
        sKok[t] = ok ? 1 : 0;
    }
    {
        const double *__restrict__ DJ = a.DJ;
        const double *__restrict__ DK = a.DK0;
        for (int e = tid; e < oEND; e += 256) {
            int r, c, ro, co;
            const double *src;
            bool ex;
            if (e < oKL) {
                r = (e - oIJ) / FJ; c = (e - oIJ) % FJ;
                ex = r / NA < nI && c / NB < nJ;
                ro = a.shell_off[fI + (ex ? r / NA : 0)] + r % NA; co = a.shell_off[fJ + (ex ? c / NB : 0)] + c % NB;
                src = DJ;
            } else if (e < oIK) {
                r = (e - oKL) / FL; c = (e - oKL) % FL;
                ex = r / NC < nK && c / ND < nL;
                ro = a.shell_off[fK + (ex ? r / NC : 0)] + r % NC; co = a.shell_off[fL + (ex ? c / ND : 0)] + c % ND;
                src = DJ;
            } else if (e < oIL) {
                r = (e - oIK) / FK; c = (e - oIK) % FK;
                ex = r / NA < nI && c / NC < nK;
                ro = a.shell_off[fI + (ex ? r / NA : 0)] + r % NA; co = a.shell_off[fK + (ex ? c / NC : 0)] + c % NC;
                src = DK;
            } else if (e < oJK) {
                r = (e - oIL) / FL; c = (e - oIL) % FL;
                ex = r / NA < nI && c / ND < nL;
                ro = a.shell_off[fI + (ex ? r / NA : 0)] + r % NA; co = a.shell_off[fL + (ex ? c / ND : 0)] + c % ND;
                src = DK;
            } else if (e < oJL) {
                r = (e - oJK) / FK; c = (e - oJK) % FK;
                ex = r / NB < nJ && c / NC < nK;
                ro = a.shell_off[fJ + (ex ? r / NB : 0)] + r % NB; co = a.shell_off[fK + (ex ? c / NC : 0)] + c % NC;
                src = DK;
            } else {
                r = (e - oJL) / FL; c = (e - oJL) % FL;
                ex = r / NB < nJ && c / ND < nL;
                ro = a.shell_off[fJ + (ex ? r / NB : 0)] + r % NB; co = a.shell_off[fL + (ex ? c / ND : 0)] + c % ND;
                src = DK;
            }
            sD[e] = ex ? src[(size_t)ro * N + co] : 0.0;
            sO[e] = 0.0;
        }
    }
    __syncthreads();
    if (tid == 0) {
        int tot = 0, mx = 1;
        for (int t = 0; t < KLP; ++t) {
            sPPoff[t] = tot;
            tot += sKref[t].npp;
            mx = max(mx, sKref[t].npp);
        }
        sStaged = tot <= S::PPCAP ? 1 : 0;
        sKmax = mx;
    }
    __syncthreads();
    const bool staged = sStaged != 0;
    if (staged && tid < KLP) {
        const PairRef r = sKref[tid];
        for (int p = 0; p < r.npp; ++p) sPP[sPPoff[tid] + p] = a.pp[r.pp0 + p];
    }
    __syncthreads();

    // ---- bra pair (fixed per lane) ----
    const bool iexist = ii < nI, jexist = jj < nJ;
    const int I = fI + (iexist ? ii : 0), J = fJ + (jexist ? jj : 0);
    const bool bvalid = iexist && jexist && (!sameBra || I >= J);
    const int Is = bvalid ? I : fI, Js = bvalid ? J : fJ;
    const int Ip = Is >= Js ? Is : Js, Jp = Is >= Js ? Js : Is;
    const int pb = Ip * (Ip + 1) / 2 + Jp;
    const double qbra = a.pair_q[pb];
    const PairRef bref = a.pairs[pb];
    PairGeom<R> bg;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        bg.A[x] = R(a.shell_xyz[3 * Is + x]);
        bg.AB[x] = R(a.shell_xyz[3 * Is + x] - a.shell_xyz[3 * Js + x]);
    }
    const int origIJ = max(a.shell_orig[Is], a.shell_orig[Js]);
    const int nbp_lane = bref.npp;
    int nbp = nbp_lane;
#pragma unroll
    for (int b = 0; b < 6; ++b) nbp = max(nbp, __shfl_xor(nbp, 1 << b, 64));
    // bra primitive pairs: contraction 1 is the common case, keep it in registers
    const PrimPair bp0 = a.pp[bref.pp0];

    double jab[NA * NB];
#pragma unroll
    for (int i = 0; i < NA * NB; ++i) jab[i] = 0.0;
    const double *__restrict__ DIJ = sD + oIJ + (ii * NA) * FJ + jj * NB;

    for (int kq = wave; kq < TK; kq += 4) {
        if (kq >= nK) break;  // uniform
        const int K = fK + kq;
        double kac[NA * NC], kbc[NB * NC];
#pragma unroll
        for (int i = 0; i < NA * NC; ++i) kac[i] = 0.0;
#pragma unroll
        for (int i = 0; i < NB * NC; ++i) kbc[i] = 0.0;
        for (int lc = 0; lc < NLSTEP; ++lc) {
            const int lq0 = lc * KSUB + ks;
            const bool lexist = lq0 < nL;
            const int lq = lexist ? lq0 : 0;
            const int t = kq * TL + lq;
            const int L = fL + lq;
            bool valid = bvalid && lexist && sKok[t] != 0;
            if (samePair) valid = valid && (I > K || (I == K && J >= L));
            const double qket = sKq[t];
            valid = valid && (qbra * qket >= a.tau);
            if (a.start_shell > 0) valid = valid && (max(origIJ, max(a.shell_orig[K], a.shell_orig[L])) >= a.start_shell);
            if (!__any(valid)) continue;  // uniform
            double wdeg = 1.0;
            if (I == J) wdeg *= 0.5;
            if (K == L) wdeg *= 0.5;
            if (I == K && J == L) wdeg *= 0.5;
            const R scale = valid ? R(wdeg) : R(0);
            const PairRef kref = sKref[t];
            // the stand-in partner of an invalid (K, L) is K itself (equal ket
            // tiles, so fK == fL) or the first shell of the L tile
            PairGeom<R> kg;
            const int lqs = (sKok[t] != 0) ? lq : (sameKet ? kq : 0);
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                kg.A[x] = R(sXK[kq][x]);
                kg.AB[x] = R(sXK[kq][x] - sXL[lqs][x]);
            }
            double jcd[NC * ND], kad[NA * ND], kbd[NB * ND];
#pragma unroll
            for (int i = 0; i < NC * ND; ++i) jcd[i] = 0.0;
#pragma unroll
            for (int i = 0; i < NA * ND; ++i) kad[i] = 0.0;
#pragma unroll
            for (int i = 0; i < NB * ND; ++i) kbd[i] = 0.0;
            const double *__restrict__ DKL = sD + oKL + (kq * NC) * FL + lq * ND;
            const double *__restrict__ DJL = sD + oJL + (jj * NB) * FL + lq * ND;
            const double *__restrict__ DIL = sD + oIL + (ii * NA) * FL + lq * ND;
            const double *__restrict__ DJK = sD + oJK + (jj * NB) * FK + kq * NC;
            const double *__restrict__ DIK = sD + oIK + (ii * NA) * FK + kq * NC;
            const int nkp = sKmax;  // uniform bound of the ket primitive loop
            const PrimPair *kpp = staged ? (sPP + sPPoff[t]) : (a.pp + kref.pp0);
            auto digest = [&](auto ia, auto ib_, auto ic, auto id, R vr) {
                constexpr int A_ = decltype(ia)::value, B_ = decltype(ib_)::value;
                constexpr int C_ = decltype(ic)::value, D_ = decltype(id)::value;
                const double v = double(vr);
                jab[A_ * NB + B_] += v * DKL[C_ * FL + D_];
                jcd[C_ * ND + D_] += v * DIJ[A_ * FJ + B_];
                kac[A_ * NC + C_] += v * DJL[B_ * FL + D_];
                kad[A_ * ND + D_] += v * DJK[B_ * FK + C_];
                kbc[B_ * NC + C_] += v * DIL[A_ * FL + D_];
                kbd[B_ * ND + D_] += v * DIK[A_ * FK + C_];
            };
            for (int ib = 0; ib < nbp; ++ib) {
                const PrimPair bp = (ib == 0) ? bp0 : a.pp[bref.pp0 + min(ib, nbp_lane - 1)];
                for (int ik = 0; ik < nkp; ++ik) {
                    const PrimPair kp = kpp[min(ik, kref.npp - 1)];
                    const R sc = (ib < nbp_lane && ik < kref.npp) ? scale : R(0);
                    if constexpr (jk3_rootwise<LA, LB, LC, LD>())
                        primitive_quartet_rootwise<R, LA, LB, LC, LD>(bp, kp, bg, kg, a.rys, sc, digest);
                    else
                        primitive_quartet<R, LA, LB, LC, LD>(bp, kp, bg, kg, a.rys, sc, digest);
                }
            }
            // per-step results: J'(KL) (one wave owns K), K'(IL), K'(JL) (shared by waves)
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const double s = lane_reduce<MASK_I | MASK_J>(jcd[c * ND + d]);
                    if ((lane & (MASK_I | MASK_J)) == 0 && lexist) sO[oKL + (kq * NC + c) * FL + lq * ND + d] += 2.0 * s;
                }
#pragma unroll
            for (int x = 0; x < NA; ++x)
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const double s = lane_reduce<MASK_J>(kad[x * ND + d]);
                    if ((lane & MASK_J) == 0 && iexist && lexist && s != 0.0)
                        __hip_atomic_fetch_add(&sO[oIL + (ii * NA + x) * FL + lq * ND + d], s, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                }
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const double s = lane_reduce<MASK_I>(kbd[b * ND + d]);
                    if ((lane & MASK_I) == 0 && jexist && lexist && s != 0.0)
                        __hip_atomic_fetch_add(&sO[oJL + (jj * NB + b) * FL + lq * ND + d], s, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                }
        }
        // per-K results: K'(IK), K'(JK) (one wave owns K)
#pragma unroll
        for (int x = 0; x < NA; ++x)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const double s = lane_reduce<MASK_J | MASK_S>(kac[x * NC + c]);
                if ((lane & (MASK_J | MASK_S)) == 0 && iexist) sO[oIK + (ii * NA + x) * FK + kq * NC + c] += s;
            }
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const double s = lane_reduce<MASK_I | MASK_S>(kbc[b * NC + c]);
                if ((lane & (MASK_I | MASK_S)) == 0 && jexist) sO[oJK + (jj * NB + b) * FK + kq * NC + c] += s;
            }
    }
    // per-item result: J'(IJ) (shared by the four waves)
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const double s = lane_reduce<MASK_S>(jab[x * NB + b]);
            if ((lane & MASK_S) == 0 && iexist && jexist && s != 0.0)
                __hip_atomic_fetch_add(&sO[oIJ + (ii * NA + x) * FJ + jj * NB + b], 2.0 * s, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    __syncthreads();

    // ---- flush the output blocks to this XCD's J'/K' copy ----
    const size_t xoff = (size_t)xcd_slot() * N * N;
    double *const Jx = a.J + xoff;
    double *const Kx = a.K0 + xoff;
    for (int e = tid; e < oEND; e += 256) {
        const double v = sO[e];
        if (v == 0.0) continue;
        int ro, co;
        double *dst;
        if (e < oKL) {
            const int r = (e - oIJ) / FJ, c = (e - oIJ) % FJ;
            ro = sOff[r / NA] + r % NA; co = sOff[TI + c / NB] + c % NB; dst = Jx;
        } else if (e < oIK) {
            const int r = (e - oKL) / FL, c = (e - oKL) % FL;
            ro = sOff[TI + TJ + r / NC] + r % NC; co = sOff[TI + TJ + TK + c / ND] + c % ND; dst = Jx;
        } else if (e < oIL) {
            const int r = (e - oIK) / FK, c = (e - oIK) % FK;
            ro = sOff[r / NA] + r % NA; co = sOff[TI + TJ + c / NC] + c % NC; dst = Kx;
        } else if (e < oJK) {
            const int r = (e - oIL) / FL, c = (e - oIL) % FL;
            ro = sOff[r / NA] + r % NA; co = sOff[TI + TJ + TK + c / ND] + c % ND; dst = Kx;
        } else if (e < oJL) {
            const int r = (e - oJK) / FK, c = (e - oJK) % FK;
            ro = sOff[TI + r / NB] + r % NB; co = sOff[TI + TJ + c / NC] + c % NC; dst = Kx;
        } else {
            const int r = (e - oJL) / FL, c = (e - oJL) % FL;
            ro = sOff[TI + r / NB] + r % NB; co = sOff[TI + TJ + TK + c / ND] + c % ND; dst = Kx;
        }
        atomic_add(&dst[(size_t)ro * N + co], v);
    }
}

}  // namespace pm
