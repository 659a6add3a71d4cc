// eri_kernels.hip -- direct J/K build, ERI export and Schwarz kernels.
//
// Work decomposition (replaces the reference's stored integral list and its
// per-integral scatter, twoints.c:228-404 + scfaux.c:428-552):
//   shells are sorted by (l, exponent, centre) and cut into tiles of T(l) shells; a
//   work item is a canonical quartet of tiles (bra tile pair >= ket tile pair)
//   and one workgroup (4 wavefronts) processes every shell quartet in it.
//   Lanes of a wavefront hold fixed ket shell pairs (K,L) laid out as
//   lane = (js*TK + kk)*TL + ll; the wavefront steps over bra shells I and
//   takes JSUB bra shells J per step.  Each (bra, ket) lane evaluates one
//   primitive quartet after another and contracts the integrals with the
//   density at once.
// J/K accumulation: per-lane register accumulators per Cartesian block,
// reduced across the lanes that share a target (XOR butterflies over the
// lane bits the target does not depend on) and added by one leader lane with
// a device-scope FP64 atomic.  Targets, with w the degeneracy weight
// 2^-([I==J]+[K==L]+[IJ==KL]) folded into the integral v:
//   J'(ab) += 2 v D(cd)   J'(cd) += 2 v D(ab)
//   K'(ac) += v D(bd)     K'(ad) += v D(bc)   K'(bc) += v D(ad)   K'(bd) += v D(ac)
// and J = J' + J'^T, K = K' + K'^T afterwards.
#pragma once
#include "eri_kernels.h"

namespace pm {

// Lane i + 2^k of the same row of 16 (DPP row_shl), for 2^k <= 8.
template <int K>
__device__ __forceinline__ double dpp_row_shl(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x100 | (1 << K), 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x100 | (1 << K), 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Lane i + 16 for lanes in even rows (v_permlane16_swap), lane i + 32 for
// lanes 0..31 (v_permlane32_swap).
__device__ __forceinline__ double lane_plus16(double v)
{
    const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
    return __hiloint2double(hi[1], lo[1]);
}
__device__ __forceinline__ double lane_plus32(double v)
{
    const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
    return __hiloint2double(hi[1], lo[1]);
}

// Sum of v over the lanes that differ from this one only in the MASK bits,
// valid in the leader lanes ((lane & MASK) == 0) only.  Every lane of the
// wavefront must be active.  Stage k adds lane i + 2^k, which for a leader
// is itself a leader of the lower stages, so the partial sums combine.
// The same for N values at once, stage by stage, so that the cross-lane
// moves of different values overlap.
template <int MASK, int N>
__device__ __forceinline__ void lane_reduce_n(double (&v)[N])
{
    if constexpr ((MASK & 1) != 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] += dpp_row_shl<0>(v[i]);
    }
    if constexpr ((MASK & 2) != 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] += dpp_row_shl<1>(v[i]);
    }
    if constexpr ((MASK & 4) != 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] += dpp_row_shl<2>(v[i]);
    }
    if constexpr ((MASK & 8) != 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] += dpp_row_shl<3>(v[i]);
    }
    if constexpr ((MASK & 16) != 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] += lane_plus16(v[i]);
    }
    if constexpr ((MASK & 32) != 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] += lane_plus32(v[i]);
    }
}

template <int MASK>
__device__ __forceinline__ double lane_reduce(double v)
{
    if constexpr ((MASK & 1) != 0) v += dpp_row_shl<0>(v);
    if constexpr ((MASK & 2) != 0) v += dpp_row_shl<1>(v);
    if constexpr ((MASK & 4) != 0) v += dpp_row_shl<2>(v);
    if constexpr ((MASK & 8) != 0) v += dpp_row_shl<3>(v);
    if constexpr ((MASK & 16) != 0) v += lane_plus16(v);
    if constexpr ((MASK & 32) != 0) v += lane_plus32(v);
    return v;
}

// J'/K' accumulate in one private copy per XCD (the XCD's L2 is the
// coherence point for its own workgroups), so the atomics are workgroup
// scope and stay in L2; the copies are summed after the kernels finish.
__device__ __forceinline__ void atomic_add(double *p, double v)
{
    if (v != 0.0) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// XCD this workgroup runs on (hardware register XCC_ID).
__device__ __forceinline__ int xcd_slot()
{
    int x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(x));
    return x & (PM_XCD_SLOTS - 1);
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

// Register budget per class: the requested waves per SIMD bound the
// registers per lane (4 -> 128, 2 -> 256, 1 -> 512 with the accumulation
// registers), chosen so the larger classes do not spill to scratch.
template <int LA, int LB, int LC, int LD>
constexpr int jk_waves_per_eu()
{
    // live doubles: J/K accumulators, preloaded density blocks and the
    // 2-D recursion arrays of all roots
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int acc = NA * NB + NC * ND + NA * NC + NA * ND + NB * NC + NB * ND;
    constexpr int pre = NA * NB + NC * ND + (NA + NB) * (NC + ND);
    constexpr int g = ((LA + LB + LC + LD) / 2 + 1) * 3 * (LA + LB + 1) * (LC + LD + 1);
    constexpr int live = acc + (pre <= 64 ? pre : 0) + g;
#ifdef PM_JK_WAVES
    return PM_JK_WAVES;
#else
    // (dd|pp) measured faster at 2 waves per SIMD
    if (LA == 2 && LB == 2 && LC == 1 && LD == 1) return 2;
    return live <= 40 ? 4 : (live <= 96 ? 2 : 1);
#endif
}

template <int LA, int LB, int LC, int LD, typename R, int NK>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(jk_waves_per_eu<LA, LB, LC, LD>()))) void jk_kernel(JKArgs a)
{
    constexpr int TI = tile_of(LA), TJ = tile_of(LB), TK = tile_of(LC), TL = tile_of(LD);
    constexpr int KLN = TK * TL, JSUB = 64 / KLN;
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int LBITS = ilog2(TL), KBITS = ilog2(TK), JBITS = ilog2(JSUB);
    constexpr int MASK_L = ((1 << LBITS) - 1);
    constexpr int MASK_K = ((1 << KBITS) - 1) << LBITS;
    constexpr int MASK_J = ((1 << JBITS) - 1) << (LBITS + KBITS);
    constexpr int NJSTEP = TJ / JSUB;
    constexpr int NSTEP = TI * NJSTEP;

    const int item = blockIdx.x;
    if (item >= a.nitems) return;
    const int4 it = a.items[item];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ll = lane & (TL - 1), kk = (lane >> LBITS) & (TK - 1), js = lane >> (LBITS + KBITS);
    const int fI = a.tile_first[it.x], nI = a.tile_size[it.x];
    const int fJ = a.tile_first[it.y], nJ = a.tile_size[it.y];
    const int fK = a.tile_first[it.z], nK = a.tile_size[it.z];
    const int fL = a.tile_first[it.w], nL = a.tile_size[it.w];
    const bool sameBra = it.x == it.y, sameKet = it.z == it.w, samePair = it.x == it.z && it.y == it.w;
    const int N = a.N;
    const size_t xoff = (size_t)xcd_slot() * N * N;
    double *const Jx = a.J + xoff;
    double *const K0x = a.K0 + xoff;
    double *const K1x = a.K1 + xoff;

    // ---- ket (fixed per lane) ----
    const bool kexist = kk < nK, lexist = ll < nL;
    const int K = fK + (kexist ? kk : 0), L = fL + (lexist ? ll : 0);
    const bool kvalid = kexist && lexist && (!sameKet || K >= L);
    const int Ks = kvalid ? K : fK, Ls = kvalid ? L : fL;  // safe indices
    const int Kp = Ks >= Ls ? Ks : Ls, Lp = Ks >= Ls ? Ls : Ks;
    const int pk = Kp * (Kp + 1) / 2 + Lp;
    const double qket = a.pair_q[pk];
    const PairRef kref = a.pairs[pk];
    PairGeom<R> kg;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        kg.A[x] = R(a.shell_xyz[3 * Ks + x]);
        kg.AB[x] = R(a.shell_xyz[3 * Ks + x] - a.shell_xyz[3 * Ls + x]);
    }
    const int offK = a.shell_off[K], offL = a.shell_off[L];
    const int origK = a.shell_orig[Ks], origL = a.shell_orig[Ls];

    double jcd[NC * ND];
#pragma unroll
    for (int i = 0; i < NC * ND; ++i) jcd[i] = 0.0;

    for (int step = wave; step < NSTEP; step += 4) {
        const int iI = step / NJSTEP, jc = step % NJSTEP;
        if (iI >= nI) continue;  // uniform
        const int I = fI + iI;
        const int jl = jc * JSUB + js;
        const bool jexist = jl < nJ;
        const int J = fJ + (jexist ? jl : 0);
        const bool jvalid = jexist && (!sameBra || I >= J);
        const int Js = jvalid ? J : fJ;
        const int Ip = I >= Js ? I : Js, Jp = I >= Js ? Js : I;
        const int pb = Ip * (Ip + 1) / 2 + Jp;
        const double qbra = a.pair_q[pb];
        bool valid = kvalid && jvalid;
        if (samePair) valid = valid && (I > K || (I == K && J >= L));
        valid = valid && (qbra * qket >= a.tau);
        if (a.start_shell > 0) {
            int mo = max(max(a.shell_orig[I], a.shell_orig[Js]), max(origK, origL));
            valid = valid && (mo >= a.start_shell);
        }
        if (!__any(valid)) continue;
        double wdeg = 1.0;
        if (I == J) wdeg *= 0.5;
        if (K == L) wdeg *= 0.5;
        if (I == K && J == L) wdeg *= 0.5;
        const R scale = valid ? R(wdeg) : R(0);
        const PairRef bref = a.pairs[pb];
        PairGeom<R> bg;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            bg.A[x] = R(a.shell_xyz[3 * I + x]);
            bg.AB[x] = R(a.shell_xyz[3 * I + x] - a.shell_xyz[3 * Js + x]);
        }
        const int offI = a.shell_off[I], offJ = a.shell_off[Js];
        const double *__restrict__ DJ = a.DJ;
        const double *__restrict__ DK0 = a.DK0;
        const double *__restrict__ DK1 = a.DK1;

        // primitive loops: bounds are made wave-uniform so the lane reductions
        // inside the component loop see every lane
        int nbp = bref.npp, nkp = kref.npp;
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            nbp = max(nbp, __shfl_xor(nbp, 1 << b, 64));
            nkp = max(nkp, __shfl_xor(nkp, 1 << b, 64));
        }
        for (int ib = 0; ib < nbp; ++ib) {
            const PrimPair bp = a.pp[bref.pp0 + min(ib, bref.npp - 1)];
            for (int ik = 0; ik < nkp; ++ik) {
                const PrimPair kp = a.pp[kref.pp0 + min(ik, kref.npp - 1)];
                const R sc = (ib < bref.npp && ik < kref.npp) ? scale : R(0);
                double jab = 0.0;
                double kac[NK][NC], kad[NK][ND], kbc[NK][NC], kbd[NK][ND];
#pragma unroll
                for (int s = 0; s < NK; ++s) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) kac[s][c] = kbc[s][c] = 0.0;
#pragma unroll
                    for (int d = 0; d < ND; ++d) kad[s][d] = kbd[s][d] = 0.0;
                }
                // per-component contraction and block flushes; IA, IB may be
                // compile-time or run-time (uniform across the wavefront)
                auto on_comp = [&](int IA, int IB, auto ic, auto id, double v) {
                        constexpr int IC = decltype(ic)::value, ID = decltype(id)::value;
                        const int ra = offI + IA, rb = offJ + IB, rc = offK + IC, rd = offL + ID;
                        jab += v * DJ[(size_t)rc * N + rd];
                        jcd[IC * ND + ID] += v * DJ[(size_t)ra * N + rb];
                        kac[0][IC] += v * DK0[(size_t)rb * N + rd];
                        kad[0][ID] += v * DK0[(size_t)rb * N + rc];
                        kbc[0][IC] += v * DK0[(size_t)ra * N + rd];
                        kbd[0][ID] += v * DK0[(size_t)ra * N + rc];
                        if constexpr (NK == 2) {
                            kac[1][IC] += v * DK1[(size_t)rb * N + rd];
                            kad[1][ID] += v * DK1[(size_t)rb * N + rc];
                            kbc[1][IC] += v * DK1[(size_t)ra * N + rd];
                            kbd[1][ID] += v * DK1[(size_t)ra * N + rc];
                        }
                        if constexpr (IC == NC - 1 && ID == ND - 1) {
                            // end of the (a,b) block
                            double t = lane_reduce<MASK_K | MASK_L>(jab);
                            if ((lane & (MASK_K | MASK_L)) == 0 && jexist)
                                atomic_add(&Jx[(size_t)ra * N + rb], 2.0 * t);
                            jab = 0.0;
#pragma unroll
                            for (int s = 0; s < NK; ++s) {
                                double *Kt = s ? K1x : K0x;
#pragma unroll
                                for (int c = 0; c < NC; ++c) {
                                    double u = lane_reduce<MASK_L>(kbc[s][c]);
                                    if ((lane & MASK_L) == 0 && jexist && kexist)
                                        atomic_add(&Kt[(size_t)rb * N + offK + c], u);
                                    kbc[s][c] = 0.0;
                                }
#pragma unroll
                                for (int d = 0; d < ND; ++d) {
                                    double u = lane_reduce<MASK_K>(kbd[s][d]);
                                    if ((lane & MASK_K) == 0 && jexist && lexist)
                                        atomic_add(&Kt[(size_t)rb * N + offL + d], u);
                                    kbd[s][d] = 0.0;
                                }
                                if (IB == NB - 1) {
#pragma unroll
                                    for (int c = 0; c < NC; ++c) {
                                        double u = lane_reduce<MASK_L | MASK_J>(kac[s][c]);
                                        if ((lane & (MASK_L | MASK_J)) == 0 && kexist)
                                            atomic_add(&Kt[(size_t)ra * N + offK + c], u);
                                        kac[s][c] = 0.0;
                                    }
#pragma unroll
                                    for (int d = 0; d < ND; ++d) {
                                        double u = lane_reduce<MASK_K | MASK_J>(kad[s][d]);
                                        if ((lane & (MASK_K | MASK_J)) == 0 && lexist)
                                            atomic_add(&Kt[(size_t)ra * N + offL + d], u);
                                        kad[s][d] = 0.0;
                                    }
                                }
                            }
                        }
                };
                if constexpr (!HeavyClass<LA, LB, LC, LD>::value) {
                    primitive_quartet<R, LA, LB, LC, LD>(bp, kp, bg, kg, a.rys, sc,
                                                         [&](auto ia, auto ib_, auto ic, auto id, R vr) {
                                                             on_comp(decltype(ia)::value, decltype(ib_)::value, ic, id,
                                                                     double(vr));
                                                         });
                } else {
                    OneDTable<R, LA, LB, LC, LD> tbl;
                    tbl.build(bp, kp, bg, kg, a.rys, sc);
#pragma unroll 1
                    for (int ia = 0; ia < NA; ++ia) {
#pragma unroll 1
                        for (int ib2 = 0; ib2 < NB; ++ib2) {
                            const double nab = cart_norm_rt<LA>(ia) * cart_norm_rt<LB>(ib2);
                            static_for<NC>([&](auto ic) {
                                static_for<ND>([&](auto id) {
                                    constexpr int IC = decltype(ic)::value, ID = decltype(id)::value;
                                    constexpr double ncd = 1.0 / csqrt(cart_dfprod(LC, IC) * cart_dfprod(LD, ID));
                                    const double v = double(tbl.template value<IC, ID>(ia, ib2)) * nab * ncd;
                                    on_comp(ia, ib2, ic, id, v);
                                });
                            });
                        }
                    }
                }
            }
        }
    }
    // J'(cd) accumulated over all bra steps of this wavefront
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double t = lane_reduce<MASK_J>(jcd[c * ND + d]);
            if ((lane & MASK_J) == 0 && kexist && lexist) atomic_add(&Jx[(size_t)(offK + c) * N + offL + d], 2.0 * t);
        }
}

// ---------------------------------------------------------------------------
// J/K kernel with bra shell pairs fixed per lane (all classes except the
// largest).  Lane layout lane = (ks*TI + ii)*TJ + jj: lane holds bra pair
// (I,J) = (fI+ii, fJ+jj) for the whole work item, and KSUB = 64/(TI*TJ)
// ket pairs are processed per step.  A wavefront owns ket shells K = wave,
// wave+4, ... and steps over the L shells of each K.  Hence:
//   J'(ab)          lane-private for the whole item (flushed once),
//   K'(ac), K'(bc)  accumulated over all L of one K (flushed per K),
//   J'(cd), K'(ad), K'(bd)  flushed per step,
// each flush being a butterfly over the lane bits its target does not
// depend on, then one FP64 atomic per target from a leader lane.  Since the
// bra class is the larger one, the per-step targets are the small ones.
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
constexpr bool jk2_preload()
{
    return ncart(LA) * ncart(LB) + ncart(LC) * ncart(LD) +
               (ncart(LA) + ncart(LB)) * (ncart(LC) + ncart(LD)) <= 64;
}

template <int LA, int LB, int LC, int LD, typename R>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(jk_waves_per_eu<LA, LB, LC, LD>()))) void jk_kernel2(JKArgs a)
{
    constexpr int TI = tile_of(LA), TJ = tile_of(LB), TK = tile_of(LC), TL = tile_of(LD);
    constexpr int IJN = TI * TJ, KSUB = 64 / IJN;
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int JBITS = ilog2(TJ), IBITS = ilog2(TI), SBITS = ilog2(KSUB);
    constexpr int MASK_J = (1 << JBITS) - 1;
    constexpr int MASK_I = ((1 << IBITS) - 1) << JBITS;
    constexpr int MASK_S = ((1 << SBITS) - 1) << (JBITS + IBITS);
    constexpr int NLSTEP = (TL + KSUB - 1) / KSUB;
    constexpr bool PRE = jk2_preload<LA, LB, LC, LD>();
    static_assert(IJN * KSUB == 64, "bra tile pair must divide a wavefront");

    const int item = blockIdx.x;
    if (item >= a.nitems) return;
    const int4 it = a.items[item];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jj = lane & (TJ - 1), ii = (lane >> JBITS) & (TI - 1), ks = lane >> (JBITS + IBITS);
    const int fI = a.tile_first[it.x], nI = a.tile_size[it.x];
    const int fJ = a.tile_first[it.y], nJ = a.tile_size[it.y];
    const int fK = a.tile_first[it.z], nK = a.tile_size[it.z];
    const int fL = a.tile_first[it.w], nL = a.tile_size[it.w];
    const bool sameBra = it.x == it.y, sameKet = it.z == it.w, samePair = it.x == it.z && it.y == it.w;
    const int N = a.N;
    const double *__restrict__ D = a.DJ;     // J and K use one density here
    const double *__restrict__ DK = a.DK0;
    const size_t xoff = (size_t)xcd_slot() * N * N;
    double *const Jx = a.J + xoff;
    double *const K0x = a.K0 + xoff;

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
    const int offI = a.shell_off[I], offJ = a.shell_off[J];
    const int origI = a.shell_orig[Is], origJ = a.shell_orig[Js];

    double jab[NA * NB];
#pragma unroll
    for (int i = 0; i < NA * NB; ++i) jab[i] = 0.0;
    double Dab[PRE ? NA * NB : 1];
    if constexpr (PRE) {
#pragma unroll
        for (int x = 0; x < NA; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y) Dab[x * NB + y] = D[(size_t)(offI + x) * N + offJ + y];
    }

    for (int kq = wave; kq < TK; kq += 4) {
        if (kq >= nK) break;  // uniform
        const int K = fK + kq;
        const int offK = a.shell_off[K];
        double kac[NA * NC], kbc[NB * NC];
#pragma unroll
        for (int i = 0; i < NA * NC; ++i) kac[i] = 0.0;
#pragma unroll
        for (int i = 0; i < NB * NC; ++i) kbc[i] = 0.0;
        for (int lc = 0; lc < NLSTEP; ++lc) {
            const int lq = lc * KSUB + ks;
            const bool lexist = lq < nL;
            const int L = fL + (lexist ? lq : 0);
            const bool kvalid = lexist && (!sameKet || K >= L);
            bool valid = bvalid && kvalid;
            if (samePair) valid = valid && (I > K || (I == K && J >= L));
            const int Ls = kvalid ? L : (sameKet ? K : fL);  // a valid partner of K for loads
            const int Kp = K >= Ls ? K : Ls, Lp = K >= Ls ? Ls : K;
            const int pk = Kp * (Kp + 1) / 2 + Lp;
            const double qket = a.pair_q[pk];
            valid = valid && (qbra * qket >= a.tau);
            if (a.start_shell > 0) {
                int mo = max(max(origI, origJ), max(a.shell_orig[K], a.shell_orig[Ls]));
                valid = valid && (mo >= a.start_shell);
            }
            if (!__any(valid)) continue;  // uniform
            double wdeg = 1.0;
            if (I == J) wdeg *= 0.5;
            if (K == L) wdeg *= 0.5;
            if (I == K && J == L) wdeg *= 0.5;
            const R scale = valid ? R(wdeg) : R(0);
            const PairRef kref = a.pairs[pk];
            PairGeom<R> kg;
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                kg.A[x] = R(a.shell_xyz[3 * K + x]);
                kg.AB[x] = R(a.shell_xyz[3 * K + x] - a.shell_xyz[3 * Ls + x]);
            }
            const int offL = a.shell_off[L];
            double jcd[NC * ND], kad[NA * ND], kbd[NB * ND];
#pragma unroll
            for (int i = 0; i < NC * ND; ++i) jcd[i] = 0.0;
#pragma unroll
            for (int i = 0; i < NA * ND; ++i) kad[i] = 0.0;
#pragma unroll
            for (int i = 0; i < NB * ND; ++i) kbd[i] = 0.0;
            // density blocks of this step (rows uniform across the ket
            // group, columns along the lane's bra shells: D is symmetric)
            double Dcd[PRE ? NC * ND : 1], Dbd[PRE ? NB * ND : 1], Dbc[PRE ? NB * NC : 1];
            double Dad[PRE ? NA * ND : 1], Dac[PRE ? NA * NC : 1];
            if constexpr (PRE) {
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < ND; ++d) Dcd[c * ND + d] = D[(size_t)(offK + c) * N + offL + d];
#pragma unroll
                for (int d = 0; d < ND; ++d) {
#pragma unroll
                    for (int b = 0; b < NB; ++b) Dbd[b * ND + d] = DK[(size_t)(offL + d) * N + offJ + b];
#pragma unroll
                    for (int x = 0; x < NA; ++x) Dad[x * ND + d] = DK[(size_t)(offL + d) * N + offI + x];
                }
#pragma unroll
                for (int c = 0; c < NC; ++c) {
#pragma unroll
                    for (int b = 0; b < NB; ++b) Dbc[b * NC + c] = DK[(size_t)(offK + c) * N + offJ + b];
#pragma unroll
                    for (int x = 0; x < NA; ++x) Dac[x * NC + c] = DK[(size_t)(offK + c) * N + offI + x];
                }
            }
            int nbp = bref.npp, nkp = kref.npp;
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                nbp = max(nbp, __shfl_xor(nbp, 1 << b, 64));
                nkp = max(nkp, __shfl_xor(nkp, 1 << b, 64));
            }
            for (int ib = 0; ib < nbp; ++ib) {
                const PrimPair bp = a.pp[bref.pp0 + min(ib, bref.npp - 1)];
                for (int ik = 0; ik < nkp; ++ik) {
                    const PrimPair kp = a.pp[kref.pp0 + min(ik, kref.npp - 1)];
                    const R sc = (ib < bref.npp && ik < kref.npp) ? scale : R(0);
                    primitive_quartet<R, LA, LB, LC, LD>(
                        bp, kp, bg, kg, a.rys, sc, [&](auto ia, auto ib_, auto ic, auto id, R vr) {
                            constexpr int A_ = decltype(ia)::value, B_ = decltype(ib_)::value;
                            constexpr int C_ = decltype(ic)::value, D_ = decltype(id)::value;
                            const double v = double(vr);
                            if constexpr (PRE) {
                                jab[A_ * NB + B_] += v * Dcd[C_ * ND + D_];
                                jcd[C_ * ND + D_] += v * Dab[A_ * NB + B_];
                                kac[A_ * NC + C_] += v * Dbd[B_ * ND + D_];
                                kad[A_ * ND + D_] += v * Dbc[B_ * NC + C_];
                                kbc[B_ * NC + C_] += v * Dad[A_ * ND + D_];
                                kbd[B_ * ND + D_] += v * Dac[A_ * NC + C_];
                            } else {
                                const size_t ra = offI + A_, rb = offJ + B_, rc = offK + C_, rd = offL + D_;
                                jab[A_ * NB + B_] += v * D[rc * N + rd];
                                jcd[C_ * ND + D_] += v * D[ra * N + rb];
                                kac[A_ * NC + C_] += v * DK[rd * N + rb];
                                kad[A_ * ND + D_] += v * DK[rc * N + rb];
                                kbc[B_ * NC + C_] += v * DK[rd * N + ra];
                                kbd[B_ * ND + D_] += v * DK[rc * N + ra];
                            }
                        });
                }
            }
            // per-step flushes: J'(cd), K'(ad), K'(bd)
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    double t = lane_reduce<MASK_I | MASK_J>(jcd[c * ND + d]);
                    if ((lane & (MASK_I | MASK_J)) == 0 && lexist)
                        atomic_add(&Jx[(size_t)(offK + c) * N + offL + d], 2.0 * t);
                }
#pragma unroll
            for (int x = 0; x < NA; ++x)
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    double t = lane_reduce<MASK_J>(kad[x * ND + d]);
                    if ((lane & MASK_J) == 0 && iexist && lexist) atomic_add(&K0x[(size_t)(offI + x) * N + offL + d], t);
                }
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    double t = lane_reduce<MASK_I>(kbd[b * ND + d]);
                    if ((lane & MASK_I) == 0 && jexist && lexist) atomic_add(&K0x[(size_t)(offJ + b) * N + offL + d], t);
                }
        }
        // per-K flushes: K'(ac), K'(bc)
#pragma unroll
        for (int x = 0; x < NA; ++x)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                double t = lane_reduce<MASK_J | MASK_S>(kac[x * NC + c]);
                if ((lane & (MASK_J | MASK_S)) == 0 && iexist) atomic_add(&K0x[(size_t)(offI + x) * N + offK + c], t);
            }
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                double t = lane_reduce<MASK_I | MASK_S>(kbc[b * NC + c]);
                if ((lane & (MASK_I | MASK_S)) == 0 && jexist) atomic_add(&K0x[(size_t)(offJ + b) * N + offK + c], t);
            }
    }
    // per-item flush: J'(ab)
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            double t = lane_reduce<MASK_S>(jab[x * NB + b]);
            if ((lane & MASK_S) == 0 && iexist && jexist) atomic_add(&Jx[(size_t)(offI + x) * N + offJ + b], 2.0 * t);
        }
}

// ---------------------------------------------------------------------------
// ERI export: all Cartesian components of listed shell quartets (tests).
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void eri_out_kernel(EriOutArgs a)
{
    constexpr int NCOMP = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.nq) return;
    const int I = a.quartets[4 * q], J = a.quartets[4 * q + 1], K = a.quartets[4 * q + 2], L = a.quartets[4 * q + 3];
    double *out = a.out + (size_t)q * NCOMP;
    for (int i = 0; i < NCOMP; ++i) out[i] = 0.0;
    const int Ip = I >= J ? I : J, Jp = I >= J ? J : I, Kp = K >= L ? K : L, Lp = K >= L ? L : K;
    const PairRef bref = a.pairs[Ip * (Ip + 1) / 2 + Jp], kref = a.pairs[Kp * (Kp + 1) / 2 + Lp];
    PairGeom<double> bg, kg;
    for (int x = 0; x < 3; ++x) {
        bg.A[x] = a.shell_xyz[3 * I + x];
        bg.AB[x] = a.shell_xyz[3 * I + x] - a.shell_xyz[3 * J + x];
        kg.A[x] = a.shell_xyz[3 * K + x];
        kg.AB[x] = a.shell_xyz[3 * K + x] - a.shell_xyz[3 * L + x];
    }
    for (int ib = 0; ib < bref.npp; ++ib) {
        PrimPair bp = a.pp[bref.pp0 + ib];
        for (int ik = 0; ik < kref.npp; ++ik) {
            PrimPair kp = a.pp[kref.pp0 + ik];
            if constexpr (!HeavyClass<LA, LB, LC, LD>::value) {
                primitive_quartet<double, LA, LB, LC, LD>(bp, kp, bg, kg, a.rys, 1.0,
                                                          [&](auto ia, auto ib_, auto ic, auto id, double v) {
                                                              constexpr int idx =
                                                                  ((decltype(ia)::value * ncart(LB) + decltype(ib_)::value) * ncart(LC) +
                                                                   decltype(ic)::value) * ncart(LD) + decltype(id)::value;
                                                              out[idx] += v;
                                                          });
            } else {
                OneDTable<double, LA, LB, LC, LD> tbl;
                tbl.build(bp, kp, bg, kg, a.rys, 1.0);
                for (int ia = 0; ia < ncart(LA); ++ia)
                    for (int ib2 = 0; ib2 < ncart(LB); ++ib2) {
                        const double nab = cart_norm_rt<LA>(ia) * cart_norm_rt<LB>(ib2);
                        static_for<ncart(LC)>([&](auto ic) {
                            static_for<ncart(LD)>([&](auto id) {
                                constexpr int IC = decltype(ic)::value, ID = decltype(id)::value;
                                constexpr double ncd = 1.0 / csqrt(cart_dfprod(LC, IC) * cart_dfprod(LD, ID));
                                out[((ia * ncart(LB) + ib2) * ncart(LC) + IC) * ncart(LD) + ID] +=
                                    tbl.template value<IC, ID>(ia, ib2) * nab * ncd;
                            });
                        });
                    }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Schwarz bounds: Q_IJ = sqrt(max_ab |(ab|ab)|) for every sorted pair I >= J.
// ---------------------------------------------------------------------------
template <int LA, int LB>
__global__ __launch_bounds__(64) void schwarz_kernel(SchwarzArgs a)
{
    constexpr int NAB = ncart(LA) * ncart(LB);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.npairs) return;
    const int pidx = a.pair_list[t];
    const int I = a.pair_I[t], J = a.pair_J[t];
    const PairRef ref = a.pairs[pidx];
    PairGeom<double> g;
    for (int x = 0; x < 3; ++x) {
        g.A[x] = a.shell_xyz[3 * I + x];
        g.AB[x] = a.shell_xyz[3 * I + x] - a.shell_xyz[3 * J + x];
    }
    double acc[NAB];
#pragma unroll
    for (int i = 0; i < NAB; ++i) acc[i] = 0.0;
    for (int ib = 0; ib < ref.npp; ++ib) {
        PrimPair bp = a.pp[ref.pp0 + ib];
        for (int ik = 0; ik < ref.npp; ++ik) {
            PrimPair kp = a.pp[ref.pp0 + ik];
            if constexpr (!HeavyClass<LA, LB, LA, LB>::value) {
                primitive_quartet<double, LA, LB, LA, LB, false>(bp, kp, g, g, a.rys, 1.0,
                                                          [&](auto ia, auto ib_, auto ic, auto id, double v) {
                                                              constexpr int IA = decltype(ia)::value, IB = decltype(ib_)::value;
                                                              if constexpr (IA == decltype(ic)::value && IB == decltype(id)::value)
                                                                  acc[IA * ncart(LB) + IB] += v;
                                                          });
            } else {
                OneDTable<double, LA, LB, LA, LB> tbl;
                tbl.build(bp, kp, g, g, a.rys, 1.0, false);
                static_for<ncart(LA)>([&](auto ic) {
                    static_for<ncart(LB)>([&](auto id) {
                        constexpr int IC = decltype(ic)::value, ID = decltype(id)::value;
                        constexpr double nn = 1.0 / (cart_dfprod(LA, IC) * cart_dfprod(LB, ID));
                        acc[IC * ncart(LB) + ID] += tbl.template value<IC, ID>(IC, ID) * nn;
                    });
                });
            }
        }
    }
    double m = 0.0;
#pragma unroll
    for (int i = 0; i < NAB; ++i) m = fmax(m, fabs(acc[i]));
    a.q[pidx] = sqrt(m);
}

}  // namespace pm

#include "eri_kernel3.h"

namespace pm {

// ---------------------------------------------------------------------------
// launch helpers
template <int LA, int LB, int LC, int LD>
inline void launch_jk_t(const JKArgs &args, int prec, int nk, hipStream_t s)
{
    // One exchange density per launch: the two-density instantiation is not
    // built (UHF runs two single-density passes).
    if (args.nitems <= 0 || nk != 1) return;
    dim3 grid(args.nitems), block(256);
    // (dd|dd): the ket-pair-per-lane kernel with run-time component loops;
    // every other class runs the LDS-staged kernel
#ifdef PM_DDDD_K3
    constexpr bool dddd_v1 = false;
#else
    constexpr bool dddd_v1 = LA == 2 && LB == 2 && LC == 2 && LD == 2;
#endif
    if constexpr (dddd_v1) {
        if (prec == 0)
            hipLaunchKernelGGL((jk_kernel<LA, LB, LC, LD, double, 1>), grid, block, 0, s, args);
        else
            hipLaunchKernelGGL((jk_kernel<LA, LB, LC, LD, float, 1>), grid, block, 0, s, args);
    } else {
#ifdef PM_USE_K2
        if (prec == 0)
            hipLaunchKernelGGL((jk_kernel2<LA, LB, LC, LD, double>), grid, block, 0, s, args);
        else
            hipLaunchKernelGGL((jk_kernel2<LA, LB, LC, LD, float>), grid, block, 0, s, args);
#else
        // jk_kernel3 may take several workgroups per item (jk3_wg_per_item)
        const dim3 grid3((unsigned)args.nitems * (unsigned)jk3_wg_per_item<LA, LB, LC, LD>());
        if (prec == 0)
            hipLaunchKernelGGL((jk_kernel3<LA, LB, LC, LD, double>), grid3, block, 0, s, args);
        else
            hipLaunchKernelGGL((jk_kernel3<LA, LB, LC, LD, float>), grid3, block, 0, s, args);
#endif
    }
}

template <int LA, int LB, int LC, int LD>
inline void launch_out_t(const EriOutArgs &args, hipStream_t s)
{
    if (args.nq <= 0) return;
    hipLaunchKernelGGL((eri_out_kernel<LA, LB, LC, LD>), dim3((args.nq + 63) / 64), dim3(64), 0, s, args);
}

template <int LA, int LB>
inline void launch_schwarz_t(const SchwarzArgs &args, hipStream_t s)
{
    if (args.npairs <= 0) return;
    hipLaunchKernelGGL((schwarz_kernel<LA, LB>), dim3((args.npairs + 63) / 64), dim3(64), 0, s, args);
}

}  // namespace pm
