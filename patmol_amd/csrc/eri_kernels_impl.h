// eri_kernels_impl.h -- the (dd|dd) J/K kernel, ERI export and Schwarz
// kernels, the shared cross-lane and fixed-point helpers, and the per-class
// launchers (jk_kernel3, the LDS-staged J/K kernel of every other class, is
// in eri_kernel3.h).
//
// Work decomposition (replaces the reference's stored integral list and its
// per-integral scatter, twoints.c:228-404 + scfaux.c:428-552):
//   shells are sorted by (l, exponent, centre) and cut into tiles of T(l) shells; a
//   work item is a canonical quartet of tiles (bra tile pair >= ket tile pair)
//   and one workgroup (4 wavefronts) processes every shell quartet in it.
// J/K targets, with w the degeneracy weight 2^-([I==J]+[K==L]+[IJ==KL])
// folded into the integral v:
//   J'(ab) += 2 v D(cd)   J'(cd) += 2 v D(ab)
//   K'(ac) += v D(bd)     K'(ad) += v D(bc)   K'(bc) += v D(ad)   K'(bd) += v D(ac)
// and J = J' + J'^T, K = K' + K'^T afterwards.  J' and K' are fixed-point
// sums (FxAcc), so their values do not depend on the order of the adds.
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

template <int K>
__device__ __forceinline__ float dpp_row_shl(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x100 | (1 << K), 0xf, 0xf, false));
}

// Lane i + 16 for lanes in even rows (v_permlane16_swap), lane i + 32 for
// lanes 0..31 (v_permlane32_swap).
__device__ __forceinline__ float lane_plus16(float v)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_int(v), __float_as_int(v), false, false);
    return __int_as_float(r[1]);
}
__device__ __forceinline__ float lane_plus32(float v)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(v), __float_as_int(v), false, false);
    return __int_as_float(r[1]);
}
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
template <int MASK, typename T, int N>
__device__ __forceinline__ void lane_reduce_n(T (&v)[N])
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
// The low word is first brought into [-2^39, 2^39) (its excess carried into
// the high word) so that the global low words stay far from overflow.
__device__ __forceinline__ void fx_flush(FxAcc *p, FxAcc v)
{
    const long long carry = (v.lo + (1LL << 39)) >> 40;
    const long long lo = v.lo - carry * (1LL << 40), hi = v.hi + carry;
    if (hi != 0) __hip_atomic_fetch_add(&p->hi, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (lo != 0) __hip_atomic_fetch_add(&p->lo, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// one value straight into this XCD's global copy
__device__ __forceinline__ void fx_add_global(FxAcc *p, double v)
{
    if (v == 0.0) return;
    const FxAcc s = fx_split(v);
    if (s.hi != 0) __hip_atomic_fetch_add(&p->hi, s.hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (s.lo != 0) __hip_atomic_fetch_add(&p->lo, s.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// XCD this workgroup runs on (hardware register XCC_ID).
__device__ __forceinline__ int xcd_slot()
{
    int x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(x));
    return x & (PM_XCD_SLOTS - 1);
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

// (dd|dd) J/K kernel (ket shell pair fixed per lane).  Lanes of a
// wavefront hold fixed ket shell pairs (K,L) laid out as
// lane = (js*TK + kk)*TL + ll; the wavefront steps over bra shells I and
// takes JSUB bra shells J per step.  Each (bra, ket) lane evaluates one
// primitive quartet after another; the per-primitive-quartet 1-D integrals
// are tabulated once (OneDTable) and the bra components looped over at run
// time, which keeps the code size of the 1296-component class bounded.  The
// integrals are contracted with the density at once and the per-block sums
// reduced across the lanes that share a target (XOR butterflies over the lane
// bits the target does not depend on) and added by one leader lane into
// this XCD's fixed-point copy.  One wave per SIMD (the whole register file).
template <int LA, int LB, int LC, int LD, typename R>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1))) void jk_kernel(JKArgs a)
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
    static_assert(HeavyClass<LA, LB, LC, LD>::value, "jk_kernel runs the tabulated (heavy) classes only");

    const int item = blockIdx.x;
    if (item >= a.nitems) return;
    if (a.mol_active && !a.mol_active[a.item_mol[item]]) return;   // whole workgroup
    const int4 it = a.items[item];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ll = lane & (TL - 1), kk = (lane >> LBITS) & (TK - 1), js = lane >> (LBITS + KBITS);
    const int fI = a.tile_first[it.x], nI = a.tile_size[it.x];
    const int fJ = a.tile_first[it.y], nJ = a.tile_size[it.y];
    const int fK = a.tile_first[it.z], nK = a.tile_size[it.z];
    const int fL = a.tile_first[it.w], nL = a.tile_size[it.w];
    const bool sameBra = it.x == it.y, sameKet = it.z == it.w, samePair = it.x == it.z && it.y == it.w;
    int N = a.N, sb = 0;
    long long pbase = 0, dadj = 0;
    if (a.item_mol) {
        const MolDesc m = a.mols[a.item_mol[item]];
        N = m.N;
        sb = m.sbase;
        pbase = m.pbase;
        dadj = m.dadj;
    }
    auto pidx = [&](int hi, int lo) -> long long {
        const long long h = hi - sb, l = lo - sb;
        return pbase + h * (h + 1) / 2 + l;
    };
    const long long xoff = (long long)xcd_slot() * a.slot_stride + dadj;
    FxAcc *const Jx = a.J + xoff;
    FxAcc *const Kx = a.K + xoff;

    // ---- ket (fixed per lane) ----
    const bool kexist = kk < nK, lexist = ll < nL;
    const int K = fK + (kexist ? kk : 0), L = fL + (lexist ? ll : 0);
    const bool kvalid = kexist && lexist && (!sameKet || K >= L);
    const int Ks = kvalid ? K : fK, Ls = kvalid ? L : fL;  // safe indices
    const int Kp = Ks >= Ls ? Ks : Ls, Lp = Ks >= Ls ? Ls : Ks;
    const long long pk = pidx(Kp, Lp);
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
        const long long pb = pidx(Ip, Jp);
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
        const double *__restrict__ DJ = a.DJ + dadj;
        const double *__restrict__ DK = a.DK + dadj;

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
                OneDTable<R, LA, LB, LC, LD> tbl;
                tbl.build(bp, kp, bg, kg, a.rys, sc);
                double kac[NC], kad[ND];
#pragma unroll
                for (int c = 0; c < NC; ++c) kac[c] = 0.0;
#pragma unroll
                for (int d = 0; d < ND; ++d) kad[d] = 0.0;
#pragma unroll 1
                for (int ia = 0; ia < NA; ++ia) {
#pragma unroll 1
                    for (int ib2 = 0; ib2 < NB; ++ib2) {
                        const double nab = cart_norm_rt<LA>(ia) * cart_norm_rt<LB>(ib2);
                        const int ra = offI + ia, rb = offJ + ib2;
                        double jab = 0.0, kbc[NC], kbd[ND];
#pragma unroll
                        for (int c = 0; c < NC; ++c) kbc[c] = 0.0;
#pragma unroll
                        for (int d = 0; d < ND; ++d) kbd[d] = 0.0;
                        static_for<NC>([&](auto ic) {
                            static_for<ND>([&](auto id) {
                                constexpr int IC = decltype(ic)::value, ID = decltype(id)::value;
                                constexpr double ncd = 1.0 / csqrt(cart_dfprod(LC, IC) * cart_dfprod(LD, ID));
                                const double v = double(tbl.template value<IC, ID>(ia, ib2)) * nab * ncd;
                                const int rc = offK + IC, rd = offL + ID;
                                jab += v * DJ[(size_t)rc * N + rd];
                                jcd[IC * ND + ID] += v * DJ[(size_t)ra * N + rb];
                                kac[IC] += v * DK[(size_t)rb * N + rd];
                                kad[ID] += v * DK[(size_t)rb * N + rc];
                                kbc[IC] += v * DK[(size_t)ra * N + rd];
                                kbd[ID] += v * DK[(size_t)ra * N + rc];
                            });
                        });
                        // end of the (a,b) block: J'(ab), K'(bc), K'(bd)
                        const double t = lane_reduce<MASK_K | MASK_L>(jab);
                        if ((lane & (MASK_K | MASK_L)) == 0 && jexist) fx_add_global(&Jx[(size_t)ra * N + rb], 2.0 * t);
#pragma unroll
                        for (int c = 0; c < NC; ++c) {
                            const double u = lane_reduce<MASK_L>(kbc[c]);
                            if ((lane & MASK_L) == 0 && jexist && kexist) fx_add_global(&Kx[(size_t)rb * N + offK + c], u);
                        }
#pragma unroll
                        for (int d = 0; d < ND; ++d) {
                            const double u = lane_reduce<MASK_K>(kbd[d]);
                            if ((lane & MASK_K) == 0 && jexist && lexist) fx_add_global(&Kx[(size_t)rb * N + offL + d], u);
                        }
                    }
                    // end of the a row: K'(ac), K'(ad)
                    const int ra = offI + ia;
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const double u = lane_reduce<MASK_L | MASK_J>(kac[c]);
                        if ((lane & (MASK_L | MASK_J)) == 0 && kexist) fx_add_global(&Kx[(size_t)ra * N + offK + c], u);
                        kac[c] = 0.0;
                    }
#pragma unroll
                    for (int d = 0; d < ND; ++d) {
                        const double u = lane_reduce<MASK_K | MASK_J>(kad[d]);
                        if ((lane & (MASK_K | MASK_J)) == 0 && lexist) fx_add_global(&Kx[(size_t)ra * N + offL + d], u);
                        kad[d] = 0.0;
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
            const double t = lane_reduce<MASK_J>(jcd[c * ND + d]);
            if ((lane & MASK_J) == 0 && kexist && lexist) fx_add_global(&Jx[(size_t)(offK + c) * N + offL + d], 2.0 * t);
        }
}

// ---------------------------------------------------------------------------
// ERI export: all Cartesian components of listed shell quartets (tests).
// ---------------------------------------------------------------------------
// R = float: the FP32 arithmetic of the mixed-precision J/K kernels (for the
// per-class error of FP32 integrals), summed over primitives in FP64
template <int LA, int LB, int LC, int LD, typename R>
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
    PairGeom<R> bg, kg;
    for (int x = 0; x < 3; ++x) {
        bg.A[x] = R(a.shell_xyz[3 * I + x]);
        bg.AB[x] = R(a.shell_xyz[3 * I + x] - a.shell_xyz[3 * J + x]);
        kg.A[x] = R(a.shell_xyz[3 * K + x]);
        kg.AB[x] = R(a.shell_xyz[3 * K + x] - a.shell_xyz[3 * L + x]);
    }
    for (int ib = 0; ib < bref.npp; ++ib) {
        PrimPair bp = a.pp[bref.pp0 + ib];
        for (int ik = 0; ik < kref.npp; ++ik) {
            PrimPair kp = a.pp[kref.pp0 + ik];
            if constexpr (!HeavyClass<LA, LB, LC, LD>::value) {
                primitive_quartet<R, LA, LB, LC, LD>(bp, kp, bg, kg, a.rys, R(1),
                                                     [&](auto ia, auto ib_, auto ic, auto id, R v) {
                                                         constexpr int idx =
                                                             ((decltype(ia)::value * ncart(LB) + decltype(ib_)::value) * ncart(LC) +
                                                              decltype(ic)::value) * ncart(LD) + decltype(id)::value;
                                                         out[idx] += double(v);
                                                     });
            } else {
                OneDTable<R, LA, LB, LC, LD> tbl;
                tbl.build(bp, kp, bg, kg, a.rys, R(1));
                for (int ia = 0; ia < ncart(LA); ++ia)
                    for (int ib2 = 0; ib2 < ncart(LB); ++ib2) {
                        const double nab = cart_norm_rt<LA>(ia) * cart_norm_rt<LB>(ib2);
                        static_for<ncart(LC)>([&](auto ic) {
                            static_for<ncart(LD)>([&](auto id) {
                                constexpr int IC = decltype(ic)::value, ID = decltype(id)::value;
                                constexpr double ncd = 1.0 / csqrt(cart_dfprod(LC, IC) * cart_dfprod(LD, ID));
                                out[((ia * ncart(LB) + ib2) * ncart(LC) + IC) * ncart(LD) + ID] +=
                                    double(tbl.template value<IC, ID>(ia, ib2)) * nab * ncd;
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
#include "md_kernel.h"

namespace pm {

// ---------------------------------------------------------------------------
// launch helpers
template <int LA, int LB, int LC, int LD>
inline void launch_jk_t(const JKArgs &args, int prec, hipStream_t s)
{
    if (args.nitems <= 0) return;
    const dim3 block(256);
    if constexpr (MdClass<LA, LB, LC, LD>::value) {
        // FP64 items of the classes with the most components: the
        // McMurchie-Davidson kernel on the matrix cores, when the tables exist
        if (prec == 0 && args.md_slot) {
            hipLaunchKernelGGL((jk_md_kernel<LA, LB, LC, LD>), dim3((unsigned)args.nitems), block, 0, s, args);
            return;
        }
    }
    if constexpr (HeavyClass<LA, LB, LC, LD>::value && LC == 2 && LD == 2) {
        // (dd|dd): the ket-pair-per-lane kernel with run-time component loops
        const dim3 grid(args.nitems);
        if (prec == 0)
            hipLaunchKernelGGL((jk_kernel<LA, LB, LC, LD, double>), grid, block, 0, s, args);
        else
            hipLaunchKernelGGL((jk_kernel<LA, LB, LC, LD, float>), grid, block, 0, s, args);
    } else {
        // every other class: the LDS-staged kernel, possibly several
        // workgroups per item (jk3_wg_per_item)
        if (prec == 0)
            hipLaunchKernelGGL((jk_kernel3<LA, LB, LC, LD, double>),
                               dim3((unsigned)args.nitems * (unsigned)jk3_wg_per_item<LA, LB, LC, LD, double>()), block,
                               0, s, args);
        else
            hipLaunchKernelGGL((jk_kernel3<LA, LB, LC, LD, float>),
                               dim3((unsigned)args.nitems * (unsigned)jk3_wg_per_item<LA, LB, LC, LD, float>()), block,
                               0, s, args);
    }
}

template <int LA, int LB, int LC, int LD>
inline void launch_out_t(const EriOutArgs &args, hipStream_t s)
{
    if (args.nq <= 0) return;
    if (args.prec == 0)
        hipLaunchKernelGGL((eri_out_kernel<LA, LB, LC, LD, double>), dim3((args.nq + 63) / 64), dim3(64), 0, s, args);
    else
        hipLaunchKernelGGL((eri_out_kernel<LA, LB, LC, LD, float>), dim3((args.nq + 63) / 64), dim3(64), 0, s, args);
}

template <int LA, int LB>
inline void launch_schwarz_t(const SchwarzArgs &args, hipStream_t s)
{
    if (args.npairs <= 0) return;
    hipLaunchKernelGGL((schwarz_kernel<LA, LB>), dim3((args.npairs + 63) / 64), dim3(64), 0, s, args);
}

}  // namespace pm
