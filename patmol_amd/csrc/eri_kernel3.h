// eri_kernel3.h -- J/K kernel with the tile quartet staged in LDS.
//
// One workgroup per tile quartet; lane = (ket sub-slot, bra shell I, bra
// shell J), the bra pair fixed per lane for the whole item; wave w owns ket
// shells K = w, w+4, ...; each wave steps over L in chunks of KSUB.
// Everything a step reads or writes lives in LDS for the duration of the
// item:
//   * the six density blocks D(IJ), D(KL), D(IK), D(IL), D(JK), D(JL) of the
//     tile quartet (gathered once, coalesced),
//   * the ket pair table (pair reference, Schwarz bound, validity, centres)
//     and, when they fit, the ket primitive pairs,
//   * the six output blocks J'(IJ), J'(KL), K'(IK), K'(IL), K'(JK), K'(JL)
//     (FP64): cross-lane reductions land there, every wave that can write an
//     element writing its own copy of the block; at the end of the item the
//     copies are summed in copy order, converted to fixed point and moved by
//     one pass of global integer atomics into the per-XCD J'/K' copies.
//     Neither the order in which waves run nor the order of the atomics
//     changes the sums, so a G build is reproducible bit for bit.
// Classes whose 2-D tables for all roots would not fit the register file
// digest one Rys root at a time (primitive_quartet_rootwise); that choice, the
// density-read fence and the register budget per class are measured
// (jk3_code).
#pragma once
#include <type_traits>
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

// Per-class launch shape, one code per class and precision:
// code = rys_global * 256 + rys_lds * 128 + ksplit * 64 + (split_b - 1) * 32
// + (split_a - 1) * 16 + fence * 8 + waves * 2 + rootwise
// (rootwise: digest one Rys root at a time; fence: the density-read fence of
// the rootwise digestion; waves: waves per SIMD the register budget is set
// for; split_a / split_b: the bra components split over the waves; ksplit:
// every item is two workgroups, each taking half of the ket tile's K shells;
// rys_lds: the Rys coefficients of the class's root count are staged in LDS
// also above three roots; rys_global: they are read from global memory even
// at three roots or fewer, which keeps two workgroups' LDS within a CU's
// 160 KB).  The codes are the ones patmol_amd/autotune.py measured fastest
// per class, written to jk3_tuned.inc (found through the include path, so a
// tuning build can put its candidate table first).
template <int LA, int LB, int LC, int LD, typename R>
constexpr int jk3_code()
{
    const int t = jk3_tuned_shape(jk3_class_index(LA, LB, LC, LD), std::is_same<R, float>::value);
    return t >= 0 ? t : 1 * 2 + 1;
}

template <int LA, int LB, int LC, int LD, typename R>
constexpr bool jk3_rootwise()
{
    return jk3_code<LA, LB, LC, LD, R>() & 1;
}

// The Rys coefficient section of the class's root count is kept in LDS for
// the classes with up to three roots (7 / 17 / 30 KB) unless their shape
// says otherwise, and for the classes whose shape asks for it (4 roots:
// 47 KB)
template <int LA, int LB, int LC, int LD, typename R>
constexpr bool jk3_rys_lds()
{
    constexpr int code = jk3_code<LA, LB, LC, LD, R>();
    return ((LA + LB + LC + LD) / 2 + 1 <= 3 && !((code >> 8) & 1)) || ((code >> 7) & 1);
}

// Workgroups per work item (ksplit): the items of the largest classes take
// hundreds of microseconds each and there are only a few hundred of them,
// so halving them evens out the end of the launch.
template <int LA, int LB, int LC, int LD, typename R>
constexpr int jk3_wg_per_item()
{
    return ((jk3_code<LA, LB, LC, LD, R>() >> 6) & 1) ? 2 : 1;
}

template <int LA, int LB, int LC, int LD, typename R>
constexpr bool jk3_block_fence()
{
    return (jk3_code<LA, LB, LC, LD, R>() >> 3) & 1;
}

template <int LA, int LB, int LC, int LD, typename R>
constexpr int jk3_split_a()
{
    return ((jk3_code<LA, LB, LC, LD, R>() >> 4) & 1) + 1;
}

template <int LA, int LB, int LC, int LD, typename R>
constexpr int jk3_split_b()
{
    return ((jk3_code<LA, LB, LC, LD, R>() >> 5) & 1) + 1;
}

template <int LA, int LB, int LC, int LD, typename R>
constexpr int jk3_waves_per_eu()
{
    return (jk3_code<LA, LB, LC, LD, R>() >> 1) & 3;
}

// Cross-lane sums: lane bits 0-3 by DPP inside rows of 16; bits 4-5 (other
// rows) are left to LDS atomics from one leader per row, which costs less
// than the permlane swaps for the few values a step produces.
constexpr int red_lo(int mask) { return mask & 15; }
constexpr bool red_hi(int mask) { return (mask & ~red_lo(mask)) != 0; }

// v into an LDS output element: with ATOMIC lanes of this wave's instruction
// share the element (one LDS atomic instruction), otherwise this lane is its
// only writer
template <bool ATOMIC>
__device__ __forceinline__ void lds_acc(double *o, double v)
{
    if constexpr (ATOMIC) {
        if (v != 0.0) __hip_atomic_fetch_add(o, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        *o += v;
    }
}

template <int LA, int LB, int LC, int LD, typename R>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(jk3_waves_per_eu<LA, LB, LC, LD, R>()))) void jk_kernel3(JKArgs a)
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

    // Bra-component split: SA x SB waves share every K, each taking a block
    // of NAH bra-a components x NBH bra-b components of every quartet (and
    // only that block's accumulators); without a split the four waves own
    // different K.  Waves w with the same w / SPLIT (the K group) share K.
    constexpr int SA = jk3_split_a<LA, LB, LC, LD, R>(), SB = jk3_split_b<LA, LB, LC, LD, R>();
    constexpr int SPLIT = SA * SB, KSTRIDE = 4 / SPLIT;
    constexpr int NAH = NA / SA, NBH = NB / SB;
    static_assert(NA % SA == 0 && NB % SB == 0 && 4 % SPLIT == 0, "bra component split");

    // block rows padded to an odd number of doubles so that rows of
    // different bra shells fall into different LDS banks
    constexpr int SJ = FJ | 1, SK = FK | 1, SL = FL | 1;
    constexpr int ZIJ = FI * SJ, ZKL = FK * SL, ZIK = FI * SK, ZIL = FI * SL, ZJK = FJ * SK, ZJL = FJ * SL;
    // Contraction type: the FP32 instantiation contracts its FP32 integrals
    // with an FP32 copy of the density blocks and sums a step's products
    // across lanes in FP32; the sums land in the FP64 output blocks, and the
    // per-item J'(IJ) sum is FP64 in both (the FP32 part of it is one K
    // shell's products).
    using A = R;
    __shared__ A sD[ZIJ + ZKL + ZIK + ZIL + ZJK + ZJL];
    // Output blocks.  An element that several waves can write gets one copy
    // per writing wave: J'(KL) one per wave of a K group (SPLIT), K'(IK) per
    // b block (SB), K'(JK) per a block (SA), K'(IL) per (K group, b block),
    // K'(JL) per (K group, a block).  J'(IJ) has one copy; the K groups add
    // into it one after the other at the end of the item.  Lanes of one wave
    // that share an element add with one LDS atomic instruction, whose lanes
    // the LDS applies in a fixed order.
    constexpr int CKL = SPLIT, CIK = SB, CJK = SA, CIL = KSTRIDE * SB, CJL = KSTRIDE * SA;
    constexpr int pIJ = 0, pKL = pIJ + ZIJ, pIK = pKL + CKL * ZKL, pIL = pIK + CIK * ZIK, pJK = pIL + CIL * ZIL,
                  pJL = pJK + CJK * ZJK, pEND = pJL + CJL * ZJL;
    __shared__ double sO[pEND];
    __shared__ int sOff[TI + TJ + TK + TL];
    __shared__ double sXK[TK][3], sXL[TL][3];
    __shared__ PairRef sKref[KLP];
    __shared__ double sKq[KLP];
    __shared__ int sKok[KLP];
    __shared__ int sPPoff[KLP];
    __shared__ PrimPair sPP[S::PPCAP];
    __shared__ int sStaged, sKmax;
    // The Rys coefficient section of this class's root count kept in LDS
    // (the lookups are otherwise global-memory latency on the critical path
    // of every step)
    constexpr bool RYS_LDS = std::is_same<R, double>::value && jk3_rys_lds<LA, LB, LC, LD, R>();
    // only the NR-root section of the table (rys_eval<NR, true>)
    constexpr int RYS_LDS_N = RYS_LDS ? rys_tmax(S::NR) * rys_stride(S::NR) : 2;
    __shared__ __attribute__((aligned(16))) double sRys[RYS_LDS_N];
    // density block offsets inside sD
    constexpr int dIJ = 0, dKL = dIJ + ZIJ, dIK = dKL + ZKL, dIL = dIK + ZIK, dJK = dIL + ZIL, dJL = dJK + ZJK;

    // with KS workgroups per item, workgroups KS i .. KS i + KS - 1 share
    // item i and each takes one range of the K shells of the ket tile
    constexpr int KS = jk3_wg_per_item<LA, LB, LC, LD, R>();
    const int item = (int)(blockIdx.x / (unsigned)KS), ksub = (int)(blockIdx.x % (unsigned)KS);
    if (item >= a.nitems) return;
    if (a.mol_active && !a.mol_active[a.item_mol[item]]) return;   // whole workgroup
    const int4 it = a.items[item];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int jj = lane & (TJ - 1), ii = (lane >> JBITS) & (TI - 1), ks = lane >> (JBITS + IBITS);
    const int fI = a.tile_first[it.x], nI = a.tile_size[it.x];
    const int fJ = a.tile_first[it.y], nJ = a.tile_size[it.y];
    const int fK = a.tile_first[it.z], nK = a.tile_size[it.z];
    const int fL = a.tile_first[it.w], nL = a.tile_size[it.w];
    const bool sameBra = it.x == it.y, sameKet = it.z == it.w, samePair = it.x == it.z && it.y == it.w;
    // the item's molecule (batches) or the single basis
    int N = a.N, sb = 0;
    long long pbase = 0, dadj = 0;
    if (a.item_mol) {
        const MolDesc m = a.mols[a.item_mol[item]];
        N = m.N;
        sb = m.sbase;
        pbase = m.pbase;
        dadj = m.dadj;
    }
    // pair table index of sorted shells hi >= lo
    auto pidx = [&](int hi, int lo) -> long long {
        const long long h = hi - sb, l = lo - sb;
        return pbase + h * (h + 1) / 2 + l;
    };

    // ---- stage A: function offsets, centres, ket pairs, Rys table ----
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
    if (wave == 2) {
        // one lane per ket pair: reference, bound, validity, and (when the
        // primitive pairs of the whole ket tile pair fit) the primitive pairs
        const int t = lane, kq = t / TL, lq = t % TL;
        const bool here = t < KLP;
        const bool kex = kq < nK, lex = lq < nL;
        const int K = fK + (kex ? kq : 0), L = fL + (lex ? lq : 0);
        const bool ok = here && kex && lex && (!sameKet || K >= L);
        // a structurally valid stand-in pair for lanes whose (K, L) does not exist
        const int Ls = ok ? L : (sameKet ? K : fL);
        const int Kp = K >= Ls ? K : Ls, Lp = K >= Ls ? Ls : K;
        const long long pk = pidx(Kp, Lp);
        PairRef r = {0, 0};
        if (here) {
            r = a.pairs[pk];
            sKref[t] = r;
            sKq[t] = a.pair_q[pk];
            sKok[t] = ok ? 1 : 0;
        }
        // exclusive prefix sum and maximum of the primitive counts; with one
        // primitive pair per ket pair (uncontracted shells) they are known
        int off, tot, mx;
        if (__all(!here || r.npp == 1)) {
            off = t;
            tot = KLP;
            mx = 1;
        } else {
            int incl = r.npp;
            mx = r.npp;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int u = __shfl_up(incl, d, 64);
                if (lane >= d) incl += u;
                mx = max(mx, __shfl_xor(mx, d, 64));
            }
            tot = __shfl(incl, 63, 64);
            off = incl - r.npp;
        }
        const bool staged_w = tot <= S::PPCAP;
        if (here) {
            sPPoff[t] = off;
            if (staged_w)
                for (int p = 0; p < r.npp; ++p) sPP[off + p] = a.pp[r.pp0 + p];
        }
        if (lane == 0) {
            sStaged = staged_w ? 1 : 0;
            sKmax = max(mx, 1);
        }
    }
    if constexpr (RYS_LDS) {
        const double2 *__restrict__ src = reinterpret_cast<const double2 *>(a.rys.cheb + rys_off(S::NR));
        double2 *dst = reinterpret_cast<double2 *>(sRys);
        for (int e = tid; e < RYS_LDS_N / 2; e += 256) dst[e] = src[e];
    }
    RysTables rt = a.rys;
    if constexpr (RYS_LDS) rt.cheb = sRys;
    __syncthreads();
    // ---- stage B: density blocks of the tile quartet, zeroed output blocks ----
    {
        // one block: rows of tile (rsh: first entry in sOff, rn shells, nrc
        // components each), columns of tile (csh, cn, ncc)
        auto gather = [&](int o, int rows, int stride, int rsh, int rn, int nrc, int csh, int cn, int ncc,
                          const double *__restrict__ src) {
            for (int e = tid; e < rows * stride; e += 256) {
                const int r = e / stride, c = e % stride;
                const int rs = r / nrc, cs = c / ncc;
                const bool ex = c < cn * ncc && rs < rn && cs < cn;
                double v = 0.0;
                if (ex) v = src[dadj + (long long)(sOff[rsh + rs] + r % nrc) * N + sOff[csh + cs] + c % ncc];
                sD[o + e] = A(v);
            }
        };
        for (int e = tid; e < pEND; e += 256) sO[e] = 0.0;
        gather(dIJ, FI, SJ, 0, nI, NA, TI, nJ, NB, a.DJ);
        gather(dKL, FK, SL, TI + TJ, nK, NC, TI + TJ + TK, nL, ND, a.DJ);
        gather(dIK, FI, SK, 0, nI, NA, TI + TJ, nK, NC, a.DK);
        gather(dIL, FI, SL, 0, nI, NA, TI + TJ + TK, nL, ND, a.DK);
        gather(dJK, FJ, SK, TI, nJ, NB, TI + TJ, nK, NC, a.DK);
        gather(dJL, FJ, SL, TI, nJ, NB, TI + TJ + TK, nL, ND, a.DK);
    }
    __syncthreads();
    const bool staged = sStaged != 0;

    // ---- bra pair (fixed per lane) ----
    const bool iexist = ii < nI, jexist = jj < nJ;
    const int I = fI + (iexist ? ii : 0), J = fJ + (jexist ? jj : 0);
    const bool bvalid = iexist && jexist && (!sameBra || I >= J);
    const int Is = bvalid ? I : fI, Js = bvalid ? J : fJ;
    const int Ip = Is >= Js ? Is : Js, Jp = Is >= Js ? Js : Is;
    const long long pb = pidx(Ip, Jp);
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

    const int sub = wave % SPLIT, kgrp = wave / SPLIT;
    const int ha = sub % SA, hb = sub / SA;   // wave-uniform
    const int a0 = ha * NAH, b0 = hb * NBH;
    // this wave's copies of the shared output blocks
    double *const oKL = sO + pKL + sub * ZKL;
    double *const oIK = sO + pIK + hb * ZIK;
    double *const oJK = sO + pJK + ha * ZJK;
    double *const oIL = sO + pIL + (kgrp * SB + hb) * ZIL;
    double *const oJL = sO + pJL + (kgrp * SA + ha) * ZJL;

    double jab[NAH * NBH];
#pragma unroll
    for (int i = 0; i < NAH * NBH; ++i) jab[i] = 0.0;
    const A *__restrict__ DIJ = sD + dIJ + (ii * NA) * SJ + jj * NB;

    constexpr int KPER = (TK + KS - 1) / KS;
    const int kq_end = min(TK, (ksub + 1) * KPER);
    for (int kq = ksub * KPER + wave / SPLIT; kq < kq_end; kq += KSTRIDE) {
        if (kq >= nK) break;  // uniform
        const int K = fK + kq;
        // FP32: this K shell's J'(IJ) products, added to the FP64 jab after
        // the K shell (FP64 adds straight into jab)
        constexpr bool A64 = std::is_same<A, double>::value;
        constexpr int NJK = A64 ? 1 : NAH * NBH;
        A kac[NAH * NC], kbc[NBH * NC], jabk[NJK];
#pragma unroll
        for (int i = 0; i < NAH * NC; ++i) kac[i] = A(0);
#pragma unroll
        for (int i = 0; i < NBH * NC; ++i) kbc[i] = A(0);
#pragma unroll
        for (int i = 0; i < NJK; ++i) jabk[i] = A(0);
        for (int lc = 0; lc < NLSTEP; ++lc) {
            const int lq0 = lc * KSUB + ks;
            const bool lexist = lq0 < nL;
            const int lq = lexist ? lq0 : 0;
            const int t = kq * TL + lq;
            const int L = fL + lq;
            // validity from unconditional loads (t is always a staged slot)
            // and bitwise logic, so that it compiles to selects, not branches
            const int kok = sKok[t];
            const double qket = sKq[t];
            bool valid = bvalid & lexist & (kok != 0) & (qbra * qket >= a.tau);
            if (samePair) valid = valid & ((I > K) | ((I == K) & (J >= L)));
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
            const int lqs = (kok != 0) ? lq : (sameKet ? kq : 0);
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                kg.A[x] = R(sXK[kq][x]);
                kg.AB[x] = R(sXK[kq][x] - sXL[lqs][x]);
            }
            A jcd[NC * ND], kad[NAH * ND], kbd[NBH * ND];
#pragma unroll
            for (int i = 0; i < NC * ND; ++i) jcd[i] = A(0);
#pragma unroll
            for (int i = 0; i < NAH * ND; ++i) kad[i] = A(0);
#pragma unroll
            for (int i = 0; i < NBH * ND; ++i) kbd[i] = A(0);
            const A *__restrict__ DKL = sD + dKL + (kq * NC) * SL + lq * ND;
            const A *__restrict__ DJL = sD + dJL + (jj * NB) * SL + lq * ND;
            const A *__restrict__ DIL = sD + dIL + (ii * NA) * SL + lq * ND;
            const A *__restrict__ DJK = sD + dJK + (jj * NB) * SK + kq * NC;
            const A *__restrict__ DIK = sD + dIK + (ii * NA) * SK + kq * NC;
            const int nkp = sKmax;  // uniform bound of the ket primitive loop
            const PrimPair *kpp = staged ? (sPP + sPPoff[t]) : (a.pp + kref.pp0);
            // the primitive loops for one (a, b) component block
            auto run_block = [&](auto HA, auto HB) {
                constexpr int A0 = decltype(HA)::value * NAH, B0 = decltype(HB)::value * NBH;
                auto digest = [&](auto ia, auto ib_, auto ic, auto id, R vr) {
                    constexpr int A_ = decltype(ia)::value, B_ = decltype(ib_)::value;
                    constexpr int C_ = decltype(ic)::value, D_ = decltype(id)::value;
                    constexpr int Al = A_ - A0, Bl = B_ - B0;
                    const A v = A(vr);
                    if constexpr (A64) jab[Al * NBH + Bl] += v * DKL[C_ * SL + D_];
                    else jabk[Al * NBH + Bl] += v * DKL[C_ * SL + D_];
                    jcd[C_ * ND + D_] += v * DIJ[A_ * SJ + B_];
                    kac[Al * NC + C_] += v * DJL[B_ * SL + D_];
                    kad[Al * ND + D_] += v * DJK[B_ * SK + C_];
                    kbc[Bl * NC + C_] += v * DIL[A_ * SL + D_];
                    kbd[Bl * ND + D_] += v * DIK[A_ * SK + C_];
                };
                auto quartet = [&](const PrimPair &bp, const PrimPair &kp, R sc) {
                    if constexpr (jk3_rootwise<LA, LB, LC, LD, R>())
                        primitive_quartet_rootwise<R, LA, LB, LC, LD, jk3_block_fence<LA, LB, LC, LD, R>(), A0, A0 + NAH,
                                                   B0, B0 + NBH, RYS_LDS>(bp, kp, bg, kg, rt, sc, digest);
                    else
                        primitive_quartet<R, LA, LB, LC, LD, true, A0, A0 + NAH, B0, B0 + NBH, RYS_LDS>(bp, kp, bg, kg,
                                                                                                    rt, sc, digest);
                };
                if (nbp == 1 && staged) {
                    // uncontracted bra pairs and ket primitive pairs in LDS
                    // (the common case): no per-iteration selects of the bra
                    // pair, and LDS reads instead of generic-address loads
                    const PrimPair *__restrict__ kst = sPP + sPPoff[t];
                    for (int ik = 0; ik < nkp; ++ik) {
                        const PrimPair kp = kst[min(ik, kref.npp - 1)];
                        quartet(bp0, kp, ik < kref.npp ? scale : R(0));
                    }
                } else {
                    for (int ib = 0; ib < nbp; ++ib) {
                        const PrimPair bp = (ib == 0) ? bp0 : a.pp[bref.pp0 + min(ib, nbp_lane - 1)];
                        for (int ik = 0; ik < nkp; ++ik) {
                            const PrimPair kp = kpp[min(ik, kref.npp - 1)];
                            quartet(bp, kp, (ib < nbp_lane && ik < kref.npp) ? scale : R(0));
                        }
                    }
                }
            };
            using Z = std::integral_constant<int, 0>;
            using O = std::integral_constant<int, 1>;
            if constexpr (SPLIT == 1) {
                run_block(Z{}, Z{});
            } else if constexpr (SA == 2 && SB == 1) {
                if (ha == 0) run_block(Z{}, Z{});
                else run_block(O{}, Z{});
            } else if constexpr (SA == 1 && SB == 2) {
                if (hb == 0) run_block(Z{}, Z{});
                else run_block(Z{}, O{});
            } else {
                if (ha == 0 && hb == 0) run_block(Z{}, Z{});
                else if (ha == 1 && hb == 0) run_block(O{}, Z{});
                else if (ha == 0) run_block(Z{}, O{});
                else run_block(O{}, O{});
            }
            // per-step results: J'(KL) (owned by the waves of this K),
            // K'(IL), K'(JL) (shared by waves).  Lane bits 0-3 are summed
            // with DPP; lanes that differ in bits 4-5 add into LDS.
            lane_reduce_n<red_lo(MASK_I | MASK_J)>(jcd);
            lane_reduce_n<red_lo(MASK_J)>(kad);
            lane_reduce_n<red_lo(MASK_I)>(kbd);
            if ((lane & red_lo(MASK_I | MASK_J)) == 0 && lexist) {
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < ND; ++d) {
                        lds_acc<red_hi(MASK_I | MASK_J)>(&oKL[(kq * NC + c) * SL + lq * ND + d],
                                                          2.0 * double(jcd[c * ND + d]));
                    }
            }
            if ((lane & red_lo(MASK_J)) == 0 && iexist && lexist) {
#pragma unroll
                for (int x = 0; x < NAH; ++x)
#pragma unroll
                    for (int d = 0; d < ND; ++d)
                        lds_acc<red_hi(MASK_J)>(&oIL[(ii * NA + a0 + x) * SL + lq * ND + d], double(kad[x * ND + d]));
            }
            if ((lane & red_lo(MASK_I)) == 0 && jexist && lexist) {
#pragma unroll
                for (int b = 0; b < NBH; ++b)
#pragma unroll
                    for (int d = 0; d < ND; ++d)
                        lds_acc<red_hi(MASK_I)>(&oJL[(jj * NB + b0 + b) * SL + lq * ND + d], double(kbd[b * ND + d]));
            }
        }
        // per-K results: K'(IK), K'(JK) (owned by the waves of this K)
        lane_reduce_n<red_lo(MASK_J | MASK_S)>(kac);
        lane_reduce_n<red_lo(MASK_I | MASK_S)>(kbc);
        if ((lane & red_lo(MASK_J | MASK_S)) == 0 && iexist) {
#pragma unroll
            for (int x = 0; x < NAH; ++x)
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    lds_acc<red_hi(MASK_J | MASK_S)>(&oIK[(ii * NA + a0 + x) * SK + kq * NC + c], double(kac[x * NC + c]));
                }
        }
        if ((lane & red_lo(MASK_I | MASK_S)) == 0 && jexist) {
#pragma unroll
            for (int b = 0; b < NBH; ++b)
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    lds_acc<red_hi(MASK_I | MASK_S)>(&oJK[(jj * NB + b0 + b) * SK + kq * NC + c], double(kbc[b * NC + c]));
                }
        }
        if constexpr (!A64) {
#pragma unroll
            for (int i = 0; i < NAH * NBH; ++i) jab[i] += double(jabk[i]);
        }
    }
    // per-item result: J'(IJ), added by one K group after the other (the
    // waves of one K group hold disjoint (a, b) blocks)
    lane_reduce_n<red_lo(MASK_S)>(jab);
    for (int g = 0; g < KSTRIDE; ++g) {
        if (kgrp == g && (lane & red_lo(MASK_S)) == 0 && iexist && jexist) {
#pragma unroll
            for (int x = 0; x < NAH; ++x)
#pragma unroll
                for (int b = 0; b < NBH; ++b)
                    lds_acc<red_hi(MASK_S)>(&sO[pIJ + (ii * NA + a0 + x) * SJ + jj * NB + b0 + b], 2.0 * jab[x * NBH + b]);
        }
        __syncthreads();
    }

    // ---- flush: copies summed in copy order, fixed point, this XCD's J'/K' ----
    const long long xoff = (long long)xcd_slot() * a.slot_stride + dadj;
    FxAcc *const Jx = a.J + xoff;
    FxAcc *const Kx = a.K + xoff;
    auto flush = [&](int o, int copies, int size, int rows, int stride, int cols, int rsh, int nrc, int csh, int ncc,
                     FxAcc *dst) {
        for (int e = tid; e < rows * stride; e += 256) {
            const int r = e / stride, c = e % stride;
            if (c >= cols) continue;
            double v = sO[o + e];
            for (int k = 1; k < copies; ++k) v += sO[o + k * size + e];
            if (v == 0.0) continue;
            fx_flush(&dst[(long long)(sOff[rsh + r / nrc] + r % nrc) * N + sOff[csh + c / ncc] + c % ncc], fx_split(v));
        }
    };
    flush(pIJ, 1, ZIJ, FI, SJ, FJ, 0, NA, TI, NB, Jx);
    flush(pKL, CKL, ZKL, FK, SL, FL, TI + TJ, NC, TI + TJ + TK, ND, Jx);
    flush(pIK, CIK, ZIK, FI, SK, FK, 0, NA, TI + TJ, NC, Kx);
    flush(pIL, CIL, ZIL, FI, SL, FL, 0, NA, TI + TJ + TK, ND, Kx);
    flush(pJK, CJK, ZJK, FJ, SK, FK, TI, NB, TI + TJ, NC, Kx);
    flush(pJL, CJL, ZJL, FJ, SL, FL, TI, NB, TI + TJ + TK, ND, Kx);
}

}  // namespace pm
