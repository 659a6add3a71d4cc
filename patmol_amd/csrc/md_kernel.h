// md_kernel.h -- McMurchie-Davidson J/K kernel on the FP64 matrix cores for
// the d classes with the most Cartesian components, and the tables of
// Hermite expansion coefficients it reads.
//
// For one primitive quartet the integrals of every component are
//   (ab|cd) = sr sum_h sum_h' E_ab[ab][h] R[h + h'] E~_cd[cd][h']
// (Helgaker, Jorgensen, Olsen, Molecular Electronic-Structure Theory, 9.9):
// E_ab[ab][h] are the Hermite expansion coefficients of the bra product
// E^{ax bx}_t E^{ay by}_u E^{az bz}_v (recursion 9.5.6-9.5.7, without the
// Gaussian product factor), E~_cd the ket's times (-1)^(tau+nu+phi),
// R[t,u,v] the Hermite Coulomb integrals d^t/dX^t d^u/dY^u d^v/dZ^v
// F_0(alpha |PQ|^2) at PQ = P - Q, and sr the prefactor of the Rys path
// (2 pi^(5/2) s_ab s_cd c_ab c_cd / (p q sqrt(p + q)) times the degeneracy
// weight).  The reference computes the same integrals by Rys quadrature
// (twoints.c:23-186, rys.c:2269-2335); this path is a different algorithm
// for the same values.
//
// R from Rys roots: with x_r, w_r the roots (in t^2) and weights of the
// NR = Lt/2 + 1 point rule at T = alpha |PQ|^2 (sum_r w_r x_r^n = F_n(T)),
//   R[t,u,v] = sum_r w_r h_t(X; a_r) h_u(Y; a_r) h_v(Z; a_r),  a_r = alpha x_r,
//   h_0 = 1, h_1 = -2 a X, h_(t+1) = -2 a (X h_t + t h_(t-1)),
// because d^t/dX^t exp(-a X^2) = h_t exp(-a X^2) and F_0(alpha R^2) =
// int_0^1 exp(-alpha s^2 R^2) ds; the product is a polynomial of degree
// t + u + v <= Lt in s^2, which the rule integrates exactly.
//
// The two contractions run on v_mfma_f64_16x16x4f64 (lane l supplies
// A(l % 16, k + l / 16), B(k + l / 16, l % 16) and receives
// D(l / 16 + 4 q, l % 16), q = 0..3):
//   Y[h][cd] = sum_h' R[h + h'] E~_cd[cd][h']     (M = h, N = cd, K = h')
//   I[ab][cd] = sum_h E_ab[ab][h] Y[h][cd]          (M = ab, N = cd, K = h)
// The accumulator registers of Y are the B operand of the second product
// as they stand: its k-step (block bm, register q) takes h = 16 bm + 4 q +
// l / 16 from lane l, and the bra table supplies E_ab in that same order.
//
// Work: one workgroup (4 waves) per tile quartet; each wave takes bra shell
// pairs round-robin and runs every ket shell pair of the item against it,
// one shell quartet at a time (a screened-out quartet costs the wave
// nothing).  Per quartet: the NR roots and the 1-D h tables (3 NR lanes),
// the R cube (lanes over (t,u,v)), the two products, the integral block
// written to this wave's LDS buffer, and the six J/K contractions with the
// tile quartet's density blocks (one target element per lane).  J'(IJ)
// stays in a register over the ket loop, K'(IK) and K'(JK) over the L
// shells of one K; every sum is split into fixed point and added by
// integer atomics into this XCD's copy of J'/K' (order-independent, so the
// build stays reproducible bit for bit).
#pragma once
#include "eri_kernels.h"

namespace pm {

typedef double md_d4 __attribute__((ext_vector_type(4)));

// classes that run on this kernel (FP64 items)
template <int LA, int LB, int LC, int LD>
struct MdClass {
    static constexpr bool value = (LA == 2 && LB == 2 && LC == 2 && LD == 2) || (LA == 2 && LB == 2 && LC == 2 && LD == 1) ||
                                  (LA == 2 && LB == 1 && LC == 2 && LD == 1) || (LA == 2 && LB == 2 && LC == 1 && LD == 1);
};

// ---------------------------------------------------------------------------
// E tables: one 64-thread block per shell pair of class (LA, LB); every
// primitive pair of it gets a slot (slot0 + k) in the bra and/or ket table.
// ---------------------------------------------------------------------------

template <int LA, int LB>
__global__ __launch_bounds__(64) void md_etab_kernel(MdTabArgs a)
{
    constexpr int L = LA + LB, NH = herm_count(L), NA = ncart(LA), NB = ncart(LB), PB = md_pad(NB);
    constexpr int MBA = md_blocks(md_nab(LA, LB)), KSA = (NH + 3) / 4;
    __shared__ double E[3][LA + 1][LB + 1][L + 1];
    const int b = blockIdx.x, lane = threadIdx.x;
    if (b >= a.npairs) return;
    const int I = a.pair_I[b], J = a.pair_J[b];
    const PairRef r = a.pairs[a.pair_list[b]];
    for (int k = 0; k < r.npp; ++k) {
        const PrimPair q = a.pp[r.pp0 + k];
        if (lane < 3) {
            const int x = lane;
            const double PA = q.P[x] - a.shell_xyz[3 * I + x], PB = q.P[x] - a.shell_xyz[3 * J + x];
            const double h2p = 0.5 * q.ip;
            for (int i = 0; i <= LA; ++i)
                for (int j = 0; j <= LB; ++j)
                    for (int t = 0; t <= L; ++t) E[x][i][j][t] = 0.0;
            E[x][0][0][0] = 1.0;
            for (int j = 0; j <= LB; ++j)
                for (int i = 0; i <= LA; ++i) {
                    if (i == 0 && j == 0) continue;
                    // from (i-1, j) with X_PA, or (0, j-1) with X_PB
                    const int pi = i > 0 ? i - 1 : 0, pj = i > 0 ? j : j - 1;
                    const double X = i > 0 ? PA : PB;
                    for (int t = 0; t <= i + j; ++t) {
                        double v = X * E[x][pi][pj][t];
                        if (t > 0) v += h2p * E[x][pi][pj][t - 1];
                        if (t + 1 <= pi + pj) v += (t + 1) * E[x][pi][pj][t + 1];
                        E[x][i][j][t] = v;
                    }
                }
        }
        __syncthreads();
        const long long slot = a.slot0[b] + k;
        auto value = [&](int slot, int h, bool ket) -> double {
            const int ia = slot / PB, ib = slot % PB;
            if (ia >= NA || ib >= NB || h >= NH) return 0.0;
            int t, u, v;
            herm_decode(h, t, u, v);
            const int ax = cart_x(LA, ia), ay = cart_y(LA, ia), az = cart_z(LA, ia);
            const int bx = cart_x(LB, ib), by = cart_y(LB, ib), bz = cart_z(LB, ib);
            double e = E[0][ax][bx][t] * E[1][ay][by][u] * E[2][az][bz][v];
            e /= sqrt(cart_dfprod(LA, ia) * cart_dfprod(LB, ib));
            return (ket && ((t + u + v) & 1)) ? -e : e;
        };
        if (a.bra) {
            double *dst = a.bra + slot * (long long)md_bra_regs(LA, LB) * 64;
            int s = 0;
            for (int bm = 0; 16 * bm < NH; ++bm)
                for (int qq = 0; qq < 4; ++qq) {
                    if (16 * bm + 4 * qq >= NH) continue;
                    const int h = 16 * bm + 4 * qq + (lane >> 4);
                    for (int mb = 0; mb < MBA; ++mb) dst[(s * MBA + mb) * 64 + lane] = value(16 * mb + (lane & 15), h, false);
                    ++s;
                }
        }
        if (a.ket) {
            double *dst = a.ket + slot * (long long)md_ket_regs(LA, LB) * 64;
            for (int kk = 0; kk < KSA; ++kk)
                for (int nb = 0; nb < MBA; ++nb)
                    dst[(kk * MBA + nb) * 64 + lane] = value(16 * nb + (lane & 15), 4 * kk + (lane >> 4), true);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// the J/K kernel
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
struct MdShape {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    // component pairs as slots a PB + b, c PD + d (padded to powers of two)
    static constexpr int PB = md_pad(NB), PD = md_pad(ND);
    static constexpr int NAB = NA * PB, NCD = NC * PD;
    static constexpr int LAB = LA + LB, LCD = LC + LD, LT = LAB + LCD;
    static constexpr int NH = herm_count(LAB), NHP = herm_count(LCD), NENT = herm_count(LT);
    static constexpr int NR = LT / 2 + 1, NRP = (NR + 1) & ~1;
    static constexpr int MBH = md_blocks(NH), NBC = md_blocks(NCD), KSA = (NHP + 3) / 4;
    static constexpr int MBA = md_blocks(NAB), KSB = md_bra_steps(NH);
    static constexpr int NRND = (NENT + 63) / 64;
    // R cube: R[t,u,v] at t S1 + u S2 + v, so that the entry of h + h' is at
    // lin(h) + lin(h')
    static constexpr int S2 = LT + 1, S1 = S2 * S2, NCUBE = LT * S1 + 1;
    static constexpr int NHT = 3 * (LT + 1) * NRP;   // h tables [axis][t][root]
    static constexpr int TI = tile_of(LA), TJ = tile_of(LB), TK = tile_of(LC), TL = tile_of(LD);
    static constexpr int FI = TI * NA, FJ = TJ * NB, FK = TK * NC, FL = TL * ND;
    // this wave's sums: J'(KL) over the item [FK][FL | 1]; the K' rows of the
    // current bra pair: K'(IK) [NA][FK], K'(IL) [NA][FL], K'(JK) [NB][FK],
    // K'(JL) [NB][FL]
    static constexpr int SLW = FL | 1, NJKL = FK * SLW;
    static constexpr int RIK = 0, RIL = RIK + NA * FK, RJK = RIL + NA * FL, RJL = RJK + NB * FK,
                         NROW = RJL + NB * FL;
    // (dd|dd): the bra operands of the second product wait in LDS instead
    // of 54 registers (the class's register file is full without them)
    static constexpr bool EAB_LDS = LA == 2 && LB == 2 && LC == 2 && LD == 2;
    static constexpr int NEAB = EAB_LDS ? KSB * MBA * 64 : 0;
    static constexpr int WLDS = NCUBE + NHT + NJKL + NROW + NEAB;   // doubles of one wave's LDS
    static constexpr int KLP = TK * TL, IJP = TI * TJ;
    // Lane (g = lane / 16, c16 = lane % 16) holds, in the integral block's
    // register (m, n, q), slot ab = 16 m + 4 q + g and slot cd = 16 n + c16:
    //   a = m QA + q / QB, b = 4 q % PB + g      (QA = 16 / PB, QB = PB / 4)
    //   c = n (16 / PD) + c16 / PD, d = c16 % PD
    // so a is a register index, b is register bits plus lane bits 4-5, c is
    // register bits plus the lane bits CMASK, d the lane bits DMASK.
    static constexpr int QB = PB / 4, QA = 4 / QB, CPN = 16 / PD;
    static constexpr int DMASK = PD - 1, CMASK = 15 ^ DMASK, GMASK = 48;
    static_assert(PB >= 4 && PD >= 4, "second shell of each pair at least p");
};

template <int LA, int LB, int LC, int LD>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1))) void jk_md_kernel(JKArgs a)
{
    using S = MdShape<LA, LB, LC, LD>;
    constexpr int NA = S::NA, NB = S::NB, NC = S::NC, ND = S::ND, NAB = S::NAB, NCD = S::NCD;
    constexpr int NH = S::NH, NHP = S::NHP, NENT = S::NENT, NR = S::NR, NRP = S::NRP, LT = S::LT;
    constexpr int MBH = S::MBH, NBC = S::NBC, KSA = S::KSA, MBA = S::MBA, KSB = S::KSB, NRND = S::NRND;
    constexpr int TI = S::TI, TJ = S::TJ, TK = S::TK, TL = S::TL, FI = S::FI, FJ = S::FJ, FK = S::FK, FL = S::FL;
    constexpr int KLP = S::KLP, IJP = S::IJP;
    // density block rows padded to an odd number of doubles
    constexpr int SJ = FJ | 1, SK = FK | 1, SL = FL | 1;
    constexpr int ZIJ = FI * SJ, ZKL = FK * SL, ZIK = FI * SK, ZIL = FI * SL, ZJK = FJ * SK, ZJL = FJ * SL;
    constexpr int dIJ = 0, dKL = dIJ + ZIJ, dIK = dKL + ZKL, dIL = dIK + ZIK, dJK = dIL + ZIL, dJL = dJK + ZJK;
    constexpr int ZD = dJL + ZJL;
    // per-wave table of the current bra pair's quartets, one column per ket
    // pair: prefactor, PQ, alpha, the Rys rule (roots, then weights)
    constexpr int KT_SR = 0, KT_PQ = KLP, KT_AL = 4 * KLP, KT_X = 5 * KLP, KT_W = (5 + NR) * KLP,
                  KT_N = (5 + 2 * NR) * KLP;
    __shared__ double sD[ZD];
    __shared__ __attribute__((aligned(16))) double sW[4][S::WLDS];
    __shared__ double sKT[4][KT_N];
    __shared__ int sKL[4][KLP];
    __shared__ int sOff[TI + TJ + TK + TL];
    __shared__ PairRef sKref[KLP];
    __shared__ double sKq[KLP];
    __shared__ int sKok[KLP];
    // the ket primitive pairs and their table slots (uncontracted ket pairs)
    __shared__ PrimPair sKpp[KLP];
    __shared__ int sKslot[KLP];
    __shared__ int sKcontr;

    const int item = blockIdx.x;
    if (item >= a.nitems) return;
    if (a.mol_active && !a.mol_active[a.item_mol[item]]) return;   // whole workgroup
    const int4 it = a.items[item];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
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

    // ---- stage A: function offsets, ket pair table ----
    if (tid < TI) sOff[tid] = a.shell_off[fI + (tid < nI ? tid : 0)];
    else if (tid < TI + TJ) sOff[tid] = a.shell_off[fJ + (tid - TI < nJ ? tid - TI : 0)];
    else if (tid < TI + TJ + TK) sOff[tid] = a.shell_off[fK + (tid - TI - TJ < nK ? tid - TI - TJ : 0)];
    else if (tid < TI + TJ + TK + TL) sOff[tid] = a.shell_off[fL + (tid - TI - TJ - TK < nL ? tid - TI - TJ - TK : 0)];
    if (wave == 2) {
        // one lane per ket pair (KLP <= 64)
        const int t = lane, kq = t / TL, lq = t % TL;
        const bool here = t < KLP;
        const bool kex = kq < nK, lex = lq < nL;
        const int K = fK + (kex ? kq : 0), L = fL + (lex ? lq : 0);
        const bool ok = kex && lex && (!sameKet || K >= L);
        const int Ls = ok ? L : (sameKet ? K : fL);
        const int Kp = K >= Ls ? K : Ls, Lp = K >= Ls ? Ls : K;
        PairRef r = {0, 1};
        if (here) {
            const long long pk = pidx(Kp, Lp);
            r = a.pairs[pk];
            sKref[t] = r;
            sKq[t] = a.pair_q[pk];
            sKok[t] = ok ? 1 : 0;
            if (r.npp == 1) {
                sKpp[t] = a.pp[r.pp0];
                sKslot[t] = a.md_slot[r.pp0];
            }
        }
        const bool contr = __any(here && r.npp != 1);
        if (lane == 0) sKcontr = contr ? 1 : 0;
    }
    __syncthreads();
    // ---- stage B: the six density blocks of the tile quartet ----
    {
        auto gather = [&](int o, int rows, int stride, int rsh, int rn, int nrc, int csh, int cn, int ncc,
                          const double *__restrict__ src) {
            for (int e = tid; e < rows * stride; e += 256) {
                const int r = e / stride, c = e % stride;
                const int rs = r / nrc, cs = c / ncc;
                const bool ex = c < cn * ncc && rs < rn && cs < cn;
                double v = 0.0;
                if (ex) v = src[dadj + (long long)(sOff[rsh + rs] + r % nrc) * N + sOff[csh + cs] + c % ncc];
                sD[o + e] = v;
            }
        };
        gather(dIJ, FI, SJ, 0, nI, NA, TI, nJ, NB, a.DJ);
        gather(dKL, FK, SL, TI + TJ, nK, NC, TI + TJ + TK, nL, ND, a.DJ);
        gather(dIK, FI, SK, 0, nI, NA, TI + TJ, nK, NC, a.DK);
        gather(dIL, FI, SL, 0, nI, NA, TI + TJ + TK, nL, ND, a.DK);
        gather(dJK, FJ, SK, TI, nJ, NB, TI + TJ, nK, NC, a.DK);
        gather(dJL, FJ, SL, TI, nJ, NB, TI + TJ + TK, nL, ND, a.DK);
    }
    __syncthreads();

    // ---- per-lane constants ----
    // cube offsets of the first product's A operand: rows h = 16 bm + lane %
    // 16 (bra Hermite), k-slots h' = 4 kk + lane / 16 (ket Hermite); rows and
    // slots past the lists point at R[0,0,0] (their E entries are zero)
    int linA[MBH], linB[KSA], ent[NRND], entl[NRND];
#pragma unroll
    for (int bm = 0; bm < MBH; ++bm) {
        const int h = 16 * bm + (lane & 15);
        int t = 0, u = 0, v = 0;
        if (h < NH) herm_decode(h, t, u, v);
        linA[bm] = h < NH ? t * S::S1 + u * S::S2 + v : 0;
    }
#pragma unroll
    for (int kk = 0; kk < KSA; ++kk) {
        const int h = 4 * kk + (lane >> 4);
        int t = 0, u = 0, v = 0;
        if (h < NHP) herm_decode(h, t, u, v);
        linB[kk] = h < NHP ? t * S::S1 + u * S::S2 + v : 0;
    }
#pragma unroll
    for (int k = 0; k < NRND; ++k) {
        const int e = lane + 64 * k;
        int t = 0, u = 0, v = 0;
        if (e < NENT) herm_decode(e, t, u, v);
        ent[k] = e < NENT ? (t | (u << 8) | (v << 16)) : -1;
        entl[k] = t * S::S1 + u * S::S2 + v;
    }
    double *const sR = sW[wave];
    double *const sH = sR + S::NCUBE;
    double *const sJKL = sH + S::NHT;
    double *const sRow = sJKL + S::NJKL;
    constexpr int SLW = S::SLW, NROW = S::NROW, RIK = S::RIK, RIL = S::RIL, RJK = S::RJK, RJL = S::RJL;
    for (int e = lane; e < S::NJKL; e += 64) sJKL[e] = 0.0;
    const long long xoff = (long long)xcd_slot() * a.slot_stride + dadj;
    FxAcc *const Jx = a.J + xoff;
    FxAcc *const Kx = a.K + xoff;
    const double *const tbra = a.md_bra[LA * (LA + 1) / 2 + LB];
    const double *const tket = a.md_ket[LC * (LC + 1) / 2 + LD];
    constexpr int BRA_REGS = md_bra_regs(LA, LB), KET_REGS = md_ket_regs(LC, LD);
    static_assert(BRA_REGS == KSB * MBA && KET_REGS == KSA * NBC, "table layouts");

    // this lane's part of the component indices (see MdShape)
    constexpr int PB = S::PB, QA = S::QA, QB = S::QB, CPN = S::CPN;
    constexpr int DMASK = S::DMASK, CMASK = S::CMASK, GMASK = S::GMASK;
    const int g = lane >> 4, cl = (lane & 15) / S::PD, dl = (lane & 15) % S::PD;
    const int dlc = min(dl, ND - 1);   // clamped (padded slots hold zero integrals)

    for (int bp = wave; bp < IJP; bp += 4) {
        const int ii = bp / TJ, jj = bp % TJ;
        if (ii >= nI || jj >= nJ) continue;   // uniform
        const int I = fI + ii, J = fJ + jj;
        if (sameBra && I < J) continue;
        const long long pb = pidx(I, J);
        const double qbra = a.pair_q[pb];
        const PairRef bref = a.pairs[pb];
        const int origIJ = max(a.shell_orig[I], a.shell_orig[J]);
        // bra operands of the second product: loaded once for an
        // uncontracted bra pair, per bra primitive pair otherwise
        constexpr bool EAB_LDS = S::EAB_LDS;
        double eab[EAB_LDS ? 1 : KSB][EAB_LDS ? 1 : MBA];
        double *const sEab = sRow + NROW;   // this wave's copy (EAB_LDS)
        auto load_eab = [&](int ppi) {
            const double *src = tbra + (long long)a.md_slot[ppi] * (BRA_REGS * 64) + lane;
            if constexpr (EAB_LDS) {
                for (int e = 0; e < KSB * MBA; ++e) sEab[e * 64 + lane] = src[e * 64];
            } else {
#pragma unroll
                for (int s = 0; s < KSB; ++s)
#pragma unroll
                    for (int m = 0; m < MBA; ++m) eab[s][m] = src[(s * MBA + m) * 64];
            }
        };
        if (bref.npp == 1) load_eab(bref.pp0);
        const int offI = sOff[ii], offJ = sOff[TI + jj];
        const double *DIJ = sD + dIJ + (ii * NA) * SJ + jj * NB;
        double P1[MBA * 4];   // J'(IJ) partial sums over this bra pair's quartets
#pragma unroll
        for (int m = 0; m < MBA; ++m)
#pragma unroll
            for (int q = 0; q < 4; ++q) P1[m * 4 + q] = 0.0;
        for (int e = lane; e < NROW; e += 64) sRow[e] = 0.0;
        bool anyb = false;
        const PrimPair bpp0 = a.pp[bref.pp0];
        // uncontracted bra and ket pairs (the common case): every quartet's
        // scalars and Rys rule computed at once, one lane per ket pair
        const bool fast = bref.npp == 1 && sKcontr == 0;
        double *const kt = sKT[wave];
        int *const kl = sKL[wave];
        int nv;
        {
            const int t = lane;
            bool ok = false;
            if (t < KLP) {
                const int kq = t / TL, lq = t % TL, K = fK + kq, L = fL + lq;
                ok = sKok[t] != 0 && qbra * sKq[t] >= a.tau;
                if (samePair) ok = ok && ((I > K) || (I == K && J >= L));
                if (a.start_shell > 0) ok = ok && max(origIJ, max(a.shell_orig[K], a.shell_orig[L])) >= a.start_shell;
                if (ok && fast) {
                    const PrimPair kpp = sKpp[t];
                    const double p = bpp0.p, q = kpp.p;
                    const double rpq = rsqrt_pos(p + q), ipq = rpq * rpq;
                    const double sr0 = SRTERM * (bpp0.s * bpp0.ip) * (kpp.s * kpp.ip) * rpq;
                    ok = sr0 >= PRIM_SCREEN;
                    if (ok) {
                        double wdeg = 1.0;
                        if (I == J) wdeg *= 0.5;
                        if (K == L) wdeg *= 0.5;
                        if (I == K && J == L) wdeg *= 0.5;
                        const double PQ0 = bpp0.P[0] - kpp.P[0], PQ1 = bpp0.P[1] - kpp.P[1], PQ2 = bpp0.P[2] - kpp.P[2];
                        const double alpha = p * q * ipq;
                        const double T = (PQ0 * PQ0 + PQ1 * PQ1 + PQ2 * PQ2) * alpha;
                        double x[NR], w[NR];
                        rys_eval<NR>(T, x, w, a.rys);
                        kt[KT_SR + t] = sr0 * bpp0.c * kpp.c * wdeg;
                        kt[KT_PQ + t] = PQ0;
                        kt[KT_PQ + KLP + t] = PQ1;
                        kt[KT_PQ + 2 * KLP + t] = PQ2;
                        kt[KT_AL + t] = alpha;
#pragma unroll
                        for (int r = 0; r < NR; ++r) {
                            kt[KT_X + r * KLP + t] = x[r];
                            kt[KT_W + r * KLP + t] = w[r];
                        }
                    }
                }
            }
            // the kept ket pairs in (K, L) order
            const unsigned long long m = __ballot(ok);
            nv = __popcll(m);
            if (ok) kl[__popcll(m & ((1ull << lane) - 1ull))] = t;
        }
        __builtin_amdgcn_wave_barrier();

        // one primitive quartet (its scalars and rule in column t of the
        // table, its ket operands at table slot kslot): Iacc += its integrals
        md_d4 Iacc[MBA][NBC];
        auto quartet = [&](int t, int kslot) {
            double ecd[KSA][NBC];
            {
                const double *src = tket + (long long)kslot * (KET_REGS * 64) + lane;
#pragma unroll
                for (int kk = 0; kk < KSA; ++kk)
#pragma unroll
                    for (int n = 0; n < NBC; ++n) ecd[kk][n] = src[(kk * NBC + n) * 64];
            }
            // 1-D tables h_t: lane axis * NR + r
            if (lane < 3 * NR) {
                const int ax = lane / NR, r = lane % NR;
                const double x = kt[KT_X + r * KLP + t], w = kt[KT_W + r * KLP + t];
                const double X = kt[KT_PQ + ax * KLP + t];
                const double m2a = -2.0 * kt[KT_AL + t] * x;
                double h0 = ax == 0 ? w * kt[KT_SR + t] : 1.0;
                double h1 = m2a * X * h0;
                sH[(ax * (LT + 1) + 0) * NRP + r] = h0;
                sH[(ax * (LT + 1) + 1) * NRP + r] = h1;
#pragma unroll
                for (int tt = 1; tt < LT; ++tt) {
                    const double h2 = m2a * fma(X, h1, tt * h0);
                    sH[(ax * (LT + 1) + tt + 1) * NRP + r] = h2;
                    h0 = h1;
                    h1 = h2;
                }
            }
            __builtin_amdgcn_wave_barrier();
            // the R cube
#pragma unroll
            for (int k = 0; k < NRND; ++k) {
                if (ent[k] >= 0) {
                    const int tt = ent[k] & 255, uu = (ent[k] >> 8) & 255, vv = ent[k] >> 16;
                    const double *hx = sH + tt * NRP, *hy = sH + ((LT + 1) + uu) * NRP, *hz = sH + (2 * (LT + 1) + vv) * NRP;
                    double sacc = 0.0;
#pragma unroll
                    for (int r = 0; r < NR; ++r) sacc = fma(hx[r] * hy[r], hz[r], sacc);
                    sR[entl[k]] = sacc;
                }
            }
            __builtin_amdgcn_wave_barrier();
            // Y = R E~_cd^T
            md_d4 Y[MBH][NBC];
#pragma unroll
            for (int m = 0; m < MBH; ++m)
#pragma unroll
                for (int n = 0; n < NBC; ++n) Y[m][n] = md_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kk = 0; kk < KSA; ++kk) {
#pragma unroll
                for (int m = 0; m < MBH; ++m) {
                    const double av = sR[linA[m] + linB[kk]];
#pragma unroll
                    for (int n = 0; n < NBC; ++n)
                        Y[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, ecd[kk][n], Y[m][n], 0, 0, 0);
                }
            }
            // I += E_ab Y
            int sidx = 0;
#pragma unroll
            for (int bm = 0; bm < MBH; ++bm)
#pragma unroll
                for (int qq = 0; qq < 4; ++qq) {
                    if (16 * bm + 4 * qq < NH) {
#pragma unroll
                        for (int m = 0; m < MBA; ++m)
#pragma unroll
                            for (int n = 0; n < NBC; ++n)
                                Iacc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                    EAB_LDS ? sEab[(sidx * MBA + m) * 64 + lane] : eab[EAB_LDS ? 0 : sidx][EAB_LDS ? 0 : m],
                                    Y[bm][n][qq], Iacc[m][n],
                                                                                  0, 0, 0);
                        ++sidx;
                    }
                }
        };

        // K'(IK) and K'(JK) partial sums over the L shells of one K
        double P3[MBA * QA * NBC], P5[QB * NBC];
        auto zero_k = [&]() {
#pragma unroll
            for (int m = 0; m < MBA; ++m)
#pragma unroll
                for (int x = 0; x < QA; ++x)
#pragma unroll
                    for (int n = 0; n < NBC; ++n) P3[(m * QA + x) * NBC + n] = 0.0;
#pragma unroll
            for (int x = 0; x < QB; ++x)
#pragma unroll
                for (int n = 0; n < NBC; ++n) P5[x * NBC + n] = 0.0;
        };
        // ... summed across lanes and added to the K' rows of shell kq
        auto flush_k = [&](int kq) {
            lane_reduce_n<GMASK | DMASK>(P3);
            lane_reduce_n<DMASK>(P5);
            if (g == 0 && dl == 0) {
#pragma unroll
                for (int m = 0; m < MBA; ++m)
#pragma unroll
                    for (int x = 0; x < QA; ++x)
#pragma unroll
                        for (int n = 0; n < NBC; ++n) {
                            const int aa = m * QA + x, cc = n * CPN + cl;
                            if (aa < NA && cc < NC) sRow[RIK + aa * FK + kq * NC + cc] += P3[(m * QA + x) * NBC + n];
                        }
            }
            if (dl == 0) {
#pragma unroll
                for (int x = 0; x < QB; ++x)
#pragma unroll
                    for (int n = 0; n < NBC; ++n) {
                        const int bb = 4 * x + g, cc = n * CPN + cl;
                        if (bb < NB && cc < NC) sRow[RJK + bb * FK + kq * NC + cc] += P5[x * NBC + n];
                    }
            }
        };
        zero_k();
        int kq_prev = -1;
        for (int vi = 0; vi < nv; ++vi) {
            const int t = kl[vi];   // uniform
            const int kq = t / TL, lq = t % TL;
            if (kq != kq_prev) {
                if (kq_prev >= 0) {
                    flush_k(kq_prev);
                    zero_k();
                }
                kq_prev = kq;
            }
#pragma unroll
            for (int m = 0; m < MBA; ++m)
#pragma unroll
                for (int n = 0; n < NBC; ++n) Iacc[m][n] = md_d4{0.0, 0.0, 0.0, 0.0};
            if (fast) {
                quartet(t, sKslot[t]);
            } else {
                // contracted pairs: the primitive quartets one by one, each
                // one's scalars and rule written to column t first
                const int K = fK + kq, L = fL + lq;
                double wdeg = 1.0;
                if (I == J) wdeg *= 0.5;
                if (K == L) wdeg *= 0.5;
                if (I == K && J == L) wdeg *= 0.5;
                const PairRef kref = sKref[t];
                bool anyq = false;
                for (int ib = 0; ib < bref.npp; ++ib) {
                    const PrimPair bpp = a.pp[bref.pp0 + ib];
                    if (bref.npp > 1) load_eab(bref.pp0 + ib);
                    for (int ik = 0; ik < kref.npp; ++ik) {
                        const PrimPair kpp = a.pp[kref.pp0 + ik];
                        const double p = bpp.p, q = kpp.p;
                        const double rpq = rsqrt_pos(p + q), ipq = rpq * rpq;
                        const double sr0 = SRTERM * (bpp.s * bpp.ip) * (kpp.s * kpp.ip) * rpq;
                        if (sr0 < PRIM_SCREEN) continue;   // uniform
                        anyq = true;
                        const double PQ0 = bpp.P[0] - kpp.P[0], PQ1 = bpp.P[1] - kpp.P[1], PQ2 = bpp.P[2] - kpp.P[2];
                        const double alpha = p * q * ipq;
                        const double T = (PQ0 * PQ0 + PQ1 * PQ1 + PQ2 * PQ2) * alpha;
                        if (lane < NR) {
                            double x, w;
                            rys_eval_one<NR>(T, lane, x, w, a.rys);
                            kt[KT_X + lane * KLP + t] = x;
                            kt[KT_W + lane * KLP + t] = w;
                        }
                        if (lane == 0) {
                            kt[KT_SR + t] = sr0 * bpp.c * kpp.c * wdeg;
                            kt[KT_PQ + t] = PQ0;
                            kt[KT_PQ + KLP + t] = PQ1;
                            kt[KT_PQ + 2 * KLP + t] = PQ2;
                            kt[KT_AL + t] = alpha;
                        }
                        __builtin_amdgcn_wave_barrier();
                        quartet(t, a.md_slot[kref.pp0 + ik]);
                    }
                }
                if (!anyq) continue;   // every primitive quartet under the screen (uniform)
            }
            anyb = true;
            // ---- the six J/K contractions from the integral registers ----
            const double *DKL = sD + dKL + (kq * NC) * SL + lq * ND;
            const double *DJL = sD + dJL + (jj * NB) * SL + lq * ND;
            const double *DJK = sD + dJK + (jj * NB) * SK + kq * NC;
            const double *DIL = sD + dIL + (ii * NA) * SL + lq * ND;
            const double *DIK = sD + dIK + (ii * NA) * SK + kq * NC;
            // two passes over the integral registers, each with only its
            // own density values live: J'(IJ), J'(KL), K'(IK) first ...
            double P2[NBC], P4[MBA * QA], P6[QB];
            {
                // D(IJ) at this lane's (a, b) of register (m, q)
                double dkl[NBC], djl[4], dij[MBA][4];
#pragma unroll
                for (int m = 0; m < MBA; ++m)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        dij[m][q] = DIJ[min(m * QA + q / QB, NA - 1) * SJ + min((4 * q) % PB + g, NB - 1)];
#pragma unroll
                for (int n = 0; n < NBC; ++n) dkl[n] = DKL[min(n * CPN + cl, NC - 1) * SL + dlc];
#pragma unroll
                for (int q = 0; q < 4; ++q) djl[q] = DJL[min((4 * q) % PB + g, NB - 1) * SL + dlc];
#pragma unroll
                for (int n = 0; n < NBC; ++n) P2[n] = 0.0;
#pragma unroll
                for (int m = 0; m < MBA; ++m)
#pragma unroll
                    for (int n = 0; n < NBC; ++n)
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const double v = Iacc[m][n][q];
                            const int x = q / QB;
                            P1[m * 4 + q] = fma(v, dkl[n], P1[m * 4 + q]);                          // sum over c, d
                            P2[n] = fma(v, dij[m][q], P2[n]);                                       // sum over a, b
                            P3[(m * QA + x) * NBC + n] = fma(v, djl[q], P3[(m * QA + x) * NBC + n]);  // sum over b, d
                        }
            }
            __builtin_amdgcn_sched_barrier(0);
            // ... then K'(IL), K'(JK), K'(JL)
            {
                double djk[4][NBC], dil[MBA][QA], dik[MBA][QA][NBC];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int bb = min((4 * q) % PB + g, NB - 1);
#pragma unroll
                    for (int n = 0; n < NBC; ++n) djk[q][n] = DJK[bb * SK + min(n * CPN + cl, NC - 1)];
                }
#pragma unroll
                for (int m = 0; m < MBA; ++m)
#pragma unroll
                    for (int x = 0; x < QA; ++x) {
                        const int aa = min(m * QA + x, NA - 1);
                        dil[m][x] = DIL[aa * SL + dlc];
#pragma unroll
                        for (int n = 0; n < NBC; ++n) dik[m][x][n] = DIK[aa * SK + min(n * CPN + cl, NC - 1)];
                    }
#pragma unroll
                for (int m = 0; m < MBA; ++m)
#pragma unroll
                    for (int x = 0; x < QA; ++x) P4[m * QA + x] = 0.0;
#pragma unroll
                for (int x = 0; x < QB; ++x) P6[x] = 0.0;
#pragma unroll
                for (int m = 0; m < MBA; ++m)
#pragma unroll
                    for (int n = 0; n < NBC; ++n)
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const double v = Iacc[m][n][q];
                            const int x = q / QB, y = q % QB;
                            P4[m * QA + x] = fma(v, djk[q][n], P4[m * QA + x]);     // sum over b, c
                            P5[y * NBC + n] = fma(v, dil[m][x], P5[y * NBC + n]);  // sum over a, d
                            P6[y] = fma(v, dik[m][x][n], P6[y]);                   // sum over a, c
                        }
            }
            lane_reduce_n<GMASK>(P2);
            lane_reduce_n<GMASK | CMASK>(P4);
            lane_reduce_n<CMASK>(P6);
            // into this wave's sums (one leader lane per element)
            if (g == 0) {
#pragma unroll
                for (int n = 0; n < NBC; ++n) {
                    const int cc = n * CPN + cl;
                    if (cc < NC && dl < ND) sJKL[(kq * NC + cc) * SLW + lq * ND + dl] += 2.0 * P2[n];
                }
            }
            if (g == 0 && cl == 0) {
#pragma unroll
                for (int m = 0; m < MBA; ++m)
#pragma unroll
                    for (int x = 0; x < QA; ++x) {
                        const int aa = m * QA + x;
                        if (aa < NA && dl < ND) sRow[RIL + aa * FL + lq * ND + dl] += P4[m * QA + x];
                    }
            }
            if (cl == 0) {
#pragma unroll
                for (int y = 0; y < QB; ++y) {
                    const int bb = 4 * y + g;
                    if (bb < NB && dl < ND) sRow[RJL + bb * FL + lq * ND + dl] += P6[y];
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (kq_prev >= 0) flush_k(kq_prev);
        if (!anyb) continue;   // uniform: nothing of this bra pair survived
        // this bra pair's sums into this XCD's fixed-point J'/K'
        lane_reduce_n<15>(P1);
        if ((lane & 15) == 0) {
#pragma unroll
            for (int m = 0; m < MBA; ++m)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int aa = m * QA + q / QB, bb = (4 * q) % PB + g;
                    if (aa < NA && bb < NB) fx_add_global(&Jx[(long long)(offI + aa) * N + offJ + bb], 2.0 * P1[m * 4 + q]);
                }
        }
        __builtin_amdgcn_wave_barrier();
        for (int e = lane; e < NROW; e += 64) {
            const double v = sRow[e];
            if (v == 0.0) continue;
            int row, col;   // function of the bra component, column inside the ket tile
            int ctile, ncmp;
            if (e < RIL) {
                row = offI + (e - RIK) / FK, col = (e - RIK) % FK, ctile = TI + TJ, ncmp = NC;
            } else if (e < RJK) {
                row = offI + (e - RIL) / FL, col = (e - RIL) % FL, ctile = TI + TJ + TK, ncmp = ND;
            } else if (e < RJL) {
                row = offJ + (e - RJK) / FK, col = (e - RJK) % FK, ctile = TI + TJ, ncmp = NC;
            } else {
                row = offJ + (e - RJL) / FL, col = (e - RJL) % FL, ctile = TI + TJ + TK, ncmp = ND;
            }
            fx_add_global(&Kx[(long long)row * N + sOff[ctile + col / ncmp] + col % ncmp], v);
        }
    }
    // J'(KL): the four waves' sums added in wave order, then into J'
    __syncthreads();
    for (int e = tid; e < FK * FL; e += 256) {
        const int r = e / FL, c = e % FL;
        const int o = r * SLW + c;
        constexpr int J0 = S::NCUBE + S::NHT;
        const double v = ((sW[0][J0 + o] + sW[1][J0 + o]) + sW[2][J0 + o]) + sW[3][J0 + o];
        if (v == 0.0) continue;
        fx_add_global(&Jx[(long long)(sOff[TI + TJ + r / NC] + r % NC) * N + sOff[TI + TJ + TK + c / ND] + c % ND], v);
    }
}

}  // namespace pm
