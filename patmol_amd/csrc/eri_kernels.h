// eri_kernels.h -- argument blocks and launchers of the ERI kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "eri_engine.h"

namespace pm {

constexpr int PM_LMAX = 2;
constexpr int PM_NPC = (PM_LMAX + 1) * (PM_LMAX + 2) / 2;  // pair classes
// private J'/K' copies, one per XCD (MI355X has 8; a power of two)
constexpr int PM_XCD_SLOTS = 8;

// Shells per tile for angular momentum l (a power of two; the ket tile pair
// TK*TL must divide 64).
__host__ __device__ constexpr int tile_of(int l) { return l <= 1 ? 8 : 4; }
// index of the class (la lb | lc ld) in the tuned tables
__host__ __device__ constexpr int jk3_class_index(int la, int lb, int lc, int ld) { return ((la * 3 + lb) * 3 + lc) * 3 + ld; }
// (l, l') of pair class pc = l (l + 1) / 2 + l'
__host__ __device__ constexpr int pair_class_la(int pc) { return pc >= 3 ? 2 : (pc >= 1 ? 1 : 0); }
__host__ __device__ constexpr int pair_class_lb(int pc) { return pc - pair_class_la(pc) * (pair_class_la(pc) + 1) / 2; }
}  // namespace pm
// the measured launch shapes and FP32 policy (jk3_tuned_shape, jk3_fp32_pays)
#include <jk3_tuned.inc>
namespace pm {

// A J'/K' accumulator element in fixed point: the value is
// hi * 2^-FX_HI_BITS + lo * 2^-FX_LO_BITS.  Every contribution is rounded
// once onto that grid (2^-65 absolute) and then added as integers, which are
// associative: the sums do not depend on the order in which lanes, waves
// and workgroups add them, so a G build is bitwise reproducible.
constexpr int FX_HI_BITS = 24;
constexpr int FX_LO_BITS = 64;
struct __attribute__((aligned(16))) FxAcc {
    long long hi, lo;
};

// Fixed-point words of v (FxAcc): hi = v 2^24 rounded to the nearest
// integer, lo = the remainder in units of 2^-64, rounded to nearest.  The
// roundings use the 1.5 2^52 bias: for |x| < 2^51, x + 1.5 2^52 is rounded
// to an integer by the FP64 add itself and that integer sits in the low
// mantissa bits, so the difference of the bit patterns is round(x) as an
// int64.  Every step is exact in FP64 (power-of-two scalings, x - round(x)),
// so the split is exact up to the final rounding of lo; valid for
// |v| < 2^27.
constexpr double FX_HI_SCALE = 16777216.0;           // 2^FX_HI_BITS
constexpr double FX_LO_SCALE = 1099511627776.0;      // 2^(FX_LO_BITS - FX_HI_BITS)
constexpr double FX_BIAS = 6755399441055744.0;       // 1.5 2^52
static_assert(FX_LO_BITS - FX_HI_BITS == 40, "FX_LO_SCALE");
__host__ __device__ inline long long fx_bits(double x)
{
    union {
        double d;
        long long i;
    } u = {x};
    return u.i;
}
__host__ __device__ inline FxAcc fx_split(double v)
{
    const double t = v * FX_HI_SCALE;
    const double bt = t + FX_BIAS;              // round(t) in the low mantissa bits
    const double h = bt - FX_BIAS;              // round(t) as a double
    const double bl = (t - h) * FX_LO_SCALE + FX_BIAS;   // |(t - h) 2^40| <= 2^39
    return FxAcc{fx_bits(bt) - fx_bits(FX_BIAS), fx_bits(bl) - fx_bits(FX_BIAS)};
}

__host__ __device__ inline double fx_value(long long hi, long long lo)
{
    return (double)hi * (1.0 / FX_HI_SCALE) + (double)lo * (1.0 / (FX_HI_SCALE * FX_LO_SCALE));
}

// One molecule of a batch: the batch's work items never mix molecules, and
// each carries its molecule's index.  Element (row, col) of the molecule's
// D, J', K' (rows and columns numbered by the batch's global sorted
// functions) is at dadj + row N + col; its shell pairs start at pbase.
struct MolDesc {
    int N;            // functions of the molecule (row stride of its blocks)
    int sbase;        // first sorted shell of the molecule in the batch
    long long pbase;  // its first shell pair in the pair tables
    long long dadj;   // element offset of its blocks (see above)
};

struct JKArgs {
    const int4 *items;         // (tI, tJ, tK, tL): bra tile pair first
    int nitems;
    const int *tile_first;     // first sorted shell of each tile
    const int *tile_size;
    const double *shell_xyz;   // sorted shells, 3 per shell (bohr)
    const int *shell_off;      // first function of each sorted shell, in sorted function order
    const int *shell_orig;     // original shell index of each sorted shell
    const PairRef *pairs;      // sorted pairs I >= J at I(I+1)/2 + J
    const double *pair_q;      // Schwarz bound of each pair
    const PrimPair *pp;
    const double *DJ;          // density contracted into J (full N x N)
    const double *DK;          // density contracted into K
    FxAcc *J, *K;              // accumulators J', K': PM_XCD_SLOTS copies of N x N each
    int N;
    double tau;                // Schwarz threshold on Q_IJ Q_KL
    int start_shell;           // keep quartets whose largest original shell index >= start_shell
    RysTables rys;
    // batches (null item_mol: one molecule, N x N blocks at offset 0)
    const int *item_mol;       // molecule of each item
    const MolDesc *mols;
    // batches: nonzero for the molecules whose items run (null: all); the
    // batched SCF stops building the molecules that have converged
    const int *mol_active = nullptr;
    long long slot_stride;     // elements of one per-XCD copy of J' (and of K')
    // McMurchie-Davidson tables (md_kernel.h): slot of each primitive pair
    // and the bra / ket Hermite-coefficient tables per pair class; with
    // md_slot null the classes that could use them run their Rys kernels
    const int *md_slot = nullptr;
    const double *md_bra[PM_NPC] = {};
    const double *md_ket[PM_NPC] = {};
};

struct EriOutArgs {
    const int *quartets;       // 4 sorted shell indices per quartet (kernel orientation)
    int nq;
    const double *shell_xyz;
    const PairRef *pairs;
    const PrimPair *pp;
    double *out;
    RysTables rys;
    int prec;                  // 0 FP64, 1 the FP32 arithmetic of the mixed J/K kernels
};

struct SchwarzArgs {
    const int *pair_list;      // pair indices of one pair class
    const int *pair_I, *pair_J;
    int npairs;
    const double *shell_xyz;
    const PairRef *pairs;
    const PrimPair *pp;
    double *q;
    RysTables rys;
};

// ---------------------------------------------------------------------------
// McMurchie-Davidson path (md_kernel.h): Hermite lists and the layouts of
// the per-pair-class Hermite-coefficient tables
// ---------------------------------------------------------------------------
// Hermite functions (t,u,v) with t + u + v <= L, ordered by n = t + u + v,
// then t descending, then u descending (a list for L is a prefix of the list
// for L + 1).
__host__ __device__ constexpr int herm_count(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

__host__ __device__ inline void herm_decode(int h, int &t, int &u, int &v)
{
    int n = 0, base = 0;
    while (base + (n + 1) * (n + 2) / 2 <= h) {
        base += (n + 1) * (n + 2) / 2;
        ++n;
    }
    int k = h - base;
    t = n;
    while (k >= n - t + 1) {
        k -= n - t + 1;
        --t;
    }
    u = n - t - k;
    v = k;
}

// Table layouts of one pair class (la >= lb), in the lane order the kernel
// loads them (64 doubles per register).  Component pairs are numbered with
// the second index padded to a power of two, slot = a pb + b (pb =
// md_pad(ncart(lb))), so that a lane's slots split into register and lane
// bits; padded slots hold zeros.
//   bra table: E_ab for the second product's A operand, [step][slot block][lane]
//     with slot = 16 block + lane % 16 and step (bm, q) ->
//     h = 16 bm + 4 q + lane / 16, over the steps with 16 bm + 4 q < nh;
//   ket table: E~_cd for the first product's B operand, [kstep][slot block][lane]
//     with h' = 4 kstep + lane / 16, slot = 16 block + lane % 16.
__host__ __device__ constexpr int md_pad(int n) { return n <= 1 ? 1 : (n <= 2 ? 2 : (n <= 4 ? 4 : 8)); }
__host__ __device__ constexpr int md_nab(int la, int lb) { return ncart(la) * md_pad(ncart(lb)); }
__host__ __device__ constexpr int md_blocks(int n) { return (n + 15) / 16; }
__host__ __device__ constexpr int md_bra_steps(int nh)
{
    int s = 0;
    for (int bm = 0; 16 * bm < nh; ++bm)
        for (int q = 0; q < 4; ++q)
            if (16 * bm + 4 * q < nh) ++s;
    return s;
}
__host__ __device__ constexpr int md_bra_regs(int la, int lb)
{
    return md_bra_steps(herm_count(la + lb)) * md_blocks(md_nab(la, lb));
}
__host__ __device__ constexpr int md_ket_regs(int la, int lb)
{
    return ((herm_count(la + lb) + 3) / 4) * md_blocks(md_nab(la, lb));
}

// pair classes whose tables the kernel reads: bra dd, dp; ket dd, dp, pp
__host__ __device__ constexpr bool md_bra_class(int la, int lb) { return la == 2 && lb >= 1; }
__host__ __device__ constexpr bool md_ket_class(int la, int lb) { return (la == 2 && lb >= 1) || (la == 1 && lb == 1); }

struct MdTabArgs {
    const int *pair_list;      // pair index I(I+1)/2 + J of each shell pair
    const int *pair_I, *pair_J;
    const int *slot0;          // first slot of each shell pair
    int npairs;
    const double *shell_xyz;
    const PairRef *pairs;
    const PrimPair *pp;
    double *bra;               // bra table of the class or null
    double *ket;               // ket table of the class or null
};

// prec 0: FP64 instantiation, 1: FP32 (FP64 accumulation).  Returns nonzero
// for a class without a kernel.
int launch_jk(int bra_pc, int ket_pc, const JKArgs &args, int prec, hipStream_t s);
int launch_eri_out(int bra_pc, int ket_pc, const EriOutArgs &args, hipStream_t s);
int launch_schwarz(int pc, const SchwarzArgs &args, hipStream_t s);
// Hermite-coefficient tables of pair class (la, lb) (md_tables.hip)
int launch_md_etab(int la, int lb, const MdTabArgs &args, hipStream_t s);

}  // namespace pm
