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
    const double *DK0, *DK1;   // densities contracted into K
    double *J, *K0, *K1;       // accumulators J', K': PM_XCD_SLOTS copies of N x N each
    int N;
    double tau;                // Schwarz threshold on Q_IJ Q_KL
    int start_shell;           // keep quartets whose largest original shell index >= start_shell
    RysTables rys;
};

struct EriOutArgs {
    const int *quartets;       // 4 sorted shell indices per quartet (kernel orientation)
    int nq;
    const double *shell_xyz;
    const PairRef *pairs;
    const PrimPair *pp;
    double *out;
    RysTables rys;
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

int launch_jk(int bra_pc, int ket_pc, const JKArgs &args, int prec, int nk, hipStream_t s);
int launch_eri_out(int bra_pc, int ket_pc, const EriOutArgs &args, hipStream_t s);
int launch_schwarz(int pc, const SchwarzArgs &args, hipStream_t s);

}  // namespace pm
