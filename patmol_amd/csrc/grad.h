// grad.h -- argument block and launcher of the two-electron gradient kernel.
#pragma once
#include <hip/hip_runtime.h>

#include "eri_kernels.h"

namespace pm {

struct GradArgs {
    const int4 *items;         // tile quartets of one class pair (bra tile pair first)
    int nitems;
    const int *tile_first, *tile_size;
    const double *shell_xyz;   // sorted shells, 3 per shell (bohr)
    const int *shell_off;      // first function of each sorted shell, in sorted function order
    const PairRef *pairs;      // sorted pairs I >= J at I(I+1)/2 + J
    const double *pair_q;      // Schwarz bound of each pair
    const PrimPair *pp;
    const double2 *ppab;       // the two primitive exponents (of I, of J) of each primitive pair
    const double *D;           // half density, full N x N, sorted function order
    FxAcc *gshell;             // per sorted shell, x y z: fixed-point gradient sums
    int N;
    double tau;
    RysTables rys;
};

int launch_grad(int bra_pc, int ket_pc, const GradArgs &args, hipStream_t s);

}  // namespace pm
