// eri_dispatch.cpp -- class switch over the per-class kernel objects
// (every (bra pair class >= ket pair class) up to d shells, and the
// reversed classes (ss|ps), (ss|ds), (ps|ds) for the J/K build).
#include "eri_kernels.h"

namespace pm {

int cls_00_00_jk(const JKArgs &, int, hipStream_t);
int cls_00_00_out(const EriOutArgs &, hipStream_t);
int cls_00_00_schwarz(const SchwarzArgs &, hipStream_t);
int cls_10_00_jk(const JKArgs &, int, hipStream_t);
int cls_10_00_out(const EriOutArgs &, hipStream_t);
int cls_10_10_jk(const JKArgs &, int, hipStream_t);
int cls_10_10_out(const EriOutArgs &, hipStream_t);
int cls_10_10_schwarz(const SchwarzArgs &, hipStream_t);
int cls_11_00_jk(const JKArgs &, int, hipStream_t);
int cls_11_00_out(const EriOutArgs &, hipStream_t);
int cls_11_10_jk(const JKArgs &, int, hipStream_t);
int cls_11_10_out(const EriOutArgs &, hipStream_t);
int cls_11_11_jk(const JKArgs &, int, hipStream_t);
int cls_11_11_out(const EriOutArgs &, hipStream_t);
int cls_11_11_schwarz(const SchwarzArgs &, hipStream_t);
int cls_20_00_jk(const JKArgs &, int, hipStream_t);
int cls_20_00_out(const EriOutArgs &, hipStream_t);
int cls_20_10_jk(const JKArgs &, int, hipStream_t);
int cls_20_10_out(const EriOutArgs &, hipStream_t);
int cls_20_11_jk(const JKArgs &, int, hipStream_t);
int cls_20_11_out(const EriOutArgs &, hipStream_t);
int cls_20_20_jk(const JKArgs &, int, hipStream_t);
int cls_20_20_out(const EriOutArgs &, hipStream_t);
int cls_20_20_schwarz(const SchwarzArgs &, hipStream_t);
int cls_21_00_jk(const JKArgs &, int, hipStream_t);
int cls_21_00_out(const EriOutArgs &, hipStream_t);
int cls_21_10_jk(const JKArgs &, int, hipStream_t);
int cls_21_10_out(const EriOutArgs &, hipStream_t);
int cls_21_11_jk(const JKArgs &, int, hipStream_t);
int cls_21_11_out(const EriOutArgs &, hipStream_t);
int cls_21_20_jk(const JKArgs &, int, hipStream_t);
int cls_21_20_out(const EriOutArgs &, hipStream_t);
int cls_21_21_jk(const JKArgs &, int, hipStream_t);
int cls_21_21_out(const EriOutArgs &, hipStream_t);
int cls_21_21_schwarz(const SchwarzArgs &, hipStream_t);
int cls_22_00_jk(const JKArgs &, int, hipStream_t);
int cls_22_00_out(const EriOutArgs &, hipStream_t);
int cls_22_10_jk(const JKArgs &, int, hipStream_t);
int cls_22_10_out(const EriOutArgs &, hipStream_t);
int cls_22_11_jk(const JKArgs &, int, hipStream_t);
int cls_22_11_out(const EriOutArgs &, hipStream_t);
int cls_22_20_jk(const JKArgs &, int, hipStream_t);
int cls_22_20_out(const EriOutArgs &, hipStream_t);
int cls_22_21_jk(const JKArgs &, int, hipStream_t);
int cls_22_21_out(const EriOutArgs &, hipStream_t);
int cls_22_22_jk(const JKArgs &, int, hipStream_t);
int cls_22_22_out(const EriOutArgs &, hipStream_t);
int cls_22_22_schwarz(const SchwarzArgs &, hipStream_t);

// J/K kernels of the reversed classes (bra pair class below the ket's)
// (ss|ps), (ss|ds), (ps|ds): items of those class pairs run in the
// orientation with fewer executed steps (hip_context.cpp builds the list)
int cls_00_10_jk(const JKArgs &, int, hipStream_t);
int cls_00_20_jk(const JKArgs &, int, hipStream_t);
int cls_10_20_jk(const JKArgs &, int, hipStream_t);

int launch_jk(int bra_pc, int ket_pc, const JKArgs &args, int prec, hipStream_t s)
{
    switch (bra_pc * PM_NPC + ket_pc) {
    case 0: return cls_00_00_jk(args, prec, s);
    case 1: return cls_00_10_jk(args, prec, s);
    case 3: return cls_00_20_jk(args, prec, s);
    case 6: return cls_10_00_jk(args, prec, s);
    case 9: return cls_10_20_jk(args, prec, s);
    case 7: return cls_10_10_jk(args, prec, s);
    case 12: return cls_11_00_jk(args, prec, s);
    case 13: return cls_11_10_jk(args, prec, s);
    case 14: return cls_11_11_jk(args, prec, s);
    case 18: return cls_20_00_jk(args, prec, s);
    case 19: return cls_20_10_jk(args, prec, s);
    case 20: return cls_20_11_jk(args, prec, s);
    case 21: return cls_20_20_jk(args, prec, s);
    case 24: return cls_21_00_jk(args, prec, s);
    case 25: return cls_21_10_jk(args, prec, s);
    case 26: return cls_21_11_jk(args, prec, s);
    case 27: return cls_21_20_jk(args, prec, s);
    case 28: return cls_21_21_jk(args, prec, s);
    case 30: return cls_22_00_jk(args, prec, s);
    case 31: return cls_22_10_jk(args, prec, s);
    case 32: return cls_22_11_jk(args, prec, s);
    case 33: return cls_22_20_jk(args, prec, s);
    case 34: return cls_22_21_jk(args, prec, s);
    case 35: return cls_22_22_jk(args, prec, s);
    default: return -1;
    }
}

int launch_eri_out(int bra_pc, int ket_pc, const EriOutArgs &args, hipStream_t s)
{
    switch (bra_pc * PM_NPC + ket_pc) {
    case 0: return cls_00_00_out(args, s);
    case 6: return cls_10_00_out(args, s);
    case 7: return cls_10_10_out(args, s);
    case 12: return cls_11_00_out(args, s);
    case 13: return cls_11_10_out(args, s);
    case 14: return cls_11_11_out(args, s);
    case 18: return cls_20_00_out(args, s);
    case 19: return cls_20_10_out(args, s);
    case 20: return cls_20_11_out(args, s);
    case 21: return cls_20_20_out(args, s);
    case 24: return cls_21_00_out(args, s);
    case 25: return cls_21_10_out(args, s);
    case 26: return cls_21_11_out(args, s);
    case 27: return cls_21_20_out(args, s);
    case 28: return cls_21_21_out(args, s);
    case 30: return cls_22_00_out(args, s);
    case 31: return cls_22_10_out(args, s);
    case 32: return cls_22_11_out(args, s);
    case 33: return cls_22_20_out(args, s);
    case 34: return cls_22_21_out(args, s);
    case 35: return cls_22_22_out(args, s);
    default: return -1;
    }
}

int launch_schwarz(int pc, const SchwarzArgs &args, hipStream_t s)
{
    switch (pc) {
    case 0: return cls_00_00_schwarz(args, s);
    case 1: return cls_10_10_schwarz(args, s);
    case 2: return cls_11_11_schwarz(args, s);
    case 3: return cls_20_20_schwarz(args, s);
    case 4: return cls_21_21_schwarz(args, s);
    case 5: return cls_22_22_schwarz(args, s);
    default: return -1;
    }
}

}  // namespace pm
