// batch.h -- what the device SCF needs from the G-build side (hip_context.cpp):
// molecule batches and the incremental (density-difference) build.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "patmol_types.h"

struct PatmolHipBatch;

namespace pm {

struct BatchView {
    int nmol;
    const int *N;            // functions per molecule
    const int64_t *pstart;   // first element of each molecule's packed triangle in the batch arrays
    int64_t no2tot;          // elements of all packed triangles
    double *d_P;             // the batch's packed half densities (the G build reads them)
    double *d_G;             // the batch's packed G (the G build writes them)
};

// the batch's tables, buffers and layout (false if not finalised)
bool batch_view(PatmolHipBatch *b, BatchView &v);
// G = 2J[P] - K[P] for every molecule, P and G the batch's own buffers:
// one replay of the captured graph on `s`
int batch_launch(PatmolHipBatch *b, hipStream_t s);
// which molecules the following replays build (host array, one flag per
// molecule; the rest keep their G of the last replay that built them);
// all 1 after finalisation
int batch_set_active(PatmolHipBatch *b, const int *active, hipStream_t s);

// Work-item counts of one G build
struct BuildCounts {
    int64_t items = 0;     // work items of the Schwarz-screened list
    int64_t fp32 = 0;      // launched in FP32
    int64_t skipped = 0;   // not launched (density-weighted screening of a delta build)
};

// G += 2J[dP] - K[dP] for a density difference dP (device, packed): the G
// build of patmol_hip_rhf_gmatrix_device, plus screening by the density:
// quartets whose Schwarz bound times max|dP| is below the Schwarz threshold
// are skipped (their contributions are below the threshold).  The precision
// split, when on, uses max|dP|.
int rhf_gmatrix_delta(const ERIFS *e, const double *dP, double *dG, hipStream_t s, BuildCounts *counts);

// The precision mode of the G builds that follow (patmol_hip_set_precision's
// mode: 0 FP64, 1 mixed), read and set without touching the threshold: the
// device SCF runs its last iterations in FP64.
int precision_mode();
void set_precision_mode(int mode);
// the counts of the last build of e (0 if e has no tables)
int last_build_counts(const ERIFS *e, BuildCounts *counts);

}  // namespace pm
