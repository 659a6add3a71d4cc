// batch.h -- what the device SCF needs from a molecule batch (hip_context.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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

}  // namespace pm
