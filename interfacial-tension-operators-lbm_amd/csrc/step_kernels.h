// Host-visible declarations of the step kernels' launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include "lattice.h"

namespace itolbm {

struct StepArgs {
    const double* inR;   // [19][S] post-collision populations of the previous step (own site)
    const double* inB;
    double* outR;        // [19][S] post-collision populations of this step
    double* outB;
    double* aux;         // nullptr, or [6][S]: mediator momentum gm xyz and force F xyz
    long long S;         // population stride, >= nx * ny
    int ly;              // rows per wavefront segment
    int nstrip, nseg;    // tiles: strips of 64 - 2H columns x segments of ly rows
};

// tiles of one case; returns the halo width H of the operator
int step_tile_geometry(int op, int nx, int ny, int ly, int* nstrip, int* nseg);

hipError_t launch_step_f64(const CaseConst& cc, const StepArgs& a, hipStream_t stream);

}  // namespace itolbm
