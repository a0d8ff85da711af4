// Dense-to-smooth restriction of the local potential (potential.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_lds.h"

namespace q3x {

// v_dense: device, dense box (C order i1, i2, i3), real.  dense_tw / smooth_tw: the stage-major
// complex128 twiddle tables of each axis (FftTables::tw64).  work_dense / work_smooth: device
// scratch of dense / smooth box size in complex128.  v_smooth: device, smooth box, real, C order.
hipError_t launch_restrict_potential(const double* v_dense, const int dense_dims[3], const FftPlan dense_plans[3],
                                     const void* const dense_tw[3], const int smooth_dims[3],
                                     const FftPlan smooth_plans[3], const void* const smooth_tw[3], void* work_dense,
                                     void* work_smooth, double* v_smooth, hipStream_t st);

}  // namespace q3x
