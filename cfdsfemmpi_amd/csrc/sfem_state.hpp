// sfem_state.hpp -- launchers of the SolutionData streaming kernels (sfem_state.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfem {

// The td[] entries updateIterStep reads (SolutionData.cpp:83-91).
struct TimeCoeffs {
  double td1, td2, td9, td10, td11, td12, td13, td15;
};

hipError_t launch_time_update(hipStream_t s, int64_t n, const double* soln, double* prev, double* prev2,
                              double* prev3, double* prev4, const double* dot, double* dotPrev);
hipError_t launch_extrapolate(hipStream_t s, int64_t n, const double* prev, const double* prev2, double* soln);
hipError_t launch_iter_step(hipStream_t s, int64_t n, const TimeCoeffs& c, const double* soln,
                            const double* prev, const double* prev2, const double* prev3,
                            const double* prev4, double* dot, const double* dotPrev, double* cur,
                            double* dotCur);
hipError_t launch_assign_bc(hipStream_t s, int64_t n, const int32_t* ind, const double* val,
                            double timeFact, const double* soln, double* applied);
// soln = 3 prev - 3 prev2 + prev3 (quadratic extrapolation of the last three states)
hipError_t launch_extrapolate2(hipStream_t s, int64_t n, const double* prev, const double* prev2,
                              const double* prev3, double* soln);
// x[ind[k]] = val[k] * timeFact (the prescribed values of a step written into a state)
hipError_t launch_set_bc(hipStream_t s, int64_t n, const int32_t* ind, const double* val, double timeFact,
                         double* x);
hipError_t launch_add_bc(hipStream_t s, int64_t n, const int32_t* ind, const int32_t* mult,
                         const double* applied, double* soln);
hipError_t launch_update_solution(hipStream_t s, int64_t n, const int32_t* assy, const double* du,
                                  double* soln);
hipError_t launch_zero_f64(hipStream_t s, int64_t n, double* a);
hipError_t launch_zero_f32(hipStream_t s, int64_t n, float* a);

}  // namespace sfem
