// sfem_state.hip -- streaming kernels over the SolutionData vectors.
//
// SolutionData::timeUpdate / updateIterStep (src/base/SolutionData.cpp:64-91),
// assignBoundaryConditions (src/base/StabFEM.cpp:391-410), the iteration-0
// prescribed-value update and the solution update of the Newton loop
// (src/base/StabFEM2.cpp:719-730, 785-788).  Each is one HBM-bound pass.
#include <hip/hip_runtime.h>

#include "sfem_internal.hpp"
#include "sfem_state.hpp"

namespace sfem {

constexpr int kTPB = 256;

static inline unsigned grid1(int64_t n) {
  int64_t g = (n + kTPB - 1) / kTPB;
  if (g > 65535 * 16) g = 65535 * 16;
  return (unsigned)(g < 1 ? 1 : g);
}

__global__ __launch_bounds__(kTPB) void k_time_update(int64_t n, const double* __restrict__ soln,
                                                      double* __restrict__ prev, double* __restrict__ prev2,
                                                      double* __restrict__ prev3, double* __restrict__ prev4,
                                                      const double* __restrict__ dot,
                                                      double* __restrict__ dotPrev) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB) {
    prev4[i] = prev3[i];
    prev3[i] = prev2[i];
    prev2[i] = prev[i];
    prev[i] = soln[i];
    dotPrev[i] = dot[i];
  }
}

__global__ __launch_bounds__(kTPB) void k_extrapolate(int64_t n, const double* __restrict__ prev,
                                                      const double* __restrict__ prev2,
                                                      double* __restrict__ soln) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    soln[i] = 2.0 * prev[i] - prev2[i];
}

__global__ __launch_bounds__(kTPB) void k_extrapolate2(int64_t n, const double* __restrict__ prev,
                                                       const double* __restrict__ prev2,
                                                       const double* __restrict__ prev3,
                                                       double* __restrict__ soln) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    soln[i] = 3.0 * prev[i] - 3.0 * prev2[i] + prev3[i];
}

__global__ __launch_bounds__(kTPB) void k_iter_step(int64_t n, TimeCoeffs c, const double* __restrict__ soln,
                                                    const double* __restrict__ prev,
                                                    const double* __restrict__ prev2,
                                                    const double* __restrict__ prev3,
                                                    const double* __restrict__ prev4,
                                                    double* __restrict__ dot,
                                                    const double* __restrict__ dotPrev,
                                                    double* __restrict__ cur, double* __restrict__ dotCur) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB) {
    const double s = soln[i], p = prev[i];
    const double d = c.td9 * s + c.td10 * p + c.td11 * prev2[i] + c.td12 * prev3[i] + c.td13 * prev4[i] +
                     c.td15 * dotPrev[i];
    dot[i] = d;
    cur[i] = c.td2 * s + (1.0 - c.td2) * p;
    dotCur[i] = c.td1 * d + (1.0 - c.td1) * dotPrev[i];
  }
}

__global__ __launch_bounds__(kTPB) void k_assign_bc(int64_t n, const int32_t* __restrict__ ind,
                                                    const double* __restrict__ val, double timeFact,
                                                    const double* __restrict__ soln,
                                                    double* __restrict__ applied) {
  const int64_t k = (int64_t)blockIdx.x * kTPB + threadIdx.x;
  if (k >= n) return;
  const int32_t i = ind[k];
  applied[i] = val[k] * timeFact - soln[i];
}

__global__ __launch_bounds__(kTPB) void k_set_bc(int64_t n, const int32_t* __restrict__ ind,
                                                 const double* __restrict__ val, double timeFact,
                                                 double* __restrict__ x) {
  const int64_t k = (int64_t)blockIdx.x * kTPB + threadIdx.x;
  if (k >= n) return;
  x[ind[k]] = val[k] * timeFact;
}

__global__ __launch_bounds__(kTPB) void k_add_bc(int64_t n, const int32_t* __restrict__ ind,
                                                 const int32_t* __restrict__ mult,
                                                 const double* __restrict__ applied,
                                                 double* __restrict__ soln) {
  const int64_t k = (int64_t)blockIdx.x * kTPB + threadIdx.x;
  if (k >= n) return;
  const int32_t i = ind[k];
  double s = soln[i];
  const double a = applied[i];
  for (int m = 0; m < mult[k]; ++m) s += a;  // once per listing, as the reference loop does
  soln[i] = s;
}

__global__ __launch_bounds__(kTPB) void k_update_solution(int64_t n, const int32_t* __restrict__ assy,
                                                          const double* __restrict__ du,
                                                          double* __restrict__ soln) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    soln[assy[i]] += du[i];
}

template <typename T>
__global__ __launch_bounds__(kTPB) void k_zero(int64_t n, T* __restrict__ a) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    a[i] = (T)0;
}

hipError_t launch_time_update(hipStream_t s, int64_t n, const double* soln, double* prev, double* prev2,
                              double* prev3, double* prev4, const double* dot, double* dotPrev) {
  hipLaunchKernelGGL(k_time_update, dim3(grid1(n)), dim3(kTPB), 0, s, n, soln, prev, prev2, prev3, prev4,
                     dot, dotPrev);
  return hipGetLastError();
}

hipError_t launch_extrapolate(hipStream_t s, int64_t n, const double* prev, const double* prev2, double* soln) {
  hipLaunchKernelGGL(k_extrapolate, dim3(grid1(n)), dim3(kTPB), 0, s, n, prev, prev2, soln);
  return hipGetLastError();
}

hipError_t launch_extrapolate2(hipStream_t s, int64_t n, const double* prev, const double* prev2,
                              const double* prev3, double* soln) {
  hipLaunchKernelGGL(k_extrapolate2, dim3(grid1(n)), dim3(kTPB), 0, s, n, prev, prev2, prev3, soln);
  return hipGetLastError();
}

hipError_t launch_iter_step(hipStream_t s, int64_t n, const TimeCoeffs& c, const double* soln,
                            const double* prev, const double* prev2, const double* prev3,
                            const double* prev4, double* dot, const double* dotPrev, double* cur,
                            double* dotCur) {
  hipLaunchKernelGGL(k_iter_step, dim3(grid1(n)), dim3(kTPB), 0, s, n, c, soln, prev, prev2, prev3, prev4,
                     dot, dotPrev, cur, dotCur);
  return hipGetLastError();
}

hipError_t launch_assign_bc(hipStream_t s, int64_t n, const int32_t* ind, const double* val,
                            double timeFact, const double* soln, double* applied) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_assign_bc, dim3((unsigned)((n + kTPB - 1) / kTPB)), dim3(kTPB), 0, s, n, ind, val,
                     timeFact, soln, applied);
  return hipGetLastError();
}

hipError_t launch_set_bc(hipStream_t s, int64_t n, const int32_t* ind, const double* val, double timeFact,
                         double* x) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_set_bc, dim3((unsigned)((n + kTPB - 1) / kTPB)), dim3(kTPB), 0, s, n, ind, val, timeFact, x);
  return hipGetLastError();
}

hipError_t launch_add_bc(hipStream_t s, int64_t n, const int32_t* ind, const int32_t* mult,
                         const double* applied, double* soln) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_add_bc, dim3((unsigned)((n + kTPB - 1) / kTPB)), dim3(kTPB), 0, s, n, ind, mult,
                     applied, soln);
  return hipGetLastError();
}

hipError_t launch_update_solution(hipStream_t s, int64_t n, const int32_t* assy, const double* du,
                                  double* soln) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_update_solution, dim3(grid1(n)), dim3(kTPB), 0, s, n, assy, du, soln);
  return hipGetLastError();
}

hipError_t launch_zero_f64(hipStream_t s, int64_t n, double* a) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero<double>, dim3(grid1(n)), dim3(kTPB), 0, s, n, a);
  return hipGetLastError();
}

hipError_t launch_zero_f32(hipStream_t s, int64_t n, float* a) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero<float>, dim3(grid1(n)), dim3(kTPB), 0, s, n, a);
  return hipGetLastError();
}

}  // namespace sfem
