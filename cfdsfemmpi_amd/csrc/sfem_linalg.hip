// sfem_linalg.hip -- Krylov solvers and preconditioners on the device CSR.
//
// Replaces the PETSc KSP/PC calls behind SolverPetsc::solve
// (src/base/SolverPetsc.cpp:224-277): BiCGStab (-ksp_type bcgs) and
// restarted GMRES (-ksp_type gmres), right-preconditioned so that the stopping
// test ||b - A x||_2 <= rtol ||b||_2 is on the unpreconditioned residual, with a
// zero initial guess (VecZeroEntries at :249).  FP64 solves confirm it on the true
// residual b - A x (GMRES at every restart, BiCGStab when its recurrence residual
// meets the tolerance); FP32 inner solves of the mixed path stop on the recurrence
// residual, as PETSc's bcgs does.  Preconditioners: point Jacobi,
// nodal block Jacobi (the free DOFs of one node form a block of size <= 4),
// and ILU(0), factorised and applied level by level (one level per colour with
// the multicolour node ordering) in a padded 4 x 4 node-block form of the
// matrix; with ILU(0) the whole Krylov solve runs in that block form.
//
// Reductions are two-stage (per-block partials, then one block summing the
// partials in index order), so every dot product is deterministic.  Krylov
// scalars (alpha, omega, beta, ...) stay on the device; the host reads back
// one small block of scalars per iteration for the convergence test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <mutex>
#include <vector>

#include "sfem_internal.hpp"
#include "sfem_linalg.hpp"

namespace sfem {

constexpr int kRedBlocks = 1024;  // partials per reduction
constexpr int kTPB = 256;
constexpr int kGmresMaxH = 2 * (256 + 8) + 8;  // device / pinned Hessenberg scratch (restart <= 256)

// bfloat16 storage for the mixed path's ILU(0) factors: the top 16 bits of an IEEE
// float, rounded to nearest even.  Loads widen it back to float exactly.
struct bf16 {
  uint16_t bits;
};

template <typename AT>
__device__ inline AT widen(double v) { return (AT)v; }
template <typename AT>
__device__ inline AT widen(float v) { return (AT)v; }
template <typename AT>
__device__ inline AT widen(bf16 v) { return (AT)__uint_as_float(((uint32_t)v.bits) << 16); }

__device__ inline uint16_t bf16_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
// Stored value of an FP64 result in the storage type of a factor or tile inverse.
template <typename ST>
__device__ inline ST store_as(double v) { return (ST)v; }
template <>
__device__ inline bf16 store_as<bf16>(double v) { return bf16{bf16_bits((float)v)}; }

__global__ __launch_bounds__(256) void k_to_bf16(int64_t n, const float* __restrict__ a, bf16* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    b[i].bits = bf16_bits(a[i]);
}

// scalar slots
enum {
  S_RHO0 = 0, S_RHO1, S_ALPHA, S_OMEGA, S_BETA, S_RV, S_SS, S_TT, S_TS, S_RR, S_BB, S_TOL2,
  S_FLAG, S_MISC0, S_MISC1, S_RHRH, S_VV, S_NSLOTS = 32
};

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ inline double warp_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

template <int ND>
__device__ inline void block_sum_store(double (&v)[ND], double* out /*ND per block*/) {
  __shared__ double sh[ND][kTPB / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    double s = warp_sum(v[d]);
    if (lane == 0) sh[d][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < ND) {
    double s = 0.0;
    for (int w = 0; w < kTPB / 64; ++w) s += sh[threadIdx.x][w];
    out[(size_t)blockIdx.x * ND + threadIdx.x] = s;
  }
}

template <typename T>
__global__ __launch_bounds__(kTPB) void k_dot(int64_t n, const T* __restrict__ a,
                                              const T* __restrict__ b, double* part) {
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    v[0] += (double)a[i] * (double)b[i];
  block_sum_store<1>(v, part);
}

template <typename T>
__global__ __launch_bounds__(kTPB) void k_dot2(int64_t n, const T* __restrict__ a1,
                                               const T* __restrict__ b1, const T* __restrict__ a2,
                                               const T* __restrict__ b2, double* part) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB) {
    v[0] += (double)a1[i] * (double)b1[i];
    v[1] += (double)a2[i] * (double)b2[i];
  }
  block_sum_store<2>(v, part);
}

// Partial dot products of w with up to 8 vectors V_j (stored n apart): part[b*8 + j];
// absent vectors give 0.
template <typename T>
__global__ __launch_bounds__(kTPB) void k_mdot8(int64_t n, int nv, const T* __restrict__ V,
                                                const T* __restrict__ w, double* part) {
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB) {
    const double wi = (double)w[i];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j < nv) v[j] += (double)V[(int64_t)j * n + i] * wi;
  }
  block_sum_store<8>(v, part);
}

// w = w - sum_{j<nv} h[j] V_j  (sign = -1), or w = w + sum (sign = +1); h on the device.
template <typename T>
__global__ __launch_bounds__(kTPB) void k_mupdate(int64_t n, int nv, const T* __restrict__ V,
                                                  const double* __restrict__ h, double sign,
                                                  T* __restrict__ w) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB) {
    double acc = 0.0;
    for (int j = 0; j < nv; ++j) acc += h[j] * (double)V[(int64_t)j * n + i];
    w[i] = (T)((double)w[i] + sign * acc);
  }
}

// Finalisation ops: combine summed partials into Krylov scalars.
enum { FIN_STORE = 0, FIN_RHO1, FIN_ALPHA, FIN_SS, FIN_OMEGA, FIN_RR, FIN_RR_RHO };

template <int ND>
__global__ __launch_bounds__(kTPB) void k_finalize(const double* __restrict__ part, int nb,
                                                   double* scal, int slot, int op) {
  __shared__ double sh[ND][kTPB];
  for (int d = 0; d < ND; ++d) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += kTPB) s += part[(size_t)i * ND + d];
    sh[d][threadIdx.x] = s;
  }
  __syncthreads();
  for (int w = kTPB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int d = 0; d < ND; ++d) sh[d][threadIdx.x] += sh[d][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double r[ND];
  for (int d = 0; d < ND; ++d) r[d] = sh[d][0];
  switch (op) {
    case FIN_STORE:
      for (int d = 0; d < ND; ++d) scal[slot + d] = r[d];
      break;
    case FIN_RHO1: {  // rho1 = (rh, r); beta = (rho1/rho0)(alpha/omega)
      scal[S_RHO1] = r[0];
      const double rho0 = scal[S_RHO0], om = scal[S_OMEGA];
      // near breakdown: rh (almost) orthogonal to r -> restart the shadow residual
      if (!(fabs(r[0]) > 1e-12 * sqrt(scal[S_RHRH] * scal[S_RR])) || rho0 == 0.0 || om == 0.0) {
        scal[S_FLAG] = 3.0;
        scal[S_BETA] = 0.0;
      } else {
        scal[S_BETA] = (r[0] / rho0) * (scal[S_ALPHA] / om);
      }
      break;
    }
    case FIN_ALPHA:  // alpha = rho1 / (rh, v); r[1] = (v, v)
      scal[S_RV] = r[0];
      scal[S_VV] = ND > 1 ? r[ND - 1] : 0.0;
      if (scal[S_FLAG] == 3.0 || !(fabs(r[0]) > 1e-12 * sqrt(scal[S_RHRH] * scal[S_VV]))) {
        scal[S_FLAG] = 3.0;   // step with alpha = 0: a minimal-residual step on r
        scal[S_ALPHA] = 0.0;
      } else {
        scal[S_ALPHA] = scal[S_RHO1] / r[0];
      }
      break;
    case FIN_SS:
      scal[S_SS] = r[0];
      break;
    case FIN_OMEGA: {  // omega = (t, s)/(t, t); 0 when s already meets the tolerance
      scal[S_TT] = r[0];
      scal[S_TS] = r[1];
      const bool half = scal[S_SS] <= scal[S_TOL2];
      scal[S_OMEGA] = (half || r[0] == 0.0) ? 0.0 : r[1] / r[0];
      if (half) scal[S_FLAG] = 1.0;
      break;
    }
    case FIN_RR:
      scal[S_RR] = r[0];
      scal[S_RHO0] = scal[S_RHO1];
      break;
    case FIN_RR_RHO: {  // r[0] = (r, r); r[1] = (rh, r) = the next iteration's rho1, and its beta
      scal[S_RR] = r[0];
      const double rho0 = scal[S_RHO1], om = scal[S_OMEGA], rho1 = ND > 1 ? r[1] : 0.0;
      scal[S_RHO0] = rho0;
      scal[S_RHO1] = rho1;
      if (!(fabs(rho1) > 1e-12 * sqrt(scal[S_RHRH] * r[0])) || rho0 == 0.0 || om == 0.0) {
        if (scal[S_FLAG] != 1.0) scal[S_FLAG] = 3.0;
        scal[S_BETA] = 0.0;
      } else {
        scal[S_BETA] = (rho1 / rho0) * (scal[S_ALPHA] / om);
      }
      break;
    }
  }
}

// ---------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kTPB) void k_copy(int64_t n, const T* __restrict__ a, T* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    b[i] = a[i];
}

template <typename TA, typename TB>
__global__ __launch_bounds__(kTPB) void k_convert(int64_t n, const TA* __restrict__ a,
                                                  TB* __restrict__ b, double scale) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    b[i] = (TB)((double)a[i] * scale);
}

template <typename T>
__global__ __launch_bounds__(kTPB) void k_fill(int64_t n, T* __restrict__ a, T v) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    a[i] = v;
}

// p = r + beta (p - omega v)   (first iteration: p = r)
template <typename T>
__global__ __launch_bounds__(kTPB) void k_bicg_p(int64_t n, const T* __restrict__ r, T* __restrict__ p,
                                                 const T* __restrict__ v, const double* scal, int first) {
  const T beta = first ? (T)0 : (T)scal[S_BETA];
  const T om = first ? (T)0 : (T)scal[S_OMEGA];
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    p[i] = first ? r[i] : r[i] + beta * (p[i] - om * v[i]);
}

// s = r - alpha v ; partial (s, s)
template <typename T>
__global__ __launch_bounds__(kTPB) void k_bicg_s(int64_t n, const T* __restrict__ r,
                                                 const T* __restrict__ v, T* __restrict__ s,
                                                 const double* scal, double* part) {
  const T alpha = (T)scal[S_ALPHA];
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB) {
    const T si = r[i] - alpha * v[i];
    s[i] = si;
    acc[0] += (double)si * (double)si;
  }
  block_sum_store<1>(acc, part);
}

// x += alpha ph + omega sh ; r = s - omega t ; partials (r, r) and (rh, r)
template <typename T>
__global__ __launch_bounds__(kTPB) void k_bicg_xr(int64_t n, T* __restrict__ x, const T* __restrict__ ph,
                                                  const T* __restrict__ sh, T* __restrict__ r,
                                                  const T* __restrict__ s, const T* __restrict__ t,
                                                  const T* __restrict__ rh, const double* scal, double* part) {
  const T alpha = (T)scal[S_ALPHA], om = (T)scal[S_OMEGA];
  double acc[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB) {
    x[i] += alpha * ph[i] + om * sh[i];
    const T ri = s[i] - om * t[i];
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
    acc[1] += (double)rh[i] * (double)ri;
  }
  block_sum_store<2>(acc, part);
}

// y = a*x + b*y with host scalars (GMRES)
template <typename T>
__global__ __launch_bounds__(kTPB) void k_axpby(int64_t n, double a, const T* __restrict__ x, double b,
                                                T* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    y[i] = (T)(a * (double)x[i] + b * (double)y[i]);
}

// ---------------------------------------------------------------------------
// Jacobi and nodal block Jacobi
// ---------------------------------------------------------------------------
__device__ inline double csr_lookup(const int32_t* rowptr, const int32_t* colidx, const double* vals,
                                    const float* vals32, int64_t r, int32_t c, bool& found) {
  int32_t lo = rowptr[r], hi = rowptr[r + 1];
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (colidx[mid] < c) lo = mid + 1; else hi = mid;
  }
  found = (lo < rowptr[r + 1] && colidx[lo] == c);
  if (!found) return 0.0;
  return vals ? vals[lo] : (double)vals32[lo];
}

// Inverse of each nodal diagonal block (Gauss-Jordan with partial pivoting, m <= 4).
template <typename PT>
__global__ __launch_bounds__(kTPB) void k_bjac_setup(int64_t nb, const int32_t* __restrict__ blkStart,
                                                     const int32_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ colidx,
                                                     const double* __restrict__ vals,
                                                     const float* __restrict__ vals32,
                                                     PT* __restrict__ inv, double* scal) {
  const int64_t k = (int64_t)blockIdx.x * kTPB + threadIdx.x;
  if (k >= nb) return;
  const int32_t r0 = blkStart[k];
  const int m = blkStart[k + 1] - r0;
  double a[4][8];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      bool f;
      a[i][j] = csr_lookup(rowptr, colidx, vals, vals32, r0 + i, r0 + j, f);
      a[i][m + j] = (i == j) ? 1.0 : 0.0;
    }
  bool ok = true;
  for (int c = 0; c < m; ++c) {
    int p = c;
    for (int r = c + 1; r < m; ++r)
      if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
    if (a[p][c] == 0.0) { ok = false; break; }
    if (p != c)
      for (int j = 0; j < 2 * m; ++j) { double t = a[p][j]; a[p][j] = a[c][j]; a[c][j] = t; }
    const double piv = 1.0 / a[c][c];
    for (int j = 0; j < 2 * m; ++j) a[c][j] *= piv;
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      if (f == 0.0) continue;
      for (int j = 0; j < 2 * m; ++j) a[r][j] -= f * a[c][j];
    }
  }
  if (!ok) { scal[S_MISC1] = 1.0; for (int i = 0; i < m; ++i) a[i][m + i] = 1.0; }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) inv[k * 16 + i * 4 + j] = (PT)((i < m && j < m) ? a[i][m + j] : 0.0);
}

template <typename T, typename PT>
__global__ __launch_bounds__(kTPB) void k_bjac_apply(int64_t nb, const int32_t* __restrict__ blkStart,
                                                     const PT* __restrict__ inv, const T* __restrict__ r,
                                                     T* __restrict__ z) {
  const int64_t k = (int64_t)blockIdx.x * kTPB + threadIdx.x;
  if (k >= nb) return;
  const int32_t r0 = blkStart[k];
  const int m = blkStart[k + 1] - r0;
  T rv[4];
  for (int j = 0; j < 4; ++j) rv[j] = (j < m) ? r[r0 + j] : (T)0;
  const PT* B = inv + k * 16;
  for (int i = 0; i < m; ++i) {
    T s = 0;
    for (int j = 0; j < 4; ++j) s += (T)B[i * 4 + j] * rv[j];
    z[r0 + i] = s;
  }
}

template <typename PT>
__global__ __launch_bounds__(kTPB) void k_jac_setup(int64_t n, const int32_t* __restrict__ diagpos,
                                                    const double* __restrict__ vals,
                                                    const float* __restrict__ vals32,
                                                    PT* __restrict__ dinv, double* scal) {
  const int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x;
  if (i >= n) return;
  const double d = vals ? vals[diagpos[i]] : (double)vals32[diagpos[i]];
  if (d == 0.0) scal[S_MISC1] = 1.0;
  dinv[i] = (PT)(d == 0.0 ? 1.0 : 1.0 / d);
}

template <typename T, typename PT>
__global__ __launch_bounds__(kTPB) void k_jac_apply(int64_t n, const PT* __restrict__ dinv,
                                                    const T* __restrict__ r, T* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    z[i] = (T)dinv[i] * r[i];
}

// ---------------------------------------------------------------------------
// shared helpers of the node-grouped kernels
// ---------------------------------------------------------------------------
__device__ inline int lds_find(const int32_t* cols, int len, int32_t c) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cols[mid] < c) lo = mid + 1; else hi = mid;
  }
  return (lo < len && cols[lo] == c) ? lo : -1;
}

template <int G, typename AT>
__device__ inline AT group_sum(AT v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

template <typename AT>
__device__ inline AT pick4(const AT (&v)[4], int i) {
  return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3];
}

// Node block descriptor: first row, first value of that row, row length, and
// (diagonal-block column offset) | (row count << 16).
struct __align__(16) BlkDesc {
  int32_t r0, base, rl, dqm;
};

// Ordering point for the lanes of one wavefront exchanging data through LDS: a
// workgroup of exactly one wavefront needs no s_barrier (its LDS operations complete in
// issue order), only fences that keep the compiler from moving LDS accesses across this
// point; unlike __syncthreads() it does not wait for outstanding global loads.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// acc[a] += sum_{q in [q0, q1)} V[a*rl + q] * x[col[q]] for a < m, lanes of a group of G
// stepping by G; U columns per lane are loaded before any is used, so a lane keeps
// U index loads, U gathers and U*m value loads in flight.
template <int G, int U, typename VT, typename XT, typename AT>
__device__ inline void group_dot(const int32_t* __restrict__ col, const VT* __restrict__ V, const XT* __restrict__ x,
                                 int32_t rl, int m, int32_t q0, int32_t q1, int lane, AT (&acc)[4]) {
  int32_t q = q0 + lane;
  for (; q + (U - 1) * G < q1; q += U * G) {
    int32_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = col[q + u * G];
    AT v[4][U];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int u = 0; u < U; ++u) v[a][u] = (a < m) ? widen<AT>(V[(int64_t)a * rl + q + u * G]) : (AT)0;
    AT xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = (AT)x[c[u]];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int a = 0; a < 4; ++a) acc[a] += v[a][u] * xv[u];
  }
  for (; q < q1; q += G) {
    const AT xv = (AT)x[col[q]];
#pragma unroll
    for (int a = 0; a < 4; ++a)
      if (a < m) acc[a] += widen<AT>(V[(int64_t)a * rl + q]) * xv;
  }
}

// Logical workgroup index.  Workgroups are dispatched round-robin over the 8 XCDs;
// with xcd != 0 each XCD walks one contiguous range of logical workgroups instead, so
// that neighbouring node blocks, which gather overlapping x entries, share an L2.
__device__ inline int64_t logical_wg(int xcd) {
  const uint32_t b = blockIdx.x;
  if (!xcd) return b;
  const uint32_t q = gridDim.x / 8, r = gridDim.x % 8, x = b % 8, k = b / 8;
  return (int64_t)(x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// As group_dot, with every column index and value of up to K columns per lane loaded
// before any x entry is gathered, so a lane has all its loads in flight at once (two
// dependent memory round trips per node block instead of two per U columns); columns
// beyond K*G per lane fall back to group_dot.
template <int G, int K, typename VT, typename XT, typename AT>
__device__ inline void group_dot_pf(const int32_t* __restrict__ col, const VT* __restrict__ V,
                                    const XT* __restrict__ x, int32_t rl, int m, int32_t q0, int32_t q1,
                                    int lane, AT (&acc)[4]) {
  int32_t c[K];
  AT v[4][K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int32_t q = q0 + lane + k * G;
    const bool ok = q < q1;
    c[k] = ok ? col[q] : -1;
#pragma unroll
    for (int a = 0; a < 4; ++a) v[a][k] = (ok && a < m) ? widen<AT>(V[(int64_t)a * rl + q]) : (AT)0;
  }
  AT xv[K];
#pragma unroll
  for (int k = 0; k < K; ++k) xv[k] = c[k] >= 0 ? (AT)x[c[k]] : (AT)0;
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int a = 0; a < 4; ++a) acc[a] += v[a][k] * xv[k];
  if (q0 + K * G < q1) group_dot<G, 2>(col, V, x, rl, m, q0 + K * G, q1, lane, acc);
}

// U > 0: group_dot with U columns in flight per lane; U == 0: group_dot_pf covering
// 128 columns per node block before the fallback.
template <int G, int U, typename VT, typename XT, typename AT>
__device__ inline void group_dot_v(const int32_t* __restrict__ col, const VT* __restrict__ V,
                                   const XT* __restrict__ x, int32_t rl, int m, int32_t q0, int32_t q1,
                                   int lane, AT (&acc)[4]) {
  if constexpr (U == 0) group_dot_pf<G, 128 / G>(col, V, x, rl, m, q0, q1, lane, acc);
  else group_dot<G, U>(col, V, x, rl, m, q0, q1, lane, acc);
}

// Node-grouped CSR SpMV.  The free rows of one node (a block of m <= 4 rows) share one
// column list, so G lanes walk that list once, load each column index and x entry once,
// and accumulate the m rows together; values of row a sit at base + a*rl + q.
template <typename VT, typename XT, int G, int U>
__global__ __launch_bounds__(256) void k_spmv_blk(int64_t nb, const BlkDesc* __restrict__ desc,
                                                  const int32_t* __restrict__ colidx,
                                                  const VT* __restrict__ vals, const XT* __restrict__ x,
                                                  XT* __restrict__ y, int xcd) {
  using AT = typename std::conditional<sizeof(VT) == 8, double, float>::type;
  const int64_t gt = logical_wg(xcd) * 256 + threadIdx.x;
  const int64_t b = gt / G;
  const int lane = (int)(gt % G);
  AT acc[4] = {0, 0, 0, 0};
  BlkDesc d = {0, 0, 0, 0};
  if (b < nb) {
    d = desc[b];
    group_dot_v<G, U>(colidx + d.base, vals + d.base, x, d.rl, d.dqm >> 16, 0, d.rl, lane, acc);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = group_sum<G>(acc[a]);
  if (b < nb && lane < (d.dqm >> 16)) y[d.r0 + lane] = (XT)pick4(acc, lane);
}

// ---------------------------------------------------------------------------
// ILU(0) in node-block (BSR4) form.  Every node block with a free DOF is a block row
// and a block column of width 4: its m <= 4 free DOFs occupy slots 0..m-1, slots
// m..3 are padding.  A block row p stores, for each neighbouring block (bcol, sorted
// in the plan's order), a 4 x 4 block whose row a (a < m) lies at
// vbase + a*4*nb + 4*t: so lane t of a group reads row a of its block with one
// 16-byte load (8 bytes in BF16) and the neighbour's 4 x entries with one more.
// Padding columns hold 0 and padding slots of every vector the kernels write are 0.
// ILU(0) on this block pattern equals point ILU(0) on the CSR pattern when the CSR
// pattern is a union of full node blocks (element patterns are): the padded
// columns stay 0 through the elimination.
// ---------------------------------------------------------------------------
constexpr int kBsrMaxNb = 64;  // neighbour blocks per block row handled (checked on the host)

// Block row descriptor: first value, first neighbour entry, neighbour count, and
// (position of the diagonal block) | (free rows m << 16).
struct __align__(16) BDesc {
  int32_t vbase, bptr, nb, dm;
};

template <typename AT> struct V4 { AT v[4]; };

template <typename AT>
__device__ inline V4<AT> ld4(const double* p) {
  const double4 f = *(const double4*)p;
  return {{(AT)f.x, (AT)f.y, (AT)f.z, (AT)f.w}};
}
template <typename AT>
__device__ inline V4<AT> ld4(const float* p) {
  const float4 f = *(const float4*)p;
  return {{(AT)f.x, (AT)f.y, (AT)f.z, (AT)f.w}};
}
__device__ inline void st4(double* p, double a, double b, double c, double d) {
  reinterpret_cast<double2*>(p)[0] = make_double2(a, b);
  reinterpret_cast<double2*>(p)[1] = make_double2(c, d);
}
__device__ inline void st4(float* p, float a, float b, float c, float d) {
  *reinterpret_cast<float4*>(p) = make_float4(a, b, c, d);
}

template <typename AT>
__device__ inline V4<AT> ld4(const bf16* p) {
  const uint2 u = *(const uint2*)p;
  return {{(AT)__uint_as_float(u.x << 16), (AT)__uint_as_float(u.x & 0xffff0000u),
           (AT)__uint_as_float(u.y << 16), (AT)__uint_as_float(u.y & 0xffff0000u)}};
}

// A whole 4 x 4 block stored block by block (16 consecutive values, rows a >= m zero):
// the widest loads the storage type allows (two 16-byte loads in BF16, four in FP32).
template <typename AT>
struct Blk16 {
  V4<AT> r[4];
};
template <typename AT>
__device__ inline Blk16<AT> ld_blk16(const double* p) {
  Blk16<AT> b;
#pragma unroll
  for (int a = 0; a < 4; ++a) b.r[a] = ld4<AT>(p + 4 * a);
  return b;
}
template <typename AT>
__device__ inline Blk16<AT> ld_blk16(const float* p) {
  Blk16<AT> b;
#pragma unroll
  for (int a = 0; a < 4; ++a) b.r[a] = ld4<AT>(p + 4 * a);
  return b;
}
template <typename AT>
__device__ inline Blk16<AT> ld_blk16(const bf16* p) {
  const uint4 u0 = reinterpret_cast<const uint4*>(p)[0], u1 = reinterpret_cast<const uint4*>(p)[1];
  const uint32_t w[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
  Blk16<AT> b;
#pragma unroll
  for (int a = 0; a < 4; ++a)
    b.r[a] = {{(AT)__uint_as_float(w[2 * a] << 16), (AT)__uint_as_float(w[2 * a] & 0xffff0000u),
               (AT)__uint_as_float(w[2 * a + 1] << 16), (AT)__uint_as_float(w[2 * a + 1] & 0xffff0000u)}};
  return b;
}

// acc[a] += sum over the blocks of the copy, V[16 t ..), block t against x(bc[t]), for
// t in [t0, t1): every row of a block loaded with the block's wide loads (rows a >= m
// of the copy are zero, so they add nothing).  Lanes of a group of G step over t; the
// first K blocks of every lane are loaded before any is used.
template <int G, int K, typename VT, typename XT, typename AT>
__device__ inline void blk16_dot(const int32_t* __restrict__ bc, const VT* __restrict__ V,
                                 const XT* __restrict__ x, int t0, int t1, int lane, AT (&acc)[4]) {
  int32_t c[K];
  Blk16<AT> v[K];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int t = t0 + lane + s * G;
    const bool ok = t < t1;
    c[s] = ok ? bc[t] : -1;
    if (ok) v[s] = ld_blk16<AT>(V + 16 * (int64_t)t);
  }
  V4<AT> xv[K];
#pragma unroll
  for (int s = 0; s < K; ++s) xv[s] = c[s] >= 0 ? ld4<AT>(x + 4 * (int64_t)c[s]) : V4<AT>{{0, 0, 0, 0}};
#pragma unroll
  for (int s = 0; s < K; ++s)
    if (c[s] >= 0) {
#pragma unroll
      for (int a = 0; a < 4; ++a)
        acc[a] += v[s].r[a].v[0] * xv[s].v[0] + v[s].r[a].v[1] * xv[s].v[1] + v[s].r[a].v[2] * xv[s].v[2] +
                  v[s].r[a].v[3] * xv[s].v[3];
    }
  for (int t = t0 + lane + K * G; t < t1; t += G) {
    const V4<AT> xx = ld4<AT>(x + 4 * (int64_t)bc[t]);
    const Blk16<AT> w = ld_blk16<AT>(V + 16 * (int64_t)t);
#pragma unroll
    for (int a = 0; a < 4; ++a)
      acc[a] += w.r[a].v[0] * xx.v[0] + w.r[a].v[1] * xx.v[1] + w.r[a].v[2] * xx.v[2] + w.r[a].v[3] * xx.v[3];
  }
}

// acc[a] += sum over blocks t in [t0, t1) of V(a, t) . x(bcol[t]) for a < m, row a of
// block t at V + a rs + BS t (BS = 4: a block row's rows interleaved across its blocks;
// BS = 16, rs = 4: block by block).  Lanes of a group of G step over t; the first K blocks of every lane are loaded (index, x and m
// value rows) before any is used, the rest one at a time.
template <int G, int K, int BS = 4, typename VT, typename XT, typename AT>
__device__ inline void bsr_dot(const int32_t* __restrict__ bc, const VT* __restrict__ V, int32_t rs,
                               const XT* __restrict__ x, int m, int t0, int t1, int lane, AT (&acc)[4]) {
  int32_t c[K];
  V4<AT> v[4][K];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int t = t0 + lane + s * G;
    const bool ok = t < t1;
    c[s] = ok ? bc[t] : -1;
#pragma unroll
    for (int a = 0; a < 4; ++a)
      v[a][s] = (ok && a < m) ? ld4<AT>(V + (int64_t)a * rs + BS * t) : V4<AT>{{0, 0, 0, 0}};
  }
  V4<AT> xv[K];
#pragma unroll
  for (int s = 0; s < K; ++s) xv[s] = c[s] >= 0 ? ld4<AT>(x + 4 * (int64_t)c[s]) : V4<AT>{{0, 0, 0, 0}};
#pragma unroll
  for (int s = 0; s < K; ++s)
#pragma unroll
    for (int a = 0; a < 4; ++a)
      acc[a] += v[a][s].v[0] * xv[s].v[0] + v[a][s].v[1] * xv[s].v[1] + v[a][s].v[2] * xv[s].v[2] +
                v[a][s].v[3] * xv[s].v[3];
  for (int t = t0 + lane + K * G; t < t1; t += G) {
    const V4<AT> xx = ld4<AT>(x + 4 * (int64_t)bc[t]);
#pragma unroll
    for (int a = 0; a < 4; ++a)
      if (a < m) {
        const V4<AT> w = ld4<AT>(V + (int64_t)a * rs + BS * t);
        acc[a] += w.v[0] * xx.v[0] + w.v[1] * xx.v[1] + w.v[2] * xx.v[2] + w.v[3] * xx.v[3];
      }
  }
}

// y = K x in block form; G lanes per block row, padding slots of y written as 0.  With
// rows non-null the g-th group takes block row rows[g]: visiting the rows in the
// original (spatial) node order keeps the x entries a stretch of rows gathers in L2,
// where the multicolour order would sweep all of x once per colour.
template <typename VT, typename XT, int G, int K>
__global__ __launch_bounds__(256) void k_bsr_spmv(int64_t nblk, const BDesc* __restrict__ desc,
                                                  const int32_t* __restrict__ bcol, const VT* __restrict__ V,
                                                  const XT* __restrict__ x, XT* __restrict__ y, int xcd,
                                                  const int32_t* __restrict__ rows) {
  using AT = typename std::conditional<(sizeof(VT) == 8 || sizeof(XT) == 8), double, float>::type;
  const int64_t gt = logical_wg(xcd) * 256 + threadIdx.x;
  const int64_t g = gt / G;
  const int lane = (int)(gt % G);
  AT acc[4] = {0, 0, 0, 0};
  BDesc d = {0, 0, 0, 0};
  int64_t p = g;
  if (g < nblk) {
    if (rows) p = rows[g];
    d = desc[p];
    bsr_dot<G, K>(bcol + d.bptr, V + d.vbase, 4 * d.nb, x, d.dm >> 16, 0, d.nb, lane, acc);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = group_sum<G>(acc[a]);
  if (g < nblk && lane < 4) y[4 * p + lane] = lane < (d.dm >> 16) ? (XT)pick4(acc, lane) : (XT)0;
}

// Accumulation type of the triangular solves: FP64 when the vectors or the factor are
// FP64 (an FP32- or BF16-stored factor applied to FP64 vectors accumulates in FP64).
template <typename T, typename LT>
using TriAcc = typename std::conditional<(sizeof(T) == 8 || sizeof(LT) == 8), double, float>::type;

// The diagonal block of block row d, row `lane` loaded by lane < m (before the long
// sparse product) and broadcast to the group.
template <int G, typename AT>
__device__ inline void bcast_diag(const V4<AT>& dv, AT (&D)[16]) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) D[a * 4 + b] = __shfl(dv.v[b], a, G);
}

// Forward solve L y = r over the block rows of one level (unit-lower diagonal blocks).
// Level rows are the contiguous range [first, first + count) or list[first + g].
// With fuse_u the level's rows have no upper neighbour blocks (the last colour of a
// multicolour plan), so their backward solve is the diagonal block alone and is done here.
template <typename T, typename LT, int G, int K>
__global__ __launch_bounds__(256) void k_bsr_lsolve(const int32_t* __restrict__ list, int first, int count,
                                                    const BDesc* __restrict__ desc,
                                                    const int32_t* __restrict__ bcol, const LT* __restrict__ lu,
                                                    const T* __restrict__ r, T* __restrict__ y, int xcd,
                                                    int fuse_u) {
  using AT = TriAcc<T, LT>;
  const int64_t gt = logical_wg(xcd) * 256 + threadIdx.x;
  const int64_t g = gt / G;
  const int lane = (int)(gt % G);
  AT acc[4] = {0, 0, 0, 0};
  BDesc d = {0, 0, 0, 0};
  int32_t p = 0;
  AT rv = 0;
  V4<AT> dv = {{0, 0, 0, 0}};
  const bool act = g < count;
  if (act) {
    p = list ? list[first + g] : first + (int32_t)g;
    d = desc[p];
    const int m = d.dm >> 16, td = d.dm & 0xffff;
    if (lane < 4) rv = (AT)r[4 * (int64_t)p + lane];
    if (lane < m) dv = ld4<AT>(lu + d.vbase + (int64_t)lane * 4 * d.nb + 4 * td);
    bsr_dot<G, K>(bcol + d.bptr, lu + d.vbase, 4 * d.nb, y, m, 0, td, lane, acc);
  }
  AT D[16], R[4];
  bcast_diag<G>(dv, D);
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = group_sum<G>(acc[a]);
#pragma unroll
  for (int a = 0; a < 4; ++a) R[a] = __shfl(rv, a, G);
  const int m = d.dm >> 16;
  AT v[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    AT s = R[a] - acc[a];
#pragma unroll
    for (int b = 0; b < a; ++b) s -= D[a * 4 + b] * v[b];
    v[a] = a < m ? s : (AT)0;
  }
  if (fuse_u) {
#pragma unroll
    for (int a = 3; a >= 0; --a) {
      if (a < m) {
        AT s = v[a];
#pragma unroll
        for (int b = a + 1; b < 4; ++b) s -= D[a * 4 + b] * v[b];
        v[a] = s / D[a * 4 + a];
      }
    }
  }
  if (act && lane < 4) y[4 * (int64_t)p + lane] = (T)pick4(v, lane);
}

// Backward solve U z = y over one level (levels in reverse order), in place on y.
template <typename T, typename LT, int G, int K>
__global__ __launch_bounds__(256) void k_bsr_usolve(const int32_t* __restrict__ list, int first, int count,
                                                    const BDesc* __restrict__ desc,
                                                    const int32_t* __restrict__ bcol, const LT* __restrict__ lu,
                                                    T* __restrict__ y, int xcd) {
  using AT = TriAcc<T, LT>;
  const int64_t gt = logical_wg(xcd) * 256 + threadIdx.x;
  const int64_t g = gt / G;
  const int lane = (int)(gt % G);
  AT acc[4] = {0, 0, 0, 0};
  BDesc d = {0, 0, 0, 0};
  int32_t p = 0;
  AT rv = 0;
  V4<AT> dv = {{0, 0, 0, 0}};
  const bool act = g < count;
  if (act) {
    p = list ? list[first + g] : first + (int32_t)g;
    d = desc[p];
    const int m = d.dm >> 16, td = d.dm & 0xffff;
    if (lane < 4) rv = (AT)y[4 * (int64_t)p + lane];
    if (lane < m) dv = ld4<AT>(lu + d.vbase + (int64_t)lane * 4 * d.nb + 4 * td);
    bsr_dot<G, K>(bcol + d.bptr, lu + d.vbase, 4 * d.nb, y, m, td + 1, d.nb, lane, acc);
  }
  AT D[16], R[4];
  bcast_diag<G>(dv, D);
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = group_sum<G>(acc[a]);
#pragma unroll
  for (int a = 0; a < 4; ++a) R[a] = __shfl(rv, a, G);
  const int m = d.dm >> 16;
  AT v[4] = {0, 0, 0, 0};
#pragma unroll
  for (int a = 3; a >= 0; --a) {
    if (a < m) {
      AT s = R[a] - acc[a];
#pragma unroll
      for (int b = a + 1; b < 4; ++b) s -= D[a * 4 + b] * v[b];
      v[a] = s / D[a * 4 + a];
    }
  }
  if (act && lane < 4) y[4 * (int64_t)p + lane] = (T)pick4(v, lane);
}

// Triangular solves of a tiled plan, one launch per colour and one workgroup of 8 G
// lanes per tile, G lanes per block row of the tile (row g = lanes G g .. G g + G - 1).
// Every row first sums its blocks outside the tile against the vector entries earlier
// launches wrote (lower blocks forward, upper blocks backward).  The tile's own 32 x 32
// triangle (the 4 slots of its 8 rows) is then applied as a product with its inverse,
// computed once per factorisation (k_tile_invert), so the rows of a tile need no
// sequential steps: slot i = 4 g + a (lane a < 4 of row g) takes the dot product of row
// i of the inverse with the tile's right-hand sides, staged in LDS.
constexpr int kTileMax = 8;
constexpr int kTileN = 4 * kTileMax;                 // slots of a tile
constexpr int kTileTri = kTileN * (kTileN + 1) / 2;  // packed triangle
constexpr int kTileInv = 2 * kTileTri;               // L^{-1} then U^{-1}

// Packed offsets: row i of the lower triangle holds columns 0..i, row i of the upper
// triangle columns i..kTileN-1.
__host__ __device__ constexpr int tri_lo(int i) { return i * (i + 1) / 2; }
__host__ __device__ constexpr int tri_up(int i) { return kTileTri + i * kTileN - i * (i - 1) / 2; }

// Orders the tile's LDS exchanges: a wavefront-scope barrier when the tile is one
// wavefront, the workgroup barrier otherwise.
template <int G>
__device__ inline void tile_sync() {
  if constexpr (G * kTileMax <= 64) wave_sync();  // the tile is one wavefront (or part of one)
  else __syncthreads();                            // one tile per workgroup
}

// Inverses of the triangular factors of each tile's own 32 x 32 block (padding slots and
// missing rows are identity): column j of L^{-1} (unit lower) and of U^{-1} by lane j, in
// FP64, from the factor values the sweeps read.  One 64-lane workgroup per tile.
template <typename LT, typename IT>
__global__ __launch_bounds__(64) void k_tile_invert(const int32_t* __restrict__ tstart, int ntiles,
                                                    const BDesc* __restrict__ desc, const LT* __restrict__ lu,
                                                    IT* __restrict__ tinv, double* __restrict__ scal) {
  __shared__ double A[kTileN][kTileN + 1];
  __shared__ double X[kTileN][kTileN + 1];
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  const int lane = threadIdx.x;
  const int ts = tstart[t], tsz = tstart[t + 1] - ts;
  for (int q = lane; q < kTileN * kTileN; q += 64) {
    const int i = q / kTileN, j = q % kTileN;
    A[i][j] = i == j ? 1.0 : 0.0;
  }
  wave_sync();
  // row a of block (g, h) of the tile: lane loops over (g, a, h)
  for (int q = lane; q < tsz * 4 * tsz; q += 64) {
    const int g = q / (4 * tsz), a = (q / tsz) % 4, h = q % tsz;
    const BDesc d = desc[ts + g];
    const int m = d.dm >> 16, td = d.dm & 0xffff;
    if (a >= m) continue;
    const V4<double> v = ld4<double>(lu + d.vbase + (int64_t)a * 4 * d.nb + 4 * (td - g + h));
#pragma unroll
    for (int b = 0; b < 4; ++b) A[4 * g + a][4 * h + b] = v.v[b];
  }
  wave_sync();
  // Column j of each inverse is lane j's, held in registers while it is formed; the
  // rows of A are read by every lane at once (one LDS broadcast per entry).  The sums
  // run over all earlier (later) rows, whose entries of the column are 0 before the
  // column's own row, so each entry takes the same subtractions, in the same order, as
  // substitution from the column's diagonal on.
  const int j = lane;
  double xc[kTileN];
  if (lane < kTileN) {
    // L^{-1} e_j: unit lower triangle of A
#pragma unroll
    for (int i = 0; i < kTileN; ++i) {
      double x = 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) x -= A[i][k] * xc[k];
      xc[i] = i < j ? 0.0 : (i == j ? 1.0 : x);
    }
#pragma unroll
    for (int i = 0; i < kTileN; ++i) X[i][j] = xc[i];
  }
  wave_sync();
  IT* out = tinv + (int64_t)t * kTileInv;
  for (int q = lane; q < kTileTri; q += 64) {
    int i = 0;
    while (tri_lo(i + 1) <= q) ++i;
    out[q] = store_as<IT>(X[i][q - tri_lo(i)]);
  }
  wave_sync();
  if (lane < kTileN) {
    // U^{-1} e_j: upper triangle of A with its diagonal
#pragma unroll
    for (int i = kTileN - 1; i >= 0; --i) {
      double x = 0.0;
#pragma unroll
      for (int k = i + 1; k < kTileN; ++k) x -= A[i][k] * xc[k];
      const double di = A[i][i];
      if (i <= j && di == 0.0) scal[S_MISC1] = 2.0;
      const double q = i == j ? 1.0 / (di == 0.0 ? 1.0 : di) : x / (di == 0.0 ? 1.0 : di);
      xc[i] = i > j ? 0.0 : q;
    }
#pragma unroll
    for (int i = 0; i < kTileN; ++i) X[i][j] = xc[i];
  }
  wave_sync();
  for (int q = lane; q < kTileTri; q += 64) {
    int i = 0;
    while (i + 1 < kTileN && tri_up(i + 1) - kTileTri <= q) ++i;
    out[kTileTri + q] = store_as<IT>(X[i][i + (q - (tri_up(i) - kTileTri))]);
  }
}

// One packed triangle of a tile inverse (kTileTri values) moved global -> registers ->
// LDS in 16-byte pieces by all NT threads of the workgroup: coalesced vector loads,
// issued before the sums over the blocks outside the tile and stored to LDS after them.
template <typename IT, int NT>
struct TriStage {
  static constexpr int PV = 16 / (int)sizeof(IT);
  static constexpr int NV = kTileTri / PV;
  static constexpr int R = (NV + NT - 1) / NT;
  static_assert(kTileTri % PV == 0, "packed triangle is a whole number of 16-byte pieces");
  uint4 v[R];
  __device__ inline void load(const IT* __restrict__ src, int tid) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = tid + r * NT;
      // every piece but the last is in range for every thread ((R - 1) NT <= NV)
      const int qc = (r + 1 < R || q < NV) ? q : NV - 1;
      v[r] = *reinterpret_cast<const uint4*>(src + qc * PV);
    }
  }
  __device__ inline void store(IT* dst, int tid) const {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = tid + r * NT;
      if (q < NV) *reinterpret_cast<uint4*>(dst + q * PV) = v[r];
    }
  }
};

// Row ii = 4 g + (lane & 3) of a tile's triangle times the tile's right-hand sides: the
// G lanes of group g split the 32 columns, G / 4 lanes per row (column j goes to lane
// j mod (G / 4)), and are summed by shuffles; every lane of the row gets the sum.
// `row` points at the packed row with row[j] the entry of column j; only columns with
// lo <= j <= hi are present.
template <int G, typename IT, typename AT>
__device__ inline AT tile_tri_row(const IT* row, const AT* v, int lane, int lo, int hi) {
  constexpr int S = G / 4;
  const int sub = lane >> 2;
  AT part = 0;
#pragma unroll
  for (int jj = 0; jj < kTileN / S; ++jj) {
    const int j = jj * S + sub;
    if (j >= lo && j <= hi) part += widen<AT>(row[j]) * v[j];
  }
#pragma unroll
  for (int o = G / 2; o >= 4; o >>= 1) part += __shfl_xor(part, o, G);
  return part;
}

// Forward sweep of one colour: y = L^{-1} r on its tiles.  With fuse_u the tiles have no
// upper blocks outside themselves (the last colour) and their backward sweep follows.
// Row g of tile tg in a sweep: rec = trow[tg * kTileMax + g] = {block row p (-1 past the
// tile's rows), first of its blocks outside the tile in bcol, their count | m << 16, their
// first block in the sweeps' block-by-block copy}, so a workgroup reads its rows' data with
// one load indexed by the tile.
// Tiles per workgroup: a tile takes G * kTileMax lanes.  Tiles of at most one wavefront
// are packed 256 lanes to a workgroup (G = 4: eight tiles, two per wavefront; G = 8:
// four, one per wavefront), since a CU holds a limited number of workgroups and
// one-wavefront workgroups would leave its wavefront slots half empty.  A tile never
// spans two wavefronts then, so its exchanges need only wavefront-scope ordering and a
// tile past the launch's last one leaves without stalling the others.
template <int G>
__host__ __device__ constexpr int tile_tpw() {
  return G * kTileMax <= 64 ? 256 / (G * kTileMax) : 1;
}

template <typename T, typename LT, typename IT, int G, int K>
__global__ __launch_bounds__(G * kTileMax * tile_tpw<G>()) void k_bsr_tile_lsolve(
    int tfirst, int tcount, const int4* __restrict__ trow, const int32_t* __restrict__ bcol,
    const LT* __restrict__ sw, const IT* __restrict__ tinv, const T* __restrict__ r, T* __restrict__ y,
    int fuse_u) {
  using AT = TriAcc<T, LT>;
  constexpr int TPW = tile_tpw<G>(), NT = G * kTileMax;
  __shared__ AT BvA[TPW][kTileN];
  __shared__ __align__(16) IT TriA[TPW][kTileTri];
  const int sub = threadIdx.x / NT, tid = threadIdx.x % NT;
  const int g = tid / G, lane = tid % G;
  const int tix = blockIdx.x * TPW + sub;
  if (tix >= tcount) return;
  AT* Bv = BvA[sub];
  IT* Tri = TriA[sub];
  const int tg = tfirst + tix;
  const int4 rec = trow[(int64_t)tg * kTileMax + g];
  const bool act = rec.x >= 0;
  const int32_t p = rec.x;
  const IT* inv = tinv + (int64_t)tg * kTileInv;
  const int i = 4 * g + lane;
  // L^{-1} of the tile, loaded before the sums over the blocks outside the tile
  TriStage<IT, NT> lst;
  lst.load(inv, tid);
  AT acc[4] = {0, 0, 0, 0};
  AT rv = 0;
  if (act) {
    if (lane < 4) rv = (AT)r[4 * (int64_t)p + lane];
    blk16_dot<G, K>(bcol + rec.y, sw + 16 * (int64_t)rec.w, y, 0, rec.z & 0xffff, lane, acc);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = group_sum<G>(acc[a]);
  lst.store(Tri, tid);
  if (lane < 4) Bv[4 * g + lane] = act ? rv - pick4(acc, lane) : (AT)0;
  tile_sync<G>();
  const int ii = 4 * g + (lane & 3);  // the row this lane helps with (ii == i for lane < 4)
  AT yi = tile_tri_row<G>(Tri + tri_lo(ii), (const AT*)Bv, lane, 0, ii);
  if (fuse_u) {
    // the tile's backward sweep: U^{-1} staged like L^{-1} (padding rows and columns of
    // the inverse are identity, so the padded right-hand sides, 0, add nothing)
    tile_sync<G>();  // every lane is done with L^{-1} and Bv
    for (int q = tid; q < TriStage<IT, NT>::NV; q += NT)
      reinterpret_cast<uint4*>(Tri)[q] = reinterpret_cast<const uint4*>(inv + kTileTri)[q];
    if (lane < 4) Bv[i] = yi;
    tile_sync<G>();
    yi = tile_tri_row<G>(Tri + (tri_up(ii) - kTileTri) - ii, (const AT*)Bv, lane, ii, kTileN - 1);
  }
  if (act && lane < 4) y[4 * (int64_t)p + lane] = (T)yi;
}

// Backward sweep of one colour, in place: z = U^{-1} y on its tiles.
template <typename T, typename LT, typename IT, int G, int K>
__global__ __launch_bounds__(G * kTileMax * tile_tpw<G>()) void k_bsr_tile_usolve(
    int tfirst, int tcount, const int4* __restrict__ trow, const int32_t* __restrict__ bcol,
    const LT* __restrict__ sw, const IT* __restrict__ tinv, T* __restrict__ y) {
  using AT = TriAcc<T, LT>;
  constexpr int TPW = tile_tpw<G>(), NT = G * kTileMax;
  __shared__ AT BvA[TPW][kTileN];
  __shared__ __align__(16) IT TriA[TPW][kTileTri];
  const int sub = threadIdx.x / NT, tid = threadIdx.x % NT;
  const int g = tid / G, lane = tid % G;
  const int tix = blockIdx.x * TPW + sub;
  if (tix >= tcount) return;
  AT* Bv = BvA[sub];
  IT* Tri = TriA[sub];
  const int tg = tfirst + tix;
  const int4 rec = trow[(int64_t)tg * kTileMax + g];
  const bool act = rec.x >= 0;
  const int32_t p = rec.x;
  const IT* inv = tinv + (int64_t)tg * kTileInv;
  const int i = 4 * g + lane;
  // U^{-1} of the tile, loaded before the sums over the blocks outside the tile
  TriStage<IT, NT> ust;
  ust.load(inv + kTileTri, tid);
  AT acc[4] = {0, 0, 0, 0};
  AT rv = 0;
  if (act) {
    if (lane < 4) rv = (AT)y[4 * (int64_t)p + lane];
    blk16_dot<G, K>(bcol + rec.y, sw + 16 * (int64_t)rec.w, y, 0, rec.z & 0xffff, lane, acc);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = group_sum<G>(acc[a]);
  ust.store(Tri, tid);
  if (lane < 4) Bv[4 * g + lane] = act ? rv - pick4(acc, lane) : (AT)0;
  tile_sync<G>();
  const int ii = 4 * g + (lane & 3);  // the row this lane helps with (ii == i for lane < 4)
  // row[j] = U^{-1}(ii, j) for j >= ii
  AT zi = tile_tri_row<G>(Tri + (tri_up(ii) - kTileTri) - ii, (const AT*)Bv, lane, ii, kTileN - 1);
  if (act && lane < 4) y[4 * (int64_t)p + lane] = (T)zi;
}

// The tile sweeps' block-by-block copy of a factor's blocks outside the tiles (swoff):
// one 64-lane workgroup per block row; rows a >= m of a block are written as zero.
template <typename LT>
__global__ __launch_bounds__(64) void k_tile_sweep_copy(int64_t nblk, const BDesc* __restrict__ desc,
                                                        const int4* __restrict__ swoff, const LT* __restrict__ lu,
                                                        LT* __restrict__ sw) {
  const int64_t p = blockIdx.x;
  if (p >= nblk) return;
  const BDesc d = desc[p];
  const int4 o = swoff[p];
  const int m = d.dm >> 16;
  const int total = (o.z + d.nb - o.w) * 16;
  for (int q = threadIdx.x; q < total; q += 64) {
    const int j = q >> 4, a = (q >> 2) & 3, b = q & 3;
    const bool lower = j < o.z;
    const int t = lower ? j : o.w + (j - o.z);
    const int64_t dst = ((int64_t)(lower ? o.x + j : o.y + (j - o.z))) * 16 + (q & 15);
    sw[dst] = a < m ? lu[(int64_t)d.vbase + (int64_t)a * 4 * d.nb + 4 * t + b] : LT{};
  }
}

// Tile sweep shapes: code c -> G = kTileGs[c] lanes per block row, K = kTileKs[c] blocks
// per lane loaded together.
static const int kTileGs[6] = {8, 16, 8, 16, 4, 4};
static const int kTileKs[6] = {1, 1, 2, 2, 2, 3};
constexpr int kTileCodes = 6;
constexpr int kTileDefault = 1;

template <typename T, typename LT, bool UPPER, int G, int K>
static void tile_launch_g(hipStream_t st, const IluPlan& P, int t0, int cnt, const LT* sw, const void* tinv,
                          const T* r, T* y, int fuse_u) {
  // the tile inverses are stored like the factor: FP64, FP32 or BF16
  using IT = LT;
  const IT* ti = (const IT*)tinv;
  constexpr int tpw = tile_tpw<G>();
  const unsigned gr = (unsigned)((cnt + tpw - 1) / tpw);
  if constexpr (UPPER)
    hipLaunchKernelGGL((k_bsr_tile_usolve<T, LT, IT, G, K>), dim3(gr), dim3(G * kTileMax * tpw), 0, st,
                       t0, cnt, P.d_trowU, P.d_bcol, sw, ti, y);
  else
    hipLaunchKernelGGL((k_bsr_tile_lsolve<T, LT, IT, G, K>), dim3(gr), dim3(G * kTileMax * tpw), 0, st,
                       t0, cnt, P.d_trowL, P.d_bcol, sw, ti, r, y, fuse_u);
}

template <typename T, typename LT, bool UPPER>
static void tile_launch(hipStream_t st, int code, const IluPlan& P, int t0, int cnt, const LT* sw, const void* tinv,
                        const T* r, T* y, int fuse_u = 0) {
  switch (code) {
    case 0: tile_launch_g<T, LT, UPPER, 8, 1>(st, P, t0, cnt, sw, tinv, r, y, fuse_u); break;
    case 2: tile_launch_g<T, LT, UPPER, 8, 2>(st, P, t0, cnt, sw, tinv, r, y, fuse_u); break;
    case 3: tile_launch_g<T, LT, UPPER, 16, 2>(st, P, t0, cnt, sw, tinv, r, y, fuse_u); break;
    case 4: tile_launch_g<T, LT, UPPER, 4, 2>(st, P, t0, cnt, sw, tinv, r, y, fuse_u); break;
    case 5: tile_launch_g<T, LT, UPPER, 4, 3>(st, P, t0, cnt, sw, tinv, r, y, fuse_u); break;
    default: tile_launch_g<T, LT, UPPER, 16, 1>(st, P, t0, cnt, sw, tinv, r, y, fuse_u); break;
  }
}

// Block values from the CSR values: block (p, t) is the neighbour's free columns in the
// free rows of p, found in the DOF-ordered CSR at rowptr[r0orig[p] + a] + (bcolofs &
// 0xffffff); bcolofs >> 24 is the neighbour's free-column count, the columns the
// padded 4-wide block row takes from the CSR (the rest are 0).  64 lanes per block row,
// four block rows per workgroup; every lane stores one 4-wide block row with one vector
// store.
template <typename VT>
__global__ __launch_bounds__(256) void k_bsr_gather(int64_t nblk, const BDesc* __restrict__ desc,
                                                    const int32_t* __restrict__ bcolofs,
                                                    const int32_t* __restrict__ r0orig,
                                                    const int32_t* __restrict__ rowptr, const VT* __restrict__ src,
                                                    VT* __restrict__ dst) {
  const int64_t p = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
  if (p >= nblk) return;
  const int lane = threadIdx.x % 64;
  const BDesc d = desc[p];
  const int m = d.dm >> 16;
  const int32_t r0 = r0orig[p];
  for (int idx = lane; idx < m * d.nb; idx += 64) {
    const int a = idx / d.nb, t = idx - a * d.nb;
    const int32_t oc = bcolofs[d.bptr + t];
    const int mj = oc >> 24;
    const int64_t s0 = (int64_t)rowptr[r0 + a] + (oc & 0xffffff);
    VT v[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) v[b] = b < mj ? src[s0 + b] : (VT)0;
    VT* o = dst + (int64_t)d.vbase + (int64_t)a * 4 * d.nb + 4 * t;
    if constexpr (sizeof(VT) == 4) {
      *reinterpret_cast<float4*>(o) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
      reinterpret_cast<double2*>(o)[0] = make_double2(v[0], v[1]);
      reinterpret_cast<double2*>(o)[1] = make_double2(v[2], v[3]);
    }
  }
}

// Block values from the assembled node-pair block form: block (p, t) is block kbsrc of
// kb (16 values, free rows and columns packed to the front, the rest 0).  64 lanes per
// block row, four block rows per workgroup; one 4-wide load and store per lane.
template <typename VT>
__global__ __launch_bounds__(256) void k_bsr_gather_kb(int64_t nblk, const BDesc* __restrict__ desc,
                                                       const int32_t* __restrict__ kbsrc,
                                                       const VT* __restrict__ kb, VT* __restrict__ dst) {
  const int64_t p = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
  if (p >= nblk) return;
  const int lane = threadIdx.x % 64;
  const BDesc d = desc[p];
  const int m = d.dm >> 16;
  for (int idx = lane; idx < m * d.nb; idx += 64) {
    const int a = idx / d.nb, t = idx - a * d.nb;
    const V4<VT> v = ld4<VT>(kb + 16 * (int64_t)kbsrc[d.bptr + t] + 4 * a);
    st4(dst + (int64_t)d.vbase + (int64_t)a * 4 * d.nb + 4 * t, v.v[0], v.v[1], v.v[2], v.v[3]);
  }
}

// The same gather, also forming the multiplicative coarse correction's A Z of each block
// row from the values it moves: the pressure column of every neighbour block, summed per
// coarse unknown of the neighbour's pressure DOF, neighbours in order (as
// k_coarse_az_build sums them).  Row a's values go through LDS, one wavefront per row.
template <typename VT>
__global__ __launch_bounds__(256) void k_bsr_gather_kb_az(int64_t nblk, const BDesc* __restrict__ desc,
                                                          const int32_t* __restrict__ kbsrc,
                                                          const VT* __restrict__ kb, VT* __restrict__ dst,
                                                          const int32_t* __restrict__ bcol,
                                                          const int32_t* __restrict__ cinfo,
                                                          const int32_t* __restrict__ azoff,
                                                          const int32_t* __restrict__ azbox, VT* __restrict__ az) {
  __shared__ double sv[4][4][kBsrMaxNb];
  __shared__ int sb[4][kBsrMaxNb];
  const int w = threadIdx.x / 64;
  const int64_t p = (int64_t)blockIdx.x * 4 + w;
  if (p >= nblk) return;
  const int lane = threadIdx.x % 64;
  const BDesc d = desc[p];
  const int m = d.dm >> 16;
  for (int idx = lane; idx < m * d.nb; idx += 64) {
    const int a = idx / d.nb, t = idx - a * d.nb;
    const V4<VT> v = ld4<VT>(kb + 16 * (int64_t)kbsrc[d.bptr + t] + 4 * a);
    st4(dst + (int64_t)d.vbase + (int64_t)a * 4 * d.nb + 4 * t, v.v[0], v.v[1], v.v[2], v.v[3]);
    const int info = cinfo[bcol[d.bptr + t]];
    const int b = info & 3;
    const VT pv = b == 0 ? v.v[0] : b == 1 ? v.v[1] : b == 2 ? v.v[2] : v.v[3];
    sv[w][a][t] = info >= 0 ? (double)pv : 0.0;
    if (a == 0) sb[w][t] = info >= 0 ? info >> 3 : -1;
  }
  wave_sync();
  const int off = azoff[p], cnt = azoff[p + 1] - off;
  for (int q = lane; q < 4 * cnt; q += 64) {
    const int k = q >> 2, a = q & 3;
    const int c = azbox[off + k];
    double s = 0.0;
    if (a < m)
      for (int t = 0; t < d.nb; ++t)
        if (sb[w][t] == c) s += sv[w][a][t];
    az[4 * (int64_t)(off + k) + a] = (VT)s;
  }
}

// CSR values from the node-pair block form: node block k's rows r0 + a, its pair q's
// free columns at (natofs[q] & 0xffffff) in each row (natofs[q] >> 24 of them).  One
// wavefront per node block, four per workgroup.
template <typename VT>
__global__ __launch_bounds__(256) void k_kb_to_csr(int64_t nblk, const BlkDesc* __restrict__ desc,
                                                   const int64_t* __restrict__ blkPtr,
                                                   const int32_t* __restrict__ natofs, const VT* __restrict__ kb,
                                                   VT* __restrict__ vals) {
  const int64_t k = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
  if (k >= nblk) return;
  const int lane = threadIdx.x % 64;
  const BlkDesc d = desc[k];
  const int m = d.dqm >> 16;
  const int64_t q0 = blkPtr[k];
  const int nbk = (int)(blkPtr[k + 1] - q0);
  for (int idx = lane; idx < m * nbk; idx += 64) {
    const int a = idx / nbk, t = idx - a * nbk;
    const int32_t o = natofs[q0 + t];
    const int mj = o >> 24;
    const V4<VT> v = ld4<VT>(kb + 16 * (q0 + t) + 4 * a);
    VT* row = vals + (int64_t)d.base + (int64_t)a * d.rl + (o & 0xffffff);
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (b < mj) row[b] = v.v[b];
  }
}

// LDS of one row being factorised: its values in FP64 (rows[a][4 cap]), a pivot row's
// upper blocks (urow, storage type), the row's neighbour blocks and the positions of the
// pivot's upper blocks in it, and small per-pivot scratch.
template <typename LT>
struct FacRow {
  double* rows;
  LT* urow;
  int32_t* cols;
  int32_t* pos;
  double* ub;    // [16] pivot's diagonal block
  double* lblk;  // [16] L_pc, row-major
  double* lval;  // [4]
};

template <typename LT>
__device__ inline FacRow<LT> fac_row_at(double* base, int cap) {
  FacRow<LT> f;
  f.rows = base;
  f.urow = (LT*)(f.rows + 16 * cap);
  f.cols = (int32_t*)(f.urow + 16 * cap);
  f.pos = f.cols + cap;
  f.ub = (double*)(((uintptr_t)(f.pos + cap) + 7) & ~(uintptr_t)7);
  f.lblk = f.ub + 16;
  f.lval = f.lblk + 16;
  return f;
}

// bytes of one FacRow for rows of at most cap neighbour blocks (8-byte aligned)
template <typename LT>
__host__ __device__ constexpr size_t fac_row_bytes(int cap) {
  return ((size_t)cap * (16 * sizeof(double) + 16 * sizeof(LT) + 2 * sizeof(int32_t)) + 7) / 8 * 8 +
         36 * sizeof(double);
}

// Load block row d of a0 (the block tangent) and its neighbour list into f (one wavefront).
template <typename LT>
__device__ inline void fac_load_row(const FacRow<LT>& f, const BDesc& d, const int32_t* __restrict__ bcol,
                                    const LT* __restrict__ a0, int cap, int lane) {
  const int m = d.dm >> 16, nb = d.nb, rs = 4 * nb;
  constexpr int KC = 4 * kBsrMaxNb / 64;
  // every load issued before the first LDS store
  const int32_t c = lane < nb ? bcol[d.bptr + lane] : 0;
  double v[4][KC];
#pragma unroll
  for (int k = 0; k < KC; ++k)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int q = lane + 64 * k;
      v[a][k] = (a < m && q < rs) ? (double)a0[d.vbase + (int64_t)a * rs + q] : 0.0;
    }
  if (lane < nb) f.cols[lane] = c;
#pragma unroll
  for (int k = 0; k < KC; ++k)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int q = lane + 64 * k;
      if (a < m && q < rs) f.rows[a * 4 * cap + q] = v[a][k];
    }
}

// IKJ steps of a row (one wavefront) for its lower blocks t in [t0, t1), whose rows are
// final in lu: for lower block c, c's upper blocks and diagonal block are staged in LDS
// (their global loads issued before the previous block's work and held in registers
// meanwhile), the positions of c's upper blocks in this row are searched once, L_pc =
// R_pc U_cc^{-1} is one triangular solve per row (lanes < m), and every upper block of
// c present in this row is updated, R_pj -= L_pc U_cj, one (block, row, column) entry
// per lane.
template <typename LT>
__device__ inline void fac_pivots_global(const FacRow<LT>& f, const BDesc& d, int t0, int t1,
                                         const BDesc* __restrict__ desc, const int32_t* __restrict__ bcol,
                                         const LT* __restrict__ lu, double* __restrict__ scal, int cap, int lane) {
  const int m = d.dm >> 16, nb = d.nb;
  int32_t pc = -1;
  LT pu[4][4];
  LT pd = (LT)0;
  auto issue = [&](int t) {
    const BDesc dc = desc[f.cols[t]];
    const int mc = dc.dm >> 16, tdc = dc.dm & 0xffff, rsc = 4 * dc.nb;
    const int nup = dc.nb - tdc - 1;
    if (lane < nup) {
      pc = bcol[dc.bptr + tdc + 1 + lane];
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int bp = 0; bp < 4; ++bp)
          pu[b][bp] = b < mc ? lu[dc.vbase + (int64_t)b * rsc + 4 * (tdc + 1 + lane) + bp] : (LT)0;
    }
    if (lane < 16) pd = (lane >> 2) < mc ? lu[dc.vbase + (int64_t)(lane >> 2) * rsc + 4 * tdc + (lane & 3)] : (LT)0;
  };
  if (t0 < t1) issue(t0);
  for (int t = t0; t < t1; ++t) {
    const BDesc dc = desc[f.cols[t]];
    const int mc = dc.dm >> 16, tdc = dc.dm & 0xffff;
    const int nup = dc.nb - tdc - 1;
    if (lane < nup) {
      f.pos[lane] = lds_find(f.cols, nb, pc);
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int bp = 0; bp < 4; ++bp) f.urow[b * 4 * cap + 4 * lane + bp] = pu[b][bp];
    }
    if (lane < 16) f.ub[lane] = (double)pd;
    wave_sync();
    if (t + 1 < t1) issue(t + 1);
    // L_pc: row a (lane a < m) by forward substitution against U_cc
    if (lane < m) {
      double* r = f.rows + lane * 4 * cap + 4 * t;
      double l[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if (b < mc) {
          double x = r[b];
#pragma unroll
          for (int bq = 0; bq < b; ++bq) x -= l[bq] * f.ub[bq * 4 + b];
          const double dd = f.ub[b * 4 + b];
          if (dd == 0.0) scal[S_MISC1] = 2.0;
          l[b] = x / (dd == 0.0 ? 1.0 : dd);
          r[b] = l[b];
        }
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) f.lblk[lane * 4 + b] = l[b];
    }
    wave_sync();
    // R_pj -= L_pc U_cj over the upper blocks of c that this row holds
    for (int it = lane; it < 16 * nup; it += 64) {
      const int k = it >> 4, a = (it >> 2) & 3, bp = it & 3;
      const int ps = f.pos[k];
      if (ps < 0 || a >= m) continue;
      double x = f.rows[a * 4 * cap + 4 * ps + bp];
      for (int b = 0; b < mc; ++b) x -= f.lblk[a * 4 + b] * (double)f.urow[b * 4 * cap + 4 * k + bp];
      f.rows[a * 4 * cap + 4 * ps + bp] = x;
    }
    wave_sync();
  }
}

// The row's own diagonal block (point IKJ inside it), after all its lower blocks.
template <typename LT>
__device__ inline void fac_diag(const FacRow<LT>& f, const BDesc& d, double* __restrict__ scal, int cap, int lane) {
  const int m = d.dm >> 16, td = d.dm & 0xffff, rs = 4 * d.nb;
  const int q0 = 4 * td;
  for (int a = 1; a < m; ++a) {
    for (int ap = 0; ap < a; ++ap) {
      if (lane == 0) {
        const double dd = f.rows[ap * 4 * cap + q0 + ap];
        if (dd == 0.0) scal[S_MISC1] = 2.0;
        const double l = f.rows[a * 4 * cap + q0 + ap] / (dd == 0.0 ? 1.0 : dd);
        f.rows[a * 4 * cap + q0 + ap] = l;
        f.lval[0] = l;
      }
      wave_sync();
      const double l = f.lval[0];
      for (int q = q0 + ap + 1 + lane; q < rs; q += 64) f.rows[a * 4 * cap + q] -= l * f.rows[ap * 4 * cap + q];
      wave_sync();
    }
  }
}

// Store the factored row in the factor's storage type.
template <typename LT>
__device__ inline void fac_store_row(const FacRow<LT>& f, const BDesc& d, LT* __restrict__ lu, int cap, int lane) {
  const int m = d.dm >> 16, rs = 4 * d.nb;
  for (int a = 0; a < m; ++a)
    for (int q = lane; q < rs; q += 64) lu[d.vbase + (int64_t)a * rs + q] = (LT)f.rows[a * 4 * cap + q];
}

// ILU(0) factorisation of the block rows of one level, one 64-lane workgroup (one
// wavefront) per block row, IKJ order by lower neighbour blocks, a 4 x 4 block at a time
// (fac_pivots_global), then the row's own diagonal block (fac_diag).  Point ILU(0) on a
// pattern of full node blocks is exactly this: the subtractions run in the point
// algorithm's pivot order, so the factor is the same to the last bit.  A row starts from
// its values in a0 (the block tangent) and its factor is written to lu, so no separate
// copy of the tangent seeds the factor.  Dynamic LDS: fac_row_bytes(cap) for rows of at
// most cap <= kBsrMaxNb neighbour blocks.
template <typename LT>
__global__ __launch_bounds__(64) void k_bsr_factor(const int32_t* __restrict__ list, int first, int count,
                                                   const BDesc* __restrict__ desc,
                                                   const int32_t* __restrict__ bcol, LT* __restrict__ lu,
                                                   const LT* __restrict__ a0, double* __restrict__ scal, int cap) {
  extern __shared__ double smem[];
  const int w = blockIdx.x;
  if (w >= count) return;
  const int lane = threadIdx.x;
  const int32_t p = list ? list[first + w] : first + w;
  const BDesc d = desc[p];
  const FacRow<LT> f = fac_row_at<LT>(smem, cap);
  fac_load_row<LT>(f, d, bcol, a0, cap, lane);
  wave_sync();
  fac_pivots_global<LT>(f, d, 0, d.dm & 0xffff, desc, bcol, lu, scal, cap, lane);
  fac_diag<LT>(f, d, scal, cap, lane);
  fac_store_row<LT>(f, d, lu, cap, lane);
}

// Pivot-pair schedule of the factorisation (symbolic, once per plan).  For block row p
// a pair is (lower block t of p, upper block q of pivot row c = bcol[t]) whose column is
// also a block of row p, at position ps: the pairs are exactly the block updates
// R_p,ps -= L_pt U_cq of the IKJ factorisation.  Record: t | q << 8 | ps << 16, in the
// order (t, q) ascending.  One wavefront per row; WRITE = false counts the pairs of
// each row into cnt, WRITE = true writes them from pofs[p] on.
constexpr int kFacPairsWaves = 4;
template <bool WRITE>
__global__ __launch_bounds__(64 * kFacPairsWaves) void k_fac_pairs(int64_t nblk, const BDesc* __restrict__ desc,
                                                                   const int32_t* __restrict__ bcol,
                                                                   const int64_t* __restrict__ pofs,
                                                                   int32_t* __restrict__ cnt,
                                                                   uint32_t* __restrict__ prec) {
  __shared__ int32_t s_cols[kFacPairsWaves][kBsrMaxNb];
  __shared__ int32_t s_cofs[kFacPairsWaves][kBsrMaxNb + 1];
  __shared__ int32_t s_cb[kFacPairsWaves][kBsrMaxNb];
  __shared__ int32_t s_cq[kFacPairsWaves][kBsrMaxNb];
  const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t p = (int64_t)blockIdx.x * kFacPairsWaves + wv;
  if (p >= nblk) return;
  int32_t* cols = s_cols[wv];
  int32_t* cofs = s_cofs[wv];
  const BDesc d = desc[p];
  const int nb = d.nb, td = d.dm & 0xffff;
  int32_t c = -1;
  if (lane < nb) {
    c = bcol[d.bptr + lane];
    cols[lane] = c;
  }
  int nup = 0;
  if (lane < td) {
    const BDesc dc = desc[c];
    const int tdc = dc.dm & 0xffff;
    nup = dc.nb - tdc - 1;
    s_cb[wv][lane] = dc.bptr;
    s_cq[wv][lane] = tdc + 1;
  }
  // inclusive scan of the upper-block counts of the pivots
  int inc = nup;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(inc, o, 64);
    if (lane >= o) inc += v;
  }
  if (lane == 0) cofs[0] = 0;
  if (lane < td) cofs[lane + 1] = inc;
  wave_sync();
  const int ncand = td > 0 ? cofs[td] : 0;
  int base = 0;
  for (int i0 = 0; i0 < ncand; i0 += 64) {
    const int i = i0 + lane;
    int t = 0, q = 0, ps = -1;
    if (i < ncand) {
      int lo = 0, hi = td;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cofs[mid] <= i) lo = mid; else hi = mid;
      }
      t = lo;
      q = s_cq[wv][t] + (i - cofs[t]);
      ps = lds_find(cols, nb, bcol[s_cb[wv][t] + q]);
    }
    const uint64_t mask = __ballot(ps >= 0);
    if (WRITE && ps >= 0) {
      const int r = __popcll(mask & ((1ull << lane) - 1));
      prec[pofs[p] + base + r] = (uint32_t)t | ((uint32_t)q << 8) | ((uint32_t)ps << 16);
    }
    base += __popcll(mask);
  }
  if (!WRITE && lane == 0) cnt[p] = base;
}

// v of lane (lane & ~3) | B, in every lane of each group of four (DPP quad permute)
template <int B>
__device__ inline double quad_bcast(double v) {
  constexpr int ctrl = B | (B << 2) | (B << 4) | (B << 6);
  const long long i = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)i, ctrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(i >> 32), ctrl, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

constexpr int kFacRowsPerWave = 4;  // block rows per wavefront of the scheduled factorisation

// LDS of one row of the scheduled factorisation: the row in FP64 ([4][4 nbc]), per pivot
// {first neighbour entry of the pivot row, its diagonal position | free rows << 16} and
// {first, end} of its pairs, and the row's pair records.
__host__ __device__ constexpr size_t fac_group_bytes(int nbc, int maxpairs) {
  return ((size_t)nbc * (16 * sizeof(double) + 4 * sizeof(int32_t)) + (size_t)maxpairs * sizeof(uint32_t) + 15) /
         16 * 16;
}

// ILU(0) factorisation of the block rows of one level from the pivot-pair schedule: 16
// lanes per row, lane 4 a + b' holding entry (a, b') of the 4 x 4 block being worked on,
// four rows per wavefront.  Pivot t: L_pt = R_pt U_cc^{-1} by substitution across the
// four lanes of row a (each multiplier handed on by a DPP quad broadcast), then
// R_p,ps -= L_pt U_cq for the pairs of t, U_cc and U_cq read from the block copy fu
// (one load of column b' per lane); then the diagonal block by point IKJ, and the row
// stored to lu and its diagonal and upper blocks to fu.  Every entry takes the
// subtractions of the search-based kernel (k_bsr_factor) and the point IKJ algorithm in
// the same order, so the factor is the same to the last bit.  Dynamic LDS:
// kFacRowsPerWave * fac_group_bytes(nbc, maxpairs), nbc >= every row's neighbour count,
// maxpairs >= every row's pair count.
template <typename LT>
__global__ __launch_bounds__(64) void k_bsr_factor_grp(const int32_t* __restrict__ list, int first, int count,
                                                       const BDesc* __restrict__ desc,
                                                       const int32_t* __restrict__ bcol,
                                                       const int64_t* __restrict__ pofs,
                                                       const uint32_t* __restrict__ prec, LT* __restrict__ lu,
                                                       LT* __restrict__ fu, const LT* __restrict__ a0,
                                                       double* __restrict__ scal, int nbc, int maxpairs) {
  extern __shared__ double smem[];
  const int lane = threadIdx.x, g = lane >> 4, j = lane & 15;
  const int a = j >> 2, bp = j & 3;
  const int w = blockIdx.x * kFacRowsPerWave + g;
  if (w >= count) return;  // no lane exchanges data outside its own row
  const int32_t p = list ? list[first + w] : first + w;
  const BDesc d = desc[p];
  const int m = d.dm >> 16, td = d.dm & 0xffff, nb = d.nb, rs = 4 * nb, RS = 4 * nbc;
  double* rows = (double*)((char*)smem + g * fac_group_bytes(nbc, maxpairs));
  int2* pv = (int2*)(rows + 4 * RS);
  int2* pse = pv + nbc;
  uint32_t* recs = (uint32_t*)(pse + nbc);
  const int64_t r0 = pofs[p];
  const int np = (int)(pofs[p + 1] - r0);
  for (int i = j; i < np; i += 16) recs[i] = prec[r0 + i];
  for (int t = j; t < td; t += 16) {
    const BDesc dc = desc[bcol[d.bptr + t]];
    pv[t] = make_int2(dc.bptr, dc.dm);
    pse[t] = make_int2(0, 0);
  }
  for (int ar = 0; ar < m; ++ar)
    for (int q0 = 0; q0 < rs; q0 += 128) {
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int q = q0 + j + 16 * k;
        v[k] = q < rs ? (double)a0[d.vbase + (int64_t)ar * rs + q] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int q = q0 + j + 16 * k;
        if (q < rs) rows[ar * RS + q] = v[k];
      }
    }
  wave_sync();
  // pair range of each pivot (records sorted by pivot)
  for (int i = j; i < np; i += 16) {
    const int t = recs[i] & 0xff;
    if (i == 0 || (int)(recs[i - 1] & 0xff) != t) pse[t].x = i;
    if (i == np - 1 || (int)(recs[i + 1] & 0xff) != t) pse[t].y = i + 1;
  }
  wave_sync();
  const bool act = a < m;
  for (int t = 0; t < td; ++t) {
    const int2 pvt = pv[t];
    const int mc = pvt.y >> 16;
    const int2 pr = pse[t];
    const LT* ucc = fu + 16 * ((int64_t)pvt.x + (pvt.y & 0xffff));
    // column bp of U_cc and its diagonal
    const V4<LT> uc = ld4<LT>(ucc + 4 * bp);
    double dg[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) dg[b] = (double)ucc[5 * b];
    double x = rows[a * RS + 4 * t + bp];
    double lv[4] = {0.0, 0.0, 0.0, 0.0};
    double mine = 0.0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (b < mc) {
        const double dd = dg[b];
        if (dd == 0.0) scal[S_MISC1] = 2.0;
        const double l = x / (dd == 0.0 ? 1.0 : dd);
        if (bp == b) mine = l;
        switch (b) {
          case 0: lv[0] = quad_bcast<0>(l); break;
          case 1: lv[1] = quad_bcast<1>(l); break;
          case 2: lv[2] = quad_bcast<2>(l); break;
          default: lv[3] = quad_bcast<3>(l);
        }
        if (bp > b) x -= lv[b] * (double)uc.v[b];
      }
    }
    if (act && bp < mc) rows[a * RS + 4 * t + bp] = mine;
    // R_p,ps -= L_pt U_cq, eight pairs at a time with their loads in flight together
    for (int k0 = pr.x; k0 < pr.y; k0 += 8) {
      V4<LT> u[8];
      int pos[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        pos[k] = -1;
        u[k] = V4<LT>{{(LT)0, (LT)0, (LT)0, (LT)0}};
        if (k0 + k < pr.y) {
          const uint32_t rc = recs[k0 + k];
          pos[k] = (rc >> 16) & 0xff;
          u[k] = ld4<LT>(fu + 16 * ((int64_t)pvt.x + ((rc >> 8) & 0xff)) + 4 * bp);
        }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (act && pos[k] >= 0) {
          double* e = rows + a * RS + 4 * pos[k] + bp;
          double y = *e;
#pragma unroll
          for (int b = 0; b < 4; ++b)
            if (b < mc) y -= lv[b] * (double)u[k].v[b];
          *e = y;
        }
      }
    }
    wave_sync();
  }
  // the diagonal block, point IKJ
  const int q0 = 4 * td;
  for (int ar = 1; ar < m; ++ar) {
    for (int ap = 0; ap < ar; ++ap) {
      const int col = q0 + ap;
      const double dd = rows[ap * RS + col];
      const double num = rows[ar * RS + col];
      if (dd == 0.0) scal[S_MISC1] = 2.0;
      const double l = num / (dd == 0.0 ? 1.0 : dd);
      wave_sync();
      if (j == 0) rows[ar * RS + col] = l;
      for (int q = col + 1 + j; q < rs; q += 16) rows[ar * RS + q] -= l * rows[ap * RS + q];
      wave_sync();
    }
  }
  for (int ar = 0; ar < m; ++ar)
    for (int q = j; q < rs; q += 16) lu[d.vbase + (int64_t)ar * rs + q] = (LT)rows[ar * RS + q];
  // diagonal and upper blocks into the block copy: lane j = 4 b' + b holds U[b][b']
  const int cb = j >> 2, rb = j & 3;
  for (int t = td; t < nb; ++t)
    fu[16 * ((int64_t)d.bptr + t) + j] = rb < m ? (LT)rows[rb * RS + 4 * t + cb] : (LT)0;
}

// b[i] = scale * a[padmap[i]] into the padded block layout (0 where padmap < 0)
template <typename TA, typename TB>
__global__ __launch_bounds__(kTPB) void k_pad_in(int64_t n, const int32_t* __restrict__ padmap,
                                                 const TA* __restrict__ a, TB* __restrict__ b, double scale) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB) {
    const int32_t k = padmap[i];
    b[i] = k >= 0 ? (TB)((double)a[k] * scale) : (TB)0;
  }
}

// b[padmap[i]] = scale * a[i] back to the DOF layout
template <typename TA, typename TB>
__global__ __launch_bounds__(kTPB) void k_pad_out(int64_t n, const int32_t* __restrict__ padmap,
                                                  const TA* __restrict__ a, TB* __restrict__ b, double scale) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB) {
    const int32_t k = padmap[i];
    if (k >= 0) b[k] = (TB)((double)a[i] * scale);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#define LA_CHECK(x)                                  \
  do {                                               \
    hipError_t _e = (x);                             \
    if (_e != hipSuccess) return set_hip_error(_e, #x); \
  } while (0)

static int grid_for(int64_t n) {
  int64_t g = (n + kTPB - 1) / kTPB;
  if (g > kRedBlocks) g = kRedBlocks;
  if (g < 1) g = 1;
  return (int)g;
}

template <typename T>
static hipError_t alloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  return hipMalloc((void**)p, n * sizeof(T));
}

LinAlg::~LinAlg() {
  la_direct_destroy(this);
  void* ptrs[] = {d_blkStart, d_diagpos, d_part, d_scal, r, rh, p, v, s, t, ph, sh, x, w, z, pr, pz,
                  r32, rh32, p32, v32, s32, t32, ph32, sh32, x32, du32, w32, z32, pr32, pz32,
                  dinv, dinv32, binv, binv32, lu, lu32, gm_V, gm_V32, d_desc, lu16, kp, kp32, d_h, d_mpart,
                  tinv, tinv32, tinv16, sw, sw32, sw16, fu, fu32};
  delete ilu;
  coarse_destroy(coarse);
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (h_scal) (void)hipHostFree(h_scal);
  if (h_h) (void)hipHostFree(h_h);
}

static std::vector<BlkDesc> make_desc(const std::vector<int32_t>& bs, const std::vector<int64_t>& rowptr,
                                      const std::vector<int32_t>& diagpos);

int la_create(sfem_ctx* ctx) {
  LinAlg* L = new LinAlg;
  ctx->la = L;
  const int64_t n = ctx->nfree;
  L->n = n;
  // nodal blocks: free rows of each node with at least one free dof
  std::vector<int32_t> bs;
  bs.reserve(ctx->nNode + 1);
  const int ndof = ctx->dims.ndof;
  int32_t acc = 0;
  bs.push_back(0);
  for (int64_t i = 0; i < ctx->nNode; ++i) {
    int c = 0;
    for (int d = 0; d < ndof; ++d) c += ctx->nodeDof[i * ndof + d] >= 0;
    if (c) { acc += c; bs.push_back(acc); }
  }
  L->nblk = (int64_t)bs.size() - 1;
  L->h_blkStart = bs;
  // diagonal positions
  std::vector<int32_t> dp(n);
  int maxrow = 0;
  for (int64_t r = 0; r < n; ++r) {
    const int64_t b = ctx->rowptr[r], e = ctx->rowptr[r + 1];
    maxrow = std::max(maxrow, (int)(e - b));
    const int32_t* f = std::lower_bound(&ctx->colidx[0] + b, &ctx->colidx[0] + e, (int32_t)r);
    dp[r] = (f != &ctx->colidx[0] + e && *f == r) ? (int32_t)(f - &ctx->colidx[0]) : -1;
  }
  L->maxrow = maxrow;
  L->all_diag = std::find(dp.begin(), dp.end(), -1) == dp.end();
  L->nk = n;
  L->nvec = std::max<int64_t>(n, 4 * L->nblk);  // Krylov vectors fit both layouts
  L->h_diagpos = dp;
  LA_CHECK(alloc(&L->d_blkStart, bs.size()));
  LA_CHECK(hipMemcpy(L->d_blkStart, bs.data(), bs.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  LA_CHECK(alloc(&L->d_diagpos, n));
  LA_CHECK(hipMemcpy(L->d_diagpos, dp.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
  {
    std::vector<int32_t> rowBlk(n);
    for (int64_t k = 0; k < L->nblk; ++k)
      for (int32_t r = bs[k]; r < bs[k + 1]; ++r) rowBlk[r] = (int32_t)k;
    // full-block check on the first row of every block (the rows of a block share it)
    bool full = true;
    for (int64_t k = 0; k < L->nblk && full; ++k) {
      const int64_t b = ctx->rowptr[bs[k]], e = ctx->rowptr[bs[k] + 1];
      for (int64_t q = b; q < e && full;) {
        const int32_t c = rowBlk[ctx->colidx[q]];
        const int32_t len = bs[c + 1] - bs[c];
        full = ctx->colidx[q] == bs[c] && q + len <= e;
        for (int32_t t = 0; t < len && full; ++t) full = ctx->colidx[q + t] == bs[c] + t;
        q += len;
      }
    }
    L->full_blocks = full;
  }
  {
    std::vector<BlkDesc> dd = make_desc(bs, ctx->rowptr, dp);
    LA_CHECK(alloc(&L->d_desc, dd.size()));
    if (!dd.empty()) LA_CHECK(hipMemcpy(L->d_desc, dd.data(), dd.size() * sizeof(BlkDesc), hipMemcpyHostToDevice));
  }
  LA_CHECK(alloc(&L->d_part, (size_t)kRedBlocks * 4));
  LA_CHECK(alloc(&L->d_scal, S_NSLOTS));
  LA_CHECK(hipHostMalloc((void**)&L->h_scal, S_NSLOTS * sizeof(double), hipHostMallocDefault));
  return SFEM_OK;
}

void la_destroy(sfem_ctx* ctx) {
  delete ctx->la;
  ctx->la = nullptr;
}

// --- ILU(0) plan: node ordering, (permuted) pattern, level sets (host) ---
static void bucket_levels(const std::vector<int32_t>& lv, std::vector<int32_t>& ptr, std::vector<int32_t>& list) {
  int nl = 0;
  for (int32_t x : lv) nl = std::max(nl, x + 1);
  ptr.assign(nl + 1, 0);
  for (int32_t x : lv) ptr[x + 1]++;
  for (int i = 0; i < nl; ++i) ptr[i + 1] += ptr[i];
  list.resize(lv.size());
  std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
  for (int64_t k = 0; k < (int64_t)lv.size(); ++k) list[pos[lv[k]]++] = (int32_t)k;
}

// Reverse Cuthill-McKee over the block (node) graph, starting each component from a
// pseudo-peripheral node (George & Liu, Computer Solution of Large Sparse Positive
// Definite Systems, 1981, sec. 4.3).  Returns the old block ids in the new order.
static std::vector<int32_t> rcm_order(const std::vector<int64_t>& adjPtr, const std::vector<int32_t>& adj) {
  const int64_t nb = (int64_t)adjPtr.size() - 1;
  std::vector<int32_t> deg(nb), order;
  order.reserve(nb);
  for (int64_t k = 0; k < nb; ++k) deg[k] = (int32_t)(adjPtr[k + 1] - adjPtr[k]);
  std::vector<char> seen(nb, 0);
  std::vector<int32_t> mark(nb, -1), lvl(nb, 0), queue;
  queue.reserve(nb);
  int stamp = 0;
  auto bfs_last_level = [&](int32_t s, int& ecc, std::vector<int32_t>& last) {
    ++stamp;
    queue.clear();
    queue.push_back(s);
    mark[s] = stamp;
    lvl[s] = 0;
    ecc = 0;
    for (size_t h = 0; h < queue.size(); ++h) {
      const int32_t u = queue[h];
      for (int64_t q = adjPtr[u]; q < adjPtr[u + 1]; ++q) {
        const int32_t w = adj[q];
        if (mark[w] == stamp || seen[w]) continue;
        mark[w] = stamp;
        lvl[w] = lvl[u] + 1;
        ecc = std::max(ecc, lvl[w]);
        queue.push_back(w);
      }
    }
    last.clear();
    for (int32_t u : queue)
      if (lvl[u] == ecc) last.push_back(u);
  };
  std::vector<int32_t> last, nbrs;
  for (int64_t s0 = 0; s0 < nb; ++s0) {
    if (seen[s0]) continue;
    // minimum-degree node of this component as the first candidate
    int ecc = 0;
    bfs_last_level((int32_t)s0, ecc, last);
    int32_t s = (int32_t)s0;
    for (int32_t u : queue)
      if (deg[u] < deg[s]) s = u;
    bfs_last_level(s, ecc, last);
    for (int it = 0; it < 8; ++it) {
      int32_t t = last[0];
      for (int32_t u : last)
        if (deg[u] < deg[t]) t = u;
      int ecc2 = 0;
      std::vector<int32_t> last2;
      bfs_last_level(t, ecc2, last2);
      if (ecc2 <= ecc) break;
      s = t;
      ecc = ecc2;
      last.swap(last2);
    }
    // Cuthill-McKee from s: neighbours in increasing degree
    const size_t first = order.size();
    order.push_back(s);
    seen[s] = 1;
    for (size_t h = first; h < order.size(); ++h) {
      const int32_t u = order[h];
      nbrs.clear();
      for (int64_t q = adjPtr[u]; q < adjPtr[u + 1]; ++q)
        if (!seen[adj[q]]) nbrs.push_back(adj[q]);
      std::sort(nbrs.begin(), nbrs.end(),
                [&](int32_t a, int32_t b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
      for (int32_t w : nbrs) {
        seen[w] = 1;
        order.push_back(w);
      }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// Greedy distance-1 colouring of the block graph in block order; returns the old block
// ids sorted by (colour, id).  Blocks of one colour are pairwise non-adjacent, so the
// ILU(0) factorisation and both triangular solves advance one colour per level.  On a
// structured hex grid this is the 2x2x2 parity colouring (8 colours).
// Greedy colouring visiting the blocks in increasing (reverse = false) or decreasing
// order; returns the number of colours.
static int greedy_color(const std::vector<int64_t>& adjPtr, const std::vector<int32_t>& adj, bool reverse,
                        std::vector<int32_t>& color) {
  const int64_t nb = (int64_t)adjPtr.size() - 1;
  color.assign(nb, -1);
  std::vector<int64_t> mark;
  int nc = 0;
  for (int64_t i = 0; i < nb; ++i) {
    const int64_t k = reverse ? nb - 1 - i : i;
    for (int64_t q = adjPtr[k]; q < adjPtr[k + 1]; ++q) {
      const int32_t c = color[adj[q]];
      if (c >= 0) {
        if ((int)mark.size() <= c) mark.resize(c + 1, -1);
        mark[c] = k;
      }
    }
    int c = 0;
    while (c < (int)mark.size() && mark[c] == k) ++c;
    color[k] = c;
    nc = std::max(nc, c + 1);
  }
  return nc;
}

// Multicolour block ordering: colour classes in order, blocks of a class in their
// original order.  Greedy colouring is tried in both visiting orders and the one with
// fewer colours kept: on a structured 27-point node graph with a removed corner node
// (a pinned pressure) the visit that starts at the removed corner needs 10 colours,
// the other the optimal 8.
static std::vector<int32_t> multicolor_order(const std::vector<int64_t>& adjPtr, const std::vector<int32_t>& adj,
                                             std::vector<int32_t>* colorPtr) {
  const int64_t nb = (int64_t)adjPtr.size() - 1;
  std::vector<int32_t> color, color_r;
  int nc = greedy_color(adjPtr, adj, false, color);
  const int nc_r = greedy_color(adjPtr, adj, true, color_r);
  if (nc_r < nc) {
    nc = nc_r;
    color.swap(color_r);
  }
  std::vector<int32_t> ptr(nc + 1, 0), order(nb);
  for (int64_t k = 0; k < nb; ++k) ptr[color[k] + 1]++;
  for (int c = 0; c < nc; ++c) ptr[c + 1] += ptr[c];
  *colorPtr = ptr;  // colour c occupies [ptr[c], ptr[c+1]) of the new order
  for (int64_t k = 0; k < nb; ++k) order[ptr[color[k]]++] = (int32_t)k;
  return order;
}

template <typename T>
static int upload(T** d, const std::vector<T>& h) {
  hipError_t e = alloc(d, h.size());
  if (e == hipSuccess && !h.empty()) e = hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
  return e == hipSuccess ? SFEM_OK : set_hip_error(e, "upload");
}

// Descriptors of the node blocks of a (possibly permuted) pattern.
static std::vector<BlkDesc> make_desc(const std::vector<int32_t>& bs, const std::vector<int64_t>& rowptr,
                                      const std::vector<int32_t>& diagpos) {
  const int64_t nb = (int64_t)bs.size() - 1;
  std::vector<BlkDesc> d(nb);
  for (int64_t k = 0; k < nb; ++k) {
    const int32_t r0 = bs[k];
    const int32_t base = (int32_t)rowptr[r0];
    const int32_t rl = (int32_t)(rowptr[r0 + 1] - rowptr[r0]);
    const int32_t dq0 = diagpos[r0] - base;
    d[k] = BlkDesc{r0, base, rl, dq0 | ((bs[k + 1] - r0) << 16)};
  }
  return d;
}

IluPlan::~IluPlan() {
  void* ptrs[] = {d_levL, d_levU, d_bdesc, d_bcol, d_bcolofs, d_r0orig, d_padmap, d_natrows, d_tileStart,
                  d_swoff, d_trowL, d_trowU, d_fpofs, d_fprec, d_kbsrc};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
}

// Tiled multicolour ordering.  Tiles are the node blocks of one element, taken
// greedily in element order while none of them is in a tile yet (on a structured hex
// grid with an even number of nodes per side these are the 2 x 2 x 2 node cubes of
// every other element); blocks left over form tiles of one.  The tile graph (tiles
// adjacent when any of their blocks are) is coloured greedily, visiting tiles in both
// orders and keeping the colouring with fewer colours.  New order: colour, then tile,
// then block.  Within a tile every block couples to every other, so the triangular
// solves are sequential inside a tile and parallel across the tiles of a colour; ILU(0)
// in this order needs far fewer Krylov iterations than the point multicolour order
// (nearly those of the natural order) with the same number of launches per sweep.
static std::vector<int32_t> tiled_order(sfem_ctx* ctx, const std::vector<int32_t>& rowBlk,
                                        const std::vector<int64_t>& adjPtr, const std::vector<int32_t>& adj,
                                        IluPlan& P) {
  const int64_t nb = (int64_t)adjPtr.size() - 1;
  const int npe = ctx->dims.npe, ndof = ctx->dims.ndof;
  const int64_t nNode = ctx->nNode, nElem = ctx->nElem;
  std::vector<int32_t> nodeBlk(nNode, -1);
  for (int64_t v = 0; v < nNode; ++v)
    for (int d = 0; d < ndof; ++d) {
      const int32_t r = ctx->nodeDof[v * ndof + d];
      if (r >= 0) { nodeBlk[v] = rowBlk[r]; break; }
    }
  std::vector<int32_t> tileOf(nb, -1), tptr(1, 0), tmem;
  tmem.reserve(nb);
  int32_t mem[8];
  for (int64_t e = 0; e < nElem; ++e) {
    int cnt = 0;
    bool ok = true;
    for (int a = 0; a < npe && ok; ++a) {
      const int32_t k = nodeBlk[ctx->conn[e * npe + a]];
      if (k < 0) continue;
      if (tileOf[k] >= 0) ok = false;
      else mem[cnt++] = k;
    }
    if (!ok || cnt == 0) continue;
    std::sort(mem, mem + cnt);
    const int32_t t = (int32_t)tptr.size() - 1;
    for (int i = 0; i < cnt; ++i) { tileOf[mem[i]] = t; tmem.push_back(mem[i]); }
    tptr.push_back((int32_t)tmem.size());
  }
  for (int64_t k = 0; k < nb; ++k)
    if (tileOf[k] < 0) {
      tileOf[k] = (int32_t)tptr.size() - 1;
      tmem.push_back((int32_t)k);
      tptr.push_back((int32_t)tmem.size());
    }
  const int64_t nt = (int64_t)tptr.size() - 1;
  // tile graph
  std::vector<int64_t> tadjPtr(nt + 1, 0);
  std::vector<int32_t> tadj, tmp;
  for (int64_t t = 0; t < nt; ++t) {
    tmp.clear();
    for (int32_t i = tptr[t]; i < tptr[t + 1]; ++i) {
      const int32_t k = tmem[i];
      for (int64_t q = adjPtr[k]; q < adjPtr[k + 1]; ++q)
        if (tileOf[adj[q]] != t) tmp.push_back(tileOf[adj[q]]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    tadj.insert(tadj.end(), tmp.begin(), tmp.end());
    tadjPtr[t + 1] = (int64_t)tadj.size();
  }
  std::vector<int32_t> color, color_r;
  int nc = greedy_color(tadjPtr, tadj, false, color);
  const int nc_r = greedy_color(tadjPtr, tadj, true, color_r);
  if (nc_r < nc) {
    nc = nc_r;
    color.swap(color_r);
  }
  std::vector<int32_t> cptr(nc + 1, 0), torder(nt);
  for (int64_t t = 0; t < nt; ++t) cptr[color[t] + 1]++;
  for (int c = 0; c < nc; ++c) cptr[c + 1] += cptr[c];
  P.tileColPtr = cptr;
  for (int64_t t = 0; t < nt; ++t) torder[cptr[color[t]]++] = (int32_t)t;
  std::vector<int32_t> order;
  order.reserve(nb);
  P.h_tileStart.assign(1, 0);
  for (int64_t i = 0; i < nt; ++i) {
    const int32_t t = torder[i];
    for (int32_t j = tptr[t]; j < tptr[t + 1]; ++j) order.push_back(tmem[j]);
    P.h_tileStart.push_back((int32_t)order.size());
  }
  P.tiled = true;
  return order;
}

// Block plan of ILU(0) for one node ordering: block rows in the new order, each with
// its neighbour blocks sorted by new position, the offsets of those neighbours' first
// columns in the DOF-ordered CSR row (for the value gather), block levels (one per
// colour with the multicolour ordering), and the padded-slot -> DOF map of vectors.
// Blocks of level lv of the forward (upper=false) or backward (upper=true) sweep: a list
// slice, or with a contiguous plan the block range of colour lv (reversed for backward).
static void level_range(const IluPlan& P, bool upper, int lv, const int32_t** list, int* first, int* count);


// The factorisation's pivot-pair schedule (k_fac_pairs) for the plan's block rows.
static int build_fac_pairs(sfem_ctx* ctx, IluPlan& P) {
  hipStream_t st = ctx->stream;
  const int64_t nb = P.nblk;
  if (nb == 0) return SFEM_OK;
  int32_t* d_cnt = nullptr;
  hipError_t e = alloc(&d_cnt, (size_t)nb);
  if (e != hipSuccess) return set_hip_error(e, "factorisation schedule");
  const unsigned g = (unsigned)((nb + kFacPairsWaves - 1) / kFacPairsWaves);
  hipLaunchKernelGGL((k_fac_pairs<false>), dim3(g), dim3(64 * kFacPairsWaves), 0, st, nb, P.d_bdesc, P.d_bcol,
                     nullptr, d_cnt, nullptr);
  std::vector<int32_t> cnt(nb);
  e = hipMemcpyAsync(cnt.data(), d_cnt, (size_t)nb * sizeof(int32_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d_cnt);
  if (e != hipSuccess) return set_hip_error(e, "factorisation schedule");
  std::vector<int64_t> ofs(nb + 1, 0);
  for (int64_t p = 0; p < nb; ++p) ofs[p + 1] = ofs[p] + cnt[p];
  P.fac_pairs = ofs[nb];
  int rc = upload(&P.d_fpofs, ofs);
  if (rc) return rc;
  e = alloc(&P.d_fprec, (size_t)std::max<int64_t>(P.fac_pairs, 1));
  if (e != hipSuccess) return set_hip_error(e, "factorisation schedule");
  hipLaunchKernelGGL((k_fac_pairs<true>), dim3(g), dim3(64 * kFacPairsWaves), 0, st, nb, P.d_bdesc, P.d_bcol,
                     P.d_fpofs, nullptr, P.d_fprec);
  e = hipGetLastError();
  if (e != hipSuccess) return set_hip_error(e, "factorisation schedule");
  int mx = 0;
  for (int64_t p = 0; p < nb; ++p) mx = std::max(mx, cnt[p]);
  P.fac_maxpairs = std::max(mx, 1);
  // SFEM_FACTOR_KERNEL=search: every level on the search-based kernel (a diagnostic
  // switch; both kernels give the same factor)
  const char* fk = getenv("SFEM_FACTOR_KERNEL");
  P.fac_sched = !(fk && strcmp(fk, "search") == 0) &&
                kFacRowsPerWave * fac_group_bytes(P.nbmax, P.fac_maxpairs) <= 65536;
  return SFEM_OK;
}

static int build_ilu_plan(sfem_ctx* ctx, LinAlg* L, int reorder) {
  if (!L->full_blocks)
    return set_error(SFEM_ERR_UNSUPPORTED, "ILU(0): the pattern is not a union of full node blocks");
  delete L->ilu;
  L->ilu = nullptr;
  // factor-sized buffers belong to the previous plan
  void** vals[] = {(void**)&L->kp, (void**)&L->kp32, (void**)&L->lu, (void**)&L->lu32, (void**)&L->lu16,
                   (void**)&L->tinv, (void**)&L->tinv32, (void**)&L->tinv16, (void**)&L->sw, (void**)&L->sw32,
                   (void**)&L->sw16, (void**)&L->fu, (void**)&L->fu32};
  for (void** q : vals)
    if (*q) { (void)hipFree(*q); *q = nullptr; }
  IluPlan* Pp = new IluPlan;
  L->ilu = Pp;
  ++L->plan_id;
  IluPlan& P = *Pp;
  P.reorder = reorder;
  const std::vector<int32_t>& bs = L->h_blkStart;
  const int64_t nb = L->nblk;
  std::vector<int32_t> rowBlk(L->n);
  for (int64_t k = 0; k < nb; ++k)
    for (int32_t r = bs[k]; r < bs[k + 1]; ++r) rowBlk[r] = (int32_t)k;
  // neighbour blocks of every block (self included) with their column offsets
  std::vector<int64_t> nbPtr(nb + 1, 0);
  for (int64_t k = 0; k < nb; ++k) {
    int32_t cnt = 0, prev = -1;
    for (int64_t q = ctx->rowptr[bs[k]]; q < ctx->rowptr[bs[k] + 1]; ++q) {
      const int32_t j = rowBlk[ctx->colidx[q]];
      if (j != prev) ++cnt;
      prev = j;
    }
    nbPtr[k + 1] = nbPtr[k] + cnt;
  }
  const int64_t npairs = nbPtr[nb];
  std::vector<int32_t> nbB(npairs), nbOfs(npairs);
  parallel_for(nb, [&](int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; ++k) {
      int64_t o = nbPtr[k];
      int32_t prev = -1;
      const int64_t r0 = ctx->rowptr[bs[k]];
      for (int64_t q = r0; q < ctx->rowptr[bs[k] + 1]; ++q) {
        const int32_t j = rowBlk[ctx->colidx[q]];
        if (j != prev) { nbB[o] = j; nbOfs[o] = (int32_t)(q - r0); ++o; }
        prev = j;
      }
    }
  });
  // node ordering on the graph without self loops
  std::vector<int32_t> order(nb), colorPtr;
  if (reorder == 0) {
    for (int64_t k = 0; k < nb; ++k) order[k] = (int32_t)k;
  } else {
    std::vector<int64_t> adjPtr(nb + 1, 0);
    std::vector<int32_t> adj;
    adj.reserve(npairs - nb);
    for (int64_t k = 0; k < nb; ++k) {
      for (int64_t q = nbPtr[k]; q < nbPtr[k + 1]; ++q)
        if (nbB[q] != k) adj.push_back(nbB[q]);
      adjPtr[k + 1] = (int64_t)adj.size();
    }
    if (reorder == 3) order = tiled_order(ctx, rowBlk, adjPtr, adj, P);
    else order = reorder == 2 ? multicolor_order(adjPtr, adj, &colorPtr) : rcm_order(adjPtr, adj);
  }
  std::vector<int32_t> iord(nb);
  for (int64_t p = 0; p < nb; ++p) iord[order[p]] = (int32_t)p;
  // block rows in the new order
  std::vector<BDesc> bd(nb);
  std::vector<int64_t> bptr(nb + 1, 0);
  for (int64_t p = 0; p < nb; ++p) {
    const int32_t k = order[p];
    bptr[p + 1] = bptr[p] + (nbPtr[k + 1] - nbPtr[k]);
  }
  std::vector<int32_t> bcol(npairs), bcolofs(npairs), tdiag(nb);
  int nbmax = 0;
  for (int64_t k = 0; k < nb; ++k) nbmax = std::max(nbmax, (int)(nbPtr[k + 1] - nbPtr[k]));
  if (nbmax > kBsrMaxNb) {
    delete L->ilu;
    L->ilu = nullptr;
    return set_error(SFEM_ERR_UNSUPPORTED, "ILU(0): a node couples to more than 64 node blocks");
  }
  // the assembled block form numbers a block row's pairs like nbPtr (node order)
  const bool kb = ctx->kb_on;
  if (kb && (ctx->blkPtr.size() != (size_t)nb + 1 || ctx->blkPtr[nb] != npairs)) {
    delete L->ilu;
    L->ilu = nullptr;
    return set_error(SFEM_ERR_STATE, "ILU(0): node-pair blocks do not match the pattern");
  }
  std::vector<int32_t> kbsrc(kb ? npairs : 0);
  parallel_for(nb, [&](int64_t lo, int64_t hi) {
    std::vector<std::pair<int32_t, int64_t>> tmp;
    for (int64_t p = lo; p < hi; ++p) {
      const int32_t k = order[p];
      tmp.clear();
      for (int64_t q = nbPtr[k]; q < nbPtr[k + 1]; ++q) tmp.push_back({iord[nbB[q]], q});
      std::sort(tmp.begin(), tmp.end());
      for (size_t i = 0; i < tmp.size(); ++i) {
        bcol[bptr[p] + i] = tmp[i].first;
        bcolofs[bptr[p] + i] = nbOfs[tmp[i].second];
        if (kb) kbsrc[bptr[p] + i] = (int32_t)tmp[i].second;
        if (tmp[i].first == p) tdiag[p] = (int32_t)i;
      }
    }
  });
  int64_t vb = 0;
  std::vector<int32_t> r0orig(nb), padmap(4 * nb, -1);
  for (int64_t p = 0; p < nb; ++p) {
    const int32_t k = order[p];
    const int32_t m = bs[k + 1] - bs[k];
    const int32_t cnt = (int32_t)(bptr[p + 1] - bptr[p]);
    bd[p] = BDesc{(int32_t)vb, (int32_t)bptr[p], cnt, tdiag[p] | (m << 16)};
    vb += (int64_t)m * 4 * cnt;
    if (vb >= ((int64_t)1 << 31)) {
      delete L->ilu;
      L->ilu = nullptr;
      return set_error(SFEM_ERR_UNSUPPORTED, "ILU(0): block values exceed 2^31");
    }
    r0orig[p] = bs[k];
    for (int32_t a = 0; a < m; ++a) padmap[4 * p + a] = bs[k] + a;
  }
  P.nblk = nb;
  P.npad = 4 * nb;
  P.nvals = vb;
  P.npairs = npairs;
  P.nbmax = nbmax;
  P.h_order = order;
  if (reorder == 2) {
    // one level per colour: blocks of a colour are independent and contiguous in the
    // new order; the backward solve walks the colours in reverse
    P.contiguous = true;
    P.levLPtr = colorPtr;
    P.levUPtr = colorPtr;
    // rows of the last colour have no upper neighbours (no two neighbours share a
    // colour): their backward solve is the diagonal block and joins the forward sweep
    const int nc = (int)colorPtr.size() - 1;
    bool fl = nc > 0;
    for (int64_t p = nc > 0 ? colorPtr[nc - 1] : 0; fl && p < nb; ++p) fl = tdiag[p] + 1 == (int32_t)(bptr[p + 1] - bptr[p]);
    P.fuse_last = fl;
  } else {
    // level of a block row = 1 + max level of the rows it depends on (lower
    // neighbours for the forward solve and the factorisation, upper ones backward)
    std::vector<int32_t> lvL(nb, 0), lvU(nb, 0);
    for (int64_t p = 0; p < nb; ++p) {
      int lv = 0;
      for (int64_t q = bptr[p]; q < bptr[p] + tdiag[p]; ++q) lv = std::max(lv, lvL[bcol[q]] + 1);
      lvL[p] = lv;
    }
    for (int64_t p = nb - 1; p >= 0; --p) {
      int lv = 0;
      for (int64_t q = bptr[p] + tdiag[p] + 1; q < bptr[p + 1]; ++q) lv = std::max(lv, lvU[bcol[q]] + 1);
      lvU[p] = lv;
    }
    bucket_levels(lvL, P.levLPtr, P.h_levL);
    bucket_levels(lvU, P.levUPtr, P.h_levU);
  }
  int rc;
  {
    hipError_t e = alloc(&P.d_bdesc, bd.size());
    if (e == hipSuccess) e = hipMemcpy(P.d_bdesc, bd.data(), bd.size() * sizeof(BDesc), hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_hip_error(e, "ILU block descriptors");
  }
  // the gather's column offset of each neighbour block with the neighbour's free-column
  // count in the top byte (row offsets stay below 4 * kBsrMaxNb)
  for (int64_t q = 0; q < npairs; ++q) bcolofs[q] |= (bd[bcol[q]].dm >> 16) << 24;
  if ((rc = upload(&P.d_bcol, bcol)) || (rc = upload(&P.d_bcolofs, bcolofs)) || (rc = upload(&P.d_r0orig, r0orig)) ||
      (rc = upload(&P.d_padmap, padmap)) || (rc = upload(&P.d_levL, P.h_levL)) || (rc = upload(&P.d_levU, P.h_levU)))
    return rc;
  if (reorder != 0 && (rc = upload(&P.d_natrows, iord))) return rc;  // block rows in original order
  if (kb && (rc = upload(&P.d_kbsrc, kbsrc))) return rc;
  if ((rc = build_fac_pairs(ctx, P))) return rc;
  if (P.tiled) {
    // what the tile sweeps assume: at most kTileMax rows per tile, every row coupled to
    // every other row of its tile, so that its intra-tile blocks are the ones next to its
    // diagonal block
    const int64_t ntl = (int64_t)P.h_tileStart.size() - 1;
    for (int64_t t = 0; t < ntl; ++t) {
      const int32_t ts = P.h_tileStart[t], te = P.h_tileStart[t + 1];
      bool ok = te - ts <= kTileMax;
      for (int32_t p = ts; ok && p < te; ++p) {
        const int64_t b0 = bptr[p];
        const int32_t td = tdiag[p];
        if (td < p - ts || bptr[p + 1] - b0 < td + te - p) { ok = false; break; }
        for (int32_t q = ts; ok && q < te; ++q) ok = bcol[b0 + td + (q - p)] == q;
      }
      if (!ok) {
        delete L->ilu;
        L->ilu = nullptr;
        return set_error(SFEM_ERR_UNSUPPORTED, "tiled ILU(0) plan: a tile whose rows are not all coupled");
      }
    }
    P.tile_vals = 0;
    for (int64_t t = 0; t < ntl; ++t) {
      const int32_t tsz = P.h_tileStart[t + 1] - P.h_tileStart[t];
      for (int32_t p = P.h_tileStart[t]; p < P.h_tileStart[t + 1]; ++p) P.tile_vals += (int64_t)(bd[p].dm >> 16) * 4 * tsz;
    }
    // the last colour's tiles: no upper blocks outside the tile, so their backward sweep
    // joins the forward one
    const int nc = (int)P.tileColPtr.size() - 1;
    bool fl = nc > 0;
    for (int32_t t = nc > 0 ? P.tileColPtr[nc - 1] : 0; fl && t < P.tileColPtr[nc]; ++t) {
      const int32_t tend = P.h_tileStart[t + 1];
      for (int32_t p = P.h_tileStart[t]; fl && p < tend; ++p) fl = bcol[bptr[p + 1] - 1] < tend;
    }
    P.tile_fuse_last = fl;
    if ((rc = upload(&P.d_tileStart, P.h_tileStart))) return rc;
    // offsets of the sweeps' block-by-block copy of the blocks outside the tiles
    std::vector<int4> swoff(nb);
    int64_t nlo = 0, nup = 0;
    for (int64_t t = 0; t < ntl; ++t) {
      const int32_t ts = P.h_tileStart[t], te = P.h_tileStart[t + 1];
      for (int32_t q = ts; q < te; ++q) {
        const int32_t cnt = (int32_t)(bptr[q + 1] - bptr[q]);
        const int32_t lo = tdiag[q] - (q - ts), u0 = tdiag[q] + 1 + (te - 1 - q);
        swoff[q] = make_int4((int32_t)nlo, (int32_t)nup, lo, u0);
        nlo += lo;
        nup += cnt - u0;
      }
    }
    for (auto& o : swoff) o.y += (int32_t)nlo;
    if (nlo + nup >= ((int64_t)1 << 31) / 16) {
      delete L->ilu;
      L->ilu = nullptr;
      return set_error(SFEM_ERR_UNSUPPORTED, "tiled ILU(0) plan: too many blocks outside the tiles");
    }
    P.sw_blocks = nlo + nup;
    if ((rc = upload(&P.d_swoff, swoff))) return rc;
    // the sweeps' per (tile, row) records: {p, first bcol entry, count | m << 16, first
    // block of the copy} for the blocks outside the tile, lower (forward) and upper
    // (backward); rows past a tile's size hold p = -1
    std::vector<int4> trL((size_t)ntl * kTileMax, make_int4(-1, 0, 0, 0)), trU = trL;
    for (int64_t t = 0; t < ntl; ++t) {
      const int32_t ts = P.h_tileStart[t], te = P.h_tileStart[t + 1];
      for (int32_t q = ts; q < te; ++q) {
        const int4 o = swoff[q];
        const int32_t cnt = (int32_t)(bptr[q + 1] - bptr[q]), m = bd[q].dm >> 16;
        const size_t k = (size_t)t * kTileMax + (q - ts);
        trL[k] = make_int4(q, (int32_t)bptr[q], o.z | (m << 16), o.x);
        trU[k] = make_int4(q, (int32_t)(bptr[q] + o.w), (cnt - o.w) | (m << 16), o.y);
      }
    }
    if ((rc = upload(&P.d_trowL, trL)) || (rc = upload(&P.d_trowU, trU))) return rc;
  }
  return SFEM_OK;
}

template <typename T>
static int ensure_vec(T** p, int64_t n) {
  if (*p) return SFEM_OK;
  hipError_t e = alloc(p, (size_t)n);
  if (e != hipSuccess) return set_hip_error(e, "hipMalloc krylov vector");
  return SFEM_OK;
}

// One factorisation level: the scheduled kernel, or the search-based one.
template <typename LT>
static void fac_launch(hipStream_t st, const int32_t* list, int first, int cnt, const IluPlan& P, LT* lu, LT* fu,
                       const LT* a0, double* scal, int cap) {
  if (P.fac_sched)
    hipLaunchKernelGGL((k_bsr_factor_grp<LT>), dim3((cnt + kFacRowsPerWave - 1) / kFacRowsPerWave), dim3(64),
                       kFacRowsPerWave * fac_group_bytes(P.nbmax, P.fac_maxpairs), st, list, first, cnt, P.d_bdesc,
                       P.d_bcol, P.d_fpofs, P.d_fprec, lu, fu, a0, scal, P.nbmax, P.fac_maxpairs);
  else
    hipLaunchKernelGGL((k_bsr_factor<LT>), dim3(cnt), dim3(64), fac_row_bytes<LT>(cap), st, list, first, cnt,
                       P.d_bdesc, P.d_bcol, lu, a0, scal, cap);
}

// Blocks of level lv of the forward (upper=false) or backward (upper=true) sweep: a list
// slice, or with a contiguous plan the block range of colour lv (reversed for backward).
static void level_range(const IluPlan& P, bool upper, int lv, const int32_t** list, int* first, int* count) {
  if (P.contiguous) {
    const int nl = (int)P.levLPtr.size() - 1;
    const int c = upper ? nl - 1 - lv : lv;
    *list = nullptr;
    *first = P.levLPtr[c];
    *count = P.levLPtr[c + 1] - P.levLPtr[c];
  } else {
    const std::vector<int32_t>& ptr = upper ? P.levUPtr : P.levLPtr;
    *list = upper ? P.d_levU : P.d_levL;
    *first = ptr[lv];
    *count = ptr[lv + 1] - ptr[lv];
  }
}

// Build the preconditioner from the current matrix (FP64 vals or FP32 vals32).
// refactor = false (ILU(0) only): refresh the node-block tangent the Krylov products use
// and keep the existing factor (a lagged preconditioner).
int la_setup_pc(sfem_ctx* ctx, int pc_type, bool fp32, int reorder, bool bf16f, bool refactor, bool store32) {
  LinAlg* L = ctx->la;
  hipStream_t st = ctx->stream;
  const int64_t n = L->n;
  if (store32 && (fp32 || pc_type != SFEM_PC_ILU0)) return set_error(SFEM_ERR_ARG, "FP32 storage needs FP64 ILU(0)");
  L->pc_type = pc_type;
  if (refactor) {
    // a lagged refresh (refactor = false) keeps the factor, so it keeps the precision
    // flags the factorisation set (the caller checked that they match the request)
    L->pc_fp32 = fp32;
    L->pc_store32 = false;
  }
  if (pc_type != SFEM_PC_ILU0) {
    int rc = la_ensure_csr(ctx);
    if (rc) return rc;
  }
  const double* vals = fp32 ? nullptr : ctx->d_vals;
  const float* vals32 = fp32 ? ctx->d_vals32 : nullptr;
  LA_CHECK(hipMemsetAsync(L->d_scal + S_MISC1, 0, sizeof(double), st));
  if (pc_type == SFEM_PC_JACOBI) {
    if (!L->all_diag) return set_error(SFEM_ERR_ARG, "Jacobi: missing diagonal entry");
    if (fp32) {
      int rc = ensure_vec(&L->dinv32, n); if (rc) return rc;
      hipLaunchKernelGGL((k_jac_setup<float>), dim3((unsigned)((n + kTPB - 1) / kTPB)), dim3(kTPB), 0, st,
                         n, L->d_diagpos, vals, vals32, L->dinv32, L->d_scal);
    } else {
      int rc = ensure_vec(&L->dinv, n); if (rc) return rc;
      hipLaunchKernelGGL((k_jac_setup<double>), dim3((unsigned)((n + kTPB - 1) / kTPB)), dim3(kTPB), 0, st,
                         n, L->d_diagpos, vals, vals32, L->dinv, L->d_scal);
    }
    LA_CHECK(hipGetLastError());
  } else if (pc_type == SFEM_PC_BJACOBI4) {
    const int64_t nb = L->nblk;
    const unsigned g = (unsigned)((nb + kTPB - 1) / kTPB);
    if (fp32) {
      int rc = ensure_vec(&L->binv32, nb * 16); if (rc) return rc;
      if (nb) hipLaunchKernelGGL((k_bjac_setup<float>), dim3(g), dim3(kTPB), 0, st, nb, L->d_blkStart,
                                 ctx->d_rowptr32, ctx->d_colidx, vals, vals32, L->binv32, L->d_scal);
    } else {
      int rc = ensure_vec(&L->binv, nb * 16); if (rc) return rc;
      if (nb) hipLaunchKernelGGL((k_bjac_setup<double>), dim3(g), dim3(kTPB), 0, st, nb, L->d_blkStart,
                                 ctx->d_rowptr32, ctx->d_colidx, vals, vals32, L->binv, L->d_scal);
    }
    LA_CHECK(hipGetLastError());
  } else if (pc_type == SFEM_PC_ILU0) {
    if (!L->all_diag) return set_error(SFEM_ERR_ARG, "ILU(0): missing diagonal entry");
    if (!L->ilu || L->ilu->reorder != reorder) {
      int rc = build_ilu_plan(ctx, L, reorder);
      if (rc) return rc;
    }
    const IluPlan& P = *L->ilu;
    const int64_t nv = P.nvals;
    // the gather gives the block tangent of the Krylov products; the factorisation reads
    // its starting values from it
    const unsigned gb = (unsigned)std::max<int64_t>((P.nblk + 3) / 4, 1);
    if (ctx->kb_on && ctx->kb_fp32 != fp32) return set_error(SFEM_ERR_STATE, "the tangent was assembled in the other precision");
    // the multiplicative coarse correction's A Z formed by the gather (plan already known to
    // the correction, products on the gathered tangent itself)
    const int32_t *az_ci = nullptr, *az_off = nullptr, *az_box = nullptr;
    void* az_val = nullptr;
    const bool fuse_az = ctx->kb_on && !store32 && coarse_az_targets(L, &az_ci, &az_off, &az_box, &az_val);
    if (fp32) {
      int rc;
      if ((rc = ensure_vec(&L->kp32, nv)) || (rc = ensure_vec(&L->lu32, nv))) return rc;
      if (fuse_az)
        hipLaunchKernelGGL((k_bsr_gather_kb_az<float>), dim3(gb), dim3(256), 0, st, P.nblk, P.d_bdesc, P.d_kbsrc,
                           ctx->d_kb32, L->kp32, P.d_bcol, az_ci, az_off, az_box, (float*)az_val);
      else if (ctx->kb_on)
        hipLaunchKernelGGL((k_bsr_gather_kb<float>), dim3(gb), dim3(256), 0, st, P.nblk, P.d_bdesc, P.d_kbsrc,
                           ctx->d_kb32, L->kp32);
      else
        hipLaunchKernelGGL((k_bsr_gather<float>), dim3(gb), dim3(256), 0, st, P.nblk, P.d_bdesc, P.d_bcolofs,
                           P.d_r0orig, ctx->d_rowptr32, ctx->d_vals32, L->kp32);
    } else {
      int rc;
      if ((rc = ensure_vec(&L->kp, nv)) || (rc = ensure_vec(&L->lu, nv))) return rc;
      if (fuse_az)
        hipLaunchKernelGGL((k_bsr_gather_kb_az<double>), dim3(gb), dim3(256), 0, st, P.nblk, P.d_bdesc, P.d_kbsrc,
                           ctx->d_kb, L->kp, P.d_bcol, az_ci, az_off, az_box, (double*)az_val);
      else if (ctx->kb_on)
        hipLaunchKernelGGL((k_bsr_gather_kb<double>), dim3(gb), dim3(256), 0, st, P.nblk, P.d_bdesc, P.d_kbsrc,
                           ctx->d_kb, L->kp);
      else
        hipLaunchKernelGGL((k_bsr_gather<double>), dim3(gb), dim3(256), 0, st, P.nblk, P.d_bdesc, P.d_bcolofs,
                           P.d_r0orig, ctx->d_rowptr32, ctx->d_vals, L->kp);
    }
    // FP32 storage: the block tangent (and below, the FP64-computed factor) rounded to FP32
    auto to32 = [&](const double* a, float** b) -> int {
      int rc1 = ensure_vec(b, nv);
      if (rc1) return rc1;
      hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(nv)), dim3(kTPB), 0, st, nv, a, *b, 1.0);
      return SFEM_OK;
    };
    if (store32) {
      int rc = to32(L->kp, &L->kp32);
      if (rc) return rc;
    }
    // the tangent the Krylov products use: FP32 when built in FP32 or stored in FP32
    const bool op32 = fp32 || store32 || (!refactor && L->pc_store32);
    if (fuse_az) coarse_set_az_precision(L, fp32);
    if (!refactor) {
      // lagged factor; the multiplicative coarse correction's A Z follows the new tangent
      if (L->coarse_k > 0 && !fuse_az) {
        int rc = coarse_refresh_az(ctx, op32 ? nullptr : L->kp, op32 ? L->kp32 : nullptr);
        if (rc) return rc;
      }
      LA_CHECK(hipGetLastError());
      return SFEM_OK;
    }
    const int cap = (P.nbmax + 15) / 16 * 16;
    const int nl = (int)P.levLPtr.size() - 1;
    if (P.fac_sched) {
      int rc = fp32 ? ensure_vec(&L->fu32, P.npairs * 16) : ensure_vec(&L->fu, P.npairs * 16);
      if (rc) return rc;
    }
    for (int lv = 0; lv < nl; ++lv) {
      const int32_t* list; int first, cnt;
      level_range(P, false, lv, &list, &first, &cnt);
      if (!cnt) continue;
      if (fp32) fac_launch<float>(st, list, first, cnt, P, L->lu32, L->fu32, L->kp32, L->d_scal, cap);
      else fac_launch<double>(st, list, first, cnt, P, L->lu, L->fu, L->kp, L->d_scal, cap);
    }
    if (store32) {
      int rc = to32(L->lu, &L->lu32);
      if (rc) return rc;
      L->pc_fp32 = true;  // the factor the sweeps read is the FP32 copy
      L->pc_store32 = true;
    }
    L->pc_bf16 = (fp32 || store32) && bf16f;
    if (L->pc_bf16) {
      int rc = ensure_vec(&L->lu16, nv); if (rc) return rc;
      const unsigned gz = (unsigned)std::min<int64_t>((nv + 255) / 256, 1 << 20);
      hipLaunchKernelGGL(k_to_bf16, dim3(gz), dim3(256), 0, st, nv, L->lu32, L->lu16);
    }
    if (P.tiled) {
      // inverses of every tile's triangles, from the factor the sweeps read
      const int ntl = (int)P.h_tileStart.size() - 1;
      const int64_t ni = (int64_t)ntl * kTileInv;
      int rc;
      if (L->pc_bf16) {
        if ((rc = ensure_vec(&L->tinv16, ni))) return rc;
        hipLaunchKernelGGL((k_tile_invert<bf16, bf16>), dim3(ntl), dim3(64), 0, st, P.d_tileStart, ntl, P.d_bdesc,
                           L->lu16, L->tinv16, L->d_scal);
      } else if (L->pc_fp32) {
        if ((rc = ensure_vec(&L->tinv32, ni))) return rc;
        hipLaunchKernelGGL((k_tile_invert<float, float>), dim3(ntl), dim3(64), 0, st, P.d_tileStart, ntl, P.d_bdesc,
                           L->lu32, L->tinv32, L->d_scal);
      } else {
        if ((rc = ensure_vec(&L->tinv, ni))) return rc;
        hipLaunchKernelGGL((k_tile_invert<double, double>), dim3(ntl), dim3(64), 0, st, P.d_tileStart, ntl,
                           P.d_bdesc, L->lu, L->tinv, L->d_scal);
      }
      // the sweeps' block-by-block copy of the blocks outside the tiles, from the same factor
      const int64_t nsw = 16 * std::max<int64_t>(P.sw_blocks, 1);
      const unsigned nbk = (unsigned)P.nblk;
      if (L->pc_bf16) {
        if ((rc = ensure_vec(&L->sw16, nsw))) return rc;
        hipLaunchKernelGGL((k_tile_sweep_copy<bf16>), dim3(nbk), dim3(64), 0, st, P.nblk, P.d_bdesc, P.d_swoff,
                           L->lu16, L->sw16);
      } else if (L->pc_fp32) {
        if ((rc = ensure_vec(&L->sw32, nsw))) return rc;
        hipLaunchKernelGGL((k_tile_sweep_copy<float>), dim3(nbk), dim3(64), 0, st, P.nblk, P.d_bdesc, P.d_swoff,
                           L->lu32, L->sw32);
      } else {
        if ((rc = ensure_vec(&L->sw, nsw))) return rc;
        hipLaunchKernelGGL((k_tile_sweep_copy<double>), dim3(nbk), dim3(64), 0, st, P.nblk, P.d_bdesc, P.d_swoff,
                           L->lu, L->sw);
      }
    }
    LA_CHECK(hipGetLastError());
    // the coarse pressure correction's E = Z^T A Z from the tangent of this factor
    if (L->coarse_k > 0) {
      int rc = coarse_build(ctx, fp32 ? nullptr : L->kp, fp32 ? L->kp32 : nullptr);
      if (!rc && !fuse_az) rc = coarse_refresh_az(ctx, op32 ? nullptr : L->kp, op32 ? L->kp32 : nullptr);
      if (rc) return rc;
    }
  }
  return SFEM_OK;
}

template <typename T>
static int pc_apply_impl(sfem_ctx* ctx, const T* r, T* z, const RowSubset* sub);
template <typename T>
static void spmv_in(sfem_ctx* ctx, const T* x, T* y, bool block, const RowSubset* sub = nullptr);

template <typename T>
static int pc_apply(sfem_ctx* ctx, const T* r, T* z, const RowSubset* sub = nullptr) {
  ctx->kt.begin(SFEM_KT_PC, ctx->stream);
  const int rc = pc_apply_impl<T>(ctx, r, z, sub);
  ctx->kt.end(SFEM_KT_PC, ctx->stream);
  return rc;
}

// Block kernel shapes (SpMV and triangular solves): code = vi | (xcd << 5), with
// G = kBlkG[vi] lanes per block row and K = kBlkK[vi] neighbour blocks per lane whose
// loads are issued together.
static const int kBlkG[8] = {4, 4, 8, 8, 16, 16, 32, 64};
static const int kBlkK[8] = {1, 2, 1, 2, 1, 2, 1, 1};
constexpr int kBlkCodes = 8;
constexpr int kBlkDefault = 6;  // G = 32, K = 1, no XCD remap

template <typename T, typename LT, bool UPPER, int G, int K>
static void tri_launch_g(hipStream_t st, int xcd, const int32_t* list, int first, int cnt, const IluPlan& P,
                         const LT* lu, const T* r, T* y, int fuse_u) {
  const unsigned gr = (unsigned)(((int64_t)cnt * G + 255) / 256);
  if constexpr (UPPER)
    hipLaunchKernelGGL((k_bsr_usolve<T, LT, G, K>), dim3(gr), dim3(256), 0, st, list, first, cnt, P.d_bdesc,
                       P.d_bcol, lu, y, xcd);
  else
    hipLaunchKernelGGL((k_bsr_lsolve<T, LT, G, K>), dim3(gr), dim3(256), 0, st, list, first, cnt, P.d_bdesc,
                       P.d_bcol, lu, r, y, xcd, fuse_u);
}

template <typename T, typename LT, bool UPPER>
static void tri_launch(hipStream_t st, int code, const int32_t* list, int first, int cnt, const IluPlan& P,
                       const LT* lu, const T* r, T* y, int fuse_u = 0) {
  const int xcd = code >> 5;
#define SFEM_TRI(i, G, K) \
  case i: tri_launch_g<T, LT, UPPER, G, K>(st, xcd, list, first, cnt, P, lu, r, y, fuse_u); break;
  switch (code & 31) {
    SFEM_TRI(0, 4, 1) SFEM_TRI(1, 4, 2) SFEM_TRI(2, 8, 1) SFEM_TRI(3, 8, 2)
    SFEM_TRI(4, 16, 1) SFEM_TRI(5, 16, 2) SFEM_TRI(6, 32, 1)
    default: tri_launch_g<T, LT, UPPER, 64, 1>(st, xcd, list, first, cnt, P, lu, r, y, fuse_u); break;
  }
#undef SFEM_TRI
}

// Apply the ILU(0) preconditioner: z = U^{-1} L^{-1} r in the plan's padded block
// layout, the forward solve into z and the backward solve in place.  Each level launch
// uses its autotuned shape, or `force` for all levels when force >= 0; with ev non-null
// an event is recorded after every level (ev[0] before the first), forward levels then
// backward levels.
template <typename T, typename LT>
static void ilu_solve(sfem_ctx* ctx, const IluPlan& P, const LT* lu, const T* r, T* z, int force = -1,
                      hipEvent_t* ev = nullptr, const RowSubset* sub = nullptr) {
  LinAlg* L = ctx->la;
  hipStream_t st = ctx->stream;
  const int slot = sizeof(T) == 8 ? 0 : 1;
  const std::vector<int>& tl = L->tune_L[slot];
  const std::vector<int>& tu = L->tune_U[slot];
  const int nl = (int)P.levLPtr.size() - 1;
  const int nu = (int)P.levUPtr.size() - 1;
  if (ev) (void)hipEventRecord(ev[0], st);
  if (P.tiled && !sub) {
    // one launch per colour and sweep, one workgroup per tile; with ev an event after
    // every launch (forward colours, then backward colours)
    const int nc = (int)P.tileColPtr.size() - 1;
    const int tf = P.tile_fuse_last ? 1 : 0;
    const void* ti = std::is_same<LT, double>::value  ? (const void*)L->tinv
                     : std::is_same<LT, float>::value ? (const void*)L->tinv32
                                                      : (const void*)L->tinv16;
    // the sweeps read the block-by-block copy of the factor's blocks outside the tiles
    const LT* sw = std::is_same<LT, double>::value  ? (const LT*)(const void*)L->sw
                   : std::is_same<LT, float>::value ? (const LT*)(const void*)L->sw32
                                                    : (const LT*)(const void*)L->sw16;
    for (int c = 0; c < nc; ++c) {
      const int t0 = P.tileColPtr[c], cnt = P.tileColPtr[c + 1] - t0;
      const int code = force >= 0 ? force : c < (int)tl.size() ? tl[c] : kTileDefault;
      if (cnt) tile_launch<T, LT, false>(st, code, P, t0, cnt, sw, ti, r, z, c == nc - 1 ? tf : 0);
      if (ev) (void)hipEventRecord(ev[1 + c], st);
    }
    for (int k = 0; k < nc - tf; ++k) {
      const int c = nc - 1 - tf - k;
      const int t0 = P.tileColPtr[c], cnt = P.tileColPtr[c + 1] - t0;
      const int code = force >= 0 ? force : k < (int)tu.size() ? tu[k] : kTileDefault;
      if (cnt) tile_launch<T, LT, true>(st, code, P, t0, cnt, sw, ti, z, z);
      if (ev) (void)hipEventRecord(ev[1 + nc + k], st);
    }
    return;
  }
  // the last forward level does the first backward level too when that is the same set
  // of rows without upper neighbours
  const int fuse = P.fuse_last ? 1 : 0;
  for (int lv = 0; lv < nl; ++lv) {
    const int32_t* list; int first, cnt;
    if (sub) { list = sub->d_lev; first = sub->lofs[lv]; cnt = sub->lofs[lv + 1] - first; }
    else level_range(P, false, lv, &list, &first, &cnt);
    const int code = force >= 0 ? force : lv < (int)tl.size() ? tl[lv] : kBlkDefault;
    if (cnt) tri_launch<T, LT, false>(st, code, list, first, cnt, P, lu, r, z, lv == nl - 1 ? fuse : 0);
    if (ev) (void)hipEventRecord(ev[1 + lv], st);
  }
  if (ev && fuse) (void)hipEventRecord(ev[1 + nl], st);
  for (int lv = fuse; lv < nu; ++lv) {
    const int32_t* list; int first, cnt;
    if (sub) { list = sub->d_lev; first = sub->uofs[lv]; cnt = sub->uofs[lv + 1] - first; }
    else level_range(P, true, lv, &list, &first, &cnt);
    const int code = force >= 0 ? force : lv < (int)tu.size() ? tu[lv] : kBlkDefault;
    if (cnt) tri_launch<T, LT, true>(st, code, list, first, cnt, P, lu, z, z);
    if (ev) (void)hipEventRecord(ev[1 + nl + lv], st);
  }
}

// The factor (and, for a tiled plan, the tile inverses and the sweeps' off-tile copy)
// that ilu_apply reads for the current precision flags exist.
static bool ilu_buffers_ready(const LinAlg* L) {
  if (!L->ilu) return false;
  const void* lu = L->pc_bf16 ? (const void*)L->lu16 : L->pc_fp32 ? (const void*)L->lu32 : (const void*)L->lu;
  if (!lu) return false;
  if (!L->ilu->tiled) return true;
  const void* ti = L->pc_bf16 ? (const void*)L->tinv16 : L->pc_fp32 ? (const void*)L->tinv32 : (const void*)L->tinv;
  const void* sw = L->pc_bf16 ? (const void*)L->sw16 : L->pc_fp32 ? (const void*)L->sw32 : (const void*)L->sw;
  return ti && sw;
}

// ILU(0) apply with the factor type of the current preconditioner.
template <typename T>
static void ilu_apply(sfem_ctx* ctx, const T* r, T* z, int force = -1, hipEvent_t* ev = nullptr,
                      const RowSubset* sub = nullptr) {
  LinAlg* L = ctx->la;
  const IluPlan& P = *L->ilu;
  if (L->pc_bf16) ilu_solve<T, bf16>(ctx, P, L->lu16, r, z, force, ev, sub);
  else if (L->pc_fp32) ilu_solve<T, float>(ctx, P, L->lu32, r, z, force, ev, sub);
  else ilu_solve<T, double>(ctx, P, L->lu, r, z, force, ev, sub);
}

template <typename T>
static int pc_apply_impl(sfem_ctx* ctx, const T* r, T* z, const RowSubset* sub) {
  LinAlg* L = ctx->la;
  hipStream_t st = ctx->stream;
  const int64_t n = L->n;
  const int g = grid_for(n);
  switch (L->pc_type) {
    case SFEM_PC_JACOBI:
      if (L->pc_fp32) hipLaunchKernelGGL((k_jac_apply<T, float>), dim3(g), dim3(kTPB), 0, st, n, L->dinv32, r, z);
      else hipLaunchKernelGGL((k_jac_apply<T, double>), dim3(g), dim3(kTPB), 0, st, n, L->dinv, r, z);
      break;
    case SFEM_PC_BJACOBI4: {
      const unsigned gb = (unsigned)((L->nblk + kTPB - 1) / kTPB);
      if (L->nblk == 0) break;
      if (L->pc_fp32) hipLaunchKernelGGL((k_bjac_apply<T, float>), dim3(gb), dim3(kTPB), 0, st, L->nblk, L->d_blkStart, L->binv32, r, z);
      else hipLaunchKernelGGL((k_bjac_apply<T, double>), dim3(gb), dim3(kTPB), 0, st, L->nblk, L->d_blkStart, L->binv, r, z);
      break;
    }
    case SFEM_PC_ILU0:
      if (!ilu_buffers_ready(L)) return set_error(SFEM_ERR_STATE, "ILU(0) factor buffers missing for its precision");
      if (coarse_active(L) && !L->coarse_mult) {
        // additive coarse correction: r restricted before the sweeps (z may alias r)
        coarse_restrict<T>(ctx, r);
        ilu_apply<T>(ctx, r, z, -1, nullptr, sub);
        coarse_correct<T>(ctx, z);
      } else if (coarse_active(L)) {
        // multiplicative: z = M^-1 r, then z += Z E^-1 Z^T (r - A z) (the product into the
        // correction's own buffer; BiCGStab uses the fused form in pc_op instead)
        if (r == z) return set_error(SFEM_ERR_ARG, "multiplicative coarse correction: r and z must differ");
        T* az = coarse_tmp<T>(ctx);
        ilu_apply<T>(ctx, r, z, -1, nullptr, sub);
        spmv_in<T>(ctx, z, az, true, sub);
        coarse_restrict<T>(ctx, r, az);
        coarse_correct<T>(ctx, z);
      } else {
        ilu_apply<T>(ctx, r, z, -1, nullptr, sub);
      }
      break;
    default:
      hipLaunchKernelGGL((k_copy<T>), dim3(g), dim3(kTPB), 0, st, n, r, z);
  }
  return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "preconditioner launch");
}

// (lanes per node, columns in flight per lane) of the tunable kernels
// SpMV kernel shapes: code = vi | (xcd << 5), G = kSpmvG[vi], U = kSpmvU[vi] (0: all
// of a lane's columns, up to 128 / G, loaded at once).
static const int kSpmvG[9] = {8, 8, 8, 16, 16, 16, 32, 32, 32};
static const int kSpmvU[9] = {2, 4, 0, 2, 4, 0, 2, 4, 0};
constexpr int kSpmvCodes = 9;

template <typename VT, typename T>
static void spmv_launch(hipStream_t st, int code, int64_t nb, const BlkDesc* desc, const int32_t* colidx,
                        const VT* vals, const T* x, T* y) {
  const int vi = code & 31, xcd = code >> 5;
  const unsigned g = (unsigned)((nb * kSpmvG[vi] + 255) / 256);
#define SFEM_SPMV(i, G, U)                                                                                   \
  case i:                                                                                                    \
    hipLaunchKernelGGL((k_spmv_blk<VT, T, G, U>), dim3(g), dim3(256), 0, st, nb, desc, colidx, vals, x, y, xcd); \
    break;
  switch (vi) {
    SFEM_SPMV(0, 8, 2) SFEM_SPMV(1, 8, 4) SFEM_SPMV(2, 8, 0)
    SFEM_SPMV(3, 16, 2) SFEM_SPMV(4, 16, 4) SFEM_SPMV(5, 16, 0)
    SFEM_SPMV(6, 32, 2) SFEM_SPMV(7, 32, 4)
    default: hipLaunchKernelGGL((k_spmv_blk<VT, T, 32, 0>), dim3(g), dim3(256), 0, st, nb, desc, colidx, vals, x, y, xcd); break;
  }
#undef SFEM_SPMV
}

// Block SpMV shapes: code = vi | (xcd << 5) | (spatial row order << 6).
template <typename VT, typename T>
static void bsr_spmv_launch(hipStream_t st, int code, const IluPlan& P, const VT* vals, const T* x, T* y,
                            const RowSubset* sub = nullptr) {
  const int vi = code & 31, xcd = (code >> 5) & 1;
  const int32_t* rows = sub ? sub->d_rows : (code >> 6) & 1 ? P.d_natrows : nullptr;
  const int64_t nr = sub ? sub->nrows : P.nblk;
  if (nr == 0) return;
  const unsigned g = (unsigned)((nr * kBlkG[vi] + 255) / 256);
#define SFEM_BSPMV(i, G, K)                                                                                  \
  case i:                                                                                                    \
    hipLaunchKernelGGL((k_bsr_spmv<VT, T, G, K>), dim3(g), dim3(256), 0, st, nr, P.d_bdesc, P.d_bcol, vals,     \
                       x, y, xcd, rows);                                                                     \
    break;
  switch (vi) {
    SFEM_BSPMV(0, 4, 1) SFEM_BSPMV(1, 4, 2) SFEM_BSPMV(2, 8, 1) SFEM_BSPMV(3, 8, 2)
    SFEM_BSPMV(4, 16, 1) SFEM_BSPMV(5, 16, 2) SFEM_BSPMV(6, 32, 1)
    default:
      hipLaunchKernelGGL((k_bsr_spmv<VT, T, 64, 1>), dim3(g), dim3(256), 0, st, nr, P.d_bdesc, P.d_bcol, vals,
                         x, y, xcd, rows);
      break;
  }
#undef SFEM_BSPMV
}

// y = K x in the DOF layout (block = false, CSR) or in the ILU(0) plan's padded block
// layout (block = true).
template <typename T>
static void spmv_in(sfem_ctx* ctx, const T* x, T* y, bool block, const RowSubset* sub) {
  LinAlg* L = ctx->la;
  const int slot = sizeof(T) == 8 ? 0 : 1;
  if (L->nblk == 0) return;
  ctx->kt.begin(SFEM_KT_SPMV, ctx->stream);
  if (block) {
    const int code = L->tune_bspmv[slot];
    if (sizeof(T) == 8 && !L->pc_store32 && !L->pc_fp32)
      bsr_spmv_launch<double, T>(ctx->stream, code, *L->ilu, L->kp, x, y, sub);
    else bsr_spmv_launch<float, T>(ctx->stream, code, *L->ilu, L->kp32, x, y, sub);
  } else {
    const int v = L->tune_spmv[slot];
    if (sizeof(T) == 8) spmv_launch<double, T>(ctx->stream, v, L->nblk, L->d_desc, ctx->d_colidx, ctx->d_vals, x, y);
    else spmv_launch<float, T>(ctx->stream, v, L->nblk, L->d_desc, ctx->d_colidx, ctx->d_vals32, x, y);
  }
  ctx->kt.end(SFEM_KT_SPMV, ctx->stream);
}

template <typename T>
static void spmv(sfem_ctx* ctx, const T* x, T* y) {
  spmv_in<T>(ctx, x, y, ctx->la->block_space);
}

// Autotuner, run once per precision and layout on the first solve with the context's own
// matrix (HIP events on the context stream).  The SpMV keeps the fastest of its shapes;
// the ILU(0) apply is timed level by level for every shape and each level keeps its own
// fastest, since light levels (few lower or upper neighbours) favour few lanes per block
// row and heavy ones many.
template <typename T>
static int autotune(sfem_ctx* ctx, T* x, T* y) {
  LinAlg* L = ctx->la;
  const int slot = sizeof(T) == 8 ? 0 : 1;
  hipStream_t st = ctx->stream;
  const bool on = ctx->kt.on;
  const bool block = L->block_space;
  ctx->kt.on = false;
  std::vector<hipEvent_t> ev(2, nullptr);
  auto cleanup = [&] {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    ctx->kt.on = on;
  };
  for (hipEvent_t& e : ev)
    if (hipEventCreate(&e) != hipSuccess) { e = nullptr; cleanup(); return set_error(SFEM_ERR_HIP, "autotune events"); }
  hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(L->nk)), dim3(kTPB), 0, st, L->nk, x, (T)1);
  float best = 1e30f;
  const int ncodes = block ? kBlkCodes : kSpmvCodes;
  int* cur = block ? &L->tune_bspmv[slot] : &L->tune_spmv[slot];
  int bestc = *cur;
  const int norders = block && L->ilu->d_natrows ? 2 : 1;
  for (int ord = 0; ord < norders; ++ord)
  for (int xcd = 0; xcd < 2; ++xcd)
    for (int vi = 0; vi < ncodes; ++vi) {
      *cur = vi | (xcd << 5) | (ord << 6);
      spmv<T>(ctx, x, y);  // warm
      (void)hipEventRecord(ev[0], st);
      for (int k = 0; k < 3; ++k) spmv<T>(ctx, x, y);
      (void)hipEventRecord(ev[1], st);
      (void)hipEventSynchronize(ev[1]);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
      if (ms < best) { best = ms; bestc = *cur; }
    }
  *cur = bestc;
  if (block) {
    const IluPlan& P = *L->ilu;
    // launches per sweep: levels, or colours for a tiled plan (whose last colour may do
    // its backward sweep within the forward launch)
    const int ntc = P.tiled ? (int)P.tileColPtr.size() - 1 : 0;
    const int nl = P.tiled ? ntc : (int)P.levLPtr.size() - 1;
    const int nu = P.tiled ? ntc - (P.tile_fuse_last ? 1 : 0) : (int)P.levUPtr.size() - 1;
    const int ncode = P.tiled ? kTileCodes : 2 * kBlkCodes;
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.assign(nl + nu + 1, nullptr);
    for (hipEvent_t& e : ev)
      if (hipEventCreate(&e) != hipSuccess) { e = nullptr; cleanup(); return set_error(SFEM_ERR_HIP, "autotune events"); }
    std::vector<float> bt(nl + nu, 1e30f), acc(nl + nu);
    std::vector<int> bc(nl + nu, P.tiled ? kTileDefault : kBlkDefault);
    for (int ci = 0; ci < ncode; ++ci) {
        const int code = P.tiled ? ci : (ci % kBlkCodes) | ((ci / kBlkCodes) << 5);
        ilu_apply<T>(ctx, x, y, code);  // warm
        std::fill(acc.begin(), acc.end(), 0.f);
        for (int k = 0; k < 3; ++k) {
          ilu_apply<T>(ctx, x, y, code, ev.data());
          if (hipEventSynchronize(ev.back()) != hipSuccess) { cleanup(); return set_error(SFEM_ERR_HIP, "autotune"); }
          for (int l = 0; l < nl + nu; ++l) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[l], ev[l + 1]);
            acc[l] += ms;
          }
        }
        for (int l = 0; l < nl + nu; ++l)
          if (acc[l] < bt[l]) { bt[l] = acc[l]; bc[l] = code; }
      }
    L->tune_L[slot].assign(bc.begin(), bc.begin() + nl);
    L->tune_U[slot].assign(bc.begin() + nl, bc.end());
    L->tune_plan[slot] = L->plan_id;
  } else {
    L->tuned_csr[slot] = true;
  }
  cleanup();
  return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "autotune");
}

// The current ILU(0) factor on the reordered point pattern: plan row (block p, slot a)
// is point row pbs[p] + a, pbs the running count of free rows in plan order.
int la_get_ilu(sfem_ctx* ctx, int64_t* nnz_out, int32_t* perm, int32_t* rowptr, int32_t* colidx, double* vals) {
  LinAlg* L = ctx->la;
  if (!L || L->pc_type != SFEM_PC_ILU0 || !L->ilu || L->pc_K_version < 0)
    return set_error(SFEM_ERR_STATE, "no ILU(0) factor has been built");
  const IluPlan& P = *L->ilu;
  const int64_t nb = P.nblk, nv = P.nvals;
  LA_CHECK(hipStreamSynchronize(ctx->stream));
  std::vector<BDesc> bd(nb);
  std::vector<int32_t> bcol(P.npairs), r0(nb);
  LA_CHECK(hipMemcpy(bd.data(), P.d_bdesc, nb * sizeof(BDesc), hipMemcpyDeviceToHost));
  LA_CHECK(hipMemcpy(bcol.data(), P.d_bcol, P.npairs * sizeof(int32_t), hipMemcpyDeviceToHost));
  LA_CHECK(hipMemcpy(r0.data(), P.d_r0orig, nb * sizeof(int32_t), hipMemcpyDeviceToHost));
  std::vector<double> lu(vals ? nv : 0);
  if (vals) {
    if (L->pc_bf16) {
      std::vector<uint16_t> h(nv);
      LA_CHECK(hipMemcpy(h.data(), L->lu16, nv * sizeof(uint16_t), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < nv; ++i) {
        const uint32_t u = (uint32_t)h[i] << 16;
        float f;
        std::memcpy(&f, &u, sizeof(f));
        lu[i] = f;
      }
    } else if (L->pc_fp32) {
      std::vector<float> h(nv);
      LA_CHECK(hipMemcpy(h.data(), L->lu32, nv * sizeof(float), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < nv; ++i) lu[i] = h[i];
    } else {
      LA_CHECK(hipMemcpy(lu.data(), L->lu, nv * sizeof(double), hipMemcpyDeviceToHost));
    }
  }
  std::vector<int32_t> pbs(nb + 1, 0);
  for (int64_t p = 0; p < nb; ++p) pbs[p + 1] = pbs[p] + (bd[p].dm >> 16);
  int64_t q = 0;
  if (rowptr) rowptr[0] = 0;
  for (int64_t p = 0; p < nb; ++p) {
    const BDesc& d = bd[p];
    const int m = d.dm >> 16;
    for (int a = 0; a < m; ++a) {
      const int32_t row = pbs[p] + a;
      if (perm) perm[row] = r0[p] + a;
      for (int t = 0; t < d.nb; ++t) {
        const int32_t j = bcol[d.bptr + t];
        const int mj = bd[j].dm >> 16;
        for (int b = 0; b < mj; ++b, ++q) {
          if (colidx) colidx[q] = pbs[j] + b;
          if (vals) vals[q] = lu[d.vbase + (int64_t)a * 4 * d.nb + 4 * t + b];
        }
      }
      if (rowptr) rowptr[row + 1] = (int32_t)q;
    }
  }
  if (nnz_out) *nnz_out = q;
  return SFEM_OK;
}

int la_block_info(sfem_ctx* ctx, int64_t* nblk, int64_t* npairs, int64_t* nvals) {
  const LinAlg* L = ctx->la;
  const IluPlan* P = L ? L->ilu : nullptr;
  if (nblk) *nblk = P ? P->nblk : 0;
  if (npairs) *npairs = P ? P->npairs : 0;
  if (nvals) *nvals = P ? P->nvals : 0;
  return SFEM_OK;
}

int la_tile_info(sfem_ctx* ctx, int64_t* ntiles, int64_t* ncolors, int64_t* tile_vals) {
  const LinAlg* L = ctx->la;
  const IluPlan* P = L ? L->ilu : nullptr;
  const bool t = P && P->tiled;
  if (ntiles) *ntiles = t ? (int64_t)P->h_tileStart.size() - 1 : 0;
  if (ncolors) *ncolors = t ? (int64_t)P->tileColPtr.size() - 1 : 0;
  if (tile_vals) *tile_vals = t ? P->tile_vals : 0;
  return SFEM_OK;
}

int la_tuning(sfem_ctx* ctx, int* out) {
  LinAlg* L = ctx->la;
  for (int k = 0; k < 2; ++k) {
    const bool blk = L && L->block_space && L->tune_plan[k] == L->plan_id;
    const bool csr = L && !L->block_space && L->tuned_csr[k];
    const int c = blk ? L->tune_bspmv[k] : csr ? L->tune_spmv[k] : 0;
    out[4 * k + 0] = blk ? kBlkG[c & 31] : csr ? kSpmvG[c & 31] : 0;
    out[4 * k + 1] = blk ? kBlkK[c & 31] : csr ? kSpmvU[c & 31] : 0;
    out[4 * k + 2] = (blk || csr) ? c >> 5 : 0;  // bit 0: XCD order, bit 1: spatial row order
    out[4 * k + 3] = blk ? (int)(L->tune_L[k].size() + L->tune_U[k].size()) : 0;
  }
  return SFEM_OK;
}

int la_tuning_levels(sfem_ctx* ctx, int fp32, int* out, int cap, int* nlevels) {
  LinAlg* L = ctx->la;
  const int k = fp32 ? 1 : 0;
  std::vector<int> all(L->tune_L[k]);
  all.insert(all.end(), L->tune_U[k].begin(), L->tune_U[k].end());
  *nlevels = (int)all.size();
  const bool tiled = L->ilu && L->ilu->tiled;
  for (int i = 0; i < (int)all.size() && i < cap; ++i) {
    // (lanes per block row, blocks per lane loaded together, XCD order)
    out[3 * i + 0] = tiled ? kTileGs[all[i]] : kBlkG[all[i] & 31];
    out[3 * i + 1] = tiled ? kTileKs[all[i]] : kBlkK[all[i] & 31];
    out[3 * i + 2] = tiled ? 0 : all[i] >> 5;
  }
  return SFEM_OK;
}

int la_ensure_csr(sfem_ctx* ctx) {
  if (!ctx->kb_on || ctx->csr_version == ctx->la_K_version) return SFEM_OK;
  LinAlg* L = ctx->la;
  const unsigned g = (unsigned)std::max<int64_t>((L->nblk + 3) / 4, 1);
  if (ctx->kb_fp32) {
    if (!ctx->d_vals32 || !ctx->d_kb32) return set_error(SFEM_ERR_STATE, "no FP32 tangent assembled");
    hipLaunchKernelGGL((k_kb_to_csr<float>), dim3(g), dim3(256), 0, ctx->stream, L->nblk, L->d_desc, ctx->d_blkPtr,
                       ctx->d_natofs, ctx->d_kb32, ctx->d_vals32);
  } else {
    hipLaunchKernelGGL((k_kb_to_csr<double>), dim3(g), dim3(256), 0, ctx->stream, L->nblk, L->d_desc, ctx->d_blkPtr,
                       ctx->d_natofs, ctx->d_kb, ctx->d_vals);
  }
  LA_CHECK(hipGetLastError());
  ctx->csr_version = ctx->la_K_version;
  return SFEM_OK;
}

int la_spmv64(sfem_ctx* ctx, const double* x, double* y) {
  int rc = la_ensure_csr(ctx);
  if (rc) return rc;
  spmv_in<double>(ctx, x, y, false);
  return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "spmv launch");
}
int la_spmv32(sfem_ctx* ctx, const float* x, float* y) {
  int rc = la_ensure_csr(ctx);
  if (rc) return rc;
  spmv_in<float>(ctx, x, y, false);
  return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "spmv launch");
}

// Sum of squares of a device vector into host (used for the Newton residual norm).
int la_norm2(sfem_ctx* ctx, const double* a, double* out) {
  LinAlg* L = ctx->la;
  const int g = grid_for(L->n);
  hipLaunchKernelGGL((k_dot<double>), dim3(g), dim3(kTPB), 0, ctx->stream, L->n, a, a, L->d_part);
  hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kTPB), 0, ctx->stream, L->d_part, g, L->d_scal,
                     (int)S_MISC0, (int)FIN_STORE);
  LA_CHECK(hipGetLastError());
  LA_CHECK(hipMemcpyAsync(L->h_scal + S_MISC0, L->d_scal + S_MISC0, sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
  LA_CHECK(hipStreamSynchronize(ctx->stream));
  *out = std::sqrt(L->h_scal[S_MISC0]);
  return SFEM_OK;
}

int la_residual_norm(sfem_ctx* ctx, const double* b, const double* x, double* out) {
  LinAlg* L = ctx->la;
  int rc;
  if ((rc = la_ensure_csr(ctx)) || (rc = ensure_vec(&L->w, L->nvec))) return rc;
  spmv_in<double>(ctx, x, L->w, false);
  hipLaunchKernelGGL((k_axpby<double>), dim3(grid_for(L->n)), dim3(kTPB), 0, ctx->stream, L->n, 1.0, b, -1.0, L->w);
  LA_CHECK(hipGetLastError());
  return la_norm2(ctx, L->w, out);
}

// Internal status of bicgstab with a stall window: no 10% improvement of the best
// residual for `stall` iterations.
constexpr int kStagnated = -100;
constexpr int kBicgStall = 100;  // stall window of the FP32 inner BiCGStab
constexpr int kMaxReplacements = 8;  // true-residual replacements per BiCGStab solve

// z = M^-1 r and out = A z.  With the multiplicative coarse correction the product of the
// ILU(0) output is formed once and reused: z0 = M_ILU^-1 r, out0 = A z0, y = E^-1 Z^T
// (r - out0), z = z0 + Z y, out = out0 + (A Z) y, with A Z kept per block row.
template <typename T>
static int pc_op(sfem_ctx* ctx, const T* r, T* z, T* out, const RowSubset* sub = nullptr) {
  LinAlg* L = ctx->la;
  if (!(L->pc_type == SFEM_PC_ILU0 && L->coarse_mult && coarse_active(L))) {
    int rc = pc_apply<T>(ctx, r, z, sub);
    if (rc) return rc;
    spmv_in<T>(ctx, z, out, L->block_space, sub);
    return SFEM_OK;
  }
  if (!ilu_buffers_ready(L)) return set_error(SFEM_ERR_STATE, "ILU(0) factor buffers missing for its precision");
  hipStream_t st = ctx->stream;
  ctx->kt.begin(SFEM_KT_PC, st);
  ilu_apply<T>(ctx, r, z, -1, nullptr, sub);
  ctx->kt.end(SFEM_KT_PC, st);
  spmv_in<T>(ctx, z, out, true, sub);
  ctx->kt.begin(SFEM_KT_PC, st);
  coarse_restrict<T>(ctx, r, out);
  coarse_correct_az<T>(ctx, z, out);
  ctx->kt.end(SFEM_KT_PC, st, false);
  return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "preconditioner launch");
}

template <typename T>
static int bicgstab(sfem_ctx* ctx, const T* b, T* x, double rtol, int maxit, int* iters,
                    double* relres, T* r, T* rh, T* p, T* v, T* s, T* t, T* ph, T* sh, int stall = 0) {
  LinAlg* L = ctx->la;
  hipStream_t st = ctx->stream;
  const int64_t n = L->nk;
  const int g = grid_for(n);
  double* sc = L->d_scal;
  double* hs = L->h_scal;
  *iters = 0;
  // x = 0, r = rh = b, ||b||
  hipLaunchKernelGGL((k_fill<T>), dim3(g), dim3(kTPB), 0, st, n, x, (T)0);
  hipLaunchKernelGGL((k_copy<T>), dim3(g), dim3(kTPB), 0, st, n, b, r);
  hipLaunchKernelGGL((k_copy<T>), dim3(g), dim3(kTPB), 0, st, n, b, rh);
  hipLaunchKernelGGL((k_fill<T>), dim3(g), dim3(kTPB), 0, st, n, v, (T)0);
  hipLaunchKernelGGL((k_fill<T>), dim3(g), dim3(kTPB), 0, st, n, p, (T)0);
  hipLaunchKernelGGL((k_dot<T>), dim3(g), dim3(kTPB), 0, st, n, b, b, L->d_part);
  hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, sc, (int)S_BB, (int)FIN_STORE);
  LA_CHECK(hipGetLastError());
  LA_CHECK(hipMemcpyAsync(hs + S_BB, sc + S_BB, sizeof(double), hipMemcpyDeviceToHost, st));
  LA_CHECK(hipStreamSynchronize(st));
  const double bnorm = std::sqrt(hs[S_BB]);
  if (!(bnorm > 0.0)) {
    *relres = 0.0;
    return std::isfinite(bnorm) ? SFEM_OK : set_error(SFEM_ERR_DIVERGED, "BiCGStab: non-finite rhs");
  }
  double init[S_NSLOTS];
  std::memset(init, 0, sizeof(init));
  init[S_RHO0] = 1.0; init[S_ALPHA] = 1.0; init[S_OMEGA] = 1.0;
  init[S_BB] = hs[S_BB];
  init[S_RR] = hs[S_BB];
  init[S_RHRH] = hs[S_BB];
  init[S_RHO1] = hs[S_BB];  // (rh, r) with rh = r = b
  init[S_TOL2] = (rtol * bnorm) * (rtol * bnorm);
  std::memcpy(hs, init, sizeof(init));
  LA_CHECK(hipMemcpyAsync(sc, hs, S_NSLOTS * sizeof(double), hipMemcpyHostToDevice, st));
  double rnorm = bnorm;
  bool first = true;
  int restarts = 0;
  int replacements = 0;
  double best = bnorm;
  int best_it = 0;
  for (int it = 1; it <= maxit; ++it) {
    // rho1 = (rh, r) and beta were formed with the previous iteration's residual update
    hipLaunchKernelGGL((k_bicg_p<T>), dim3(g), dim3(kTPB), 0, st, n, r, p, v, sc, first ? 1 : 0);
    first = false;
    int rc = pc_op<T>(ctx, p, ph, v); if (rc) return rc;
    hipLaunchKernelGGL((k_dot2<T>), dim3(g), dim3(kTPB), 0, st, n, rh, v, v, v, L->d_part);
    hipLaunchKernelGGL((k_finalize<2>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, sc, 0, (int)FIN_ALPHA);
    hipLaunchKernelGGL((k_bicg_s<T>), dim3(g), dim3(kTPB), 0, st, n, r, v, s, sc, L->d_part);
    hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, sc, 0, (int)FIN_SS);
    rc = pc_op<T>(ctx, s, sh, t); if (rc) return rc;
    hipLaunchKernelGGL((k_dot2<T>), dim3(g), dim3(kTPB), 0, st, n, t, t, t, s, L->d_part);
    hipLaunchKernelGGL((k_finalize<2>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, sc, 0, (int)FIN_OMEGA);
    hipLaunchKernelGGL((k_bicg_xr<T>), dim3(g), dim3(kTPB), 0, st, n, x, ph, sh, r, s, t, rh, sc, L->d_part);
    hipLaunchKernelGGL((k_finalize<2>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, sc, 0, (int)FIN_RR_RHO);
    LA_CHECK(hipGetLastError());
    LA_CHECK(hipMemcpyAsync(hs, sc, S_NSLOTS * sizeof(double), hipMemcpyDeviceToHost, st));
    LA_CHECK(hipStreamSynchronize(st));
    *iters = it;
    rnorm = std::sqrt(hs[S_RR]);
    if (!std::isfinite(rnorm)) {
      *relres = rnorm;
      return set_error(SFEM_ERR_DIVERGED, "BiCGStab: non-finite residual");
    }
    if (hs[S_FLAG] == 1.0 || rnorm <= rtol * bnorm) {
      // FP32 solves are the inner solves of a defect correction whose FP64 residual
      // decides convergence: they stop on the recurrence residual, as PETSc's bcgs does
      // (in FP32 it can sit up to an order of magnitude below b - A x).
      if (!std::is_same<T, double>::value) {
        *relres = rnorm / bnorm;
        return SFEM_OK;
      }
      // FP64 solves confirm convergence on the true residual; if that is still above
      // the tolerance it replaces the recurrence residual and the iteration restarts
      // from the current x with it as the shadow residual.
      spmv<T>(ctx, x, t);
      hipLaunchKernelGGL((k_copy<T>), dim3(g), dim3(kTPB), 0, st, n, b, r);
      hipLaunchKernelGGL((k_axpby<T>), dim3(g), dim3(kTPB), 0, st, n, -1.0, t, 1.0, r);
      hipLaunchKernelGGL((k_dot<T>), dim3(g), dim3(kTPB), 0, st, n, r, r, L->d_part);
      hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, sc, (int)S_MISC0,
                         (int)FIN_STORE);
      LA_CHECK(hipGetLastError());
      LA_CHECK(hipMemcpyAsync(hs + S_MISC0, sc + S_MISC0, sizeof(double), hipMemcpyDeviceToHost, st));
      LA_CHECK(hipStreamSynchronize(st));
      const double rr_true = hs[S_MISC0];
      rnorm = std::sqrt(rr_true);
      *relres = rnorm / bnorm;
      if (!std::isfinite(rnorm)) return set_error(SFEM_ERR_DIVERGED, "BiCGStab: non-finite residual");
      if (rnorm <= rtol * bnorm * 1.001) return SFEM_OK;
      if (++replacements > kMaxReplacements)
        return set_error(SFEM_ERR_MAXIT, "BiCGStab: true residual stays above the tolerance");
      hipLaunchKernelGGL((k_copy<T>), dim3(g), dim3(kTPB), 0, st, n, r, rh);
      hs[S_FLAG] = 0.0;
      hs[S_RHO0] = 1.0; hs[S_ALPHA] = 1.0; hs[S_OMEGA] = 1.0;
      hs[S_RR] = rr_true;
      hs[S_RHRH] = rr_true;
      hs[S_RHO1] = rr_true;  // (rh, r) with rh = r
      LA_CHECK(hipMemcpyAsync(sc, hs, S_NSLOTS * sizeof(double), hipMemcpyHostToDevice, st));
      first = true;
      continue;
    }
    if (rnorm < 0.9 * best) {
      best = rnorm;
      best_it = it;
    } else if (stall > 0 && it - best_it >= stall) {
      *relres = rnorm / bnorm;
      return kStagnated;
    }
    if (hs[S_FLAG] == 3.0) {
      // near breakdown: restart with the current residual as shadow residual
      if (++restarts > 100) {
        *relres = rnorm / bnorm;
        return set_error(SFEM_ERR_DIVERGED, "BiCGStab: repeated breakdown");
      }
      hipLaunchKernelGGL((k_copy<T>), dim3(g), dim3(kTPB), 0, st, n, r, rh);
      hs[S_FLAG] = 0.0;
      hs[S_RHO0] = 1.0; hs[S_ALPHA] = 1.0; hs[S_OMEGA] = 1.0;
      hs[S_RHRH] = hs[S_RR];
      hs[S_RHO1] = hs[S_RR];  // (rh, r) with rh = r
      LA_CHECK(hipMemcpyAsync(sc, hs, S_NSLOTS * sizeof(double), hipMemcpyHostToDevice, st));
      first = true;
    }
  }
  *relres = rnorm / bnorm;
  return set_error(SFEM_ERR_MAXIT, "BiCGStab: maximum iterations reached");
}

// Restarted GMRES(m), right preconditioned, with classical Gram-Schmidt applied twice
// (CGS2: the projections onto all earlier basis vectors are one batched reduction and
// one batched update per pass, so the orthogonalisation is 2 passes over the basis
// instead of k dependent dot products), deterministic reductions, and one host
// synchronisation per iteration to fetch the new Hessenberg column.  The small
// least-squares problem is solved on the host with Givens rotations.
template <typename T>
static int gmres(sfem_ctx* ctx, const T* b, T* x, int m, double rtol, int maxit, int* iters,
                 double* relres, T* V, T* w, T* z) {
  LinAlg* L = ctx->la;
  hipStream_t st = ctx->stream;
  const int64_t n = L->nk;
  const int g = grid_for(n);
  if (!L->d_h) {
    LA_CHECK(alloc(&L->d_h, (size_t)kGmresMaxH));
    LA_CHECK(alloc(&L->d_mpart, (size_t)kRedBlocks * 8));
    LA_CHECK(hipHostMalloc((void**)&L->h_h, kGmresMaxH * sizeof(double), hipHostMallocDefault));
  }
  if (2 * (m + 8) + 1 > kGmresMaxH) return set_error(SFEM_ERR_ARG, "GMRES: restart length too large");
  double* dh = L->d_h;             // [0, m+8): first pass, [m+8, 2m+16): second pass, [2m+16]: ||w||^2
  double* dh2 = L->d_h + (m + 8);
  double* dnn = L->d_h + 2 * (m + 8);
  double* hh = L->h_h;
  // project w against V_0..V_{k1-1} into out[0..k1) (out has room for k1 rounded up to 8)
  auto project = [&](int k1, double* out) {
    for (int j0 = 0; j0 < k1; j0 += 8) {
      const int nv = std::min(8, k1 - j0);
      hipLaunchKernelGGL((k_mdot8<T>), dim3(g), dim3(kTPB), 0, st, n, nv, V + (size_t)j0 * n, w, L->d_mpart);
      hipLaunchKernelGGL((k_finalize<8>), dim3(1), dim3(kTPB), 0, st, L->d_mpart, g, out, j0, (int)FIN_STORE);
    }
    hipLaunchKernelGGL((k_mupdate<T>), dim3(g), dim3(kTPB), 0, st, n, k1, V, out, -1.0, w);
  };
  auto norm2 = [&](const T* a, double* out) {
    hipLaunchKernelGGL((k_dot<T>), dim3(g), dim3(kTPB), 0, st, n, a, a, L->d_part);
    hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, out, 0, (int)FIN_STORE);
  };
  auto fetch = [&](const double* src, double* dst, int cnt) -> int {
    LA_CHECK(hipGetLastError());
    LA_CHECK(hipMemcpyAsync(dst, src, cnt * sizeof(double), hipMemcpyDeviceToHost, st));
    return SFEM_OK;
  };
  hipLaunchKernelGGL((k_fill<T>), dim3(g), dim3(kTPB), 0, st, n, x, (T)0);
  norm2(b, dnn);
  int rc = fetch(dnn, hh, 1); if (rc) return rc;
  LA_CHECK(hipStreamSynchronize(st));
  const double bnorm = std::sqrt(hh[0]);
  *iters = 0;
  if (!(bnorm > 0.0)) { *relres = 0.0; return std::isfinite(bnorm) ? SFEM_OK : set_error(SFEM_ERR_DIVERGED, "GMRES: non-finite rhs"); }
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m);
  int it = 0;
  double rnorm = bnorm;
  T* r = z;  // z holds the residual at a restart
  while (it < maxit) {
    spmv<T>(ctx, x, w);
    hipLaunchKernelGGL((k_copy<T>), dim3(g), dim3(kTPB), 0, st, n, b, r);
    hipLaunchKernelGGL((k_axpby<T>), dim3(g), dim3(kTPB), 0, st, n, -1.0, w, 1.0, r);
    norm2(r, dnn);
    if ((rc = fetch(dnn, hh, 1))) return rc;
    LA_CHECK(hipStreamSynchronize(st));
    const double beta = std::sqrt(hh[0]);
    rnorm = beta;
    if (!std::isfinite(beta)) return set_error(SFEM_ERR_DIVERGED, "GMRES: non-finite residual");
    if (beta <= rtol * bnorm) break;
    hipLaunchKernelGGL((k_convert<T, T>), dim3(g), dim3(kTPB), 0, st, n, r, V, 1.0 / beta);
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = beta;
    int k = 0;
    for (; k < m && it < maxit; ++k) {
      ++it;
      T* vk = V + (size_t)k * n;
      if ((rc = pc_apply<T>(ctx, vk, z))) return rc;
      spmv<T>(ctx, z, w);
      project(k + 1, dh);
      project(k + 1, dh2);
      norm2(w, dnn);
      if ((rc = fetch(dh, hh, 2 * (m + 8) + 1))) return rc;
      LA_CHECK(hipStreamSynchronize(st));
      for (int j = 0; j <= k; ++j) H[(size_t)j * m + k] = hh[j] + hh[m + 8 + j];
      const double hn = std::sqrt(std::max(hh[2 * (m + 8)], 0.0));
      H[(size_t)(k + 1) * m + k] = hn;
      if (!std::isfinite(hn)) return set_error(SFEM_ERR_DIVERGED, "GMRES: non-finite Krylov vector");
      if (hn != 0.0)
        hipLaunchKernelGGL((k_convert<T, T>), dim3(g), dim3(kTPB), 0, st, n, w, V + (size_t)(k + 1) * n, 1.0 / hn);
      for (int j = 0; j < k; ++j) {
        const double t1 = H[(size_t)j * m + k], t2 = H[(size_t)(j + 1) * m + k];
        H[(size_t)j * m + k] = cs[j] * t1 + sn[j] * t2;
        H[(size_t)(j + 1) * m + k] = -sn[j] * t1 + cs[j] * t2;
      }
      const double a = H[(size_t)k * m + k], bq = H[(size_t)(k + 1) * m + k];
      const double d = std::sqrt(a * a + bq * bq);
      cs[k] = d == 0.0 ? 1.0 : a / d;
      sn[k] = d == 0.0 ? 0.0 : bq / d;
      H[(size_t)k * m + k] = d;
      H[(size_t)(k + 1) * m + k] = 0.0;
      gv[k + 1] = -sn[k] * gv[k];
      gv[k] = cs[k] * gv[k];
      rnorm = std::fabs(gv[k + 1]);
      if (rnorm <= rtol * bnorm || hn == 0.0) { ++k; break; }
    }
    for (int j = k - 1; j >= 0; --j) {
      double sum = gv[j];
      for (int q = j + 1; q < k; ++q) sum -= H[(size_t)j * m + q] * y[q];
      y[j] = sum / H[(size_t)j * m + j];
    }
    // x += M^{-1} (V y)
    std::copy(y.begin(), y.begin() + k, hh);
    LA_CHECK(hipMemcpyAsync(dh, hh, k * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL((k_fill<T>), dim3(g), dim3(kTPB), 0, st, n, w, (T)0);
    hipLaunchKernelGGL((k_mupdate<T>), dim3(g), dim3(kTPB), 0, st, n, k, V, dh, 1.0, w);
    if ((rc = pc_apply<T>(ctx, w, z))) return rc;
    hipLaunchKernelGGL((k_axpby<T>), dim3(g), dim3(kTPB), 0, st, n, 1.0, z, 1.0, x);
    LA_CHECK(hipGetLastError());
    // the host copy of y must not be overwritten before the transfer has read it
    LA_CHECK(hipStreamSynchronize(st));
    // convergence is decided on the true residual at the top of the loop: in FP32 the
    // Givens estimate can sit below the tolerance while the true residual does not
  }
  // true residual
  spmv<T>(ctx, x, w);
  hipLaunchKernelGGL((k_copy<T>), dim3(g), dim3(kTPB), 0, st, n, b, z);
  hipLaunchKernelGGL((k_axpby<T>), dim3(g), dim3(kTPB), 0, st, n, -1.0, w, 1.0, z);
  norm2(z, dnn);
  if ((rc = fetch(dnn, hh, 1))) return rc;
  LA_CHECK(hipStreamSynchronize(st));
  rnorm = std::sqrt(hh[0]);
  *iters = it;
  *relres = rnorm / bnorm;
  if (!std::isfinite(rnorm)) return set_error(SFEM_ERR_DIVERGED, "GMRES: non-finite residual");
  // the true residual may sit a few ulps above the recurrence's estimate
  if (rnorm > rtol * bnorm * 1.001) return set_error(SFEM_ERR_MAXIT, "GMRES: maximum iterations reached");
  return SFEM_OK;
}

// Shapes are tuned on the first solve of a precision and layout, and again when an
// ILU(0) plan appears that they were not tuned for.
static bool needs_tuning(const LinAlg* L, int slot) {
  return L->block_space ? L->tune_plan[slot] != L->plan_id : !L->tuned_csr[slot];
}

constexpr int kSpmvDefault = 3;  // CSR SpMV: G = 16, U = 2

// SFEM_TUNE_FIXED: the default shapes for the current layout, with no timing, so that
// the summation order of every product (and with it the FP32 rounding) is the same on
// every run and device.
static void fixed_shapes(LinAlg* L, int slot) {
  if (L->block_space) {
    L->tune_bspmv[slot] = kBlkDefault;
    L->tune_L[slot].clear();
    L->tune_U[slot].clear();
    L->tune_plan[slot] = L->plan_id;
  } else {
    L->tune_spmv[slot] = kSpmvDefault;
    L->tuned_csr[slot] = true;
  }
}

int la_set_tuning_mode(sfem_ctx* ctx, int mode) {
  LinAlg* L = ctx->la;
  if (mode != SFEM_TUNE_TIMED && mode != SFEM_TUNE_FIXED) return set_error(SFEM_ERR_ARG, "bad tuning mode");
  L->tune_mode = mode;
  // the next solve of each precision times the shapes again or installs the defaults
  for (int k = 0; k < 2; ++k) {
    L->tune_plan[k] = -1;
    L->tuned_csr[k] = false;
  }
  return SFEM_OK;
}

// Table of one precision: {CSR SpMV code, block SpMV code, forward launches nL, backward
// launches nU, nL forward codes, nU backward codes}.
int la_get_tuning_table(sfem_ctx* ctx, int fp32, int* table, int cap, int* n) {
  LinAlg* L = ctx->la;
  const int k = fp32 ? 1 : 0;
  std::vector<int> t = {L->tune_spmv[k], L->tune_bspmv[k], (int)L->tune_L[k].size(), (int)L->tune_U[k].size()};
  t.insert(t.end(), L->tune_L[k].begin(), L->tune_L[k].end());
  t.insert(t.end(), L->tune_U[k].begin(), L->tune_U[k].end());
  *n = (int)t.size();
  for (int i = 0; i < (int)t.size() && i < cap; ++i) table[i] = t[i];
  return SFEM_OK;
}

int la_set_tuning_table(sfem_ctx* ctx, int fp32, const int* table, int n) {
  LinAlg* L = ctx->la;
  const int k = fp32 ? 1 : 0;
  if (n < 4 || table[2] < 0 || table[3] < 0 || n != 4 + table[2] + table[3])
    return set_error(SFEM_ERR_ARG, "tuning table: bad length");
  const bool spmv_ok = (table[0] & 31) < kSpmvCodes && (table[0] >> 5) <= 1 &&
                       (table[1] & 31) < kBlkCodes && (table[1] >> 5) <= 3;
  if (!spmv_ok) return set_error(SFEM_ERR_ARG, "tuning table: bad SpMV shape code");
  const int nl = table[2], nu = table[3];
  if (nl + nu > 0) {
    if (!L->ilu) return set_error(SFEM_ERR_STATE, "tuning table: set up the ILU(0) solver first");
    const IluPlan& P = *L->ilu;
    const int ntc = P.tiled ? (int)P.tileColPtr.size() - 1 : 0;
    const int pl = P.tiled ? ntc : (int)P.levLPtr.size() - 1;
    const int pu = P.tiled ? ntc - (P.tile_fuse_last ? 1 : 0) : (int)P.levUPtr.size() - 1;
    if (nl != pl || nu != pu) return set_error(SFEM_ERR_ARG, "tuning table: launch counts differ from the ILU(0) plan");
    for (int i = 4; i < n; ++i) {
      const int c = table[i];
      const bool ok = P.tiled ? (c >= 0 && c < kTileCodes) : (c >= 0 && (c & 31) < kBlkCodes && (c >> 5) <= 1);
      if (!ok) return set_error(SFEM_ERR_ARG, "tuning table: bad sweep shape code");
    }
  }
  L->tune_mode = SFEM_TUNE_FIXED;
  L->tune_spmv[k] = table[0];
  L->tuned_csr[k] = true;
  L->tune_bspmv[k] = table[1];
  L->tune_L[k].assign(table + 4, table + 4 + nl);
  L->tune_U[k].assign(table + 4 + nl, table + n);
  L->tune_plan[k] = L->ilu ? L->plan_id : -1;
  return SFEM_OK;
}

// Symbolic set-up of the preconditioner for the given options (ILU(0): ordering,
// block pattern and level sets), so that the first solve does not pay for it.
int la_prepare(sfem_ctx* ctx, const sfem_solver_opts* o) {
  LinAlg* L = ctx->la;
  if (o->pc_type == SFEM_PC_LU) return la_direct_setup(ctx);
  if (o->pc_type != SFEM_PC_ILU0) return SFEM_OK;
  if (L->ilu && L->ilu->reorder == o->reorder) return SFEM_OK;
  return build_ilu_plan(ctx, L, o->reorder);
}

// Solve A du = rhs.  FP64 mode: everything in double.  Mixed mode: FP32 matrix and
// Krylov vectors; the FP64 rhs is scaled by 1/||rhs|| into FP32 and the FP32 solution
// scaled back into the FP64 increment.  With ILU(0) the Krylov solve runs in the plan's
// padded block layout (rhs in, solution out through the plan's slot map).
// Preconditioner (rebuilt when the tangent or the options changed; the tangent is
// constant within a time step), Krylov layout, and autotuned kernel shapes for opts.
int la_prepare_solve(sfem_ctx* ctx, const sfem_solver_opts* o) {
  LinAlg* L = ctx->la;
  const bool fp32 = ctx->precision == SFEM_PREC_MIXED;
  // FP64 Krylov over an FP32-stored tangent and factor (ILU(0) only; other
  // preconditioners run the plain FP64 path)
  const bool store32 = ctx->precision == SFEM_PREC_STORE32 && o->pc_type == SFEM_PC_ILU0;
  int rc;
  const int reorder = o->pc_type == SFEM_PC_ILU0 ? o->reorder : 0;
  const bool bf16f = (fp32 || store32) && o->pc_type == SFEM_PC_ILU0 && o->factor_bf16 != 0;
  if (o->pc_type != SFEM_PC_ILU0 && (rc = la_ensure_csr(ctx))) return rc;  // CSR products
  const bool same_pc = L->pc_K_version >= 0 && L->pc_type == o->pc_type && L->pc_fp32 == (fp32 || store32) &&
                       L->pc_store32 == store32 && L->pc_reorder == reorder && L->pc_bf16 == bf16f;
  if (!same_pc || (L->kb_version != ctx->la_K_version && ++L->lag_count >= L->pc_lag)) {
    L->pc_K_version = L->kb_version = -1;
    if ((rc = la_setup_pc(ctx, o->pc_type, fp32, reorder, bf16f, true, store32))) return rc;
    L->pc_reorder = reorder;
    L->pc_K_version = L->kb_version = ctx->la_K_version;
    L->lag_count = 0;
  } else if (L->kb_version != ctx->la_K_version) {
    // lagged preconditioner: the products use the new tangent, the factor stays
    if (o->pc_type == SFEM_PC_ILU0 && (rc = la_setup_pc(ctx, o->pc_type, fp32, reorder, bf16f, false, store32)))
      return rc;
    L->kb_version = ctx->la_K_version;
  }
  L->block_space = L->pc_type == SFEM_PC_ILU0;
  L->nk = L->block_space ? L->ilu->npad : L->n;
  if (fp32) {
    if ((rc = ensure_vec(&L->t32, L->nvec)) || (rc = ensure_vec(&L->v32, L->nvec))) return rc;
    if (needs_tuning(L, 1)) {
      if (L->tune_mode == SFEM_TUNE_FIXED) fixed_shapes(L, 1);
      else if ((rc = autotune<float>(ctx, L->t32, L->v32))) return rc;
    }
  } else {
    if ((rc = ensure_vec(&L->t, L->nvec)) || (rc = ensure_vec(&L->v, L->nvec))) return rc;
    if (needs_tuning(L, 0)) {
      if (L->tune_mode == SFEM_TUNE_FIXED) fixed_shapes(L, 0);
      else if ((rc = autotune<double>(ctx, L->t, L->v))) return rc;
    }
  }
  return SFEM_OK;
}

int la_apply_pc32(sfem_ctx* ctx, const float* r, float* z, const RowSubset* sub) {
  if (sub && ctx->la->pc_type != SFEM_PC_ILU0) return set_error(SFEM_ERR_UNSUPPORTED, "row subsets need ILU(0)");
  return pc_apply<float>(ctx, r, z, sub);
}
int la_apply_pc64(sfem_ctx* ctx, const double* r, double* z, const RowSubset* sub) {
  if (sub && ctx->la->pc_type != SFEM_PC_ILU0) return set_error(SFEM_ERR_UNSUPPORTED, "row subsets need ILU(0)");
  return pc_apply<double>(ctx, r, z, sub);
}
// z = M^{-1} r with the preconditioner of the last solve, host vectors in DOF order
// (ILU(0) runs in the plan's padded block layout, as inside the Krylov solves).
int la_pc_apply_host(sfem_ctx* ctx, const double* r, double* z) {
  LinAlg* L = ctx->la;
  if (!L || L->pc_K_version < 0) return set_error(SFEM_ERR_STATE, "no preconditioner has been built");
  hipStream_t st = ctx->stream;
  const int64_t n = L->n;
  const bool blk = L->pc_type == SFEM_PC_ILU0 && L->ilu;
  const int64_t nk = blk ? L->ilu->npad : n;
  double *dr = nullptr, *dz = nullptr, *pr = nullptr, *pz = nullptr;
  auto cleanup = [&] {
    double* ps[] = {dr, dz, pr, pz};
    for (double* q : ps)
      if (q) (void)hipFree(q);
  };
  if (alloc(&dr, (size_t)std::max<int64_t>(n, 1)) != hipSuccess || alloc(&dz, (size_t)std::max<int64_t>(n, 1)) != hipSuccess ||
      (blk && (alloc(&pr, (size_t)nk) != hipSuccess || alloc(&pz, (size_t)nk) != hipSuccess))) {
    cleanup();
    return set_error(SFEM_ERR_HIP, "pc apply buffers");
  }
  int rc = SFEM_OK;
  if (hipMemcpy(dr, r, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = set_error(SFEM_ERR_HIP, "copy r");
  if (!rc && blk) {
    const int32_t* padmap = L->ilu->d_padmap;
    hipLaunchKernelGGL((k_pad_in<double, double>), dim3(grid_for(nk)), dim3(kTPB), 0, st, nk, padmap, dr, pr, 1.0);
    rc = pc_apply<double>(ctx, pr, pz);
    if (!rc)
      hipLaunchKernelGGL((k_pad_out<double, double>), dim3(grid_for(nk)), dim3(kTPB), 0, st, nk, padmap, pz, dz, 1.0);
  } else if (!rc) {
    rc = pc_apply<double>(ctx, dr, dz);
  }
  if (!rc && (hipStreamSynchronize(st) != hipSuccess || hipMemcpy(z, dz, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
    rc = set_error(SFEM_ERR_HIP, "pc apply");
  cleanup();
  return rc;
}

// ---------------------------------------------------------------------------
// Allocation-free preconditioner application (sfem_pc_apply_noalloc).  The staging
// vectors that la_pc_apply_host allocates on every call live instead in one file-scope
// device array, `scratch`, of kNoAllocDoubles doubles (64 MiB), which every GPU the
// library runs on holds until the process ends.  Host code never passes the array's
// name to a kernel or a copy (on the host it is only a stand-in symbol): each device's
// address of its own copy is looked up once, with that device current, when the first
// context on it is created (la_scratch_init), and that address is used instead.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocDoubles = int64_t(8) << 20;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kMaxScratchDevices = 64;
static double* g_scratch_ptr[kMaxScratchDevices] = {nullptr};
static std::mutex g_scratch_mu;

int la_scratch_init(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return set_error(SFEM_ERR_UNSUPPORTED, "device index too large");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return SFEM_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return set_hip_error(e, "hipGetSymbolAddress(scratch)");
  g_scratch_ptr[device] = (double*)ptr;
  return SFEM_OK;
}

int64_t la_noalloc_capacity() { return kNoAllocDoubles; }

// z = M^{-1} r as la_pc_apply_host, with dr, dz (n each) and, for ILU(0), pr, pz (npad
// each) carved from this device's scratch array.  Every slot a kernel reads is written
// earlier in the same call (k_pad_in writes all npad slots, the sweeps every slot of each
// block row), so nothing is cleared.  Calls that overlap on one GPU share the array and
// race; this call returns after its work on the context's stream has finished.
int la_pc_apply_noalloc(sfem_ctx* ctx, const double* r, double* z) {
  LinAlg* L = ctx->la;
  if (!L || L->pc_K_version < 0) return set_error(SFEM_ERR_STATE, "no preconditioner has been built");
  double* base = (ctx->device >= 0 && ctx->device < kMaxScratchDevices) ? g_scratch_ptr[ctx->device] : nullptr;
  if (!base) return set_error(SFEM_ERR_STATE, "allocation-free scratch not set up on this device");
  hipStream_t st = ctx->stream;
  const int64_t n = L->n;
  const bool blk = L->pc_type == SFEM_PC_ILU0 && L->ilu;
  const int64_t nk = blk ? L->ilu->npad : 0;
  const int64_t need = 2 * n + 2 * nk;
  if (need > kNoAllocDoubles) {
    char msg[160];
    std::snprintf(msg, sizeof msg, "allocation-free apply needs %lld doubles of scratch, the maximum is %lld",
                  (long long)need, (long long)kNoAllocDoubles);
    return set_error(SFEM_ERR_ARG, msg);
  }
  double* dr = base;
  double* dz = base + n;
  double* pr = base + 2 * n;
  double* pz = pr + nk;
  LA_CHECK(hipMemcpyAsync(dr, r, n * sizeof(double), hipMemcpyHostToDevice, st));
  int rc;
  if (blk) {
    const int32_t* padmap = L->ilu->d_padmap;
    hipLaunchKernelGGL((k_pad_in<double, double>), dim3(grid_for(nk)), dim3(kTPB), 0, st, nk, padmap, dr, pr, 1.0);
    rc = pc_apply<double>(ctx, pr, pz);
    if (!rc)
      hipLaunchKernelGGL((k_pad_out<double, double>), dim3(grid_for(nk)), dim3(kTPB), 0, st, nk, padmap, pz, dz, 1.0);
  } else {
    rc = pc_apply<double>(ctx, dr, dz);
  }
  if (rc) {
    (void)hipStreamSynchronize(st);
    return rc;
  }
  LA_CHECK(hipGetLastError());
  LA_CHECK(hipMemcpyAsync(z, dz, n * sizeof(double), hipMemcpyDeviceToHost, st));
  LA_CHECK(hipStreamSynchronize(st));
  return SFEM_OK;
}

int la_apply_op32(sfem_ctx* ctx, const float* x, float* y, const RowSubset* sub) {
  if (sub && !ctx->la->block_space) return set_error(SFEM_ERR_UNSUPPORTED, "row subsets need the block layout");
  spmv_in<float>(ctx, x, y, ctx->la->block_space, sub);
  return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "spmv launch");
}
// The coarse step's buffers (FP64 padded vectors), allocated before any capture.
int la_coarse_step_prepare(sfem_ctx* ctx) {
  LinAlg* L = ctx->la;
  int rc;
  if ((rc = ensure_vec(&L->pr, L->nvec)) || (rc = ensure_vec(&L->pz, L->nvec))) return rc;
  return SFEM_OK;
}

// The coarse step's kernels alone (capturable): increment = Z E^{-1} Z^T rhs in FP64
// (DOF layout), with the correction already built and la_coarse_step_prepare run.
int la_coarse_step_kernels(sfem_ctx* ctx) {
  LinAlg* L = ctx->la;
  if (!coarse_active(L) || !L->pr || !L->pz) return set_error(SFEM_ERR_STATE, "coarse step: correction not built");
  hipStream_t st = ctx->stream;
  const int64_t nk = L->nk;
  const int32_t* padmap = L->ilu->d_padmap;
  hipLaunchKernelGGL((k_pad_in<double, double>), dim3(grid_for(nk)), dim3(kTPB), 0, st, nk, padmap, ctx->d_rhs, L->pr,
                     1.0);
  hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(nk)), dim3(kTPB), 0, st, nk, L->pz, 0.0);
  coarse_restrict<double>(ctx, L->pr);
  coarse_correct<double>(ctx, L->pz);
  hipLaunchKernelGGL((k_pad_out<double, double>), dim3(grid_for(nk)), dim3(kTPB), 0, st, nk, padmap, L->pz, ctx->d_du,
                     1.0);
  return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "coarse step launch");
}

// The Newton loop's coarse step: increment = Z E^{-1} Z^T rhs in FP64 (DOF layout), with
// the preconditioner (and E) brought up to date for opts first.
int la_coarse_step(sfem_ctx* ctx, const sfem_solver_opts* o) {
  LinAlg* L = ctx->la;
  if (L->coarse_k <= 0 || o->pc_type != SFEM_PC_ILU0 || ctx->batch)
    return set_error(SFEM_ERR_STATE, "coarse step: needs the coarse correction of ILU(0) on a single mesh");
  int rc = la_prepare_solve(ctx, o);
  if (rc) return rc;
  if (!coarse_active(L)) return set_error(SFEM_ERR_STATE, "coarse step: correction not built");
  hipStream_t st = ctx->stream;
  const int64_t nk = L->nk;
  const int32_t* padmap = L->ilu->d_padmap;
  if ((rc = ensure_vec(&L->pr, L->nvec)) || (rc = ensure_vec(&L->pz, L->nvec))) return rc;
  hipLaunchKernelGGL((k_pad_in<double, double>), dim3(grid_for(nk)), dim3(kTPB), 0, st, nk, padmap, ctx->d_rhs, L->pr,
                     1.0);
  hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(nk)), dim3(kTPB), 0, st, nk, L->pz, 0.0);
  coarse_restrict<double>(ctx, L->pr);
  coarse_correct<double>(ctx, L->pz);
  hipLaunchKernelGGL((k_pad_out<double, double>), dim3(grid_for(nk)), dim3(kTPB), 0, st, nk, padmap, L->pz, ctx->d_du,
                     1.0);
  return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "coarse step launch");
}

int la_apply_pc_op32(sfem_ctx* ctx, const float* r, float* z, float* out, const RowSubset* sub) {
  if (sub && ctx->la->pc_type != SFEM_PC_ILU0) return set_error(SFEM_ERR_UNSUPPORTED, "row subsets need ILU(0)");
  return pc_op<float>(ctx, r, z, out, sub);
}
int la_apply_pc_op64(sfem_ctx* ctx, const double* r, double* z, double* out, const RowSubset* sub) {
  if (sub && ctx->la->pc_type != SFEM_PC_ILU0) return set_error(SFEM_ERR_UNSUPPORTED, "row subsets need ILU(0)");
  return pc_op<double>(ctx, r, z, out, sub);
}

int la_apply_op64(sfem_ctx* ctx, const double* x, double* y, const RowSubset* sub) {
  if (sub && !ctx->la->block_space) return set_error(SFEM_ERR_UNSUPPORTED, "row subsets need the block layout");
  spmv_in<double>(ctx, x, y, ctx->la->block_space, sub);
  return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "spmv launch");
}

RowSubset::~RowSubset() {
  if (d_rows) (void)hipFree(d_rows);
  if (d_lev) (void)hipFree(d_lev);
}

int la_subset_build(sfem_ctx* ctx, const uint8_t* keep, RowSubset* S, const RowSubset* from) {
  LinAlg* L = ctx->la;
  if (!L->ilu) return set_error(SFEM_ERR_STATE, "row subset without an ILU(0) plan");
  const IluPlan& P = *L->ilu;
  if (from && (from == S || from->plan_id != L->plan_id)) return set_error(SFEM_ERR_ARG, "bad source subset");
  std::vector<int32_t> rows, lev, lofs, uofs;
  if (from) {
    for (int32_t p : from->h_rows)
      if (keep[p]) rows.push_back(p);
    auto add = [&](const std::vector<int32_t>& fo, std::vector<int32_t>& ofs) {
      ofs.assign(1, (int32_t)lev.size());
      for (size_t lv = 0; lv + 1 < fo.size(); ++lv) {
        for (int32_t i = fo[lv]; i < fo[lv + 1]; ++i)
          if (keep[from->h_lev[i]]) lev.push_back(from->h_lev[i]);
        ofs.push_back((int32_t)lev.size());
      }
    };
    add(from->lofs, lofs);
    add(from->uofs, uofs);
  } else {
    for (int64_t p = 0; p < P.nblk; ++p)
      if (keep[p]) rows.push_back((int32_t)p);
    auto add = [&](bool upper, std::vector<int32_t>& ofs) {
      const int nlev = (int)(upper ? P.levUPtr.size() : P.levLPtr.size()) - 1;
      const std::vector<int32_t>& hl = upper ? P.h_levU : P.h_levL;
      ofs.assign(1, (int32_t)lev.size());
      for (int lv = 0; lv < nlev; ++lv) {
        const int32_t* dl; int first, cnt;
        level_range(P, upper, lv, &dl, &first, &cnt);
        for (int i = 0; i < cnt; ++i) {
          const int32_t p = dl ? hl[first + i] : first + i;
          if (keep[p]) lev.push_back(p);
        }
        ofs.push_back((int32_t)lev.size());
      }
    };
    add(false, lofs);
    add(true, uofs);
  }
  for (int32_t** q : {&S->d_rows, &S->d_lev})
    if (*q) { (void)hipFree(*q); *q = nullptr; }
  S->nrows = (int64_t)rows.size();
  int rc;
  if ((rc = upload(&S->d_rows, rows)) || (rc = upload(&S->d_lev, lev))) return rc;
  S->h_rows.swap(rows);
  S->h_lev.swap(lev);
  S->lofs.swap(lofs);
  S->uofs.swap(uofs);
  S->plan_id = L->plan_id;
  return SFEM_OK;
}

int la_solve(sfem_ctx* ctx, const sfem_solver_opts* o_in, int* iters, double* relres) {
  LinAlg* L = ctx->la;
  if (o_in->pc_type == SFEM_PC_LU) return la_direct_solve(ctx, iters, relres);
  // PETSc's default test stops at ||r|| <= max(rtol ||b||, atol): with a zero initial
  // guess that is the relative test at rtol' = max(rtol, atol / ||b||)
  sfem_solver_opts oa = *o_in;
  const sfem_solver_opts* o = &oa;
  if (o_in->atol > 0.0) {
    double b0;
    int rc0 = la_norm2(ctx, ctx->d_rhs, &b0);
    if (rc0) return rc0;
    if (b0 <= o_in->atol) {
      hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(L->n)), dim3(kTPB), 0, ctx->stream, L->n, ctx->d_du, 0.0);
      *iters = 0;
      *relres = 0.0;
      return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "fill launch");
    }
    oa.rtol = std::max(o_in->rtol, o_in->atol / b0);
  }
  const int64_t n = L->n;
  const bool fp32 = ctx->precision == SFEM_PREC_MIXED;
  hipStream_t st = ctx->stream;
  int rc;
  if ((rc = la_prepare_solve(ctx, o))) return rc;
  const int32_t* padmap = L->block_space ? L->ilu->d_padmap : nullptr;
  const int64_t nk = L->nk;
  const int gk = grid_for(nk), gn = grid_for(n);
  const int m = std::max(1, o->restart);
  if (!fp32) {
    const double* b = ctx->d_rhs;
    double* x = ctx->d_du;
    if (padmap) {
      if ((rc = ensure_vec(&L->pr, L->nvec)) || (rc = ensure_vec(&L->pz, L->nvec))) return rc;
      hipLaunchKernelGGL((k_pad_in<double, double>), dim3(gk), dim3(kTPB), 0, st, nk, padmap, ctx->d_rhs, L->pr, 1.0);
      b = L->pr;
      x = L->pz;
    }
    if (o->ksp_type == SFEM_KSP_GMRES) {
      if ((rc = ensure_vec(&L->w, L->nvec)) || (rc = ensure_vec(&L->z, L->nvec))) return rc;
      if (!L->gm_V || L->gm_m < m) {
        if (L->gm_V) (void)hipFree(L->gm_V);
        L->gm_V = nullptr;
        if ((rc = ensure_vec(&L->gm_V, (int64_t)(m + 1) * L->nvec))) return rc;
        L->gm_m = m;
      }
      rc = gmres<double>(ctx, b, x, m, o->rtol, o->max_it, iters, relres, L->gm_V, L->w, L->z);
    } else {
      double** vs[] = {&L->r, &L->rh, &L->p, &L->v, &L->s, &L->t, &L->ph, &L->sh};
      for (double** q : vs)
        if ((rc = ensure_vec(q, L->nvec))) return rc;
      rc = bicgstab<double>(ctx, b, x, o->rtol, o->max_it, iters, relres, L->r, L->rh, L->p, L->v, L->s,
                            L->t, L->ph, L->sh);
    }
    // FP32-stored operator: the solve is one step of a defect correction, as in mixed mode
    if (L->pc_store32 && rc == SFEM_ERR_MAXIT && std::isfinite(*relres) && *relres < 1.0) {
      ++L->partial;
      rc = SFEM_OK;
    }
    if (padmap) {
      hipLaunchKernelGGL((k_pad_out<double, double>), dim3(gk), dim3(kTPB), 0, st, nk, padmap, x, ctx->d_du, 1.0);
      if (hipGetLastError() != hipSuccess) return set_error(SFEM_ERR_HIP, "permute launch");
    }
    return rc;
  }
  // mixed precision
  float** vs[] = {&L->r32, &L->rh32, &L->p32, &L->v32, &L->s32, &L->t32, &L->ph32, &L->sh32, &L->x32, &L->du32};
  for (float** q : vs)
    if ((rc = ensure_vec(q, L->nvec))) return rc;
  double bn;
  if ((rc = la_norm2(ctx, ctx->d_rhs, &bn))) return rc;
  if (!(bn > 0.0)) {
    hipLaunchKernelGGL((k_fill<double>), dim3(gn), dim3(kTPB), 0, st, n, ctx->d_du, 0.0);
    *iters = 0; *relres = 0.0;
    return SFEM_OK;
  }
  // b32 = rhs / ||rhs|| (x32 holds it; the solver overwrites its other vectors)
  float* b32 = L->x32;
  if (padmap)
    hipLaunchKernelGGL((k_pad_in<double, float>), dim3(gk), dim3(kTPB), 0, st, nk, padmap, ctx->d_rhs, b32, 1.0 / bn);
  else
    hipLaunchKernelGGL((k_convert<double, float>), dim3(gn), dim3(kTPB), 0, st, n, ctx->d_rhs, b32, 1.0 / bn);
  if (o->ksp_type == SFEM_KSP_GMRES) {
    if ((rc = ensure_vec(&L->w32, L->nvec)) || (rc = ensure_vec(&L->z32, L->nvec))) return rc;
    if (!L->gm_V32 || L->gm_m32 < m) {
      if (L->gm_V32) (void)hipFree(L->gm_V32);
      L->gm_V32 = nullptr;
      if ((rc = ensure_vec(&L->gm_V32, (int64_t)(m + 1) * L->nvec))) return rc;
      L->gm_m32 = m;
    }
    rc = gmres<float>(ctx, b32, L->du32, m, o->rtol, o->max_it, iters, relres, L->gm_V32, L->w32, L->z32);
  } else {
    // FP32 BiCGStab can stall where FP64 BiCGStab converges (a multicolour-ordered ILU(0)
    // on small, strongly graded meshes makes the short recurrences sensitive to FP32
    // rounding): after a stall window the solve restarts from zero as FP32 GMRES.
    rc = bicgstab<float>(ctx, b32, L->du32, o->rtol, o->max_it, iters, relres, L->r32, L->rh32, L->p32,
                         L->v32, L->s32, L->t32, L->ph32, L->sh32, kBicgStall);
    if (rc == kStagnated) {
      const int used = *iters;
      const int mg = std::max(1, std::min(o->restart > 0 ? o->restart : 30, 64));
      if ((rc = ensure_vec(&L->w32, L->nvec)) || (rc = ensure_vec(&L->z32, L->nvec))) return rc;
      if (!L->gm_V32 || L->gm_m32 < mg) {
        if (L->gm_V32) (void)hipFree(L->gm_V32);
        L->gm_V32 = nullptr;
        if ((rc = ensure_vec(&L->gm_V32, (int64_t)(mg + 1) * L->nvec))) return rc;
        L->gm_m32 = mg;
      }
      ++L->fallbacks;
      rc = gmres<float>(ctx, b32, L->du32, mg, o->rtol, std::max(1, o->max_it - used), iters, relres, L->gm_V32,
                        L->w32, L->z32);
      *iters += used;
    }
  }
  // Mixed precision is defect correction: an FP32 solve that reached max_it with a
  // reduced, finite residual still moves the FP64 Newton iterate toward the solution
  // (whose own test, ||F|| < tol in FP64, decides convergence), so its iterate is used.
  if (rc == SFEM_ERR_MAXIT && std::isfinite(*relres) && *relres < 1.0) {
    ++L->partial;
    rc = SFEM_OK;
  }
  if (padmap)
    hipLaunchKernelGGL((k_pad_out<float, double>), dim3(gk), dim3(kTPB), 0, st, nk, padmap, L->du32, ctx->d_du, bn);
  else
    hipLaunchKernelGGL((k_convert<float, double>), dim3(gn), dim3(kTPB), 0, st, n, L->du32, ctx->d_du, bn);
  if (hipGetLastError() != hipSuccess) return set_error(SFEM_ERR_HIP, "convert launch");
  return rc;
}

}  // namespace sfem
