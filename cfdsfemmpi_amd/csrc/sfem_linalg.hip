// sfem_linalg.hip -- Krylov solvers and preconditioners on the device CSR.
//
// Replaces the PETSc KSP/PC calls behind SolverPetsc::solve
// (src/base/SolverPetsc.cpp:224-277): BiCGStab (-ksp_type bcgs) and
// restarted GMRES (-ksp_type gmres), right-preconditioned so that the stopping
// test ||b - A x||_2 <= rtol ||b||_2 is on the true residual, with a zero
// initial guess (VecZeroEntries at :249).  Preconditioners: point Jacobi,
// nodal block Jacobi (the free DOFs of one node form a block of size <= 4),
// and ILU(0) on the CSR pattern, factorised and applied by level scheduling
// over nodes (all free rows of a node are handled by one wavefront).
//
// Reductions are two-stage (per-block partials, then one block summing the
// partials in index order), so every dot product is deterministic.  Krylov
// scalars (alpha, omega, beta, ...) stay on the device; the host reads back
// one small block of scalars per iteration for the convergence test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
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

__global__ __launch_bounds__(256) void k_to_bf16(int64_t n, const float* __restrict__ a, bf16* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const uint32_t u = __float_as_uint(a[i]);
    b[i].bits = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
  }
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
enum { FIN_STORE = 0, FIN_RHO1, FIN_ALPHA, FIN_SS, FIN_OMEGA, FIN_RR };

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

// x += alpha ph + omega sh ; r = s - omega t ; partial (r, r)
template <typename T>
__global__ __launch_bounds__(kTPB) void k_bicg_xr(int64_t n, T* __restrict__ x, const T* __restrict__ ph,
                                                  const T* __restrict__ sh, T* __restrict__ r,
                                                  const T* __restrict__ s, const T* __restrict__ t,
                                                  const double* scal, double* part) {
  const T alpha = (T)scal[S_ALPHA], om = (T)scal[S_OMEGA];
  double acc[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB) {
    x[i] += alpha * ph[i] + om * sh[i];
    const T ri = s[i] - om * t[i];
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
  }
  block_sum_store<1>(acc, part);
}

// dst[i] = src[map[i]]
template <typename T>
__global__ __launch_bounds__(kTPB) void k_gather(int64_t n, const int32_t* __restrict__ map,
                                                 const T* __restrict__ src, T* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    dst[i] = src[map[i]];
}

// dst0[i] = dst1[i] = src[map[i]]  (the permuted tangent and the factor's initial values)
template <typename T>
__global__ __launch_bounds__(kTPB) void k_gather2(int64_t n, const int32_t* __restrict__ map,
                                                  const T* __restrict__ src, T* __restrict__ dst0,
                                                  T* __restrict__ dst1) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB) {
    const T v = src[map[i]];
    dst0[i] = v;
    dst1[i] = v;
  }
}

// b[i] = scale * a[perm[i]]  (into the solver ordering)
template <typename TA, typename TB>
__global__ __launch_bounds__(kTPB) void k_permute_in(int64_t n, const int32_t* __restrict__ perm,
                                                     const TA* __restrict__ a, TB* __restrict__ b, double scale) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    b[i] = (TB)((double)a[perm[i]] * scale);
}

// b[perm[i]] = scale * a[i]  (back to the DOF ordering)
template <typename TA, typename TB>
__global__ __launch_bounds__(kTPB) void k_permute_out(int64_t n, const int32_t* __restrict__ perm,
                                                      const TA* __restrict__ a, TB* __restrict__ b, double scale) {
  for (int64_t i = (int64_t)blockIdx.x * kTPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTPB)
    b[perm[i]] = (TB)((double)a[i] * scale);
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
// ILU(0), level-scheduled over nodal blocks; one 64-lane workgroup per block.
// All rows of a block share one column list (the pattern is the union of full
// node blocks), so the block's rows are staged in LDS against that list.
// ---------------------------------------------------------------------------
constexpr int kIluMaxRow = 256;  // longest row handled in LDS (checked on the host)

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

// ILU(0) factorisation of the blocks of one level (one 64-lane workgroup per block).
// IKJ order per row: external lower columns (earlier blocks, already final) ascending,
// then the block's own lower columns.  The m rows of the block share the column list, so
// the position of every upper entry of an external row is looked up once for all rows.
template <typename LT>
__global__ __launch_bounds__(64) void k_ilu_factor_blk(const int32_t* __restrict__ list, int first, int count,
                                                       const int32_t* __restrict__ blkStart,
                                                       const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ colidx,
                                                       const int32_t* __restrict__ diagpos,
                                                       LT* __restrict__ lu, double* scal) {
  __shared__ int32_t cols[kIluMaxRow];
  __shared__ double rows[4][kIluMaxRow];
  __shared__ double lval[4];
  const int w = blockIdx.x;
  if (w >= count) return;
  const int lane = threadIdx.x;
  const int32_t blk = list ? list[first + w] : first + w;
  const int32_t r0 = blkStart[blk];
  const int m = blkStart[blk + 1] - r0;
  const int32_t base = rowptr[r0];
  const int rl = rowptr[r0 + 1] - base;
  const int dq0 = diagpos[r0] - base;
  for (int q = lane; q < rl; q += 64) cols[q] = colidx[base + q];
  for (int a = 0; a < m; ++a)
    for (int q = lane; q < rl; q += 64) rows[a][q] = (double)lu[base + a * rl + q];
  __syncthreads();
  for (int q = 0; q < dq0; ++q) {
    const int32_t c = cols[q];
    if (lane < m) {
      const double d = (double)lu[diagpos[c]];
      if (d == 0.0) scal[S_MISC1] = 2.0;
      const double l = rows[lane][q] / (d == 0.0 ? 1.0 : d);
      rows[lane][q] = l;
      lval[lane] = l;
    }
    __syncthreads();
    const int32_t cb = diagpos[c] + 1, ce = rowptr[c + 1];
    for (int32_t j = cb + lane; j < ce; j += 64) {
      const int pos = lds_find(cols, rl, colidx[j]);
      if (pos >= 0) {
        const double u = (double)lu[j];
        for (int a = 0; a < m; ++a) rows[a][pos] -= lval[a] * u;
      }
    }
    __syncthreads();
  }
  for (int a = 1; a < m; ++a) {
    for (int ap = 0; ap < a; ++ap) {
      if (lane == 0) {
        const double d = rows[ap][dq0 + ap];
        if (d == 0.0) scal[S_MISC1] = 2.0;
        const double l = rows[a][dq0 + ap] / (d == 0.0 ? 1.0 : d);
        rows[a][dq0 + ap] = l;
        lval[0] = l;
      }
      __syncthreads();
      const double l = lval[0];
      for (int q = dq0 + ap + 1 + lane; q < rl; q += 64) rows[a][q] -= l * rows[ap][q];
      __syncthreads();
    }
  }
  for (int a = 0; a < m; ++a)
    for (int q = lane; q < rl; q += 64) lu[base + a * rl + q] = (LT)rows[a][q];
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

// ILU(0) factorisation of the node blocks of one level, one 64-lane workgroup (one
// wavefront) per block, organised by lower node blocks: the columns of row block i fall
// into groups, one per neighbouring node block, and every row of a lower block c shares
// c's column list.  So per lower block the positions (in i's list) of c's upper columns
// are searched once, c's m_c upper rows are staged in LDS, and the m_c eliminations of
// IKJ order run from LDS; the global loads of the next lower block are issued before the
// eliminations of the current one and land in registers meanwhile.  The result equals
// the row-by-row IKJ factorisation (same updates in the same order).  Dynamic LDS:
// factor_nb_lds(cap) bytes for rows of at most cap (a multiple of 64, <= kIluMaxRow)
// entries.
template <typename LT>
__host__ __device__ constexpr size_t factor_nb_lds(int cap) {
  return (size_t)cap * (4 * sizeof(double) + 4 * sizeof(LT) + 3 * sizeof(int32_t) + sizeof(BlkDesc));
}

template <typename LT>
__global__ __launch_bounds__(64) void k_ilu_factor_nb(const int32_t* __restrict__ list, int first, int count,
                                                      const BlkDesc* __restrict__ desc,
                                                      const int32_t* __restrict__ colidx,
                                                      const int32_t* __restrict__ rowBlk, LT* __restrict__ lu,
                                                      double* __restrict__ scal, int cap) {
  extern __shared__ double smem[];
  double* rows = smem;                              // [4][cap], this block's rows
  LT* urow = (LT*)(rows + 4 * cap);                 // [4][cap], external upper part of c's rows
  int32_t* cols = (int32_t*)(urow + 4 * cap);       // [cap], this block's column list
  int32_t* pos = cols + cap;                        // [cap], positions of c's upper columns
  int32_t* gq = pos + cap;                          // [cap], first column of each lower group
  BlkDesc* gdesc = (BlkDesc*)(gq + cap);            // [cap], descriptor of each lower group
  __shared__ double ub[16];                         // c's diagonal block
  __shared__ double lval[4];
  __shared__ int ngroups;
  constexpr int S = kIluMaxRow / 64;                // prefetch slots per lane
  const int w = blockIdx.x;
  if (w >= count) return;
  const int lane = threadIdx.x;
  const BlkDesc d = desc[list ? list[first + w] : first + w];
  const int m = d.dqm >> 16, dq0 = d.dqm & 0xffff, rl = d.rl;
  for (int q = lane; q < rl; q += 64) cols[q] = colidx[d.base + q];
  for (int a = 0; a < m; ++a)
    for (int q = lane; q < rl; q += 64) rows[a * cap + q] = (double)lu[d.base + (int64_t)a * rl + q];
  // lower node groups: block of every lower column, group starts compacted by ballot
  for (int q = lane; q < dq0; q += 64) pos[q] = rowBlk[colidx[d.base + q]];
  if (lane == 0) ngroups = 0;
  wave_sync();
  for (int q0 = 0; q0 < dq0; q0 += 64) {
    const int q = q0 + lane;
    const bool start = q < dq0 && (q == 0 || pos[q] != pos[q - 1]);
    const uint64_t mask = __ballot(start);
    const int base = ngroups;
    if (start) {
      const int k = base + __popcll(mask & ((1ull << lane) - 1ull));
      gq[k] = q;
      gdesc[k] = desc[pos[q]];
    }
    wave_sync();
    if (lane == 0) ngroups = base + __popcll(mask);
    wave_sync();
  }
  const int ng = ngroups;
  // registers of the prefetched lower block: column indices, upper values, diagonal block
  int32_t pc[S];
  LT pu[S][4];
  LT pd = (LT)0;
  auto issue = [&](int g) {
    const BlkDesc dc = gdesc[g];
    const int mc = dc.dqm >> 16, dq0c = dc.dqm & 0xffff;
    const int ext0 = dq0c + mc, next = dc.rl - ext0;
#pragma unroll
    for (int sl = 0; sl < S; ++sl) {
      const int t = lane + 64 * sl;
      if (t < next) {
        pc[sl] = colidx[dc.base + ext0 + t];
#pragma unroll
        for (int b = 0; b < 4; ++b)
          if (b < mc) pu[sl][b] = lu[dc.base + (int64_t)b * dc.rl + ext0 + t];
      }
    }
    const int b = lane >> 2, bp = lane & 3;
    if (lane < 16 && b < mc && bp < mc) pd = lu[dc.base + (int64_t)b * dc.rl + dq0c + bp];
  };
  if (ng > 0) issue(0);
  for (int g = 0; g < ng; ++g) {
    const int q = gq[g];
    const BlkDesc dc = gdesc[g];
    const int mc = dc.dqm >> 16, dq0c = dc.dqm & 0xffff;
    const int next = dc.rl - (dq0c + mc);
#pragma unroll
    for (int sl = 0; sl < S; ++sl) {
      const int t = lane + 64 * sl;
      if (t < next) {
        pos[t] = lds_find(cols, rl, pc[sl]);
#pragma unroll
        for (int b = 0; b < 4; ++b)
          if (b < mc) urow[b * cap + t] = pu[sl][b];
      }
    }
    if (lane < 16) ub[lane] = ((lane >> 2) < mc && (lane & 3) < mc) ? (double)pd : 0.0;
    wave_sync();
    if (g + 1 < ng) issue(g + 1);
    for (int b = 0; b < mc; ++b) {
      if (lane < m) {
        const double dd = ub[b * 4 + b];
        if (dd == 0.0) scal[S_MISC1] = 2.0;
        const double l = rows[lane * cap + q + b] / (dd == 0.0 ? 1.0 : dd);
        rows[lane * cap + q + b] = l;
        lval[lane] = l;
      }
      wave_sync();
      const int nin = mc - 1 - b;  // remaining columns of block c in this row block
      for (int t = lane; t < nin + next; t += 64) {
        int p;
        double u;
        if (t < nin) {
          p = q + b + 1 + t;
          u = ub[b * 4 + b + 1 + t];
        } else {
          p = pos[t - nin];
          u = (double)urow[b * cap + t - nin];
        }
        if (p >= 0)
          for (int a = 0; a < m; ++a) rows[a * cap + p] -= lval[a] * u;
      }
      wave_sync();
    }
  }
  // the block's own diagonal block
  for (int a = 1; a < m; ++a) {
    for (int ap = 0; ap < a; ++ap) {
      if (lane == 0) {
        const double dd = rows[ap * cap + dq0 + ap];
        if (dd == 0.0) scal[S_MISC1] = 2.0;
        const double l = rows[a * cap + dq0 + ap] / (dd == 0.0 ? 1.0 : dd);
        rows[a * cap + dq0 + ap] = l;
        lval[0] = l;
      }
      wave_sync();
      const double l = lval[0];
      for (int qq = dq0 + ap + 1 + lane; qq < rl; qq += 64) rows[a * cap + qq] -= l * rows[ap * cap + qq];
      wave_sync();
    }
  }
  for (int a = 0; a < m; ++a)
    for (int qq = lane; qq < rl; qq += 64) lu[d.base + (int64_t)a * rl + qq] = (LT)rows[a * cap + qq];
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

template <typename LT> struct AccOf { using type = float; };
template <> struct AccOf<double> { using type = double; };

// The m x m diagonal block of a node block: entry e = 4a + b (row a, column b) is loaded
// by lane e % G as its (e / G)-th value (load_diag, issued before the long sparse dot
// product) and then broadcast to the whole group (bcast_diag); absent entries read 0.
template <int G> struct DiagPerLane { static constexpr int n = G >= 16 ? 1 : 16 / G; };

template <int G, typename AT, typename LT>
__device__ inline void load_diag(const LT* __restrict__ lu, const BlkDesc& d, int lane,
                                 AT (&dv)[DiagPerLane<G>::n]) {
  const int m = d.dqm >> 16, dq0 = d.dqm & 0xffff;
#pragma unroll
  for (int t = 0; t < DiagPerLane<G>::n; ++t) {
    const int e = lane + t * G;
    dv[t] = (e < 16 && (e >> 2) < m && (e & 3) < m)
                ? widen<AT>(lu[d.base + (int64_t)(e >> 2) * d.rl + dq0 + (e & 3)])
                : (AT)0;
  }
}

template <int G, typename AT>
__device__ inline void bcast_diag(const AT (&dv)[DiagPerLane<G>::n], AT (&D)[16]) {
#pragma unroll
  for (int e = 0; e < 16; ++e) D[e] = __shfl(dv[e / G], e % G, G);
}

// Forward solve L y = r over one level with G >= 4 lanes per node block: the external
// lower part as a grouped sparse dot product, then the unit-lower m x m diagonal block.
// Level blocks are either the contiguous range [first, first + count) or list[first + g].
template <typename T, typename LT, int G, int U>
__global__ __launch_bounds__(256) void k_ilu_lsolve_blk(const int32_t* __restrict__ list, int first, int count,
                                                        const BlkDesc* __restrict__ desc,
                                                        const int32_t* __restrict__ colidx,
                                                        const LT* __restrict__ lu,
                                                        const T* __restrict__ r, T* __restrict__ y, int xcd) {
  using AT = typename AccOf<LT>::type;
  const int64_t gt = logical_wg(xcd) * 256 + threadIdx.x;
  const int64_t g = gt / G;
  const int lane = (int)(gt % G);
  AT acc[4] = {0, 0, 0, 0};
  BlkDesc d = {0, 0, 0, 0};
  AT rv = 0;
  AT dv[DiagPerLane<G>::n] = {};
  const bool act = g < count;
  if (act) {
    d = desc[list ? list[first + g] : first + (int32_t)g];
    const int m = d.dqm >> 16, dq0 = d.dqm & 0xffff;
    if (lane < m) rv = (AT)r[d.r0 + lane];
    load_diag<G>(lu, d, lane, dv);
    group_dot_v<G, U>(colidx + d.base, lu + d.base, y, d.rl, m, 0, dq0, lane, acc);
  }
  AT D[16], R[4];
  bcast_diag<G>(dv, D);
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = group_sum<G>(acc[a]);
#pragma unroll
  for (int a = 0; a < 4; ++a) R[a] = __shfl(rv, a, G);
  const int m = d.dqm >> 16;
  AT v[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    AT s = R[a] - acc[a];
#pragma unroll
    for (int ap = 0; ap < a; ++ap) s -= D[a * 4 + ap] * v[ap];
    v[a] = s;
  }
  if (act && lane < m) y[d.r0 + lane] = (T)pick4(v, lane);
}

// Backward solve U z = y over one level (levels in reverse order), in place on y.
template <typename T, typename LT, int G, int U>
__global__ __launch_bounds__(256) void k_ilu_usolve_blk(const int32_t* __restrict__ list, int first, int count,
                                                        const BlkDesc* __restrict__ desc,
                                                        const int32_t* __restrict__ colidx,
                                                        const LT* __restrict__ lu, T* __restrict__ y, int xcd) {
  using AT = typename AccOf<LT>::type;
  const int64_t gt = logical_wg(xcd) * 256 + threadIdx.x;
  const int64_t g = gt / G;
  const int lane = (int)(gt % G);
  AT acc[4] = {0, 0, 0, 0};
  BlkDesc d = {0, 0, 0, 0};
  AT rv = 0;
  AT dv[DiagPerLane<G>::n] = {};
  const bool act = g < count;
  if (act) {
    d = desc[list ? list[first + g] : first + (int32_t)g];
    const int m = d.dqm >> 16, dq0 = d.dqm & 0xffff;
    if (lane < m) rv = (AT)y[d.r0 + lane];
    load_diag<G>(lu, d, lane, dv);
    group_dot_v<G, U>(colidx + d.base, lu + d.base, y, d.rl, m, dq0 + m, d.rl, lane, acc);
  }
  AT D[16], R[4];
  bcast_diag<G>(dv, D);
#pragma unroll
  for (int a = 0; a < 4; ++a) acc[a] = group_sum<G>(acc[a]);
#pragma unroll
  for (int a = 0; a < 4; ++a) R[a] = __shfl(rv, a, G);
  const int m = d.dqm >> 16;
  AT v[4] = {0, 0, 0, 0};
#pragma unroll
  for (int a = 3; a >= 0; --a) {
    if (a < m) {
      AT s = R[a] - acc[a];
#pragma unroll
      for (int ap = a + 1; ap < 4; ++ap)
        if (ap < m) s -= D[a * 4 + ap] * v[ap];
      v[a] = s / D[a * 4 + a];
    }
  }
  if (act && lane < m) y[d.r0 + lane] = (T)pick4(v, lane);
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
  void* ptrs[] = {d_blkStart, d_diagpos, d_part, d_scal, r, rh, p, v, s, t, ph, sh, x, w, z, pr, pz,
                  r32, rh32, p32, v32, s32, t32, ph32, sh32, x32, du32, w32, z32, pr32, pz32,
                  dinv, dinv32, binv, binv32, lu, lu32, gm_V, gm_V32, d_desc, lu16, kp, kp32, d_rowBlk,
                  d_h, d_mpart};
  delete ilu;
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
    LA_CHECK(alloc(&L->d_rowBlk, rowBlk.size()));
    if (n) LA_CHECK(hipMemcpy(L->d_rowBlk, rowBlk.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
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

// Level of a block = 1 + max level of the blocks it depends on (lower blocks for the
// forward solve and the factorisation, upper blocks for the backward solve).
static void build_levels(const std::vector<int64_t>& rowptr, const std::vector<int32_t>& colidx,
                         const std::vector<int32_t>& bs, IluPlan& P) {
  const int64_t nb = (int64_t)bs.size() - 1;
  std::vector<int32_t> rowBlk(bs.back());
  for (int64_t k = 0; k < nb; ++k)
    for (int32_t r = bs[k]; r < bs[k + 1]; ++r) rowBlk[r] = (int32_t)k;
  std::vector<int32_t> lvL(nb, 0), lvU(nb, 0);
  for (int64_t k = 0; k < nb; ++k) {
    int lv = 0;
    for (int64_t q = rowptr[bs[k]]; q < rowptr[bs[k] + 1]; ++q) {
      const int32_t bk = rowBlk[colidx[q]];
      if (bk < k) lv = std::max(lv, lvL[bk] + 1);
    }
    lvL[k] = lv;
  }
  for (int64_t k = nb - 1; k >= 0; --k) {
    int lv = 0;
    for (int64_t q = rowptr[bs[k]]; q < rowptr[bs[k] + 1]; ++q) {
      const int32_t bk = rowBlk[colidx[q]];
      if (bk > k) lv = std::max(lv, lvU[bk] + 1);
    }
    lvU[k] = lv;
  }
  bucket_levels(lvL, P.levLPtr, P.h_levL);
  bucket_levels(lvU, P.levUPtr, P.h_levU);
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
  void* ptrs[] = {d_levL, d_levU, d_rowptr, d_colidx, d_diagpos, d_blkStart, d_pmap, d_perm, d_desc, d_rowBlk};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
}

static int build_ilu_plan(sfem_ctx* ctx, LinAlg* L, int reorder) {
  delete L->ilu;
  L->ilu = new IluPlan;
  ++L->plan_id;
  IluPlan& P = *L->ilu;
  P.reorder = reorder;
  P.identity = (reorder == 0);
  const std::vector<int32_t>& bs = L->h_blkStart;
  const int64_t nb = L->nblk;
  if (P.identity) {
    build_levels(ctx->rowptr, ctx->colidx, bs, P);
    P.d_desc = nullptr;  // the context's descriptors are used
    P.rowptr = ctx->d_rowptr32;
    P.colidx = ctx->d_colidx;
    P.diagpos = L->d_diagpos;
    P.blkStart = L->d_blkStart;
    P.rowBlk = L->d_rowBlk;
  } else {
    // block graph from the first row of each block
    std::vector<int32_t> rowBlk(L->n);
    for (int64_t k = 0; k < nb; ++k)
      for (int32_t r = bs[k]; r < bs[k + 1]; ++r) rowBlk[r] = (int32_t)k;
    std::vector<int64_t> adjPtr(nb + 1, 0);
    std::vector<int32_t> adj;
    for (int64_t k = 0; k < nb; ++k) {
      int32_t prev = -1;
      for (int64_t q = ctx->rowptr[bs[k]]; q < ctx->rowptr[bs[k] + 1]; ++q) {
        const int32_t b = rowBlk[ctx->colidx[q]];
        if (b != prev && b != k) adj.push_back(b);
        prev = b;
      }
      adjPtr[k + 1] = (int64_t)adj.size();
    }
    std::vector<int32_t> colorPtr;
    const std::vector<int32_t> order = reorder == 2 ? multicolor_order(adjPtr, adj, &colorPtr) : rcm_order(adjPtr, adj);
    std::vector<int32_t> nbs(nb + 1, 0), perm(L->n), iperm(L->n);
    for (int64_t p = 0; p < nb; ++p) {
      const int32_t k = order[p];
      nbs[p + 1] = nbs[p] + (bs[k + 1] - bs[k]);
      for (int32_t a = 0; a < bs[k + 1] - bs[k]; ++a) perm[nbs[p] + a] = bs[k] + a;
    }
    for (int64_t i = 0; i < L->n; ++i) iperm[perm[i]] = (int32_t)i;
    // permuted rows: every row of a block shares its column list, so sort once per block
    std::vector<int64_t> rp(L->n + 1, 0);
    for (int64_t p = 0; p < nb; ++p) {
      const int32_t k = order[p];
      const int64_t len = ctx->rowptr[bs[k] + 1] - ctx->rowptr[bs[k]];
      for (int32_t r = nbs[p]; r < nbs[p + 1]; ++r) rp[r + 1] = rp[r] + len;
    }
    std::vector<int32_t> ci(ctx->nnz), pmap(ctx->nnz), dp(L->n, -1);
    parallel_for(nb, [&](int64_t lo, int64_t hi) {
      std::vector<std::pair<int32_t, int32_t>> tmp;
      for (int64_t p = lo; p < hi; ++p) {
        const int32_t k = order[p];
        const int64_t b0 = ctx->rowptr[bs[k]];
        const int64_t len = ctx->rowptr[bs[k] + 1] - b0;
        tmp.resize(len);
        for (int64_t q = 0; q < len; ++q) tmp[q] = {iperm[ctx->colidx[b0 + q]], (int32_t)q};
        std::sort(tmp.begin(), tmp.end());
        for (int32_t a = 0; a < bs[k + 1] - bs[k]; ++a) {
          const int32_t rn = nbs[p] + a;
          const int64_t ob = ctx->rowptr[bs[k] + a];
          for (int64_t q = 0; q < len; ++q) {
            ci[rp[rn] + q] = tmp[q].first;
            pmap[rp[rn] + q] = (int32_t)(ob + tmp[q].second);
            if (tmp[q].first == rn) dp[rn] = (int32_t)(rp[rn] + q);
          }
        }
      }
    });
    if (reorder == 2) {
      // one level per colour: blocks of a colour are independent and contiguous in the
      // new order; the backward solve walks the colours in reverse
      P.contiguous = true;
      P.levLPtr = colorPtr;
      P.levUPtr = colorPtr;
      P.h_levL.clear();
      P.h_levU.clear();
    } else {
      build_levels(rp, ci, nbs, P);
    }
    {
      std::vector<BlkDesc> dd = make_desc(nbs, rp, dp);
      hipError_t e = alloc(&P.d_desc, dd.size());
      if (e == hipSuccess) e = hipMemcpy(P.d_desc, dd.data(), dd.size() * sizeof(BlkDesc), hipMemcpyHostToDevice);
      if (e != hipSuccess) return set_hip_error(e, "ILU descriptors");
    }
    std::vector<int32_t> rp32(rp.begin(), rp.end());
    std::vector<int32_t> rb(L->n);
    for (int64_t p = 0; p < nb; ++p)
      for (int32_t r = nbs[p]; r < nbs[p + 1]; ++r) rb[r] = (int32_t)p;
    int rc;
    if ((rc = upload(&P.d_rowptr, rp32)) || (rc = upload(&P.d_colidx, ci)) || (rc = upload(&P.d_diagpos, dp)) ||
        (rc = upload(&P.d_blkStart, nbs)) || (rc = upload(&P.d_pmap, pmap)) || (rc = upload(&P.d_perm, perm)) ||
        (rc = upload(&P.d_rowBlk, rb)))
      return rc;
    P.rowBlk = P.d_rowBlk;
    P.rowptr = P.d_rowptr;
    P.colidx = P.d_colidx;
    P.diagpos = P.d_diagpos;
    P.blkStart = P.d_blkStart;
  }
  int rc;
  if ((rc = upload(&P.d_levL, P.h_levL)) || (rc = upload(&P.d_levU, P.h_levU))) return rc;
  return SFEM_OK;
}

template <typename T>
static int ensure_vec(T** p, int64_t n) {
  if (*p) return SFEM_OK;
  hipError_t e = alloc(p, (size_t)n);
  if (e != hipSuccess) return set_hip_error(e, "hipMalloc krylov vector");
  return SFEM_OK;
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
int la_setup_pc(sfem_ctx* ctx, int pc_type, bool fp32, int reorder, bool bf16f) {
  LinAlg* L = ctx->la;
  hipStream_t st = ctx->stream;
  const int64_t n = L->n;
  L->pc_type = pc_type;
  L->pc_fp32 = fp32;
  const double* vals = fp32 ? nullptr : ctx->d_vals;
  const float* vals32 = fp32 ? ctx->d_vals32 : nullptr;
  LA_CHECK(hipMemsetAsync(L->d_scal + S_MISC1, 0, sizeof(double), st));
  if (pc_type == SFEM_PC_JACOBI) {
    for (int64_t r = 0; r < n; ++r)
      if (L->h_diagpos[r] < 0) return set_error(SFEM_ERR_ARG, "Jacobi: missing diagonal entry");
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
    if (L->maxrow > kIluMaxRow) return set_error(SFEM_ERR_UNSUPPORTED, "ILU(0): row longer than 256 entries");
    for (int64_t r = 0; r < n; ++r)
      if (L->h_diagpos[r] < 0) return set_error(SFEM_ERR_ARG, "ILU(0): missing diagonal entry");
    if (!L->ilu || L->ilu->reorder != reorder) {
      int rc = build_ilu_plan(ctx, L, reorder);
      if (rc) return rc;
    }
    const IluPlan& P = *L->ilu;
    const int64_t nnz = ctx->nnz;
    const unsigned gz = (unsigned)std::min<int64_t>((nnz + kTPB - 1) / kTPB, 1 << 20);
    // reordered plan: one gather gives the permuted tangent (for the products of the
    // Krylov solve) and the factor's starting values
    if (fp32) {
      int rc = ensure_vec(&L->lu32, nnz); if (rc) return rc;
      if (P.identity) {
        LA_CHECK(hipMemcpyAsync(L->lu32, ctx->d_vals32, nnz * sizeof(float), hipMemcpyDeviceToDevice, st));
      } else {
        if ((rc = ensure_vec(&L->kp32, nnz))) return rc;
        hipLaunchKernelGGL((k_gather2<float>), dim3(gz), dim3(kTPB), 0, st, nnz, P.d_pmap, ctx->d_vals32,
                           L->kp32, L->lu32);
      }
    } else {
      int rc = ensure_vec(&L->lu, nnz); if (rc) return rc;
      if (P.identity) {
        LA_CHECK(hipMemcpyAsync(L->lu, ctx->d_vals, nnz * sizeof(double), hipMemcpyDeviceToDevice, st));
      } else {
        if ((rc = ensure_vec(&L->kp, nnz))) return rc;
        hipLaunchKernelGGL((k_gather2<double>), dim3(gz), dim3(kTPB), 0, st, nnz, P.d_pmap, ctx->d_vals,
                           L->kp, L->lu);
      }
    }
    const int nl = (int)P.levLPtr.size() - 1;
    for (int lv = 0; lv < nl; ++lv) {
      const int32_t* list; int first, cnt;
      level_range(P, false, lv, &list, &first, &cnt);
      if (!cnt) continue;
      const BlkDesc* desc = P.identity ? L->d_desc : P.d_desc;
      if (L->full_blocks) {
        const int cap = std::max(64, (L->maxrow + 63) / 64 * 64);
        if (fp32)
          hipLaunchKernelGGL((k_ilu_factor_nb<float>), dim3(cnt), dim3(64), factor_nb_lds<float>(cap), st, list,
                             first, cnt, desc, P.colidx, P.rowBlk, L->lu32, L->d_scal, cap);
        else
          hipLaunchKernelGGL((k_ilu_factor_nb<double>), dim3(cnt), dim3(64), factor_nb_lds<double>(cap), st,
                             list, first, cnt, desc, P.colidx, P.rowBlk, L->lu, L->d_scal, cap);
      } else if (fp32)
        hipLaunchKernelGGL((k_ilu_factor_blk<float>), dim3(cnt), dim3(64), 0, st, list, first, cnt,
                           P.blkStart, P.rowptr, P.colidx, P.diagpos, L->lu32, L->d_scal);
      else
        hipLaunchKernelGGL((k_ilu_factor_blk<double>), dim3(cnt), dim3(64), 0, st, list, first, cnt,
                           P.blkStart, P.rowptr, P.colidx, P.diagpos, L->lu, L->d_scal);
    }
    L->pc_bf16 = fp32 && bf16f;
    if (L->pc_bf16) {
      int rc = ensure_vec(&L->lu16, nnz); if (rc) return rc;
      hipLaunchKernelGGL(k_to_bf16, dim3(gz), dim3(256), 0, st, nnz, L->lu32, L->lu16);
    }
    LA_CHECK(hipGetLastError());
  }
  return SFEM_OK;
}

template <typename T>
static int pc_apply_impl(sfem_ctx* ctx, const T* r, T* z);

template <typename T>
static int pc_apply(sfem_ctx* ctx, const T* r, T* z) {
  ctx->kt.begin(SFEM_KT_PC, ctx->stream);
  const int rc = pc_apply_impl<T>(ctx, r, z);
  ctx->kt.end(SFEM_KT_PC, ctx->stream);
  return rc;
}

// Triangular-solve kernel shapes: code = vi | (xcd << 5), with G = kTriG[vi] lanes per
// node block and U = kTriU[vi] columns in flight per lane (0: all of a lane's columns,
// up to 128 / G, loaded at once).
static const int kTriG[15] = {4, 4, 4, 8, 8, 8, 8, 16, 16, 16, 16, 32, 32, 32, 32};
static const int kTriU[15] = {1, 2, 4, 1, 2, 4, 0, 1, 2, 4, 0, 1, 2, 4, 0};
constexpr int kTriCodes = 15;
constexpr int kTriDefault = 8;  // G = 16, U = 2, no XCD remap

template <typename T, typename LT, bool UPPER, int G, int U>
static void tri_launch_g(hipStream_t st, int xcd, const int32_t* list, int first, int cnt, const BlkDesc* desc,
                         const int32_t* col, const LT* lu, const T* r, T* y) {
  const unsigned gr = (unsigned)(((int64_t)cnt * G + 255) / 256);
  if constexpr (UPPER)
    hipLaunchKernelGGL((k_ilu_usolve_blk<T, LT, G, U>), dim3(gr), dim3(256), 0, st, list, first, cnt, desc, col,
                       lu, y, xcd);
  else
    hipLaunchKernelGGL((k_ilu_lsolve_blk<T, LT, G, U>), dim3(gr), dim3(256), 0, st, list, first, cnt, desc, col,
                       lu, r, y, xcd);
}

template <typename T, typename LT, bool UPPER>
static void tri_launch(hipStream_t st, int code, const int32_t* list, int first, int cnt, const BlkDesc* desc,
                       const int32_t* col, const LT* lu, const T* r, T* y) {
  const int xcd = code >> 5;
#define SFEM_TRI(i, G, U) \
  case i: tri_launch_g<T, LT, UPPER, G, U>(st, xcd, list, first, cnt, desc, col, lu, r, y); break;
  switch (code & 31) {
    SFEM_TRI(0, 4, 1) SFEM_TRI(1, 4, 2) SFEM_TRI(2, 4, 4)
    SFEM_TRI(3, 8, 1) SFEM_TRI(4, 8, 2) SFEM_TRI(5, 8, 4) SFEM_TRI(6, 8, 0)
    SFEM_TRI(7, 16, 1) SFEM_TRI(8, 16, 2) SFEM_TRI(9, 16, 4) SFEM_TRI(10, 16, 0)
    SFEM_TRI(11, 32, 1) SFEM_TRI(12, 32, 2) SFEM_TRI(13, 32, 4)
    default: tri_launch_g<T, LT, UPPER, 32, 0>(st, xcd, list, first, cnt, desc, col, lu, r, y); break;
  }
#undef SFEM_TRI
}

// Apply the ILU(0) preconditioner: z = U^{-1} L^{-1} r, the forward solve into z and the
// backward solve in place.  r and z are in the plan's ordering (the Krylov solve runs
// there when the plan is reordered).  Each level launch uses its autotuned shape, or
// `force` for all levels when force >= 0; with ev non-null an event is recorded after
// every level (ev[0] before the first), forward levels then backward levels.
template <typename T, typename LT>
static void ilu_solve(sfem_ctx* ctx, const IluPlan& P, const LT* lu, const T* r, T* z, int force = -1,
                      hipEvent_t* ev = nullptr) {
  LinAlg* L = ctx->la;
  hipStream_t st = ctx->stream;
  const int slot = sizeof(T) == 8 ? 0 : 1;
  const BlkDesc* desc = P.identity ? L->d_desc : P.d_desc;
  const std::vector<int>& tl = L->tune_L[slot];
  const std::vector<int>& tu = L->tune_U[slot];
  const int nl = (int)P.levLPtr.size() - 1;
  const int nu = (int)P.levUPtr.size() - 1;
  if (ev) (void)hipEventRecord(ev[0], st);
  for (int lv = 0; lv < nl; ++lv) {
    const int32_t* list; int first, cnt;
    level_range(P, false, lv, &list, &first, &cnt);
    const int code = force >= 0 ? force : lv < (int)tl.size() ? tl[lv] : kTriDefault;
    if (cnt) tri_launch<T, LT, false>(st, code, list, first, cnt, desc, P.colidx, lu, r, z);
    if (ev) (void)hipEventRecord(ev[1 + lv], st);
  }
  for (int lv = 0; lv < nu; ++lv) {
    const int32_t* list; int first, cnt;
    level_range(P, true, lv, &list, &first, &cnt);
    const int code = force >= 0 ? force : lv < (int)tu.size() ? tu[lv] : kTriDefault;
    if (cnt) tri_launch<T, LT, true>(st, code, list, first, cnt, desc, P.colidx, lu, z, z);
    if (ev) (void)hipEventRecord(ev[1 + nl + lv], st);
  }
}

// ILU(0) apply with the factor type of the current preconditioner.
template <typename T>
static void ilu_apply(sfem_ctx* ctx, const T* r, T* z, int force = -1, hipEvent_t* ev = nullptr) {
  LinAlg* L = ctx->la;
  const IluPlan& P = *L->ilu;
  if (L->pc_bf16) ilu_solve<T, bf16>(ctx, P, L->lu16, r, z, force, ev);
  else if (L->pc_fp32) ilu_solve<T, float>(ctx, P, L->lu32, r, z, force, ev);
  else ilu_solve<T, double>(ctx, P, L->lu, r, z, force, ev);
}

template <typename T>
static int pc_apply_impl(sfem_ctx* ctx, const T* r, T* z) {
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
      ilu_apply<T>(ctx, r, z);
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

template <typename T>
static void spmv_in(sfem_ctx* ctx, const T* x, T* y, bool perm) {
  LinAlg* L = ctx->la;
  const int64_t nb = L->nblk;
  if (nb == 0) return;
  ctx->kt.begin(SFEM_KT_SPMV, ctx->stream);
  const int v = L->tune_spmv[sizeof(T) == 8 ? 0 : 1];
  const BlkDesc* desc = perm ? L->ilu->d_desc : L->d_desc;
  const int32_t* col = perm ? L->ilu->colidx : ctx->d_colidx;
  if (sizeof(T) == 8) spmv_launch<double, T>(ctx->stream, v, nb, desc, col, perm ? L->kp : ctx->d_vals, x, y);
  else spmv_launch<float, T>(ctx->stream, v, nb, desc, col, perm ? L->kp32 : ctx->d_vals32, x, y);
  ctx->kt.end(SFEM_KT_SPMV, ctx->stream);
}

template <typename T>
static void spmv(sfem_ctx* ctx, const T* x, T* y) {
  spmv_in<T>(ctx, x, y, ctx->la->perm_space);
}

// Autotuner, run once per precision on the first solve with the context's own matrix
// (HIP events on the context stream).  The SpMV keeps the fastest of its shapes; the
// ILU(0) apply is timed level by level for every shape and each level keeps its own
// fastest, since light levels (few lower or upper neighbours) favour few lanes per node
// block and heavy ones many.
template <typename T>
static int autotune(sfem_ctx* ctx, T* x, T* y) {
  LinAlg* L = ctx->la;
  const int slot = sizeof(T) == 8 ? 0 : 1;
  hipStream_t st = ctx->stream;
  const bool on = ctx->kt.on;
  ctx->kt.on = false;
  std::vector<hipEvent_t> ev(2);
  auto cleanup = [&] {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    ctx->kt.on = on;
  };
  for (hipEvent_t& e : ev) e = nullptr;
  for (hipEvent_t& e : ev)
    if (hipEventCreate(&e) != hipSuccess) { e = nullptr; cleanup(); return set_error(SFEM_ERR_HIP, "autotune events"); }
  hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(L->n)), dim3(kTPB), 0, st, L->n, x, (T)1);
  float best = 1e30f;
  for (int xcd = 0; xcd < 2; ++xcd)
    for (int vi = 0; vi < kSpmvCodes; ++vi) {
      L->tune_spmv[slot] = vi | (xcd << 5);
      spmv<T>(ctx, x, y);  // warm
      (void)hipEventRecord(ev[0], st);
      for (int k = 0; k < 3; ++k) spmv<T>(ctx, x, y);
      (void)hipEventRecord(ev[1], st);
      (void)hipEventSynchronize(ev[1]);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
      if (ms < best) { best = ms; L->tune_best_spmv[slot] = vi | (xcd << 5); }
    }
  L->tune_spmv[slot] = L->tune_best_spmv[slot];
  if (L->pc_type == SFEM_PC_ILU0) {
    const IluPlan& P = *L->ilu;
    const int nl = (int)P.levLPtr.size() - 1, nu = (int)P.levUPtr.size() - 1;
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.assign(nl + nu + 1, nullptr);
    for (hipEvent_t& e : ev)
      if (hipEventCreate(&e) != hipSuccess) { e = nullptr; cleanup(); return set_error(SFEM_ERR_HIP, "autotune events"); }
    std::vector<float> bt(nl + nu, 1e30f), acc(nl + nu);
    std::vector<int> bc(nl + nu, kTriDefault);
    for (int xcd = 0; xcd < 2; ++xcd)
      for (int vi = 0; vi < kTriCodes; ++vi) {
        const int code = vi | (xcd << 5);
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
  }
  cleanup();
  L->tuned[slot] = true;
  return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "autotune");
}

// The current ILU(0) factor in its plan's ordering: row permutation (plan row -> DOF
// row), pattern and values (FP32 / BF16 factors widened to double).
int la_get_ilu(sfem_ctx* ctx, int64_t* nnz_out, int32_t* perm, int32_t* rowptr, int32_t* colidx, double* vals) {
  LinAlg* L = ctx->la;
  if (!L || L->pc_type != SFEM_PC_ILU0 || !L->ilu || L->pc_K_version < 0)
    return set_error(SFEM_ERR_STATE, "no ILU(0) factor has been built");
  const IluPlan& P = *L->ilu;
  const int64_t n = L->n, nnz = ctx->nnz;
  if (nnz_out) *nnz_out = nnz;
  LA_CHECK(hipStreamSynchronize(ctx->stream));
  if (perm) {
    if (P.identity) for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    else LA_CHECK(hipMemcpy(perm, P.d_perm, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  }
  if (rowptr) LA_CHECK(hipMemcpy(rowptr, P.rowptr, (n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (colidx) LA_CHECK(hipMemcpy(colidx, P.colidx, nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (vals) {
    if (L->pc_bf16) {
      std::vector<uint16_t> h(nnz);
      LA_CHECK(hipMemcpy(h.data(), L->lu16, nnz * sizeof(uint16_t), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < nnz; ++i) {
        const uint32_t u = (uint32_t)h[i] << 16;
        float f;
        std::memcpy(&f, &u, sizeof(f));
        vals[i] = f;
      }
    } else if (L->pc_fp32) {
      std::vector<float> h(nnz);
      LA_CHECK(hipMemcpy(h.data(), L->lu32, nnz * sizeof(float), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < nnz; ++i) vals[i] = h[i];
    } else {
      LA_CHECK(hipMemcpy(vals, L->lu, nnz * sizeof(double), hipMemcpyDeviceToHost));
    }
  }
  return SFEM_OK;
}

int la_tuning(sfem_ctx* ctx, int* out) {
  LinAlg* L = ctx->la;
  for (int k = 0; k < 2; ++k) {
    const bool t = L && L->tuned[k];
    const int c = t ? L->tune_spmv[k] : 0;
    out[4 * k + 0] = t ? kSpmvG[c & 31] : 0;
    out[4 * k + 1] = t ? kSpmvU[c & 31] : 0;
    out[4 * k + 2] = t ? c >> 5 : 0;
    out[4 * k + 3] = t ? (int)(L->tune_L[k].size() + L->tune_U[k].size()) : 0;
  }
  return SFEM_OK;
}

int la_tuning_levels(sfem_ctx* ctx, int fp32, int* out, int cap, int* nlevels) {
  LinAlg* L = ctx->la;
  const int k = fp32 ? 1 : 0;
  std::vector<int> all(L->tune_L[k]);
  all.insert(all.end(), L->tune_U[k].begin(), L->tune_U[k].end());
  *nlevels = (int)all.size();
  for (int i = 0; i < (int)all.size() && i < cap; ++i) {
    out[3 * i + 0] = kTriG[all[i] & 31];
    out[3 * i + 1] = kTriU[all[i] & 31];
    out[3 * i + 2] = all[i] >> 5;
  }
  return SFEM_OK;
}

int la_spmv64(sfem_ctx* ctx, const double* x, double* y) {
  spmv_in<double>(ctx, x, y, false);
  return hipGetLastError() == hipSuccess ? SFEM_OK : set_error(SFEM_ERR_HIP, "spmv launch");
}
int la_spmv32(sfem_ctx* ctx, const float* x, float* y) {
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

// Internal status of bicgstab with a stall window: no 10% improvement of the best
// residual for `stall` iterations.
constexpr int kStagnated = -100;
constexpr int kBicgStall = 100;  // stall window of the FP32 inner BiCGStab

template <typename T>
static int bicgstab(sfem_ctx* ctx, const T* b, T* x, double rtol, int maxit, int* iters,
                    double* relres, T* r, T* rh, T* p, T* v, T* s, T* t, T* ph, T* sh, int stall = 0) {
  LinAlg* L = ctx->la;
  hipStream_t st = ctx->stream;
  const int64_t n = L->n;
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
  init[S_TOL2] = (rtol * bnorm) * (rtol * bnorm);
  std::memcpy(hs, init, sizeof(init));
  LA_CHECK(hipMemcpyAsync(sc, hs, S_NSLOTS * sizeof(double), hipMemcpyHostToDevice, st));
  double rnorm = bnorm;
  bool first = true;
  int restarts = 0;
  double best = bnorm;
  int best_it = 0;
  for (int it = 1; it <= maxit; ++it) {
    hipLaunchKernelGGL((k_dot<T>), dim3(g), dim3(kTPB), 0, st, n, rh, r, L->d_part);
    hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, sc, 0, (int)FIN_RHO1);
    hipLaunchKernelGGL((k_bicg_p<T>), dim3(g), dim3(kTPB), 0, st, n, r, p, v, sc, first ? 1 : 0);
    first = false;
    int rc = pc_apply<T>(ctx, p, ph); if (rc) return rc;
    spmv<T>(ctx, ph, v);
    hipLaunchKernelGGL((k_dot2<T>), dim3(g), dim3(kTPB), 0, st, n, rh, v, v, v, L->d_part);
    hipLaunchKernelGGL((k_finalize<2>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, sc, 0, (int)FIN_ALPHA);
    hipLaunchKernelGGL((k_bicg_s<T>), dim3(g), dim3(kTPB), 0, st, n, r, v, s, sc, L->d_part);
    hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, sc, 0, (int)FIN_SS);
    rc = pc_apply<T>(ctx, s, sh); if (rc) return rc;
    spmv<T>(ctx, sh, t);
    hipLaunchKernelGGL((k_dot2<T>), dim3(g), dim3(kTPB), 0, st, n, t, t, t, s, L->d_part);
    hipLaunchKernelGGL((k_finalize<2>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, sc, 0, (int)FIN_OMEGA);
    hipLaunchKernelGGL((k_bicg_xr<T>), dim3(g), dim3(kTPB), 0, st, n, x, ph, sh, r, s, t, sc, L->d_part);
    hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kTPB), 0, st, L->d_part, g, sc, 0, (int)FIN_RR);
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
      *relres = rnorm / bnorm;
      return SFEM_OK;
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
  const int64_t n = L->n;
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

// Shapes are tuned on the first solve of a precision and again when an ILU(0) plan
// appears that they were not tuned for.
static bool needs_tuning(const LinAlg* L, int slot) {
  if (!L->tuned[slot]) return true;
  return L->pc_type == SFEM_PC_ILU0 && L->tune_plan[slot] != L->plan_id;
}

// Symbolic set-up of the preconditioner for the given options (ILU(0): ordering,
// permuted pattern and level sets), so that the first solve does not pay for it.
int la_prepare(sfem_ctx* ctx, const sfem_solver_opts* o) {
  LinAlg* L = ctx->la;
  if (o->pc_type != SFEM_PC_ILU0) return SFEM_OK;
  if (L->maxrow > kIluMaxRow) return set_error(SFEM_ERR_UNSUPPORTED, "ILU(0): row longer than 256 entries");
  if (L->ilu && L->ilu->reorder == o->reorder) return SFEM_OK;
  return build_ilu_plan(ctx, L, o->reorder);
}

// Solve A du = rhs.  FP64 mode: everything in double.  Mixed mode: FP32
// matrix and Krylov vectors; the FP64 rhs is scaled by 1/||rhs|| into FP32 and
// the FP32 solution scaled back into the FP64 increment.
int la_solve(sfem_ctx* ctx, const sfem_solver_opts* o, int* iters, double* relres) {
  LinAlg* L = ctx->la;
  const int64_t n = L->n;
  const bool fp32 = ctx->precision == SFEM_PREC_MIXED;
  int rc;
  // the tangent is constant within a time step: rebuild the preconditioner only when it changed
  const int reorder = o->pc_type == SFEM_PC_ILU0 ? o->reorder : 0;
  const bool bf16f = fp32 && o->pc_type == SFEM_PC_ILU0 && o->factor_bf16 != 0;
  if (L->pc_K_version != ctx->la_K_version || L->pc_type != o->pc_type || L->pc_fp32 != fp32 ||
      L->pc_reorder != reorder || L->pc_bf16 != bf16f) {
    if ((rc = la_setup_pc(ctx, o->pc_type, fp32, reorder, bf16f))) return rc;
    L->pc_reorder = reorder;
    L->pc_K_version = ctx->la_K_version;
  }
  // a reordered ILU(0) plan moves the whole Krylov solve into its ordering
  L->perm_space = L->pc_type == SFEM_PC_ILU0 && L->ilu && !L->ilu->identity;
  const int32_t* perm = L->perm_space ? L->ilu->d_perm : nullptr;
  const int gn = grid_for(n);
  const int m = std::max(1, o->restart);
  if (!fp32) {
    const double* b = ctx->d_rhs;
    double* x = ctx->d_du;
    if (perm) {
      if ((rc = ensure_vec(&L->pr, n)) || (rc = ensure_vec(&L->pz, n))) return rc;
      hipLaunchKernelGGL((k_permute_in<double, double>), dim3(gn), dim3(kTPB), 0, ctx->stream, n, perm,
                         ctx->d_rhs, L->pr, 1.0);
      b = L->pr;
      x = L->pz;
    }
    if (o->ksp_type == SFEM_KSP_GMRES) {
      if ((rc = ensure_vec(&L->w, n)) || (rc = ensure_vec(&L->z, n))) return rc;
      if (!L->gm_V || L->gm_m < m) {
        if (L->gm_V) (void)hipFree(L->gm_V);
        L->gm_V = nullptr;
        if ((rc = ensure_vec(&L->gm_V, (int64_t)(m + 1) * n))) return rc;
        L->gm_m = m;
      }
      rc = gmres<double>(ctx, b, x, m, o->rtol, o->max_it, iters, relres, L->gm_V, L->w, L->z);
    } else {
      double** vs[] = {&L->r, &L->rh, &L->p, &L->v, &L->s, &L->t, &L->ph, &L->sh};
      for (double** q : vs)
        if ((rc = ensure_vec(q, n))) return rc;
      if (needs_tuning(L, 0) && (rc = autotune<double>(ctx, L->t, L->v))) return rc;
      rc = bicgstab<double>(ctx, b, x, o->rtol, o->max_it, iters, relres, L->r, L->rh, L->p, L->v, L->s,
                            L->t, L->ph, L->sh);
    }
    if (perm) {
      hipLaunchKernelGGL((k_permute_out<double, double>), dim3(gn), dim3(kTPB), 0, ctx->stream, n, perm, x,
                         ctx->d_du, 1.0);
      if (hipGetLastError() != hipSuccess) return set_error(SFEM_ERR_HIP, "permute launch");
    }
    return rc;
  }
  // mixed precision
  float** vs[] = {&L->r32, &L->rh32, &L->p32, &L->v32, &L->s32, &L->t32, &L->ph32, &L->sh32, &L->x32};
  for (float** q : vs)
    if ((rc = ensure_vec(q, n))) return rc;
  if (needs_tuning(L, 1) && (rc = autotune<float>(ctx, L->t32, L->v32))) return rc;
  double bn;
  if ((rc = la_norm2(ctx, ctx->d_rhs, &bn))) return rc;
  const int g = grid_for(n);
  if (!(bn > 0.0)) {
    hipLaunchKernelGGL((k_fill<double>), dim3(g), dim3(kTPB), 0, ctx->stream, n, ctx->d_du, 0.0);
    *iters = 0; *relres = 0.0;
    return SFEM_OK;
  }
  // b32 = rhs / ||rhs|| into the FP32 "s" slot's sibling (rh32 is overwritten by the solver, use x32 as b)
  float* b32 = L->x32;
  if (perm)
    hipLaunchKernelGGL((k_permute_in<double, float>), dim3(g), dim3(kTPB), 0, ctx->stream, n, perm, ctx->d_rhs,
                       b32, 1.0 / bn);
  else
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(kTPB), 0, ctx->stream, n, ctx->d_rhs, b32,
                       1.0 / bn);
  if (!L->du32) { if ((rc = ensure_vec(&L->du32, n))) return rc; }
  if (o->ksp_type == SFEM_KSP_GMRES) {
    if ((rc = ensure_vec(&L->w32, n)) || (rc = ensure_vec(&L->z32, n))) return rc;
    if (!L->gm_V32 || L->gm_m32 < m) {
      if (L->gm_V32) (void)hipFree(L->gm_V32);
      L->gm_V32 = nullptr;
      if ((rc = ensure_vec(&L->gm_V32, (int64_t)(m + 1) * n))) return rc;
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
      if ((rc = ensure_vec(&L->w32, n)) || (rc = ensure_vec(&L->z32, n))) return rc;
      if (!L->gm_V32 || L->gm_m32 < mg) {
        if (L->gm_V32) (void)hipFree(L->gm_V32);
        L->gm_V32 = nullptr;
        if ((rc = ensure_vec(&L->gm_V32, (int64_t)(mg + 1) * n))) return rc;
        L->gm_m32 = mg;
      }
      ++L->fallbacks;
      rc = gmres<float>(ctx, b32, L->du32, mg, o->rtol, std::max(1, o->max_it - used), iters, relres, L->gm_V32,
                        L->w32, L->z32);
      *iters += used;
    }
  }
  if (perm)
    hipLaunchKernelGGL((k_permute_out<float, double>), dim3(g), dim3(kTPB), 0, ctx->stream, n, perm, L->du32,
                       ctx->d_du, bn);
  else
    hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(kTPB), 0, ctx->stream, n, L->du32, ctx->d_du, bn);
  if (hipGetLastError() != hipSuccess) return set_error(SFEM_ERR_HIP, "convert launch");
  return rc;
}

}  // namespace sfem
