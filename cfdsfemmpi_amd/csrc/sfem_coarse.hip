// sfem_coarse.hip -- coarse pressure correction of the ILU(0) preconditioner.
//
// No counterpart in the reference: it preconditions every Newton system with PETSc's
// ILU(0) (SolverPetsc::factoriseAndSolve, src/base/SolverPetsc.cpp:284-298), whose
// BiCGStab iteration count grows with the mesh because ILU(0) only damps the pressure
// error locally.  This adds a two-level correction, additive or multiplicative:
//
//   z = M_ILU^{-1} r + Z E^{-1} Z^T r,   or   z0 = M_ILU^{-1} r, z = z0 + Z E^{-1} Z^T (r - A z0),
//   E = Z^T A Z,
//
// with Z the piecewise-constant pressure on k^d boxes of the mesh's bounding box (one
// column per box holding at least one free pressure DOF; velocity DOFs get no coarse
// component).  A batch of independent meshes (sfem_set_batch) gets k^d boxes per mesh,
// on the mesh's own bounding box: E is then block diagonal, one block per mesh.  The
// preconditioned operator changes, the linear system and its tolerance do not.
//
//   * Layout: the ILU(0) plan's padded block vectors (4 slots per node block).  The free
//     pressure slots are listed sorted by (mesh, box, slot) and cut into chunks of at most
//     kChunk slots that never straddle a box; every sum below runs over chunks and lanes
//     in a fixed order, so E, its inverse and every correction are the same from run to
//     run.
//   * E: at each preconditioner rebuild, one wavefront per chunk reads the pressure row
//     of each of its block rows in the plan's block tangent (the tangent the Krylov
//     products use) and sums its pressure columns by the box offset (-1, 0, +1 per
//     direction) of the neighbour's box.  Boxes at least one element wide keep every
//     neighbour within one box in each direction (checked on the host over the elements).
//   * E's diagonal: the pressure-pressure block of the stabilised tangent is the PSPG
//     term, integral of tau grad N_i . grad N_j, whose rows sum to zero (the shape
//     functions sum to one) except where a neighbour's pressure is prescribed.  Summing
//     a box's own block directly cancels tens of thousands of entries down to its
//     couplings with the neighbouring boxes and the prescribed pressures, which FP32
//     entries cannot resolve; so E(c, c) is formed as the row sums of the rows next to a
//     prescribed pressure minus the box's couplings to the other boxes,
//     sum_{i in c} rowsum_i - sum_{c' != c} E(c, c'), the same value in exact arithmetic.
//   * Inverse: Gauss-Jordan elimination without pivoting (E is symmetric and definite),
//     one launch per step with one workgroup per row for a single mesh, one workgroup per
//     mesh for a batch; FP32 Krylov solves use the inverse without the global constant
//     pressure of each mesh, whose correction their vectors cannot hold beside the rest
//     (the FP64 Newton loop takes it as a coarse step).
//   * A mesh with no prescribed pressure (a traction-free outflow fixes the pressure level
//     through the momentum rows instead): every row of its E sums to zero, the constant is
//     its null vector, and the correction is E's pseudo-inverse, formed as the inverse of
//     E + (m / n) 1 1^T (m the mean diagonal, n the mesh's unknowns) without the constant,
//     B = F - (F 1)(1^T F) / (1^T F 1) with F that inverse; the constant pressure is left
//     to the Krylov iteration.
//   * Apply: restriction (chunk sums of r at the pressure slots, before the ILU(0)
//     sweeps, so z may alias r), the coarse solve (box sums, then y = E^{-1} c), and the
//     prolongation z[slot] += y[box] after the sweeps; the multiplicative form restricts
//     r - A z0 and corrects the product A z0 with A Z y.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "sfem_internal.hpp"
#include "sfem_linalg.hpp"

namespace sfem {

constexpr int kChunk = 512;       // pressure slots per chunk (8 per lane of a wavefront)
constexpr int kOffs = 27;         // box offsets (-1, 0, +1)^3
constexpr int kSelf = 13;         // offset (0, 0, 0)
constexpr int kOffsPad = 29;      // LDS row stride of the per-lane offset sums (odd: no bank conflicts)
constexpr int kMaxCoarse = 1024;  // coarse unknowns handled

#define CO_CHECK(x)                                     \
  do {                                                  \
    hipError_t _e = (x);                                \
    if (_e != hipSuccess) return set_hip_error(_e, #x); \
  } while (0)

struct Coarse {
  int k = 0;
  int plan_id = -1;
  int nc = 0;             // coarse unknowns (boxes with a free pressure slot)
  int64_t npres = 0;      // free pressure slots
  int nchunk = 0;
  bool ready = false;     // E^{-1} built for the current factor
  bool fp32 = false;      // built from the FP32 tangent (mixed precision)
  int32_t* d_slot = nullptr;      // npres: padded slot 4p + a, sorted by (box, slot)
  int32_t* d_chunk = nullptr;     // nchunk + 1: first entry of d_slot per chunk
  int32_t* d_chunkbox = nullptr;  // nchunk: coarse unknown of the chunk
  int32_t* d_boxchunk = nullptr;  // nc + 1: chunks of each coarse unknown
  int32_t* d_cinfo = nullptr;     // per block row: coarse unknown << 3 | (next to a prescribed
                                  // pressure) << 2 | pressure slot, -1 without pressure
  int32_t* d_cpos = nullptr;      // per coarse unknown: box coordinates bx | by << 10 | bz << 20
  int32_t* d_boxnb = nullptr;     // nc * 27: coarse unknown of the box at each offset, -1 if none
  // meshes (one for a single mesh): unknowns of mesh g are [goff[g], goff[g+1]); its
  // blocks of the Gauss-Jordan workspace (ng x 2ng) and of the inverses (ng x ng) start at
  // gjbase[g] and ebase[g]; the coarse solve's workgroups are {mesh, first row} records
  int ngroup = 1;
  std::vector<int32_t> h_goff;
  std::vector<int64_t> h_ebase;
  int32_t* d_cgrp = nullptr;
  int32_t* d_goff = nullptr;
  int64_t* d_gjbase = nullptr;
  int64_t* d_ebase = nullptr;
  int2* d_swg = nullptr;
  int nswg = 0;
  int maxng = 0;
  double* d_epart = nullptr;      // nchunk * 27 chunk sums of E's rows by offset
  double* d_gj = nullptr;         // per mesh ng x 2ng Gauss-Jordan workspace [E | I]
  double* d_einv = nullptr;       // per mesh ng x ng: E^{-1}, row-major
  double* d_einvb = nullptr;      // the correction without the global constant (FP32 Krylov
                                  // vectors), or NULL
  double* d_cpart = nullptr;      // nchunk restricted chunk sums
  double* d_y = nullptr;          // nc coarse solution
  int32_t* d_flag = nullptr;      // 1: E was singular (correction switched off)
  int32_t* d_gfloat = nullptr;    // per mesh, 1: no prescribed pressure (E 1 = 0)
  bool any_float = false;
  // A Z for the multiplicative form: per block row the coarse unknowns of its neighbours'
  // pressure DOFs (d_azoff[p] .. d_azoff[p+1] in d_azbox, ascending) and, per entry, the 4
  // rows' sums of A over that unknown's pressure columns (FP64 or FP32, like the tangent)
  int64_t naz = 0;
  int32_t* d_azoff = nullptr;
  int32_t* d_azbox = nullptr;
  void* d_az = nullptr;
  bool az_fp32 = false;
  void* d_tmp = nullptr;          // A z of a standalone multiplicative application (npad x 8 bytes)
  ~Coarse() {
    void* ptrs[] = {d_slot, d_chunk, d_chunkbox, d_boxchunk, d_cinfo, d_cpos, d_boxnb,
                    d_epart, d_gj, d_einv, d_einvb, d_cpart, d_y, d_flag, d_azoff, d_azbox, d_az, d_tmp,
                    d_gfloat, d_cgrp, d_goff, d_gjbase, d_ebase, d_swg};
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
  }
};

void coarse_destroy(Coarse* C) { delete C; }

// ---------------------------------------------------------------------------
// Device kernels
// ---------------------------------------------------------------------------

__device__ inline int box_offset(int32_t a, int32_t b) {
  const int dx = (b & 1023) - (a & 1023);
  const int dy = ((b >> 10) & 1023) - ((a >> 10) & 1023);
  const int dz = ((b >> 20) & 1023) - ((a >> 20) & 1023);
  return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
}

// Chunk sums of E's rows: for the block rows of a chunk (one wavefront, each lane its
// rows in order), the pressure row's entries in pressure columns of other boxes, summed
// per box offset of the column's box in neighbour order, and at the box's own offset the
// full row sums of rows next to a prescribed pressure; then per offset the 64 lanes in
// order.
template <typename VT>
__global__ void __launch_bounds__(64) k_coarse_rows(const int32_t* __restrict__ slot, const int32_t* __restrict__ chunk,
                                                     const int4* __restrict__ bdesc, const int32_t* __restrict__ bcol,
                                                     const VT* __restrict__ kp, const int32_t* __restrict__ cinfo,
                                                     const int32_t* __restrict__ cpos, double* __restrict__ epart) {
  __shared__ double acc[64 * kOffsPad];
  const int lane = threadIdx.x, ch = blockIdx.x;
  double* my = acc + lane * kOffsPad;
  for (int o = 0; o < kOffs; ++o) my[o] = 0.0;
  const int i0 = chunk[ch], i1 = chunk[ch + 1];
  for (int i = i0 + lane; i < i1; i += 64) {
    const int s = slot[i];
    const int p = s >> 2, a = s & 3;
    const int4 d = bdesc[p];
    const int vbase = d.x, bptr = d.y, nb = d.z;
    const int32_t own = cinfo[p];
    const int32_t pos = cpos[own >> 3];
    const VT* row = kp + vbase + (int64_t)a * 4 * nb;
    double rs = 0.0;
    for (int t = 0; t < nb; ++t) {
      const int info = cinfo[bcol[bptr + t]];
      if (info < 0) continue;
      const double v = (double)row[4 * t + (info & 3)];
      const int o = box_offset(pos, cpos[info >> 3]);
      if (o != kSelf && (unsigned)o < (unsigned)kOffs) my[o] += v;
      rs += v;
    }
    if (own & 4) my[kSelf] += rs;
  }
  __syncthreads();
  if (lane < kOffs) {
    double s = 0.0;
    for (int l = 0; l < 64; ++l) s += acc[l * kOffsPad + lane];
    epart[(int64_t)ch * kOffs + lane] = s;
  }
}

// [E | I] from the chunk sums, one thread per row of E, in its mesh's block: each
// (unknown, offset) cell is one entry of E, summed over the unknown's chunks in order.
__global__ void __launch_bounds__(64) k_coarse_fill(int nc, const int32_t* __restrict__ cgrp,
                                                     const int32_t* __restrict__ goff,
                                                     const int64_t* __restrict__ gjbase,
                                                     const int32_t* __restrict__ boxchunk,
                                                     const int32_t* __restrict__ boxnb,
                                                     const double* __restrict__ epart, double* __restrict__ gj) {
  const int c = blockIdx.x * 64 + threadIdx.x;
  if (c >= nc) return;
  const int g = cgrp[c], g0 = goff[g], ng = goff[g + 1] - g0, lc = c - g0;
  const int w = 2 * ng;
  double* row = gj + gjbase[g] + (int64_t)lc * w;
  for (int j = 0; j < w; ++j) row[j] = j == ng + lc ? 1.0 : 0.0;
  double own = 0.0;
  for (int o = 0; o < kOffs; ++o) {
    const int cc = boxnb[c * kOffs + o];
    double s = 0.0;
    for (int q = boxchunk[c]; q < boxchunk[c + 1]; ++q) s += epart[(int64_t)q * kOffs + o];
    if (o == kSelf) own += s;       // row sums next to prescribed pressures
    else if (cc >= 0) { row[cc - g0] = s; own -= s; }
  }
  row[lc] = own;  // E(c, c) = those row sums - the couplings to the other boxes
}

// Meshes without a prescribed pressure (gfloat[g] = 1), one workgroup each: E += (m / n)
// 1 1^T, m the mean of E's diagonal summed in row order, so that the elimination sees a
// definite matrix whose inverse, without the constant, is E's pseudo-inverse.
__global__ void __launch_bounds__(256) k_coarse_float(const int32_t* __restrict__ goff,
                                                      const int64_t* __restrict__ gjbase,
                                                      const int32_t* __restrict__ gfloat, double* __restrict__ gj) {
  __shared__ double s_shift;
  const int g = blockIdx.x;
  if (!gfloat[g]) return;
  const int ng = goff[g + 1] - goff[g], w = 2 * ng;
  double* a = gj + gjbase[g];
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < ng; ++i) t += a[(int64_t)i * w + i];
    s_shift = t / ((double)ng * (double)ng);
  }
  __syncthreads();
  const double sh = s_shift;
  for (int e = threadIdx.x; e < ng * ng; e += 256) {
    const int i = e / ng, j = e - i * ng;
    a[(int64_t)i * w + j] += sh;
  }
}

// Gauss-Jordan step j without row scaling: row i -= (E(i, j) / E(j, j)) row j for every
// i != j, one workgroup per row, over columns j+1 .. nc+j (row j is zero beyond them;
// earlier columns are not read again).  Row j itself is left alone, so no workgroup
// writes what another reads.  E is symmetric and definite (the pressure block of the
// tangent is a weighted Laplacian), so no pivoting is needed; the finishing kernel checks
// the pivots.
__global__ void __launch_bounds__(256) k_gj_step(int nc, int j, double* __restrict__ gj) {
  const int i = blockIdx.x;
  if (i == j) return;
  const int w = 2 * nc;
  double* ri = gj + (int64_t)i * w;
  const double* rj = gj + (int64_t)j * w;
  const double f = ri[j] / rj[j];
  if (f == 0.0) return;
  for (int c = j + 1 + threadIdx.x; c <= nc + j; c += 256) ri[c] -= f * rj[c];
}

// All Gauss-Jordan steps of one mesh's block (the same steps as k_gj_step), one
// workgroup per mesh: the batches' blocks are small (k^dim unknowns).
__global__ void __launch_bounds__(256) k_gj_group(const int32_t* __restrict__ goff,
                                                  const int64_t* __restrict__ gjbase, double* __restrict__ gj) {
  const int g = blockIdx.x;
  const int ng = goff[g + 1] - goff[g], w = 2 * ng;
  double* a = gj + gjbase[g];
  for (int j = 0; j < ng; ++j) {
    const double* rj = a + (int64_t)j * w;
    for (int e = threadIdx.x; e < ng * ng; e += 256) {
      const int i = e / ng, c = j + 1 + (e - i * ng);
      if (i == j) continue;
      double* ri = a + (int64_t)i * w;
      const double f = ri[j] / rj[j];
      if (f != 0.0) ri[c] -= f * rj[c];
    }
    __syncthreads();
  }
}

// Per mesh (one workgroup each): E^{-1} (row-major) = the right half scaled by the
// diagonal the steps left; with border = 1 also B = E^{-1} - (E^{-1} 1)(1^T E^{-1}) /
// (1^T E^{-1} 1), the leading block of [E b1; b1^T 0]^{-1}: the correction whose output
// sums to zero.  A mesh without a prescribed pressure (gfloat[g]) gets that form in both
// (its E was shifted by a multiple of 1 1^T: the result is E's pseudo-inverse).  A pivot
// below 1e-15 of the mesh's largest (or not finite) zeroes both for that mesh and sets
// flag[g].  Sums in a fixed order.
__global__ void __launch_bounds__(256) k_coarse_finish(int border, const int32_t* __restrict__ gfloat,
                                                       const int32_t* __restrict__ goff,
                                                       const int64_t* __restrict__ gjbase,
                                                       const int64_t* __restrict__ ebase, const double* __restrict__ gjall,
                                                       double* __restrict__ einvall, double* __restrict__ einvball,
                                                       int32_t* __restrict__ flag) {
  __shared__ double red[256];
  __shared__ double u[kMaxCoarse], v[kMaxCoarse];
  __shared__ double s_sum;
  __shared__ int s_bad;
  const int g = blockIdx.x;
  const int nc = goff[g + 1] - goff[g];
  const double* gj = gjall + gjbase[g];
  double* einv = einvall + ebase[g];
  double* einvb = border ? einvball + ebase[g] : nullptr;
  const bool fl = gfloat[g] != 0;
  const int tid = threadIdx.x;
  const int w = 2 * nc;
  double dmax = 0.0;
  for (int i = tid; i < nc; i += 256) dmax = fmax(dmax, fabs(gj[(int64_t)i * w + i]));
  red[tid] = dmax;
  if (tid == 0) s_bad = 0;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] = fmax(red[tid], red[tid + h]);
    __syncthreads();
  }
  const double tol = 1e-15 * red[0];
  for (int i = tid; i < nc; i += 256) {
    const double d = gj[(int64_t)i * w + i];
    if (!(fabs(d) > tol) || !isfinite(d)) s_bad = 1;
  }
  __syncthreads();
  for (int e = tid; e < nc * nc; e += 256) {
    const int i = e / nc, c = e - i * nc;
    const double x = gj[(int64_t)i * w + nc + c] / gj[(int64_t)i * w + i];
    einv[e] = x;
    if (!isfinite(x)) s_bad = 1;
  }
  __syncthreads();
  if (border || fl) {
    for (int i = tid; i < nc; i += 256) {
      double a = 0.0, b = 0.0;
      for (int c = 0; c < nc; ++c) {
        a += einv[(int64_t)i * nc + c];   // row sums: E^{-1} 1
        b += einv[(int64_t)c * nc + i];   // column sums: 1^T E^{-1}
      }
      u[i] = a;
      v[i] = b;
    }
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int i = 0; i < nc; ++i) t += u[i];
      s_sum = t;
      if (!(fabs(t) > 0.0) || !isfinite(t)) s_bad = 1;
    }
    __syncthreads();
    const double sinv = 1.0 / s_sum;
    for (int e = tid; e < nc * nc; e += 256) {
      const int i = e / nc, c = e - i * nc;
      const double b = einv[e] - u[i] * v[c] * sinv;
      if (border) einvb[e] = b;
      if (fl) einv[e] = b;
    }
    __syncthreads();
  }
  if (s_bad) {
    for (int e = tid; e < nc * nc; e += 256) {
      einv[e] = 0.0;
      if (border) einvb[e] = 0.0;
    }
  }
  if (tid == 0) flag[g] = s_bad;
}

// Restriction: chunk sums of r (or of r - w) over the chunk's pressure slots (lanes in
// order, then a fixed tree over the wavefront).
template <typename T>
__global__ void __launch_bounds__(64) k_coarse_restrict(const int32_t* __restrict__ slot,
                                                         const int32_t* __restrict__ chunk, const T* __restrict__ r,
                                                         const T* __restrict__ w, double* __restrict__ cpart) {
  const int lane = threadIdx.x, ch = blockIdx.x;
  double s = 0.0;
  for (int i = chunk[ch] + lane; i < chunk[ch + 1]; i += 64) {
    const int q = slot[i];
    s += w ? (double)r[q] - (double)w[q] : (double)r[q];
  }
  for (int h = 32; h > 0; h >>= 1) s += __shfl_xor(s, h);
  if (lane == 0) cpart[ch] = s;
}

// A Z: for block row p and each coarse unknown of its neighbours' pressure DOFs, the 4
// rows' sums of the block tangent over that unknown's pressure columns, neighbours in
// order.  One thread per block row (once per preconditioner rebuild).
template <typename VT>
__global__ void __launch_bounds__(256) k_coarse_az_build(int64_t nblk, const int4* __restrict__ bdesc,
                                                          const int32_t* __restrict__ bcol, const VT* __restrict__ kp,
                                                          const int32_t* __restrict__ cinfo,
                                                          const int32_t* __restrict__ azoff,
                                                          const int32_t* __restrict__ azbox, VT* __restrict__ az) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= nblk) return;
  const int4 d = bdesc[p];
  const int vbase = d.x, bptr = d.y, nb = d.z, m = d.w >> 16;
  for (int k = azoff[p]; k < azoff[p + 1]; ++k) {
    const int c = azbox[k];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int t = 0; t < nb; ++t) {
      const int info = cinfo[bcol[bptr + t]];
      if (info < 0 || (info >> 3) != c) continue;
      const VT* col = kp + vbase + 4 * t + (info & 3);
      a0 += (double)col[0];
      if (m > 1) a1 += (double)col[4 * nb];
      if (m > 2) a2 += (double)col[8 * nb];
      if (m > 3) a3 += (double)col[12 * nb];
    }
    VT* o = az + 4 * (int64_t)k;
    o[0] = (VT)a0;
    o[1] = (VT)a1;
    o[2] = (VT)a2;
    o[3] = (VT)a3;
  }
}

// The multiplicative form's tail in one pass over the block rows: out += A Z y (rows
// a >= m carry 0, so padding slots stay 0) and z[slot] += y[box] at the row's pressure
// slot (as k_coarse_prolong; each slot is one row's, so the order does not matter).
template <typename T, typename VT>
__global__ void __launch_bounds__(256) k_coarse_post(int64_t nblk, const int32_t* __restrict__ cinfo,
                                                      const int32_t* __restrict__ azoff,
                                                      const int32_t* __restrict__ azbox, const VT* __restrict__ az,
                                                      const double* __restrict__ y, T* __restrict__ z,
                                                      T* __restrict__ v) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= nblk) return;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int k = azoff[p]; k < azoff[p + 1]; ++k) {
    const double yk = y[azbox[k]];
    const VT* o = az + 4 * (int64_t)k;
    a0 += (double)o[0] * yk;
    a1 += (double)o[1] * yk;
    a2 += (double)o[2] * yk;
    a3 += (double)o[3] * yk;
  }
  T* vp = v + 4 * p;
  vp[0] = (T)((double)vp[0] + a0);
  vp[1] = (T)((double)vp[1] + a1);
  vp[2] = (T)((double)vp[2] + a2);
  vp[3] = (T)((double)vp[3] + a3);
  const int info = cinfo[p];
  if (info >= 0) {
    T* zs = z + 4 * p + (info & 3);
    *zs = (T)((double)*zs + y[info >> 3]);
  }
}

// Coarse solve, per workgroup record {mesh g, first row r0}: the mesh's box sums of the
// chunk sums (one lane per box, chunks in order), then y = E_g^{-1} c for rows r0 .. r0+3
// of the mesh, one wavefront per row (lanes over columns in order, a fixed shuffle tree).
__global__ void __launch_bounds__(256) k_coarse_solve(const int2* __restrict__ swg, const int32_t* __restrict__ goff,
                                                      const int64_t* __restrict__ ebase,
                                                      const int32_t* __restrict__ boxchunk,
                                                      const double* __restrict__ cpart,
                                                      const double* __restrict__ einv, double* __restrict__ y) {
  __shared__ double c[kMaxCoarse];
  const int tid = threadIdx.x;
  const int2 rec = swg[blockIdx.x];
  const int g0 = goff[rec.x], ng = goff[rec.x + 1] - g0;
  for (int b = tid; b < ng; b += 256) {
    double s = 0.0;
    for (int q = boxchunk[g0 + b]; q < boxchunk[g0 + b + 1]; ++q) s += cpart[q];
    c[b] = s;
  }
  __syncthreads();
  const int i = rec.y + (tid >> 6), lane = tid & 63;
  if (i >= ng) return;
  const double* row = einv + ebase[rec.x] + (int64_t)i * ng;
  double s = 0.0;
  for (int j = lane; j < ng; j += 64) s += row[j] * c[j];
  for (int h = 32; h > 0; h >>= 1) s += __shfl_xor(s, h);
  if (lane == 0) y[g0 + i] = s;
}

// Prolongation: z[slot] += y[box of the slot's chunk].
template <typename T>
__global__ void __launch_bounds__(64) k_coarse_prolong(const int32_t* __restrict__ slot,
                                                        const int32_t* __restrict__ chunk,
                                                        const int32_t* __restrict__ chunkbox,
                                                        const double* __restrict__ y, T* __restrict__ z) {
  const int lane = threadIdx.x, ch = blockIdx.x;
  const double v = y[chunkbox[ch]];
  for (int i = chunk[ch] + lane; i < chunk[ch + 1]; i += 64) {
    const int s = slot[i];
    z[s] = (T)((double)z[s] + v);
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

template <typename X>
static int upload(X** d, const std::vector<X>& h) {
  if (*d) (void)hipFree(*d);
  *d = nullptr;
  CO_CHECK(hipMalloc((void**)d, std::max<size_t>(h.size(), 1) * sizeof(X)));
  if (!h.empty()) CO_CHECK(hipMemcpy(*d, h.data(), h.size() * sizeof(X), hipMemcpyHostToDevice));
  return SFEM_OK;
}

// Mesh of every node (0 for a single mesh; a batch's meshes, nodes of no element -1).
static int node_groups(const sfem_ctx* ctx, std::vector<int32_t>& grp) {
  const std::vector<int32_t>* nm = batch_node_mesh(ctx);
  int ng = 1;
  if (nm) {
    grp = *nm;
    for (int32_t m : grp) ng = std::max(ng, m + 1);
  } else {
    grp.assign(ctx->nNode, 0);
  }
  return ng;
}

// Box coordinates of every node on its mesh's bounding box:
// b_d = min(k - 1, floor(k (x_d - min_d) / extent_d)).
static void node_boxes(const sfem_ctx* ctx, int k, const std::vector<int32_t>& grp, int ngroup,
                       std::vector<int32_t>& pos) {
  const int nd = ctx->dims.ndim;
  const int64_t nn = ctx->nNode;
  std::vector<double> lo((size_t)ngroup * 3, 1e300), hi((size_t)ngroup * 3, -1e300);
  for (int64_t v = 0; v < nn; ++v) {
    const int g = grp[v];
    if (g < 0) continue;
    for (int d = 0; d < nd; ++d) {
      lo[g * 3 + d] = std::min(lo[g * 3 + d], ctx->coords[v * nd + d]);
      hi[g * 3 + d] = std::max(hi[g * 3 + d], ctx->coords[v * nd + d]);
    }
  }
  pos.assign(nn, 0);
  for (int64_t v = 0; v < nn; ++v) {
    const int g = grp[v];
    if (g < 0) continue;
    int32_t code = 0;
    for (int d = 0; d < nd; ++d) {
      const double ext = hi[g * 3 + d] - lo[g * 3 + d];
      int b = ext > 0 ? (int)std::floor((ctx->coords[v * nd + d] - lo[g * 3 + d]) / ext * k) : 0;
      b = std::min(std::max(b, 0), k - 1);
      code |= b << (10 * d);
    }
    pos[v] = code;
  }
}

// k^dim boxes (per mesh), and every element's nodes within one box of each other in each
// direction, so that the neighbours of a block row sit at box offsets in {-1, 0, 1}^dim.
static int coarse_check(const sfem_ctx* ctx, int k) {
  const int nd = ctx->dims.ndim, npe = ctx->dims.npe;
  int64_t nbox = 1;
  for (int d = 0; d < nd; ++d) nbox *= k;
  if (k < 1 || nbox > kMaxCoarse) return set_error(SFEM_ERR_ARG, "coarse boxes: k^dim must lie in [1, 1024]");
  std::vector<int32_t> grp, npos;
  const int ngroup = node_groups(ctx, grp);
  node_boxes(ctx, k, grp, ngroup, npos);
  for (int64_t e = 0; e < ctx->nElem; ++e) {
    const int32_t* en = &ctx->conn[e * npe];
    for (int d = 0; d < nd; ++d) {
      int mn = 1 << 20, mx = -1;
      for (int a = 0; a < npe; ++a) {
        const int b = (npos[en[a]] >> (10 * d)) & 1023;
        mn = std::min(mn, b);
        mx = std::max(mx, b);
      }
      if (mx - mn > 1) return set_error(SFEM_ERR_ARG, "coarse boxes narrower than an element: use fewer boxes");
    }
  }
  return SFEM_OK;
}

// Host lists for the current ILU(0) plan: pressure slots by box, chunks, the per-block
// coarse info and the box neighbourhoods.
static int coarse_plan(sfem_ctx* ctx, LinAlg* L, Coarse* C) {
  const IluPlan& P = *L->ilu;
  const int k = C->k;
  const int nd = ctx->dims.ndim, ndof = ctx->dims.ndof, npe = ctx->dims.npe;
  std::vector<int32_t> grp, npos;
  const int ngroup = node_groups(ctx, grp);
  node_boxes(ctx, k, grp, ngroup, npos);
  const int64_t nb = P.nblk;
  const std::vector<int32_t>& bs = L->h_blkStart;
  // nodes sharing an element with a node whose pressure is prescribed: their pressure
  // rows do not sum to zero
  std::vector<uint8_t> nearfix(ctx->nNode, 0);
  std::vector<int32_t> gfloat(ngroup, 1);
  for (int64_t n = 0; n < ctx->nNode; ++n)
    if (grp[n] >= 0 && ctx->nodeDof[n * ndof + ndof - 1] < 0) gfloat[grp[n]] = 0;
  for (int64_t e = 0; e < ctx->nElem; ++e) {
    const int32_t* en = &ctx->conn[e * npe];
    bool fix = false;
    for (int a = 0; a < npe; ++a) fix = fix || ctx->nodeDof[(int64_t)en[a] * ndof + ndof - 1] < 0;
    if (fix)
      for (int a = 0; a < npe; ++a) nearfix[en[a]] = 1;
  }
  // pressure slot of each block row (plan position), its box (mesh << 30 | box
  // coordinates), and its node's mark
  std::vector<int32_t> pslot(nb, -1);
  std::vector<int64_t> pbox(nb, 0);
  std::vector<uint8_t> pfix(nb, 0);
  int64_t npres = 0;
  for (int64_t p = 0; p < nb; ++p) {
    const int32_t b = P.h_order[p];
    for (int32_t r = bs[b]; r < bs[b + 1]; ++r) {
      const int32_t g = ctx->assyForSoln[r];
      if (g % ndof == ndof - 1) {
        pslot[p] = r - bs[b];
        pbox[p] = (int64_t)std::max(grp[g / ndof], 0) << 30 | npos[g / ndof];
        pfix[p] = nearfix[g / ndof];
        ++npres;
      }
    }
  }
  if (npres == 0) return set_error(SFEM_ERR_ARG, "coarse correction: no free pressure DOF");
  // coarse unknowns: non-empty boxes in (mesh, box) order, x fastest
  auto key = [](int64_t gb) {
    const int32_t pos = (int32_t)(gb & ((1 << 30) - 1));
    return (gb >> 30) << 30 | (int64_t)((pos >> 20) & 1023) << 20 | (int64_t)((pos >> 10) & 1023) << 10 | (pos & 1023);
  };
  std::vector<int64_t> bkeys;
  bkeys.reserve(npres);
  for (int64_t p = 0; p < nb; ++p)
    if (pslot[p] >= 0) bkeys.push_back(key(pbox[p]));
  std::sort(bkeys.begin(), bkeys.end());
  bkeys.erase(std::unique(bkeys.begin(), bkeys.end()), bkeys.end());
  const int nc = (int)bkeys.size();
  std::vector<int32_t> boxes(nc), cgrp(nc);
  for (int c = 0; c < nc; ++c) {
    const int64_t kk = bkeys[c];
    const int32_t bx = kk & 1023, by = (kk >> 10) & 1023, bz = (kk >> 20) & 1023;
    boxes[c] = bx | by << 10 | bz << 20;
    cgrp[c] = (int32_t)(kk >> 30);
  }
  auto coarse_of = [&](int64_t gb) -> int {
    const auto it = std::lower_bound(bkeys.begin(), bkeys.end(), key(gb));
    return it != bkeys.end() && *it == key(gb) ? (int)(it - bkeys.begin()) : -1;
  };
  std::vector<int32_t> cinfo(nb, -1);
  std::vector<std::pair<int32_t, int32_t>> order;  // (coarse unknown, slot)
  order.reserve(npres);
  for (int64_t p = 0; p < nb; ++p)
    if (pslot[p] >= 0) {
      const int c = coarse_of(pbox[p]);
      cinfo[p] = c << 3 | pfix[p] << 2 | pslot[p];
      order.push_back({c, (int32_t)(4 * p + pslot[p])});
    }
  std::sort(order.begin(), order.end());
  std::vector<int32_t> slot(npres), chunk, chunkbox, boxchunk(nc + 1, 0);
  for (int64_t i = 0; i < npres; ++i) slot[i] = order[i].second;
  for (int64_t i = 0; i < npres;) {
    const int c = order[i].first;
    int64_t e = i;
    while (e < npres && order[e].first == c) ++e;
    boxchunk[c] = (int32_t)chunkbox.size();
    for (int64_t s = i; s < e; s += kChunk) {
      chunk.push_back((int32_t)s);
      chunkbox.push_back(c);
    }
    i = e;
  }
  chunk.push_back((int32_t)npres);
  boxchunk[nc] = (int32_t)chunkbox.size();
  std::vector<int32_t> boxnb((size_t)nc * kOffs, -1);
  const int ky = nd > 1 ? k : 1, kz = nd > 2 ? k : 1;
  for (int c = 0; c < nc; ++c) {
    const int32_t pos = boxes[c];
    const int bx = pos & 1023, by = (pos >> 10) & 1023, bz = (pos >> 20) & 1023;
    for (int o = 0; o < kOffs; ++o) {
      const int x = bx + o % 3 - 1, y = by + (o / 3) % 3 - 1, z = bz + o / 9 - 1;
      if (x < 0 || y < 0 || z < 0 || x >= k || y >= ky || z >= kz) continue;
      boxnb[(size_t)c * kOffs + o] = coarse_of((int64_t)cgrp[c] << 30 | (x | y << 10 | z << 20));
    }
  }
  // per-mesh blocks: unknown ranges, workspace and inverse offsets, solve records
  std::vector<int32_t> goff(ngroup + 1, 0);
  for (int c = 0; c < nc; ++c) goff[cgrp[c] + 1]++;
  for (int g = 0; g < ngroup; ++g) goff[g + 1] += goff[g];
  std::vector<int64_t> gjbase(ngroup), ebase(ngroup);
  std::vector<int2> swg;
  int64_t gjtot = 0, etot = 0;
  int maxng = 0;
  for (int g = 0; g < ngroup; ++g) {
    const int ng = goff[g + 1] - goff[g];
    gjbase[g] = gjtot;
    ebase[g] = etot;
    gjtot += 2 * (int64_t)ng * ng;
    etot += (int64_t)ng * ng;
    maxng = std::max(maxng, ng);
    for (int r0 = 0; r0 < ng; r0 += 4) swg.push_back(int2{g, r0});
  }
  // A Z's sparsity: per block row the coarse unknowns of its neighbours' pressure DOFs
  std::vector<int4> hbd(nb);
  std::vector<int32_t> hbcol(P.npairs);
  CO_CHECK(hipMemcpy(hbd.data(), P.d_bdesc, nb * sizeof(int4), hipMemcpyDeviceToHost));
  CO_CHECK(hipMemcpy(hbcol.data(), P.d_bcol, P.npairs * sizeof(int32_t), hipMemcpyDeviceToHost));
  std::vector<int32_t> azoff(nb + 1, 0), azbox;
  azbox.reserve(nb + nb / 2);
  for (int64_t p = 0; p < nb; ++p) {
    int32_t cs[64];
    int ncs = 0;
    for (int t = 0; t < hbd[p].z; ++t) {
      const int32_t info = cinfo[hbcol[hbd[p].y + t]];
      if (info < 0) continue;
      const int32_t c = info >> 3;
      bool seen = false;
      for (int u = 0; u < ncs; ++u) seen = seen || cs[u] == c;
      if (!seen && ncs < 64) cs[ncs++] = c;
    }
    std::sort(cs, cs + ncs);
    azbox.insert(azbox.end(), cs, cs + ncs);
    azoff[p + 1] = (int32_t)azbox.size();
  }
  int rc;
  if ((rc = upload(&C->d_azoff, azoff)) || (rc = upload(&C->d_azbox, azbox))) return rc;
  C->naz = (int64_t)azbox.size();
  if ((rc = upload(&C->d_cgrp, cgrp)) || (rc = upload(&C->d_goff, goff)) || (rc = upload(&C->d_gjbase, gjbase)) ||
      (rc = upload(&C->d_ebase, ebase)) || (rc = upload(&C->d_swg, swg)) || (rc = upload(&C->d_gfloat, gfloat)))
    return rc;
  C->any_float = std::find(gfloat.begin(), gfloat.end(), 1) != gfloat.end();
  C->ngroup = ngroup;
  C->h_goff = goff;
  C->h_ebase = ebase;
  C->nswg = (int)swg.size();
  C->maxng = maxng;
  if ((rc = upload(&C->d_slot, slot)) || (rc = upload(&C->d_chunk, chunk)) || (rc = upload(&C->d_chunkbox, chunkbox)) ||
      (rc = upload(&C->d_boxchunk, boxchunk)) || (rc = upload(&C->d_cinfo, cinfo)) || (rc = upload(&C->d_cpos, boxes)) ||
      (rc = upload(&C->d_boxnb, boxnb)))
    return rc;
  C->nc = nc;
  C->npres = npres;
  C->nchunk = (int)chunkbox.size();
  void* old[] = {C->d_epart, C->d_gj, C->d_einv, C->d_einvb, C->d_cpart, C->d_y, C->d_flag, C->d_az, C->d_tmp};
  C->d_einvb = nullptr;
  for (void* q : old)
    if (q) (void)hipFree(q);
  CO_CHECK(hipMalloc((void**)&C->d_epart, (size_t)C->nchunk * kOffs * sizeof(double)));
  CO_CHECK(hipMalloc((void**)&C->d_gj, std::max<size_t>((size_t)gjtot, 1) * sizeof(double)));
  CO_CHECK(hipMalloc((void**)&C->d_einv, std::max<size_t>((size_t)etot, 1) * sizeof(double)));
  CO_CHECK(hipMalloc((void**)&C->d_cpart, (size_t)C->nchunk * sizeof(double)));
  CO_CHECK(hipMalloc((void**)&C->d_y, (size_t)nc * sizeof(double)));
  CO_CHECK(hipMalloc((void**)&C->d_flag, ngroup * sizeof(int32_t)));
  CO_CHECK(hipMalloc(&C->d_az, std::max<size_t>((size_t)C->naz, 1) * 4 * sizeof(double)));
  CO_CHECK(hipMalloc(&C->d_tmp, (size_t)P.npad * sizeof(double)));
  CO_CHECK(hipMemset(C->d_tmp, 0, (size_t)P.npad * sizeof(double)));
  C->plan_id = L->plan_id;
  return SFEM_OK;
}

int coarse_set(sfem_ctx* ctx, int k, int mult) {
  LinAlg* L = ctx->la;
  L->coarse_mult = k > 0 && mult != 0;
  if (k == 0) {
    delete L->coarse;
    L->coarse = nullptr;
    L->coarse_k = 0;
    return SFEM_OK;
  }
  int rc = coarse_check(ctx, k);
  if (rc) return rc;
  // planned again at the next preconditioner build (boxes, meshes and form may differ)
  delete L->coarse;
  L->coarse = nullptr;
  L->coarse_k = k;
  return SFEM_OK;
}

// E and E^{-1} from the plan's block tangent (FP64 kp or FP32 kp32); called after each
// ILU(0) factorisation.
int coarse_build(sfem_ctx* ctx, const double* kp, const float* kp32) {
  LinAlg* L = ctx->la;
  if (L->coarse_k <= 0) return SFEM_OK;
  if (!L->coarse) {
    L->coarse = new Coarse;
    L->coarse->k = L->coarse_k;
  }
  Coarse* C = L->coarse;
  C->ready = false;
  if (C->plan_id != L->plan_id) {
    int rc = coarse_plan(ctx, L, C);
    if (rc) return rc;
  }
  const IluPlan& P = *L->ilu;
  hipStream_t st = ctx->stream;
  const int4* bd = reinterpret_cast<const int4*>(P.d_bdesc);
  if (kp32)
    hipLaunchKernelGGL((k_coarse_rows<float>), dim3(C->nchunk), dim3(64), 0, st, C->d_slot, C->d_chunk, bd, P.d_bcol,
                       kp32, C->d_cinfo, C->d_cpos, C->d_epart);
  else
    hipLaunchKernelGGL((k_coarse_rows<double>), dim3(C->nchunk), dim3(64), 0, st, C->d_slot, C->d_chunk, bd, P.d_bcol,
                       kp, C->d_cinfo, C->d_cpos, C->d_epart);
  const int nc = C->nc;
  hipLaunchKernelGGL(k_coarse_fill, dim3((nc + 63) / 64), dim3(64), 0, st, nc, C->d_cgrp, C->d_goff, C->d_gjbase,
                     C->d_boxchunk, C->d_boxnb, C->d_epart, C->d_gj);
  if (C->any_float)
    hipLaunchKernelGGL(k_coarse_float, dim3(C->ngroup), dim3(256), 0, st, C->d_goff, C->d_gjbase, C->d_gfloat, C->d_gj);
  if (C->ngroup == 1)  // one large block: one launch per step, one workgroup per row
    for (int j = 0; j < nc; ++j) hipLaunchKernelGGL(k_gj_step, dim3(nc), dim3(256), 0, st, nc, j, C->d_gj);
  else                 // a batch's small blocks: one workgroup per mesh
    hipLaunchKernelGGL(k_gj_group, dim3(C->ngroup), dim3(256), 0, st, C->d_goff, C->d_gjbase, C->d_gj);
  // FP32 Krylov vectors: their correction leaves the global constant pressure to the FP64
  // coarse step of the Newton loop
  const int border = kp32 != nullptr;
  const size_t etot = (size_t)(C->h_ebase.back() + (int64_t)(C->h_goff[C->ngroup] - C->h_goff[C->ngroup - 1]) *
                                                       (C->h_goff[C->ngroup] - C->h_goff[C->ngroup - 1]));
  if (border && !C->d_einvb) CO_CHECK(hipMalloc((void**)&C->d_einvb, std::max<size_t>(etot, 1) * sizeof(double)));
  hipLaunchKernelGGL(k_coarse_finish, dim3(C->ngroup), dim3(256), 0, st, border, C->d_gfloat, C->d_goff, C->d_gjbase,
                     C->d_ebase, C->d_gj, C->d_einv, C->d_einvb, C->d_flag);
  C->fp32 = border;
  CO_CHECK(hipGetLastError());
  C->ready = true;
  return SFEM_OK;
}

// A Z of the multiplicative form from the block tangent the Krylov products use (kp32 if
// given, else kp), after every new tangent: the fused form's corrected product must be the
// current operator's, even when the preconditioner (E) is lagged.
int coarse_refresh_az(sfem_ctx* ctx, const double* kp, const float* kp32) {
  LinAlg* L = ctx->la;
  Coarse* C = L->coarse;
  if (!L->coarse_mult || !C || C->plan_id != L->plan_id) return SFEM_OK;
  const IluPlan& P = *L->ilu;
  const int4* bd = reinterpret_cast<const int4*>(P.d_bdesc);
  const unsigned gz = (unsigned)((P.nblk + 255) / 256);
  if (kp32)
    hipLaunchKernelGGL((k_coarse_az_build<float>), dim3(gz), dim3(256), 0, ctx->stream, P.nblk, bd, P.d_bcol, kp32,
                       C->d_cinfo, C->d_azoff, C->d_azbox, (float*)C->d_az);
  else
    hipLaunchKernelGGL((k_coarse_az_build<double>), dim3(gz), dim3(256), 0, ctx->stream, P.nblk, bd, P.d_bcol, kp,
                       C->d_cinfo, C->d_azoff, C->d_azbox, (double*)C->d_az);
  C->az_fp32 = kp32 != nullptr;
  CO_CHECK(hipGetLastError());
  return SFEM_OK;
}

// The arrays a gather that forms A Z itself needs (false when the correction is additive
// or not yet planned for the current ILU(0) plan).
bool coarse_az_targets(const LinAlg* L, const int32_t** cinfo, const int32_t** azoff, const int32_t** azbox,
                       void** az) {
  const Coarse* C = L->coarse;
  if (!L->coarse_mult || !C || C->plan_id != L->plan_id) return false;
  *cinfo = C->d_cinfo;
  *azoff = C->d_azoff;
  *azbox = C->d_azbox;
  *az = C->d_az;
  return true;
}

void coarse_set_az_precision(LinAlg* L, bool fp32) {
  if (L->coarse) L->coarse->az_fp32 = fp32;
}

bool coarse_active(const LinAlg* L) { return L->coarse && L->coarse->ready && L->coarse->plan_id == L->plan_id; }

template <typename T>
int coarse_restrict(sfem_ctx* ctx, const T* r, const T* w) {
  Coarse* C = ctx->la->coarse;
  hipLaunchKernelGGL((k_coarse_restrict<T>), dim3(C->nchunk), dim3(64), 0, ctx->stream, C->d_slot, C->d_chunk, r, w,
                     C->d_cpart);
  return SFEM_OK;
}

// The multiplicative form's coarse solve and tail: y = E^{-1} c (c from the last
// restriction), then z += Z y and out += A Z y in one pass.
template <typename T>
int coarse_correct_az(sfem_ctx* ctx, T* z, T* out) {
  Coarse* C = ctx->la->coarse;
  const double* inv = sizeof(T) == 4 && C->d_einvb && C->fp32 ? C->d_einvb : C->d_einv;
  hipLaunchKernelGGL(k_coarse_solve, dim3(C->nswg), dim3(256), 0, ctx->stream, C->d_swg, C->d_goff, C->d_ebase,
                     C->d_boxchunk, C->d_cpart, inv, C->d_y);
  const int64_t nb = ctx->la->ilu->nblk;
  const unsigned g = (unsigned)((nb + 255) / 256);
  if (C->az_fp32)
    hipLaunchKernelGGL((k_coarse_post<T, float>), dim3(g), dim3(256), 0, ctx->stream, nb, C->d_cinfo, C->d_azoff,
                       C->d_azbox, (const float*)C->d_az, C->d_y, z, out);
  else
    hipLaunchKernelGGL((k_coarse_post<T, double>), dim3(g), dim3(256), 0, ctx->stream, nb, C->d_cinfo, C->d_azoff,
                       C->d_azbox, (const double*)C->d_az, C->d_y, z, out);
  return SFEM_OK;
}

template int coarse_correct_az<float>(sfem_ctx*, float*, float*);
template int coarse_correct_az<double>(sfem_ctx*, double*, double*);

template <typename T>
T* coarse_tmp(sfem_ctx* ctx) {
  return (T*)ctx->la->coarse->d_tmp;
}

template float* coarse_tmp<float>(sfem_ctx*);
template double* coarse_tmp<double>(sfem_ctx*);

template <typename T>
int coarse_correct(sfem_ctx* ctx, T* z) {
  Coarse* C = ctx->la->coarse;
  const double* inv = sizeof(T) == 4 && C->d_einvb && C->fp32 ? C->d_einvb : C->d_einv;
  hipLaunchKernelGGL(k_coarse_solve, dim3(C->nswg), dim3(256), 0, ctx->stream, C->d_swg, C->d_goff, C->d_ebase,
                     C->d_boxchunk, C->d_cpart, inv, C->d_y);
  hipLaunchKernelGGL((k_coarse_prolong<T>), dim3(C->nchunk), dim3(64), 0, ctx->stream, C->d_slot, C->d_chunk,
                     C->d_chunkbox, C->d_y, z);
  return SFEM_OK;
}

template int coarse_restrict<float>(sfem_ctx*, const float*, const float*);
template int coarse_restrict<double>(sfem_ctx*, const double*, const double*);
template int coarse_correct<float>(sfem_ctx*, float*);
template int coarse_correct<double>(sfem_ctx*, double*);

int coarse_info(sfem_ctx* ctx, int64_t* out) {
  LinAlg* L = ctx->la;
  const Coarse* C = L->coarse;
  out[0] = L->coarse_k;
  out[1] = C ? C->nc : 0;
  out[2] = C ? C->npres : 0;
  out[3] = C ? C->nchunk : 0;
  int32_t flag = 0;
  if (C && C->d_flag && C->ready) {
    std::vector<int32_t> f(C->ngroup);
    CO_CHECK(hipStreamSynchronize(ctx->stream));
    CO_CHECK(hipMemcpy(f.data(), C->d_flag, f.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int32_t v : f) flag |= v;
  }
  out[4] = flag;
  out[5] = C && C->ready ? 1 : 0;
  return SFEM_OK;
}

// E^{-1} (which = 0) or the FP32 vectors' inverse without the global constant (which = 1),
// nc x nc row-major (block diagonal for a batch, one block per mesh), of the last build.
int coarse_get_inverse(sfem_ctx* ctx, int which, double* out, int cap) {
  const Coarse* C = ctx->la->coarse;
  if (!C || !C->ready) return set_error(SFEM_ERR_STATE, "coarse correction not built");
  if (which == 1 && !(C->fp32 && C->d_einvb)) return set_error(SFEM_ERR_STATE, "no FP32 coarse inverse");
  const int64_t nc = C->nc;
  if (cap < nc * nc) return set_error(SFEM_ERR_ARG, "buffer too small");
  const int64_t ng1 = C->h_goff[C->ngroup] - C->h_goff[C->ngroup - 1];
  std::vector<double> blk((size_t)(C->h_ebase.back() + ng1 * ng1));
  CO_CHECK(hipStreamSynchronize(ctx->stream));
  CO_CHECK(hipMemcpy(blk.data(), which ? C->d_einvb : C->d_einv, blk.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::fill(out, out + nc * nc, 0.0);
  for (int g = 0; g < C->ngroup; ++g) {
    const int g0 = C->h_goff[g], ng = C->h_goff[g + 1] - g0;
    for (int i = 0; i < ng; ++i)
      for (int j = 0; j < ng; ++j) out[(g0 + i) * nc + g0 + j] = blk[C->h_ebase[g] + (int64_t)i * ng + j];
  }
  return SFEM_OK;
}

}  // namespace sfem
