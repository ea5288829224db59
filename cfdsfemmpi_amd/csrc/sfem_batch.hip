// sfem_batch.hip -- many independent meshes solved together in one context.
//
// The meshes are one disjoint union (nodes and elements mesh by mesh), so assembly,
// the tangent, its node-block form and the ILU(0) factor are block diagonal and every
// existing kernel handles all meshes in one launch.  What is per mesh: the viscosity
// (a per-element array read by the element kernel), the Newton test (each mesh stops
// at ||F_m|| < tol, StabFEM2.cpp:740-749, and is frozen from then on), and the Krylov
// solve: BiCGStab with per-mesh scalars (alpha, omega, rho, ...), per-mesh convergence
// ||r_m|| <= rtol ||b_m|| and per-mesh breakdown restarts, so that each mesh follows
// the iteration it would follow alone.  Reductions are per mesh and deterministic
// (fixed chunk order).  A time step runs its Newton iterations as HIP graphs
// (step_batch_graph): the residual assembly, per-mesh norms and Newton test, the start
// and end of each solve and the solution update are captured once per step, the BiCGStab
// iterations of all meshes replay in graphs of a few iterations, and the host reads one
// status word per replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sfem_internal.hpp"
#include "sfem_linalg.hpp"
#include "sfem_state.hpp"

namespace sfem {

namespace {
constexpr int kTB = 256;          // threads per block of the per-chunk kernels
constexpr int kChunk = 4096;      // Krylov slots per chunk (at most)
// per-mesh Krylov scalars
enum {
  B_RHO0 = 0, B_RHO1, B_ALPHA, B_OMEGA, B_BETA, B_RR, B_BB, B_TOL2, B_RHRH, B_SS,
  B_FIRST, B_DONE, B_HALF, B_RESTART, B_ITS, B_BAD,
  B_BEST, B_BESTIT, B_SAVED, B_SAVE, B_RESTORE, B_STALL, B_N = 24
};
constexpr int kStallWindow = 100;  // FP32: no 10% residual gain in this many iterations
}  // namespace

struct Batch {
  int nmesh = 0;
  std::vector<int64_t> dof_ptr;   // free-DOF range of each mesh (DOF order)
  std::vector<int32_t> node_mesh; // mesh of every node (-1: in no element)
  int64_t* d_dof_ptr = nullptr;
  int plan_id = -1;               // ILU plan the chunk lists were built for
  int nchunk = 0;
  int4* d_chunk = nullptr;        // {slot begin, slot end, mesh, chunk id}
  int4* d_cchunk = nullptr;       // chunks of the meshes still iterating (capacity nchunk)
  std::vector<int4> h_chunk;
  std::vector<int32_t> blk_mesh;  // mesh of every block row (plan position)
  std::vector<int64_t> mesh_nblk; // block rows of every mesh
  std::vector<uint8_t> keep;      // scratch: block rows of the next row subset
  RowSubset sub[2];               // row subsets of a shrinking solve, used alternately
  int64_t subsets_built = 0;
  int32_t* d_mptr = nullptr;      // chunks of mesh m: d_mlist[d_mptr[m] .. d_mptr[m+1])
  int32_t* d_mlist = nullptr;
  double* d_sc = nullptr;         // [nmesh][B_N]
  double* d_part = nullptr;       // [nchunk][2]
  double* d_norm = nullptr;       // per-mesh ||F_m||
  double* h_norm = nullptr;       // pinned
  int32_t* d_active = nullptr;    // per-mesh Newton activity
  int32_t* d_ndone = nullptr;     // finished meshes of the current solve
  int32_t* h_ndone = nullptr;     // pinned
  double* h_sc = nullptr;         // pinned copy of d_sc
  void* vec[11] = {};             // r, rh, p, v, s, t, ph, sh, x, b, best x
  int64_t stalled = 0;            // FP32 mesh solves ended on a stall (best iterate kept)
  size_t vec_bytes = 0;
  bool use_graph = true;
  bool narrowing = true;
  // graph-captured Newton step: per-mesh Newton iterations, Krylov iterations, residual
  // norms and the step status (k_b_step_reset), on the device
  int32_t* d_newton = nullptr;
  int32_t* d_kry = nullptr;
  double* d_rhsn = nullptr;
  int32_t* d_status = nullptr;
  int32_t* h_status = nullptr;    // pinned
  int32_t* d_cmask = nullptr;     // per mesh: takes the FP64 coarse step this Newton iteration
  int its_per_graph = 4;          // BiCGStab iterations per replayed graph in the graph-captured step
  ~Batch() {
    void* ptrs[] = {d_dof_ptr, d_chunk, d_cchunk, d_mptr, d_mlist, d_sc, d_part, d_norm, d_active, d_ndone,
                    d_newton, d_kry, d_rhsn, d_status, d_cmask};
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
    for (void* q : vec)
      if (q) (void)hipFree(q);
    if (h_norm) (void)hipHostFree(h_norm);
    if (h_ndone) (void)hipHostFree(h_ndone);
    if (h_sc) (void)hipHostFree(h_sc);
    if (h_status) (void)hipHostFree(h_status);
  }
};

void batch_destroy(sfem_ctx* ctx) {
  delete ctx->batch;
  ctx->batch = nullptr;
}

int64_t batch_stalled(const sfem_ctx* ctx) { return ctx->batch ? ctx->batch->stalled : 0; }

const std::vector<int32_t>* batch_node_mesh(const sfem_ctx* ctx) {
  return ctx->batch ? &ctx->batch->node_mesh : nullptr;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__device__ inline double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
  __syncthreads();
  return s;  // valid in thread 0
}

// ||a||^2 over each mesh's DOF range, one block per mesh
__global__ __launch_bounds__(kTB) void k_seg_norm2(const int64_t* __restrict__ ptr, const double* __restrict__ a,
                                                   double* __restrict__ out) {
  __shared__ double sh[kTB / 64];
  const int m = blockIdx.x;
  double s = 0.0;
  for (int64_t i = ptr[m] + threadIdx.x; i < ptr[m + 1]; i += kTB) s += a[i] * a[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[m] = sqrt(s);
}

// chunk-parallel kernels: block b handles slots [c.x, c.y) of mesh c.z, c = ch[b], and
// writes its reduction partials to part[2 c.w], c.w being the chunk's id in the full list
template <typename T>
__global__ __launch_bounds__(kTB) void k_b_pad_in(const int4* __restrict__ ch, const int32_t* __restrict__ padmap,
                                                  const double* __restrict__ rhs, const double* __restrict__ norm,
                                                  const int32_t* __restrict__ active, T* __restrict__ b) {
  const int4 c = ch[blockIdx.x];
  const double nm = norm[c.z];
  const double sc = (active[c.z] && nm > 0.0) ? 1.0 / nm : 0.0;
  for (int i = c.x + threadIdx.x; i < c.y; i += kTB) {
    const int32_t k = padmap[i];
    b[i] = k >= 0 ? (T)(rhs[k] * sc) : (T)0;
  }
}

template <typename T>
__global__ __launch_bounds__(kTB) void k_b_pad_out(const int4* __restrict__ ch, const int32_t* __restrict__ padmap,
                                                   const T* __restrict__ x, const double* __restrict__ norm,
                                                   const int32_t* __restrict__ active, double* __restrict__ du) {
  const int4 c = ch[blockIdx.x];
  const double sc = active[c.z] ? norm[c.z] : 0.0;
  for (int i = c.x + threadIdx.x; i < c.y; i += kTB) {
    const int32_t k = padmap[i];
    if (k >= 0) du[k] = (double)x[i] * sc;
  }
}

// x = 0, r = rh = b, p = v = 0; partial (b, b)
template <typename T>
__global__ __launch_bounds__(kTB) void k_b_init(const int4* __restrict__ ch, const T* __restrict__ b, T* __restrict__ x,
                                                T* __restrict__ r, T* __restrict__ rh, T* __restrict__ p,
                                                T* __restrict__ v, T* __restrict__ xb, double* __restrict__ part) {
  __shared__ double sh[kTB / 64];
  const int4 c = ch[blockIdx.x];
  double s = 0.0;
  for (int i = c.x + threadIdx.x; i < c.y; i += kTB) {
    const T bi = b[i];
    x[i] = (T)0; r[i] = bi; rh[i] = bi; p[i] = (T)0; v[i] = (T)0; xb[i] = (T)0;
    s += (double)bi * (double)bi;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) { part[2 * c.w] = s; part[2 * c.w + 1] = 0.0; }
}

// p = r + beta (p - omega v), or p = r on a (re)start.  The chunks of a finished mesh
// return at once here and in k_b_s / k_b_xr: its x and r stay as converged, and its p
// does not keep growing under a stale beta while other meshes iterate (it would
// overflow, and x += 0 * inf is NaN).
template <typename T>
__global__ __launch_bounds__(kTB) void k_b_p(const int4* __restrict__ ch, const double* __restrict__ sc,
                                             const T* __restrict__ r, T* __restrict__ p, const T* __restrict__ v) {
  const int4 c = ch[blockIdx.x];
  const double* S = sc + (int64_t)c.z * B_N;
  if (S[B_DONE] != 0.0) return;  // finished meshes keep their vectors
  const bool first = S[B_FIRST] != 0.0;
  const T beta = first ? (T)0 : (T)S[B_BETA], om = first ? (T)0 : (T)S[B_OMEGA];
  for (int i = c.x + threadIdx.x; i < c.y; i += kTB) p[i] = first ? r[i] : r[i] + beta * (p[i] - om * v[i]);
}

// partials (a1, b1), (a2, b2)
template <typename T>
__global__ __launch_bounds__(kTB) void k_b_dot2(const int4* __restrict__ ch, const T* __restrict__ a1,
                                                const T* __restrict__ b1, const T* __restrict__ a2,
                                                const T* __restrict__ b2, double* __restrict__ part) {
  __shared__ double sh[kTB / 64];
  const int4 c = ch[blockIdx.x];
  double s1 = 0.0, s2 = 0.0;
  for (int i = c.x + threadIdx.x; i < c.y; i += kTB) {
    s1 += (double)a1[i] * (double)b1[i];
    s2 += (double)a2[i] * (double)b2[i];
  }
  s1 = block_sum(s1, sh);
  s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) { part[2 * c.w] = s1; part[2 * c.w + 1] = s2; }
}

// s = r - alpha v ; partial (s, s)
template <typename T>
__global__ __launch_bounds__(kTB) void k_b_s(const int4* __restrict__ ch, const double* __restrict__ sc,
                                             const T* __restrict__ r, const T* __restrict__ v, T* __restrict__ s,
                                             double* __restrict__ part) {
  __shared__ double sh[kTB / 64];
  const int4 c = ch[blockIdx.x];
  if (sc[(int64_t)c.z * B_N + B_DONE] != 0.0) {
    if (threadIdx.x == 0) { part[2 * c.w] = 0.0; part[2 * c.w + 1] = 0.0; }
    return;
  }
  const T alpha = (T)sc[(int64_t)c.z * B_N + B_ALPHA];
  double a = 0.0;
  for (int i = c.x + threadIdx.x; i < c.y; i += kTB) {
    const T si = r[i] - alpha * v[i];
    s[i] = si;
    a += (double)si * (double)si;
  }
  a = block_sum(a, sh);
  if (threadIdx.x == 0) { part[2 * c.w] = a; part[2 * c.w + 1] = 0.0; }
}

// x += alpha ph + omega sh ; r = s - omega t ; partials (r, r), (rh, r)
template <typename T>
__global__ __launch_bounds__(kTB) void k_b_xr(const int4* __restrict__ ch, const double* __restrict__ sc,
                                              T* __restrict__ x, const T* __restrict__ ph, const T* __restrict__ shv,
                                              T* __restrict__ r, const T* __restrict__ s, const T* __restrict__ t,
                                              const T* __restrict__ rh, double* __restrict__ part) {
  __shared__ double sh[kTB / 64];
  const int4 c = ch[blockIdx.x];
  const double* S = sc + (int64_t)c.z * B_N;
  if (S[B_DONE] != 0.0) {
    if (threadIdx.x == 0) { part[2 * c.w] = 0.0; part[2 * c.w + 1] = 0.0; }
    return;
  }
  const T alpha = (T)S[B_ALPHA], om = (T)S[B_OMEGA];
  double a = 0.0, b = 0.0;
  for (int i = c.x + threadIdx.x; i < c.y; i += kTB) {
    x[i] += alpha * ph[i] + om * shv[i];
    const T ri = s[i] - om * t[i];
    r[i] = ri;
    a += (double)ri * (double)ri;
    b += (double)rh[i] * (double)ri;
  }
  a = block_sum(a, sh);
  b = block_sum(b, sh);
  if (threadIdx.x == 0) { part[2 * c.w] = a; part[2 * c.w + 1] = b; }
}

// end of an iteration, per mesh as flagged: restart of a broken-down mesh (rh = r),
// saving a new best iterate (xb = x) or, for a stalled FP32 mesh, taking it back (x = xb)
template <typename T>
__global__ __launch_bounds__(kTB) void k_b_post(const int4* __restrict__ ch, const double* __restrict__ sc,
                                                const T* __restrict__ r, T* __restrict__ rh, T* __restrict__ x,
                                                T* __restrict__ xb) {
  const int4 c = ch[blockIdx.x];
  const double* S = sc + (int64_t)c.z * B_N;
  if (S[B_RESTART] != 0.0)
    for (int i = c.x + threadIdx.x; i < c.y; i += kTB) rh[i] = r[i];
  if (S[B_SAVE] != 0.0)
    for (int i = c.x + threadIdx.x; i < c.y; i += kTB) xb[i] = x[i];
  if (S[B_RESTORE] != 0.0)
    for (int i = c.x + threadIdx.x; i < c.y; i += kTB) x[i] = xb[i];
}

enum { F_INIT = 0, F_ALPHA, F_SS, F_OMEGA, F_RR_RHO, F_GIVEUP };

// Per-mesh scalar updates from the chunk partials (summed in chunk order); one thread
// per mesh, one block.  F_RR_RHO also counts the finished meshes.
__global__ __launch_bounds__(kTB) void k_b_final(int nmesh, int op, const int32_t* __restrict__ mptr,
                                                 const int32_t* __restrict__ mlist, const double* __restrict__ part,
                                                 double* __restrict__ sc, double rtol, int stall,
                                                 int32_t* __restrict__ ndone) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int mine = 0;
  for (int m = threadIdx.x; m < nmesh; m += blockDim.x) {
    double a = 0.0, b = 0.0;
    for (int k = mptr[m]; k < mptr[m + 1]; ++k) {
      a += part[2 * mlist[k]];
      b += part[2 * mlist[k] + 1];
    }
    double* S = sc + (int64_t)m * B_N;
    switch (op) {
      case F_INIT:
        for (int q = 0; q < B_N; ++q) S[q] = 0.0;
        S[B_BB] = a; S[B_RR] = a; S[B_RHRH] = a; S[B_RHO1] = a;
        S[B_RHO0] = 1.0; S[B_ALPHA] = 1.0; S[B_OMEGA] = 1.0; S[B_FIRST] = 1.0;
        S[B_TOL2] = rtol * rtol * a;
        S[B_DONE] = (a > 0.0) ? 0.0 : 1.0;
        S[B_BEST] = a; S[B_SAVED] = a;
        break;
      case F_ALPHA:  // a = (rh, v), b = (v, v)
        S[B_RESTART] = 0.0;
        if (S[B_DONE] != 0.0 || !(fabs(a) > 1e-12 * sqrt(S[B_RHRH] * b))) {
          S[B_ALPHA] = 0.0;
          if (S[B_DONE] == 0.0) S[B_RESTART] = 1.0;  // near breakdown: restart after this step
        } else {
          S[B_ALPHA] = S[B_RHO1] / a;
        }
        break;
      case F_SS:  // a = (s, s)
        S[B_SS] = a;
        S[B_HALF] = (S[B_DONE] == 0.0 && a <= S[B_TOL2]) ? 1.0 : 0.0;
        break;
      case F_GIVEUP:  // iteration limit: unfinished meshes take their best iterate back
        S[B_SAVE] = 0.0;
        S[B_RESTART] = 0.0;
        S[B_RESTORE] = S[B_DONE] == 0.0 ? 1.0 : 0.0;
        if (S[B_DONE] == 0.0) { S[B_STALL] = 1.0; S[B_DONE] = 1.0; }
        break;
      case F_OMEGA:  // a = (t, t), b = (t, s)
        S[B_OMEGA] = (S[B_DONE] != 0.0 || S[B_HALF] != 0.0 || a == 0.0) ? 0.0 : b / a;
        break;
      case F_RR_RHO: {  // a = (r, r), b = (rh, r)
        S[B_SAVE] = 0.0;
        S[B_RESTORE] = 0.0;
        if (S[B_DONE] == 0.0) {
          S[B_ITS] += 1.0;
          S[B_RR] = a;
          if (a < 0.81 * S[B_BEST]) { S[B_BEST] = a; S[B_BESTIT] = S[B_ITS]; }
          if (!isfinite(a) && !stall) {
            S[B_BAD] = 1.0;
            S[B_DONE] = 1.0;
          } else if (S[B_HALF] != 0.0 || a <= S[B_TOL2]) {
            S[B_DONE] = 1.0;
          } else if (!isfinite(a) || (stall && S[B_ITS] - S[B_BESTIT] >= stall)) {
            // stalled or diverged FP32 solve: keep the best iterate saved so far (the
            // outer FP64 Newton iteration corrects from there)
            S[B_STALL] = 1.0;
            S[B_RESTORE] = 1.0;
            S[B_DONE] = 1.0;
          } else {
            if (stall && a < 0.25 * S[B_SAVED]) { S[B_SAVED] = a; S[B_SAVE] = 1.0; }
            const double rho0 = S[B_RHO1], om = S[B_OMEGA];
            if (S[B_RESTART] != 0.0 || !(fabs(b) > 1e-12 * sqrt(S[B_RHRH] * a)) || rho0 == 0.0 || om == 0.0) {
              // restart with the current residual as shadow residual (rh = r follows)
              S[B_RESTART] = 1.0;
              S[B_RHO0] = 1.0; S[B_ALPHA] = 1.0; S[B_OMEGA] = 1.0;
              S[B_RHRH] = a; S[B_RHO1] = a; S[B_FIRST] = 1.0;
            } else {
              S[B_RHO0] = rho0;
              S[B_RHO1] = b;
              S[B_BETA] = (b / rho0) * (S[B_ALPHA] / om);
              S[B_FIRST] = 0.0;
            }
          }
        }
        if (S[B_DONE] != 0.0) { S[B_ALPHA] = 0.0; S[B_OMEGA] = 0.0; S[B_RESTART] = 0.0; }
        mine += S[B_DONE] != 0.0 ? 1 : 0;
        break;
      }
    }
  }
  if (op == F_RR_RHO || op == F_INIT) {
    if (op == F_INIT)
      for (int m = threadIdx.x; m < nmesh; m += blockDim.x) mine += sc[(int64_t)m * B_N + B_DONE] != 0.0 ? 1 : 0;
    atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) *ndone = cnt;
  }
}

// Per-step device state of the graph-captured Newton step: status = {finished meshes of
// the current solve (written by k_b_final), meshes still iterating Newton, first failed
// mesh (INT_MAX: none), Newton iterations done}.
enum { ST_NDONE = 0, ST_NACTIVE, ST_ERR, ST_ITER, ST_NSTALL, ST_N = 8 };

// start of a step: every mesh active, counters cleared
__global__ __launch_bounds__(kTB) void k_b_step_reset(int nmesh, int max_newton, int32_t* __restrict__ active,
                                                      int32_t* __restrict__ newton, int32_t* __restrict__ kry,
                                                      double* __restrict__ rhsn, int32_t* __restrict__ status) {
  for (int m = threadIdx.x; m < nmesh; m += blockDim.x) {
    active[m] = 1;
    newton[m] = max_newton;
    kry[m] = 0;
    rhsn[m] = 0.0;
  }
  if (threadIdx.x == 0) {
    status[ST_NDONE] = 0;
    status[ST_NACTIVE] = nmesh;
    status[ST_ERR] = 0x7fffffff;
    status[ST_ITER] = 0;
    status[ST_NSTALL] = 0;
  }
}

// Newton test of every active mesh on the device (||F_m|| < tol ends its iterations,
// StabFEM2.cpp:740-749): its residual norm and iteration count are recorded, and the
// meshes still iterating are counted.
__global__ __launch_bounds__(kTB) void k_b_activity(int nmesh, const double* __restrict__ norm, double tol,
                                                    int32_t* __restrict__ active, int32_t* __restrict__ newton,
                                                    double* __restrict__ rhsn, int32_t* __restrict__ status) {
  __shared__ int cnt;
  const int it = status[ST_ITER];
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int mine = 0;
  for (int m = threadIdx.x; m < nmesh; m += blockDim.x) {
    if (!active[m]) continue;
    rhsn[m] = norm[m];
    if (norm[m] < tol) {
      active[m] = 0;
      newton[m] = it + 1;
    } else {
      ++mine;
    }
  }
  atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) {
    status[ST_NACTIVE] = cnt;
    status[ST_ITER] = it + 1;
  }
}

// meshes that take the FP64 coarse step: still iterating and not yet below the tolerance
__global__ __launch_bounds__(kTB) void k_b_coarse_mask(int nmesh, const double* __restrict__ norm, double tol,
                                                       const int32_t* __restrict__ active, int32_t* __restrict__ cmask) {
  for (int m = threadIdx.x; m < nmesh; m += blockDim.x) cmask[m] = active[m] && !(norm[m] < tol);
}

// the coarse step's increment kept only for those meshes, one block per mesh
__global__ __launch_bounds__(kTB) void k_b_mask_du(const int64_t* __restrict__ ptr, const int32_t* __restrict__ cmask,
                                                   double* __restrict__ du) {
  const int m = blockIdx.x;
  if (cmask[m]) return;
  for (int64_t i = ptr[m] + threadIdx.x; i < ptr[m + 1]; i += kTB) du[i] = 0.0;
}

// end of a batched solve: Krylov iterations per mesh, stalled FP32 mesh solves, and the
// first mesh that broke down or did not converge
__global__ __launch_bounds__(kTB) void k_b_tail(int nmesh, const double* __restrict__ sc, int32_t* __restrict__ kry,
                                                int32_t* __restrict__ status) {
  for (int m = threadIdx.x; m < nmesh; m += blockDim.x) {
    const double* S = sc + (int64_t)m * B_N;
    kry[m] += (int32_t)S[B_ITS];
    if (S[B_STALL] != 0.0) atomicAdd(&status[ST_NSTALL], 1);
    if (S[B_BAD] != 0.0 || S[B_DONE] == 0.0) atomicMin(&status[ST_ERR], m);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#define BA_CHECK(x)                                   \
  do {                                                \
    hipError_t _e = (x);                              \
    if (_e != hipSuccess) return set_hip_error(_e, #x); \
  } while (0)
#define HIP_OK(x) BA_CHECK(x)

template <typename T>
static hipError_t balloc(T** p, size_t n) {
  *p = nullptr;
  return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}

// Chunks of the Krylov layout: consecutive node blocks of one mesh, at most kChunk slots.
static int build_chunks(sfem_ctx* ctx, Batch& B) {
  LinAlg* L = ctx->la;
  const IluPlan& P = *L->ilu;
  std::vector<int32_t> r0(P.nblk);
  BA_CHECK(hipMemcpy(r0.data(), P.d_r0orig, P.nblk * sizeof(int32_t), hipMemcpyDeviceToHost));
  auto mesh_of = [&](int32_t dof) {
    return (int)(std::upper_bound(B.dof_ptr.begin(), B.dof_ptr.end(), (int64_t)dof) - B.dof_ptr.begin()) - 1;
  };
  std::vector<int4> ch;
  B.blk_mesh.resize(P.nblk);
  B.mesh_nblk.assign(B.nmesh, 0);
  for (int64_t p = 0; p < P.nblk; ++p) {
    B.blk_mesh[p] = mesh_of(r0[p]);
    B.mesh_nblk[B.blk_mesh[p]]++;
  }
  int64_t p = 0;
  while (p < P.nblk) {
    const int m = B.blk_mesh[p];
    int64_t q = p + 1;
    while (q < P.nblk && q - p < kChunk / 4 && B.blk_mesh[q] == m) ++q;
    ch.push_back(int4{(int)(4 * p), (int)(4 * q), m, (int)ch.size()});
    p = q;
  }
  std::vector<int32_t> mptr(B.nmesh + 1, 0), mlist(ch.size());
  for (const int4& c : ch) mptr[c.z + 1]++;
  for (int m = 0; m < B.nmesh; ++m) mptr[m + 1] += mptr[m];
  std::vector<int32_t> pos(mptr.begin(), mptr.end() - 1);
  for (size_t k = 0; k < ch.size(); ++k) mlist[pos[ch[k].z]++] = (int32_t)k;
  for (void* q : {(void*)B.d_chunk, (void*)B.d_cchunk, (void*)B.d_mptr, (void*)B.d_mlist, (void*)B.d_part})
    if (q) (void)hipFree(q);
  B.nchunk = (int)ch.size();
  B.keep.assign(P.nblk, 0);
  BA_CHECK(balloc(&B.d_chunk, ch.size()));
  BA_CHECK(balloc(&B.d_cchunk, ch.size()));
  BA_CHECK(balloc(&B.d_mptr, mptr.size()));
  BA_CHECK(balloc(&B.d_mlist, mlist.size()));
  BA_CHECK(balloc(&B.d_part, 2 * ch.size()));
  BA_CHECK(hipMemcpy(B.d_chunk, ch.data(), ch.size() * sizeof(int4), hipMemcpyHostToDevice));
  BA_CHECK(hipMemcpy(B.d_mptr, mptr.data(), mptr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  BA_CHECK(hipMemcpy(B.d_mlist, mlist.data(), mlist.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  B.h_chunk.swap(ch);
  B.plan_id = L->plan_id;
  return SFEM_OK;
}

// One batched BiCGStab iteration (all launches on the context stream, no host work).
template <typename T>
static int bb_iteration(sfem_ctx* ctx, Batch& B, double rtol, int stall, const int4* ch, int nch,
                        const RowSubset* sub) {
  hipStream_t st = ctx->stream;
  T** V = (T**)B.vec;
  T *r = V[0], *rh = V[1], *p = V[2], *v = V[3], *s = V[4], *t = V[5], *ph = V[6], *sh = V[7], *x = V[8];
  T* xb = V[10];
  const dim3 g(nch), tb(kTB);
  // z = M^-1 a and y = A z (one fused operation with the multiplicative coarse correction)
  auto pcop = [&](const T* a, T* z, T* y) {
    return sizeof(T) == 8 ? la_apply_pc_op64(ctx, (const double*)a, (double*)z, (double*)y, sub)
                          : la_apply_pc_op32(ctx, (const float*)a, (float*)z, (float*)y, sub);
  };
  auto fin = [&](int o) {
    hipLaunchKernelGGL(k_b_final, dim3(1), dim3(kTB), 0, st, B.nmesh, o, B.d_mptr, B.d_mlist, B.d_part, B.d_sc,
                       rtol, stall, B.d_ndone);
  };
  int rc;
  hipLaunchKernelGGL((k_b_p<T>), g, tb, 0, st, ch, B.d_sc, r, p, v);
  if ((rc = pcop(p, ph, v))) return rc;
  hipLaunchKernelGGL((k_b_dot2<T>), g, tb, 0, st, ch, rh, v, v, v, B.d_part);
  fin(F_ALPHA);
  hipLaunchKernelGGL((k_b_s<T>), g, tb, 0, st, ch, B.d_sc, r, v, s, B.d_part);
  fin(F_SS);
  if ((rc = pcop(s, sh, t))) return rc;
  hipLaunchKernelGGL((k_b_dot2<T>), g, tb, 0, st, ch, t, t, t, s, B.d_part);
  fin(F_OMEGA);
  hipLaunchKernelGGL((k_b_xr<T>), g, tb, 0, st, ch, B.d_sc, x, ph, sh, r, s, t, rh, B.d_part);
  fin(F_RR_RHO);
  hipLaunchKernelGGL((k_b_post<T>), g, tb, 0, st, ch, B.d_sc, r, rh, x, xb);
  BA_CHECK(hipGetLastError());
  return SFEM_OK;
}

// Batched solve of A du = rhs with per-mesh right-preconditioned BiCGStab; meshes with
// active[m] = 0 get du_m = 0.  In pieces: solve_prepare (preconditioner and Krylov
// layout; host work), solve_begin and solve_end (kernel launches only, capturable), and
// solve_iterate (BiCGStab iterations of all meshes, replayed as graphs of `per_graph`
// iterations until every mesh is done; the host reads the count of finished meshes
// after each replay).  A finished mesh's vectors and scalars are left alone by every
// kernel of an iteration, so replaying more iterations than a mesh needs changes
// nothing for it.
template <typename T>
static int solve_prepare(sfem_ctx* ctx, const sfem_solver_opts* o) {
  Batch& B = *ctx->batch;
  LinAlg* L = ctx->la;
  int rc;
  if ((rc = la_prepare_solve(ctx, o))) return rc;
  if (!L->block_space) return set_error(SFEM_ERR_UNSUPPORTED, "batched solve: ILU(0) preconditioner required");
  if (B.plan_id != L->plan_id && (rc = build_chunks(ctx, B))) return rc;
  const size_t vb = (size_t)L->nvec * sizeof(T);
  if (B.vec_bytes != vb) {
    for (void*& q : B.vec)
      if (q) { (void)hipFree(q); q = nullptr; }
    for (void*& q : B.vec) BA_CHECK(hipMalloc(&q, vb));
    B.vec_bytes = vb;
  }
  return SFEM_OK;
}

template <typename T>
static int solve_begin(sfem_ctx* ctx, const sfem_solver_opts* o) {
  Batch& B = *ctx->batch;
  hipStream_t st = ctx->stream;
  T** V = (T**)B.vec;
  const dim3 g(B.nchunk), tb(kTB);
  const IluPlan& P = *ctx->la->ilu;
  const int stall = sizeof(T) == 4 ? kStallWindow : 0;
  hipLaunchKernelGGL((k_b_pad_in<T>), g, tb, 0, st, B.d_chunk, P.d_padmap, ctx->d_rhs, B.d_norm, B.d_active, V[9]);
  hipLaunchKernelGGL((k_b_init<T>), g, tb, 0, st, B.d_chunk, V[9], V[8], V[0], V[1], V[2], V[3], V[10], B.d_part);
  hipLaunchKernelGGL(k_b_final, dim3(1), dim3(kTB), 0, st, B.nmesh, (int)F_INIT, B.d_mptr, B.d_mlist, B.d_part,
                     B.d_sc, o->rtol, stall, B.d_ndone);
  BA_CHECK(hipGetLastError());
  return SFEM_OK;
}

// giveup: FP32 meshes still iterating take their best iterate back
template <typename T>
static int solve_end(sfem_ctx* ctx, const sfem_solver_opts* o, bool giveup) {
  Batch& B = *ctx->batch;
  hipStream_t st = ctx->stream;
  T** V = (T**)B.vec;
  const dim3 g(B.nchunk), tb(kTB);
  const IluPlan& P = *ctx->la->ilu;
  const int stall = sizeof(T) == 4 ? kStallWindow : 0;
  if (stall && giveup) {
    hipLaunchKernelGGL(k_b_final, dim3(1), dim3(kTB), 0, st, B.nmesh, (int)F_GIVEUP, B.d_mptr, B.d_mlist, B.d_part,
                       B.d_sc, o->rtol, stall, B.d_ndone);
    hipLaunchKernelGGL((k_b_post<T>), g, tb, 0, st, B.d_chunk, B.d_sc, V[0], V[1], V[8], V[10]);
  }
  hipLaunchKernelGGL((k_b_pad_out<T>), g, tb, 0, st, B.d_chunk, P.d_padmap, V[8], B.d_norm, B.d_active, ctx->d_du);
  BA_CHECK(hipGetLastError());
  return SFEM_OK;
}

// Iterations until every mesh is done or max_it; returns with *B.h_ndone current.
template <typename T>
static int solve_iterate(sfem_ctx* ctx, const sfem_solver_opts* o, int per_graph) {
  Batch& B = *ctx->batch;
  LinAlg* L = ctx->la;
  hipStream_t st = ctx->stream;
  const IluPlan& P = *L->ilu;
  const int stall = sizeof(T) == 4 ? kStallWindow : 0;
  // While meshes finish, the iteration is narrowed to the chunks and block rows of those
  // still iterating (re-captured each time the remaining block rows fall below half of
  // those the iteration covers).
  const int4* ch = B.d_chunk;
  int nch = B.nchunk;
  const RowSubset* sub = nullptr;
  int64_t cover = P.nblk;
  hipGraphExec_t exec = nullptr;
  const bool kt_on = ctx->kt.on;
  auto iterations = [&](int k) -> int {
    for (int q = 0; q < k; ++q) {
      int rc1 = bb_iteration<T>(ctx, B, o->rtol, stall, ch, nch, sub);
      if (rc1) return rc1;
    }
    return SFEM_OK;
  };
  auto capture = [&]() -> int {
    if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
    if (!B.use_graph) return SFEM_OK;
    ctx->kt.on = false;
    hipGraph_t graph = nullptr;
    BA_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc1 = iterations(per_graph);
    hipError_t ec = hipStreamEndCapture(st, &graph);
    ctx->kt.on = kt_on;
    if (rc1) { if (graph) (void)hipGraphDestroy(graph); return rc1; }
    if (ec != hipSuccess) return set_hip_error(ec, "batched iteration capture");
    ec = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ec != hipSuccess) { exec = nullptr; return set_hip_error(ec, "batched iteration instantiate"); }
    return SFEM_OK;
  };
  // narrow to the unfinished meshes (per-mesh flags in h_sc)
  auto narrow = [&]() -> int {
    std::vector<uint8_t> live(B.nmesh);
    int64_t left = 0;
    for (int m = 0; m < B.nmesh; ++m) {
      live[m] = B.h_sc[(size_t)m * B_N + B_DONE] == 0.0;
      if (live[m]) left += B.mesh_nblk[m];
    }
    if (!B.narrowing || 2 * left >= cover) return SFEM_OK;
    RowSubset* nx = sub == &B.sub[0] ? &B.sub[1] : &B.sub[0];
    if (sub) {
      for (int32_t q : sub->h_rows) B.keep[q] = live[B.blk_mesh[q]];
    } else {
      for (int64_t q = 0; q < P.nblk; ++q) B.keep[q] = live[B.blk_mesh[q]];
    }
    int rc1 = la_subset_build(ctx, B.keep.data(), nx, sub);
    if (rc1) return rc1;
    std::vector<int4> cc;
    for (const int4& c : B.h_chunk)
      if (live[c.z]) cc.push_back(c);
    if (!cc.empty()) BA_CHECK(hipMemcpy(B.d_cchunk, cc.data(), cc.size() * sizeof(int4), hipMemcpyHostToDevice));
    ch = B.d_cchunk;
    nch = (int)cc.size();
    sub = nx;
    cover = left;
    ++B.subsets_built;
    return capture();
  };
  int rc;
  hipError_t e = hipMemcpyAsync(B.h_ndone, B.d_ndone, sizeof(int32_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return set_hip_error(e, "batched solve start");
  if (*B.h_ndone >= B.nmesh) return SFEM_OK;
  if ((rc = capture())) return rc;
  int seen = -1;
  for (int it = 0; it < o->max_it; it += per_graph) {
    if (exec) {
      e = hipGraphLaunch(exec, st);
      if (e != hipSuccess) { (void)hipGraphExecDestroy(exec); return set_hip_error(e, "batched iteration launch"); }
    } else if ((rc = iterations(per_graph))) {
      return rc;
    }
    e = hipMemcpyAsync(B.h_ndone, B.d_ndone, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { if (exec) (void)hipGraphExecDestroy(exec); return set_hip_error(e, "batched iteration"); }
    if (*B.h_ndone >= B.nmesh) break;
    if (*B.h_ndone != seen && nch > 0) {
      seen = *B.h_ndone;
      e = hipMemcpyAsync(B.h_sc, B.d_sc, (size_t)B.nmesh * B_N * sizeof(double), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) { if (exec) (void)hipGraphExecDestroy(exec); return set_hip_error(e, "batched flags"); }
      if ((rc = narrow())) { if (exec) (void)hipGraphExecDestroy(exec); return rc; }
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  return SFEM_OK;
}

static int batch_error(Batch& B, int bad, const int32_t* its) {
  const double* S = B.h_sc + (size_t)bad * B_N;
  char msg[200];
  snprintf(msg, sizeof msg, "batched BiCGStab: mesh %d %s after %d iterations (|r|^2 = %.3e, target %.3e)", bad,
           S[B_BAD] != 0.0 ? "broke down (non-finite residual)" : "did not converge", its ? its[bad] : (int)S[B_ITS],
           S[B_RR], S[B_TOL2]);
  return set_error(SFEM_ERR_MAXIT, msg);
}

// Capture `body` (kernel launches only) on the context stream into an executable graph.
template <class F>
static int capture_graph(sfem_ctx* ctx, F body, hipGraphExec_t* exec) {
  hipStream_t st = ctx->stream;
  const bool kt_on = ctx->kt.on;
  ctx->kt.on = false;
  hipGraph_t graph = nullptr;
  BA_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc1 = body();
  hipError_t ec = hipStreamEndCapture(st, &graph);
  ctx->kt.on = kt_on;
  if (rc1) { if (graph) (void)hipGraphDestroy(graph); return rc1; }
  if (ec != hipSuccess) return set_hip_error(ec, "Newton step capture");
  ec = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ec != hipSuccess) { *exec = nullptr; return set_hip_error(ec, "Newton step instantiate"); }
  return SFEM_OK;
}

// One time step of every mesh with the Newton iteration as graphs (StabFEM2.cpp:654-799
// per mesh): iteration 0 assembles the tangent and builds the preconditioner eagerly;
// every later iteration replays a captured head (residual assembly, per-mesh norms, the
// Newton test on the device, start of the solve), the BiCGStab iterations replay in
// graphs of B.its_per_graph iterations until every mesh is done, and a captured tail
// ends the solve (best iterates, increment, iteration counts, failure flags) and updates
// the solution.  The host reads one status word per head and per replay.
template <typename T>
static int step_batch_graph(sfem_ctx* c, const sfem_solver_opts* o, double tol, int max_newton,
                            int32_t* newton_iters, int32_t* krylov_iters, double* rhs_norm) {
  Batch& B = *c->batch;
  hipStream_t st = c->stream;
  const int nm = B.nmesh;
  const int stall = sizeof(T) == 4 ? kStallWindow : 0;
  hipLaunchKernelGGL(k_b_step_reset, dim3(1), dim3(kTB), 0, st, nm, max_newton, B.d_active, B.d_newton, B.d_kry,
                     B.d_rhsn, B.d_status);
  BA_CHECK(hipGetLastError());
  // FP32 inner solves with the coarse correction: every Newton iteration first takes the
  // FP64 coarse step x += Z E^-1 Z^T F on the meshes still iterating, then the residual
  // again (their inner solves leave each mesh's global constant pressure to it)
  const bool cstep = sizeof(T) == 4 && c->la->coarse_k > 0;
  auto coarse = [&]() -> int {
    hipLaunchKernelGGL(k_seg_norm2, dim3(nm), dim3(kTB), 0, st, B.d_dof_ptr, c->d_rhs, B.d_norm);
    hipLaunchKernelGGL(k_b_coarse_mask, dim3(1), dim3(kTB), 0, st, nm, B.d_norm, tol, B.d_active, B.d_cmask);
    int rc1 = la_coarse_step_kernels(c);
    if (rc1) return rc1;
    hipLaunchKernelGGL(k_b_mask_du, dim3(nm), dim3(kTB), 0, st, B.d_dof_ptr, B.d_cmask, c->d_du);
    BA_CHECK(hipGetLastError());
    BA_CHECK(launch_update_solution(st, c->nfree, c->d_assy, c->d_du, c->d_soln));
    return asm_residual_kernels(c);
  };
  auto head = [&]() -> int {
    if (cstep) {
      int rc1 = coarse();
      if (rc1) return rc1;
    }
    hipLaunchKernelGGL(k_seg_norm2, dim3(nm), dim3(kTB), 0, st, B.d_dof_ptr, c->d_rhs, B.d_norm);
    hipLaunchKernelGGL(k_b_activity, dim3(1), dim3(kTB), 0, st, nm, B.d_norm, tol, B.d_active, B.d_newton, B.d_rhsn,
                       B.d_status);
    BA_CHECK(hipGetLastError());
    return solve_begin<T>(c, o);
  };
  auto tail = [&]() -> int {
    int rc1 = solve_end<T>(c, o, stall != 0);
    if (rc1) return rc1;
    hipLaunchKernelGGL(k_b_tail, dim3(1), dim3(kTB), 0, st, nm, B.d_sc, B.d_kry, B.d_status);
    BA_CHECK(hipGetLastError());
    BA_CHECK(launch_update_solution(st, c->nfree, c->d_assy, c->d_du, c->d_soln));
    return SFEM_OK;
  };
  auto read_status = [&]() -> int {
    BA_CHECK(hipMemcpyAsync(B.h_status, B.d_status, ST_N * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    BA_CHECK(hipStreamSynchronize(st));
    return SFEM_OK;
  };
  hipGraphExec_t head_exec = nullptr, tail_exec = nullptr;
  auto cleanup = [&]() {
    if (head_exec) (void)hipGraphExecDestroy(head_exec);
    if (tail_exec) (void)hipGraphExecDestroy(tail_exec);
  };
  int rc = SFEM_OK;
  for (int iter = 0; iter < max_newton && rc == SFEM_OK; ++iter) {
    if (iter == 0) {
      double gnorm;
      if ((rc = sfem_assemble(c, 0, 1, &gnorm)) || (rc = solve_prepare<T>(c, o)) || (rc = asm_prepare_residual(c)) ||
          (cstep && (rc = la_coarse_step_prepare(c))) || (rc = head()))
        break;
    } else if (B.use_graph) {
      if (!head_exec &&
          (rc = capture_graph(c, [&]() -> int { int r = asm_residual_kernels(c); return r ? r : head(); }, &head_exec)))
        break;
      hipError_t e = hipGraphLaunch(head_exec, st);
      if (e != hipSuccess) { rc = set_hip_error(e, "Newton step head"); break; }
    } else if ((rc = asm_residual_kernels(c)) || (rc = head())) {
      break;
    }
    if ((rc = read_status())) break;
    if (B.h_status[ST_ERR] != 0x7fffffff) break;  // a mesh failed in the previous solve
    if (B.h_status[ST_NACTIVE] == 0) break;
    if ((rc = solve_iterate<T>(c, o, B.use_graph ? B.its_per_graph : 1))) break;
    if (B.use_graph) {
      if (!tail_exec && (rc = capture_graph(c, tail, &tail_exec))) break;
      hipError_t e = hipGraphLaunch(tail_exec, st);
      if (e != hipSuccess) { rc = set_hip_error(e, "Newton step tail"); break; }
    } else if ((rc = tail())) {
      break;
    }
  }
  cleanup();
  if (rc) return rc;
  std::vector<double> rn(nm);
  BA_CHECK(hipMemcpyAsync(newton_iters, B.d_newton, nm * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  BA_CHECK(hipMemcpyAsync(krylov_iters, B.d_kry, nm * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  BA_CHECK(hipMemcpyAsync(rhs_norm, B.d_rhsn, nm * sizeof(double), hipMemcpyDeviceToHost, st));
  if ((rc = read_status())) return rc;
  B.stalled += B.h_status[ST_NSTALL];
  const int bad = B.h_status[ST_ERR];
  if (bad != 0x7fffffff) {
    BA_CHECK(hipMemcpy(B.h_sc, B.d_sc, (size_t)nm * B_N * sizeof(double), hipMemcpyDeviceToHost));
    return batch_error(B, bad, nullptr);
  }
  return SFEM_OK;
}

}  // namespace sfem

using namespace sfem;

extern "C" {

int sfem_set_batch(sfem_ctx* c, int32_t nmesh, const int32_t* elem_mesh, const double* mu) {
  if (!c || nmesh <= 0 || !elem_mesh || !mu) return set_error(SFEM_ERR_ARG, "bad batch arguments");
  if (!c->la) return set_error(SFEM_ERR_STATE, "build the pattern before sfem_set_batch");
  HIP_OK(hipSetDevice(c->device));
  const int npe = c->dims.npe, ndof = c->dims.ndof;
  // mesh of every node; nodes (and so free DOFs) must come mesh by mesh
  std::vector<int32_t> nodeMesh(c->nNode, -1);
  for (int64_t e = 0; e < c->nElem; ++e) {
    const int32_t m = elem_mesh[e];
    if (m < 0 || m >= nmesh) return set_error(SFEM_ERR_ARG, "element mesh id out of range");
    for (int a = 0; a < npe; ++a) {
      int32_t& nm = nodeMesh[c->conn[e * npe + a]];
      if (nm >= 0 && nm != m) return set_error(SFEM_ERR_ARG, "a node is shared by two meshes");
      nm = m;
    }
  }
  std::vector<int64_t> ptr(nmesh + 1, 0);
  int32_t last = 0;
  for (int64_t i = 0; i < c->nNode; ++i) {
    const int32_t m = nodeMesh[i];
    if (m < 0) continue;
    if (m < last) return set_error(SFEM_ERR_ARG, "nodes must be numbered mesh by mesh");
    last = m;
    for (int d = 0; d < ndof; ++d)
      if (c->nodeDof[i * ndof + d] >= 0) ptr[m + 1]++;
  }
  for (int m = 0; m < nmesh; ++m) ptr[m + 1] += ptr[m];
  std::vector<double> emu(c->nElem);
  for (int64_t e = 0; e < c->nElem; ++e) emu[e] = mu[elem_mesh[e]];
  if (c->batch) batch_destroy(c);
  Batch* B = new Batch;
  c->batch = B;
  B->nmesh = nmesh;
  B->dof_ptr = ptr;
  B->node_mesh = nodeMesh;
  if (!c->d_elem_mu && hipMalloc((void**)&c->d_elem_mu, c->nElem * sizeof(double)) != hipSuccess)
    return set_error(SFEM_ERR_HIP, "hipMalloc elem mu");
  HIP_OK(hipMemcpy(c->d_elem_mu, emu.data(), c->nElem * sizeof(double), hipMemcpyHostToDevice));
  if (balloc(&B->d_dof_ptr, ptr.size()) != hipSuccess || balloc(&B->d_sc, (size_t)nmesh * B_N) != hipSuccess ||
      balloc(&B->d_norm, nmesh) != hipSuccess || balloc(&B->d_active, nmesh) != hipSuccess ||
      balloc(&B->d_ndone, 1) != hipSuccess || balloc(&B->d_newton, nmesh) != hipSuccess ||
      balloc(&B->d_kry, nmesh) != hipSuccess || balloc(&B->d_rhsn, nmesh) != hipSuccess ||
      balloc(&B->d_status, ST_N) != hipSuccess || balloc(&B->d_cmask, nmesh) != hipSuccess)
    return set_error(SFEM_ERR_HIP, "hipMalloc batch state");
  HIP_OK(hipMemcpy(B->d_dof_ptr, ptr.data(), ptr.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  HIP_OK(hipHostMalloc((void**)&B->h_norm, nmesh * sizeof(double), hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&B->h_ndone, sizeof(int32_t), hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&B->h_sc, (size_t)nmesh * B_N * sizeof(double), hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&B->h_status, ST_N * sizeof(int32_t), hipHostMallocDefault));
  // a coarse correction set before the batch is planned again per mesh
  if (c->la->coarse_k > 0) {
    int rc = coarse_set(c, c->la->coarse_k, c->la->coarse_mult ? 1 : 0);
    if (rc) return rc;
    c->la->pc_K_version = -1;
  }
  return SFEM_OK;
}

int sfem_batch_use_graph(sfem_ctx* c, int enable) {
  if (!c || !c->batch) return set_error(SFEM_ERR_STATE, "no batch set");
  c->batch->use_graph = enable != 0;
  return SFEM_OK;
}

int sfem_batch_narrowing(sfem_ctx* c, int enable) {
  if (!c || !c->batch) return set_error(SFEM_ERR_STATE, "no batch set");
  c->batch->narrowing = enable != 0;
  return SFEM_OK;
}

int sfem_step_batch(sfem_ctx* c, double timeNow, const sfem_solver_opts* opts, double newton_tol, int max_newton,
                    int32_t* newton_iters, int32_t* krylov_iters, double* rhs_norm) {
  if (!c || !opts || !c->batch) return set_error(SFEM_ERR_ARG, "no batch set or null argument");
  if (opts->pc_type != SFEM_PC_ILU0 || opts->ksp_type != SFEM_KSP_BCGS)
    return set_error(SFEM_ERR_UNSUPPORTED, "batched steps use BiCGStab + ILU(0)");
  if (opts->atol != 0.0) return set_error(SFEM_ERR_UNSUPPORTED, "batched steps use the relative test only (atol = 0)");
  HIP_OK(hipSetDevice(c->device));
  int rc;
  if ((rc = sfem_time_update(c, timeNow))) return rc;
  return c->precision == SFEM_PREC_MIXED
             ? step_batch_graph<float>(c, opts, newton_tol, max_newton, newton_iters, krylov_iters, rhs_norm)
             : step_batch_graph<double>(c, opts, newton_tol, max_newton, newton_iters, krylov_iters, rhs_norm);
}

}  // extern "C"
