// sfem_direct.hip -- device direct solver for the `-ksp_type preonly -pc_type lu` decks.
//
// The reference's test decks solve every Newton system with MUMPS LU through PETSc
// (testdata/ldc2dquad10/petsc_options.dat:1-3, SolverPetsc::solve at
// src/base/SolverPetsc.cpp:224-277).  This is a banded LU with partial pivoting on the
// device, for the small 2D decks (tens of thousands of DOFs):
//
//   * ordering: reverse Cuthill-McKee (Cuthill & McKee 1969, George 1971) on the
//     free-DOF graph of the CSR pattern, so that every nonzero lies within kl
//     sub-diagonals and ku super-diagonals of the reordered matrix;
//   * storage: the band in the layout of LAPACK's xGBTRF, column by column, column j
//     holding rows j - kl - ku .. j + kl (kv = kl + ku super-diagonals leave room for
//     the fill that row interchanges bring in);
//   * factorisation: right-looking by panels of kPanel columns, with the arithmetic of
//     LAPACK's unblocked DGBTF2: partial pivoting by the largest magnitude in the column
//     (the first one on ties), multipliers kept in the order they are computed (a row
//     interchange is applied to the columns from the pivot column on, not to earlier
//     multipliers); one workgroup factors a panel, then one workgroup per column to the
//     right applies the panel's interchanges and updates in order;
//   * solve: the interchanges interleaved with the columns of L (as DGBTRS does), then
//     the upper band.
//
// FP64 throughout; the tangent is factorised once per new tangent (the affine residual
// keeps it constant within a time step) and every Newton iteration solves with it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sfem_internal.hpp"
#include "sfem_linalg.hpp"

namespace sfem {

constexpr int kPanel = 16;       // columns per panel
constexpr int kPanelThreads = 256;
constexpr int kSolveThreads = 64;  // one wavefront walks the sequential steps of the triangular solves

#define DL_CHECK(x)                                  \
  do {                                               \
    hipError_t _e = (x);                             \
    if (_e != hipSuccess) return set_hip_error(_e, #x); \
  } while (0)

struct DirectLU {
  int64_t n = 0;
  int kl = 0, ku = 0, kv = 0;
  int64_t ldab = 0;
  int64_t K_version = -1;             // tangent version of the factor
  std::vector<int32_t> perm;          // new position -> DOF row
  int32_t* d_perm = nullptr;
  int32_t* d_iperm = nullptr;         // DOF row -> new position
  double* d_ab = nullptr;             // band, column-major, ldab per column
  int32_t* d_ipiv = nullptr;          // row interchanged with row j at step j
  double* d_w = nullptr;              // right-hand side / solution in the new order
  int32_t* d_flag = nullptr;          // first zero pivot (n if none)
  ~DirectLU() {
    void* ptrs[] = {d_perm, d_iperm, d_ab, d_ipiv, d_w, d_flag};
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
  }
};

// element (i, j) of the band
__device__ inline int64_t band_at(int64_t i, int64_t j, int kv, int64_t ldab) { return j * ldab + kv + i - j; }

__global__ __launch_bounds__(256) void k_lu_fill0(int64_t n, double* __restrict__ a) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) a[i] = 0.0;
}

// CSR values (DOF order) into the band of the reordered matrix: one thread per row.
__global__ __launch_bounds__(256) void k_lu_scatter(int64_t n, const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ colidx,
                                                    const double* __restrict__ vals,
                                                    const int32_t* __restrict__ iperm, int kv, int64_t ldab,
                                                    double* __restrict__ ab) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const int64_t i = iperm[r];
  for (int32_t q = rowptr[r]; q < rowptr[r + 1]; ++q) ab[band_at(i, iperm[colidx[q]], kv, ldab)] = vals[q];
}

// Factorisation of panel columns j0 .. j0 + nb - 1 (one workgroup): for each column, the
// pivot (largest |a|, first on ties) among rows j .. j + km, the interchange of rows j and
// p within the panel's columns j .. j0 + nb - 1, the multipliers, and the rank-1 update of
// the panel's columns to the right.  flag receives the first zero pivot.
__global__ __launch_bounds__(kPanelThreads) void k_lu_panel(int64_t n, int kl, int kv, int64_t ldab, int64_t j0,
                                                            int nb, double* __restrict__ ab,
                                                            int32_t* __restrict__ ipiv, int32_t* __restrict__ flag) {
  __shared__ double sv[kPanelThreads];
  __shared__ int si[kPanelThreads];
  __shared__ int piv_row;
  __shared__ double piv_val;
  const int tid = threadIdx.x;
  for (int jj = 0; jj < nb; ++jj) {
    const int64_t j = j0 + jj;
    const int km = (int)std::min<int64_t>(kl, n - 1 - j);
    // pivot search
    double best = -1.0;
    int bi = 0;
    for (int r = tid; r <= km; r += kPanelThreads) {
      const double a = fabs(ab[band_at(j + r, j, kv, ldab)]);
      if (a > best) { best = a; bi = r; }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int o = kPanelThreads / 2; o > 0; o >>= 1) {
      if (tid < o) {
        const double b2 = sv[tid + o];
        const int i2 = si[tid + o];
        if (b2 > sv[tid] || (b2 == sv[tid] && i2 < si[tid])) { sv[tid] = b2; si[tid] = i2; }
      }
      __syncthreads();
    }
    if (tid == 0) {
      piv_row = si[0];
      piv_val = ab[band_at(j + si[0], j, kv, ldab)];
      ipiv[j] = (int32_t)(j + si[0]);
      if (piv_val == 0.0 && *flag == (int32_t)n) *flag = (int32_t)j;
    }
    __syncthreads();
    const int p = piv_row;
    const double pv = piv_val;
    // interchange rows j and j + p in the panel's columns j .. j0 + nb - 1
    if (p != 0) {
      for (int c = tid; c < nb - jj && c <= kv; c += kPanelThreads) {
        const int64_t col = j + c;
        double* x = ab + band_at(j, col, kv, ldab);
        double* y = ab + band_at(j + p, col, kv, ldab);
        const double t = *x;
        *x = *y;
        *y = t;
      }
    }
    __syncthreads();
    if (pv != 0.0) {
      const double inv = 1.0 / pv;
      for (int r = 1 + tid; r <= km; r += kPanelThreads) ab[band_at(j + r, j, kv, ldab)] *= inv;
    }
    __syncthreads();
    // rank-1 update of the panel's columns right of j
    const int ncol = std::min(nb - jj - 1, kv);  // columns of the panel that row j reaches
    if (pv != 0.0 && ncol > 0) {
      for (int it = tid; it < km * ncol; it += kPanelThreads) {
        const int c = 1 + it / km, r = 1 + it % km;
        const int64_t col = j + c;
        const double u = ab[band_at(j, col, kv, ldab)];
        if (u != 0.0) ab[band_at(j + r, col, kv, ldab)] -= ab[band_at(j + r, j, kv, ldab)] * u;
      }
    }
    __syncthreads();
  }
}

// The columns right of the panel (one workgroup per column c), the panel's columns
// applied in order as the unblocked algorithm applies them: at step j the interchange
// of rows j and ipiv[j], then the rows below j updated with column j's multipliers
// (the multipliers are not interchanged by later steps, so the steps cannot be
// regrouped into one triangular solve and one product).  A step whose row j lies above
// the column's band (c > j + kl + ku) touches only zeros and is skipped.
__global__ __launch_bounds__(kPanelThreads) void k_lu_update(int64_t n, int kl, int kv, int64_t ldab, int64_t j0,
                                                             int nb, int64_t c0, double* __restrict__ ab,
                                                             const int32_t* __restrict__ ipiv) {
  const int64_t c = c0 + blockIdx.x;
  if (c >= n) return;
  const int tid = threadIdx.x;
  for (int jj = 0; jj < nb; ++jj) {
    const int64_t j = j0 + jj;
    if (c - kv > j) continue;  // the same for the whole workgroup
    const int km = (int)std::min<int64_t>(kl, n - 1 - j);
    if (tid == 0) {
      const int64_t p = ipiv[j];
      if (p != j) {
        double* x = ab + band_at(j, c, kv, ldab);
        double* y = ab + band_at(p, c, kv, ldab);
        const double t = *x;
        *x = *y;
        *y = t;
      }
    }
    __syncthreads();
    const double u = ab[band_at(j, c, kv, ldab)];
    if (u != 0.0)
      for (int r = 1 + tid; r <= km; r += kPanelThreads) ab[band_at(j + r, c, kv, ldab)] -= ab[band_at(j + r, j, kv, ldab)] * u;
    __syncthreads();
  }
}

// w[i] = b[perm[i]]  /  x[perm[i]] = w[i]
__global__ __launch_bounds__(256) void k_lu_gather(int64_t n, const int32_t* __restrict__ perm,
                                                   const double* __restrict__ b, double* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) w[i] = b[perm[i]];
}
__global__ __launch_bounds__(256) void k_lu_scatter_out(int64_t n, const int32_t* __restrict__ perm,
                                                        const double* __restrict__ w, double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[perm[i]] = w[i];
}

// Forward sweep with the interchanges interleaved (DGBTRS, no transpose), then the
// upper band, by one workgroup of one wavefront: each step's entry is settled, then the
// wavefront updates the entries the column reaches; the workgroup barriers between the
// steps make each step's global writes visible to the next.
__global__ __launch_bounds__(kSolveThreads) void k_lu_solve(int64_t n, int kl, int kv, int64_t ldab,
                                                            const double* __restrict__ ab,
                                                            const int32_t* __restrict__ ipiv,
                                                            double* __restrict__ w) {
  __shared__ double xs;
  const int lane = threadIdx.x;
  for (int64_t j = 0; j < n; ++j) {
    const int km = (int)std::min<int64_t>(kl, n - 1 - j);
    if (lane == 0) {
      const int64_t p = ipiv[j];
      const double a = w[j];
      double b = a;
      if (p != j) {
        b = w[p];
        w[p] = a;
        w[j] = b;
      }
      xs = b;
    }
    __syncthreads();
    const double bj = xs;
    if (bj != 0.0)
      for (int r = 1 + lane; r <= km; r += kSolveThreads) w[j + r] -= ab[band_at(j + r, j, kv, ldab)] * bj;
    __syncthreads();
  }
  for (int64_t j = n - 1; j >= 0; --j) {
    if (lane == 0) {
      const double x = w[j] / ab[band_at(j, j, kv, ldab)];
      w[j] = x;
      xs = x;
    }
    __syncthreads();
    const double xj = xs;
    const int64_t i0 = std::max<int64_t>(0, j - kv);
    if (xj != 0.0)
      for (int64_t i = i0 + lane; i < j; i += kSolveThreads) w[i] -= ab[band_at(i, j, kv, ldab)] * xj;
    __syncthreads();
  }
}

// Reverse Cuthill-McKee order of the free-DOF graph (the CSR pattern, structurally
// symmetric): breadth-first from a pseudo-peripheral row of the smallest degree, the
// neighbours of each row visited by increasing degree, every connected component in
// turn, the whole order reversed.
static std::vector<int32_t> rcm_dofs(int64_t n, const std::vector<int64_t>& rp, const std::vector<int32_t>& ci) {
  std::vector<int32_t> deg(n), order;
  order.reserve(n);
  for (int64_t r = 0; r < n; ++r) deg[r] = (int32_t)(rp[r + 1] - rp[r]);
  std::vector<char> seen(n, 0);
  std::vector<int32_t> lvl(n, -1), nb;
  auto bfs_far = [&](int32_t s) {  // last row reached by a breadth-first search from s
    std::vector<int32_t> q{s};
    std::vector<int32_t> mark;
    lvl[s] = 0;
    mark.push_back(s);
    size_t h = 0;
    int32_t far = s;
    while (h < q.size()) {
      const int32_t r = q[h++];
      if (lvl[r] > lvl[far] || (lvl[r] == lvl[far] && deg[r] < deg[far])) far = r;
      for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
        const int32_t c = ci[k];
        if (lvl[c] < 0) { lvl[c] = lvl[r] + 1; q.push_back(c); mark.push_back(c); }
      }
    }
    for (int32_t m : mark) lvl[m] = -1;
    return far;
  };
  for (int64_t s0 = 0; s0 < n; ++s0) {
    if (seen[s0]) continue;
    // start: the smallest-degree unvisited row, moved to a far end of its component
    int32_t s = (int32_t)s0;
    for (int it = 0; it < 2; ++it) s = bfs_far(s);
    size_t h = order.size();
    order.push_back(s);
    seen[s] = 1;
    while (h < order.size()) {
      const int32_t r = order[h++];
      nb.clear();
      for (int64_t k = rp[r]; k < rp[r + 1]; ++k)
        if (!seen[ci[k]]) { nb.push_back(ci[k]); seen[ci[k]] = 1; }
      std::stable_sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) { return deg[a] < deg[b]; });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

int la_direct_setup(sfem_ctx* ctx) {
  LinAlg* L = ctx->la;
  if (L->direct) return SFEM_OK;
  const int64_t n = ctx->nfree;
  auto D = new DirectLU;
  D->n = n;
  D->perm = rcm_dofs(n, ctx->rowptr, ctx->colidx);
  std::vector<int32_t> iperm(n);
  for (int64_t i = 0; i < n; ++i) iperm[D->perm[i]] = (int32_t)i;
  int kl = 0, ku = 0;
  for (int64_t r = 0; r < n; ++r)
    for (int64_t k = ctx->rowptr[r]; k < ctx->rowptr[r + 1]; ++k) {
      const int d = iperm[r] - iperm[ctx->colidx[k]];
      kl = std::max(kl, d);
      ku = std::max(ku, -d);
    }
  D->kl = kl;
  D->ku = ku;
  D->kv = kl + ku;
  D->ldab = 2 * (int64_t)kl + ku + 1;
  const size_t nab = (size_t)std::max<int64_t>(n, 1) * D->ldab;
  hipError_t e = hipMalloc((void**)&D->d_ab, nab * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&D->d_perm, std::max<int64_t>(n, 1) * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&D->d_iperm, std::max<int64_t>(n, 1) * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&D->d_ipiv, std::max<int64_t>(n, 1) * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&D->d_w, std::max<int64_t>(n, 1) * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&D->d_flag, sizeof(int32_t));
  if (e == hipSuccess && n)
    e = hipMemcpy(D->d_perm, D->perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess && n) e = hipMemcpy(D->d_iperm, iperm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete D;
    return set_hip_error(e, "direct solver buffers");
  }
  L->direct = D;
  return SFEM_OK;
}

void la_direct_destroy(LinAlg* L) {
  delete L->direct;
  L->direct = nullptr;
}

int la_direct_info(sfem_ctx* ctx, int* kl, int* ku) {
  LinAlg* L = ctx->la;
  if (!L || !L->direct) { *kl = *ku = 0; return SFEM_OK; }
  *kl = L->direct->kl;
  *ku = L->direct->ku;
  return SFEM_OK;
}

static int direct_factor(sfem_ctx* ctx) {
  DirectLU* D = ctx->la->direct;
  hipStream_t st = ctx->stream;
  int rc0 = la_ensure_csr(ctx);  // the band is filled from the CSR values
  if (rc0) return rc0;
  const int64_t n = D->n;
  const int64_t nab = n * D->ldab;
  hipLaunchKernelGGL(k_lu_fill0, dim3((unsigned)std::min<int64_t>((nab + 255) / 256, 1 << 20)), dim3(256), 0, st, nab,
                     D->d_ab);
  hipLaunchKernelGGL(k_lu_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, ctx->d_rowptr32,
                     ctx->d_colidx, ctx->d_vals, D->d_iperm, D->kv, D->ldab, D->d_ab);
  const int32_t none = (int32_t)n;
  DL_CHECK(hipMemcpyAsync(D->d_flag, &none, sizeof(int32_t), hipMemcpyHostToDevice, st));
  for (int64_t j0 = 0; j0 < n; j0 += kPanel) {
    const int nb = (int)std::min<int64_t>(kPanel, n - j0);
    hipLaunchKernelGGL(k_lu_panel, dim3(1), dim3(kPanelThreads), 0, st, n, D->kl, D->kv, D->ldab, j0, nb, D->d_ab,
                       D->d_ipiv, D->d_flag);
    const int64_t c0 = j0 + nb, c1 = std::min<int64_t>(n - 1, j0 + nb - 1 + D->kv);
    if (c1 >= c0)
      hipLaunchKernelGGL(k_lu_update, dim3((unsigned)(c1 - c0 + 1)), dim3(kPanelThreads), 0, st, n, D->kl, D->kv,
                         D->ldab, j0, nb, c0, D->d_ab, D->d_ipiv);
  }
  DL_CHECK(hipGetLastError());
  int32_t flag = 0;
  DL_CHECK(hipMemcpyAsync(&flag, D->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  DL_CHECK(hipStreamSynchronize(st));
  if (flag != (int32_t)n) {
    D->K_version = -1;
    return set_error(SFEM_ERR_DIVERGED, "LU: zero pivot (singular tangent)");
  }
  D->K_version = ctx->la_K_version;
  return SFEM_OK;
}

// du = K^{-1} rhs with the banded LU of the current tangent (factorised when the tangent
// changed); *relres receives ||rhs - K du|| / ||rhs||.
int la_direct_solve(sfem_ctx* ctx, int* iters, double* relres) {
  LinAlg* L = ctx->la;
  if (ctx->precision != SFEM_PREC_FP64)
    return set_error(SFEM_ERR_UNSUPPORTED, "the direct solver factorises the FP64 tangent (precision FP64 only)");
  int rc;
  if (!L->direct && (rc = la_direct_setup(ctx))) return rc;
  DirectLU* D = L->direct;
  if (D->K_version != ctx->la_K_version && (rc = direct_factor(ctx))) return rc;
  hipStream_t st = ctx->stream;
  const int64_t n = D->n;
  const unsigned g = (unsigned)std::max<int64_t>((n + 255) / 256, 1);
  *iters = 1;
  if (n == 0) { *relres = 0.0; return SFEM_OK; }
  hipLaunchKernelGGL(k_lu_gather, dim3(g), dim3(256), 0, st, n, D->d_perm, ctx->d_rhs, D->d_w);
  hipLaunchKernelGGL(k_lu_solve, dim3(1), dim3(kSolveThreads), 0, st, n, D->kl, D->kv, D->ldab, D->d_ab, D->d_ipiv,
                     D->d_w);
  hipLaunchKernelGGL(k_lu_scatter_out, dim3(g), dim3(256), 0, st, n, D->d_perm, D->d_w, ctx->d_du);
  DL_CHECK(hipGetLastError());
  // true residual of the solve
  double bn = 0.0, rn = 0.0;
  if ((rc = la_norm2(ctx, ctx->d_rhs, &bn))) return rc;
  if ((rc = la_residual_norm(ctx, ctx->d_rhs, ctx->d_du, &rn))) return rc;
  *relres = bn > 0.0 ? rn / bn : 0.0;
  if (!std::isfinite(rn)) return set_error(SFEM_ERR_DIVERGED, "LU: non-finite solution");
  return SFEM_OK;
}

}  // namespace sfem
