// sfem_linalg.hpp -- Krylov workspace and entry points (implemented in sfem_linalg.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <vector>

#include "sfem_internal.hpp"

namespace sfem {

// Error reporting shared by all translation units (sfem_api.cpp).
int set_error(int code, const char* msg);
int set_hip_error(hipError_t e, const char* what);

// ILU(0) plan for one node ordering, in node-block form: block rows in the new order
// with their neighbour blocks, the gather offsets into the DOF-ordered CSR, level sets,
// and the map from the padded block layout of vectors (4 slots per block) to DOFs.
struct BlkDesc;
struct BDesc;
struct bf16;

struct IluPlan {
  int reorder = -1;
  bool contiguous = false;   // level l is the block range [levPtr[l], levPtr[l+1])
  bool fuse_last = false;    // the last forward level also does the first backward level
  int64_t nblk = 0, npad = 0, nvals = 0, npairs = 0;
  int nbmax = 0;
  std::vector<int32_t> levLPtr, levUPtr, h_levL, h_levU;
  int32_t *d_levL = nullptr, *d_levU = nullptr;
  BDesc* d_bdesc = nullptr;
  int32_t* d_bcol = nullptr;     // neighbour blocks (new positions), per block row
  int32_t* d_bcolofs = nullptr;  // their first column's offset in the DOF-ordered CSR row
  int32_t* d_kbsrc = nullptr;    // their node-pair block in the assembled block form (ctx->kb_on)
  int32_t* d_r0orig = nullptr;   // first DOF row of each block row
  int32_t* d_padmap = nullptr;   // padded slot -> DOF row, -1 for padding
  int32_t* d_natrows = nullptr;  // block rows in the original node order (reordered plans)
  std::vector<int32_t> h_order;  // new position -> original block
  // tiled multicolour plan (reorder 3): the block rows of a tile (the node blocks of one
  // element, at most 8) are consecutive, tiles of a colour share no neighbour, and the
  // sweeps run one launch per colour with one wavefront per tile
  bool tiled = false;
  bool tile_fuse_last = false;         // the last colour's tiles have no upper blocks outside themselves
  std::vector<int32_t> h_tileStart;    // tile t = block rows [h_tileStart[t], h_tileStart[t+1])
  std::vector<int32_t> tileColPtr;     // colour c = tiles [tileColPtr[c], tileColPtr[c+1])
  int64_t tile_vals = 0;               // factor values inside the tiles (applied through the inverses)
  int32_t* d_tileStart = nullptr;
  // the tile sweeps' copy of the factor blocks outside the tiles, block by block (16
  // values, rows a >= m zero): per block row {first lower block, first upper block,
  // lower count, first upper block's position in the row}; lower parts of all rows
  // in plan order, then the upper parts
  int4* d_swoff = nullptr;
  int64_t sw_blocks = 0;
  // per (tile, row) records of the sweeps, kTileMax per tile: {block row (-1 past the
  // tile), first of its blocks outside the tile in d_bcol, their count | m << 16, their
  // first block in the copy}, forward (lower) and backward (upper)
  int4* d_trowL = nullptr;
  int4* d_trowU = nullptr;
  // the factorisation's pivot-pair schedule: per block row the updates R_p,ps -= L_pt U_cq
  // (t | q << 8 | ps << 16, from d_fpofs[p]); fac_sched false runs the search-based kernel
  int64_t* d_fpofs = nullptr;
  uint32_t* d_fprec = nullptr;
  int64_t fac_pairs = 0;
  int fac_maxpairs = 1;
  bool fac_sched = false;
  ~IluPlan();
};

struct DirectLU;  // banded LU of the direct solver (sfem_direct.hip)
struct Coarse;    // coarse pressure correction of ILU(0) (sfem_coarse.hip)

struct LinAlg {
  int64_t n = 0;             // free DOFs
  int64_t nblk = 0;          // nodal blocks (nodes with >= 1 free DOF)
  int maxrow = 0;
  int64_t nk = 0;             // length of the Krylov vectors of the current solve
  int64_t nvec = 0;           // allocated length of every Krylov vector
  bool block_space = false;   // the current solve runs in the ILU(0) plan's block layout
  // autotuned kernel variants, [0] FP64 and [1] FP32: CSR SpMV, block SpMV, and the
  // block triangular solves per level of the forward / backward sweep
  int tune_spmv[2] = {3, 3}, tune_bspmv[2] = {6, 6};
  std::vector<int> tune_L[2], tune_U[2];
  bool tuned_csr[2] = {false, false};
  int tune_mode = SFEM_TUNE_TIMED;  // SFEM_TUNE_FIXED: default or installed shapes, no timing
  int64_t fallbacks = 0;     // FP32 BiCGStab solves finished by the GMRES fallback
  int64_t partial = 0;       // FP32 solves ended above rtol, best iterate handed to Newton
  int plan_id = 0, tune_plan[2] = {-1, -1};  // ILU plan the per-level shapes were tuned for
  int pc_type = SFEM_PC_NONE;
  bool pc_fp32 = false;
  int64_t pc_K_version = -1;  // tangent version the preconditioner was built from
  int64_t kb_version = -1;    // tangent version of the node-block copy the products use
  int pc_lag = 1;             // rebuild the preconditioner every pc_lag-th new tangent
  int lag_count = 0;
  std::vector<int32_t> h_blkStart, h_diagpos;
  bool all_diag = false;      // every row's pattern holds its diagonal (checked once per pattern)
  int pc_reorder = 0;
  bool pc_bf16 = false;
  bool pc_store32 = false;    // FP64-built ILU(0) whose factor and block tangent are kept in FP32
  bf16* lu16 = nullptr;
  IluPlan* ilu = nullptr;
  int32_t *d_blkStart = nullptr, *d_diagpos = nullptr;
  BlkDesc* d_desc = nullptr;   // per node block, original order
  // every lower neighbour's rows appear as one contiguous group of its full size in each
  // row (true of element patterns); enables the node-grouped factorisation
  bool full_blocks = false;
  double* d_part = nullptr;  // reduction partials
  double* d_scal = nullptr;  // Krylov scalars
  double* h_scal = nullptr;  // pinned mirror
  double* d_h = nullptr;     // GMRES Hessenberg column scratch
  double* h_h = nullptr;     // pinned mirror
  double* d_mpart = nullptr; // batched reduction partials (8 per block)
  // FP64 Krylov vectors
  double *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr, *s = nullptr, *t = nullptr,
         *ph = nullptr, *sh = nullptr, *x = nullptr, *w = nullptr, *z = nullptr, *gm_V = nullptr,
         *pr = nullptr, *pz = nullptr;
  int gm_m = 0;
  // FP32 Krylov vectors
  float *r32 = nullptr, *rh32 = nullptr, *p32 = nullptr, *v32 = nullptr, *s32 = nullptr,
        *t32 = nullptr, *ph32 = nullptr, *sh32 = nullptr, *x32 = nullptr, *du32 = nullptr,
        *w32 = nullptr, *z32 = nullptr, *gm_V32 = nullptr, *pr32 = nullptr, *pz32 = nullptr;
  int gm_m32 = 0;
  // preconditioners
  double* dinv = nullptr;
  float* dinv32 = nullptr;
  double* binv = nullptr;
  float* binv32 = nullptr;
  double* lu = nullptr;
  float* lu32 = nullptr;
  // tiled plans: per tile the inverses of the tile's unit-lower and upper triangular
  // factors (kTileInv values: two packed 32 x 32 triangles), computed in FP64 and stored
  // like the factor (FP64, FP32 or BF16)
  double* tinv = nullptr;
  float* tinv32 = nullptr;
  bf16* tinv16 = nullptr;
  // the factorisation's copy of every row's diagonal and upper blocks, 16 values per
  // block (transposed: column b' then row b), at the block's neighbour entry (bcol index)
  double* fu = nullptr;
  float* fu32 = nullptr;
  double* sw = nullptr;    // tile sweeps' off-tile factor blocks (d_swoff), FP64 / FP32 / BF16
  float* sw32 = nullptr;
  bf16* sw16 = nullptr;
  // the tangent in the ILU(0) plan's block form, for the products of the Krylov solve
  double* kp = nullptr;
  float* kp32 = nullptr;
  DirectLU* direct = nullptr;  // direct solver (pc_type LU), set up on first use
  int coarse_k = 0;            // coarse pressure boxes per direction (0: no correction)
  bool coarse_mult = false;    // multiplicative form z + Z E^-1 Z^T (r - A z) instead of additive
  Coarse* coarse = nullptr;
  ~LinAlg();
};

// A subset of the ILU(0) plan's block rows (plan positions) that products and
// preconditioner applications can be restricted to: the rows of the subset (SpMV) and
// the subset's rows of every forward then backward level (triangular solves), as device
// lists with host offsets.  Rows outside the subset are neither read as outputs nor
// written; the subset must be closed under coupling (a union of whole meshes).
struct RowSubset {
  int plan_id = -1;
  int64_t nrows = 0;
  int32_t* d_rows = nullptr;
  int32_t* d_lev = nullptr;
  std::vector<int32_t> lofs, uofs;  // forward level l: d_lev[lofs[l] .. lofs[l+1]), backward: uofs
  std::vector<int32_t> h_rows, h_lev;
  ~RowSubset();
};
// S = the plan positions p with keep[p] != 0, taken from `from` (a subset of the same
// plan, so that shrinking costs the size of the current subset) or from all block rows
int la_subset_build(sfem_ctx* ctx, const uint8_t* keep, RowSubset* S, const RowSubset* from = nullptr);

// Products and preconditioner application inside a solve (sfem_batch.hip drives its own
// Krylov iteration through these): the current layout and preconditioner of the context,
// over all block rows or over a row subset.
int la_apply_pc32(sfem_ctx* ctx, const float* r, float* z, const RowSubset* sub = nullptr);
int la_apply_pc64(sfem_ctx* ctx, const double* r, double* z, const RowSubset* sub = nullptr);
int la_apply_op32(sfem_ctx* ctx, const float* x, float* y, const RowSubset* sub = nullptr);
int la_apply_op64(sfem_ctx* ctx, const double* x, double* y, const RowSubset* sub = nullptr);
// z = M^{-1} r and out = A z together (the multiplicative coarse correction reuses the
// product of the ILU(0) output)
int la_apply_pc_op32(sfem_ctx* ctx, const float* r, float* z, float* out, const RowSubset* sub = nullptr);
int la_apply_pc_op64(sfem_ctx* ctx, const double* r, double* z, double* out, const RowSubset* sub = nullptr);
// Bring the preconditioner and the Krylov layout up to date for opts (as la_solve does).
int la_prepare_solve(sfem_ctx* ctx, const sfem_solver_opts* o);

int la_create(sfem_ctx* ctx);
void la_destroy(sfem_ctx* ctx);
int la_setup_pc(sfem_ctx* ctx, int pc_type, bool fp32, int reorder, bool bf16f, bool refactor = true,
                bool store32 = false);
int la_solve(sfem_ctx* ctx, const sfem_solver_opts* o, int* iters, double* relres);
int la_prepare(sfem_ctx* ctx, const sfem_solver_opts* o);
int la_tuning(sfem_ctx* ctx, int* out);
int la_block_info(sfem_ctx* ctx, int64_t* nblk, int64_t* npairs, int64_t* nvals);
int la_get_ilu(sfem_ctx* ctx, int64_t* nnz_out, int32_t* perm, int32_t* rowptr, int32_t* colidx, double* vals);
int la_tile_info(sfem_ctx* ctx, int64_t* ntiles, int64_t* ncolors, int64_t* tile_vals);
int la_pc_apply_host(sfem_ctx* ctx, const double* r, double* z);
int la_scratch_init(int device);
int64_t la_noalloc_capacity();
int la_pc_apply_noalloc(sfem_ctx* ctx, const double* r, double* z);
int la_tuning_levels(sfem_ctx* ctx, int fp32, int* out, int cap, int* nlevels);
int la_set_tuning_mode(sfem_ctx* ctx, int mode);
int la_get_tuning_table(sfem_ctx* ctx, int fp32, int* table, int cap, int* n);
int la_set_tuning_table(sfem_ctx* ctx, int fp32, const int* table, int n);
int la_norm2(sfem_ctx* ctx, const double* a, double* out);
// ||b - K x|| with the assembled FP64 CSR tangent
int la_residual_norm(sfem_ctx* ctx, const double* b, const double* x, double* out);
// direct solver (sfem_direct.hip)
int la_direct_setup(sfem_ctx* ctx);
int la_direct_solve(sfem_ctx* ctx, int* iters, double* relres);
int la_direct_info(sfem_ctx* ctx, int* kl, int* ku);
void la_direct_destroy(LinAlg* L);
// coarse pressure correction (sfem_coarse.hip)
int coarse_set(sfem_ctx* ctx, int k, int mult);
int coarse_build(sfem_ctx* ctx, const double* kp, const float* kp32);
int coarse_refresh_az(sfem_ctx* ctx, const double* kp, const float* kp32);
bool coarse_az_targets(const LinAlg* L, const int32_t** cinfo, const int32_t** azoff, const int32_t** azbox,
                       void** az);
void coarse_set_az_precision(LinAlg* L, bool fp32);
bool coarse_active(const LinAlg* L);
template <typename T> int coarse_restrict(sfem_ctx* ctx, const T* r, const T* w = nullptr);
template <typename T> int coarse_correct_az(sfem_ctx* ctx, T* z, T* out);
template <typename T> T* coarse_tmp(sfem_ctx* ctx);
template <typename T> int coarse_correct(sfem_ctx* ctx, T* z);
int coarse_info(sfem_ctx* ctx, int64_t* out);
int coarse_get_inverse(sfem_ctx* ctx, int which, double* out, int cap);
int la_coarse_step(sfem_ctx* ctx, const sfem_solver_opts* o);
int la_coarse_step_prepare(sfem_ctx* ctx);
int la_coarse_step_kernels(sfem_ctx* ctx);
void coarse_destroy(Coarse* C);
// CSR values from the assembled node-pair block form when the tangent changed since
// the last conversion (element loops that assemble the block form, ctx->kb_on)
int la_ensure_csr(sfem_ctx* ctx);
int la_spmv64(sfem_ctx* ctx, const double* x, double* y);
int la_spmv32(sfem_ctx* ctx, const float* x, float* y);

// Batched meshes (sfem_batch.hip)
void batch_destroy(sfem_ctx* ctx);
// FP32 mesh solves of batched steps that stalled and kept their best iterate
int64_t batch_stalled(const sfem_ctx* ctx);
// mesh of every node of a batch (NULL without one)
const std::vector<int32_t>* batch_node_mesh(const sfem_ctx* ctx);

}  // namespace sfem
