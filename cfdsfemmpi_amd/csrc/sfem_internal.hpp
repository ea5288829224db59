// sfem_internal.hpp -- context and kernel-launcher declarations shared by the
// host ABI (sfem_api.cpp) and the HIP kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sfem.h"

namespace sfem {

// Run f(lo, hi) over [0, n) on the host's cores.
template <typename F>
inline void parallel_for(int64_t n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096 || nt == 1) {
    f((int64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=] { f(lo, hi); });
  }
  for (auto& x : th) x.join();
}

// Element families, keyed as in the reference's element factory (StabFEM.cpp:351-366).
enum ElemType { TRIA3 = 0, QUAD4 = 1, TETRA4 = 2, HEXA8 = 3 };

// Per-element-type constants (constructor of each LagrangeElem*.cpp).
struct ElemDims {
  int ndim, npe, ndof, ngp;
};
inline ElemDims elem_dims(int et) {
  switch (et) {
    case TRIA3: return {2, 3, 3, 1};
    case QUAD4: return {2, 4, 3, 4};
    case TETRA4: return {3, 4, 4, 1};
    default: return {3, 8, 4, 8};
  }
}

// Parameters passed by value to the element kernels.
struct ElemParams {
  double rho, mu, af, acceFact;   // elemData[0..1], td[2], td[8]
  double stab[3];                 // elemData[8..10]
};

// Device-side view of the mesh / pattern / state used by the assembly kernels.
struct AsmView {
  const double* coords;      // nNode*ndim
  const int32_t* conn;       // nElem*npe
  const double* charlen;     // nElem
  const double* solnPrev;    // nNode*ndof
  const double* solnCur;
  const double* solnDotCur;
  const double* solnApplied;
  const int32_t* nodeInfo;   // per node: {rowStart, rowlen, free-DOF mask, first free row}
  const int32_t* colofs;     // [slot][npe][npe] in colour-slot order, -1 where node j has no free dof
  const int32_t* sconn;      // [slot][npe]: connectivity in colour-slot order
  double* vals;              // CSR values (FP64) or NULL
  float* vals32;             // CSR values (FP32) or NULL
  // node-pair block form of the tangent (the Hexa8 loop adds into it instead of the CSR):
  // block q = 16 values, free rows then free columns packed (row a at 4 a), at the
  // block's index in the node-block pattern; blkofs [slot][npe][npe] gives q per element
  // node pair (-1 where a node has no free dof), bit 30 set for the block's first
  // element in colour-slot order, which stores the block instead of adding to it
  const int32_t* blkofs;
  double* kb;
  float* kb32;
  double* rhs;               // nfree
  double* probeK;            // element probe outputs (NULL in normal runs)
  double* probeF;
  const double* elem_mu;     // per-element viscosity (batched meshes) or NULL: ElemParams::mu
};

// Launchers implemented in sfem_elements.hip
hipError_t launch_prepare(int et, hipStream_t s, int64_t nElem, const double* coords,
                          const int32_t* conn, double* charlen, double* vol, int64_t* badElem);
// mode: 0 = K64 + F64, 1 = F64 only, 2 = K32 + F64.  lift: apply the iteration-0
// Dirichlet lift.  elems: element ids (a colour class) of length n, or NULL with
// probe_e >= 0 for the single-element probe.
// variant (element types other than Hexa8): 0 = 256 lanes per workgroup, 1 = 128, 2 = 64,
// one group of elements per workgroup; 3-5 = the same sizes with workgroups that walk
// their colour class (a grid of what the GPU holds at once)
constexpr int kAsmVariants = 6;
// residual-only assembly of Newton iterations after the first as kernel launches only
// (capturable); asm_prepare_residual must have run first (sfem_api.cpp)
int asm_prepare_residual(sfem_ctx* c);
int asm_residual_kernels(sfem_ctx* c);
hipError_t launch_assemble(int et, hipStream_t s, const AsmView& v, const ElemParams& p,
                           const int32_t* elems, int64_t n, int mode, int lift, int64_t probe_e, int variant = 0);

// Launchers implemented in sfem_linalg.hip (see there for semantics)
struct LinAlg;
struct Batch;

// Event-pair timing of launch groups by kernel class (SFEM_KT_*).
struct KernelTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;          // free events
  std::vector<hipEvent_t> open;          // begin events awaiting their end, by kind
  struct Pending { int kind; hipEvent_t a, b; bool counted; };
  std::vector<Pending> pending;
  int64_t count[SFEM_KT_NKINDS] = {0, 0, 0};
  double ms[SFEM_KT_NKINDS] = {0.0, 0.0, 0.0};
  hipEvent_t begin_ev[SFEM_KT_NKINDS] = {nullptr, nullptr, nullptr};

  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  void begin(int kind, hipStream_t s) {
    if (!on) return;
    begin_ev[kind] = get();
    (void)hipEventRecord(begin_ev[kind], s);
  }
  // counted = false adds the bracket's time to the previous counted one of its kind (a
  // second part of the same operation)
  void end(int kind, hipStream_t s, bool counted = true) {
    if (!on || !begin_ev[kind]) return;
    hipEvent_t b = get();
    (void)hipEventRecord(b, s);
    pending.push_back({kind, begin_ev[kind], b, counted});
    begin_ev[kind] = nullptr;
  }
  // call only when the stream has drained past every recorded event
  void resolve() {
    for (auto& p : pending) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
        ms[p.kind] += t;
        if (p.counted) count[p.kind] += 1;
      }
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  void reset() {
    resolve();
    for (int k = 0; k < SFEM_KT_NKINDS; ++k) { count[k] = 0; ms[k] = 0.0; }
  }
  ~KernelTimer() {
    for (auto& p : pending) { pool.push_back(p.a); pool.push_back(p.b); }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace sfem

// Opaque ABI context.
struct sfem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int et = 0;
  sfem::ElemDims dims{};
  int64_t nNode = 0, nElem = 0;
  int precision = SFEM_PREC_FP64;
  int predictor = SFEM_PRED_PREVIOUS;
  double inner_theta = 0.0;      // sfem_set_inner_tolerance: 0 = fixed inner rtol
  double inner_rtol_max = 0.5;
  int status = 0;  // bit flags: 1 prepared, 2 pattern, 4 params, 8 assembled

  // host mesh
  std::vector<double> coords;
  std::vector<int32_t> conn;
  // Dirichlet data: raw list (file order) and the unique dofs with last value and multiplicity
  std::vector<int32_t> dbc_ind_raw;
  std::vector<double> dbc_val_raw;
  std::vector<int32_t> dbc_ind;   // unique node*ndof+dof
  std::vector<double> dbc_val;    // value of the last listing
  std::vector<int32_t> dbc_mult;  // number of listings

  // DOF numbering and pattern
  int64_t nfree = 0, nnz = 0;
  std::vector<int32_t> nodeDof;       // nNode*ndof, -1 prescribed
  std::vector<int32_t> assyForSoln;   // nfree
  std::vector<int64_t> rowptr;        // nfree+1
  std::vector<int32_t> colidx;        // nnz
  std::vector<int32_t> colorPtr;      // ncolors+1
  int ncolors = 0;

  // parameters
  double elemData[11] = {0};
  int tis = 0;
  double rhoInf = 0.0, dt = 0.0;
  double td[16] = {0};

  // device buffers
  double* d_coords = nullptr;
  int32_t* d_conn = nullptr;
  double* d_charlen = nullptr;
  double* d_vol = nullptr;
  int64_t* d_bad = nullptr;
  int32_t* d_nodeInfo = nullptr;
  int32_t* d_colofs = nullptr;      // element -> nonzero map in colour-slot order
  int32_t* d_sconn = nullptr;       // connectivity in colour-slot order
  int32_t* d_elemByColor = nullptr;
  int32_t* d_rowptr32 = nullptr;   // int32 copy for kernels (nnz < 2^31 is checked)
  int32_t* d_colidx = nullptr;
  int32_t* d_assy = nullptr;
  int32_t* d_dbc_ind = nullptr;
  double* d_dbc_val = nullptr;
  int32_t* d_dbc_mult = nullptr;
  double* d_vals = nullptr;
  float* d_vals32 = nullptr;
  // node-pair block form of the tangent (element types whose loop assembles it: kb_on);
  // the CSR values are then derived from it on demand (csr_version = la_K_version of the
  // last conversion)
  bool kb_on = false;
  int64_t nbpairs = 0;                  // node-pair blocks of the pattern
  std::vector<int64_t> blkPtr;          // node block -> its first pair (nodes with free dofs, node order)
  int32_t* d_blkofs = nullptr;          // element node pair -> block, colour-slot order
  int32_t* d_natofs = nullptr;          // per block: CSR column offset in its rows | free columns << 24
  int64_t* d_blkPtr = nullptr;
  double* d_kb = nullptr;
  float* d_kb32 = nullptr;
  int64_t csr_version = -1;
  bool kb_fp32 = false;                 // the last tangent went to d_kb32 (mixed precision)
  int asm_variant[3] = {-1, -1, -1};    // element-loop variant per assembly mode (-1: not chosen)
  double* d_rhs = nullptr;
  double* d_du = nullptr;
  // SolutionData vectors (nNode*ndof each)
  double *d_soln = nullptr, *d_prev = nullptr, *d_prev2 = nullptr, *d_prev3 = nullptr,
         *d_prev4 = nullptr, *d_dot = nullptr, *d_dotPrev = nullptr, *d_cur = nullptr,
         *d_dotCur = nullptr, *d_applied = nullptr;
  double* d_elem_mu = nullptr;     // per-element viscosity of a batch of meshes, or NULL
  double* d_probeK = nullptr;
  double* d_probeF = nullptr;

  sfem::LinAlg* la = nullptr;   // Krylov workspace and preconditioners
  sfem::Batch* batch = nullptr; // many independent meshes in one context (sfem_set_batch)
  int64_t la_K_version = 0;      // bumped whenever the tangent is re-assembled

  sfem::KernelTimer kt;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  double last_asm_ms = 0.0, last_solve_ms = 0.0;
  double* h_pinned = nullptr;   // small pinned host buffer for scalar readback

  // learned initial iterate (sfem_learned.hip): candidate buffer and the judge's record
  sfem_learner* learner = nullptr;
  double* d_cand = nullptr;               // 2 x nNode*ndof: learned, quadratic candidates
  int64_t pred_count[3] = {0, 0, 0};       // steps started from the linear / quadratic / learned start
  double pred_norms[3] = {0.0, 0.0, 0.0};  // their ||F|| at the last step
};

namespace sfem {
int learned_candidate(sfem_ctx* c, const double* s0, const double* s1, const double* s2, double* out);
}
