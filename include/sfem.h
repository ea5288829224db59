/*
 * sfem.h -- C ABI of the MI355X-native stabilised-FEM hot path (libsfem_hip.so).
 *
 * The reference (chennachaos/cfdsfemmpi) has two seams on this path, both
 * called from StabFEM: the element seam (ElementBase::prepareElemData /
 * calcStiffnessAndResidual, src/base/ElementBase.h:81-88) and the solver seam
 * (SolverPetsc, src/base/SolverPetsc.h:15-65).  This ABI replaces both behind
 * one opaque context that owns all device data; each entry point below cites
 * the reference interface it stands in for.
 *
 * Conventions: every int-returning call returns SFEM_OK (0) or a negative
 * SFEM_ERR_* code and never throws; buffers are caller-owned host memory;
 * node and element ids are 0-based; solution vectors are node-major AoS
 * (node*ndof + dof), as SolutionData holds them.  One context drives one
 * device stream and is not thread-safe.  sfem_last_error() gives the message
 * of the most recent failure on the calling thread.
 */
#ifndef SFEM_H
#define SFEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sfem_ctx sfem_ctx;

enum {
  SFEM_OK = 0,
  SFEM_ERR_ARG = -1,        /* bad argument or size */
  SFEM_ERR_STATE = -2,      /* call out of order (SolverPetsc status machine, SolverPetsc.h:13) */
  SFEM_ERR_NEGJAC = -3,     /* negative Jacobian (the reference exit()s, Hexa8Node.cpp:79-84) */
  SFEM_ERR_HIP = -4,        /* HIP runtime error */
  SFEM_ERR_DIVERGED = -5,   /* Krylov breakdown or non-finite residual (SolverPetsc.cpp:264-270) */
  SFEM_ERR_MAXIT = -6,      /* Krylov hit max iterations before rtol */
  SFEM_ERR_NOMEM = -7,      /* device or host allocation failed */
  SFEM_ERR_NOTCONV = -8,    /* Newton loop hit its iteration cap (the reference moves on silently) */
  SFEM_ERR_UNSUPPORTED = -9 /* element type / option not available */
};

/* Precision of the tangent and Krylov solve (residual and update are always FP64).
 * SFEM_PREC_MIXED: FP32 element tangents, FP32 CSR, FP32 Krylov vectors and ILU(0).
 * SFEM_PREC_STORE32: FP64 tangent and FP64 ILU(0) factorisation, the node-block tangent
 * and the factor then kept in FP32 (factor_bf16: BF16) for the Krylov products, which
 * run on FP64 vectors with FP64 accumulation -- half the matrix and factor bytes per
 * iteration without FP32 rounding of the Krylov vectors, for systems too ill-conditioned
 * for FP32 Krylov (ILU(0) only; with other preconditioners it is the FP64 path). */
enum { SFEM_PREC_FP64 = 0, SFEM_PREC_MIXED = 1, SFEM_PREC_STORE32 = 2 };

/* Krylov methods and preconditioners (vocabulary of the PETSc options files).
 * SFEM_KSP_PREONLY with SFEM_PC_LU is the direct solve of the reference's test decks
 * (-ksp_type preonly -pc_type lu, testdata/ldc2dquad10/petsc_options.dat:1-3, MUMPS in
 * the reference): a banded LU with partial pivoting on the device, in the reverse
 * Cuthill-McKee order of the free-DOF graph, factorised once per new tangent (FP64
 * only; sized for small decks: the band takes nfree x (2 kl + ku + 1) doubles). */
enum { SFEM_KSP_BCGS = 0, SFEM_KSP_GMRES = 1, SFEM_KSP_PREONLY = 2 };
enum { SFEM_PC_NONE = 0, SFEM_PC_JACOBI = 1, SFEM_PC_BJACOBI4 = 2, SFEM_PC_ILU0 = 3, SFEM_PC_LU = 4 };

typedef struct sfem_solver_opts {
  int ksp_type;      /* SFEM_KSP_* (-ksp_type) */
  int pc_type;       /* SFEM_PC_*  (-pc_type / -sub_pc_type) */
  double rtol;       /* -ksp_rtol: stop when ||b - A x||_2 <= rtol ||b||_2 */
  int max_it;        /* -ksp_max_it */
  int restart;       /* -ksp_gmres_restart */
  int reorder;       /* ILU(0) node ordering: 0 the matrix's own order, 1 reverse Cuthill-McKee
                        (PETSc's -pc_factor_mat_ordering_type rcm), 2 greedy multicolour
                        (one fully parallel level per colour), 3 tiled multicolour (the node
                        blocks of one element form a tile, tiles are coloured; one launch per
                        colour and sweep, sequential inside a tile) */
  int factor_bf16;   /* SFEM_PREC_MIXED / SFEM_PREC_STORE32 only: keep the ILU(0) factors in
                        bfloat16 (computed in FP32 / FP64, rounded to nearest even) to halve
                        the preconditioner's bytes */
  double atol;       /* -ksp_atol: the solve also stops once ||b - A x||_2 <= atol (PETSc's
                        default test, max(rtol ||b||, atol)); 0 = relative test only.
                        sfem_solve only (the batched step rejects atol > 0) */
} sfem_solver_opts;

typedef struct sfem_step_stats {
  int newton_iters;      /* Newton iterations taken (assemblies) */
  int krylov_iters;      /* total Krylov iterations in the step */
  double rhs_norm0;      /* ||rhs_free||_2 at iteration 0 */
  double rhs_norm;       /* last ||rhs_free||_2 (the convergence measure) */
  double t_assembly_ms;  /* device time of the element loops */
  double t_solver_ms;    /* device time of the Krylov solves */
} sfem_step_stats;

/* Library version string. */
const char* sfem_version(void);
/* Message of the most recent error on this thread ("" if none). */
const char* sfem_last_error(void);

/* Create a context from a mesh.  Replaces StabFEM::readInputData's in-memory
 * result + prepareInputData (StabFEM.cpp:80-264, 328-386): coords is
 * nNode*ndim, conn nElem*npElem (0-based).  Supported (ndim, npElem):
 * (2,3) Tria3, (2,4) Quad4, (3,4) Tetra4, (3,8) Hexa8.  device < 0 picks
 * the current device.  Returns NULL on failure (see sfem_last_error). */
sfem_ctx* sfem_create(int ndim, int npElem, int64_t nNode, int64_t nElem, const double* coords,
                      const int32_t* conn, int device);
void sfem_destroy(sfem_ctx* ctx);

/* Dirichlet list in file order (StabFEM.cpp:209-225 DirichletBCs): node,
 * dof, value; duplicates are honoured as the reference does (the value of
 * the last listing, the iteration-0 update applied once per listing). */
int sfem_set_dirichlet(sfem_ctx* ctx, int64_t nDBC, const int32_t* node, const int32_t* dof,
                       const double* value);

/* prepareElemData for all elements (Hexa8Node.cpp:31-106 and siblings):
 * volumes, characteristic lengths, negative-Jacobian check on the device. */
int sfem_prepare_elements(sfem_ctx* ctx);

/* DOF numbering + CSR pattern + element-to-nonzero map + element colouring
 * (StabFEM2.cpp:34-319 prepareMatrixPattern, SolverPetsc::initialise
 * SolverPetsc.cpp:30-116).  Returns nfree and nnz. */
int sfem_build_pattern(sfem_ctx* ctx, int64_t* nfree, int64_t* nnz);

/* Copy out the pattern for bit-exact checks: rowptr[nfree+1], colidx[nnz],
 * assyForSoln[nfree] (StabFEM2.cpp:172-183).  Any pointer may be NULL. */
int sfem_get_pattern(sfem_ctx* ctx, int64_t* rowptr, int32_t* colidx, int32_t* assyForSoln);

/* Material and time-integration parameters (StabFEM::readControlParameters,
 * StabFEM.cpp:268-322): elemData[11] (0 rho, 1 mu, 8-10 SUPG/PSPG/LSIC
 * multipliers), tis / rhoInf / dt for SetTimeParametersFluid
 * (UtilitiesGeneral.cpp:5-209). */
int sfem_set_params(sfem_ctx* ctx, const double* elemData, int tis, double rhoInf, double dt);

/* SFEM_PREC_FP64 (default), SFEM_PREC_MIXED or SFEM_PREC_STORE32. */
int sfem_set_precision(sfem_ctx* ctx, int precision);

/* Newton initial iterate of each step: SFEM_PRED_PREVIOUS (the previous step's
 * solution, as the reference), SFEM_PRED_EXTRAPOLATE (2 solnPrev - solnPrev2, the
 * solnExtrap that SolutionData::timeUpdate computes and the reference leaves unused,
 * SolutionData.cpp:75), SFEM_PRED_QUADRATIC (3 solnPrev - 3 solnPrev2 + solnPrev3; no
 * counterpart in the reference) or SFEM_PRED_LEARNED (below).  It changes the
 * Newton/Krylov iteration counts, not the answer. */
enum { SFEM_PRED_PREVIOUS = 0, SFEM_PRED_EXTRAPOLATE = 1, SFEM_PRED_LEARNED = 2, SFEM_PRED_QUADRATIC = 3 };
int sfem_set_predictor(sfem_ctx* ctx, int predictor);

/* ---- Learned initial iterate (2D meshes: u, v, p per node).  No counterpart in the
 * reference, whose Newton loop starts from the previous solution (StabFEM2.cpp:593-832).
 * A per-node MLP 12 -> 64 -> 64 -> 3 (SiLU) over the last three states of a node and the
 * mean of its 1-ring node neighbours predicts the next increment's correction over the
 * quadratic extrapolation for u and v (the pressure keeps the quadratic extrapolation;
 * the third output is not used and not trained), or for u, v and p
 * (sfem_learner_set_components): start = s_n + 2 d1 - d2 + sigma * tau * lambda * MLP(x / lambda),
 * d1 = s_n - s_{n-1}, d2 = s_{n-1} - s_{n-2}, x the 12 sigma-scaled inputs, sigma the
 * per-component RMS of d1 over all nodes, lambda the node's activity scale (RMS of its six
 * sigma-scaled own increments + 1e-3), tau a fixed output scale set before the first Adam
 * step to the RMS of (d0 - 2 d1 + d2) / (sigma lambda) over 65,536 hashed samples
 * (minibatch tag 2^32 - 1), so that the outputs are O(1).  With SFEM_PRED_LEARNED each
 * step evaluates the residual norm at the linearly extrapolated, the quadratically
 * extrapolated and the learned start (one residual-only element pass each; the residual
 * is affine in the state, so these are the iteration-0 norms) and starts Newton from the
 * smallest: the predictor changes iteration counts, and the Newton test decides
 * convergence as before. */
#define SFEM_LEARNER_NPARAM 5190 /* W1 64x12, b1 64, W2 64x64, b2 64, W3 3x64, b3 3 (row-major),
                                     tau 3: the fixed output scale (not trained) */
typedef struct sfem_learner sfem_learner;
/* A learner for the mesh of ctx (pattern built; the node graph and the free-DOF mask are
 * taken from it) with room for `capacity` recorded states.  NULL on failure. */
sfem_learner* sfem_learner_create(sfem_ctx* ctx, int64_t capacity, uint64_t seed);
void sfem_learner_destroy(sfem_learner* learner);
/* Append ctx's current solution to the training store; new_trajectory != 0 starts a new
 * trajectory (windows of four consecutive states never cross trajectories). */
int sfem_learner_record(sfem_learner* learner, sfem_ctx* ctx, int new_trajectory);
int sfem_learner_info(sfem_learner* learner, int64_t* snapshots, int64_t* windows, int64_t* adam_steps);
/* iters Adam steps (lr, betas 0.9 / 0.999, eps 1e-8) on minibatches of `batch` (window,
 * node) samples; loss[iters] (may be NULL) receives each minibatch's loss, the mean over
 * samples and components of (y - t)^2 with t = (d0 - 2 d1 + d2) / (sigma lambda tau), d0 the
 * recorded next increment (prescribed components, and the pressure unless it is corrected,
 * count as zero).  Sample k of Adam step a is drawn by
 * h = splitmix64(seed ^ splitmix64(a * 2^32 + k)): window (h >> 32) mod nwindows, node
 * (h mod 2^32) mod nNode. */
int sfem_learner_train(sfem_learner* learner, int iters, int batch, double lr, double* loss);
/* Gradient (SFEM_LEARNER_NPARAM floats) and loss of an explicit list of samples:
 * window[k] is the index of the newest of the three input states in the store. */
int sfem_learner_gradient(sfem_learner* learner, int n, const int32_t* window, const int32_t* node,
                          float* grad, double* loss);
/* Read (get) and/or replace (set) the FP32 master weights and tau; set also resets Adam
 * and fixes tau. */
int sfem_learner_weights(sfem_learner* learner, float* get, const float* set);
/* Components the network corrects and is trained on: 2 (u, v; the default, the pressure
 * keeps the quadratic extrapolation) or 3 (u, v, p).  A change takes the output scales
 * tau again before the next Adam step (sfem_learner_weights fixes them, as before). */
int sfem_learner_set_components(sfem_learner* learner, int ncomp);
/* Attach a learner (NULL detaches; the learner must outlive its use by ctx).  A learner
 * is not thread-safe: record, train and the steps of contexts it is attached to are
 * called from one host thread (each call returns with its device work finished). */
int sfem_set_learner(sfem_ctx* ctx, sfem_learner* learner);
/* The learned start computed from ctx's current solution and its two predecessors
 * (nNode*3 doubles, host). */
int sfem_learned_candidate(sfem_ctx* ctx, double* out);
/* Under SFEM_PRED_LEARNED: counts[3] = steps started from the linear, the quadratic and
 * the learned start, and last_norms[3] = their ||F|| at the last step (either may be NULL). */
int sfem_predictor_stats(sfem_ctx* ctx, int64_t* counts, double* last_norms);

/* Rebuild the preconditioner only on every lag-th new tangent (lag >= 1; 1, the default,
 * rebuilds it whenever the tangent changes, as PCSetUp does after each assembly in
 * SolverPetsc::factoriseAndSolve, SolverPetsc.cpp:284-298).  With lag > 1 the Krylov
 * products still use the current tangent; only the ILU(0) factor (or Jacobi diagonal) is
 * reused.  No counterpart in the reference. */
int sfem_set_pc_lag(sfem_ctx* ctx, int lag);

/* Coarse pressure correction of the ILU(0) preconditioner, additive
 * (z = M_ILU^{-1} r + Z E^{-1} Z^T r) or with multiplicative != 0 multiplicative
 * (z0 = M_ILU^{-1} r, z = z0 + Z E^{-1} Z^T (r - A z0); inside BiCGStab the product A z0 is
 * the one the iteration takes anyway, corrected by A Z y), with Z the piecewise-constant pressure on k boxes
 * per direction of the mesh's bounding box (k^dim <= 1024; every box at least one element
 * wide) and E = Z^T A Z formed on the device from the tangent at each preconditioner
 * rebuild.  k = 0 (the default) switches it off.  It changes the preconditioned operator,
 * not the linear system or its tolerance; it cuts BiCGStab iterations on meshes whose
 * pressure error ILU(0) damps only locally.
 * Under SFEM_PREC_MIXED the FP32 Krylov vectors cannot hold the global constant pressure
 * mode (E's smallest eigenvalue, set by the prescribed pressures alone, lies orders of
 * magnitude below the others) beside the rest of a preconditioned vector, and the FP32
 * tangent does not resolve it either: there the correction uses B, the leading block of
 * [E b1; b1^T 0]^{-1} (its outputs sum to zero), and sfem_step (and the host driver)
 * adds a coarse step in FP64 before each inner solve -- x += Z E^{-1} Z^T F(x), then the
 * residual again (sfem_coarse_step).  On a batch (sfem_set_batch, before or after this
 * call) every mesh gets its own k boxes per direction on its own bounding box, so E is
 * block diagonal, and sfem_step_batch takes the coarse step per mesh.  No counterpart in
 * the reference, which preconditions with PETSc's ILU(0) alone
 * (SolverPetsc::factoriseAndSolve, SolverPetsc.cpp:284-298). */
int sfem_set_coarse_correction(sfem_ctx* ctx, int boxes_per_dim, int multiplicative);
/* info[6] = {k, coarse unknowns, free pressure DOFs, restriction chunks, 1 if the last E
 * was singular (correction off until the next rebuild), 1 if built}. */
int sfem_coarse_info(sfem_ctx* ctx, int64_t* info);
/* which = 0: E^{-1}; which = 1: B (mixed precision only); of the last rebuild, row-major
 * (cap >= unknowns^2 doubles; block diagonal on a batch). */
int sfem_coarse_inverse(sfem_ctx* ctx, int which, double* out, int cap);
/* The Newton loop's coarse step: the increment (sfem_get_increment, applied by
 * sfem_update_solution) = Z E^{-1} Z^T rhs of the last assembly, in FP64, with the
 * preconditioner brought up to date for opts first.  Needs the coarse correction. */
int sfem_coarse_step(sfem_ctx* ctx, const sfem_solver_opts* opts);

/* Inexact Newton for the mixed-precision paths (SFEM_PREC_MIXED, SFEM_PREC_STORE32): in
 * sfem_step each Newton iteration's inner solve stops at
 *   rtol_k = min(rtol_max, max(opts->rtol, theta * newton_tol / ||F_k||)),
 * so that no solve is asked for more reduction than the Newton test still needs, and
 * none for more than opts->rtol, the deepest reduction an FP32 solve is trusted with
 * (theta < 1 leaves room for the FP32 solve's true residual lagging its recurrence
 * residual; a solve that stops short is followed by another Newton iteration, as
 * before).  0 < rtol_max < 1 keeps every solve reducing its residual.  theta = 0 (the
 * default) solves every iteration to opts->rtol.  The FP64 path ignores it.  No
 * counterpart in the reference, whose KSP tolerance is fixed by the options file. */
int sfem_set_inner_tolerance(sfem_ctx* ctx, double theta, double rtol_max);

/* Start a time step (StabFEM2.cpp:632-648): td[] from SetTimeParametersFluid,
 * SolutionData::timeUpdate (SolutionData.cpp:64-78), and
 * assignBoundaryConditions (StabFEM.cpp:391-410) with the lid ramp
 * timeFact = timeNow <= 1 ? timeNow : 1 (StabFEM2.cpp:636-644); timeNow is the
 * caller's accumulated time (timeNow += dt, as the reference accumulates it). */
int sfem_time_update(sfem_ctx* ctx, double timeNow);

/* One element loop of the Newton iteration `iter` (StabFEM2.cpp:656-740):
 * updateIterStep, zeroMtx, calcStiffnessAndResidual for every element, the
 * iteration-0 BC lift, scatter-add into the device CSR and RHS
 * (assembleMatrixAndVectorSerial, SolverPetsc.cpp:304-314), the iteration-0
 * soln[dbc] += solnApplied, and ||rhs_free||_2 into *rhs_norm.
 * want_K = 0 assembles the residual only (valid for iter > 0). */
int sfem_assemble(sfem_ctx* ctx, int iter, int want_K, double* rhs_norm);

/* Element parity probe: Ke (row-major nsize x nsize) and Fe of element e at
 * the current state, exactly as the assembly kernel computes them before the
 * BC lift (calcStiffnessAndResidual, ElementBase.h:87-88). */
int sfem_element_KF(sfem_ctx* ctx, int64_t e, double* Ke, double* Fe);

/* Copy out the assembled CSR values[nnz] and rhs[nfree] (FP64 mode; in mixed
 * mode the FP32 values widened to double).  Either pointer may be NULL. */
int sfem_get_system(sfem_ctx* ctx, double* vals, double* rhs);

/* Krylov solve of the assembled system into the device increment
 * (SolverPetsc::factoriseAndSolve / solve, SolverPetsc.cpp:224-298; zero
 * initial guess as VecZeroEntries at :249). */
int sfem_solve(sfem_ctx* ctx, const sfem_solver_opts* opts, int* iters, double* relres);

/* Symbolic set-up of the solver for these options ahead of the time loop (the
 * ILU(0) ordering, its permuted pattern and level sets) -- the part of
 * PCSetUp that depends only on the pattern (SolverPetsc::initialise,
 * SolverPetsc.cpp:95-111).  Optional: sfem_solve does it on first use. */
int sfem_setup_solver(sfem_ctx* ctx, const sfem_solver_opts* opts);

/* soln[assyForSoln[i]] += du[i] (StabFEM2.cpp:785-788). */
int sfem_update_solution(sfem_ctx* ctx);

/* One full time step: sfem_time_update + the Newton loop of
 * StabFEM2.cpp:654-799 (at most max_iter iterations, stop when
 * ||rhs_free||_2 < newton_tol before solving).  Returns SFEM_ERR_NOTCONV if
 * the cap is hit (stats are still filled). */
int sfem_step(sfem_ctx* ctx, double timeNow, const sfem_solver_opts* opts, double newton_tol,
              int max_iter, sfem_step_stats* stats);

/* Solution vectors (node-major, nNode*ndof): soln, and the solver increment du[nfree]. */
int sfem_get_solution(sfem_ctx* ctx, double* soln);
int sfem_set_solution(sfem_ctx* ctx, const double* soln);
int sfem_get_increment(sfem_ctx* ctx, double* du);

/* Set the full SolutionData state (each nNode*ndof; NULL leaves a vector
 * unchanged) -- used to probe the element kernels at arbitrary states. */
int sfem_set_state(sfem_ctx* ctx, const double* soln, const double* solnPrev,
                   const double* solnCur, const double* solnDotCur, const double* solnDot,
                   const double* solnDotPrev);

/* Zero every SolutionData vector (soln, solnPrev..solnPrev4, solnDot, solnDotPrev,
 * solnCur, solnDotCur, solnApplied) -- SolutionData::initialise (SolutionData.cpp:16-38). */
int sfem_reset_state(sfem_ctx* ctx);

/* Per-kernel-class device timing with HIP events recorded on the context's
 * stream around each launch group of that class.  Enabling (or re-enabling)
 * clears the counters.  No counterpart in the reference (its timers are the
 * MPI_Wtime phase totals of StabFEM.cpp:711-727). */
enum { SFEM_KT_ASSEMBLY = 0, SFEM_KT_SPMV = 1, SFEM_KT_PC = 2, SFEM_KT_NKINDS = 3 };
int sfem_enable_kernel_timing(sfem_ctx* ctx, int enable);
int sfem_kernel_timing(sfem_ctx* ctx, int kind, int64_t* launches, double* total_ms);

/* The ILU(0) factor of the last solve (pc = ILU0), in the factor's node ordering:
 * perm[i] = DOF row of factor row i (n entries), rowptr (n+1) and colidx (nnz) of the
 * reordered pattern, and vals (nnz): L strictly below the diagonal (unit diagonal
 * implied), U on and above it, FP32/BF16 factors widened.  Any output may be NULL;
 * *nnz receives the count.  The reference's factor lives inside PETSc's PCILU
 * (SolverPetsc.cpp:234-277) and is not exposed; this exists for parity checks. */
int sfem_get_ilu(sfem_ctx* ctx, int64_t* nnz, int32_t* perm, int32_t* rowptr, int32_t* colidx, double* vals);

/* z = M^{-1} r with the preconditioner of the last solve (r, z: nfree entries in DOF
 * order; FP32/BF16 factors are applied to FP64 vectors).  The reference applies its
 * preconditioner only inside KSPSolve (PCApply, SolverPetsc.cpp:234-277); this exists
 * for parity checks of the triangular solves. */
int sfem_pc_apply(sfem_ctx* ctx, const double* r, double* z);

/* sfem_pc_apply without allocating: no counterpart in the reference.  For callers that
 * cannot allocate device memory (tight serving loops, code that must not call hipMalloc):
 * the same z = M^{-1} r as sfem_pc_apply, whose staging vectors (r and z on the device,
 * and for ILU(0) their padded block-layout copies) come from a fixed 64 MiB device array
 * that each GPU holds for the life of the process, instead of hipMalloc / hipFree per
 * call.  Limit: 2 nfree + 2 npad <= 8,388,608 doubles (npad = 4 x node blocks for
 * ILU(0), else 0), about 2M free DOFs; larger systems get SFEM_ERR_ARG.  Calls that
 * overlap on one GPU share the array and race; each call returns after its device work. */
int sfem_pc_apply_noalloc(sfem_ctx* ctx, const double* r, double* z);

/* Batched meshes.  The context's mesh is a disjoint union of nmesh meshes, nodes and
 * elements numbered mesh by mesh; elem_mesh[e] is the mesh of element e and mu[m] the
 * viscosity of mesh m (elemData[1] of its control file).  Call after
 * sfem_build_pattern.  No counterpart in the reference, which runs one mesh per job. */
int sfem_set_batch(sfem_ctx* ctx, int32_t nmesh, const int32_t* elem_mesh, const double* mu);

/* One time step of every mesh of the batch: the reference's Newton loop
 * (StabFEM2.cpp:654-799) per mesh, a mesh stopping when ||F_m|| < newton_tol and staying
 * frozen while the others iterate; each solve is BiCGStab + ILU(0) with per-mesh scalars
 * and per-mesh convergence (||r_m|| <= rtol ||b_m||), one iteration of all meshes being
 * one replayed HIP graph.  Per mesh: newton_iters (assemblies), krylov_iters (summed
 * over the step's solves) and the last ||F_m||.  In mixed precision a mesh whose FP32
 * solve stalls (no 10% residual gain in 100 iterations), diverges or reaches max_it
 * hands its best iterate to the next FP64 Newton iteration (counted by
 * sfem_solver_stats); in FP64 such a mesh fails the step with SFEM_ERR_MAXIT. */
int sfem_step_batch(sfem_ctx* ctx, double timeNow, const sfem_solver_opts* opts, double newton_tol,
                    int max_newton, int32_t* newton_iters, int32_t* krylov_iters, double* rhs_norm);

/* Replay the batched BiCGStab iteration as a HIP graph (1, default) or launch it kernel
 * by kernel (0). */
int sfem_batch_use_graph(sfem_ctx* ctx, int enable);

/* Narrow the batched iteration to the meshes still iterating (1, default): each time the
 * block rows of the unfinished meshes fall below half of those the iteration covers, its
 * products, ILU(0) sweeps and vector updates are restricted to them (identical results,
 * less work when mesh sizes or iteration counts differ).  0: every iteration covers
 * every mesh. */
int sfem_batch_narrowing(sfem_ctx* ctx, int enable);

/* Sizes of the ILU(0) node-block form (0 before an ILU(0) solver is set up): block
 * rows, stored neighbour blocks, and stored values (padded 4-wide blocks, m rows each).
 * No counterpart in the reference. */
int sfem_block_info(sfem_ctx* ctx, int64_t* nblk, int64_t* npairs, int64_t* nvals);

/* Tiles of a tiled multicolour ILU(0) plan (reorder = 3; zeros otherwise): number of
 * tiles, tile colours, and the factor values inside tiles, which the sweeps apply through
 * the tiles' inverse triangles instead.  No counterpart in the reference (PETSc's PCILU
 * keeps its ordering internal); for bandwidth accounting. */
int sfem_tile_info(sfem_ctx* ctx, int64_t* ntiles, int64_t* ncolors, int64_t* tile_vals);

/* Mixed-precision solver events since the context was created.  fp32_fallbacks: solves
 * in which the FP32 BiCGStab stalled (no 10% gain of its best residual over 100
 * iterations) and which FP32 GMRES finished from a zero start.  fp32_partial (may be
 * NULL): FP32 solves that ended above rtol (max_it reached with a reduced, finite
 * residual; in batched steps also a stalled or diverged mesh solve, which keeps its best
 * iterate) and handed their iterate to the FP64 Newton iteration, whose own test decides
 * convergence.  No counterpart in the reference (PETSc reports a diverged KSP). */
int sfem_solver_stats(sfem_ctx* ctx, int64_t* fp32_fallbacks, int64_t* fp32_partial);

/* Kernel shapes chosen by the solver's autotuner, as out[8] = {SpMV lanes per node
 * (block row) G, columns (neighbour blocks) in flight per lane, order flags (bit 0:
 * XCD-contiguous workgroups, bit 1: block rows visited in the original node order),
 * number of tuned ILU(0) levels} for FP64 then FP32; 0 where not tuned yet.  No
 * counterpart in the reference. */
int sfem_tuning(sfem_ctx* ctx, int* out);

/* Per-level shapes of the ILU(0) apply for one precision (fp32 = 0 FP64, 1 FP32):
 * out[3*l .. 3*l+2] = {G, U, XCD order} for the forward levels, then the backward
 * levels; *nlevels receives the count, at most cap entries are written.  No
 * counterpart in the reference. */
int sfem_tuning_levels(sfem_ctx* ctx, int fp32, int* out, int cap, int* nlevels);

/* The element loop's variant per assembly mode (out[0] tangent + residual FP64, out[1]
 * residual only, out[2] FP32 tangent + FP64 residual), -1 where not chosen yet.  For
 * Tria3, Quad4 and Tetra4 the first assembly of a mode under SFEM_TUNE_TIMED times six
 * variants over the whole colour loop and keeps the fastest: 256, 128 or 64 lanes per
 * workgroup (16, 8 or 4 elements side by side), each with one group of elements per
 * workgroup (0-2) or with workgroups that walk their colour class (3-5).  Under
 * SFEM_TUNE_FIXED, variant 0.  All variants give the same result (no two elements of a
 * colour class add to the same entry).  The Hexa8 loop has one form (0).  No
 * counterpart in the reference (PETSc assembly is serial per rank). */
int sfem_element_tuning(sfem_ctx* ctx, int* out);

/* How the solver picks the kernel shapes of its SpMV and ILU(0) sweeps (lanes per block
 * row, blocks in flight per lane, workgroup order).  SFEM_TUNE_TIMED (the default): on
 * the first solve of a precision and ILU(0) plan every shape is timed with HIP events on
 * the context's stream and the fastest kept, per sweep launch.  SFEM_TUNE_FIXED: the
 * library's default shapes (or a table installed by sfem_set_tuning_table), no timing.
 * A shape fixes the summation order of the products, and FP32 Krylov iterates (and
 * their iteration counts) follow that rounding: mixed-precision results are the same
 * from run to run and device to device only with fixed shapes.  Setting a mode makes the
 * next solve of each precision install its shapes again.  No counterpart in the
 * reference (PETSc's kernels have no shape parameters). */
enum { SFEM_TUNE_TIMED = 0, SFEM_TUNE_FIXED = 1 };
int sfem_set_tuning_mode(sfem_ctx* ctx, int mode);

/* The shape table of one precision (fp32 = 0 FP64 solves, 1 FP32 solves) as integer
 * codes: {CSR SpMV, block SpMV, forward sweep launches nL, backward launches nU, nL
 * forward codes, nU backward codes}.  get copies it out (*n receives its length, at most
 * cap entries are written); set installs it for the current ILU(0) plan (the launch
 * counts must match it) and switches to SFEM_TUNE_FIXED, so that a table recorded by one
 * run is replayed exactly by another.  No counterpart in the reference. */
int sfem_get_tuning_table(sfem_ctx* ctx, int fp32, int* table, int cap, int* n);
int sfem_set_tuning_table(sfem_ctx* ctx, int fp32, const int* table, int n);

/* Bandwidths of the direct solver's reordered tangent (kl sub- and ku super-diagonals;
 * zeros before an SFEM_PC_LU solve was set up).  No counterpart in the reference. */
int sfem_direct_info(sfem_ctx* ctx, int* kl, int* ku);

/* Sizes and counts of the context. */
int sfem_info(sfem_ctx* ctx, int64_t* nfree, int64_t* nnz, int* ncolors, int* ndof);

/* Device time (ms) of the last assembly's element kernels and of the last
 * sfem_solve, measured with HIP events on the context's stream. */
int sfem_last_timings(sfem_ctx* ctx, double* assembly_ms, double* solve_ms);

#ifdef __cplusplus
}
#endif
#endif
