/*
 * sfem_oracle.h -- CPU restatement of the cfdsfemmpi stabilised-FEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path (cfdsfemmpi_amd/csrc).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker or as
 * the reported CPU baseline -- never as the thing measured or shipped.
 *
 * Every function follows a named reference file:line (paths relative to the
 * reference repository root).  Arithmetic is written loop-for-loop in the
 * same order as the reference so that the restatement can be read against it.
 *
 * Element types use the reference's node counts: Tria3 (ndim 2, npElem 3),
 * Quad4 (2, 4), Tetra4 (3, 4), Hexa8 (3, 8).
 */
#ifndef SFEM_ORACLE_H
#define SFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_TRIA3 = 0, ORC_QUAD4 = 1, ORC_TETRA4 = 2, ORC_HEXA8 = 3 };

/* element type from (ndim, npElem); -1 if unsupported (StabFEM.cpp:351-366) */
int orc_elem_type(int ndim, int npElem);
/* nGP of the element type (ctor of each LagrangeElem*.cpp) */
int orc_elem_ngp(int etype);

/* SetTimeParametersFluid (UtilitiesGeneral.cpp:5-209).  td has 16 entries
 * (the reference sizes it 100; entries >= 16 are never written). */
int orc_time_params(int tis, double rhoInf, double dt, double* td);

/* Gauss rule of the element type: points gp[ngp*3] (unused coords 0), weights gw[ngp]
 * (UtilitiesQuadrature.cpp:10-28, 224-301, 437-455; Tria3 uses the quad n=1 rule,
 * LagrangeElem2DNavierStokesTria3Node.cpp:56). Returns ngp. */
int orc_gauss(int etype, double* gp, double* gw);

/* Degree-1 shape functions and parametric derivatives at one point
 * (BasisFunctionsLagrange.cpp:517-546, 687-732, 873-911, 1019-1080). */
void orc_shape(int etype, const double* xi, double* N, double* dN1, double* dN2, double* dN3);

/* prepareElemData for all elements (Quad4 :36-113, Tria3 :36-113, Tetra4 :31-112,
 * Hexa8 :31-106).  Outputs per element e, GP g, local node i at
 * [(e*ngp + g)*nlbf + i]; dvol at [e*ngp + g]; charlen at [e].
 * Returns 0, or -(e+1) for the first element with a negative Jacobian. */
int64_t orc_prepare(int etype, int64_t nElem, const double* coords, const int32_t* conn,
                    double* N, double* dNdx, double* dNdy, double* dNdz,
                    double* dvol, double* charlen);

/* calcStiffnessAndResidual of one element (Hexa8 :764-1074, Tetra4 :769-1078,
 * Quad4 :683-970, Tria3 :683-968).  Ke is row-major nsize x nsize, Fe nsize.
 * elemData has 11 entries (StabFEM.cpp:284-315); td as from orc_time_params. */
void orc_element_KF(int etype, int64_t e, const int32_t* conn,
                    const double* N, const double* dNdx, const double* dNdy, const double* dNdz,
                    const double* dvol, const double* charlen,
                    const double* elemData, const double* td,
                    const double* solnPrev, const double* solnCur, const double* solnDotCur,
                    double* Ke, double* Fe);

/* DOF numbering (StabFEM2.cpp:65-81) and assyForSoln (:172-183).
 * nodeDof[nNode*ndof] = free id or -1.  Returns nfree. */
int64_t orc_dof_numbering(int64_t nNode, int ndof, int64_t nDBC, const int32_t* dbcNode,
                          const int32_t* dbcDof, int32_t* nodeDof, int32_t* assyForSoln);

/* CSR pattern by brute force, as the per-row std::set (StabFEM2.cpp:192-227).
 * Pass colidx = NULL to get nnz only (rowptr filled).  Returns nnz. */
int64_t orc_pattern(int64_t nElem, int npElem, int ndof, const int32_t* conn,
                    const int32_t* nodeDof, int64_t nfree, int64_t* rowptr, int32_t* colidx);

/* SolutionData::timeUpdate (SolutionData.cpp:64-78) */
void orc_time_update(int64_t n, double* soln, double* solnPrev, double* solnPrev2,
                     double* solnPrev3, double* solnPrev4, double* solnDot, double* solnDotPrev);
/* SolutionData::updateIterStep (SolutionData.cpp:83-91) */
void orc_update_iter_step(int64_t n, const double* td, const double* soln, const double* solnPrev,
                          const double* solnPrev2, const double* solnPrev3, const double* solnPrev4,
                          double* solnDot, const double* solnDotPrev, double* solnCur, double* solnDotCur);

/* Element loop of one Newton iteration (StabFEM2.cpp:669-708): Ke/Fe per
 * element, BC lift at iter 0 (:678-703), then MatSetValues/VecSetValues ADD
 * with a per-row binary search (SolverPetsc.cpp:304-314).  vals/rhs are
 * zeroed first (zeroMtx, SolverPetsc.cpp:127-145).  Elements [e0, e1).
 * vals may be NULL to assemble the residual only. */
void orc_assemble(int etype, int64_t e0, int64_t e1, int iter, const int32_t* conn,
                  const double* N, const double* dNdx, const double* dNdy, const double* dNdz,
                  const double* dvol, const double* charlen,
                  const double* elemData, const double* td,
                  const double* solnPrev, const double* solnCur, const double* solnDotCur,
                  const double* solnApplied, const int32_t* nodeDof,
                  const int64_t* rowptr, const int32_t* colidx, int64_t nfree,
                  double* vals, double* rhs, int zero_first);

/* y = A x (CSR) */
void orc_spmv(int64_t n, const int64_t* rowptr, const int32_t* colidx, const double* vals,
              const double* x, double* y);

/* Preconditioners.  type 0 none, 1 point Jacobi, 2 nodal block Jacobi (blocks
 * given by blk_ptr: rows [blk_ptr[k], blk_ptr[k+1]) form block k, size <= 4),
 * 3 ILU(0) on the CSR pattern (textbook IKJ variant, Saad, Iterative Methods,
 * 2nd ed., Alg. 10.4). */
typedef struct orc_pc orc_pc;
orc_pc* orc_pc_create(int type, int64_t n, const int64_t* rowptr, const int32_t* colidx,
                      const double* vals, int64_t nblk, const int64_t* blk_ptr, int* status);
void orc_pc_apply(const orc_pc* pc, const double* r, double* z);
void orc_pc_destroy(orc_pc* pc);
/* Copy out the ILU(0) factor values (L strictly below, U on and above the
 * diagonal, on the CSR pattern).  Returns the count, or -1 if not ILU. */
int64_t orc_pc_lu(const orc_pc* pc, double* out);

/* Right-preconditioned BiCGStab (van der Vorst 1992; Saad Alg. 9.21 with right
 * preconditioning), x0 = 0, stop on ||b - A x||_2 <= rtol ||b||_2.
 * Returns 0 converged, 1 max iterations, 2 breakdown. */
int orc_bicgstab(int64_t n, const int64_t* rowptr, const int32_t* colidx, const double* vals,
                 const orc_pc* pc, const double* b, double* x, double rtol, int maxit,
                 int* iters, double* relres);

/* Right-preconditioned restarted GMRES(m) with modified Gram-Schmidt and Givens
 * rotations (Saad Alg. 6.9/9.5), x0 = 0, same stopping rule. */
int orc_gmres(int64_t n, const int64_t* rowptr, const int32_t* colidx, const double* vals,
              const orc_pc* pc, const double* b, double* x, int restart, double rtol, int maxit,
              int* iters, double* relres);

/* Host threads of the parallel paths (default 1: everything sequential, in the
 * reference's element and row order).  With n > 1: elements run colour class by colour
 * class, the ILU(0) factorisation and triangular solves level by level over the
 * dependency levels of their rows, the pattern node by node, products and vector
 * updates split over rows, dot products as fixed chunks summed in index order. */
void orc_set_threads(int n);
int orc_get_threads(void);

/* Count of floating-point operations executed by orc_element_KF for one
 * element with K and F (want_K=1) or F only (want_K=0), counted on the
 * reference's arithmetic as written. */
int64_t orc_flops_per_element(int etype, int want_K);

#ifdef __cplusplus
}
#endif
#endif
