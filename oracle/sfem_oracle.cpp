// sfem_oracle.cpp -- CPU restatement of the cfdsfemmpi hot path (test infrastructure).
//
// TEST INFRASTRUCTURE ONLY: the parity checker and the reported CPU baseline.
// The product path (cfdsfemmpi_amd/csrc) never links or calls this file.
//
// Each function cites the reference file:line it restates (paths relative to the
// reference repository root, src/base/ unless stated).  The element kernel is
// templated on its scalar type so that the same code, instantiated with a
// counting scalar, yields the algorithmic flop count per element.
//
// Parity pinning (what fixes each function independently of this text):
//   orc_time_params   closed forms of the generalised-alpha / BDF coefficients
//   orc_update_iter_step  the generalised-alpha update in its trapezoidal form
//                     (rho_inf 0, 0.5, 1); BDF1-4 exact on polynomials of their order
//   orc_time_update   the history shift, entry by entry
//   orc_gauss         literal tables (incl. the 0.577350269189626 literal)
//   orc_shape         partition of unity, sum dN = 0, exact linear reproduction,
//                     and the Hexa8 derivative table checked entry by entry
//   orc_prepare       closed-form Jacobians/volumes of affine elements
//   orc_element_KF    affine invariant Fe(u+d)-Fe(u) = -Ke d; zero state; an
//                     independent Gauss-quadrature evaluation of the weak form;
//                     end-to-end reference solutions (Quad4 only)
//   orc_pattern       closed-form nnz counts of structured meshes
//   orc_assemble      reference solutions of the two 2D lid-driven cavity tests
//   Krylov / PC       converged answer vs a direct solve (iteration counts and
//                     residual histories of PETSc's KSP are parity unpinned)
#include "sfem_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// Host threads of the parallel paths (orc_set_threads).  1 (the default) runs every
// function sequentially in the reference's order: elements in index order, rows of the
// triangular solves one after another.  With more threads the element loop runs colour
// class by colour class (no two elements of a class share a node, so the ADD assembly
// has no conflicting writes), the ILU(0) factorisation and triangular solves run level
// by level over the dependency levels of their rows, and products and vector updates
// are split over rows; dot products then sum fixed-size chunks in index order, so the
// result does not depend on the thread count.
int g_threads = 1;
constexpr int64_t kDotChunk = 8192;

// PI as written in headersBasic.h:28 (rounds to the nearest double of pi)
const double kPI = 3.141592653589793238462643383279502884197169399375105;

struct ElemInfo {
  int ndim, npElem, nlbf, ndof, nsize, nGP;
};

ElemInfo elem_info(int etype) {
  switch (etype) {
    case ORC_TRIA3: return {2, 3, 3, 3, 9, 1};    // Tria3Node.cpp:20-28
    case ORC_QUAD4: return {2, 4, 4, 3, 12, 4};   // Quad4Node.cpp:19-27
    case ORC_TETRA4: return {3, 4, 4, 4, 16, 1};  // Tetra4Node.cpp:13-21
    default: return {3, 8, 8, 4, 32, 8};          // Hexa8Node.cpp:12-20
  }
}

// ---------------------------------------------------------------------------
// Counting scalar: every +, -, *, /, sqrt, pow on it increments a counter.
// ---------------------------------------------------------------------------
struct CountingScalar {
  double v;
  static int64_t count;
  CountingScalar(double x = 0.0) : v(x) {}
};
int64_t CountingScalar::count = 0;
inline CountingScalar operator+(CountingScalar a, CountingScalar b) { ++CountingScalar::count; return a.v + b.v; }
inline CountingScalar operator-(CountingScalar a, CountingScalar b) { ++CountingScalar::count; return a.v - b.v; }
inline CountingScalar operator*(CountingScalar a, CountingScalar b) { ++CountingScalar::count; return a.v * b.v; }
inline CountingScalar operator/(CountingScalar a, CountingScalar b) { ++CountingScalar::count; return a.v / b.v; }
inline CountingScalar operator-(CountingScalar a) { return -a.v; }
inline CountingScalar& operator+=(CountingScalar& a, CountingScalar b) { a = a + b; return a; }
inline CountingScalar& operator-=(CountingScalar& a, CountingScalar b) { a = a - b; return a; }
inline CountingScalar& operator*=(CountingScalar& a, CountingScalar b) { a = a * b; return a; }
inline CountingScalar sqrt_(CountingScalar a) { ++CountingScalar::count; return std::sqrt(a.v); }
inline double sqrt_(double a) { return std::sqrt(a); }

// ---------------------------------------------------------------------------
// evaluateStabParams_algo3 (stabilisationRoutines.h:128-143) with the G of the
// element kernels: G = (4/charlen^2) I on the first ndim diagonal entries of a
// 3x3 matrix (Quad4Node.cpp:722-726, Hexa8Node.cpp:786-791).
// ---------------------------------------------------------------------------
template <typename T>
void stab_algo3(int ndim, const T* u /*3*/, T fact, T rho, T mu, T CI, T* tau) {
  // G*u, then u.dot(G*u)
  T Gu[3];
  for (int a = 0; a < 3; ++a) Gu[a] = (a < ndim) ? fact * u[a] : T(0.0);
  T uGu = u[0] * Gu[0] + u[1] * Gu[1] + u[2] * Gu[2];
  // G.cwiseAbs2().sum() and G.trace() over the ndim nonzero diagonal entries
  T G2 = fact * fact;
  T trG = fact;
  for (int a = 1; a < ndim; ++a) { G2 = G2 + fact * fact; trG = trG + fact; }
  tau[0] = uGu + CI * (mu * mu / rho / rho) * G2;
  tau[0] = T(1.0) / sqrt_(tau[0]);
  tau[2] = T(1.0) / (tau[0] * trG);
  tau[0] = tau[0] / rho;
  tau[1] = tau[0];
}

// ---------------------------------------------------------------------------
// calcStiffnessAndResidual, live "semi-implicit type B" kernels.
// ---------------------------------------------------------------------------
template <typename T>
void element_KF(int etype, int64_t e, const int32_t* conn, const double* Nall, const double* dNdxall,
                const double* dNdyall, const double* dNdzall, const double* dvolall,
                const double* charlenall, const double* elemData, const double* td,
                const double* solnPrev, const double* solnCur, const double* solnDotCur, T* K,
                T* F, bool wantK) {
  const ElemInfo ei = elem_info(etype);
  const int ndim = ei.ndim, nlbf = ei.nlbf, ndof = ei.ndof, nsize = ei.nsize, nGP = ei.nGP;
  const int32_t* nodeNums = conn + e * ei.npElem;
  // Hexa8 negates the continuity and PSPG rows relative to the other three
  // (Hexa8Node.cpp:984-986,1027-1030,1052,1060 vs Tetra4Node.cpp:990-992,1033-1036,1058,1066)
  const double sgnP = (etype == ORC_HEXA8) ? -1.0 : 1.0;
  const bool supg = (ndim == 2);  // SUPG live in 2D only (Quad4Node.cpp:906-918,945-950)
  const T CI = (ndim == 2) ? T(4.0) : T(10.0);

  const T charlen = charlenall[e];
  T fact = T(4.0) / charlen / charlen;

  const T rho = elemData[0];
  const T mu = elemData[1];
  const T af = td[2];
  const T acceFact = td[8];
  const T muTaf = mu * af;

  for (int i = 0; i < nsize * nsize; ++i) K[i] = T(0.0);
  for (int i = 0; i < nsize; ++i) F[i] = T(0.0);

  for (int gp = 0; gp < nGP; ++gp) {
    const double* Nv = Nall + (e * nGP + gp) * nlbf;
    const double* dN[3] = {dNdxall + (e * nGP + gp) * nlbf, dNdyall + (e * nGP + gp) * nlbf,
                           ndim == 3 ? dNdzall + (e * nGP + gp) * nlbf : nullptr};
    T vel[3] = {0, 0, 0}, velPrev[3] = {0, 0, 0}, velDot[3] = {0, 0, 0};
    T grad[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gradPrev[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    T pres = 0.0, dp[3] = {0, 0, 0};

    // interpolation at the GP (Hexa8Node.cpp:828-889, Quad4Node.cpp:748-795)
    for (int ii = 0; ii < nlbf; ++ii) {
      const int64_t TI = (int64_t)nodeNums[ii] * ndof;
      for (int a = 0; a < ndim; ++a) {
        T b = solnPrev[TI + a];
        velPrev[a] += b * T(Nv[ii]);
        for (int c = 0; c < ndim; ++c) gradPrev[a][c] += b * T(dN[c][ii]);
      }
      for (int a = 0; a < ndim; ++a) {
        T b = solnCur[TI + a];
        vel[a] += b * T(Nv[ii]);
        for (int c = 0; c < ndim; ++c) grad[a][c] += b * T(dN[c][ii]);
      }
      T b4 = solnCur[TI + ndim];
      pres += b4 * T(Nv[ii]);
      for (int c = 0; c < ndim; ++c) dp[c] += b4 * T(dN[c][ii]);
      for (int a = 0; a < ndim; ++a) velDot[a] += T(solnDotCur[TI + a]) * T(Nv[ii]);
    }

    // pseudo-stress mu*grad - p I (Hexa8Node.cpp:892-895)
    T stress[3][3];
    for (int a = 0; a < ndim; ++a)
      for (int c = 0; c < ndim; ++c) stress[a][c] = mu * grad[a][c];
    for (int a = 0; a < ndim; ++a) stress[a][a] -= pres;

    // gradTvel = gradPrev*vel + grad*velPrev - gradPrev*velPrev (:899)
    T gradTvel[3];
    for (int a = 0; a < ndim; ++a) {
      T s1 = gradPrev[a][0] * vel[0], s2 = grad[a][0] * velPrev[0], s3 = gradPrev[a][0] * velPrev[0];
      for (int c = 1; c < ndim; ++c) {
        s1 += gradPrev[a][c] * vel[c];
        s2 += grad[a][c] * velPrev[c];
        s3 += gradPrev[a][c] * velPrev[c];
      }
      gradTvel[a] = s1 + s2 - s3;
    }
    // res2 = rho*(velDot + gradTvel - force), force = 0 (:901-903); rStab (:905-907)
    T res2[3], rStab[3];
    for (int a = 0; a < ndim; ++a) res2[a] = rho * (velDot[a] + gradTvel[a] - T(0.0));
    for (int a = 0; a < ndim; ++a) rStab[a] = res2[a] - mu * T(0.0) + dp[a];

    // tau (:911-914); 2D passes (velPrev0, velPrev1, 0) (Quad4Node.cpp:817-825)
    T u3[3] = {velPrev[0], velPrev[1], ndim == 3 ? velPrev[2] : T(0.0)};
    T tau[3];
    stab_algo3<T>(ndim, u3, fact, rho, mu, CI, tau);
    tau[0] *= T(elemData[8]);
    tau[1] *= T(elemData[9]);
    tau[2] *= T(elemData[10]);

    const T dvol = dvolall[e * nGP + gp];
    T trace = grad[0][0];
    for (int a = 1; a < ndim; ++a) trace = trace + grad[a][a];

    for (int ii = 0; ii < nlbf; ++ii) {
      const int TI = ndof * ii;
      T bb[3] = {0, 0, 0};
      for (int c = 0; c < ndim; ++c) bb[c] = T(dN[c][ii]) * dvol;  // b1, b2, b3
      T b4 = T(Nv[ii]) * dvol;
      T bm[3];
      for (int c = 0; c < ndim; ++c) bm[c] = muTaf * bb[c];  // b5, b6, b7
      T b8 = af * b4;
      T Da = velPrev[0] * bb[0];
      for (int c = 1; c < ndim; ++c) Da = Da + velPrev[c] * bb[c];
      Da = rho * Da * tau[0];

      if (wantK) {
        for (int jj = 0; jj < nlbf; ++jj) {
          const int TJ = ndof * jj;
          T fact2 = rho * acceFact * T(Nv[jj]);
          // time acceleration + diffusion (:938-946)
          T f = b4 * fact2;
          T dif = bm[0] * T(dN[0][jj]);
          for (int c = 1; c < ndim; ++c) dif = dif + bm[c] * T(dN[c][jj]);
          f += dif;
          for (int a = 0; a < ndim; ++a) K[(TI + a) * nsize + TJ + a] += f;

          // convection, semi-implicit type B (:950-971)
          T gradN[3][3];
          T rN = rho * T(Nv[jj]);
          for (int a = 0; a < ndim; ++a)
            for (int c = 0; c < ndim; ++c) gradN[a][c] = gradPrev[a][c] * rN;
          T Db = velPrev[0] * T(dN[0][jj]);
          for (int c = 1; c < ndim; ++c) Db = Db + velPrev[c] * T(dN[c][jj]);
          Db = rho * Db;
          for (int a = 0; a < ndim; ++a) gradN[a][a] += Db;
          for (int a = 0; a < ndim; ++a)
            for (int c = 0; c < ndim; ++c) K[(TI + a) * nsize + TJ + c] += (b8 * gradN[a][c]);

          // pressure term (:974-976)
          for (int a = 0; a < ndim; ++a) K[(TI + a) * nsize + TJ + ndim] -= (bb[a] * af * T(Nv[jj]));

          // continuity equation (:979-981 Hexa8 "-"; Tetra4/Quad4/Tria3 "+")
          for (int c = 0; c < ndim; ++c) {
            if (sgnP > 0) K[(TI + ndim) * nsize + TJ + c] += (b8 * T(dN[c][jj]));
            else          K[(TI + ndim) * nsize + TJ + c] -= (b8 * T(dN[c][jj]));
          }

          // Dj (:988-1003)
          for (int a = 0; a < ndim; ++a)
            for (int c = 0; c < ndim; ++c) gradN[a][c] *= af;
          T Dj[3][4];
          for (int a = 0; a < ndim; ++a) {
            for (int c = 0; c < ndim; ++c) Dj[a][c] = gradN[a][c];
            Dj[a][a] = gradN[a][a] + fact2;
            Dj[a][ndim] = af * T(dN[a][jj]);
          }

          // SUPG, 2D only (Quad4Node.cpp:906-914)
          if (supg) {
            for (int a = 0; a < ndim; ++a)
              for (int c = 0; c <= ndim; ++c) K[(TI + a) * nsize + TJ + c] += Da * Dj[a][c];
          }
          // PSPG (Hexa8Node.cpp:1027-1030 "-"; Quad4Node.cpp:917-919 "+")
          for (int c = 0; c <= ndim; ++c) {
            T s = bb[0] * Dj[0][c];
            for (int a = 1; a < ndim; ++a) s = s + bb[a] * Dj[a][c];
            if (sgnP > 0) K[(TI + ndim) * nsize + TJ + c] += s * tau[1];
            else          K[(TI + ndim) * nsize + TJ + c] -= s * tau[1];
          }
        }
      }

      // residual (Hexa8Node.cpp:1049-1060, Quad4Node.cpp:936-950)
      for (int a = 0; a < ndim; ++a) {
        T s = b4 * res2[a];
        for (int c = 0; c < ndim; ++c) s = s + bb[c] * stress[a][c];
        F[TI + a] -= s;
      }
      if (sgnP > 0) F[TI + ndim] -= (b4 * trace);
      else          F[TI + ndim] += (b4 * trace);
      if (supg) {
        for (int a = 0; a < ndim; ++a) F[TI + a] -= Da * rStab[a];
      }
      T s = bb[0] * rStab[0];
      for (int a = 1; a < ndim; ++a) s = s + bb[a] * rStab[a];
      if (sgnP > 0) F[TI + ndim] -= (tau[1] * s);
      else          F[TI + ndim] += (tau[1] * s);
    }
  }
}

// 3x3 determinant and inverse by cofactors (the reference calls Eigen's
// determinant()/inverse() on a dynamic 3x3, BasisFunctionsLagrange.cpp:172-173;
// results agree to rounding).
void det_inv3(const double B[3][3], double& det, double Binv[3][3]) {
  double c00 = B[1][1] * B[2][2] - B[1][2] * B[2][1];
  double c01 = B[1][2] * B[2][0] - B[1][0] * B[2][2];
  double c02 = B[1][0] * B[2][1] - B[1][1] * B[2][0];
  det = B[0][0] * c00 + B[0][1] * c01 + B[0][2] * c02;
  double id = 1.0 / det;
  Binv[0][0] = c00 * id;
  Binv[1][0] = c01 * id;
  Binv[2][0] = c02 * id;
  Binv[0][1] = (B[0][2] * B[2][1] - B[0][1] * B[2][2]) * id;
  Binv[1][1] = (B[0][0] * B[2][2] - B[0][2] * B[2][0]) * id;
  Binv[2][1] = (B[0][1] * B[2][0] - B[0][0] * B[2][1]) * id;
  Binv[0][2] = (B[0][1] * B[1][2] - B[0][2] * B[1][1]) * id;
  Binv[1][2] = (B[0][2] * B[1][0] - B[0][0] * B[1][2]) * id;
  Binv[2][2] = (B[0][0] * B[1][1] - B[0][1] * B[1][0]) * id;
}

}  // namespace

extern "C" {

int orc_elem_type(int ndim, int npElem) {
  if (ndim == 2 && npElem == 3) return ORC_TRIA3;
  if (ndim == 2 && npElem == 4) return ORC_QUAD4;
  if (ndim == 3 && npElem == 4) return ORC_TETRA4;
  if (ndim == 3 && npElem == 8) return ORC_HEXA8;
  return -1;
}

int orc_elem_ngp(int etype) { return elem_info(etype).nGP; }

int orc_time_params(int tis, double rho, double dt, double* td) {
  for (int i = 0; i < 16; ++i) td[i] = 0.0;
  td[0] = dt;
  double alpf, alpm, gamm;
  switch (tis) {
    case 0:  // quasi static (:15-26)
      td[1] = 1.0; td[2] = 1.0; td[3] = 1.0; td[4] = 1.0;
      break;
    case 1:  // generalised midpoint (:28-50)
      alpf = 1.0 / (1.0 + rho); alpm = 1.0; gamm = 1.0;
      td[1] = alpm; td[2] = alpf; td[3] = alpm; td[4] = gamm;
      td[5] = alpf * dt; td[6] = alpm / gamm; td[7] = 1.0 - alpm / gamm; td[8] = 0.0;
      td[9] = 1.0 / dt; td[10] = -td[9]; td[11] = 0.0; td[12] = 0.0; td[15] = 0.0;
      break;
    case 2:  // generalised alpha (:52-74)
    case 4:  // generalised alpha, identical coefficients (:101-124)
      alpf = 1.0 / (1.0 + rho);
      alpm = 0.5 * (3.0 - rho) / (1.0 + rho);
      gamm = 0.5 + alpm - alpf;
      td[1] = alpm; td[2] = alpf; td[3] = alpm; td[4] = gamm;
      td[5] = alpf * dt; td[6] = alpm / gamm; td[7] = 1.0 - alpm / gamm; td[8] = alpm / gamm / dt;
      td[9] = 1.0 / gamm / dt; td[10] = -td[9]; td[11] = 0.0; td[12] = 0.0; td[13] = 0.0;
      td[15] = 1.0 - 1.0 / gamm;
      break;
    case 3:  // backward Euler / BDF1 (:76-99)
      td[1] = 1.0; td[2] = 1.0; td[3] = 1.0; td[4] = 1.0;
      td[5] = dt; td[6] = 1.0; td[7] = 0.0; td[8] = 1.0 / dt;
      td[9] = 1.0 / dt; td[10] = -1.0 / dt;
      break;
    case 5:  // BDF2 (:126-149)
      td[1] = 1.0; td[2] = 1.0; td[3] = 1.0; td[4] = 1.0;
      td[5] = dt; td[6] = 1.0; td[7] = 0.0; td[8] = 1.0 / dt;
      td[9] = 1.5 / dt; td[10] = -2.0 / dt; td[11] = 0.5 / dt;
      break;
    case 6:  // BDF3 (:151-174)
      td[1] = 1.0; td[2] = 1.0; td[3] = 1.0; td[4] = 1.0;
      td[5] = dt; td[6] = 1.0; td[7] = 0.0; td[8] = 1.0 / dt;
      td[9] = (11.0 / 6.0) / dt; td[10] = -3.0 / dt; td[11] = 1.5 / dt; td[12] = -(1.0 / 3.0) / dt;
      break;
    case 7:  // BDF4 (:176-199)
      td[1] = 1.0; td[2] = 1.0; td[3] = 1.0; td[4] = 1.0;
      td[5] = dt; td[6] = 1.0; td[7] = 0.0; td[8] = 1.0 / dt;
      td[9] = (25.0 / 12.0) / dt; td[10] = -4.0 / dt; td[11] = 3.0 / dt;
      td[12] = -(4.0 / 3.0) / dt; td[13] = (1.0 / 4.0) / dt;
      break;
    default:
      return -1;  // the reference prints an error and leaves td zero (:201-205)
  }
  return 0;
}

int orc_gauss(int etype, double* gp, double* gw) {
  const double g2[2] = {-0.577350269189626, 0.577350269189626};  // UtilitiesQuadrature.cpp:25-26
  switch (etype) {
    case ORC_TRIA3:  // getGaussPointsQuad(1): point (0,0), weight 2*2 (:17-21, :224-257)
      gp[0] = 0.0; gp[1] = 0.0; gp[2] = 0.0; gw[0] = 4.0;
      return 1;
    case ORC_QUAD4: {  // x fastest (:245-253)
      int ind = 0;
      for (int jj = 0; jj < 2; ++jj)
        for (int ii = 0; ii < 2; ++ii) {
          gp[3 * ind] = g2[ii]; gp[3 * ind + 1] = g2[jj]; gp[3 * ind + 2] = 0.0;
          gw[ind] = 1.0 * 1.0;
          ++ind;
        }
      return 4;
    }
    case ORC_TETRA4:  // (:448-455)
      gp[0] = 0.25; gp[1] = 0.25; gp[2] = 0.25; gw[0] = 1.0 / 6.0;
      return 1;
    default: {  // Hexa, x fastest then y then z (:285-297)
      int ind = 0;
      for (int kk = 0; kk < 2; ++kk)
        for (int jj = 0; jj < 2; ++jj)
          for (int ii = 0; ii < 2; ++ii) {
            gp[3 * ind] = g2[ii]; gp[3 * ind + 1] = g2[jj]; gp[3 * ind + 2] = g2[kk];
            gw[ind] = 1.0 * 1.0 * 1.0;
            ++ind;
          }
      return 8;
    }
  }
}

void orc_shape(int etype, const double* xi, double* N, double* dN1, double* dN2, double* dN3) {
  switch (etype) {
    case ORC_TRIA3: {  // BasisFunctionsLagrange.cpp:517-546
      double xi3 = 1.0 - xi[0] - xi[1];
      N[0] = xi3; N[1] = xi[0]; N[2] = xi[1];
      dN1[0] = -1.0; dN1[1] = 1.0; dN1[2] = 0.0;
      dN2[0] = -1.0; dN2[1] = 0.0; dN2[2] = 1.0;
      break;
    }
    case ORC_QUAD4: {  // :687-732
      double b1 = 0.5 * (1.0 - xi[1]), b2 = 0.5 * (1.0 + xi[1]);
      double db1 = -0.5, db2 = 0.5;
      double a1 = 0.5 * (1.0 - xi[0]), a2 = 0.5 * (1.0 + xi[0]);
      double da1 = -0.5, da2 = 0.5;
      N[0] = b1 * a1; N[1] = b1 * a2; N[2] = b2 * a2; N[3] = b2 * a1;
      dN1[0] = b1 * da1; dN1[1] = b1 * da2; dN1[2] = b2 * da2; dN1[3] = b2 * da1;
      dN2[0] = db1 * a1; dN2[1] = db1 * a2; dN2[2] = db2 * a2; dN2[3] = db2 * a1;
      break;
    }
    case ORC_TETRA4: {  // :873-911
      double xi4 = 1.0 - xi[0] - xi[1] - xi[2];
      N[0] = xi4; N[1] = xi[0]; N[2] = xi[1]; N[3] = xi[2];
      dN1[0] = -1.0; dN1[1] = 1.0; dN1[2] = 0.0; dN1[3] = 0.0;
      dN2[0] = -1.0; dN2[1] = 0.0; dN2[2] = 1.0; dN2[3] = 0.0;
      dN3[0] = -1.0; dN3[1] = 0.0; dN3[2] = 0.0; dN3[3] = 1.0;
      break;
    }
    default: {  // Hexa8, :1019-1080.  N[] is the textbook trilinear set; the
      // xi2- and xi3-derivatives of nodes 2<->3 and 6<->7 are as the reference
      // writes them (:1056-1074), i.e. swapped relative to N[].
      double v11 = 1.0 - xi[0], v12 = 1.0 + xi[0];
      double v21 = 1.0 - xi[1], v22 = 1.0 + xi[1];
      double v31 = 1.0 - xi[2], v32 = 1.0 + xi[2];
      N[0] = 0.125 * v11 * v21 * v31; N[1] = 0.125 * v12 * v21 * v31;
      N[2] = 0.125 * v12 * v22 * v31; N[3] = 0.125 * v11 * v22 * v31;
      N[4] = 0.125 * v11 * v21 * v32; N[5] = 0.125 * v12 * v21 * v32;
      N[6] = 0.125 * v12 * v22 * v32; N[7] = 0.125 * v11 * v22 * v32;
      dN1[0] = -0.125 * v21 * v31; dN1[1] = 0.125 * v21 * v31;
      dN1[2] = 0.125 * v22 * v31;  dN1[3] = -0.125 * v22 * v31;
      dN1[4] = -0.125 * v21 * v32; dN1[5] = 0.125 * v21 * v32;
      dN1[6] = 0.125 * v22 * v32;  dN1[7] = -0.125 * v22 * v32;
      dN2[0] = -0.125 * v11 * v31; dN2[1] = -0.125 * v12 * v31;
      dN2[2] = 0.125 * v11 * v31;  dN2[3] = 0.125 * v12 * v31;
      dN2[4] = -0.125 * v11 * v32; dN2[5] = -0.125 * v12 * v32;
      dN2[6] = 0.125 * v11 * v32;  dN2[7] = 0.125 * v12 * v32;
      dN3[0] = -0.125 * v11 * v21; dN3[1] = -0.125 * v12 * v21;
      dN3[2] = -0.125 * v11 * v22; dN3[3] = -0.125 * v12 * v22;
      dN3[4] = 0.125 * v11 * v21;  dN3[5] = 0.125 * v12 * v21;
      dN3[6] = 0.125 * v11 * v22;  dN3[7] = 0.125 * v12 * v22;
      break;
    }
  }
}

int64_t orc_prepare(int etype, int64_t nElem, const double* coords, const int32_t* conn,
                    double* Nout, double* dNdx, double* dNdy, double* dNdz, double* dvolout,
                    double* charlen) {
  const ElemInfo ei = elem_info(etype);
  const int ndim = ei.ndim, nlbf = ei.nlbf, nGP = ei.nGP;
  double gps[3 * 8], gws[8];
  orc_gauss(etype, gps, gws);
  int64_t bad = nElem;  // first element with a negative Jacobian
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static) reduction(min : bad)
  for (int64_t e = 0; e < nElem; ++e) {
    double xN[8], yN[8], zN[8];
    for (int ii = 0; ii < ei.npElem; ++ii) {
      int64_t n = conn[e * ei.npElem + ii];
      xN[ii] = coords[n * ndim];
      yN[ii] = coords[n * ndim + 1];
      zN[ii] = (ndim == 3) ? coords[n * ndim + 2] : 0.0;
    }
    double elemVol = 0.0;
    for (int gp = 0; gp < nGP; ++gp) {
      double N[8], d1[8], d2[8], d3[8], Jac;
      orc_shape(etype, gps + 3 * gp, N, d1, d2, d3);
      const int64_t o = (e * nGP + gp) * nlbf;
      if (ndim == 2) {  // computeBasisFunctions2D (BasisFunctionsLagrange.cpp:64-95)
        double B00 = 0, B10 = 0, B01 = 0, B11 = 0;
        for (int ii = 0; ii < nlbf; ++ii) {
          B00 += xN[ii] * d1[ii]; B10 += xN[ii] * d2[ii];
          B01 += yN[ii] * d1[ii]; B11 += yN[ii] * d2[ii];
        }
        Jac = B00 * B11 - B01 * B10;
        double detinv = 1.0 / Jac;
        double Bi00 = B11 * detinv, Bi01 = -B01 * detinv, Bi10 = -B10 * detinv, Bi11 = B00 * detinv;
        for (int ii = 0; ii < nlbf; ++ii) {
          Nout[o + ii] = N[ii];
          dNdx[o + ii] = d1[ii] * Bi00 + d2[ii] * Bi01;
          dNdy[o + ii] = d1[ii] * Bi10 + d2[ii] * Bi11;
        }
      } else {  // computeBasisFunctions3D (:150-182): B(r,c) = sum_i x_c(i) dN_i/dxi_r
        double B[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Bi[3][3];
        for (int ii = 0; ii < nlbf; ++ii) {
          const double xc[3] = {xN[ii], yN[ii], zN[ii]};
          for (int c = 0; c < 3; ++c) {
            B[0][c] += xc[c] * d1[ii];
            B[1][c] += xc[c] * d2[ii];
            B[2][c] += xc[c] * d3[ii];
          }
        }
        det_inv3(B, Jac, Bi);
        for (int ii = 0; ii < nlbf; ++ii) {
          Nout[o + ii] = N[ii];
          dNdx[o + ii] = d1[ii] * Bi[0][0] + d2[ii] * Bi[0][1] + d3[ii] * Bi[0][2];
          dNdy[o + ii] = d1[ii] * Bi[1][0] + d2[ii] * Bi[1][1] + d3[ii] * Bi[1][2];
          dNdz[o + ii] = d1[ii] * Bi[2][0] + d2[ii] * Bi[2][1] + d3[ii] * Bi[2][2];
        }
      }
      if (Jac < 0.0) bad = std::min(bad, e);  // the reference exits (Hexa8Node.cpp:79-84)
      double dvol = gws[gp] * Jac;
      dvolout[e * nGP + gp] = dvol;
      elemVol += dvol;
    }
    charlen[e] = (ndim == 2) ? std::sqrt(4.0 * std::fabs(elemVol) / kPI)
                             : std::pow(6.0 * std::fabs(elemVol) / kPI, 1.0 / 3.0);
  }
  return bad < nElem ? -(bad + 1) : 0;
}

void orc_element_KF(int etype, int64_t e, const int32_t* conn, const double* N, const double* dNdx,
                    const double* dNdy, const double* dNdz, const double* dvol,
                    const double* charlen, const double* elemData, const double* td,
                    const double* solnPrev, const double* solnCur, const double* solnDotCur,
                    double* Ke, double* Fe) {
  element_KF<double>(etype, e, conn, N, dNdx, dNdy, dNdz, dvol, charlen, elemData, td, solnPrev,
                     solnCur, solnDotCur, Ke, Fe, true);
}

int64_t orc_flops_per_element(int etype, int want_K) {
  // one element of unit size with nonzero state
  const ElemInfo ei = elem_info(etype);
  std::vector<double> coords, Nv, dx, dy, dz, dv, cl(1);
  std::vector<int32_t> conn(ei.npElem);
  const double hex[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const double tet[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < ei.npElem; ++i) {
    conn[i] = i;
    for (int c = 0; c < ei.ndim; ++c)
      coords.push_back(etype == ORC_TETRA4 || etype == ORC_TRIA3 ? tet[i][c] : hex[i][c]);
  }
  const int ng = ei.nGP, nl = ei.nlbf;
  Nv.resize(ng * nl); dx.resize(ng * nl); dy.resize(ng * nl); dz.resize(ng * nl); dv.resize(ng);
  orc_prepare(etype, 1, coords.data(), conn.data(), Nv.data(), dx.data(), dy.data(), dz.data(),
              dv.data(), cl.data());
  std::vector<double> s(ei.npElem * ei.ndof);
  for (size_t i = 0; i < s.size(); ++i) s[i] = 0.1 + 0.01 * i;
  double ed[11] = {1.0, 0.01, 0, 0, 0, 0, 0, 0, 1.0, 1.0, 0.0};
  double td[16];
  orc_time_params(2, 0.0, 0.1, td);
  std::vector<CountingScalar> K(ei.nsize * ei.nsize), F(ei.nsize);
  CountingScalar::count = 0;
  element_KF<CountingScalar>(etype, 0, conn.data(), Nv.data(), dx.data(), dy.data(), dz.data(),
                             dv.data(), cl.data(), ed, td, s.data(), s.data(), s.data(), K.data(),
                             F.data(), want_K != 0);
  return CountingScalar::count;
}

int64_t orc_dof_numbering(int64_t nNode, int ndof, int64_t nDBC, const int32_t* dbcNode,
                          const int32_t* dbcDof, int32_t* nodeDof, int32_t* assyForSoln) {
  std::vector<char> fixed((size_t)nNode * ndof, 0);
  for (int64_t i = 0; i < nDBC; ++i) fixed[(size_t)dbcNode[i] * ndof + dbcDof[i]] = 1;
  int64_t n = 0;
  for (int64_t i = 0; i < nNode; ++i)
    for (int j = 0; j < ndof; ++j) {
      if (!fixed[i * ndof + j]) {
        nodeDof[i * ndof + j] = (int32_t)n;
        if (assyForSoln) assyForSoln[n] = (int32_t)(i * ndof + j);
        ++n;
      } else {
        nodeDof[i * ndof + j] = -1;
      }
    }
  return n;
}

namespace {
// The pattern of orc_pattern built node by node (threads > 1): the columns of every free
// row of node i are the free DOFs of all nodes of the elements containing i, in
// increasing DOF id -- the same set the per-row insertion of the reference collects.
// Two passes over the nodes (row lengths, then columns), each split over threads.
int64_t pattern_by_node(int64_t nElem, int npElem, int ndof, const int32_t* conn, const int32_t* nodeDof,
                        int64_t nfree, int64_t* rowptr, int32_t* colidx) {
  int64_t nNode = 0;
  for (int64_t q = 0; q < nElem * npElem; ++q) nNode = std::max<int64_t>(nNode, (int64_t)conn[q] + 1);
  std::vector<int64_t> ptr(nNode + 1, 0);
  for (int64_t q = 0; q < nElem * npElem; ++q) ptr[conn[q] + 1]++;
  for (int64_t i = 0; i < nNode; ++i) ptr[i + 1] += ptr[i];
  std::vector<int32_t> elems(nElem * npElem);
  {
    std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
    for (int64_t e = 0; e < nElem; ++e)
      for (int a = 0; a < npElem; ++a) elems[pos[conn[e * npElem + a]]++] = (int32_t)e;
  }
  std::vector<int32_t> nfreeOf(nNode, 0);
  for (int64_t i = 0; i < nNode; ++i)
    for (int d = 0; d < ndof; ++d) nfreeOf[i] += nodeDof[i * ndof + d] >= 0;
  auto neighbours = [&](int64_t i, std::vector<int32_t>& nb) {
    nb.clear();
    for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q)
      for (int a = 0; a < npElem; ++a) nb.push_back(conn[(int64_t)elems[q] * npElem + a]);
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
  };
  std::vector<int64_t> rowlen(nNode, 0);
#pragma omp parallel num_threads(g_threads)
  {
    std::vector<int32_t> nb;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t i = 0; i < nNode; ++i) {
      if (!nfreeOf[i]) continue;
      neighbours(i, nb);
      int64_t len = 0;
      for (int32_t j : nb) len += nfreeOf[j];
      rowlen[i] = len;
    }
  }
  std::vector<int64_t> first(nNode, 0);
  rowptr[0] = 0;
  int64_t r = 0;
  for (int64_t i = 0; i < nNode; ++i) {
    first[i] = rowptr[r];
    for (int k = 0; k < nfreeOf[i]; ++k, ++r) rowptr[r + 1] = rowptr[r] + rowlen[i];
  }
  if (colidx) {
#pragma omp parallel num_threads(g_threads)
    {
      std::vector<int32_t> nb;
#pragma omp for schedule(dynamic, 1024)
      for (int64_t i = 0; i < nNode; ++i) {
        if (!nfreeOf[i]) continue;
        neighbours(i, nb);
        int64_t q = first[i];
        for (int32_t j : nb)
          for (int d = 0; d < ndof; ++d)
            if (nodeDof[(int64_t)j * ndof + d] >= 0) colidx[q++] = nodeDof[(int64_t)j * ndof + d];
        for (int k = 1; k < nfreeOf[i]; ++k)
          std::copy(colidx + first[i], colidx + first[i] + rowlen[i], colidx + first[i] + k * rowlen[i]);
      }
    }
  }
  return rowptr[nfree];
}
}  // namespace

int64_t orc_pattern(int64_t nElem, int npElem, int ndof, const int32_t* conn,
                    const int32_t* nodeDof, int64_t nfree, int64_t* rowptr, int32_t* colidx) {
  if (g_threads > 1) return pattern_by_node(nElem, npElem, ndof, conn, nodeDof, nfree, rowptr, colidx);
  // forAssyMat[r].insert(c) for every free (r, c) pair of every element
  const int nsize = npElem * ndof;
  std::vector<std::vector<int32_t>> rows((size_t)nfree);
  std::vector<int32_t> tt(nsize);
  for (int64_t e = 0; e < nElem; ++e) {
    for (int i = 0; i < npElem; ++i)
      for (int d = 0; d < ndof; ++d) tt[i * ndof + d] = nodeDof[(int64_t)conn[e * npElem + i] * ndof + d];
    for (int ii = 0; ii < nsize; ++ii) {
      if (tt[ii] == -1) continue;
      for (int jj = 0; jj < nsize; ++jj)
        if (tt[jj] != -1) rows[tt[ii]].push_back(tt[jj]);
    }
  }
  rowptr[0] = 0;
  for (int64_t r = 0; r < nfree; ++r) {
    auto& v = rows[r];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    rowptr[r + 1] = rowptr[r] + (int64_t)v.size();
    if (colidx) std::copy(v.begin(), v.end(), colidx + rowptr[r]);
  }
  return rowptr[nfree];
}

void orc_time_update(int64_t n, double* soln, double* solnPrev, double* solnPrev2, double* solnPrev3,
                     double* solnPrev4, double* solnDot, double* solnDotPrev) {
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    solnPrev4[i] = solnPrev3[i];
    solnPrev3[i] = solnPrev2[i];
    solnPrev2[i] = solnPrev[i];
    solnPrev[i] = soln[i];
    solnDotPrev[i] = solnDot[i];
  }
}

void orc_update_iter_step(int64_t n, const double* td, const double* soln, const double* solnPrev,
                          const double* solnPrev2, const double* solnPrev3, const double* solnPrev4,
                          double* solnDot, const double* solnDotPrev, double* solnCur,
                          double* solnDotCur) {
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    solnDot[i] = td[9] * soln[i] + td[10] * solnPrev[i] + td[11] * solnPrev2[i] +
                 td[12] * solnPrev3[i] + td[13] * solnPrev4[i] + td[15] * solnDotPrev[i];
    solnCur[i] = td[2] * soln[i] + (1.0 - td[2]) * solnPrev[i];
    solnDotCur[i] = td[1] * solnDot[i] + (1.0 - td[1]) * solnDotPrev[i];
  }
}

void orc_assemble(int etype, int64_t e0, int64_t e1, int iter, const int32_t* conn, const double* N,
                  const double* dNdx, const double* dNdy, const double* dNdz, const double* dvol,
                  const double* charlen, const double* elemData, const double* td,
                  const double* solnPrev, const double* solnCur, const double* solnDotCur,
                  const double* solnApplied, const int32_t* nodeDof, const int64_t* rowptr,
                  const int32_t* colidx, int64_t nfree, double* vals, double* rhs, int zero_first) {
  const ElemInfo ei = elem_info(etype);
  const int nsize = ei.nsize, ndof = ei.ndof;
  if (zero_first) {
    if (vals) std::memset(vals, 0, sizeof(double) * (size_t)rowptr[nfree]);
    std::memset(rhs, 0, sizeof(double) * (size_t)nfree);
  }
  const bool wantK = (vals != nullptr) || iter == 0;
  auto element = [&](int64_t e, double* K, double* F, int32_t* fav, int64_t* gdof) {
    element_KF<double>(etype, e, conn, N, dNdx, dNdy, dNdz, dvol, charlen, elemData, td, solnPrev,
                       solnCur, solnDotCur, K, F, wantK);
    for (int i = 0; i < ei.npElem; ++i)
      for (int d = 0; d < ndof; ++d) {
        int64_t n = conn[e * ei.npElem + i];
        fav[i * ndof + d] = nodeDof[n * ndof + d];   // forAssyVec (StabFEM2.cpp:154-166)
        gdof[i * ndof + d] = n * ndof + d;           // globalDOFnums
      }
    if (iter == 0) {  // BC lift (StabFEM2.cpp:678-703)
      for (int ii = 0; ii < nsize; ++ii) {
        if (fav[ii] != -1) continue;
        double fact = solnApplied[gdof[ii]];
        if (std::fabs(fact) > 1.0e-10) {
          for (int jj = 0; jj < nsize; ++jj)
            if (fav[jj] != -1) F[jj] -= K[jj * nsize + ii] * fact;
        }
      }
    }
    // VecSetValues / MatSetValues ADD_VALUES, negative indices ignored
    for (int ii = 0; ii < nsize; ++ii) {
      const int32_t r = fav[ii];
      if (r < 0) continue;
      rhs[r] += F[ii];
      if (!vals) continue;
      const int32_t* cb = colidx + rowptr[r];
      const int32_t* ce = colidx + rowptr[r + 1];
      for (int jj = 0; jj < nsize; ++jj) {
        const int32_t c = fav[jj];
        if (c < 0) continue;
        const int32_t* p = std::lower_bound(cb, ce, c);
        vals[p - colidx] += K[ii * nsize + jj];
      }
    }
  };
  if (g_threads <= 1) {
    std::vector<double> K(nsize * nsize), F(nsize);
    std::vector<int32_t> fav(nsize);
    std::vector<int64_t> gdof(nsize);
    for (int64_t e = e0; e < e1; ++e) element(e, K.data(), F.data(), fav.data(), gdof.data());
    return;
  }
  // greedy element colouring in index order: an element takes the lowest colour none of
  // its nodes has yet
  const int npe = ei.npElem;
  int64_t nNode = 0;
  for (int64_t q = e0 * npe; q < e1 * npe; ++q) nNode = std::max<int64_t>(nNode, (int64_t)conn[q] + 1);
  std::vector<uint64_t> used(nNode, 0);
  std::vector<std::vector<int64_t>> cls;
  std::vector<int64_t> rest;  // elements for which all 64 colours are taken: run one by one
  for (int64_t e = e0; e < e1; ++e) {
    uint64_t u = 0;
    for (int a = 0; a < npe; ++a) u |= used[conn[e * npe + a]];
    if (u == ~0ull) { rest.push_back(e); continue; }
    int col = 0;
    while ((u >> col) & 1ull) ++col;
    if ((int)cls.size() <= col) cls.resize(col + 1);
    cls[col].push_back(e);
    for (int a = 0; a < npe; ++a) used[conn[e * npe + a]] |= 1ull << col;
  }
  if (!rest.empty()) {
    std::vector<double> K(nsize * nsize), F(nsize);
    std::vector<int32_t> fav(nsize);
    std::vector<int64_t> gdof(nsize);
    for (int64_t e : rest) element(e, K.data(), F.data(), fav.data(), gdof.data());
  }
  for (const auto& c : cls) {
#pragma omp parallel num_threads(g_threads)
    {
      std::vector<double> K(nsize * nsize), F(nsize);
      std::vector<int32_t> fav(nsize);
      std::vector<int64_t> gdof(nsize);
#pragma omp for schedule(dynamic, 256)
      for (int64_t k = 0; k < (int64_t)c.size(); ++k) element(c[k], K.data(), F.data(), fav.data(), gdof.data());
    }
  }
}

void orc_spmv(int64_t n, const int64_t* rowptr, const int32_t* colidx, const double* vals,
              const double* x, double* y) {
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    double s = 0.0;
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) s += vals[k] * x[colidx[k]];
    y[r] = s;
  }
}

struct orc_pc {
  int type;
  int64_t n;
  std::vector<double> dinv;             // Jacobi
  std::vector<int64_t> bptr;            // block Jacobi
  std::vector<double> binv;             // 16 per block, row-major inverse
  std::vector<int64_t> rowptr;          // ILU(0)
  std::vector<int32_t> colidx;
  std::vector<double> lu;
  std::vector<int64_t> diag;
  // threads > 1: rows grouped by dependency level of the forward (lower) and the
  // backward (upper) solve, levPtr[l] .. levPtr[l+1] in levRows
  std::vector<int64_t> levLPtr, levUPtr;
  std::vector<int32_t> levL, levU;
};

namespace {
// Gauss-Jordan inverse with partial pivoting of an m x m block (m <= 4).
bool small_inverse(int m, const double* A, double* Ai) {
  double a[4][8];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < 2 * m; ++j) a[i][j] = (j < m) ? A[i * m + j] : (j - m == i ? 1.0 : 0.0);
  for (int c = 0; c < m; ++c) {
    int p = c;
    for (int r = c + 1; r < m; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
    if (a[p][c] == 0.0) return false;
    if (p != c)
      for (int j = 0; j < 2 * m; ++j) std::swap(a[p][j], a[c][j]);
    double piv = 1.0 / a[c][c];
    for (int j = 0; j < 2 * m; ++j) a[c][j] *= piv;
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      double f = a[r][c];
      if (f == 0.0) continue;
      for (int j = 0; j < 2 * m; ++j) a[r][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) Ai[i * m + j] = a[i][m + j];
  return true;
}
}  // namespace

orc_pc* orc_pc_create(int type, int64_t n, const int64_t* rowptr, const int32_t* colidx,
                      const double* vals, int64_t nblk, const int64_t* blk_ptr, int* status) {
  orc_pc* pc = new orc_pc;
  pc->type = type;
  pc->n = n;
  *status = 0;
  if (type == 1) {
    pc->dinv.assign(n, 1.0);
    for (int64_t r = 0; r < n; ++r) {
      const int32_t* p = std::lower_bound(colidx + rowptr[r], colidx + rowptr[r + 1], (int32_t)r);
      double d = (p != colidx + rowptr[r + 1] && *p == r) ? vals[p - colidx] : 0.0;
      if (d == 0.0) { *status = -1; d = 1.0; }
      pc->dinv[r] = 1.0 / d;
    }
  } else if (type == 2) {
    pc->bptr.assign(blk_ptr, blk_ptr + nblk + 1);
    pc->binv.assign(16 * nblk, 0.0);
    for (int64_t k = 0; k < nblk; ++k) {
      const int64_t r0 = blk_ptr[k];
      const int m = (int)(blk_ptr[k + 1] - r0);
      double A[16];
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
          const int32_t c = (int32_t)(r0 + j);
          const int32_t* p = std::lower_bound(colidx + rowptr[r0 + i], colidx + rowptr[r0 + i + 1], c);
          A[i * m + j] = (p != colidx + rowptr[r0 + i + 1] && *p == c) ? vals[p - colidx] : 0.0;
        }
      if (!small_inverse(m, A, &pc->binv[16 * k])) *status = -1;
    }
  } else if (type == 3) {
    pc->rowptr.assign(rowptr, rowptr + n + 1);
    pc->colidx.assign(colidx, colidx + rowptr[n]);
    pc->lu.assign(vals, vals + rowptr[n]);
    pc->diag.assign(n, -1);
    for (int64_t r = 0; r < n; ++r) {
      const int32_t* p = std::lower_bound(colidx + rowptr[r], colidx + rowptr[r + 1], (int32_t)r);
      if (p == colidx + rowptr[r + 1] || *p != r) { *status = -1; continue; }
      pc->diag[r] = p - colidx;
    }
    if (*status == 0) {
      // IKJ ILU(0) (Saad Alg. 10.4) with a dense work index per row
      double* a = pc->lu.data();
      auto factor_row = [&](int64_t i, std::vector<int64_t>& iw) -> int {
        int st = 0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) iw[colidx[k]] = k;
        for (int64_t k = rowptr[i]; k < pc->diag[i]; ++k) {
          const int32_t j = colidx[k];
          double piv = a[pc->diag[j]];
          if (piv == 0.0) { st = -2; break; }
          a[k] /= piv;
          const double lij = a[k];
          for (int64_t q = pc->diag[j] + 1; q < rowptr[j + 1]; ++q) {
            const int64_t w = iw[colidx[q]];
            if (w >= 0) a[w] -= lij * a[q];
          }
        }
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) iw[colidx[k]] = -1;
        if (a[pc->diag[i]] == 0.0) st = -2;
        return st;
      };
      if (g_threads <= 1) {
        std::vector<int64_t> iw(n, -1);
        for (int64_t i = 0; i < n; ++i)
          if (int st = factor_row(i, iw)) *status = st;
      } else {
        // dependency levels: row i after every row j < i of its lower part (forward
        // solve and factorisation), after every row j > i of its upper part (backward)
        std::vector<int32_t> lv(n, 0);
        int nlL = 0, nlU = 0;
        for (int64_t i = 0; i < n; ++i) {
          int l = 0;
          for (int64_t k = rowptr[i]; k < pc->diag[i]; ++k) l = std::max(l, lv[colidx[k]] + 1);
          lv[i] = l;
          nlL = std::max(nlL, l + 1);
        }
        auto bucket = [&](int nl, std::vector<int64_t>& ptr, std::vector<int32_t>& rows) {
          ptr.assign(nl + 1, 0);
          for (int64_t i = 0; i < n; ++i) ptr[lv[i] + 1]++;
          for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
          rows.resize(n);
          std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
          for (int64_t i = 0; i < n; ++i) rows[pos[lv[i]]++] = (int32_t)i;
        };
        bucket(nlL, pc->levLPtr, pc->levL);
        for (int64_t i = n - 1; i >= 0; --i) {
          int l = 0;
          for (int64_t k = pc->diag[i] + 1; k < rowptr[i + 1]; ++k) l = std::max(l, lv[colidx[k]] + 1);
          lv[i] = l;  // rows j > i already hold their upper level
          nlU = std::max(nlU, l + 1);
        }
        bucket(nlU, pc->levUPtr, pc->levU);
        int st_all = 0;
#pragma omp parallel num_threads(g_threads) reduction(min : st_all)
        {
          std::vector<int64_t> iw(n, -1);
          for (int l = 0; l < nlL; ++l) {
#pragma omp for schedule(dynamic, 64)
            for (int64_t q = pc->levLPtr[l]; q < pc->levLPtr[l + 1]; ++q)
              st_all = std::min(st_all, factor_row(pc->levL[q], iw));
          }
        }
        if (st_all) *status = st_all;
      }
    }
  }
  return pc;
}

void orc_pc_apply(const orc_pc* pc, const double* r, double* z) {
  const int64_t n = pc->n;
  if (pc->type == 1) {
    for (int64_t i = 0; i < n; ++i) z[i] = pc->dinv[i] * r[i];
  } else if (pc->type == 2) {
    const int64_t nb = (int64_t)pc->bptr.size() - 1;
    for (int64_t k = 0; k < nb; ++k) {
      const int64_t r0 = pc->bptr[k];
      const int m = (int)(pc->bptr[k + 1] - r0);
      const double* B = &pc->binv[16 * k];
      for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int j = 0; j < m; ++j) s += B[i * m + j] * r[r0 + j];
        z[r0 + i] = s;
      }
    }
  } else if (pc->type == 3) {
    const int64_t* rp = pc->rowptr.data();
    const int32_t* ci = pc->colidx.data();
    const double* a = pc->lu.data();
    auto lower = [&](int64_t i) {  // L y = r, unit lower
      double s = r[i];
      for (int64_t k = rp[i]; k < pc->diag[i]; ++k) s -= a[k] * z[ci[k]];
      z[i] = s;
    };
    auto upper = [&](int64_t i) {  // U z = y
      double s = z[i];
      for (int64_t k = pc->diag[i] + 1; k < rp[i + 1]; ++k) s -= a[k] * z[ci[k]];
      z[i] = s / a[pc->diag[i]];
    };
    if (pc->levLPtr.empty()) {
      for (int64_t i = 0; i < n; ++i) lower(i);
      for (int64_t i = n - 1; i >= 0; --i) upper(i);
    } else {
      const int nlL = (int)pc->levLPtr.size() - 1, nlU = (int)pc->levUPtr.size() - 1;
#pragma omp parallel num_threads(g_threads)
      {
        for (int l = 0; l < nlL; ++l) {
#pragma omp for schedule(static)
          for (int64_t q = pc->levLPtr[l]; q < pc->levLPtr[l + 1]; ++q) lower(pc->levL[q]);
        }
        for (int l = 0; l < nlU; ++l) {
#pragma omp for schedule(static)
          for (int64_t q = pc->levUPtr[l]; q < pc->levUPtr[l + 1]; ++q) upper(pc->levU[q]);
        }
      }
    }
  } else {
    std::memcpy(z, r, sizeof(double) * n);
  }
}

void orc_pc_destroy(orc_pc* pc) { delete pc; }

int64_t orc_pc_lu(const orc_pc* pc, double* out) {
  if (pc->type != 3) return -1;
  if (out) std::copy(pc->lu.begin(), pc->lu.end(), out);
  return (int64_t)pc->lu.size();
}

namespace {
double dot(int64_t n, const double* a, const double* b) {
  if (g_threads <= 1) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
  }
  const int64_t nc = (n + kDotChunk - 1) / kDotChunk;
  std::vector<double> part(nc);
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int64_t c = 0; c < nc; ++c) {
    double s = 0.0;
    const int64_t hi = std::min(n, (c + 1) * kDotChunk);
    for (int64_t i = c * kDotChunk; i < hi; ++i) s += a[i] * b[i];
    part[c] = s;
  }
  double s = 0.0;
  for (int64_t c = 0; c < nc; ++c) s += part[c];
  return s;
}

}  // namespace
extern "C++" {
namespace {
// f(i) for i in [0, n), split over the threads when there are several
template <typename F>
void for_rows(int64_t n, F f) {
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1) schedule(static)
  for (int64_t i = 0; i < n; ++i) f(i);
}
}  // namespace
}

void orc_set_threads(int n) { g_threads = std::max(1, n); }
int orc_get_threads(void) { return g_threads; }

int orc_bicgstab(int64_t n, const int64_t* rowptr, const int32_t* colidx, const double* vals,
                 const orc_pc* pc, const double* b, double* x, double rtol, int maxit, int* iters,
                 double* relres) {
  std::vector<double> r(n), rh(n), p(n, 0.0), v(n, 0.0), s(n), t(n), ph(n), sh(n);
  for_rows(n, [&](int64_t i) { x[i] = 0.0; r[i] = b[i]; rh[i] = b[i]; });
  const double bnorm = std::sqrt(dot(n, b, b));
  *iters = 0;
  if (bnorm == 0.0) { *relres = 0.0; return 0; }
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  double rnorm = bnorm;
  for (int it = 1; it <= maxit; ++it) {
    const double rho1 = dot(n, rh.data(), r.data());
    if (rho1 == 0.0) { *iters = it - 1; *relres = rnorm / bnorm; return 2; }
    if (it == 1) {
      for_rows(n, [&](int64_t i) { p[i] = r[i]; });
    } else {
      const double beta = (rho1 / rho) * (alpha / omega);
      for_rows(n, [&](int64_t i) { p[i] = r[i] + beta * (p[i] - omega * v[i]); });
    }
    orc_pc_apply(pc, p.data(), ph.data());
    orc_spmv(n, rowptr, colidx, vals, ph.data(), v.data());
    const double rv = dot(n, rh.data(), v.data());
    if (rv == 0.0) { *iters = it - 1; *relres = rnorm / bnorm; return 2; }
    alpha = rho1 / rv;
    for_rows(n, [&](int64_t i) { s[i] = r[i] - alpha * v[i]; });
    const double snorm = std::sqrt(dot(n, s.data(), s.data()));
    if (snorm <= rtol * bnorm) {
      for_rows(n, [&](int64_t i) { x[i] += alpha * ph[i]; });
      *iters = it; *relres = snorm / bnorm; return 0;
    }
    orc_pc_apply(pc, s.data(), sh.data());
    orc_spmv(n, rowptr, colidx, vals, sh.data(), t.data());
    const double tt = dot(n, t.data(), t.data());
    omega = (tt == 0.0) ? 0.0 : dot(n, t.data(), s.data()) / tt;
    for_rows(n, [&](int64_t i) {
      x[i] += alpha * ph[i] + omega * sh[i];
      r[i] = s[i] - omega * t[i];
    });
    rnorm = std::sqrt(dot(n, r.data(), r.data()));
    rho = rho1;
    *iters = it;
    if (rnorm <= rtol * bnorm) { *relres = rnorm / bnorm; return 0; }
    if (omega == 0.0) { *relres = rnorm / bnorm; return 2; }
  }
  *relres = rnorm / bnorm;
  return 1;
}

int orc_gmres(int64_t n, const int64_t* rowptr, const int32_t* colidx, const double* vals,
              const orc_pc* pc, const double* b, double* x, int m, double rtol, int maxit,
              int* iters, double* relres) {
  std::vector<double> V((size_t)(m + 1) * n), H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1),
      w(n), z(n), r(n), y(m);
  for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
  const double bnorm = std::sqrt(dot(n, b, b));
  *iters = 0;
  if (bnorm == 0.0) { *relres = 0.0; return 0; }
  int it = 0;
  double rnorm = bnorm;
  while (it < maxit) {
    orc_spmv(n, rowptr, colidx, vals, x, w.data());
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - w[i];
    double beta = std::sqrt(dot(n, r.data(), r.data()));
    rnorm = beta;
    if (beta <= rtol * bnorm) break;
    for (int64_t i = 0; i < n; ++i) V[i] = r[i] / beta;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    for (; k < m && it < maxit; ++k) {
      ++it;
      orc_pc_apply(pc, &V[(size_t)k * n], z.data());
      orc_spmv(n, rowptr, colidx, vals, z.data(), w.data());
      for (int j = 0; j <= k; ++j) {
        double h = dot(n, w.data(), &V[(size_t)j * n]);
        H[(size_t)j * m + k] = h;
        for (int64_t i = 0; i < n; ++i) w[i] -= h * V[(size_t)j * n + i];
      }
      double hn = std::sqrt(dot(n, w.data(), w.data()));
      H[(size_t)(k + 1) * m + k] = hn;
      if (hn != 0.0)
        for (int64_t i = 0; i < n; ++i) V[(size_t)(k + 1) * n + i] = w[i] / hn;
      for (int j = 0; j < k; ++j) {
        double t1 = H[(size_t)j * m + k], t2 = H[(size_t)(j + 1) * m + k];
        H[(size_t)j * m + k] = cs[j] * t1 + sn[j] * t2;
        H[(size_t)(j + 1) * m + k] = -sn[j] * t1 + cs[j] * t2;
      }
      double a = H[(size_t)k * m + k], bb = H[(size_t)(k + 1) * m + k];
      double d = std::sqrt(a * a + bb * bb);
      cs[k] = (d == 0.0) ? 1.0 : a / d;
      sn[k] = (d == 0.0) ? 0.0 : bb / d;
      H[(size_t)k * m + k] = d;
      H[(size_t)(k + 1) * m + k] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      rnorm = std::fabs(g[k + 1]);
      if (rnorm <= rtol * bnorm || hn == 0.0) { ++k; break; }
    }
    for (int j = k - 1; j >= 0; --j) {
      double s = g[j];
      for (int q = j + 1; q < k; ++q) s -= H[(size_t)j * m + q] * y[q];
      y[j] = s / H[(size_t)j * m + j];
    }
    std::fill(w.begin(), w.end(), 0.0);
    for (int j = 0; j < k; ++j)
      for (int64_t i = 0; i < n; ++i) w[i] += y[j] * V[(size_t)j * n + i];
    orc_pc_apply(pc, w.data(), z.data());
    for (int64_t i = 0; i < n; ++i) x[i] += z[i];
    if (rnorm <= rtol * bnorm) break;
  }
  orc_spmv(n, rowptr, colidx, vals, x, w.data());
  for (int64_t i = 0; i < n; ++i) r[i] = b[i] - w[i];
  rnorm = std::sqrt(dot(n, r.data(), r.data()));
  *iters = it;
  *relres = rnorm / bnorm;
  return (rnorm <= rtol * bnorm * 1.0000001) ? 0 : 1;
}

}  // extern "C"
