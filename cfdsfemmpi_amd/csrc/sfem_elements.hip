// sfem_elements.hip -- element geometry and the stabilised Navier-Stokes element
// kernels (residual Fe and tangent Ke) with scatter-add into the device CSR.
//
// One "team" of lanes per element: a team of TEAM lanes (64 for Hexa8 = one
// wavefront, 16 for the 2D / Tetra4 elements = a quarter wavefront) owns one
// element.  Lane (i, j) of the team owns the ndof x ndof block of node pair
// (i, j) of Ke and keeps it in registers across all Gauss points; the
// per-Gauss-point data (shape gradients, interpolated state, stabilisation
// parameters) is computed cooperatively by lane (gp, k) and staged in LDS.
//
// The arithmetic restates calcStiffnessAndResidual of the reference
// (src/base/LagrangeElem3DNavierStokesHexa8Node.cpp:764-1074,
// ...Tetra4Node.cpp:769-1078, ...Quad4Node.cpp:683-970, ...Tria3Node.cpp:683-968),
// regrouped so that every (i, j, gp) term is a handful of FMAs on per-node and
// per-Gauss-point factors:
//   K_uu(a,b) += (w Ni Nj) A_ab + d_ab [c w Ni Nj + mu af w dNi.dNj + w Ni u_j]
//   K_up(a)   -= af (w dNi_a) Nj
//   K_pu(b)   += s af w Ni dNj_b + Nj g_i(b) + s tauP (w dNi_b) d_j     (PSPG)
//   K_pp      += s tauP af w dNi.dNj
//   2D SUPG:    K_u(a). += Da_i Dj(a, .)
// with A = af rho gradPrev, u_j = af rho (uPrev.dNj), d_j = u_j + rho td8 Nj,
// g_i(b) = s tauP sum_a (w dNi_a) A_ab, s = -1 for Hexa8 and +1 otherwise
// (Hexa8Node.cpp:984-986,1027-1030 vs Tetra4Node.cpp:990-992,1033-1036).
//
// Scatter: elements are processed one colour class per launch; no two
// elements of a class share a node, so every CSR entry and RHS entry has at
// most one writer per launch and plain read-modify-write stores are exact and
// deterministic (summation order = colour order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "sfem_internal.hpp"

namespace sfem {

template <int ET> struct Traits;
template <> struct Traits<TRIA3> {
  static constexpr int NDIM = 2, NPE = 3, NDOF = 3, NGP = 1, TEAM = 16;
  static constexpr double SGN = 1.0, CI = 4.0;
  static constexpr bool SUPG = true;
};
template <> struct Traits<QUAD4> {
  static constexpr int NDIM = 2, NPE = 4, NDOF = 3, NGP = 4, TEAM = 16;
  static constexpr double SGN = 1.0, CI = 4.0;
  static constexpr bool SUPG = true;
};
template <> struct Traits<TETRA4> {
  static constexpr int NDIM = 3, NPE = 4, NDOF = 4, NGP = 1, TEAM = 16;
  static constexpr double SGN = 1.0, CI = 10.0;
  static constexpr bool SUPG = false;
};
template <> struct Traits<HEXA8> {
  static constexpr int NDIM = 3, NPE = 8, NDOF = 4, NGP = 8, TEAM = 64;
  static constexpr double SGN = -1.0, CI = 10.0;
  static constexpr bool SUPG = false;
};

// PI of the reference (headersBasic.h:28) -- the nearest double to pi.
constexpr double kPi = 3.14159265358979323846;
// Two-point Gauss abscissa as the reference writes it (UtilitiesQuadrature.cpp:25-26).
constexpr double kG2 = 0.577350269189626;

// Gauss point gp of the element's rule: x-fastest ordering (UtilitiesQuadrature.cpp:245-253,
// 285-297); Tria3 uses the quad 1-point rule at (0,0) with weight 4 (Tria3Node.cpp:56).
template <int ET>
__host__ __device__ inline void gauss_point(int gp, double* xi, double& wgt) {
  if constexpr (ET == HEXA8) {
    xi[0] = (gp & 1) ? kG2 : -kG2;
    xi[1] = (gp & 2) ? kG2 : -kG2;
    xi[2] = (gp & 4) ? kG2 : -kG2;
    wgt = 1.0;
  } else if constexpr (ET == QUAD4) {
    xi[0] = (gp & 1) ? kG2 : -kG2;
    xi[1] = (gp & 2) ? kG2 : -kG2;
    xi[2] = 0.0;
    wgt = 1.0;
  } else if constexpr (ET == TETRA4) {
    xi[0] = 0.25; xi[1] = 0.25; xi[2] = 0.25;
    wgt = 1.0 / 6.0;
  } else {
    xi[0] = 0.0; xi[1] = 0.0; xi[2] = 0.0;
    wgt = 4.0;
  }
}

// Degree-1 shape functions and parametric derivatives (BasisFunctionsLagrange.cpp:517-546,
// 687-732, 873-911, 1019-1080).  For Hexa8 the xi2/xi3 derivatives of nodes 2,3 and 6,7
// are taken exactly as the reference tabulates them (:1056-1074), which pairs them
// with the opposite node of each x-edge.
template <int ET>
__host__ __device__ inline void shape(const double* xi, double* N, double* d1, double* d2, double* d3) {
  if constexpr (ET == HEXA8) {
    const double v11 = 1.0 - xi[0], v12 = 1.0 + xi[0];
    const double v21 = 1.0 - xi[1], v22 = 1.0 + xi[1];
    const double v31 = 1.0 - xi[2], v32 = 1.0 + xi[2];
    N[0] = 0.125 * v11 * v21 * v31; N[1] = 0.125 * v12 * v21 * v31;
    N[2] = 0.125 * v12 * v22 * v31; N[3] = 0.125 * v11 * v22 * v31;
    N[4] = 0.125 * v11 * v21 * v32; N[5] = 0.125 * v12 * v21 * v32;
    N[6] = 0.125 * v12 * v22 * v32; N[7] = 0.125 * v11 * v22 * v32;
    d1[0] = -0.125 * v21 * v31; d1[1] = 0.125 * v21 * v31;
    d1[2] = 0.125 * v22 * v31;  d1[3] = -0.125 * v22 * v31;
    d1[4] = -0.125 * v21 * v32; d1[5] = 0.125 * v21 * v32;
    d1[6] = 0.125 * v22 * v32;  d1[7] = -0.125 * v22 * v32;
    d2[0] = -0.125 * v11 * v31; d2[1] = -0.125 * v12 * v31;
    d2[2] = 0.125 * v11 * v31;  d2[3] = 0.125 * v12 * v31;
    d2[4] = -0.125 * v11 * v32; d2[5] = -0.125 * v12 * v32;
    d2[6] = 0.125 * v11 * v32;  d2[7] = 0.125 * v12 * v32;
    d3[0] = -0.125 * v11 * v21; d3[1] = -0.125 * v12 * v21;
    d3[2] = -0.125 * v11 * v22; d3[3] = -0.125 * v12 * v22;
    d3[4] = 0.125 * v11 * v21;  d3[5] = 0.125 * v12 * v21;
    d3[6] = 0.125 * v11 * v22;  d3[7] = 0.125 * v12 * v22;
  } else if constexpr (ET == QUAD4) {
    const double b1 = 0.5 * (1.0 - xi[1]), b2 = 0.5 * (1.0 + xi[1]);
    const double a1 = 0.5 * (1.0 - xi[0]), a2 = 0.5 * (1.0 + xi[0]);
    N[0] = b1 * a1; N[1] = b1 * a2; N[2] = b2 * a2; N[3] = b2 * a1;
    d1[0] = b1 * -0.5; d1[1] = b1 * 0.5; d1[2] = b2 * 0.5; d1[3] = b2 * -0.5;
    d2[0] = -0.5 * a1; d2[1] = -0.5 * a2; d2[2] = 0.5 * a2; d2[3] = 0.5 * a1;
  } else if constexpr (ET == TETRA4) {
    N[0] = 1.0 - xi[0] - xi[1] - xi[2]; N[1] = xi[0]; N[2] = xi[1]; N[3] = xi[2];
    d1[0] = -1.0; d1[1] = 1.0; d1[2] = 0.0; d1[3] = 0.0;
    d2[0] = -1.0; d2[1] = 0.0; d2[2] = 1.0; d2[3] = 0.0;
    d3[0] = -1.0; d3[1] = 0.0; d3[2] = 0.0; d3[3] = 1.0;
  } else {
    N[0] = 1.0 - xi[0] - xi[1]; N[1] = xi[0]; N[2] = xi[1];
    d1[0] = -1.0; d1[1] = 1.0; d1[2] = 0.0;
    d2[0] = -1.0; d2[1] = 0.0; d2[2] = 1.0;
  }
}

// Geometry at one Gauss point (computeBasisFunctions2D/3D, BasisFunctionsLagrange.cpp:18-186):
// B(r,c) = sum_n x_c(n) dN_n/dxi_r, Jac = det B, dN/dx_c = sum_r dN/dxi_r Binv(c,r).
// X is [NPE][3].  Writes the physical gradients of node k only when k >= 0.
template <int ET>
__host__ __device__ inline void gp_geometry(const double (*X)[3], int gp, int k, double& Nk,
                                            double* dNk, double& dvol) {
  using C = Traits<ET>;
  double xi[3], wgt;
  gauss_point<ET>(gp, xi, wgt);
  double N[C::NPE], d1[C::NPE], d2[C::NPE], d3[C::NPE];
  shape<ET>(xi, N, d1, d2, d3);
  double Jac;
  if constexpr (C::NDIM == 2) {
    double B00 = 0, B10 = 0, B01 = 0, B11 = 0;
    for (int n = 0; n < C::NPE; ++n) {
      B00 += X[n][0] * d1[n]; B10 += X[n][0] * d2[n];
      B01 += X[n][1] * d1[n]; B11 += X[n][1] * d2[n];
    }
    Jac = B00 * B11 - B01 * B10;
    if (k >= 0) {
      const double di = 1.0 / Jac;
      const double i00 = B11 * di, i01 = -B01 * di, i10 = -B10 * di, i11 = B00 * di;
      Nk = N[k];
      dNk[0] = d1[k] * i00 + d2[k] * i01;
      dNk[1] = d1[k] * i10 + d2[k] * i11;
      dNk[2] = 0.0;
    }
  } else {
    double B[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int n = 0; n < C::NPE; ++n)
      for (int c = 0; c < 3; ++c) {
        B[0][c] += X[n][c] * d1[n];
        B[1][c] += X[n][c] * d2[n];
        B[2][c] += X[n][c] * d3[n];
      }
    const double c00 = B[1][1] * B[2][2] - B[1][2] * B[2][1];
    const double c01 = B[1][2] * B[2][0] - B[1][0] * B[2][2];
    const double c02 = B[1][0] * B[2][1] - B[1][1] * B[2][0];
    Jac = B[0][0] * c00 + B[0][1] * c01 + B[0][2] * c02;
    if (k >= 0) {
      const double id = 1.0 / Jac;
      // rows c of Binv: Binv(c, r)
      const double i00 = c00 * id, i01 = (B[0][2] * B[2][1] - B[0][1] * B[2][2]) * id,
                   i02 = (B[0][1] * B[1][2] - B[0][2] * B[1][1]) * id;
      const double i10 = c01 * id, i11 = (B[0][0] * B[2][2] - B[0][2] * B[2][0]) * id,
                   i12 = (B[0][2] * B[1][0] - B[0][0] * B[1][2]) * id;
      const double i20 = c02 * id, i21 = (B[0][1] * B[2][0] - B[0][0] * B[2][1]) * id,
                   i22 = (B[0][0] * B[1][1] - B[0][1] * B[1][0]) * id;
      Nk = N[k];
      dNk[0] = d1[k] * i00 + d2[k] * i01 + d3[k] * i02;
      dNk[1] = d1[k] * i10 + d2[k] * i11 + d3[k] * i12;
      dNk[2] = d1[k] * i20 + d2[k] * i21 + d3[k] * i22;
    }
  }
  dvol = wgt * Jac;
}

// ---------------------------------------------------------------------------
// prepareElemData: volume, charlen (2D sqrt(4|A|/pi), 3D (6|V|/pi)^(1/3)),
// negative-Jacobian detection (Hexa8Node.cpp:31-106, Quad4Node.cpp:36-113).
// ---------------------------------------------------------------------------
template <int ET>
__global__ __launch_bounds__(256) void k_prepare(int64_t nElem, const double* __restrict__ coords,
                                                 const int32_t* __restrict__ conn,
                                                 double* __restrict__ charlen,
                                                 double* __restrict__ vol, int64_t* bad) {
  using C = Traits<ET>;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nElem) return;
  double X[C::NPE][3];
  for (int n = 0; n < C::NPE; ++n) {
    const int64_t node = conn[e * C::NPE + n];
    for (int c = 0; c < 3; ++c) X[n][c] = (c < C::NDIM) ? coords[node * C::NDIM + c] : 0.0;
  }
  double v = 0.0;
  bool neg = false;
  for (int gp = 0; gp < C::NGP; ++gp) {
    double Nk, dNk[3], dvol;
    gp_geometry<ET>(X, gp, -1, Nk, dNk, dvol);
    if (dvol < 0.0) neg = true;  // weights are positive, so sign(dvol) = sign(Jac)
    v += dvol;
  }
  if (neg) atomicMin((unsigned long long*)bad, (unsigned long long)e);
  vol[e] = v;
  charlen[e] = (C::NDIM == 2) ? sqrt(4.0 * fabs(v) / kPi) : pow(6.0 * fabs(v) / kPi, 1.0 / 3.0);
}

// Interpolated Gauss-point quantities, in the order of the reference's loop
// (Hexa8Node.cpp:828-889): uPrev, gradPrev, u, grad, p, grad p, uDot.
// src: 0 solnPrev, 1 solnCur, 2 solnDotCur; der: 0 = N, 1+c = d/dx_c.
template <int NDIM>
__device__ inline void qdesc(int t, int& src, int& comp, int& der) {
  constexpr int D = NDIM, DD = NDIM * NDIM;
  if (t < D) { src = 0; comp = t; der = 0; return; }
  t -= D;
  if (t < DD) { src = 0; comp = t / D; der = 1 + t % D; return; }
  t -= DD;
  if (t < D) { src = 1; comp = t; der = 0; return; }
  t -= D;
  if (t < DD) { src = 1; comp = t / D; der = 1 + t % D; return; }
  t -= DD;
  if (t == 0) { src = 1; comp = D; der = 0; return; }
  t -= 1;
  if (t < D) { src = 1; comp = D; der = 1 + t; return; }
  t -= D;
  src = 2; comp = t; der = 0;
}

template <int ET>
struct TeamLds {
  using C = Traits<ET>;
  static constexpr int NQ = 2 * C::NDIM + 2 * C::NDIM * C::NDIM + 1 + 2 * C::NDIM;
  double X[C::NPE][3];
  double S[3][C::NPE][4];
  double app[C::NPE][4];
  double phi[C::NGP][C::NPE][4];
  double W[C::NGP];
  double Q[C::NGP][NQ];
  double U[C::NGP][10];              // A (NDIM x NDIM, row-major in 3x3 slots), s*tauP
  double P[C::NGP][C::NPE][6];       // g_k(0..2), u_k, d_k, Da_k
  double Fk[C::NGP][C::NPE][4];
  double lift[C::NPE][C::NPE][4];
  int32_t info[C::NPE][4];           // rowStart, rowlen, mask, firstRow
};

// Per-wavefront staging of the pipelined Hexa8 loop: the residual and lift partial sums
// are reduced across lanes in registers, so only the shared per-element data is here.
struct HexLds {
  using C = Traits<HEXA8>;
  static constexpr int NQ = TeamLds<HEXA8>::NQ;
  double X[8][3];
  double S[3][8][4];
  double app[8][4];
  double phi[8][8][4];
  double W[8];
  double Q[8][NQ];
  double U[8][10];
  double P[8][8][6];
  int32_t info[8][4];
};

// P2: shape gradients of node k at Gauss point gp, lane l = gp * NPE + k (l < NGP * NPE).
template <int ET, class LDS>
__device__ inline void phase_geometry(LDS& L, int l) {
  using C = Traits<ET>;
  const int gp = l / C::NPE, k = l % C::NPE;
  double Nk, dNk[3], dvol;
  gp_geometry<ET>(L.X, gp, k, Nk, dNk, dvol);
  L.phi[gp][k][0] = Nk;
  L.phi[gp][k][1] = dNk[0];
  L.phi[gp][k][2] = dNk[1];
  L.phi[gp][k][3] = dNk[2];
  if (k == 0) L.W[gp] = dvol;
}

// P3: interpolated state at each Gauss point, summed over nodes in node order.
template <int ET, class LDS>
__device__ inline void phase_interp(LDS& L, int l) {
  using C = Traits<ET>;
  constexpr int NQ = TeamLds<ET>::NQ;
  const int gp = l / C::NPE, k = l % C::NPE;
  for (int t = k; t < NQ; t += C::NPE) {
    int src, comp, der;
    qdesc<C::NDIM>(t, src, comp, der);
    double s = 0.0;
    for (int nn = 0; nn < C::NPE; ++nn) s += L.S[src][nn][comp] * L.phi[gp][nn][der];
    L.Q[gp][t] = s;
  }
}

// P4: per-Gauss-point derived data (stabilisation parameters of
// evaluateStabParams_algo3, stabilisationRoutines.h:128-143, with G = 4/h^2 I), the
// residual contribution Fk of node k at gp, and the per-(node, gp) factors of the
// tangent, for lane l = gp * NPE + k.
template <int ET, bool WANT_K = true, class LDS>
__device__ inline void phase_gauss(LDS& L, int l, double h, const ElemParams& p, double (&Fk)[4]) {
  using C = Traits<ET>;
  constexpr int D = C::NDIM;
  const int gp = l / C::NPE, k = l % C::NPE;
  const double* Q = L.Q[gp];
  const double* vp = Q;
  const double* gpv = Q + D;                  // gradPrev row-major
  const double* vel = Q + D + D * D;
  const double* gr = Q + 2 * D + D * D;       // grad row-major
  const double pres = Q[2 * D + 2 * D * D];
  const double* dp = Q + 2 * D + 2 * D * D + 1;
  const double* vd = Q + 3 * D + 2 * D * D + 1;
  const double rho = p.rho, mu = p.mu, af = p.af;

  double res2[3], rStab[3];
  for (int a = 0; a < D; ++a) {
    double s1 = gpv[a * D] * vel[0], s2 = gr[a * D] * vp[0], s3 = gpv[a * D] * vp[0];
    for (int c = 1; c < D; ++c) {
      s1 += gpv[a * D + c] * vel[c];
      s2 += gr[a * D + c] * vp[c];
      s3 += gpv[a * D + c] * vp[c];
    }
    res2[a] = rho * (vd[a] + (s1 + s2 - s3));
    rStab[a] = res2[a] + dp[a];
  }
  const double fact = 4.0 / h / h;
  double uGu = 0.0;
  for (int a = 0; a < D; ++a) uGu += vp[a] * (fact * vp[a]);
  const double G2 = D * (fact * fact);
  double tau0 = 1.0 / sqrt(uGu + C::CI * (mu * mu / rho / rho) * G2);
  tau0 = tau0 / rho;
  const double tauS = tau0 * p.stab[0];
  const double tauP = tau0 * p.stab[1];
  double trace = gr[0];
  for (int a = 1; a < D; ++a) trace += gr[a * D + a];

  const double w = L.W[gp];
  const double Nk = L.phi[gp][k][0];
  double b[3];
  for (int c = 0; c < D; ++c) b[c] = L.phi[gp][k][1 + c] * w;
  const double b4 = Nk * w;
  double Da = vp[0] * b[0];
  for (int c = 1; c < D; ++c) Da += vp[c] * b[c];
  Da = rho * Da * tauS;
  // residual of node k at gp (Hexa8Node.cpp:1049-1060, Quad4Node.cpp:936-950)
  for (int a = 0; a < D; ++a) {
    double s = b4 * res2[a];
    for (int c = 0; c < D; ++c) s += b[c] * (mu * gr[a * D + c] - (a == c ? pres : 0.0));
    double f = -s;
    if constexpr (C::SUPG) f -= Da * rStab[a];
    Fk[a] = f;
  }
  double br = b[0] * rStab[0];
  for (int a = 1; a < D; ++a) br += b[a] * rStab[a];
  Fk[D] = C::SGN * (-(b4 * trace) - tauP * br);
  for (int a = D + 1; a < 4; ++a) Fk[a] = 0.0;
  if constexpr (!WANT_K) return;
  // tangent factors
  const double stp = C::SGN * tauP;
  for (int c = 0; c < D; ++c) {
    double g = 0.0;
    for (int a = 0; a < D; ++a) g += b[a] * (af * rho * gpv[a * D + c]);
    L.P[gp][k][c] = stp * g;
  }
  double ud = vp[0] * L.phi[gp][k][1];
  for (int c = 1; c < D; ++c) ud += vp[c] * L.phi[gp][k][1 + c];
  const double uk = af * rho * ud;
  L.P[gp][k][3] = uk;
  L.P[gp][k][4] = uk + rho * p.acceFact * Nk;
  L.P[gp][k][5] = Da;
  if (k == 0) {
    for (int a = 0; a < D; ++a)
      for (int c = 0; c < D; ++c) L.U[gp][3 * a + c] = af * rho * gpv[a * D + c];
    L.U[gp][9] = stp;
  }
}

// Ke accumulation of lane (i, j) in scalar type T (double, or float for the mixed path).
// GPU Gauss points are unrolled together.
template <int ET, typename T, int GPU = Traits<ET>::NGP, class LDS>
__device__ inline void accumulate_block(const LDS& L, int i, int j, const ElemParams& p,
                                        T (&K)[Traits<ET>::NDOF][Traits<ET>::NDOF]) {
  using C = Traits<ET>;
  constexpr int D = C::NDIM;
  const T af = (T)p.af, cst = (T)(p.rho * p.acceFact), muaf = (T)(p.mu * p.af);
  const T saf = (T)(C::SGN * p.af);
  for (int a = 0; a < C::NDOF; ++a)
    for (int b = 0; b < C::NDOF; ++b) K[a][b] = (T)0;
#pragma unroll GPU
  for (int gp = 0; gp < C::NGP; ++gp) {
    const T w = (T)L.W[gp];
    const T Ni = (T)L.phi[gp][i][0], Nj = (T)L.phi[gp][j][0];
    T wdNi[D], dNj[D];
    for (int c = 0; c < D; ++c) {
      wdNi[c] = w * (T)L.phi[gp][i][1 + c];
      dNj[c] = (T)L.phi[gp][j][1 + c];
    }
    const T wNi = w * Ni;
    const T s1 = wNi * Nj;
    T dd = wdNi[0] * dNj[0];
    for (int c = 1; c < D; ++c) dd += wdNi[c] * dNj[c];
    const T uj = (T)L.P[gp][j][3], dj = (T)L.P[gp][j][4];
    const T s2 = cst * s1 + muaf * dd + wNi * uj;
    const T stp = (T)L.U[gp][9];
    for (int a = 0; a < D; ++a) {
      for (int b = 0; b < D; ++b) K[a][b] += s1 * (T)L.U[gp][3 * a + b];
      K[a][a] += s2;
      K[a][D] -= af * wdNi[a] * Nj;
    }
    for (int b = 0; b < D; ++b)
      K[D][b] += saf * wNi * dNj[b] + Nj * (T)L.P[gp][i][b] + stp * wdNi[b] * dj;
    K[D][D] += stp * af * dd;
    if constexpr (C::SUPG) {
      const T Da = (T)L.P[gp][i][5];
      for (int a = 0; a < D; ++a) {
        for (int b = 0; b < D; ++b) K[a][b] += Da * (T)L.U[gp][3 * a + b] * Nj;
        K[a][a] += Da * dj;
        K[a][D] += Da * af * dNj[a];
      }
    }
  }
}

// Read-modify-write of one CSR row segment of a node-pair block: a full block (all
// NDOF columns of node j free) goes through one vector access of NDOF consecutive
// values (the segment is only 8- or 4-byte aligned), a partial one value by value.
template <int NDOF, typename T>
__device__ inline void add_row_segment(T* __restrict__ row, const T (&k)[NDOF], int mj) {
  constexpr int FULL = (1 << NDOF) - 1;
  if (mj == FULL) {
    typedef T vec_t __attribute__((ext_vector_type(NDOF == 3 ? 4 : NDOF), aligned(sizeof(T))));
    if constexpr (NDOF == 4) {
      vec_t r = *reinterpret_cast<const vec_t*>(row);
      r.x += k[0]; r.y += k[1]; r.z += k[2]; r.w += k[3];
      *reinterpret_cast<vec_t*>(row) = r;
      return;
    }
  }
  int rb = 0;
#pragma unroll
  for (int bq = 0; bq < NDOF; ++bq) {
    if (!((mj >> bq) & 1)) continue;
    row[rb] += k[bq];
    ++rb;
  }
}

// Generic element loop (all element types; the Hexa8 colour loops use the pipelined
// kernel below and this one serves the per-element probe): TEAMS elements per 256-lane
// workgroup, one per team.
// WGS lanes per workgroup (WGS / TEAM elements side by side); PERSIST: the grid is what
// the GPU holds at once and each workgroup walks the slots blockIdx.x, blockIdx.x +
// gridDim.x, ... (in groups of WGS / TEAM), instead of one group per workgroup.  Every
// variant does the same arithmetic per element, and no two elements of a colour class
// add to the same entry, so all variants give the same result.
template <int ET, int MODE, int WGS = 256, bool PERSIST = false>
__global__ __launch_bounds__(WGS) void k_assemble(AsmView v, ElemParams p0,
                                                  const int32_t* __restrict__ elems, int64_t n,
                                                  int lift, int64_t probe_e) {
  using C = Traits<ET>;
  constexpr int TEAMS = WGS / C::TEAM;
  constexpr int D = C::NDIM, NPE = C::NPE, NGP = C::NGP, NDOF = C::NDOF;
  __shared__ TeamLds<ET> lds[TEAMS];
  const int team = threadIdx.x / C::TEAM;
  const int l = threadIdx.x % C::TEAM;
  TeamLds<ET>& L = lds[team];
  const int64_t limit = probe_e >= 0 ? 1 : n;
  for (int64_t grp = blockIdx.x; grp * TEAMS < limit; grp += gridDim.x) {
  ElemParams p = p0;
  const int64_t slot = grp * TEAMS + team;
  const bool probe = probe_e >= 0;
  const bool active = probe ? (slot == 0) : (slot < n);
  const int64_t e = active ? (probe ? probe_e : (int64_t)elems[slot]) : 0;
  const bool doLift = (lift != 0) && !probe;
  if (v.elem_mu && active) p.mu = v.elem_mu[e];  // batched meshes carry their own viscosity

  // P1: nodal data of the element
  if (active && l < NPE) {
    const int64_t node = v.conn[e * NPE + l];
    for (int c = 0; c < 3; ++c) L.X[l][c] = (c < D) ? v.coords[node * D + c] : 0.0;
    const int64_t o = node * NDOF;
    for (int c = 0; c < NDOF; ++c) {
      L.S[0][l][c] = v.solnPrev[o + c];
      L.S[1][l][c] = v.solnCur[o + c];
      L.S[2][l][c] = v.solnDotCur[o + c];
      L.app[l][c] = doLift ? v.solnApplied[o + c] : 0.0;
    }
    const int4 inf = reinterpret_cast<const int4*>(v.nodeInfo)[node];
    L.info[l][0] = inf.x; L.info[l][1] = inf.y; L.info[l][2] = inf.z; L.info[l][3] = inf.w;
  }
  __syncthreads();
  if (active && l < NGP * NPE) phase_geometry<ET>(L, l);
  __syncthreads();
  if (active && l < NGP * NPE) phase_interp<ET>(L, l);
  __syncthreads();
  if (active && l < NGP * NPE) {
    double Fk[4];
    phase_gauss<ET>(L, l, v.charlen[e], p, Fk);
    const int gp = l / NPE, k = l % NPE;
    for (int a = 0; a < NDOF; ++a) L.Fk[gp][k][a] = Fk[a];
  }
  __syncthreads();

  // P5: tangent block of node pair (i, j), lift partial sums, CSR scatter
  if constexpr (MODE != 1) {
    if (active && l < NPE * NPE) {
      const int i = l / NPE, j = l % NPE;
      const int mi = L.info[i][2], mj = L.info[j][2];
      if constexpr (MODE == 0) {
        double K[NDOF][NDOF];
        accumulate_block<ET, double>(L, i, j, p, K);
        if (probe) {
          for (int a = 0; a < NDOF; ++a)
            for (int bq = 0; bq < NDOF; ++bq)
              v.probeK[(i * NDOF + a) * (NPE * NDOF) + j * NDOF + bq] = K[a][bq];
        } else {
          if (doLift) {
            for (int a = 0; a < NDOF; ++a) {
              double s = 0.0;
              for (int bq = 0; bq < NDOF; ++bq) {
                const double dlt = L.app[j][bq];
                if (!((mj >> bq) & 1) && fabs(dlt) > 1.0e-10) s += K[a][bq] * dlt;
              }
              L.lift[i][j][a] = s;
            }
          }
          if (mi != 0 && mj != 0) {
            const int32_t co = v.colofs[(slot * NPE + i) * NPE + j];
            const int32_t rs = L.info[i][0], rl = L.info[i][1];
            int ra = 0;
            for (int a = 0; a < NDOF; ++a) {
              if (!((mi >> a) & 1)) continue;
              add_row_segment<NDOF, double>(v.vals + (int64_t)rs + (int64_t)ra * rl + co, K[a], mj);
              ++ra;
            }
          }
        }
      } else {
        float K[NDOF][NDOF];
        accumulate_block<ET, float>(L, i, j, p, K);
        if (probe) {
          for (int a = 0; a < NDOF; ++a)
            for (int bq = 0; bq < NDOF; ++bq)
              v.probeK[(i * NDOF + a) * (NPE * NDOF) + j * NDOF + bq] = (double)K[a][bq];
        } else if (mi != 0 && mj != 0) {
          const int32_t co = v.colofs[(slot * NPE + i) * NPE + j];
          const int32_t rs = L.info[i][0], rl = L.info[i][1];
          int ra = 0;
          for (int a = 0; a < NDOF; ++a) {
            if (!((mi >> a) & 1)) continue;
            add_row_segment<NDOF, float>(v.vals32 + (int64_t)rs + (int64_t)ra * rl + co, K[a], mj);
            ++ra;
          }
        }
      }
    }
    __syncthreads();
  }

  // P6: residual of node i summed over Gauss points, lift, RHS scatter
  if (active && l < NPE) {
    const int i = l;
    const int mi = L.info[i][2];
    double F[NDOF];
    for (int a = 0; a < NDOF; ++a) {
      double s = 0.0;
      for (int gp = 0; gp < NGP; ++gp) s += L.Fk[gp][i][a];
      F[a] = s;
    }
    if (probe) {
      for (int a = 0; a < NDOF; ++a) v.probeF[i * NDOF + a] = F[a];
    } else {
      if (MODE == 0 && doLift) {
        for (int a = 0; a < NDOF; ++a)
          for (int j = 0; j < NPE; ++j) F[a] -= L.lift[i][j][a];
      }
      if (mi != 0) {
        const int32_t r0 = L.info[i][3];
        int ra = 0;
        for (int a = 0; a < NDOF; ++a) {
          if (!((mi >> a) & 1)) continue;
          v.rhs[r0 + ra] += F[a];
          ++ra;
        }
      }
    }
  }
  if constexpr (!PERSIST) break;  // the grid covers every group
  __syncthreads();                // the team's LDS is reused by the next group
  }
}

// Orders the LDS traffic of one wavefront (the whole workgroup of the pipelined kernel):
// the compiler may not move LDS accesses across it, and it waits for the issued ones.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Hexa8 element loop of one colour class, pipelined: one wavefront per workgroup walks
// the slots blockIdx.x, blockIdx.x + gridDim.x, ... of the class (the grid is at most
// what the GPU holds at once).  Lane l is node pair (i, j) = (l / 8, l % 8) of the
// tangent, Gauss point / node (gp, k) = (l / 8, l % 8) of the geometry and residual
// phases, and loads quantity q = l / 8 of node l % 8 of the next element:
//   q = 0 coordinates, 1 solnPrev, 2 solnCur, 3 solnDotCur, 4 solnApplied (lift only),
//   5 node info {rowStart, rowlen, free mask, first row}.
// While one element is computed, the next element's connectivity, element -> nonzero
// offsets and length scale are in flight (issued before the geometry phase), then its
// nodal data (issued before the tangent phase); the CSR and RHS values an element adds
// to are loaded before its tangent is accumulated (no other element of the class owns
// them).  The residual sums over Gauss points and the lift sums over neighbour nodes are
// butterfly reductions across lanes.
// Register budget per mode: FP64 tangent (0), two wavefronts per SIMD, the tangent
// phase holding the 4 x 4 block, the block values it adds to (loaded before the
// accumulation) and two Gauss points' worth of LDS operands; FP32 tangent (2), three
// wavefronts per SIMD, one Gauss point at a time and the block values loaded after the
// accumulation; residual only (1), three wavefronts per SIMD without the tangent factors.
template <int MODE>
struct HexCfg {
  static constexpr int occ = MODE == 0 ? 2 : 3;
  static constexpr int unroll = MODE == 2 ? 1 : 2;  // Gauss points of the tangent loop unrolled together
  static constexpr bool late_old = MODE == 2;
};

template <int MODE>
__global__ __launch_bounds__(64, HexCfg<MODE>::occ) void k_assemble_hex8(AsmView v, ElemParams p,
                                                      const int32_t* __restrict__ elems, int64_t n,
                                                      int lift) {
  using T = typename std::conditional<MODE == 2, float, double>::type;
  constexpr int NPE = 8, NDOF = 4;
  typedef T vec_t __attribute__((ext_vector_type(4), aligned(sizeof(T))));
  __shared__ HexLds L;
  const int l0 = threadIdx.x;
  const bool doLift = (MODE == 0) && (lift != 0);
  int nn = l0 & 7, q = l0 >> 3;
  int64_t s = blockIdx.x;
  if (s >= n) return;

  // staging registers of the element to come
  int32_t co = MODE != 1 ? v.blkofs[s * 64 + l0] : -1;
  int32_t eid = elems[s];
  double h = v.charlen[eid];
  double mu = v.elem_mu ? v.elem_mu[eid] : p.mu;
  double rv[4] = {0.0, 0.0, 0.0, 0.0};
  int4 ri = {0, 0, 0, 0};
  int cnt = 0;
  auto issue_node = [&](int32_t node) {
    const double* src = nullptr;
    cnt = 0;
    if (q == 0) { src = v.coords + (int64_t)node * 3; cnt = 3; }
    else if (q <= 3) { src = (q == 1 ? v.solnPrev : q == 2 ? v.solnCur : v.solnDotCur) + (int64_t)node * 4; cnt = 4; }
    else if (q == 4 && doLift) { src = v.solnApplied + (int64_t)node * 4; cnt = 4; }
#pragma unroll
    for (int c = 0; c < 4; ++c) rv[c] = c < cnt ? src[c] : 0.0;
    if (q == 5) ri = reinterpret_cast<const int4*>(v.nodeInfo)[node];
  };
  issue_node(v.sconn[s * NPE + nn]);

  for (;;) {
    // The lane index is re-derived in every iteration (through an empty asm the compiler
    // cannot see through), so that per-lane LDS addresses of the phases are recomputed
    // instead of being hoisted out of the loop and held in registers across it.
    int l = l0;
    asm volatile("" : "+v"(l));
    nn = l & 7;
    q = l >> 3;
    const int i = l >> 3, j = l & 7;
    // stage the element's nodal data
    if (q < 5) {
      double* dst = q == 0 ? &L.X[nn][0] : q <= 3 ? &L.S[q - 1][nn][0] : &L.app[nn][0];
      const int w = q == 0 ? 3 : 4;  // app is zero without lift
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c < w) dst[c] = rv[c];
    } else if (q == 5) {
      L.info[nn][0] = ri.x; L.info[nn][1] = ri.y; L.info[nn][2] = ri.z; L.info[nn][3] = ri.w;
    }
    const int32_t co_cur = co;
    ElemParams pe = p;
    pe.mu = mu;
    const double h_cur = h;
    wave_sync();
    // slot-ordered data of the next element
    const int64_t sn = s + gridDim.x;
    const bool more = sn < n;
    int32_t nid_n = 0;
    if (more) {
      nid_n = v.sconn[sn * NPE + nn];
      if (MODE != 1) co = v.blkofs[sn * 64 + l0];
      eid = elems[sn];
    }
    phase_geometry<HEXA8>(L, l);
    wave_sync();
    phase_interp<HEXA8>(L, l);
    wave_sync();
    double Fk[4];
    phase_gauss<HEXA8, MODE != 1>(L, l, h_cur, pe, Fk);
    // nodal data and length scale of the next element
    if (more) {
      issue_node(nid_n);
      h = v.charlen[eid];
      if (v.elem_mu) mu = v.elem_mu[eid];
    }
    wave_sync();

    const int mi = L.info[i][2], mj = L.info[j][2];
    // the RHS entries of node i (lanes j == 0) and, below, the CSR segments of block (i, j)
    const bool rhs_lane = (j == 0) && (mi != 0);
    const int32_t r0 = L.info[i][3];
    double rold[4] = {0.0, 0.0, 0.0, 0.0};
    if (rhs_lane) {
      int ra = 0;
#pragma unroll
      for (int a = 0; a < NDOF; ++a)
        if ((mi >> a) & 1) rold[a] = v.rhs[r0 + ra++];
    }
    double lsum[4] = {0.0, 0.0, 0.0, 0.0};
    if constexpr (MODE != 1) {
      // block (i, j) of the node-pair block form: 16 values, one aligned run, loaded
      // before the tangent is accumulated (no other element of the class adds to it)
      typedef T blk_t __attribute__((ext_vector_type(4)));
      T* kbv = MODE == 2 ? reinterpret_cast<T*>(v.kb32) : reinterpret_cast<T*>(v.kb);
      // (bit 30 of the map: this element is the block's first, and stores it whole)
      const bool blk = mi != 0 && mj != 0;
      const bool first = (co_cur >> 30) & 1;
      blk_t* bq = reinterpret_cast<blk_t*>(kbv + (int64_t)(blk ? (co_cur & ((1 << 30) - 1)) : 0) * 16);
      blk_t old[NDOF];
      if constexpr (!HexCfg<MODE>::late_old) {
#pragma unroll
        for (int a = 0; a < NDOF; ++a) old[a] = (blk && !first) ? bq[a] : blk_t{(T)0, (T)0, (T)0, (T)0};
      }
      T K[NDOF][NDOF];
      accumulate_block<HEXA8, T, HexCfg<MODE>::unroll>(L, i, j, pe, K);
      if constexpr (HexCfg<MODE>::late_old) {
#pragma unroll
        for (int a = 0; a < NDOF; ++a) old[a] = (blk && !first) ? bq[a] : blk_t{(T)0, (T)0, (T)0, (T)0};
      }
      if (doLift) {
#pragma unroll
        for (int a = 0; a < NDOF; ++a) {
          double sl = 0.0;
#pragma unroll
          for (int b = 0; b < NDOF; ++b) {
            const double dlt = L.app[j][b];
            if (!((mj >> b) & 1) && fabs(dlt) > 1.0e-10) sl += (double)K[a][b] * dlt;
          }
          lsum[a] = sl;
        }
      }
      if (blk) {
        if (mi == 15 && mj == 15) {
#pragma unroll
          for (int a = 0; a < NDOF; ++a) {
            blk_t r = old[a];
            r.x += K[a][0]; r.y += K[a][1]; r.z += K[a][2]; r.w += K[a][3];
            bq[a] = r;
          }
        } else {
          // free rows and columns packed to the front of the block
          T nv[NDOF][NDOF];
#pragma unroll
          for (int a = 0; a < NDOF; ++a) {
            nv[a][0] = old[a].x; nv[a][1] = old[a].y; nv[a][2] = old[a].z; nv[a][3] = old[a].w;
          }
          int ra = 0;
#pragma unroll
          for (int a = 0; a < NDOF; ++a) {
            if (!((mi >> a) & 1)) continue;
            int rb = 0;
#pragma unroll
            for (int b = 0; b < NDOF; ++b) {
              if (!((mj >> b) & 1)) continue;
              // rows and columns are only ever packed towards the front: ra <= a, rb <= b
#pragma unroll
              for (int x = 0; x < NDOF; ++x)
#pragma unroll
                for (int y = 0; y < NDOF; ++y)
                  if (x == ra && y == rb) nv[x][y] += K[a][b];
              ++rb;
            }
            ++ra;
          }
#pragma unroll
          for (int a = 0; a < NDOF; ++a) bq[a] = blk_t{nv[a][0], nv[a][1], nv[a][2], nv[a][3]};
        }
      }
    }
    // residual: sum over Gauss points (lane bits 3-5), then node i's sum to lane 8i
#pragma unroll
    for (int a = 0; a < NDOF; ++a) {
      double f = Fk[a];
      f += __shfl_xor(f, 8);
      f += __shfl_xor(f, 16);
      f += __shfl_xor(f, 32);
      Fk[a] = __shfl(f, i);
    }
    if (doLift) {
#pragma unroll
      for (int a = 0; a < NDOF; ++a) {
        double t = lsum[a];
        t += __shfl_xor(t, 1);
        t += __shfl_xor(t, 2);
        t += __shfl_xor(t, 4);
        Fk[a] -= t;
      }
    }
    if (rhs_lane) {
      int ra = 0;
#pragma unroll
      for (int a = 0; a < NDOF; ++a)
        if ((mi >> a) & 1) { v.rhs[r0 + ra] = rold[a] + Fk[a]; ++ra; }
    }
    wave_sync();
    if (!more) break;
    s = sn;
  }
}

template <int ET>
static hipError_t prepare_t(hipStream_t s, int64_t nElem, const double* coords, const int32_t* conn,
                            double* charlen, double* vol, int64_t* bad) {
  if (nElem == 0) return hipSuccess;
  const int64_t grid = (nElem + 255) / 256;
  hipLaunchKernelGGL((k_prepare<ET>), dim3((unsigned)grid), dim3(256), 0, s, nElem, coords, conn,
                     charlen, vol, bad);
  return hipGetLastError();
}

hipError_t launch_prepare(int et, hipStream_t s, int64_t nElem, const double* coords,
                          const int32_t* conn, double* charlen, double* vol, int64_t* bad) {
  switch (et) {
    case TRIA3: return prepare_t<TRIA3>(s, nElem, coords, conn, charlen, vol, bad);
    case QUAD4: return prepare_t<QUAD4>(s, nElem, coords, conn, charlen, vol, bad);
    case TETRA4: return prepare_t<TETRA4>(s, nElem, coords, conn, charlen, vol, bad);
    default: return prepare_t<HEXA8>(s, nElem, coords, conn, charlen, vol, bad);
  }
}

// One launch of variant (WGS, PERSIST) of the generic element loop.
template <int ET, int MODE, int WGS, bool PERSIST>
static hipError_t assemble_v(hipStream_t s, const AsmView& v, const ElemParams& p, const int32_t* elems, int64_t n,
                             int lift, int64_t probe_e) {
  constexpr int TEAMS = WGS / Traits<ET>::TEAM;
  const int64_t cnt = probe_e >= 0 ? 1 : n;
  if (cnt == 0) return hipSuccess;
  int64_t grid = (cnt + TEAMS - 1) / TEAMS;
  if (PERSIST) {
    static int64_t resident[16];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) dev = 0;
    if (resident[dev] == 0) {
      int cus = 0, per = 0;
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_assemble<ET, MODE, WGS, PERSIST>, WGS, 0) !=
              hipSuccess || per < 1)
        per = 1;
      resident[dev] = (int64_t)(cus > 0 ? cus : 256) * per;
    }
    grid = std::min(grid, resident[dev]);
  }
  hipLaunchKernelGGL((k_assemble<ET, MODE, WGS, PERSIST>), dim3((unsigned)grid), dim3(WGS), 0, s, v, p, elems, n,
                     lift, probe_e);
  return hipGetLastError();
}

template <int ET, int MODE>
static hipError_t assemble_m(hipStream_t s, const AsmView& v, const ElemParams& p, const int32_t* elems, int64_t n,
                             int lift, int64_t probe_e, int variant) {
  switch (variant) {
    case 1: return assemble_v<ET, MODE, 128, false>(s, v, p, elems, n, lift, probe_e);
    case 2: return assemble_v<ET, MODE, 64, false>(s, v, p, elems, n, lift, probe_e);
    case 3: return assemble_v<ET, MODE, 256, true>(s, v, p, elems, n, lift, probe_e);
    case 4: return assemble_v<ET, MODE, 128, true>(s, v, p, elems, n, lift, probe_e);
    case 5: return assemble_v<ET, MODE, 64, true>(s, v, p, elems, n, lift, probe_e);
    default: return assemble_v<ET, MODE, 256, false>(s, v, p, elems, n, lift, probe_e);
  }
}

template <int ET>
static hipError_t assemble_t(hipStream_t s, const AsmView& v, const ElemParams& p,
                             const int32_t* elems, int64_t n, int mode, int lift, int64_t probe_e, int variant) {
  if (probe_e >= 0) variant = 0;
  if (mode == 0) return assemble_m<ET, 0>(s, v, p, elems, n, lift, probe_e, variant);
  if (mode == 1) return assemble_m<ET, 1>(s, v, p, elems, n, lift, probe_e, variant);
  return assemble_m<ET, 2>(s, v, p, elems, n, lift, probe_e, variant);
}

// Workgroups of the pipelined Hexa8 kernel the current device holds at once (each walks
// its share of a colour class).
template <int MODE>
static hipError_t hex8_launch(hipStream_t s, const AsmView& v, const ElemParams& p, const int32_t* elems,
                              int64_t n, int lift) {
  static int64_t resident[16];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) dev = 0;
  if (resident[dev] == 0) {
    int cus = 0, per = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_assemble_hex8<MODE>, 64, 0) != hipSuccess || per < 1)
      per = 4 * HexCfg<MODE>::occ;
    resident[dev] = (int64_t)(cus > 0 ? cus : 256) * per;
  }
  const unsigned grid = (unsigned)std::min<int64_t>(n, resident[dev]);
  hipLaunchKernelGGL((k_assemble_hex8<MODE>), dim3(grid), dim3(64), 0, s, v, p, elems, n, lift);
  return hipGetLastError();
}

hipError_t launch_assemble(int et, hipStream_t s, const AsmView& v, const ElemParams& p,
                           const int32_t* elems, int64_t n, int mode, int lift, int64_t probe_e, int variant) {
  if (et == HEXA8 && probe_e < 0) {
    if (n == 0) return hipSuccess;
    if (mode == 0) return hex8_launch<0>(s, v, p, elems, n, lift);
    if (mode == 1) return hex8_launch<1>(s, v, p, elems, n, lift);
    return hex8_launch<2>(s, v, p, elems, n, lift);
  }
  switch (et) {
    case TRIA3: return assemble_t<TRIA3>(s, v, p, elems, n, mode, lift, probe_e, variant);
    case QUAD4: return assemble_t<QUAD4>(s, v, p, elems, n, mode, lift, probe_e, variant);
    case TETRA4: return assemble_t<TETRA4>(s, v, p, elems, n, mode, lift, probe_e, variant);
    default: return assemble_t<HEXA8>(s, v, p, elems, n, mode, lift, probe_e, 0);
  }
}

}  // namespace sfem
