// sfem_elements.hip -- element geometry and the stabilised Navier-Stokes element
// kernels (residual Fe and tangent Ke) with scatter-add into the device CSR.
//
// Tria3 / Quad4 / Tetra4 (and the per-element probe): one "team" of TEAM lanes per
// element (16, a quarter wavefront; 64 for the Hexa8 probe).  Lane (i, j) of the team owns the ndof x ndof
// block of node pair (i, j) of Ke and keeps it in registers across all Gauss points;
// the per-Gauss-point data (shape gradients, interpolated state, stabilisation
// parameters) is computed cooperatively by lane (gp, k) and staged in LDS.
// Hexa8 colour loops: the batched kernel k_hex8_batch (four elements per wavefront
// iteration, Gauss-point lanes for the residual, node-pair lanes for the tangent).
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

// Generic element loop (all element types; the Hexa8 colour loops use the batched
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

// Orders the LDS traffic of one wavefront (the whole workgroup of the batched kernel):
// the compiler may not move LDS accesses across it, and it waits for the issued ones.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Batched Hexa8 element loop: four elements of a colour class per wavefront iteration.
//
// Phase G, lane (e, g, h) = (l / 16, (l / 2) % 8, l % 2): Gauss point g of element e.
// Both lanes of a pair form the Jacobian at g; the shape gradients of the eight nodes
// go to LDS; both lanes interpolate the state (31 sums over the nodes, in node order)
// and form the stabilisation parameter; lane h then forms the residual of nodes h, h + 2,
// h + 4, h + 6 at g.  The residual is summed over Gauss points through LDS by lane
// (e, k, c) = (l / 16, (l / 2) % 8, l % 2), which ends with components c and c + 2 of
// node k.  For the tangent, the per-(Gauss point) record and, for the FP32 tangent, the
// per-(Gauss point, node) records follow.
// Phase K, lane (i, j) = (l / 8, l % 8): block (i, j) of each of the four elements in
// turn, summed over the Gauss points from the records, added to the node-pair block form
// (the element whose map entry carries bit 30 stores the block instead).
// The next batch's connectivity, block offsets and element ids are loaded while phase G
// runs, its nodal data while the residual is summed and phase K runs.
// ---------------------------------------------------------------------------
constexpr int kHexEB = 4;  // elements per batch

template <int MODE>
struct HexBCfg {
  using T = typename std::conditional<MODE == 2, float, double>::type;  // tangent type
  static constexpr bool KM = MODE != 1;
  // Per (element, Gauss point) coefficients of the tangent, with the weight w folded in
  // (A = af rho gradPrev, s = -1, the Hexa8 sign):
  //   0-8 w A, 9 w rho td8, 10 w mu af, 11 w af, 12 w s af, 13 w s tauP, 14 w s tauP af,
  //   15 w; the FP64 tangent adds 16-18 af rho uPrev and 19 s tauP.
  // Rows are padded so that consecutive Gauss points start in different LDS banks.
  static constexpr int CW = MODE == 2 ? 20 : 22;
  static constexpr int occ = 2;
};

// Per-(element, Gauss point) rows of 8 nodes x 4 values, padded by 16 bytes so that the
// eight Gauss points of an element start in different LDS banks.
constexpr int kHexRow = 34;   // doubles: dN/dx, dN/dy, dN/dz, N of each node
constexpr int kHexRow32 = 68; // floats: the FP32 tangent's 8 values of each node

template <int MODE>
struct HexBLds {
  using Cfg = HexBCfg<MODE>;
  using T = typename Cfg::T;
  union {
    struct {
      double X[kHexEB][8][4];     // coordinates
      double S[3][kHexEB][8][4];  // solnPrev, solnCur, solnDotCur
    } nd;
    // residual (element, Gauss point, node), components 0, 2, 1, 3 (the residual-only
    // loop keeps it over the shape gradients instead)
    double R[MODE == 1 ? 1 : kHexEB][8][kHexRow];
    struct {
      T cg[Cfg::KM ? kHexEB : 1][8][Cfg::CW];
      double lift[MODE == 0 ? kHexEB : 1][8][4];  // lift sums of node i of each element
    } k;
  } a;
  union {
    // per (element, Gauss point, node): dN/dx, dN/dy, dN/dz, N (the FP64 tangent's record)
    double gd[kHexEB][8][kHexRow];
    // FP32 tangent, per (element, Gauss point, node): dN/dx, dN/dy, dN/dz, N,
    // g(0..2) = s tauP sum_a w dN_a A_a., u = af rho uPrev.dN; each node's 32 bytes are
    // written over its 32 bytes of gd once they are read
    float rec[kHexEB][8][kHexRow32];
  } b;
  double app[MODE == 0 ? kHexEB : 1][8][4];  // prescribed increments (iteration-0 lift)
  int32_t info[kHexEB][8][2];                 // free-DOF mask, first free row
};

// The residual rows: in the residual-only loop they overlay the shape gradients (each
// lane reads a node's gradients before writing that node's residual at the same place).
template <int MODE>
__device__ inline double* hexb_res_row(HexBLds<MODE>& L, int e, int g) {
  if constexpr (MODE == 1) return L.b.gd[e][g];
  else return L.a.R[e][g];
}

// Adds lane (i, j)'s block K to block form entry co (free rows and columns packed to the
// front); old = its previous values (zero when this element stores the block).
template <typename T>
__device__ inline void hex_block_store(T* __restrict__ kbv, int32_t co, int mi, int mj, const T (&K)[4][4],
                                       const T (&old)[4][4]) {
  typedef T blk_t __attribute__((ext_vector_type(4)));
  blk_t* bq = reinterpret_cast<blk_t*>(kbv + (int64_t)(co & ((1 << 30) - 1)) * 16);
  if (mi == 15 && mj == 15) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
      bq[a] = blk_t{old[a][0] + K[a][0], old[a][1] + K[a][1], old[a][2] + K[a][2], old[a][3] + K[a][3]};
    return;
  }
  T nv[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) nv[a][b] = old[a][b];
  int ra = 0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if (!((mi >> a) & 1)) continue;
    int rb = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (!((mj >> b) & 1)) continue;
      // rows and columns are only ever packed towards the front: ra <= a, rb <= b
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
          if (x == ra && y == rb) nv[x][y] += K[a][b];
      ++rb;
    }
    ++ra;
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) bq[a] = blk_t{nv[a][0], nv[a][1], nv[a][2], nv[a][3]};
}

// Per-Gauss-point results of phase G kept in registers for the records.
struct HexGp {
  double w, tauP, A[3][3], vp[3];
};

// Phase G of lane (e, g, h), up to the residual written to the residual rows.
template <int MODE>
__device__ inline void hexb_gauss(HexBLds<MODE>& L, int e, int g, int h, double hlen, double mu,
                                  const ElemParams& p, HexGp& out) {
  constexpr double SGN = Traits<HEXA8>::SGN, CI = Traits<HEXA8>::CI;
  typedef double d2 __attribute__((ext_vector_type(2)));
  double xi[3];
  xi[0] = (g & 1) ? kG2 : -kG2;
  xi[1] = (g & 2) ? kG2 : -kG2;
  xi[2] = (g & 4) ? kG2 : -kG2;
  double w;
  {
    double N[8], d1[8], d2v[8], d3[8];
    shape<HEXA8>(xi, N, d1, d2v, d3);
    // Jacobian B(r, c) = sum_n x_c(n) dN_n/dxi_r and its inverse (computeBasisFunctions3D)
    double B[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      const double* x = L.a.nd.X[e][n];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        B[0][c] += x[c] * d1[n];
        B[1][c] += x[c] * d2v[n];
        B[2][c] += x[c] * d3[n];
      }
    }
    const double c00 = B[1][1] * B[2][2] - B[1][2] * B[2][1];
    const double c01 = B[1][2] * B[2][0] - B[1][0] * B[2][2];
    const double c02 = B[1][0] * B[2][1] - B[1][1] * B[2][0];
    const double Jac = B[0][0] * c00 + B[0][1] * c01 + B[0][2] * c02;
    const double id = 1.0 / Jac;
    const double i00 = c00 * id, i01 = (B[0][2] * B[2][1] - B[0][1] * B[2][2]) * id,
                 i02 = (B[0][1] * B[1][2] - B[0][2] * B[1][1]) * id;
    const double i10 = c01 * id, i11 = (B[0][0] * B[2][2] - B[0][2] * B[2][0]) * id,
                 i12 = (B[0][2] * B[1][0] - B[0][0] * B[1][2]) * id;
    const double i20 = c02 * id, i21 = (B[0][1] * B[2][0] - B[0][0] * B[2][1]) * id,
                 i22 = (B[0][0] * B[1][1] - B[0][1] * B[1][0]) * id;
    w = 1.0 * Jac;  // weight 1 (UtilitiesQuadrature.cpp:285-297)
    // shape gradients of nodes 4h .. 4h + 3 to LDS (the pair's other lane writes the rest)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if ((k >> 2) != h) continue;
      const double g0 = d1[k] * i00 + d2v[k] * i01 + d3[k] * i02;
      const double g1 = d1[k] * i10 + d2v[k] * i11 + d3[k] * i12;
      const double g2 = d1[k] * i20 + d2v[k] * i21 + d3[k] * i22;
      d2* dst = reinterpret_cast<d2*>(&L.b.gd[e][g][4 * k]);
      dst[0] = d2{g0, g1};
      dst[1] = d2{g2, N[k]};
    }
  }
  wave_sync();
  // interpolated state (Hexa8Node.cpp:828-889), summed in node order
  double vp[3] = {0, 0, 0}, gpv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double vel[3] = {0, 0, 0}, gr[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double pres = 0.0, dp[3] = {0, 0, 0}, vd[3] = {0, 0, 0};
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const double* sp = L.a.nd.S[0][e][n];
    const double* sc = L.a.nd.S[1][e][n];
    const double* sd = L.a.nd.S[2][e][n];
    const double* gn = &L.b.gd[e][g][4 * n];
    const double Nn = gn[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      vp[a] += sp[a] * Nn;
      vel[a] += sc[a] * Nn;
      vd[a] += sd[a] * Nn;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        gpv[a][c] += sp[a] * gn[c];
        gr[a][c] += sc[a] * gn[c];
      }
    }
    pres += sc[3] * Nn;
#pragma unroll
    for (int c = 0; c < 3; ++c) dp[c] += sc[3] * gn[c];
    __builtin_amdgcn_sched_barrier(0);  // one node's LDS operands in flight at a time
  }
  // stabilisation (evaluateStabParams_algo3, stabilisationRoutines.h:128-143, G = 4/h^2 I)
  const double rho = p.rho;
  double res2[3], rStab[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double s1 = gpv[a][0] * vel[0], s2 = gr[a][0] * vp[0], s3 = gpv[a][0] * vp[0];
#pragma unroll
    for (int c = 1; c < 3; ++c) {
      s1 += gpv[a][c] * vel[c];
      s2 += gr[a][c] * vp[c];
      s3 += gpv[a][c] * vp[c];
    }
    res2[a] = rho * (vd[a] + (s1 + s2 - s3));
    rStab[a] = res2[a] + dp[a];
  }
  const double fact = 4.0 / hlen / hlen;
  double uGu = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) uGu += vp[a] * (fact * vp[a]);
  const double G2 = 3 * (fact * fact);
  double tau0 = 1.0 / sqrt(uGu + CI * (mu * mu / rho / rho) * G2);
  tau0 = tau0 / rho;
  const double tauP = tau0 * p.stab[1];
  // residual (Hexa8Node.cpp:1049-1060): component a of node k is -w (N_k X_a + dN_k . Y_a)
  // with X_a = res2_a, Y_a = mu grad_a - p e_a for a < 3, X_3 = s trace, Y_3 = s tauP rStab
  double Y[4][3], X[4];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    X[a] = res2[a];
#pragma unroll
    for (int c = 0; c < 3; ++c) Y[a][c] = mu * gr[a][c] - (a == c ? pres : 0.0);
  }
  X[3] = SGN * (gr[0][0] + gr[1][1] + gr[2][2]);
#pragma unroll
  for (int c = 0; c < 3; ++c) Y[3][c] = SGN * tauP * rStab[c];
  out.w = w;
  out.tauP = tauP;
  const double afr = p.af * rho;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    out.vp[a] = vp[a];
#pragma unroll
    for (int c = 0; c < 3; ++c) out.A[a][c] = afr * gpv[a][c];
  }
  wave_sync();  // the interpolation's reads of the nodal data are done (R overlays it)
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    const int k = 2 * kk + h;
    const d2* src = reinterpret_cast<const d2*>(&L.b.gd[e][g][4 * k]);
    const d2 ga = src[0], gb = src[1];
    double F[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) F[a] = -w * (gb.y * X[a] + ga.x * Y[a][0] + ga.y * Y[a][1] + gb.x * Y[a][2]);
    d2* dst = reinterpret_cast<d2*>(hexb_res_row<MODE>(L, e, g) + 4 * k);
    dst[0] = d2{F[0], F[2]};
    dst[1] = d2{F[1], F[3]};
  }
}

// Phase G records of lane (e, g, h) for the tangent (after the residual sums).
template <int MODE>
__device__ inline void hexb_records(HexBLds<MODE>& L, int e, int g, int h, double mu, const ElemParams& p,
                                    const HexGp& q) {
  using T = typename HexBCfg<MODE>::T;
  constexpr double SGN = Traits<HEXA8>::SGN;
  const double stp = SGN * q.tauP, afr = p.af * p.rho;
  if constexpr (MODE == 2) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int k = 2 * kk + h;
      const d2* src = reinterpret_cast<const d2*>(&L.b.gd[e][g][4 * k]);
      const d2 ga = src[0], gb = src[1];
      const double dk[3] = {ga.x, ga.y, gb.x};
      double gk[3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        gk[c] = stp * ((dk[0] * q.w) * q.A[0][c] + (dk[1] * q.w) * q.A[1][c] + (dk[2] * q.w) * q.A[2][c]);
      const double ud = q.vp[0] * dk[0] + q.vp[1] * dk[1] + q.vp[2] * dk[2];
      f4* r = reinterpret_cast<f4*>(&L.b.rec[e][g][8 * k]);
      r[0] = f4{(float)dk[0], (float)dk[1], (float)dk[2], (float)gb.y};
      r[1] = f4{(float)gk[0], (float)gk[1], (float)gk[2], (float)(afr * ud)};
    }
  }
  if (h == 0) {
    T* c = L.a.k.cg[e][g];
    const double w = q.w;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) c[3 * a + b] = (T)(w * q.A[a][b]);
    c[9] = (T)(w * (p.rho * p.acceFact));
    c[10] = (T)(w * (mu * p.af));
    c[11] = (T)(w * p.af);
    c[12] = (T)(w * (SGN * p.af));
    c[13] = (T)(w * stp);
    c[14] = (T)(w * stp * p.af);
    c[15] = (T)w;
    if constexpr (MODE == 0) {
#pragma unroll
      for (int b = 0; b < 3; ++b) c[16 + b] = afr * q.vp[b];
      c[19] = stp;
    }
  }
}

// Phase K: block (i, j) of element e summed over the Gauss points from the records.
// With p = Ni Nj, d = dNi.dNj and the weight folded into the coefficients c:
//   K_uu(a,b) += p wA_ab + delta_ab [w rho td8 p + w mu af d + w Ni u_j]
//   K_up(a)   -= (w af Nj) dNi_a
//   K_pu(b)   += (w s af Ni) dNj_b + Nj g_i(b) + (w s tauP d_j) dNi_b
//   K_pp      += w s tauP af d
// with d_j = u_j + rho td8 Nj (accumulate_block's terms, regrouped).
template <int MODE>
__device__ inline void hexb_block(const HexBLds<MODE>& L, int e, int i, int j, const ElemParams& p,
                                  typename HexBCfg<MODE>::T (&K)[4][4]) {
  using T = typename HexBCfg<MODE>::T;
  const T cst = (T)(p.rho * p.acceFact);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) K[a][b] = (T)0;
#pragma unroll
  for (int g = 0; g < 8; ++g) {
    const T* c = L.a.k.cg[e][g];
    T dNi[3], dNj[3], Ni, Nj, gi[3], uj;
    if constexpr (MODE == 2) {
      const float* ri = &L.b.rec[e][g][8 * i];
      const float* rj = &L.b.rec[e][g][8 * j];
#pragma unroll
      for (int x = 0; x < 3; ++x) { dNi[x] = ri[x]; dNj[x] = rj[x]; gi[x] = ri[4 + x]; }
      Ni = ri[3]; Nj = rj[3];
      uj = rj[7];
    } else {
      const double* ri = &L.b.gd[e][g][4 * i];
      const double* rj = &L.b.gd[e][g][4 * j];
#pragma unroll
      for (int x = 0; x < 3; ++x) { dNi[x] = ri[x]; dNj[x] = rj[x]; }
      Ni = ri[3]; Nj = rj[3];
#pragma unroll
      for (int b = 0; b < 3; ++b) gi[b] = c[19] * (dNi[0] * c[b] + dNi[1] * c[3 + b] + dNi[2] * c[6 + b]);
      uj = c[16] * dNj[0] + c[17] * dNj[1] + c[18] * dNj[2];
    }
    const T pp = Ni * Nj;
    const T dd = dNi[0] * dNj[0] + dNi[1] * dNj[1] + dNi[2] * dNj[2];
    const T dj = uj + cst * Nj;
    const T s2 = c[9] * pp + c[10] * dd + (c[15] * Ni) * uj;
    const T c1 = c[11] * Nj, c2 = c[12] * Ni, c3 = c[13] * dj;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int b = 0; b < 3; ++b) K[a][b] += pp * c[3 * a + b];
      K[a][a] += s2;
      K[a][3] -= c1 * dNi[a];
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) K[3][b] += c2 * dNj[b] + Nj * gi[b] + c3 * dNi[b];
    K[3][3] += c[14] * dd;
    __builtin_amdgcn_sched_barrier(0);  // one Gauss point's LDS operands in flight at a time
  }
}

template <int MODE>
__global__ __launch_bounds__(64, HexBCfg<MODE>::occ) void k_hex8_batch(AsmView v, ElemParams p,
                                                                const int32_t* __restrict__ elems, int64_t n,
                                                                int lift) {
  using Cfg = HexBCfg<MODE>;
  using T = typename Cfg::T;
  __shared__ __align__(16) HexBLds<MODE> L;
  const int l0 = threadIdx.x;
  const bool doLift = (MODE == 0) && (lift != 0);
  const int64_t nb = (n + kHexEB - 1) / kHexEB;
  int64_t b = blockIdx.x;
  if (b >= nb) return;
  // lane (e, k, c) of staging and of the residual sums; e is also phase G's element
  const int se = l0 >> 4, sk = (l0 >> 1) & 7, sc = l0 & 1;

  // Staging registers of the batch to come.  Every lane issues the same loads, with no
  // branch around them (slots past the class end read its last element, whose results
  // are not stored), so the compiler can wait for each group of loads by count.
  int32_t node = 0, eid = 0;
  int32_t co[kHexEB];
  auto issue_ids = [&](int64_t bb) {
    const int64_t sl = min<int64_t>(bb * kHexEB + se, n - 1);
    node = v.sconn[sl * 8 + sk];
    eid = elems[sl];
#pragma unroll
    for (int q = 0; q < kHexEB; ++q) {
      const int64_t s2 = min<int64_t>(bb * kHexEB + q, n - 1);
      co[q] = Cfg::KM ? v.blkofs[s2 * 64 + l0] : 0;
    }
  };
  typedef double d4 __attribute__((ext_vector_type(4)));
  d4 rA, rB;
  double rX[3];
  int2 ri;
  double hlen, mu;
  auto issue_nodal = [&]() {
    const int64_t o = (int64_t)node * 4;
    // lane c = 0: solnPrev and (lift) solnApplied; c = 1: solnCur and solnDotCur
    rA = *reinterpret_cast<const d4*>((sc ? v.solnCur : v.solnPrev) + o);
    rB = *reinterpret_cast<const d4*>((sc == 0 && doLift ? v.solnApplied : v.solnDotCur) + o);
    const double* x = v.coords + (int64_t)node * 3;
    rX[0] = x[0]; rX[1] = x[1]; rX[2] = x[2];
    ri = reinterpret_cast<const int2*>(v.nodeInfo)[2 * (int64_t)node + 1];  // free mask, first row
    hlen = v.charlen[eid];
    mu = v.elem_mu ? v.elem_mu[eid] : p.mu;
  };
  issue_ids(b);
  issue_nodal();

  for (;;) {
    const int64_t slot0 = b * kHexEB;
    const int nv = (int)min<int64_t>(kHexEB, n - slot0);  // elements of this batch
    // stage the batch's nodal data
    if (sc == 0) {
      L.a.nd.X[se][sk][0] = rX[0]; L.a.nd.X[se][sk][1] = rX[1]; L.a.nd.X[se][sk][2] = rX[2];
      *reinterpret_cast<d4*>(L.a.nd.S[0][se][sk]) = rA;
      if (doLift) *reinterpret_cast<d4*>(L.app[se][sk]) = rB;
    } else {
      *reinterpret_cast<d4*>(L.a.nd.S[1][se][sk]) = rA;
      *reinterpret_cast<d4*>(L.a.nd.S[2][se][sk]) = rB;
      L.info[se][sk][0] = ri.x;
      L.info[se][sk][1] = ri.y;
    }
    const double hlen_cur = hlen, mu_cur = mu;
    int32_t co_cur[kHexEB];
#pragma unroll
    for (int q = 0; q < kHexEB; ++q) co_cur[q] = co[q];
    wave_sync();
    // connectivity, element ids and block offsets of the next batch
    const int64_t bn = b + gridDim.x;
    const bool more = bn < nb;
    issue_ids(more ? bn : b);
    // phase G.  The lane's Gauss point is re-derived through an empty asm in every
    // iteration, so that its shape values are recomputed here instead of being hoisted
    // out of the loop and held in registers across it.
    int lg = l0;
    asm volatile("" : "+v"(lg));
    HexGp gq;
    hexb_gauss<MODE>(L, se, (lg >> 1) & 7, sc, hlen_cur, mu_cur, p, gq);
    wave_sync();
    // residual summed over Gauss points: lane (e, k, c) ends with components c, c + 2 of node k
    double fsum[2];
    {
      typedef double d2 __attribute__((ext_vector_type(2)));
      d2 s = *reinterpret_cast<const d2*>(hexb_res_row<MODE>(L, se, 0) + 4 * sk + 2 * sc);
#pragma unroll
      for (int g = 1; g < 8; ++g) s += *reinterpret_cast<const d2*>(hexb_res_row<MODE>(L, se, g) + 4 * sk + 2 * sc);
      fsum[0] = s.x;
      fsum[1] = s.y;
    }
    // RHS entries of node k this lane adds to (loaded before the tangent phase)
    const int mk = L.info[se][sk][0], rk = L.info[se][sk][1];
    const bool own = se < nv;
    int row[2];
    double rold[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int a = sc + 2 * r;
      row[r] = (own && ((mk >> a) & 1)) ? rk + __popc(mk & ((1 << a) - 1)) : -1;
      rold[r] = v.rhs[row[r] >= 0 ? row[r] : 0];  // row 0 exists whenever this kernel runs
    }
    issue_nodal();  // (the last batch reloads its own data, unused)
    if constexpr (Cfg::KM) {
      T* kbv = MODE == 2 ? reinterpret_cast<T*>(v.kb32) : reinterpret_cast<T*>(v.kb);
      typedef T blk_t __attribute__((ext_vector_type(4)));
      // Block values an element adds to, loaded one element ahead (the first element's
      // while the records are formed).  Lanes that add nothing, and lanes whose element
      // stores the block (bit 30), read block 0 instead (always present, and one cached
      // line for all of them); the last element's reload past the batch end is not used.
      auto blk_load = [&](int e, blk_t (&o)[4]) {
        int lk = l0;
        asm volatile("" : "+v"(lk));
        const int32_t coe = e == 0 ? co_cur[0] : e == 1 ? co_cur[1] : e == 2 ? co_cur[2] : co_cur[3];
        const bool keep = L.info[e][lk >> 3][0] != 0 && L.info[e][lk & 7][0] != 0 && !((coe >> 30) & 1);
        const blk_t* bq = reinterpret_cast<const blk_t*>(kbv + (int64_t)(keep ? (coe & ((1 << 30) - 1)) : 0) * 16);
#pragma unroll
        for (int a = 0; a < 4; ++a) o[a] = bq[a];
      };
      // (the FP64 loop has no registers to spare for the look-ahead: it loads each
      // element's block values as it starts on the element)
      // The element loop is unrolled, the two look-ahead buffers used in turn (a copy
      // from one to the other at the loop's end would wait for the load just issued).
      constexpr bool ahead = MODE == 2;
      blk_t ob[2][4];
      if constexpr (ahead) blk_load(0, ob[0]);
      wave_sync();  // the residual sums have read the residual rows (the coefficients overlay them)
      hexb_records<MODE>(L, se, sk, sc, mu_cur, p, gq);
      wave_sync();
#pragma unroll
      for (int e = 0; e < kHexEB; ++e) {
        if (e >= nv) break;
        blk_t (&ocur)[4] = ob[e & 1];
        if constexpr (ahead) {
          if (e + 1 < kHexEB) blk_load(e + 1 < nv ? e + 1 : e, ob[(e + 1) & 1]);
        } else {
          blk_load(e, ocur);
        }
        // (re-derived per element, so that the per-lane LDS addresses are not hoisted)
        int lk = l0;
        asm volatile("" : "+v"(lk));
        const int i = lk >> 3, j = lk & 7;
        const int32_t coe = e == 0 ? co_cur[0] : e == 1 ? co_cur[1] : e == 2 ? co_cur[2] : co_cur[3];
        const int mi = L.info[e][i][0], mj = L.info[e][j][0];
        const bool blk = mi != 0 && mj != 0;
        const bool keep = blk && !((coe >> 30) & 1);  // bit 30: this element stores the block
        T K[4][4];
        hexb_block<MODE>(L, e, i, j, p, K);
        // the block is formed here, before the branches that use it (sunk into the store
        // branch, it left every Gauss point's LDS operands live at once)
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int bb = 0; bb < 4; ++bb) asm volatile("" : "+v"(K[a][bb]));
        T old[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          old[a][0] = keep ? ocur[a].x : (T)0; old[a][1] = keep ? ocur[a].y : (T)0;
          old[a][2] = keep ? ocur[a].z : (T)0; old[a][3] = keep ? ocur[a].w : (T)0;
        }
        if (doLift) {
          double ls[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            double sl = 0.0;
#pragma unroll
            for (int bb = 0; bb < 4; ++bb) {
              const double dlt = L.app[e][j][bb];
              if (!((mj >> bb) & 1) && fabs(dlt) > 1.0e-10) sl += (double)K[a][bb] * dlt;
            }
            sl += __shfl_xor(sl, 1);
            sl += __shfl_xor(sl, 2);
            sl += __shfl_xor(sl, 4);
            ls[a] = sl;
          }
          if (j == 0) {
#pragma unroll
            for (int a = 0; a < 4; ++a) L.a.k.lift[e][i][a] = ls[a];
          }
        }
        if (blk) hex_block_store<T>(kbv, coe, mi, mj, K, old);
      }
      if (doLift) {
        wave_sync();
        if (own) {
          fsum[0] -= L.a.k.lift[se][sk][sc];
          fsum[1] -= L.a.k.lift[se][sk][sc + 2];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
      if (row[r] >= 0) v.rhs[row[r]] = rold[r] + fsum[r];
    wave_sync();
    if (!more) break;
    b = bn;
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

// Workgroups of the batched Hexa8 kernel the current device holds at once.
template <int MODE>
static hipError_t hex8_batch_launch(hipStream_t s, const AsmView& v, const ElemParams& p, const int32_t* elems,
                                    int64_t n, int lift) {
  static int64_t resident[16];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) dev = 0;
  if (resident[dev] == 0) {
    int cus = 0, per = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_hex8_batch<MODE>, 64, 0) != hipSuccess || per < 1)
      per = 4 * HexBCfg<MODE>::occ;
    resident[dev] = (int64_t)(cus > 0 ? cus : 256) * per;
  }
  const int64_t nb = (n + kHexEB - 1) / kHexEB;
  const unsigned grid = (unsigned)std::min<int64_t>(nb, resident[dev]);
  hipLaunchKernelGGL((k_hex8_batch<MODE>), dim3(grid), dim3(64), 0, s, v, p, elems, n, lift);
  return hipGetLastError();
}

hipError_t launch_assemble(int et, hipStream_t s, const AsmView& v, const ElemParams& p,
                           const int32_t* elems, int64_t n, int mode, int lift, int64_t probe_e, int variant) {
  if (et == HEXA8 && probe_e < 0) {
    if (n == 0) return hipSuccess;
    if (mode == 0) return hex8_batch_launch<0>(s, v, p, elems, n, lift);
    if (mode == 1) return hex8_batch_launch<1>(s, v, p, elems, n, lift);
    return hex8_batch_launch<2>(s, v, p, elems, n, lift);
  }
  switch (et) {
    case TRIA3: return assemble_t<TRIA3>(s, v, p, elems, n, mode, lift, probe_e, variant);
    case QUAD4: return assemble_t<QUAD4>(s, v, p, elems, n, mode, lift, probe_e, variant);
    case TETRA4: return assemble_t<TETRA4>(s, v, p, elems, n, mode, lift, probe_e, variant);
    default: return assemble_t<HEXA8>(s, v, p, elems, n, mode, lift, probe_e, 0);
  }
}

}  // namespace sfem
