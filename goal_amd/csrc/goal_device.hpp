// Device-side math for the fused element kernels: quadrature, tet shape
// functions, 3x3 tensor helpers and the constitutive models.
//
// Every model returns, at one quadrature point, the generalized stress
//   sig = [ P (9, row-major P_iJ), r1, s3 (3) ]
// where the element residual is
//   R^u_{a,i} = sum_q (sum_J P_iJ dN_a/dX_J) w dv                 (src/goal_mresidual.cpp:26-32)
//   R^p_a     = sum_q (r1 N_a + sum_J s3_J dN_a/dX_J) w dv        (src/goal_presidual.cpp:53-77,
//                                                                  src/goal_stabilization.cpp:68-81)
// and, for the Jacobian, the hand-derived tangent D = d sig / d[F, p, grad p]
// (13 x 13).  The tangent is not produced by automatic differentiation: each
// column is the closed-form directional derivative of the reference formulas
// along one input direction, written out term by term below.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

#include "goal_kernels.hpp"

namespace goal {
namespace dev {

// ---------------------------------------------------------------- quadrature
// order 1: the centroid with weight 1/6 (apf integration order 1 on a tet,
// src/goal_assembly.cpp:78-80).  order 2: 4-point rule, barycentric
// (a,b,b,b) permutations, a = (5+3 sqrt5)/20, b = (5-sqrt5)/20, weight 1/24.
template <int NQ>
struct Quad;
template <>
struct Quad<1> {
  __host__ __device__ static inline void point(int, double xi[3], double& w) {
    xi[0] = xi[1] = xi[2] = 0.25;
    w = 1.0 / 6.0;
  }
  // xi[c] of point q, as point() sets it
  __host__ __device__ static inline double coord(int, int) { return 0.25; }
};
template <>
struct Quad<4> {
  __host__ __device__ static inline void point(int q, double xi[3], double& w) {
    const double a = 0.5854101966249685, b = 0.1381966011250105;
    xi[0] = q == 1 ? a : b;
    xi[1] = q == 2 ? a : b;
    xi[2] = q == 3 ? a : b;
    w = 1.0 / 24.0;
  }
  __host__ __device__ static inline double coord(int q, int c) {
    const double a = 0.5854101966249685, b = 0.1381966011250105;
    return q == c + 1 ? a : b;
  }
};

// ------------------------------------------------------------ tet geometry
// dX/dxi columns X1-X0, X2-X0, X3-X0; grad lambda_{c+1} = row c of its inverse.
__host__ __device__ inline double tet_gradlambda(const double X[4][3], double gl[4][3]) {
  double A[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[r][c] = X[c + 1][r] - X[0][r];
  double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
               A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
               A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  double id = 1.0 / det;
  // inverse rows
  gl[1][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * id;
  gl[1][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  gl[1][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  gl[2][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * id;
  gl[2][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  gl[2][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  gl[3][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * id;
  gl[3][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  gl[3][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
  for (int c = 0; c < 3; ++c) gl[0][c] = -(gl[1][c] + gl[2][c] + gl[3][c]);
  return det;
}

// Shape function a (value and X-gradient) at barycentric point lam.
// P1: N = lambda_a.  P2 (apf Lagrange quadratic): vertices lambda(2 lambda-1),
// edges (i,j) in order 01,12,20,03,13,23: 4 lambda_i lambda_j.
template <int NN>
__host__ __device__ inline void shape(int a, const double lam[4], const double gl[4][3], double& N, double dN[3]) {
  if (NN == 4 || a < 4) {
    if (NN == 4) {
      N = lam[a];
      for (int c = 0; c < 3; ++c) dN[c] = gl[a][c];
    } else {
      N = lam[a] * (2.0 * lam[a] - 1.0);
      double f = 4.0 * lam[a] - 1.0;
      for (int c = 0; c < 3; ++c) dN[c] = f * gl[a][c];
    }
  } else {
    const int ei[6] = {0, 1, 2, 0, 1, 2};
    const int ej[6] = {1, 2, 0, 3, 3, 3};
    int i = ei[a - 4], j = ej[a - 4];
    N = 4.0 * lam[i] * lam[j];
    for (int c = 0; c < 3; ++c) dN[c] = 4.0 * (lam[i] * gl[j][c] + lam[j] * gl[i][c]);
  }
}

// --------------------------------------------------------------- 3x3 tools
__host__ __device__ inline double det3(const double* F) {
  return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
         F[2] * (F[3] * F[7] - F[4] * F[6]);
}
__host__ __device__ inline void inv3(const double* F, double det, double* Fi) {
  double id = 1.0 / det;
  Fi[0] = (F[4] * F[8] - F[5] * F[7]) * id;
  Fi[1] = (F[2] * F[7] - F[1] * F[8]) * id;
  Fi[2] = (F[1] * F[5] - F[2] * F[4]) * id;
  Fi[3] = (F[5] * F[6] - F[3] * F[8]) * id;
  Fi[4] = (F[0] * F[8] - F[2] * F[6]) * id;
  Fi[5] = (F[2] * F[3] - F[0] * F[5]) * id;
  Fi[6] = (F[3] * F[7] - F[4] * F[6]) * id;
  Fi[7] = (F[1] * F[6] - F[0] * F[7]) * id;
  Fi[8] = (F[0] * F[4] - F[1] * F[3]) * id;
}

// P = J sigma F^{-T}:  P_iJ = J sum_m sigma_im Finv_Jm
__host__ __device__ inline void piola(const double* sig, const double* Fi, double J, double* P) {
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c)
      P[3 * i + c] = J * (sig[3 * i] * Fi[3 * c] + sig[3 * i + 1] * Fi[3 * c + 1] + sig[3 * i + 2] * Fi[3 * c + 2]);
}

// Mixed<T>::at_point: sigma_ii += p - tr(sigma)/3   (src/goal_mixed.cpp:33-46)
__host__ __device__ inline void mixed(double* sig, double p) {
  double pbar = (sig[0] + sig[4] + sig[8]) / 3.0;
  sig[0] += p - pbar;
  sig[4] += p - pbar;
  sig[8] += p - pbar;
}

// Matrix exponential by scaling and squaring of a degree-16 Taylor series
// (used for the plastic deformation update Fp = exp(dgam N) Fp_old).
__host__ __device__ inline void expm3(const double* A, double* E) {
  double nrm = 0.0;
  for (int k = 0; k < 9; ++k) nrm = fmax(nrm, fabs(A[k]));
  int s = 0;
  while (nrm * 3.0 > 0.5 && s < 60) {
    nrm *= 0.5;
    ++s;
  }
  double sc = ldexp(1.0, -s);
  double B[9], T[9], tmp[9];
  for (int k = 0; k < 9; ++k) {
    B[k] = A[k] * sc;
    E[k] = (k % 4 == 0) ? 1.0 : 0.0;
    T[k] = E[k];
  }
  for (int n = 1; n <= 16; ++n) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        tmp[3 * i + j] = (T[3 * i] * B[j] + T[3 * i + 1] * B[3 + j] + T[3 * i + 2] * B[6 + j]) / n;
    for (int k = 0; k < 9; ++k) {
      T[k] = tmp[k];
      E[k] += T[k];
    }
  }
  for (int r = 0; r < s; ++r) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        tmp[3 * i + j] = E[3 * i] * E[j] + E[3 * i + 1] * E[3 + j] + E[3 * i + 2] * E[6 + j];
    for (int k = 0; k < 9; ++k) E[k] = tmp[k];
  }
}

// Inputs shared by every model at one quadrature point.
struct QP {
  double F[9], Fi[9], J;  // deformation gradient, inverse, determinant
  double H[9];            // displacement gradient (small-strain model)
  double p, g[3];         // pressure and its X-gradient
};

// Output of one directional derivative.
// v0, v1 or v2 by a runtime index without indexing a register array (the
// directional derivatives below pick rows and columns of the point's 3 x 3
// arrays this way, so a loop over directions need not be unrolled)
__host__ __device__ inline double sel3(int i, double v0, double v1, double v2) {
  return i == 0 ? v0 : (i == 1 ? v1 : v2);
}

struct Dir {
  double dsig[9];  // Cauchy stress after Mixed (QoI derivatives)
  double dP[9];
  double dr1;
  double ds3[3];
};

// Stabilization flux s3_J = tau J sum_I Cinv_IJ g_I with C = F^T F
// (src/goal_stabilization.cpp:68-81); Cinv = Finv Finv^T.
struct Stab {
  double tau, Cinv[9], JCg[3];
  __host__ __device__ inline void init(const QP& q, double tau_) {
    tau = tau_;
    for (int I = 0; I < 3; ++I)
      for (int K = 0; K < 3; ++K)
        Cinv[3 * I + K] = q.Fi[3 * I] * q.Fi[3 * K] + q.Fi[3 * I + 1] * q.Fi[3 * K + 1] + q.Fi[3 * I + 2] * q.Fi[3 * K + 2];
    for (int j = 0; j < 3; ++j)
      JCg[j] = q.J * (Cinv[j] * q.g[0] + Cinv[3 + j] * q.g[1] + Cinv[6 + j] * q.g[2]);
  }
  __host__ __device__ inline void value(double* s3) const {
    for (int j = 0; j < 3; ++j) s3[j] = tau * JCg[j];
  }
  // direction dF = e_k e_L^T:
  //   dJ = J Finv_Lk,  dCinv_IJ = -Finv_Ik Cinv_LJ - Finv_Jk Cinv_IL
  __host__ __device__ inline void dir(const QP& q, int k, int L, double* ds3) const {
    double FiK[3], CL[3], CcL[3];  // column k of Finv, row L and column L of Cinv
    for (int m = 0; m < 3; ++m) {
      FiK[m] = sel3(k, q.Fi[3 * m], q.Fi[3 * m + 1], q.Fi[3 * m + 2]);
      CL[m] = sel3(L, Cinv[m], Cinv[3 + m], Cinv[6 + m]);
      CcL[m] = sel3(L, Cinv[3 * m], Cinv[3 * m + 1], Cinv[3 * m + 2]);
    }
    double dJJ = sel3(L, FiK[0], FiK[1], FiK[2]);  // Finv_Lk = dJ / J
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int I = 0; I < 3; ++I) {
        double dC = -FiK[I] * CL[j] - FiK[j] * CcL[I];
        acc += dC * q.g[I];
      }
      ds3[j] = tau * (dJJ * JCg[j] + q.J * acc);
    }
  }
  // direction d(grad p) = e_I: ds3_j = tau J Cinv_Ij
  __host__ __device__ inline void dir_g(const QP& q, int I, double* ds3) const {
    for (int j = 0; j < 3; ++j) ds3[j] = tau * q.J * Cinv[3 * I + j];
  }
};

// dP for a direction of F given sigma, dsigma:
//   dP_iJ = (dJ/J) P_iJ + J (dsigma F^{-T})_iJ - Finv_Jk P_iL
__host__ __device__ inline void dpiola(const QP& q, const double* P, const double* dsig, int k, int L, double* dP) {
  double FiK[3], PL[3];  // column k of Finv, column L of P
  for (int m = 0; m < 3; ++m) {
    FiK[m] = sel3(k, q.Fi[3 * m], q.Fi[3 * m + 1], q.Fi[3 * m + 2]);
    PL[m] = sel3(L, P[3 * m], P[3 * m + 1], P[3 * m + 2]);
  }
  double dJJ = sel3(L, FiK[0], FiK[1], FiK[2]);
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) {
      double t = dsig[3 * i] * q.Fi[3 * c] + dsig[3 * i + 1] * q.Fi[3 * c + 1] + dsig[3 * i + 2] * q.Fi[3 * c + 2];
      dP[3 * i + c] = dJJ * P[3 * i + c] + q.J * t - FiK[c] * PL[i];
    }
}

// ----------------------------------------------------------- neo-Hookean
// sigma = mu J^{-5/3} dev(F F^T) + 1/2 kappa (J - 1/J) I   (src/goal_neohookean.cpp:59-72)
// then Mixed; P = J sigma F^{-T}                          (src/goal_neohookean.cpp:79-86)
// r1 = p/kappa - 1/2 (J - 1/J)                             (src/goal_presidual.cpp:53-59)
struct NeoHookean {
  double mu, kappa;
  double Jm53, devb[9], sig[9], P[9];
  __host__ __device__ inline void init(const QP& q, const MatParams& m) {
    mu = m.mu;
    kappa = m.kappa;
    double Jm13 = 1.0 / cbrt(q.J);
    double Jm23 = Jm13 * Jm13;
    Jm53 = Jm23 * Jm23 * Jm13;
    double b[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        b[3 * i + j] = q.F[3 * i] * q.F[3 * j] + q.F[3 * i + 1] * q.F[3 * j + 1] + q.F[3 * i + 2] * q.F[3 * j + 2];
    double trb = (b[0] + b[4] + b[8]) / 3.0;
    double pJ = 0.5 * kappa * (q.J - 1.0 / q.J);
    for (int k = 0; k < 9; ++k) {
      devb[k] = b[k] - ((k % 4 == 0) ? trb : 0.0);
      sig[k] = mu * Jm53 * devb[k] + ((k % 4 == 0) ? pJ : 0.0);
    }
    mixed(sig, q.p);
    piola(sig, q.Fi, q.J, P);
  }
  __host__ __device__ inline double r1(const QP& q) const { return q.p / kappa - 0.5 * (q.J - 1.0 / q.J); }
  // dF = e_k e_L^T:  dJ = J Finv_Lk, d(J^{-5/3}) = -5/3 J^{-5/3} dJ/J,
  // db_ij = d_ik F_jL + F_iL d_jk, d(pJ) = 1/2 kappa (1 + 1/J^2) dJ.
  __host__ __device__ inline void dir(const QP& q, int k, int L, Dir& d) const {
    double dJJ = q.Fi[3 * L + k];
    double dJ = q.J * dJJ;
    double dJm53 = -(5.0 / 3.0) * Jm53 * dJJ;
    double dtrb = 2.0 * q.F[3 * k + L] / 3.0;
    double dpJ = 0.5 * kappa * (1.0 + 1.0 / (q.J * q.J)) * dJ;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double db = (i == k ? q.F[3 * j + L] : 0.0) + (j == k ? q.F[3 * i + L] : 0.0);
        double ddev = db - (i == j ? dtrb : 0.0);
        d.dsig[3 * i + j] = mu * (dJm53 * devb[3 * i + j] + Jm53 * ddev) + (i == j ? dpJ : 0.0);
      }
    double dpbar = (d.dsig[0] + d.dsig[4] + d.dsig[8]) / 3.0;
    d.dsig[0] -= dpbar;
    d.dsig[4] -= dpbar;
    d.dsig[8] -= dpbar;
    dpiola(q, P, d.dsig, k, L, d.dP);
    d.dr1 = -0.5 * (1.0 + 1.0 / (q.J * q.J)) * dJ;
  }
};

// ----------------------------------------------- finite-strain J2 plasticity
// Radial return of src/goal_J2.cpp:71-143 with linear isotropic hardening:
//   be = J^{-2/3} F Cp^{-1} F^T, s = mu dev(be), mubar = mu tr(be)/3,
//   f = |s| - sqrt(2/3)(Y + K eqps_old); if f > 1e-12 the scalar Newton of the
//   reference runs for dgam, s -= 2 mubar dgam N, eqps = eqps_old + sqrt(2/3)
//   dgam, Fp = exp(dgam N) Fp_old.  sigma = s/J + 1/2 kappa (J - 1/J) I.
// The tangent carries the derivative of every Newton iterate, so it equals
// the derivative of the iterated map the reference executes.
struct J2Model {
  double mu, kappa, K, Y;
  double Jm23, bebar[9], str[9], smag, mubar, f, eqps_old;
  bool plastic;
  int iters;
  double dgam, alpha, sig[9], P[9], s[9];
  __host__ __device__ inline void init(const QP& q, const MatParams& m, const double* Fp_old, double eqps_o) {
    mu = m.mu;
    kappa = m.kappa;
    K = m.K;
    Y = m.Y;
    eqps_old = eqps_o;
    Jm23 = pow(q.J, -2.0 / 3.0);
    double Fpi[9];
    inv3(Fp_old, det3(Fp_old), Fpi);
    double Cpi[9], G[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        Cpi[3 * i + j] = Fpi[3 * i] * Fpi[3 * j] + Fpi[3 * i + 1] * Fpi[3 * j + 1] + Fpi[3 * i + 2] * Fpi[3 * j + 2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        G[3 * i + j] = q.F[3 * i] * Cpi[j] + q.F[3 * i + 1] * Cpi[3 + j] + q.F[3 * i + 2] * Cpi[6 + j];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        bebar[3 * i + j] = G[3 * i] * q.F[3 * j] + G[3 * i + 1] * q.F[3 * j + 1] + G[3 * i + 2] * q.F[3 * j + 2];
    double trbe = Jm23 * (bebar[0] + bebar[4] + bebar[8]);
    double ss = 0.0;
    for (int k = 0; k < 9; ++k) {
      str[k] = mu * (Jm23 * bebar[k] - ((k % 4 == 0) ? trbe / 3.0 : 0.0));
      ss += str[k] * str[k];
    }
    mubar = trbe * mu / 3.0;
    smag = sqrt(ss);
    const double sq23 = sqrt(2.0 / 3.0);
    f = smag - sq23 * (Y + K * eqps_old);
    plastic = f > 1.0e-12;
    dgam = 0.0;
    alpha = eqps_old;
    iters = 0;
    for (int k = 0; k < 9; ++k) s[k] = str[k];
    if (plastic) {
      double dX;
      iters = newton(0.0, 0.0, dgam, dX);
      alpha = eqps_old + sq23 * dgam;
      for (int k = 0; k < 9; ++k) s[k] = str[k] - 2.0 * mubar * dgam * (str[k] / smag);
    }
    double pJ = 0.5 * kappa * (q.J - 1.0 / q.J);
    for (int k = 0; k < 9; ++k) sig[k] = s[k] / q.J + ((k % 4 == 0) ? pJ : 0.0);
    mixed(sig, q.p);
    piola(sig, q.Fi, q.J, P);
  }
  // The scalar Newton of src/goal_J2.cpp:98-121 together with the derivative
  // of each iterate along a direction in which |s| and mubar change by
  // dsmag and dmubar (eqps_old, Y, K are fixed).
  __host__ __device__ inline int newton(double dsmag, double dmubar, double& X, double& dX) const {
    const double sq23 = sqrt(2.0 / 3.0);
    X = 0.0;
    dX = 0.0;
    double R = f, dR = dsmag;
    double dRdX = -2.0 * mubar * (1.0 + 0.0 / (3.0 * mubar));
    double ddRdX = -2.0 * dmubar;
    int iter = 0;
    bool converged = false;
    while (!converged && iter < 30) {
      ++iter;
      double Xn = X - R / dRdX;
      dX = dX - (dR * dRdX - R * ddRdX) / (dRdX * dRdX);
      X = Xn;
      double al = eqps_old + sq23 * X;
      double H = K * al;
      R = smag - (2.0 * mubar * X + sq23 * (Y + H));
      dR = dsmag - (2.0 * dmubar * X + 2.0 * mubar * dX + sq23 * K * sq23 * dX);
      dRdX = -2.0 * mubar * (1.0 + K / (3.0 * mubar));
      ddRdX = -2.0 * dmubar;
      double res = fabs(R);
      if ((res < 1.0e-11) || (res / Y < 1.0e-11) || (res / f < 1.0e-11)) converged = true;
    }
    return iter;
  }
  __host__ __device__ inline bool failed() const { return plastic && iters >= 30; }
  __host__ __device__ inline void new_Fp(const double* Fp_old, double* Fpn) const {
    double A[9], E[9];
    for (int k = 0; k < 9; ++k) A[k] = dgam * (str[k] / smag);
    expm3(A, E);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        Fpn[3 * i + j] = E[3 * i] * Fp_old[j] + E[3 * i + 1] * Fp_old[3 + j] + E[3 * i + 2] * Fp_old[6 + j];
  }
  __host__ __device__ inline double r1(const QP& q) const { return q.p / kappa - 0.5 * (q.J - 1.0 / q.J); }
  // dF = e_k e_L^T with bebar = F Cp^{-1} F^T and G = F Cp^{-1}:
  //   d bebar_ij = d_ik G_jL + G_iL d_jk, with G_jL = (bebar F^{-T})_jL
  __host__ __device__ inline void dir(const QP& q, int k, int L, Dir& d) const {
    double FiL[3];  // row L of Finv
    for (int m = 0; m < 3; ++m) FiL[m] = sel3(L, q.Fi[m], q.Fi[3 + m], q.Fi[6 + m]);
    double dJJ = sel3(k, FiL[0], FiL[1], FiL[2]);
    double dJ = q.J * dJJ;
    double dJm23 = -(2.0 / 3.0) * Jm23 * dJJ;
    // G = F Cp^{-1} recovered from bebar: G = bebar F^{-T}
    double GL[3];
    for (int j = 0; j < 3; ++j)
      GL[j] = bebar[3 * j] * FiL[0] + bebar[3 * j + 1] * FiL[1] + bebar[3 * j + 2] * FiL[2];
    double dbe[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double db = (i == k ? GL[j] : 0.0) + (j == k ? GL[i] : 0.0);
        dbe[3 * i + j] = dJm23 * bebar[3 * i + j] + Jm23 * db;
      }
    double dtr = dbe[0] + dbe[4] + dbe[8];
    double dstr[9], dsmag = 0.0;
    for (int m = 0; m < 9; ++m) {
      dstr[m] = mu * (dbe[m] - ((m % 4 == 0) ? dtr / 3.0 : 0.0));
      dsmag += str[m] * dstr[m];
    }
    dsmag /= smag;
    double dmubar = dtr * mu / 3.0;
    double ds[9];
    if (plastic) {
      double X, dX;
      (void)newton(dsmag, dmubar, X, dX);
      for (int m = 0; m < 9; ++m) {
        double N = str[m] / smag;
        double dN = dstr[m] / smag - str[m] * dsmag / (smag * smag);
        ds[m] = dstr[m] - 2.0 * (dmubar * dgam * N + mubar * dX * N + mubar * dgam * dN);
      }
    } else {
      for (int m = 0; m < 9; ++m) ds[m] = dstr[m];
    }
    double dpJ = 0.5 * kappa * (1.0 + 1.0 / (q.J * q.J)) * dJ;
    for (int m = 0; m < 9; ++m)
      d.dsig[m] = ds[m] / q.J - s[m] * dJ / (q.J * q.J) + ((m % 4 == 0) ? dpJ : 0.0);
    double dpbar = (d.dsig[0] + d.dsig[4] + d.dsig[8]) / 3.0;
    d.dsig[0] -= dpbar;
    d.dsig[4] -= dpbar;
    d.dsig[8] -= dpbar;
    dpiola(q, P, d.dsig, k, L, d.dP);
    d.dr1 = -0.5 * (1.0 + 1.0 / (q.J * q.J)) * dJ;
  }
};

// ------------------------------------------------------ small-strain elastic
// sigma = 2 mu sym(H) + lambda tr(H) I, Mixed, P = sigma   (src/goal_elastic.cpp:61-70)
// r1 = p - (2 mu/3 + lambda) tr(H)                         (src/goal_presidual.cpp:61-69)
struct Elastic {
  double mu, lambda, sig[9], P[9];
  __host__ __device__ inline void init(const QP& q, const MatParams& m) {
    mu = m.mu;
    lambda = m.lambda;
    double tr = q.H[0] + q.H[4] + q.H[8];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        sig[3 * i + j] = 2.0 * mu * 0.5 * (q.H[3 * i + j] + q.H[3 * j + i]) + (i == j ? lambda * tr : 0.0);
    mixed(sig, q.p);
    for (int k = 0; k < 9; ++k) P[k] = sig[k];
  }
  __host__ __device__ inline double r1(const QP& q) const {
    return q.p - (2.0 * mu / 3.0 + lambda) * (q.H[0] + q.H[4] + q.H[8]);
  }
  __host__ __device__ inline void dir(const QP&, int k, int L, Dir& d) const {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        d.dsig[3 * i + j] = mu * ((i == k && j == L ? 1.0 : 0.0) + (j == k && i == L ? 1.0 : 0.0)) +
                            (i == j && k == L ? lambda : 0.0);
    double dpbar = (d.dsig[0] + d.dsig[4] + d.dsig[8]) / 3.0;
    d.dsig[0] -= dpbar;
    d.dsig[4] -= dpbar;
    d.dsig[8] -= dpbar;
    for (int m = 0; m < 9; ++m) d.dP[m] = d.dsig[m];
    d.dr1 = (k == L) ? -(2.0 * mu / 3.0 + lambda) : 0.0;
  }
};

// ------------------------------------------- neo-Hookean tangent, closed form
// After Mixed the neo-Hookean stress is sigma = mu J^{-5/3} dev(F F^T) + p I,
// so P = c A + p J F^{-T} with c = mu J^{-2/3} and A = F - tr(F F^T)/3 F^{-T}.
// Differentiating along dF = e_k e_L^T (dJ = J Finv_Lk, dF^{-T}_iJ =
// -Finv_Jk Finv_Li, d tr(F F^T) = 2 F_kL):
//   dP_iJ = Finv_Lk (P_iJ - 5/3 c A_iJ) - Finv_Jk P_iL
//           + c (d_ik d_JL + F_iL Finv_Jk - 2/3 F_kL Finv_Ji)
//   dr1   = -1/2 (1 + 1/J^2) J Finv_Lk
// and for the stabilization flux s3_j = tau J (Cinv g)_j, with h = Finv^T g
// and v = Cinv g:
//   ds3_j = tau (Finv_Lk J v_j - J (h_k Cinv_Lj + Finv_jk v_L)).
// The reference differentiates the same formulas before simplification
// (the bulk term 1/2 kappa (J - 1/J) is added by the model and removed again
// by Mixed), so the two agree to rounding.
// S: the type the point's tangent base is kept in (double; float for FP32
// element matrices, whose columns are evaluated in float anyway).  The base
// keeps the point's combinations that every column shares: A = c F - P,
// Q = P - 5/3 c A, tau J Cinv and tau J v, so a displacement column costs
// about 74 operations instead of 90.
template <typename S>
struct NHBaseT {
  S F[9], Fi[9], A[9], Q[9], tCinv[9], tv[3], h[3];
  S J, c, c23, beta, ikappa;  // c23 = 2/3 c, beta = -1/2 (1 + 1/J^2) J
};
using NHBase = NHBaseT<double>;

template <typename S>
__host__ __device__ inline void nh_base(const QP& q, const NeoHookean& m, const Stab& st, NHBaseT<S>& b) {
  double trb = 0.0;
  for (int k = 0; k < 9; ++k) trb += q.F[k] * q.F[k];
  const double c = m.mu * q.J * m.Jm53;
  const double tJ = st.tau * q.J;
  for (int k = 0; k < 9; ++k) {
    b.F[k] = q.F[k];
    b.Fi[k] = q.Fi[k];
    b.A[k] = c * q.F[k] - m.P[k];
    b.tCinv[k] = tJ * st.Cinv[k];
  }
  // Q = P - 5/3 c A,  A = F - (tr b / 3) F^{-T}
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      b.Q[3 * i + j] = m.P[3 * i + j] - (5.0 / 3.0) * c * (q.F[3 * i + j] - (trb / 3.0) * q.Fi[3 * j + i]);
  for (int j = 0; j < 3; ++j) {
    b.tv[j] = tJ * (st.Cinv[3 * j] * q.g[0] + st.Cinv[3 * j + 1] * q.g[1] + st.Cinv[3 * j + 2] * q.g[2]);
    b.h[j] = q.Fi[j] * q.g[0] + q.Fi[3 + j] * q.g[1] + q.Fi[6 + j] * q.g[2];
  }
  b.J = q.J;
  b.c = c;
  b.c23 = (2.0 / 3.0) * c;
  b.beta = -0.5 * (1.0 + 1.0 / (q.J * q.J)) * q.J;
  b.ikappa = 1.0 / m.kappa;
}

// The 13 tangent rows (P_iJ, r1, s3_j) in one direction of the element DOFs,
// contracted with the shape function of node b.
//   Displacement DOF k of node b (dF = e_k (x) G, G = dN_b/dX):
//     dP_iJ = g_k Q_iJ + Finv_Jk (c (F G)_i - (P G)_i) + c d_ik G_J - 2/3 c (F G)_k Finv_Ji
//     dr1   = beta g_k,   with g_k = sum_L G_L Finv_Lk
//     ds3_j = tau J (g_k v_j - h_k (Cinv G)_j - Finv_jk (v . G))
//   Pressure DOF of node b (dp = N_b, d grad p = G), Mixed and PResidual:
//     dP_iJ = N_b J Finv_Ji,   dr1 = N_b / kappa,   ds3_j = tau J (Cinv G)_j
// These are the columns of the point tangent summed against G; scaling G and
// N_b by w dv scales the result.  T = float evaluates the FP32 tangent from the
// FP64 point state.
// Displacement DOF k (< 3) of node b.
template <typename T, typename S>
__host__ __device__ inline void nh_direction_u(const NHBaseT<S>& b, int k, const T* G, T* C) {
  const T c = (T)b.c;
  T t[3], CG[3];  // t = (c F - P) G,  CG = tau J Cinv G
  for (int i = 0; i < 3; ++i) {
    t[i] = (T)b.A[3 * i] * G[0] + (T)b.A[3 * i + 1] * G[1] + (T)b.A[3 * i + 2] * G[2];
    CG[i] = (T)b.tCinv[3 * i] * G[0] + (T)b.tCinv[3 * i + 1] * G[1] + (T)b.tCinv[3 * i + 2] * G[2];
  }
  // k-indexed quantities come from LDS (computed address), not from
  // register arrays, so no per-lane selects
  const T gk = (T)b.Fi[k] * G[0] + (T)b.Fi[3 + k] * G[1] + (T)b.Fi[6 + k] * G[2];
  const T FGk = (T)b.F[3 * k] * G[0] + (T)b.F[3 * k + 1] * G[1] + (T)b.F[3 * k + 2] * G[2];
  const T u = (T)b.c23 * FGk;
  for (int i = 0; i < 3; ++i) {
    const T dik = (i == k) ? c : (T)0;
    for (int jj = 0; jj < 3; ++jj)
      C[3 * i + jj] = gk * (T)b.Q[3 * i + jj] + (T)b.Fi[3 * jj + k] * t[i] - u * (T)b.Fi[3 * jj + i] + dik * G[jj];
  }
  C[9] = (T)b.beta * gk;
  const T vG = (T)b.tv[0] * G[0] + (T)b.tv[1] * G[1] + (T)b.tv[2] * G[2];
  for (int j = 0; j < 3; ++j) C[10 + j] = gk * (T)b.tv[j] - (T)b.h[k] * CG[j] - (T)b.Fi[3 * j + k] * vG;
}

// Pressure DOF of node b.
template <typename T, typename S>
__host__ __device__ inline void nh_direction_p(const NHBaseT<S>& b, T N, const T* G, T* C) {
  const T NJ = N * (T)b.J;
  for (int i = 0; i < 3; ++i)
    for (int jj = 0; jj < 3; ++jj) C[3 * i + jj] = NJ * (T)b.Fi[3 * jj + i];
  C[9] = N * (T)b.ikappa;
  for (int j = 0; j < 3; ++j)
    C[10 + j] = (T)b.tCinv[3 * j] * G[0] + (T)b.tCinv[3 * j + 1] * G[1] + (T)b.tCinv[3 * j + 2] * G[2];
}

template <typename T, typename S>
__host__ __device__ inline void nh_direction(const NHBaseT<S>& b, int k, T N, const T* G, T* C) {
  if (k < 3) nh_direction_u<T>(b, k, G, C);
  else nh_direction_p<T>(b, N, G, C);
}

// tau = 1/2 c0 h^2 / mu with h^2 the mean squared edge length
// (src/goal_stabilization.cpp:59-73).
__host__ __device__ inline double stab_tau(const double X[4][3], const MatParams& m) {
  const int ei[6] = {0, 1, 2, 0, 1, 2};
  const int ej[6] = {1, 2, 0, 3, 3, 3};
  double h = 0.0;
  for (int e = 0; e < 6; ++e) {
    double l2 = 0.0;
    for (int c = 0; c < 3; ++c) {
      double d = X[ej[e]][c] - X[ei[e]][c];
      l2 += d * d;
    }
    h += l2;
  }
  h = sqrt(h / 6.0);
  return 0.5 * m.c0 * h * h / m.mu;
}

}  // namespace dev
}  // namespace goal
