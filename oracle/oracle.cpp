// ORACLE -- TEST INFRASTRUCTURE ONLY.  Loaded by tests/, by the smoke check
// and by bench.py's cpu_baseline leg, always as the reference values a result
// is compared with or as the timed CPU baseline, never as part of the product
// path.
//
// A plain C++ restatement of goal's element assembly and nested-adjoint DWR
// path, written from the reference source as text and following its
// evaluator chain literally (one element at a time, the hooks' order of
// operations, forward-mode dual numbers for the Jacobian, per-element scatter
// into an FP64 CSR with a column search).  The reference itself cannot be
// built here (it needs Trilinos, SCOREC core and MPI), so this restatement is
// the oracle.  Function-by-function pins (checks that a plausible mistake
// would fail) are listed in tests/test_oracle.py; see the "pinned by" notes.
//
// Reference files followed:
//   assembly loop            src/goal_assembly.cpp:65-88
//   chain order              src/goal_mechanics.cpp:97-146, 169-218
//   gather / interpolation   src/goal_displacement.cpp:48-72,139-172; src/goal_pressure.cpp:49-71,136-165
//   weights                  src/goal_vector_weight.cpp:13-28; src/goal_scalar_weight.cpp:16-31
//   adjoint weights          src/goal_displacement_adjoint.cpp:37-53; src/goal_pressure_adjoint.cpp:38-49
//   kinematics               src/goal_kinematics.cpp:17-25
//   neo-Hookean              src/goal_neohookean.cpp:59-86
//   J2                       src/goal_J2.cpp:71-157
//   elastic                  src/goal_elastic.cpp:61-70
//   mixed                    src/goal_mixed.cpp:33-46
//   momentum residual        src/goal_mresidual.cpp:26-32
//   pressure residual        src/goal_presidual.cpp:53-77
//   stabilization            src/goal_stabilization.cpp:59-81
//   scatter                  src/goal_displacement.cpp:77-86,177-214; src/goal_pressure.cpp:76-83,170-203
//   graph                    src/goal_disc.cpp:307-332 (clique union, sorted columns)
//   Dirichlet rows           src/goal_dbcs.cpp:39-99
//   QoI                      src/goal_qoi.cpp:20-81; src/goal_avg_disp.cpp:16-21;
//                            src/goal_avg_vm.cpp:51-61; src/goal_von_mises.cpp:6-17
//   DWR transfers            src/goal_nested.cpp:369-412; src/goal_nested_adjoint.cpp:182-197
//   error indicators         src/goal_error.cpp:7-56
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <set>
#include <thread>
#include <type_traits>
#include <vector>

#include "fad.hpp"

using oracle::Fad;

namespace {

// ------------------------------------------------------------ tensors
template <typename T>
struct T3 {
  T a[3][3];
  T& operator()(int i, int j) { return a[i][j]; }
  const T& operator()(int i, int j) const { return a[i][j]; }
};

template <typename T>
T3<T> zeros() {
  T3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r(i, j) = 0.0;
  return r;
}
template <typename T>
T3<T> eye() {
  T3<T> r = zeros<T>();
  for (int i = 0; i < 3; ++i) r(i, i) = 1.0;
  return r;
}
template <typename T>
T3<T> mul(const T3<T>& A, const T3<T>& B) {
  T3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T s = A(i, 0) * B(0, j);
      s += A(i, 1) * B(1, j);
      s += A(i, 2) * B(2, j);
      r(i, j) = s;
    }
  return r;
}
template <typename T>
T3<T> transpose(const T3<T>& A) {
  T3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r(i, j) = A(j, i);
  return r;
}
template <typename T, typename S>
T3<T> scale(const S& s, const T3<T>& A) {
  T3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r(i, j) = s * A(i, j);
  return r;
}
template <typename T>
T3<T> add(const T3<T>& A, const T3<T>& B) {
  T3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r(i, j) = A(i, j) + B(i, j);
  return r;
}
template <typename T>
T3<T> sub(const T3<T>& A, const T3<T>& B) {
  T3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r(i, j) = A(i, j) - B(i, j);
  return r;
}
template <typename T>
T trace(const T3<T>& A) {
  T s = A(0, 0);
  s += A(1, 1);
  s += A(2, 2);
  return s;
}
// rule of Sarrus
template <typename T>
T det(const T3<T>& A) {
  return A(0, 0) * A(1, 1) * A(2, 2) + A(0, 1) * A(1, 2) * A(2, 0) + A(0, 2) * A(1, 0) * A(2, 1) -
         A(0, 0) * A(1, 2) * A(2, 1) - A(0, 1) * A(1, 0) * A(2, 2) - A(0, 2) * A(1, 1) * A(2, 0);
}
// adjugate / determinant
template <typename T>
T3<T> inverse(const T3<T>& A) {
  T d = det(A);
  T3<T> c;
  c(0, 0) = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
  c(0, 1) = A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2);
  c(0, 2) = A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1);
  c(1, 0) = A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2);
  c(1, 1) = A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0);
  c(1, 2) = A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2);
  c(2, 0) = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
  c(2, 1) = A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1);
  c(2, 2) = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c(i, j) = c(i, j) / d;
  return c;
}
template <typename T>
T3<T> dev(const T3<T>& A) {
  T3<T> r = A;
  T m = trace(A) / 3.0;
  for (int i = 0; i < 3; ++i) r(i, i) = r(i, i) - m;
  return r;
}
template <typename T>
T norm(const T3<T>& A) {
  T s = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s += A(i, j) * A(i, j);
  return sqrt(s);
}
using std::sqrt;
using std::cbrt;
using std::pow;
using oracle::sqrt;
using oracle::cbrt;
using oracle::pow;

// exp of a symmetric 3x3 matrix by cyclic Jacobi eigen-decomposition:
// exp(A) = Q diag(e^l) Q^T.
void expm_sym(const double A[3][3], double E[3][3]) {
  double a[3][3], V[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      a[i][j] = 0.5 * (A[i][j] + A[j][i]);
      V[i][j] = i == j ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if (off < 1e-300) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  double ev[3] = {std::exp(a[0][0]), std::exp(a[1][1]), std::exp(a[2][2])};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) E[i][j] = V[i][0] * ev[0] * V[j][0] + V[i][1] * ev[1] * V[j][1] + V[i][2] * ev[2] * V[j][2];
}

// ------------------------------------------------------------ inputs
struct Mat {
  double E, nu, K, Y, c0;
};

struct Problem {
  int order, npe, nq, model, stab, save;
  int64_t nn, ne;
  const double* coords;
  const int32_t* conn;
  const int32_t* elem_set;
  const Mat* mats;
  const double* x;
  const int64_t* row_ptr;
  const int32_t* cols;
  const double* Fp_old;
  const double* eqps_old;
  double* Fp;
  double* eqps;
  double* sigma;
  const double* z_diff;
  const double* z_coarse;
  // Rank-local rows (partitioned assembly): global node -> local node, and the
  // offset of a local node's four CSR rows in the part's value array.  Null:
  // R and vals are the global arrays.
  const int32_t* lnode = nullptr;
  const int64_t* loff = nullptr;
};

// quadrature (parent tet, weights sum to 1/6)
void quad_point(int nq, int q, double xi[3], double& w) {
  if (nq == 1) {
    xi[0] = xi[1] = xi[2] = 0.25;
    w = 1.0 / 6.0;
    return;
  }
  const double a = 0.5854101966249685, b = 0.1381966011250105;
  xi[0] = q == 1 ? a : b;
  xi[1] = q == 2 ? a : b;
  xi[2] = q == 3 ? a : b;
  w = 1.0 / 24.0;
}

// apf-style shape evaluation: local gradients times the inverse Jacobian.
struct Shape {
  double BF[10];
  double GBF[10][3];
  double dv;
};
void eval_shape(const Problem& P, int64_t e, const double xi[3], Shape& s) {
  const int32_t* ec = P.conn + e * P.npe;
  double X[4][3];
  for (int v = 0; v < 4; ++v)
    for (int c = 0; c < 3; ++c) X[v][c] = P.coords[3 * (int64_t)ec[v] + c];
  double lam[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
  const double dlam[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double dloc[10][3];
  if (P.npe == 4) {
    for (int a = 0; a < 4; ++a) {
      s.BF[a] = lam[a];
      for (int r = 0; r < 3; ++r) dloc[a][r] = dlam[a][r];
    }
  } else {
    const int ei[6] = {0, 1, 2, 0, 1, 2}, ej[6] = {1, 2, 0, 3, 3, 3};
    for (int a = 0; a < 4; ++a) {
      s.BF[a] = lam[a] * (2.0 * lam[a] - 1.0);
      for (int r = 0; r < 3; ++r) dloc[a][r] = (4.0 * lam[a] - 1.0) * dlam[a][r];
    }
    for (int k = 0; k < 6; ++k) {
      int i = ei[k], j = ej[k];
      s.BF[4 + k] = 4.0 * lam[i] * lam[j];
      for (int r = 0; r < 3; ++r) dloc[4 + k][r] = 4.0 * (lam[i] * dlam[j][r] + lam[j] * dlam[i][r]);
    }
  }
  // Jm[r][c] = dX_c / dxi_r
  T3<double> Jm;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Jm(r, c) = X[r + 1][c] - X[0][c];
  s.dv = det(Jm);
  T3<double> Ji = inverse(Jm);
  for (int a = 0; a < P.npe; ++a)
    for (int c = 0; c < 3; ++c) s.GBF[a][c] = Ji(c, 0) * dloc[a][0] + Ji(c, 1) * dloc[a][1] + Ji(c, 2) * dloc[a][2];
}

double elem_size(const Problem& P, int64_t e) {
  const int32_t* ec = P.conn + e * P.npe;
  const int ei[6] = {0, 1, 2, 0, 1, 2}, ej[6] = {1, 2, 0, 3, 3, 3};
  double h = 0.0;
  for (int k = 0; k < 6; ++k) {
    double l2 = 0.0;
    for (int c = 0; c < 3; ++c) {
      double d = P.coords[3 * (int64_t)ec[ej[k]] + c] - P.coords[3 * (int64_t)ec[ei[k]] + c];
      l2 += d * d;
    }
    double l = std::sqrt(l2);
    h += l * l;
  }
  return std::sqrt(h / 6.0);
}

// element outputs
template <typename T>
struct ElemOut {
  T ru[10][3];
  T rp[10];
};

enum { K_RESID = 0, K_JAC = 1, K_ADJ = 2, K_LOC = 3 };

// The evaluator chain for one element.
template <typename T>
void element(const Problem& P, int64_t e, int kind, ElemOut<T>& out) {
  const int nn = P.npe, neq = 4, nd = nn * neq;
  const int32_t* ec = P.conn + e * nn;
  const Mat& m = P.mats[P.elem_set ? P.elem_set[e] : 0];
  const double mu = m.E / (2.0 * (1.0 + m.nu));
  const double kappa = m.E / (3.0 * (1.0 - 2.0 * m.nu));
  const double lambda = m.E * m.nu / ((1.0 + m.nu) * (1.0 - 2.0 * m.nu));
  const bool fad = (kind == K_JAC || kind == K_ADJ);
  // gather (FAD seeds n*neq + d for u, n*neq + dim for p)
  T un[10][3], pn[10];
  for (int n = 0; n < nn; ++n) {
    for (int d = 0; d < 3; ++d) {
      un[n][d] = P.x[4 * (int64_t)ec[n] + d];
      if constexpr (!std::is_same<T, double>::value)
        if (fad) {
          double v = un[n][d].val();
          un[n][d].diff(n * neq + d, nd);
          un[n][d].v = v;
        }
      out.ru[n][d] = 0.0;
    }
    pn[n] = P.x[4 * (int64_t)ec[n] + 3];
    if constexpr (!std::is_same<T, double>::value)
      if (fad) {
        double v = pn[n].val();
        pn[n].diff(n * neq + 3, nd);
        pn[n].v = v;
      }
    out.rp[n] = 0.0;
  }
  // adjoint weight fields
  double zd[10][4], zc[10];
  if (kind == K_LOC)
    for (int n = 0; n < nn; ++n) {
      for (int c = 0; c < 4; ++c) zd[n][c] = P.z_diff[4 * (int64_t)ec[n] + c];
      zc[n] = P.z_coarse[4 * (int64_t)ec[n] + 3];
    }
  const double h = elem_size(P, e);
  for (int q = 0; q < P.nq; ++q) {
    double xi[3], ipw;
    quad_point(P.nq, q, xi, ipw);
    Shape s;
    eval_shape(P, e, xi, s);
    const double dv = s.dv;
    // Displacement::at_point
    T uval[3];
    T3<T> ugrad;
    for (int i = 0; i < 3; ++i) {
      uval[i] = un[0][i] * s.BF[0];
      for (int n = 1; n < nn; ++n) uval[i] += un[n][i] * s.BF[n];
      for (int j = 0; j < 3; ++j) {
        ugrad(i, j) = un[0][i] * s.GBF[0][j];
        for (int n = 1; n < nn; ++n) ugrad(i, j) += un[n][i] * s.GBF[n][j];
      }
    }
    // Pressure::at_point
    T pval = pn[0] * s.BF[0];
    for (int n = 1; n < nn; ++n) pval += pn[n] * s.BF[n];
    T pgrad[3];
    for (int i = 0; i < 3; ++i) {
      pgrad[i] = pn[0] * s.GBF[0][i];
      for (int n = 1; n < nn; ++n) pgrad[i] += pn[n] * s.GBF[n][i];
    }
    // weights: uw, pw (primal) or DisplacementAdjoint / PressureAdjoint
    double wv[10][3], wg[10][3][3], pwv[10], pwg[10][3], pcv[10], pcg[10][3];
    if (kind != K_LOC) {
      for (int n = 0; n < nn; ++n) {
        for (int i = 0; i < 3; ++i) {
          wv[n][i] = s.BF[n];
          for (int j = 0; j < 3; ++j) wg[n][i][j] = s.GBF[n][j];
        }
        pwv[n] = pcv[n] = s.BF[n];
        for (int j = 0; j < 3; ++j) pwg[n][j] = pcg[n][j] = s.GBF[n][j];
      }
    } else {
      double z[4] = {0, 0, 0, 0}, gz[4][3] = {{0}}, zcv = 0.0, gzc[3] = {0, 0, 0};
      for (int n = 0; n < nn; ++n) {
        for (int c = 0; c < 4; ++c) {
          z[c] += zd[n][c] * s.BF[n];
          for (int j = 0; j < 3; ++j) gz[c][j] += zd[n][c] * s.GBF[n][j];
        }
        zcv += zc[n] * s.BF[n];
        for (int j = 0; j < 3; ++j) gzc[j] += zc[n] * s.GBF[n][j];
      }
      for (int n = 0; n < nn; ++n) {
        for (int i = 0; i < 3; ++i) {
          wv[n][i] = z[i] * s.BF[n];
          for (int j = 0; j < 3; ++j) wg[n][i][j] = gz[i][j] * s.BF[n] + z[i] * s.GBF[n][j];
        }
        pwv[n] = z[3] * s.BF[n];
        for (int j = 0; j < 3; ++j) pwg[n][j] = gz[3][j] * s.BF[n] + z[3] * s.GBF[n][j];
        pcv[n] = zcv * s.BF[n];
        for (int j = 0; j < 3; ++j) pcg[n][j] = gzc[j] * s.BF[n] + zcv * s.GBF[n][j];
      }
    }
    (void)pcv;
    // Kinematics
    T3<T> F = ugrad;
    for (int i = 0; i < 3; ++i) F(i, i) += 1.0;
    T J = det(F);
    // Model
    T3<T> sigma;
    const int64_t sidx = e * P.nq + q, npts = P.ne * P.nq;
    if (P.model == 0) {
      T3<T> eps;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) eps(i, j) = 0.5 * (ugrad(i, j) + ugrad(j, i));
      sigma = add(scale(2.0 * mu, eps), scale(lambda * trace(eps), eye<T>()));
    } else if (P.model == 1) {
      T Jm13 = 1.0 / cbrt(J);
      T Jm23 = Jm13 * Jm13;
      T Jm53 = Jm23 * Jm23 * Jm13;
      T3<T> b = mul(F, transpose(F));
      T pp = 0.5 * kappa * (J - 1.0 / J);
      sigma = add(scale(mu * Jm53, dev(b)), scale(pp, eye<T>()));
    } else {
      const double sq23 = std::sqrt(2.0 / 3.0);
      T Jm23 = pow(J, -2.0 / 3.0);
      T3<T> Fp;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Fp(i, j) = P.Fp_old[(3 * i + j) * npts + sidx];
      T3<T> Fpinv = inverse(Fp);
      T3<T> Cpinv = mul(Fpinv, transpose(Fpinv));
      T3<T> be = scale(Jm23, mul(mul(F, Cpinv), transpose(F)));
      T3<T> s_ = scale(mu, dev(be));
      T mubar = trace(be) * mu / 3.0;
      T smag = norm(s_);
      T eqps = P.eqps_old[sidx];
      T f = smag - sq23 * (m.Y + m.K * eqps);
      if (f > 1.0e-12) {
        int iter = 0;
        bool converged = false;
        T H(0.0), dH(0.0), alpha(0.0), res(0.0);
        T X = 0.0;
        T R = f;
        T dRdX = -2.0 * mubar * (1.0 + H / (3.0 * mubar));
        while ((!converged) && (iter < 30)) {
          iter++;
          X = X - R / dRdX;
          alpha = eqps + sq23 * X;
          H = m.K * alpha;
          dH = m.K;
          R = smag - (2.0 * mubar * X + sq23 * (m.Y + H));
          dRdX = -2.0 * mubar * (1.0 + dH / (3.0 * mubar));
          res = oracle::abs(R);
          if ((res < 1.0e-11) || (oracle::val(res) / m.Y < 1.0e-11) || (oracle::val(res) / oracle::val(f) < 1.0e-11))
            converged = true;
          if (iter == 30) throw 3;
        }
        T dgam = X;
        T3<T> N = scale(1.0 / smag, s_);
        s_ = sub(s_, scale(2.0 * mubar * dgam, N));
        if (P.save) P.eqps[sidx] = oracle::val(alpha);
        double A[3][3], Ex[3][3];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) A[i][j] = oracle::val(dgam) * oracle::val(N(i, j));
        expm_sym(A, Ex);
        if (P.save)
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              double v = 0.0;
              for (int k = 0; k < 3; ++k) v += Ex[i][k] * oracle::val(Fp(k, j));
              P.Fp[(3 * i + j) * npts + sidx] = v;
            }
      } else if (P.save) {
        P.eqps[sidx] = oracle::val(eqps);
      }
      T pp = 0.5 * kappa * (J - 1.0 / J);
      sigma = add(scale(1.0 / J, s_), scale(pp, eye<T>()));
    }
    // Mixed
    {
      T pbar = 0.0;
      for (int i = 0; i < 3; ++i) pbar += sigma(i, i);
      pbar /= 3.0;
      for (int i = 0; i < 3; ++i) sigma(i, i) += pval - pbar;
      if (P.save)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) P.sigma[(3 * i + j) * npts + sidx] = oracle::val(sigma(i, j));
    }
    // MResidual (get_first_pk: P = J sigma F^{-T}; elastic: P = sigma)
    T3<T> Pk;
    if (P.model == 0) Pk = sigma;
    else Pk = scale(J, mul(sigma, transpose(inverse(F))));
    for (int n = 0; n < nn; ++n)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out.ru[n][i] += Pk(i, j) * wg[n][i][j] * ipw * dv;
    // PResidual
    if (P.model == 0) {
      T pbar = 0.0;
      for (int i = 0; i < 3; ++i) pbar += ugrad(i, i);
      pbar *= (2.0 * mu / 3.0 + lambda);
      for (int n = 0; n < nn; ++n) out.rp[n] += (pval - pbar) * pwv[n] * ipw * dv;
    } else {
      T dUdJ = 0.5 * (J - 1.0 / J);
      for (int n = 0; n < nn; ++n) out.rp[n] += ((pval / kappa) - dUdJ) * pwv[n] * ipw * dv;
    }
    // Stabilization (weight pwc in the error chain)
    if (P.stab) {
      double tau = 0.5 * m.c0 * h * h / mu;
      T3<T> Cinv = inverse(mul(transpose(F), F));
      for (int n = 0; n < nn; ++n)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) out.rp[n] += tau * J * Cinv(i, j) * pgrad[i] * pcg[n][j] * ipw * dv;
    }
  }
}

int64_t find_col(const Problem& P, int64_t row, int32_t col) {
  const int32_t* b = P.cols + P.row_ptr[row];
  const int32_t* e = P.cols + P.row_ptr[row + 1];
  const int32_t* it = std::lower_bound(b, e, col);
  if (it == e || *it != col) return -1;
  return P.row_ptr[row] + (it - b);
}

template <int ND>
int assemble_range(const Problem& P, int kind, int64_t e0, int64_t e1, double* R, double* vals) {
  const int nn = P.npe;
  for (int64_t e = e0; e < e1; ++e) {
    const int32_t* ec = P.conn + e * nn;
    int lids[40];
    for (int n = 0; n < nn; ++n)
      for (int q = 0; q < 4; ++q) lids[4 * n + q] = 4 * ec[n] + q;
    // where node n's rows live: R row base and CSR value shift
    int64_t rbase[10], vshift[10];
    for (int n = 0; n < nn; ++n) {
      if (P.lnode) {
        const int32_t ln = P.lnode[ec[n]];
        rbase[n] = 4 * (int64_t)ln;
        vshift[n] = P.loff[ln] - P.row_ptr[4 * (int64_t)ec[n]];
      } else {
        rbase[n] = 4 * (int64_t)ec[n];
        vshift[n] = 0;
      }
    }
    if (kind == K_JAC || kind == K_ADJ) {
      ElemOut<Fad<ND>> o;
      element(P, e, kind, o);
      for (int n = 0; n < nn; ++n)
        for (int i = 0; i < 4; ++i) {
          const Fad<ND>& v = i < 3 ? o.ru[n][i] : o.rp[n];
          int row = lids[4 * n + i];
          if (R) R[rbase[n] + i] += v.val();
          if (!vals) continue;
          for (int c = 0; c < 4 * nn; ++c) {
            if (kind == K_JAC) {
              int64_t slot = find_col(P, row, lids[c]);
              if (slot < 0) return 2;
              vals[slot + vshift[n]] += v.dx(c);
            } else {
              int64_t slot = find_col(P, lids[c], row);
              if (slot < 0) return 2;
              vals[slot + vshift[c >> 2]] += v.dx(c);
            }
          }
        }
    } else {
      ElemOut<double> o;
      element(P, e, kind, o);
      for (int n = 0; n < nn; ++n)
        for (int i = 0; i < 4; ++i) R[rbase[n] + i] += i < 3 ? o.ru[n][i] : o.rp[n];
    }
  }
  return 0;
}

int assemble_any(const Problem& P, int kind, int64_t e0, int64_t e1, double* R, double* vals) {
  try {
    if (P.npe == 4) return assemble_range<16>(P, kind, e0, e1, R, vals);
    return assemble_range<40>(P, kind, e0, e1, R, vals);
  } catch (int code) {
    return code;
  }
}

}  // namespace

extern "C" {

// Clique-union graph with std::set per row: rows in lid order, columns
// ascending.  Returns nnz; fills row_ptr [nn*4+1] always and cols when
// cols_cap >= nnz.
int64_t oracle_graph(int npe, int64_t nn, int64_t ne, const int32_t* conn, int64_t* row_ptr, int32_t* cols,
                     int64_t cols_cap) {
  std::vector<std::set<int32_t>> rows(nn * 4);
  for (int64_t e = 0; e < ne; ++e)
    for (int i = 0; i < npe; ++i)
      for (int j = 0; j < 4; ++j) {
        int64_t row = 4 * (int64_t)conn[e * npe + i] + j;
        for (int k = 0; k < npe; ++k)
          for (int l = 0; l < 4; ++l) rows[row].insert(4 * conn[e * npe + k] + l);
      }
  int64_t nnz = 0;
  row_ptr[0] = 0;
  for (int64_t r = 0; r < nn * 4; ++r) {
    nnz += (int64_t)rows[r].size();
    row_ptr[r + 1] = nnz;
  }
  if (cols && cols_cap >= nnz) {
    int64_t k = 0;
    for (int64_t r = 0; r < nn * 4; ++r)
      for (int32_t c : rows[r]) cols[k++] = c;
  }
  return nnz;
}

// Element assembly.  kind: 0 residual (ST, PRIMAL), 1 Jacobian (FADT,
// PRIMAL), 2 Jacobian transpose (FADT, ADJOINT), 3 localized residual (ST,
// DWR weights).  mats: [num_sets][5] = E, nu, K, Y, c0.  States are
// [comp][ne*nq].  threads > 1 splits the elements into contiguous ranges
// assembled into private copies of R / vals that are summed at the end (one
// partition per core, the reference's rank-per-part model).
// Returns 0, 2 (column missing from graph) or 3 (J2 return map failed).
int oracle_assemble(int kind, int order, int quad_order, int model, int stab, int save, int64_t nn, int64_t ne,
                    const double* coords, const int32_t* conn, const int32_t* elem_set, const double* mats,
                    const double* x, const int64_t* row_ptr, const int32_t* cols, double* R, double* vals,
                    const double* Fp_old, const double* eqps_old, double* Fp, double* eqps, double* sigma,
                    const double* z_diff, const double* z_coarse, int threads) {
  Problem P;
  P.order = order;
  P.npe = order == 1 ? 4 : 10;
  P.nq = quad_order == 1 ? 1 : 4;
  P.model = model;
  P.stab = stab;
  P.save = save;
  P.nn = nn;
  P.ne = ne;
  P.coords = coords;
  P.conn = conn;
  P.elem_set = elem_set;
  P.mats = reinterpret_cast<const Mat*>(mats);
  P.x = x;
  P.row_ptr = row_ptr;
  P.cols = cols;
  P.Fp_old = Fp_old;
  P.eqps_old = eqps_old;
  P.Fp = Fp;
  P.eqps = eqps;
  P.sigma = sigma;
  P.z_diff = z_diff;
  P.z_coarse = z_coarse;
  if (threads <= 1) return assemble_any(P, kind, 0, ne, R, vals);
  const int64_t ndof = nn * 4, nnz = row_ptr ? row_ptr[ndof] : 0;
  std::vector<std::vector<double>> pr(threads), pv(threads);
  std::vector<int> st(threads, 0);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      int64_t e0 = ne * t / threads, e1 = ne * (t + 1) / threads;
      if (R) pr[t].assign(ndof, 0.0);
      if (vals) pv[t].assign(nnz, 0.0);
      st[t] = assemble_any(P, kind, e0, e1, R ? pr[t].data() : nullptr, vals ? pv[t].data() : nullptr);
    });
  }
  for (auto& th : pool) th.join();
  // gather (Export ADD), parallel over entries
  std::vector<std::thread> merge;
  for (int t = 0; t < threads; ++t)
    merge.emplace_back([&, t] {
      if (R) {
        int64_t a = ndof * t / threads, b = ndof * (t + 1) / threads;
        for (int s = 0; s < threads; ++s)
          for (int64_t i = a; i < b; ++i) R[i] += pr[s][i];
      }
      if (vals) {
        int64_t a = nnz * t / threads, b = nnz * (t + 1) / threads;
        for (int s = 0; s < threads; ++s)
          for (int64_t i = a; i < b; ++i) vals[i] += pv[s][i];
      }
    });
  for (auto& th : merge) th.join();
  for (int s : st)
    if (s) return s;
  return 0;
}

// ------------------------------------------------ rank-per-part assembly
// The reference's parallel model on one host: one part per core, each owning
// a contiguous range of elements (a slab of the mesh, since elements are
// numbered along the mesh) and holding only the rows of the nodes its
// elements touch -- its owned and ghost rows, as the ghost SolInfo of one MPI
// rank (src/goal_sol_info.cpp:6-31).  An assembly is, per part, zero its
// rows, assemble its elements into them (scatter, src/goal_displacement.cpp:
// 177-194), then the Export ADD of the ghost rows into their owners
// (src/goal_sol_info.cpp:33-43): every node is owned by the first part that
// touches it, which sums the node's rows over all touching parts in part order.
// The plan (which rows each part holds) is built once, as the reference's
// Disc builds its maps once per mesh.
struct OracleParts {
  int threads = 1;
  int64_t nn = 0, ne = 0;
  const int64_t* row_ptr = nullptr;
  std::vector<int64_t> e0;                   // [threads + 1]
  std::vector<std::vector<int32_t>> nodes;   // per part: touched nodes, ascending
  std::vector<std::vector<int32_t>> lnode;   // per part: node -> local index or -1
  std::vector<std::vector<int64_t>> loff;    // per part: local node -> value offset [nodes + 1]
  std::vector<std::vector<double>> R, V;     // per part: local rows
  std::vector<int64_t> tptr;                 // node -> touching parts (CSR)
  std::vector<int32_t> tpart, tloc;
  std::vector<std::vector<int32_t>> owned;   // per part: nodes it owns
};

void* oracle_parts_create(int order, int64_t nn, int64_t ne, const int32_t* conn, const int64_t* row_ptr,
                          int threads) {
  const int npe = order == 1 ? 4 : 10;
  auto* P = new OracleParts;
  P->threads = std::max(1, threads);
  P->nn = nn;
  P->ne = ne;
  P->row_ptr = row_ptr;
  const int T = P->threads;
  P->e0.resize(T + 1);
  for (int t = 0; t <= T; ++t) P->e0[t] = ne * t / T;
  P->nodes.resize(T);
  P->lnode.resize(T);
  P->loff.resize(T);
  P->R.resize(T);
  P->V.resize(T);
  std::vector<std::thread> pool;
  for (int t = 0; t < T; ++t)
    pool.emplace_back([&, t] {
      std::vector<int32_t>& ln = P->lnode[t];
      ln.assign(nn, -1);
      for (int64_t e = P->e0[t]; e < P->e0[t + 1]; ++e)
        for (int k = 0; k < npe; ++k) ln[conn[e * npe + k]] = 0;
      std::vector<int32_t>& nd = P->nodes[t];
      for (int64_t a = 0; a < nn; ++a)
        if (ln[a] == 0) nd.push_back((int32_t)a);
      std::vector<int64_t>& off = P->loff[t];
      off.assign(nd.size() + 1, 0);
      for (size_t k = 0; k < nd.size(); ++k) {
        ln[nd[k]] = (int32_t)k;
        off[k + 1] = off[k] + (row_ptr[4 * (int64_t)nd[k] + 4] - row_ptr[4 * (int64_t)nd[k]]);
      }
      P->R[t].assign(4 * nd.size(), 0.0);
      P->V[t].assign(off.back(), 0.0);
    });
  for (auto& th : pool) th.join();
  // touching parts of every node, in part order; the first one owns it
  P->tptr.assign(nn + 1, 0);
  for (int t = 0; t < T; ++t)
    for (int32_t a : P->nodes[t]) ++P->tptr[a + 1];
  for (int64_t a = 0; a < nn; ++a) P->tptr[a + 1] += P->tptr[a];
  P->tpart.resize(P->tptr[nn]);
  P->tloc.resize(P->tptr[nn]);
  std::vector<int64_t> fill(P->tptr.begin(), P->tptr.end() - 1);
  for (int t = 0; t < T; ++t)
    for (size_t k = 0; k < P->nodes[t].size(); ++k) {
      const int32_t a = P->nodes[t][k];
      P->tpart[fill[a]] = t;
      P->tloc[fill[a]++] = (int32_t)k;
    }
  P->owned.resize(T);
  for (int64_t a = 0; a < nn; ++a)
    if (P->tptr[a + 1] > P->tptr[a]) P->owned[P->tpart[P->tptr[a]]].push_back((int32_t)a);
  return P;
}

void oracle_parts_destroy(void* h) { delete static_cast<OracleParts*>(h); }

// Rows held by each part and its ghost share: out[0] = sum of rows held over
// parts, out[1] = rows of the mesh (a measure of the ghost overhead).
void oracle_parts_rows(void* h, int64_t* out) {
  auto* P = static_cast<OracleParts*>(h);
  int64_t held = 0;
  for (auto& nd : P->nodes) held += 4 * (int64_t)nd.size();
  out[0] = held;
  out[1] = 4 * P->nn;
}

// One assembly over the parts (arguments as oracle_assemble); R and vals
// (global) receive the exported sums, added to what they hold.  seconds[3]:
// the longest part's zeroing, element assembly and export times.
int oracle_assemble_parts(void* h, int kind, int order, int quad_order, int model, int stab, int save,
                          const double* coords, const int32_t* conn, const int32_t* elem_set, const double* mats,
                          const double* x, const int32_t* cols, double* R, double* vals, const double* Fp_old,
                          const double* eqps_old, double* Fp, double* eqps, double* sigma, double* seconds) {
  auto* H = static_cast<OracleParts*>(h);
  const int T = H->threads;
  std::vector<int> st(T, 0);
  std::vector<double> tz(T), ta(T), tx(T);
  std::vector<std::thread> pool;
  std::atomic<int> arrived{0};
  for (int t = 0; t < T; ++t)
    pool.emplace_back([&, t] {
      using clk = std::chrono::steady_clock;
      auto c0 = clk::now();
      std::fill(H->R[t].begin(), H->R[t].end(), 0.0);
      if (vals) std::fill(H->V[t].begin(), H->V[t].end(), 0.0);
      auto c1 = clk::now();
      Problem P;
      P.order = order;
      P.npe = order == 1 ? 4 : 10;
      P.nq = quad_order == 1 ? 1 : 4;
      P.model = model;
      P.stab = stab;
      P.save = save;
      P.nn = H->nn;
      P.ne = H->ne;
      P.coords = coords;
      P.conn = conn;
      P.elem_set = elem_set;
      P.mats = reinterpret_cast<const Mat*>(mats);
      P.x = x;
      P.row_ptr = H->row_ptr;
      P.cols = cols;
      P.Fp_old = Fp_old;
      P.eqps_old = eqps_old;
      P.Fp = Fp;
      P.eqps = eqps;
      P.sigma = sigma;
      P.z_diff = nullptr;
      P.z_coarse = nullptr;
      P.lnode = H->lnode[t].data();
      P.loff = H->loff[t].data();
      st[t] = assemble_any(P, kind, H->e0[t], H->e0[t + 1], H->R[t].data(), vals ? H->V[t].data() : nullptr);
      auto c2 = clk::now();
      // every part's rows are complete before any export reads them
      arrived.fetch_add(1);
      while (arrived.load() < T) std::this_thread::yield();
      auto c2b = clk::now();
      for (int32_t a : H->owned[t]) {
        const int64_t r0 = H->row_ptr[4 * (int64_t)a];
        const int64_t len = H->row_ptr[4 * (int64_t)a + 4] - r0;
        for (int64_t k = H->tptr[a]; k < H->tptr[a + 1]; ++k) {
          const int p = H->tpart[k], l = H->tloc[k];
          for (int i = 0; i < 4; ++i) R[4 * (int64_t)a + i] += H->R[p][4 * (int64_t)l + i];
          if (vals) {
            const double* src = H->V[p].data() + H->loff[p][l];
            for (int64_t j = 0; j < len; ++j) vals[r0 + j] += src[j];
          }
        }
      }
      auto c3 = clk::now();
      tz[t] = std::chrono::duration<double>(c1 - c0).count();
      ta[t] = std::chrono::duration<double>(c2 - c1).count();
      tx[t] = std::chrono::duration<double>(c3 - c2b).count();
    });
  for (auto& th : pool) th.join();
  if (seconds) {
    seconds[0] = *std::max_element(tz.begin(), tz.end());
    seconds[1] = *std::max_element(ta.begin(), ta.end());
    seconds[2] = *std::max_element(tx.begin(), tx.end());
  }
  for (int v : st)
    if (v) return v;
  return 0;
}

// QoI value (ST chain) and optional derivative (FADT chain) into dMdu.
// kind 0 AvgDisp: per qp (sum_i u_i w dv)/3.  kind 1 AvgVM over element set
// qoi_set (-1: all): vm(sigma after Mixed) w dv.  elem_values [ne] optional.
int oracle_qoi(int kind, int order, int quad_order, int model, int64_t nn, int64_t ne, const double* coords,
               const int32_t* conn, const int32_t* elem_set, const double* mats, const double* x,
               const double* Fp_old, const double* eqps_old, int qoi_set, double* value, double* elem_values,
               double* dMdu) {
  (void)nn;
  const int npe = order == 1 ? 4 : 10, nq = quad_order == 1 ? 1 : 4;
  const Mat* M = reinterpret_cast<const Mat*>(mats);
  double total = 0.0;
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t* ec = conn + e * npe;
    const int set = elem_set ? elem_set[e] : 0;
    double ev = 0.0;
    if (kind == 1 && qoi_set >= 0 && set != qoi_set) {
      if (elem_values) elem_values[e] = 0.0;
      continue;
    }
    Problem P{};
    P.order = order;
    P.npe = npe;
    P.nq = nq;
    P.model = model;
    P.ne = ne;
    P.coords = coords;
    P.conn = conn;
    P.elem_set = elem_set;
    P.mats = M;
    P.x = x;
    P.Fp_old = Fp_old;
    P.eqps_old = eqps_old;
    for (int q = 0; q < nq; ++q) {
      double xi[3], w;
      quad_point(nq, q, xi, w);
      Shape s;
      eval_shape(P, e, xi, s);
      if (kind == 0) {
        // Displacement<FADT> values with seeds; elem_value += u_i w dv; /= dim
        Fad<40> qv = 0.0;
        Fad<40> u[3];
        for (int i = 0; i < 3; ++i) {
          u[i] = 0.0;
          for (int n = 0; n < npe; ++n) {
            Fad<40> un;
            un.diff(4 * n + i, 4 * npe);
            un.v = x[4 * (int64_t)ec[n] + i];
            u[i] += un * s.BF[n];
          }
        }
        for (int i = 0; i < 3; ++i) qv += u[i] * w * s.dv;
        qv /= 3.0;
        ev += qv.val();
        if (dMdu)
          for (int c = 0; c < 4 * npe; ++c) dMdu[4 * (int64_t)ec[c / 4] + c % 4] += qv.dx(c);
        continue;
      }
      // AvgVM: the model chain with FAD seeds on all element DOFs
      const Mat& m = M[set];
      const double mu = m.E / (2.0 * (1.0 + m.nu));
      const double kappa = m.E / (3.0 * (1.0 - 2.0 * m.nu));
      const double lambda = m.E * m.nu / ((1.0 + m.nu) * (1.0 - 2.0 * m.nu));
      using T = Fad<40>;
      T3<T> ug = zeros<T>();
      T pv = 0.0;
      for (int n = 0; n < npe; ++n) {
        for (int i = 0; i < 3; ++i) {
          T un;
          un.diff(4 * n + i, 4 * npe);
          un.v = x[4 * (int64_t)ec[n] + i];
          for (int j = 0; j < 3; ++j) ug(i, j) += un * s.GBF[n][j];
        }
        T pn;
        pn.diff(4 * n + 3, 4 * npe);
        pn.v = x[4 * (int64_t)ec[n] + 3];
        pv += pn * s.BF[n];
      }
      T3<T> F = ug;
      for (int i = 0; i < 3; ++i) F(i, i) += 1.0;
      T J = det(F);
      T3<T> sigma;
      const int64_t sidx = e * nq + q, npts = ne * nq;
      if (model == 0) {
        T3<T> eps;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) eps(i, j) = 0.5 * (ug(i, j) + ug(j, i));
        sigma = add(scale(2.0 * mu, eps), scale(lambda * trace(eps), eye<T>()));
      } else if (model == 1) {
        T Jm13 = 1.0 / cbrt(J);
        T Jm23 = Jm13 * Jm13;
        T Jm53 = Jm23 * Jm23 * Jm13;
        T3<T> b = mul(F, transpose(F));
        sigma = add(scale(mu * Jm53, dev(b)), scale(0.5 * kappa * (J - 1.0 / J), eye<T>()));
      } else {
        const double sq23 = std::sqrt(2.0 / 3.0);
        T Jm23 = pow(J, -2.0 / 3.0);
        T3<T> Fp;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) Fp(i, j) = Fp_old[(3 * i + j) * npts + sidx];
        T3<T> Fpinv = inverse(Fp);
        T3<T> Cpinv = mul(Fpinv, transpose(Fpinv));
        T3<T> be = scale(Jm23, mul(mul(F, Cpinv), transpose(F)));
        T3<T> s_ = scale(mu, dev(be));
        T mubar = trace(be) * mu / 3.0;
        T smag = norm(s_);
        T eqps = eqps_old[sidx];
        T f = smag - sq23 * (m.Y + m.K * eqps);
        if (f > 1.0e-12) {
          int iter = 0;
          bool converged = false;
          T H(0.0), dH(0.0), alpha(0.0), res(0.0);
          T X = 0.0, R = f;
          T dRdX = -2.0 * mubar * (1.0 + H / (3.0 * mubar));
          while ((!converged) && (iter < 30)) {
            iter++;
            X = X - R / dRdX;
            alpha = eqps + sq23 * X;
            H = m.K * alpha;
            dH = m.K;
            R = smag - (2.0 * mubar * X + sq23 * (m.Y + H));
            dRdX = -2.0 * mubar * (1.0 + dH / (3.0 * mubar));
            res = oracle::abs(R);
            if ((res < 1.0e-11) || (res.val() / m.Y < 1.0e-11) || (res.val() / f.val() < 1.0e-11)) converged = true;
            if (iter == 30) return 3;
          }
          T3<T> N = scale(1.0 / smag, s_);
          s_ = sub(s_, scale(2.0 * mubar * X, N));
        }
        sigma = add(scale(1.0 / J, s_), scale(0.5 * kappa * (J - 1.0 / J), eye<T>()));
      }
      T pbar = 0.0;
      for (int i = 0; i < 3; ++i) pbar += sigma(i, i);
      pbar /= 3.0;
      for (int i = 0; i < 3; ++i) sigma(i, i) += pv - pbar;
      T s1 = (sigma(0, 0) - sigma(1, 1)) * (sigma(0, 0) - sigma(1, 1));
      T s2 = (sigma(1, 1) - sigma(2, 2)) * (sigma(1, 1) - sigma(2, 2));
      T s3 = (sigma(2, 2) - sigma(0, 0)) * (sigma(2, 2) - sigma(0, 0));
      T s4 = sigma(0, 1) * sigma(0, 1);
      T s5 = sigma(1, 2) * sigma(1, 2);
      T s6 = sigma(2, 0) * sigma(2, 0);
      T s7 = 0.5 * (s1 + s2 + s3 + 6.0 * (s4 + s5 + s6));
      T vm = sqrt(s7);
      T qv = vm * w * s.dv;
      ev += qv.val();
      if (dMdu)
        for (int c = 0; c < 4 * npe; ++c) dMdu[4 * (int64_t)ec[c / 4] + c % 4] += qv.dx(c);
    }
    if (elem_values) elem_values[e] = ev;
    total += ev;
  }
  *value = total;
  return 0;
}

// set_resid_dbcs / set_jac_dbcs for one BC entry.
void oracle_dbc(int64_t n, const int32_t* nodes, const double* values, int comp, const double* x, double* R,
                double* dMdu, const int64_t* row_ptr, const int32_t* cols, double* vals) {
  for (int64_t k = 0; k < n; ++k) {
    int64_t row = 4 * (int64_t)nodes[k] + comp;
    double sol = x[row];
    R[row] = sol - values[k];
    if (vals) {
      if (dMdu) dMdu[row] = 0.0;
      for (int64_t j = row_ptr[row]; j < row_ptr[row + 1]; ++j) vals[j] = 0.0;
      for (int64_t j = row_ptr[row]; j < row_ptr[row + 1]; ++j)
        if (cols[j] == row) vals[j] = 1.0;
    }
  }
}

// set_tbcs, one entry (src/goal_tbcs.cpp:30-66).  Faces come as node lists:
// 3 vertices, then (quadratic) the nodes of edges v0v1, v1v2, v2v0.  The
// reference integrates each face with apf's one-point triangle rule: the
// centroid xi = (1/3, 1/3), weight 1/2, dv = |dx/dxi| = twice the area.  T is
// the traction at the centroid per face; z (optional DOF vector) makes the
// weight z_i(centroid) N_a (DisplacementAdjoint) instead of N_a.
void oracle_traction(int order, int64_t nf, const int32_t* fnodes, const double* coords, const double* T,
                     const double* z, double* R) {
  const int nfn = order == 1 ? 3 : 6;
  // triangle shape functions at the centroid, from the barycentric
  // coordinates l = (1/3, 1/3, 1/3): linear N = l_k; quadratic corner
  // l_k (2 l_k - 1), edge 4 l_j l_k
  const double l = 1.0 / 3.0;
  double N[6];
  for (int k = 0; k < 3; ++k) N[k] = order == 1 ? l : l * (2.0 * l - 1.0);
  for (int k = 3; k < 6; ++k) N[k] = 4.0 * l * l;
  const double ipw = 0.5;
  for (int64_t f = 0; f < nf; ++f) {
    const int32_t* fn = &fnodes[f * nfn];
    const double* X0 = &coords[3 * (int64_t)fn[0]];
    const double* X1 = &coords[3 * (int64_t)fn[1]];
    const double* X2 = &coords[3 * (int64_t)fn[2]];
    double a[3], b[3];
    for (int i = 0; i < 3; ++i) {
      a[i] = X1[i] - X0[i];
      b[i] = X2[i] - X0[i];
    }
    const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
    const double dv = std::sqrt(cx * cx + cy * cy + cz * cz);
    double zc[3] = {1.0, 1.0, 1.0};
    if (z)
      for (int i = 0; i < 3; ++i) {
        zc[i] = 0.0;
        for (int k = 0; k < nfn; ++k) zc[i] += z[4 * (int64_t)fn[k] + i] * N[k];
      }
    for (int k = 0; k < nfn; ++k)
      for (int i = 0; i < 3; ++i) R[4 * (int64_t)fn[k] + i] -= T[3 * f + i] * zc[i] * N[k] * ipw * dv;
  }
}

// Nested::set_coarse (in place): new vertex = mean of its two parents.
// Pinned by tests/test_oracle.py against the base mesh's P1 interpolant
// evaluated at each nested vertex by barycentric coordinates, and by affine
// fields coming back unchanged.
void oracle_set_coarse(double* z, const int32_t* edge_parents, int64_t num_old, int64_t num_new) {
  for (int64_t k = 0; k < num_new; ++k) {
    int64_t v = num_old + k, a = edge_parents[2 * k], b = edge_parents[2 * k + 1];
    for (int c = 0; c < 4; ++c) z[4 * v + c] = (z[4 * a + c] + z[4 * b + c]) * 0.5;
  }
}

// compute_error: interpolate e_u, e_p at the centroid, |sum|.
void oracle_compute_error(int order, int64_t ne, const int32_t* conn, const double* e, double* eta) {
  const int npe = order == 1 ? 4 : 10;
  for (int64_t el = 0; el < ne; ++el) {
    double lam[4] = {0.25, 0.25, 0.25, 0.25};
    double N[10];
    if (npe == 4)
      for (int a = 0; a < 4; ++a) N[a] = lam[a];
    else {
      const int ei[6] = {0, 1, 2, 0, 1, 2}, ej[6] = {1, 2, 0, 3, 3, 3};
      for (int a = 0; a < 4; ++a) N[a] = lam[a] * (2.0 * lam[a] - 1.0);
      for (int k = 0; k < 6; ++k) N[4 + k] = 4.0 * lam[ei[k]] * lam[ej[k]];
    }
    double u[3] = {0, 0, 0}, p = 0.0;
    for (int a = 0; a < npe; ++a) {
      int64_t b = 4 * (int64_t)conn[el * npe + a];
      for (int d = 0; d < 3; ++d) u[d] += N[a] * e[b + d];
      p += N[a] * e[b + 3];
    }
    double total = 0.0;
    for (int d = 0; d < 3; ++d) total += u[d];
    total += p;
    eta[el] = std::fabs(total);
  }
}

double oracle_sum_contribs(int64_t nv, const int32_t* vertex_nodes, const double* e) {
  double sum = 0.0;
  for (int64_t k = 0; k < nv; ++k) {
    int64_t b = 4 * (int64_t)vertex_nodes[k];
    double tmp = 0.0;
    for (int d = 0; d < 3; ++d) tmp += e[b + d];
    tmp += e[b + 3];
    sum += std::fabs(tmp);
  }
  return sum;
}

// Nested::set_error: base[parent[child]] += child value, in child order.
void oracle_set_error(int64_t num_fine, const int32_t* parent, const double* child, int64_t num_base, double* base) {
  for (int64_t E = 0; E < num_base; ++E) base[E] = 0.0;
  for (int64_t c = 0; c < num_fine; ++c) base[parent[c]] += child[c];
}

}  // extern "C"
