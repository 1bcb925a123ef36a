// Forward-mode dual number of static capacity N with an active length n:
// the oracle's stand-in for Sacado::Fad::SLFad<double, GOAL_FAD_SIZE>
// (src/goal_scalar_types.hpp:8-9).  TEST INFRASTRUCTURE ONLY.
#pragma once

#include <cmath>

namespace oracle {

template <int N>
struct Fad {
  double v = 0.0;
  double d[N];
  int n = 0;

  Fad() {}
  Fad(double x) : v(x), n(0) {}
  Fad& operator=(double x) {
    v = x;
    n = 0;
    return *this;
  }
  double val() const { return v; }
  double dx(int i) const { return i < n ? d[i] : 0.0; }
  // Sacado's diff(i, n): value kept, derivative array of length n set to e_i
  void diff(int i, int len) {
    n = len;
    for (int k = 0; k < len; ++k) d[k] = 0.0;
    d[i] = 1.0;
  }
};

template <int N>
inline int len2(const Fad<N>& a, const Fad<N>& b) {
  return a.n > b.n ? a.n : b.n;
}

template <int N>
inline Fad<N> operator+(const Fad<N>& a, const Fad<N>& b) {
  Fad<N> r;
  r.v = a.v + b.v;
  r.n = len2(a, b);
  for (int i = 0; i < r.n; ++i) r.d[i] = a.dx(i) + b.dx(i);
  return r;
}
template <int N>
inline Fad<N> operator-(const Fad<N>& a, const Fad<N>& b) {
  Fad<N> r;
  r.v = a.v - b.v;
  r.n = len2(a, b);
  for (int i = 0; i < r.n; ++i) r.d[i] = a.dx(i) - b.dx(i);
  return r;
}
template <int N>
inline Fad<N> operator*(const Fad<N>& a, const Fad<N>& b) {
  Fad<N> r;
  r.v = a.v * b.v;
  r.n = len2(a, b);
  for (int i = 0; i < r.n; ++i) r.d[i] = a.dx(i) * b.v + a.v * b.dx(i);
  return r;
}
template <int N>
inline Fad<N> operator/(const Fad<N>& a, const Fad<N>& b) {
  Fad<N> r;
  r.v = a.v / b.v;
  r.n = len2(a, b);
  const double ib2 = 1.0 / (b.v * b.v);
  for (int i = 0; i < r.n; ++i) r.d[i] = (a.dx(i) * b.v - a.v * b.dx(i)) * ib2;
  return r;
}
template <int N>
inline Fad<N> operator-(const Fad<N>& a) {
  Fad<N> r;
  r.v = -a.v;
  r.n = a.n;
  for (int i = 0; i < r.n; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N> inline Fad<N> operator+(const Fad<N>& a, double b) { return a + Fad<N>(b); }
template <int N> inline Fad<N> operator+(double a, const Fad<N>& b) { return Fad<N>(a) + b; }
template <int N> inline Fad<N> operator-(const Fad<N>& a, double b) { return a - Fad<N>(b); }
template <int N> inline Fad<N> operator-(double a, const Fad<N>& b) { return Fad<N>(a) - b; }
template <int N> inline Fad<N> operator*(const Fad<N>& a, double b) {
  Fad<N> r;
  r.v = a.v * b;
  r.n = a.n;
  for (int i = 0; i < r.n; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int N> inline Fad<N> operator*(double a, const Fad<N>& b) { return b * a; }
template <int N> inline Fad<N> operator/(const Fad<N>& a, double b) { return a / Fad<N>(b); }
template <int N> inline Fad<N> operator/(double a, const Fad<N>& b) { return Fad<N>(a) / b; }
template <int N, typename T> inline Fad<N>& operator+=(Fad<N>& a, const T& b) { return a = a + b; }
template <int N, typename T> inline Fad<N>& operator-=(Fad<N>& a, const T& b) { return a = a - b; }
template <int N, typename T> inline Fad<N>& operator*=(Fad<N>& a, const T& b) { return a = a * b; }
template <int N, typename T> inline Fad<N>& operator/=(Fad<N>& a, const T& b) { return a = a / b; }

template <int N> inline bool operator>(const Fad<N>& a, double b) { return a.v > b; }
template <int N> inline bool operator<(const Fad<N>& a, double b) { return a.v < b; }

template <int N>
inline Fad<N> chain(const Fad<N>& a, double fv, double fp) {
  Fad<N> r;
  r.v = fv;
  r.n = a.n;
  for (int i = 0; i < r.n; ++i) r.d[i] = fp * a.d[i];
  return r;
}
template <int N> inline Fad<N> sqrt(const Fad<N>& a) {
  double s = std::sqrt(a.v);
  return chain(a, s, 0.5 / s);
}
template <int N> inline Fad<N> cbrt(const Fad<N>& a) {
  double c = std::cbrt(a.v);
  return chain(a, c, c / (3.0 * a.v));
}
template <int N> inline Fad<N> pow(const Fad<N>& a, double e) {
  double p = std::pow(a.v, e);
  return chain(a, p, e * std::pow(a.v, e - 1.0));
}
template <int N> inline Fad<N> abs(const Fad<N>& a) { return a.v < 0 ? -a : a; }

inline double val(double x) { return x; }
inline double abs(double x) { return std::fabs(x); }
template <int N> inline double val(const Fad<N>& x) { return x.v; }

}  // namespace oracle
