// Host-side structure of the aggregation multigrid hierarchy (see
// goal_amg.hpp).
#include "goal_amg.hpp"

#include <algorithm>
#include <cmath>

namespace goal {

AmgCoarsening amg_coarsen(int64_t nf, const int64_t* fptr, const int32_t* fcol) {
  AmgCoarsening c;
  c.nf = nf;
  c.agg.assign(nf, -1);
  int32_t na = 0;
  // roots: every neighbour still free
  for (int64_t i = 0; i < nf; ++i) {
    if (c.agg[i] >= 0) continue;
    bool free = true;
    for (int64_t k = fptr[i]; k < fptr[i + 1] && free; ++k) free = c.agg[fcol[k]] < 0;
    if (!free) continue;
    for (int64_t k = fptr[i]; k < fptr[i + 1]; ++k) c.agg[fcol[k]] = na;
    c.agg[i] = na++;
  }
  // leftovers: join the first aggregated neighbour, or start an aggregate
  for (int64_t i = 0; i < nf; ++i) {
    if (c.agg[i] >= 0) continue;
    int32_t a = -1;
    for (int64_t k = fptr[i]; k < fptr[i + 1] && a < 0; ++k)
      if (fcol[k] != i) a = c.agg[fcol[k]];
    c.agg[i] = a >= 0 ? a : na++;
  }
  c.nc = na;
  c.aptr.assign(na + 1, 0);
  for (int64_t i = 0; i < nf; ++i) ++c.aptr[c.agg[i] + 1];
  for (int64_t a = 0; a < na; ++a) c.aptr[a + 1] += c.aptr[a];
  c.amem.resize(nf);
  {
    std::vector<int64_t> fill(c.aptr.begin(), c.aptr.end() - 1);
    for (int64_t i = 0; i < nf; ++i) c.amem[fill[c.agg[i]]++] = (int32_t)i;
  }
  // coarse graph: the aggregates reached from each aggregate's members
  std::vector<int32_t> mark(na, -1);
  c.cptr.assign(na + 1, 0);
  std::vector<int32_t> row;
  for (int64_t a = 0; a < na; ++a) {
    row.clear();
    for (int64_t m = c.aptr[a]; m < c.aptr[a + 1]; ++m) {
      const int32_t i = c.amem[m];
      for (int64_t k = fptr[i]; k < fptr[i + 1]; ++k) {
        const int32_t b = c.agg[fcol[k]];
        if (mark[b] != (int32_t)a) {
          mark[b] = (int32_t)a;
          row.push_back(b);
        }
      }
    }
    std::sort(row.begin(), row.end());
    c.ccol.insert(c.ccol.end(), row.begin(), row.end());
    c.cptr[a + 1] = (int64_t)c.ccol.size();
  }
  // Galerkin lists: fine block (i, fcol[k]) -> coarse block (agg i, agg fcol[k])
  const int64_t nfb = fptr[nf], ncb = (int64_t)c.ccol.size();
  std::vector<int64_t> slot(nfb);
  for (int64_t i = 0; i < nf; ++i) {
    const int32_t a = c.agg[i];
    const int32_t* lo = c.ccol.data() + c.cptr[a];
    const int32_t* hi = c.ccol.data() + c.cptr[a + 1];
    for (int64_t k = fptr[i]; k < fptr[i + 1]; ++k)
      slot[k] = c.cptr[a] + (std::lower_bound(lo, hi, c.agg[fcol[k]]) - lo);
  }
  c.gptr.assign(ncb + 1, 0);
  for (int64_t k = 0; k < nfb; ++k) ++c.gptr[slot[k] + 1];
  for (int64_t s = 0; s < ncb; ++s) c.gptr[s + 1] += c.gptr[s];
  c.gidx.resize(nfb);
  {
    std::vector<int64_t> fill(c.gptr.begin(), c.gptr.end() - 1);
    for (int64_t k = 0; k < nfb; ++k) c.gidx[fill[slot[k]]++] = k;
  }
  return c;
}

bool dense_inverse(int n, std::vector<double>& a) {
  std::vector<double> inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)p * n + c])) p = r;
    if (a[(size_t)p * n + c] == 0.0) {
      a.assign((size_t)n * n, 0.0);
      for (int i = 0; i < n; ++i) a[(size_t)i * n + i] = 1.0;
      return false;
    }
    if (p != c)
      for (int j = 0; j < n; ++j) {
        std::swap(a[(size_t)c * n + j], a[(size_t)p * n + j]);
        std::swap(inv[(size_t)c * n + j], inv[(size_t)p * n + j]);
      }
    const double d = 1.0 / a[(size_t)c * n + c];
    for (int j = 0; j < n; ++j) {
      a[(size_t)c * n + j] *= d;
      inv[(size_t)c * n + j] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        a[(size_t)r * n + j] -= f * a[(size_t)c * n + j];
        inv[(size_t)r * n + j] -= f * inv[(size_t)c * n + j];
      }
    }
  }
  a.swap(inv);
  return true;
}

std::vector<std::complex<double>> hessenberg_eigenvalues(int n, const std::vector<double>& h) {
  using C = std::complex<double>;
  std::vector<C> a((size_t)n * n);
  for (size_t k = 0; k < a.size(); ++k) a[k] = h[k];
  auto A = [&](int i, int j) -> C& { return a[(size_t)i * n + j]; };
  std::vector<C> ev(n);
  std::vector<double> gc(n);
  std::vector<C> gs(n);
  int hi = n - 1, its = 0;
  while (hi >= 0) {
    if (hi == 0) {
      ev[0] = A(0, 0);
      break;
    }
    // the active block [lo, hi]: every subdiagonal inside it is non-negligible
    int lo = hi;
    while (lo > 0) {
      const double scale = std::abs(A(lo - 1, lo - 1)) + std::abs(A(lo, lo));
      if (std::abs(A(lo, lo - 1)) <= 1e-15 * (scale > 0.0 ? scale : 1.0)) break;
      --lo;
    }
    if (lo == hi || its >= 200) {  // deflate (or give up on this one after 200 steps)
      ev[hi] = A(hi, hi);
      --hi;
      its = 0;
      continue;
    }
    // Wilkinson shift: the eigenvalue of the trailing 2 x 2 nearer A(hi, hi)
    const C p = A(hi - 1, hi - 1), q = A(hi - 1, hi), r = A(hi, hi - 1), t = A(hi, hi);
    const C half = 0.5 * (p - t), disc = std::sqrt(half * half + q * r);
    C mu = t - q * r / (half + (std::abs(half + disc) >= std::abs(half - disc) ? disc : -disc));
    if (!std::isfinite(mu.real()) || !std::isfinite(mu.imag())) mu = t;
    if (its % 11 == 10) mu += std::abs(A(hi, hi - 1)) * 0.75;  // exceptional shift
    ++its;
    for (int k = lo; k <= hi; ++k) A(k, k) -= mu;
    // A - mu = QR by Givens rotations on the subdiagonal, then A = RQ + mu
    for (int k = lo; k < hi; ++k) {
      const C x = A(k, k), y = A(k + 1, k);
      const double ax = std::abs(x), nr = std::hypot(ax, std::abs(y));
      double c;
      C sg;
      if (nr == 0.0) {
        c = 1.0;
        sg = 0.0;
      } else if (ax == 0.0) {
        c = 0.0;
        sg = 1.0;
      } else {
        c = ax / nr;
        sg = (x / ax) * std::conj(y) / nr;
      }
      gc[k] = c;
      gs[k] = sg;
      for (int j = k; j <= hi; ++j) {
        const C u = A(k, j), v = A(k + 1, j);
        A(k, j) = c * u + sg * v;
        A(k + 1, j) = -std::conj(sg) * u + c * v;
      }
    }
    for (int k = lo; k < hi; ++k) {
      const double c = gc[k];
      const C sg = gs[k];
      for (int i = lo; i <= std::min(k + 2, hi); ++i) {
        const C u = A(i, k), v = A(i, k + 1);
        A(i, k) = c * u + std::conj(sg) * v;
        A(i, k + 1) = -sg * u + c * v;
      }
    }
    for (int k = lo; k <= hi; ++k) A(k, k) += mu;
  }
  return ev;
}

void chebyshev_ellipse(const std::vector<std::complex<double>>& ritz, double* theta, double* c2) {
  double top = 0.0;
  for (const auto& z : ritz) top = std::max(top, std::max(z.real(), z.real() > 0.0 ? std::abs(z) : 0.0));
  if (!(top > 0.0) || !std::isfinite(top)) {
    *theta = 0.0;
    *c2 = 0.0;
    return;
  }
  const double hi = 1.1 * top, lo = hi / 10.0, d = 0.5 * (hi + lo), a = 0.5 * (hi - lo);
  double b = 0.0;
  for (const auto& z : ritz) {
    const double u = (z.real() - d) / a;
    if (std::fabs(u) < 1.0) b = std::max(b, std::fabs(z.imag()) / std::sqrt(1.0 - u * u));
  }
  b *= 1.1;
  *theta = d;
  *c2 = a * a - b * b;
}

}  // namespace goal
