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

}  // namespace goal
