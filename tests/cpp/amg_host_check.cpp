// Host checks of the multigrid set-up's dense helpers (goal_amd/csrc/goal_amg.cpp).
// Reads from stdin: a count, then per case "n" and n*n row-major entries of an
// upper Hessenberg matrix; prints per case n lines "re im" of its eigenvalues
// from goal::hessenberg_eigenvalues, then for the same eigenvalues the
// Chebyshev ellipse "theta c2" of goal::chebyshev_ellipse.
#include <cstdio>
#include <vector>

#include "goal_amg.hpp"

int main() {
  int cases = 0;
  if (std::scanf("%d", &cases) != 1) return 2;
  for (int c = 0; c < cases; ++c) {
    int n = 0;
    if (std::scanf("%d", &n) != 1 || n < 1 || n > 64) return 2;
    std::vector<double> h((size_t)n * n);
    for (auto& v : h)
      if (std::scanf("%lf", &v) != 1) return 2;
    const auto ev = goal::hessenberg_eigenvalues(n, h);
    for (const auto& z : ev) std::printf("%.17g %.17g\n", z.real(), z.imag());
    double theta = 0.0, c2 = 0.0;
    goal::chebyshev_ellipse(ev, &theta, &c2);
    std::printf("%.17g %.17g\n", theta, c2);
  }
  return 0;
}
