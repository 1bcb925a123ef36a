// Structure of an aggregation multigrid hierarchy over node-block (4x4) sparse
// matrices: the part that depends only on the sparsity graph, built once per
// discretization on the host.  The values (coarse matrices, smoother
// parameters) are recomputed on the device for every matrix.
//
// It stands in for the MueLu preconditioner of goal's linear solve
// (src/goal_linear_solve.cpp:53-70, "number of equations" = 4 in the example
// inputs): aggregates of nodes, each node's 4 equations kept together.
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

namespace goal {

// One coarsening step: fine node graph (CSR over nodes, sorted columns, the
// diagonal included) -> aggregates -> coarse node graph.
struct AmgCoarsening {
  int64_t nf = 0, nc = 0;
  std::vector<int32_t> agg;    // [nf] fine node -> aggregate (coarse node)
  std::vector<int64_t> aptr;   // [nc + 1] aggregate -> its fine nodes
  std::vector<int32_t> amem;   //          (ascending)
  std::vector<int64_t> cptr;   // [nc + 1] coarse graph, columns sorted
  std::vector<int32_t> ccol;
  // Galerkin sum of the plain-aggregation prolongator, P^T A P: coarse block
  // s is the sum of the fine blocks gidx[gptr[s] .. gptr[s+1]) (fine block
  // index into the fine graph, ascending, so the sum has a fixed order).
  std::vector<int64_t> gptr;   // [ccol.size() + 1]
  std::vector<int64_t> gidx;
};

// Greedy aggregation in node order: a node whose neighbours are all free
// becomes the root of an aggregate of itself and its neighbours; nodes left
// over join the aggregate of their first aggregated neighbour (or start one).
AmgCoarsening amg_coarsen(int64_t nf, const int64_t* fptr, const int32_t* fcol);

// In-place Gauss-Jordan inverse of a dense n x n row-major matrix with
// partial pivoting.  Returns false (and leaves the identity) if singular.
bool dense_inverse(int n, std::vector<double>& a);

// Eigenvalues of a real upper Hessenberg n x n matrix (row-major, n small:
// the Arnoldi matrix of a smoother's spectrum estimate) by the shifted QR
// algorithm in complex arithmetic: Givens rotations, Wilkinson shifts,
// deflation at negligible subdiagonal entries.
std::vector<std::complex<double>> hessenberg_eigenvalues(int n, const std::vector<double>& h);

// The Chebyshev smoother's ellipse from Ritz values of Binv A: real extent
// [hi / 10, hi] with hi = 1.1 max(Re, |.|) over the Ritz values, imaginary
// semi-axis just large enough (times 1.1) to enclose the Ritz values whose
// real part lies inside that extent.  Returns the centre theta and the squared
// focal half-distance c2 = a^2 - b^2 (negative when the ellipse is taller
// than wide); theta = 0 when no Ritz value has a positive real part.
void chebyshev_ellipse(const std::vector<std::complex<double>>& ritz, double* theta, double* c2);

}  // namespace goal
