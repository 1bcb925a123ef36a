// Nested-adjoint DWR field transfers and element error indicators.
//   set_coarse   Nested::set_coarse          src/goal_nested.cpp:369-393
//   subtract     NestedAdjoint::subtract     src/goal_nested_adjoint.cpp:182-197
//   compute_error                            src/goal_error.cpp:7-35
//   sum_contribs                             src/goal_error.cpp:37-56
//   set_error    Nested::set_error           src/goal_nested.cpp:395-412
// All are streaming kernels over nodal fields laid out like the solution
// (4 doubles per node: u_x u_y u_z p) or over element arrays.
#include "goal_kernels.hpp"

namespace goal {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

// new vertex := mean of its two parent vertices, all 4 components
__global__ __launch_bounds__(kBlock) void set_coarse_kernel(double* z, const int32_t* ep, int64_t num_old,
                                                            int64_t num_new) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < num_new * 4; t += (int64_t)gridDim.x * kBlock) {
    const int64_t k = t >> 2;
    const int c = (int)(t & 3);
    const int64_t p0 = ep[2 * k], p1 = ep[2 * k + 1];
    z[4 * (num_old + k) + c] = (z[4 * p0 + c] + z[4 * p1 + c]) * 0.5;
  }
}

__global__ __launch_bounds__(kBlock) void subtract_kernel(const double* a, const double* b, double* o, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) o[i] = a[i] - b[i];
}

// eta_e = | sum_d e_u,d(xi_c) + e_p(xi_c) | at the centroid xi_c.  Centroid
// values of the shape functions: P1 1/4; P2 vertices -1/8, edges 1/4.
template <int NN>
__global__ __launch_bounds__(kBlock) void compute_error_kernel(const int32_t* conn, int64_t ne, const double* e,
                                                               double* eta) {
  for (int64_t el = (int64_t)blockIdx.x * kBlock + threadIdx.x; el < ne; el += (int64_t)gridDim.x * kBlock) {
    double u[4] = {0.0, 0.0, 0.0, 0.0};
    for (int n = 0; n < NN; ++n) {
      const double N = (NN == 4) ? 0.25 : (n < 4 ? -0.125 : 0.25);
      const int64_t b = 4 * (int64_t)conn[el * NN + n];
      for (int c = 0; c < 4; ++c) u[c] += e[b + c] * N;
    }
    double total = 0.0;
    for (int d = 0; d < 3; ++d) total += u[d];
    total += u[3];
    eta[el] = fabs(total);
  }
}

__device__ inline double block_sum(double v) {
  __shared__ double part[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) part[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kBlock / 64; ++k) s += part[k];
  __syncthreads();
  return s;
}

// sum over mesh vertices of | sum_d e_u,d + e_p |, two deterministic passes
__global__ __launch_bounds__(kBlock) void contribs1_kernel(const int32_t* vn, int64_t nv, const double* e,
                                                           double* scratch) {
  double acc = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < nv; k += (int64_t)gridDim.x * kBlock) {
    const int64_t b = 4 * (int64_t)vn[k];
    double tmp = 0.0;
    for (int d = 0; d < 3; ++d) tmp += e[b + d];
    tmp += e[b + 3];
    acc += fabs(tmp);
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) scratch[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void contribs2_kernel(const double* scratch, int nb, double* out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) acc += scratch[i];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) out[0] = s;
}

// base element E owns children 8E..8E+7
__global__ __launch_bounds__(kBlock) void set_error_kernel(const double* child, int64_t nb, double* base) {
  for (int64_t E = (int64_t)blockIdx.x * kBlock + threadIdx.x; E < nb; E += (int64_t)gridDim.x * kBlock) {
    double v = 0.0;
    for (int c = 0; c < 8; ++c) v += child[8 * E + c];
    base[E] = v;
  }
}

// child QP state := parent QP state (same quadrature point index)
__global__ __launch_bounds__(kBlock) void transfer_state_kernel(const double* base, double* fine, int ncomp,
                                                                int64_t num_fine, int nq, int64_t base_pts) {
  const int64_t fine_pts = num_fine * nq;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < fine_pts * ncomp;
       t += (int64_t)gridDim.x * kBlock) {
    const int64_t k = t / fine_pts, pt = t % fine_pts;
    const int64_t child = pt / nq, q = pt % nq;
    fine[k * fine_pts + pt] = base[k * base_pts + (child >> 3) * nq + q];
  }
}

}  // namespace

hipError_t launch_transfer_state(const double* base, double* fine, int ncomp, int64_t num_fine, int nq,
                                 int64_t base_pts, hipStream_t s) {
  if (num_fine <= 0) return hipSuccess;
  hipLaunchKernelGGL(transfer_state_kernel, dim3(grid_for(num_fine * nq * ncomp)), dim3(kBlock), 0, s, base, fine,
                     ncomp, num_fine, nq, base_pts);
  return hipGetLastError();
}

// ---- uniform refinement on the device (the nested mesh, 8 children per tet)
// Edge numbering: the edges (lo, hi), lo < hi, in lexicographic order, as the
// host refinement numbers them.  With the base mesh's sorted node adjacency
// (the DOF CSR: row 4a holds 4b for every neighbour b of a, ascending), edge
// (lo, hi) is number base[lo] + (rank of hi among the neighbours above lo),
// base[] being the exclusive prefix sum of the neighbour counts above each node.
__device__ inline int64_t upper_begin(const int64_t* row_ptr, const int32_t* cols, int64_t a, int64_t* deg) {
  const int64_t b0 = row_ptr[4 * a];
  *deg = (row_ptr[4 * a + 1] - b0) / 4;
  // first neighbour above a (a itself is a neighbour)
  int64_t lo = 0, hi = *deg;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (cols[b0 + 4 * mid] <= 4 * a) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void upper_degree_kernel(const int64_t* row_ptr, const int32_t* cols, int64_t nv,
                                                              int64_t* ud) {
  for (int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x; a < nv; a += (int64_t)gridDim.x * kBlock) {
    int64_t deg;
    const int64_t k = upper_begin(row_ptr, cols, a, &deg);
    ud[a] = deg - k;
  }
}

__global__ __launch_bounds__(kBlock) void refine_edges_kernel(const int64_t* row_ptr, const int32_t* cols,
                                                              const int64_t* ebase, const double* X, int64_t nv,
                                                              int32_t* edge_parents, double* fX) {
  for (int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x; a < nv; a += (int64_t)gridDim.x * kBlock) {
    int64_t deg;
    const int64_t k0 = upper_begin(row_ptr, cols, a, &deg);
    const int64_t b0 = row_ptr[4 * a];
    for (int64_t k = k0; k < deg; ++k) {
      const int64_t b = cols[b0 + 4 * k] / 4, id = ebase[a] + (k - k0);
      edge_parents[2 * id] = (int32_t)a;
      edge_parents[2 * id + 1] = (int32_t)b;
      for (int c = 0; c < 3; ++c) fX[3 * (nv + id) + c] = 0.5 * (X[3 * a + c] + X[3 * b + c]);
    }
  }
}

__device__ inline double tet_det4(const double* x0, const double* x1, const double* x2, const double* x3) {
  double u[3], v[3], w[3];
  for (int i = 0; i < 3; ++i) {
    u[i] = x1[i] - x0[i];
    v[i] = x2[i] - x0[i];
    w[i] = x3[i] - x0[i];
  }
  return u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) + u[2] * (v[0] * w[1] - v[1] * w[0]);
}

__global__ __launch_bounds__(kBlock) void refine_elems_kernel(const int32_t* conn, const int64_t* row_ptr,
                                                              const int32_t* cols, const int64_t* ebase,
                                                              const double* fX, int64_t ne, int64_t nv,
                                                              int32_t* fconn, int32_t* status) {
  constexpr int E[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};  // edges 01,12,20,03,13,23
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < ne; e += (int64_t)gridDim.x * kBlock) {
    int32_t v[4], m[6];
    for (int k = 0; k < 4; ++k) v[k] = conn[4 * e + k];
    for (int ed = 0; ed < 6; ++ed) {
      const int32_t p = v[E[ed][0]], q = v[E[ed][1]];
      const int64_t lo = p < q ? p : q, hi = p < q ? q : p;
      int64_t deg;
      const int64_t k0 = upper_begin(row_ptr, cols, lo, &deg);
      const int64_t b0 = row_ptr[4 * lo];
      int64_t l = k0, h = deg;  // hi is among the neighbours above lo
      while (l < h) {
        const int64_t mid = (l + h) / 2;
        if (cols[b0 + 4 * mid] < 4 * hi) l = mid + 1;
        else h = mid;
      }
      if (l >= deg || cols[b0 + 4 * l] != 4 * hi) {  // not an edge of the graph: report, stay in bounds
        atomicOr(status, 1);
        l = k0;
      }
      m[ed] = (int32_t)(nv + ebase[lo] + (l - k0));
    }
    const int32_t kids[8][4] = {{v[0], m[0], m[2], m[3]}, {m[0], v[1], m[1], m[4]}, {m[2], m[1], v[2], m[5]},
                                {m[3], m[4], m[5], v[3]}, {m[0], m[2], m[3], m[4]}, {m[0], m[2], m[1], m[4]},
                                {m[2], m[3], m[4], m[5]}, {m[2], m[1], m[4], m[5]}};
    for (int c = 0; c < 8; ++c) {
      int32_t cc[4] = {kids[c][0], kids[c][1], kids[c][2], kids[c][3]};
      const double d = tet_det4(&fX[3 * (int64_t)cc[0]], &fX[3 * (int64_t)cc[1]], &fX[3 * (int64_t)cc[2]],
                                &fX[3 * (int64_t)cc[3]]);
      if (d < 0) {
        const int32_t t = cc[2];
        cc[2] = cc[3];
        cc[3] = t;
      }
      for (int k = 0; k < 4; ++k) fconn[4 * (8 * e + c) + k] = cc[k];
    }
  }
}

hipError_t launch_upper_degree(const int64_t* row_ptr, const int32_t* cols, int64_t nv, int64_t* ud, hipStream_t s) {
  if (nv <= 0) return hipSuccess;
  hipLaunchKernelGGL(upper_degree_kernel, dim3(grid_for(nv)), dim3(kBlock), 0, s, row_ptr, cols, nv, ud);
  return hipGetLastError();
}

hipError_t launch_refine(const int32_t* conn, const double* X, const int64_t* row_ptr, const int32_t* cols,
                         const int64_t* ebase, int64_t ne, int64_t nv, int32_t* edge_parents, double* fX,
                         int32_t* fconn, int32_t* status, hipStream_t s) {
  if (nv <= 0 || ne <= 0) return hipSuccess;
  hipLaunchKernelGGL(refine_edges_kernel, dim3(grid_for(nv)), dim3(kBlock), 0, s, row_ptr, cols, ebase, X, nv,
                     edge_parents, fX);
  hipLaunchKernelGGL(refine_elems_kernel, dim3(grid_for(ne)), dim3(kBlock), 0, s, conn, row_ptr, cols, ebase, fX,
                     ne, nv, fconn, status);
  return hipGetLastError();
}

hipError_t launch_set_coarse(double* z, const int32_t* edge_parents, int64_t num_old, int64_t num_new,
                             hipStream_t s) {
  if (num_new <= 0) return hipSuccess;
  hipLaunchKernelGGL(set_coarse_kernel, dim3(grid_for(num_new * 4)), dim3(kBlock), 0, s, z, edge_parents, num_old,
                     num_new);
  return hipGetLastError();
}

hipError_t launch_subtract(const double* a, const double* b, double* out, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(subtract_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, a, b, out, n);
  return hipGetLastError();
}

hipError_t launch_compute_error(int order, const int32_t* conn, int64_t ne, const double* e, double* eta,
                                hipStream_t s) {
  if (ne <= 0) return hipSuccess;
  if (order == 1)
    hipLaunchKernelGGL(compute_error_kernel<4>, dim3(grid_for(ne)), dim3(kBlock), 0, s, conn, ne, e, eta);
  else
    hipLaunchKernelGGL(compute_error_kernel<10>, dim3(grid_for(ne)), dim3(kBlock), 0, s, conn, ne, e, eta);
  return hipGetLastError();
}

hipError_t launch_sum_contribs(const int32_t* vertex_nodes, int64_t nv, const double* e, double* scratch,
                               double* out, hipStream_t s) {
  int nb = (int)std::min<int64_t>(kReduceBlocks, std::max<int64_t>(1, (nv + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(contribs1_kernel, dim3(nb), dim3(kBlock), 0, s, vertex_nodes, nv, e, scratch);
  hipLaunchKernelGGL(contribs2_kernel, dim3(1), dim3(kBlock), 0, s, scratch, nb, out);
  return hipGetLastError();
}

hipError_t launch_set_error(const double* eta_child, int64_t num_base, double* eta_base, hipStream_t s) {
  if (num_base <= 0) return hipSuccess;
  hipLaunchKernelGGL(set_error_kernel, dim3(grid_for(num_base)), dim3(kBlock), 0, s, eta_child, num_base, eta_base);
  return hipGetLastError();
}

}  // namespace goal
