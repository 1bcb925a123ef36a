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
