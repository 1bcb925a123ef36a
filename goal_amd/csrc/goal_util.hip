// Streaming helper kernels: fills, copies, axpy, deterministic reductions,
// CSR expansion from the node adjacency, Dirichlet row fix-up and boundary
// loads.
#include "goal_kernels.hpp"

namespace goal {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int per_thread = 1) {
  int64_t g = (n + (int64_t)kBlock * per_thread - 1) / ((int64_t)kBlock * per_thread);
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

__global__ __launch_bounds__(kBlock) void fill_kernel(double* p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = v;
}

__global__ __launch_bounds__(kBlock) void copy_kernel(double* d, const double* s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) d[i] = s[i];
}

__global__ __launch_bounds__(kBlock) void axpy_kernel(double* y, const double* x, double a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] += a * x[i];
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

__global__ __launch_bounds__(kBlock) void reduce1_kernel(int op, const double* x, const double* y, int64_t n,
                                                         double* scratch) {
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double v = x[i];
    acc += op == 0 ? v : op == 1 ? v * v : op == 2 ? v * y[i] : fabs(v);
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) scratch[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void reduce2_kernel(const double* scratch, int nb, double* out) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) acc += scratch[i];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) out[0] = s;
}

// One wavefront per node: the node's 4 DOF rows are
//   cols = { 4*b + j : b in adj(a), j = 0..3 },  row (a,i) starts at 16*adj_ptr[a] + 4*i*deg(a)
__global__ __launch_bounds__(kBlock) void expand_csr_kernel(const int64_t* adj_ptr, const int32_t* adj, int64_t nn,
                                                            int64_t* row_ptr, int32_t* cols) {
  const int lane = threadIdx.x % 64;
  for (int64_t a = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / 64; a < nn; a += (int64_t)gridDim.x * (kBlock / 64)) {
    const int64_t p0 = adj_ptr[a], deg = adj_ptr[a + 1] - p0;
    const int64_t base = 16 * p0;
    if (lane < 4) row_ptr[4 * a + lane] = base + 4 * lane * deg;
    if (a == nn - 1 && lane == 0) row_ptr[4 * nn] = 16 * adj_ptr[nn];
    const int64_t per_row = 4 * deg;
    for (int64_t t = lane; t < 4 * per_row; t += 64) {
      const int64_t i = t / per_row, r = t % per_row;
      cols[base + i * per_row + r] = 4 * adj[p0 + r / 4] + (int32_t)(r % 4);
    }
  }
}

// One wavefront per Dirichlet node (src/goal_dbcs.cpp:39-99):
//   R[row] = u[row] - g;  if vals: row := 0, diagonal := 1;  if dMdu: dMdu[row] = 0
__global__ __launch_bounds__(kBlock) void dbc_kernel(const int32_t* nodes, const double* values, int64_t n, int comp,
                                                     const double* x, double* R, double* dMdu,
                                                     const int64_t* row_ptr, const int32_t* cols, double* vals) {
  const int lane = threadIdx.x % 64;
  for (int64_t k = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / 64; k < n; k += (int64_t)gridDim.x * (kBlock / 64)) {
    const int64_t row = 4 * (int64_t)nodes[k] + comp;
    if (lane == 0) {
      if (R) R[row] = x[row] - values[k];
      if (dMdu) dMdu[row] = 0.0;
    }
    if (vals) {
      const int64_t b = row_ptr[row], e = row_ptr[row + 1];
      for (int64_t j = b + lane; j < e; j += 64) vals[j] = (cols[j] == row) ? 1.0 : 0.0;
    }
  }
}

// Boundary loads (src/goal_tbcs.cpp:30-66): R[rows[k]] -= values[k].  The
// host sums every face's share of a row first, so rows are distinct.
__global__ __launch_bounds__(kBlock) void sub_rows_kernel(const int64_t* rows, const double* values, int64_t n,
                                                          double* R) {
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock)
    R[rows[k]] -= values[k];
}

__global__ __launch_bounds__(kBlock) void take_rows_kernel(const int64_t* rows, const double* src, int64_t n,
                                                           double* out) {
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock)
    out[k] = src[rows[k]];
}

}  // namespace

hipError_t launch_take_rows(const int64_t* rows, const double* src, int64_t n, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(take_rows_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, rows, src, n, out);
  return hipGetLastError();
}

hipError_t launch_sub_rows(const int64_t* rows, const double* values, int64_t n, double* R, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(sub_rows_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, rows, values, n, R);
  return hipGetLastError();
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n, 4)), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_copy(double* dst, const double* src, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n, 4)), dim3(kBlock), 0, s, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_axpy(double* y, const double* x, double a, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n, 4)), dim3(kBlock), 0, s, y, x, a, n);
  return hipGetLastError();
}

hipError_t launch_reduce(int op, const double* x, const double* y, int64_t n, double* scratch, double* out,
                         hipStream_t s) {
  int nb = (int)std::min<int64_t>(kReduceBlocks, std::max<int64_t>(1, (n + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(reduce1_kernel, dim3(nb), dim3(kBlock), 0, s, op, x, y, n, scratch);
  hipLaunchKernelGGL(reduce2_kernel, dim3(1), dim3(kBlock), 0, s, scratch, nb, out);
  return hipGetLastError();
}

hipError_t launch_expand_csr(const int64_t* node_adj_ptr, const int32_t* node_adj, int64_t num_nodes,
                             int64_t* row_ptr, int32_t* cols, hipStream_t s) {
  if (num_nodes <= 0) return hipSuccess;
  hipLaunchKernelGGL(expand_csr_kernel, dim3(grid_for(num_nodes * 64)), dim3(kBlock), 0, s, node_adj_ptr, node_adj,
                     num_nodes, row_ptr, cols);
  return hipGetLastError();
}

hipError_t launch_dbc(const int32_t* nodes, const double* values, int64_t n, int comp, const double* x,
                      double* R, double* dMdu, const int64_t* row_ptr, const int32_t* cols, double* vals,
                      hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(dbc_kernel, dim3(grid_for(n * 64)), dim3(kBlock), 0, s, nodes, values, n, comp, x, R, dMdu,
                     row_ptr, cols, vals);
  return hipGetLastError();
}

}  // namespace goal
