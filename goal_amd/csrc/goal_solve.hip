// Kernels of the device linear solve (GMRES preconditioned by aggregation
// multigrid or by node block-Jacobi) that stands in for goal's Belos/MueLu
// solve (src/goal_linear_solve.cpp:22-90).  FP64; every reduction has a fixed
// order, so a solve is bitwise reproducible.

#include "goal_kernels.hpp"

namespace goal {
namespace {

constexpr int kSolveThreads = 256;

// y = A x for a CSR matrix, one wavefront per row.
__global__ __launch_bounds__(kSolveThreads) void spmv_kernel(int64_t nrows, const int64_t* rp, const int32_t* cols,
                                                             const double* vals, const double* x, double* y) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (kSolveThreads / 64) + (threadIdx.x >> 6);
  if (row >= nrows) return;  // uniform per wave
  double s = 0.0;
  for (int64_t k = rp[row] + lane; k < rp[row + 1]; k += 64) s += vals[k] * x[cols[k]];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[row] = s;
}

// Gauss-Jordan inverse with partial pivoting of the 4x4 block in M[:, 0:4]
// (M[:, 4:8] holds the identity); writes it to out, or the identity and
// status bit 0 for a zero pivot.
__device__ inline void invert4(double (&M)[4][8], double* out, int* status) {
  bool ok = true;
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabs(M[r][c]) > fabs(M[p][c])) p = r;
    if (M[p][c] == 0.0) {
      ok = false;
      break;
    }
    if (p != c)
      for (int j = 0; j < 8; ++j) {
        const double t = M[c][j];
        M[c][j] = M[p][j];
        M[p][j] = t;
      }
    const double inv = 1.0 / M[c][c];
    for (int j = 0; j < 8; ++j) M[c][j] *= inv;
    for (int r = 0; r < 4; ++r)
      if (r != c) {
        const double f = M[r][c];
        for (int j = 0; j < 8; ++j) M[r][j] -= f * M[c][j];
      }
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[4 * i + j] = ok ? M[i][4 + j] : (i == j ? 1.0 : 0.0);
  if (!ok) atomicOr(status, 1);
}

// Inverse of each node's 4x4 diagonal block.  The block's columns 4a..4a+3
// sit contiguously in each of the node's rows.
__global__ __launch_bounds__(kSolveThreads) void block_inverse_kernel(int64_t nn, const int64_t* rp,
                                                                      const int32_t* cols, const double* vals,
                                                                      double* binv, int* status) {
  const int64_t a = (int64_t)blockIdx.x * kSolveThreads + threadIdx.x;
  if (a >= nn) return;
  double M[4][8];
  for (int i = 0; i < 4; ++i) {
    const int64_t row = 4 * a + i;
    int64_t lo = rp[row], hi = rp[row + 1];
    const int32_t c0 = (int32_t)(4 * a);
    while (lo < hi) {  // first column >= 4a
      const int64_t mid = (lo + hi) >> 1;
      if (cols[mid] < c0) lo = mid + 1;
      else hi = mid;
    }
    for (int j = 0; j < 4; ++j) {
      const int64_t k = lo + j;
      M[i][j] = (k < rp[row + 1] && cols[k] == c0 + j) ? vals[k] : 0.0;
      M[i][4 + j] = (i == j) ? 1.0 : 0.0;
    }
  }
  invert4(M, binv + 16 * a, status);
}

// ---------------------------------------------------- aggregation multigrid
// Coarse levels are node-block CSR (BSR) matrices: block k (row node a,
// column node bcol[k]) holds its 4x4 values row-major at bval[16 k].

// Inverse of each diagonal block of a BSR matrix (missing: identity, status 1).
__global__ __launch_bounds__(kSolveThreads) void bsr_block_inverse_kernel(int64_t nb, const int64_t* bptr,
                                                                          const int32_t* bcol, const double* bval,
                                                                          double* binv, int* status) {
  const int64_t a = (int64_t)blockIdx.x * kSolveThreads + threadIdx.x;
  if (a >= nb) return;
  int64_t lo = bptr[a], hi = bptr[a + 1];
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (bcol[mid] < a) lo = mid + 1;
    else hi = mid;
  }
  const bool has = lo < bptr[a + 1] && bcol[lo] == a;
  double M[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      M[i][j] = has ? bval[16 * lo + 4 * i + j] : 0.0;
      M[i][4 + j] = (i == j) ? 1.0 : 0.0;
    }
  invert4(M, binv + 16 * a, status);
}

// y = A x, or y = b - A x when b is given; one thread per scalar row.
__global__ __launch_bounds__(kSolveThreads) void bsr_spmv_kernel(int64_t nb, const int64_t* bptr,
                                                                 const int32_t* bcol, const double* bval,
                                                                 const double* x, const double* b, double* y) {
  const int64_t r = (int64_t)blockIdx.x * kSolveThreads + threadIdx.x;
  if (r >= 4 * nb) return;
  const int64_t a = r >> 2;
  const int i = (int)(r & 3);
  double s = 0.0;
  for (int64_t k = bptr[a]; k < bptr[a + 1]; ++k) {
    const double* v = bval + 16 * k + 4 * i;
    const double* xc = x + 4 * (int64_t)bcol[k];
    s += v[0] * xc[0] + v[1] * xc[1] + v[2] * xc[2] + v[3] * xc[3];
  }
  y[r] = b ? b[r] - s : s;
}

// y = b - A x for the scalar CSR of the finest level, one wavefront per row.
__global__ __launch_bounds__(kSolveThreads) void spmv_resid_kernel(int64_t nrows, const int64_t* rp,
                                                                   const int32_t* cols, const double* vals,
                                                                   const double* x, const double* b, double* y) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (kSolveThreads / 64) + (threadIdx.x >> 6);
  if (row >= nrows) return;  // uniform per wave
  double s = 0.0;
  for (int64_t k = rp[row] + lane; k < rp[row + 1]; k += 64) s += vals[k] * x[cols[k]];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[row] = b[row] - s;
}

// Galerkin sum P^T A P of the plain-aggregation prolongator: coarse block s,
// entry e, is the sum of entry e of the fine blocks gidx[gptr[s]..gptr[s+1])
// in list order.  Fine blocks of the finest level live in the scalar CSR
// (brow: each block's row node; its 4 rows start at rp[4 a] and are 4 deg(a)
// long, block k is columns 4 (k - rp[4 a] / 16) ..); others are BSR.
__global__ __launch_bounds__(kSolveThreads) void galerkin_kernel(int64_t ncb, const int64_t* gptr,
                                                                 const int64_t* gidx, const int32_t* brow,
                                                                 const int64_t* rp, const double* fvals,
                                                                 double* cvals) {
  const int64_t t = (int64_t)blockIdx.x * kSolveThreads + threadIdx.x;
  if (t >= 16 * ncb) return;
  const int64_t sidx = t >> 4;
  const int e = (int)(t & 15), i = e >> 2, j = e & 3;
  double sum = 0.0;
  for (int64_t p = gptr[sidx]; p < gptr[sidx + 1]; ++p) {
    const int64_t k = gidx[p];
    if (brow) {
      const int64_t a = brow[k];
      const int64_t base = rp[4 * a], len = rp[4 * a + 1] - base;  // 4 deg(a)
      sum += fvals[base + i * len + 4 * (k - base / 16) + j];
    } else {
      sum += fvals[16 * k + e];
    }
  }
  cvals[16 * sidx + e] = sum;
}

// bc[A] = sum of r over the aggregate's nodes (P^T r), one thread per (A, i).
__global__ __launch_bounds__(kSolveThreads) void restrict_kernel(int64_t nc, const int64_t* aptr,
                                                                 const int32_t* amem, const double* r,
                                                                 double* bc) {
  const int64_t t = (int64_t)blockIdx.x * kSolveThreads + threadIdx.x;
  if (t >= 4 * nc) return;
  const int64_t A = t >> 2;
  const int i = (int)(t & 3);
  double s = 0.0;
  for (int64_t m = aptr[A]; m < aptr[A + 1]; ++m) s += r[4 * (int64_t)amem[m] + i];
  bc[t] = s;
}

// x += P xc: each node takes its aggregate's correction.
__global__ __launch_bounds__(kSolveThreads) void prolong_kernel(int64_t nf, const int32_t* agg, const double* xc,
                                                                double* x) {
  const int64_t t = (int64_t)blockIdx.x * kSolveThreads + threadIdx.x;
  if (t >= 4 * nf) return;
  x[t] += xc[4 * (int64_t)agg[t >> 2] + (t & 3)];
}

// One Chebyshev step with the block-Jacobi preconditioner, node by node:
// d = c1 d + c2 Binv r;  x += d.
__global__ __launch_bounds__(kSolveThreads) void cheb_kernel(int64_t nn, const double* binv, const double* r,
                                                             double c1, double c2, double* d, double* x) {
  const int64_t a = (int64_t)blockIdx.x * kSolveThreads + threadIdx.x;
  if (a >= nn) return;
  const double* B = binv + 16 * a;
  const double r0 = r[4 * a], r1 = r[4 * a + 1], r2 = r[4 * a + 2], r3 = r[4 * a + 3];
  for (int i = 0; i < 4; ++i) {
    const double z = B[4 * i] * r0 + B[4 * i + 1] * r1 + B[4 * i + 2] * r2 + B[4 * i + 3] * r3;
    const double dn = c1 * d[4 * a + i] + c2 * z;
    d[4 * a + i] = dn;
    x[4 * a + i] += dn;
  }
}

// y = M x for a dense row-major n x n matrix (the coarsest level's inverse).
__global__ __launch_bounds__(kSolveThreads) void dense_matvec_kernel(int n, const double* M, const double* x,
                                                                     double* y) {
  const int i = blockIdx.x * kSolveThreads + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < n; ++j) s += M[(int64_t)i * n + j] * x[j];
  y[i] = s;
}

// y = Binv x, node by node.
__global__ __launch_bounds__(kSolveThreads) void block_apply_kernel(int64_t nn, const double* binv,
                                                                    const double* x, double* y) {
  const int64_t a = (int64_t)blockIdx.x * kSolveThreads + threadIdx.x;
  if (a >= nn) return;
  const double* B = binv + 16 * a;
  const double x0 = x[4 * a], x1 = x[4 * a + 1], x2 = x[4 * a + 2], x3 = x[4 * a + 3];
  for (int i = 0; i < 4; ++i) y[4 * a + i] = B[4 * i] * x0 + B[4 * i + 1] * x1 + B[4 * i + 2] * x2 + B[4 * i + 3] * x3;
}

// partial[j][block] = sum over the block's slice of V_j . w, j < k.  The
// vectors go eight at a time: each row's w is read once per eight vectors and
// the eight loads of a row are independent.
constexpr int kDotGroup = 8;

__global__ __launch_bounds__(kSolveThreads) void multi_dot_partial_kernel(int64_t n, int k, const double* V,
                                                                          int64_t ldv, const double* w,
                                                                          double* partial) {
  __shared__ double red[kSolveThreads / 64][kDotGroup];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t stride = (int64_t)gridDim.x * kSolveThreads;
  for (int j0 = 0; j0 < k; j0 += kDotGroup) {
    const int nj = min(kDotGroup, k - j0);
    double s[kDotGroup];
#pragma unroll
    for (int jj = 0; jj < kDotGroup; ++jj) s[jj] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kSolveThreads + tid; i < n; i += stride) {
      const double wi = w[i];
#pragma unroll
      for (int jj = 0; jj < kDotGroup; ++jj)
        if (jj < nj) s[jj] += V[(int64_t)(j0 + jj) * ldv + i] * wi;
    }
#pragma unroll
    for (int jj = 0; jj < kDotGroup; ++jj) {
      double t = s[jj];
      for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
      if (lane == 0) red[wv][jj] = t;
    }
    __syncthreads();
    if (tid < nj) {
      double t = 0.0;
      for (int q = 0; q < kSolveThreads / 64; ++q) t += red[q][tid];
      partial[(int64_t)(j0 + tid) * gridDim.x + blockIdx.x] = t;
    }
    __syncthreads();
  }
}

// out[j] = sum_b partial[j][b], one block per j.
__global__ __launch_bounds__(kSolveThreads) void multi_dot_final_kernel(int nb, const double* partial, double* out) {
  __shared__ double red[kSolveThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, j = blockIdx.x;
  double s = 0.0;
  for (int b = tid; b < nb; b += kSolveThreads) s += partial[(int64_t)j * nb + b];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) red[wv] = s;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int q = 0; q < kSolveThreads / 64; ++q) t += red[q];
    out[j] = t;
  }
}

// w += sign * sum_j c_j V_j, j < k.
__global__ __launch_bounds__(kSolveThreads) void multi_axpy_kernel(int64_t n, int k, const double* V, int64_t ldv,
                                                                   const double* c, double sign, double* w) {
  const int64_t i = (int64_t)blockIdx.x * kSolveThreads + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; ++j) s += c[j] * V[(int64_t)j * ldv + i];
  w[i] += sign * s;
}

// y = a * x  (y may alias x)
__global__ __launch_bounds__(kSolveThreads) void scale_kernel(int64_t n, double a, const double* x, double* y) {
  const int64_t i = (int64_t)blockIdx.x * kSolveThreads + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

// y = b - y
__global__ __launch_bounds__(kSolveThreads) void rsub_kernel(int64_t n, const double* b, double* y) {
  const int64_t i = (int64_t)blockIdx.x * kSolveThreads + threadIdx.x;
  if (i < n) y[i] = b[i] - y[i];
}

dim3 grid_for(int64_t n) { return dim3((unsigned)((n + kSolveThreads - 1) / kSolveThreads)); }

}  // namespace

int multi_dot_blocks(int64_t n) {
  const int64_t b = (n + kSolveThreads - 1) / kSolveThreads;
  return (int)std::min<int64_t>(std::max<int64_t>(b, 1), 1024);
}

hipError_t launch_spmv(int64_t nrows, const int64_t* rp, const int32_t* cols, const double* vals, const double* x,
                       double* y, hipStream_t s) {
  if (nrows == 0) return hipSuccess;
  const int64_t blocks = (nrows + kSolveThreads / 64 - 1) / (kSolveThreads / 64);
  hipLaunchKernelGGL(spmv_kernel, dim3((unsigned)blocks), dim3(kSolveThreads), 0, s, nrows, rp, cols, vals, x, y);
  return hipGetLastError();
}

hipError_t launch_block_inverse(int64_t nn, const int64_t* rp, const int32_t* cols, const double* vals,
                                double* binv, int* status, hipStream_t s) {
  if (nn == 0) return hipSuccess;
  hipLaunchKernelGGL(block_inverse_kernel, grid_for(nn), dim3(kSolveThreads), 0, s, nn, rp, cols, vals, binv, status);
  return hipGetLastError();
}

hipError_t launch_bsr_block_inverse(int64_t nb, const int64_t* bptr, const int32_t* bcol, const double* bval,
                                    double* binv, int* status, hipStream_t s) {
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(bsr_block_inverse_kernel, grid_for(nb), dim3(kSolveThreads), 0, s, nb, bptr, bcol, bval, binv,
                     status);
  return hipGetLastError();
}

hipError_t launch_bsr_spmv(int64_t nb, const int64_t* bptr, const int32_t* bcol, const double* bval, const double* x,
                           const double* b, double* y, hipStream_t s) {
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(bsr_spmv_kernel, grid_for(4 * nb), dim3(kSolveThreads), 0, s, nb, bptr, bcol, bval, x, b, y);
  return hipGetLastError();
}

hipError_t launch_spmv_resid(int64_t nrows, const int64_t* rp, const int32_t* cols, const double* vals,
                             const double* x, const double* b, double* y, hipStream_t s) {
  if (nrows == 0) return hipSuccess;
  const int64_t blocks = (nrows + kSolveThreads / 64 - 1) / (kSolveThreads / 64);
  hipLaunchKernelGGL(spmv_resid_kernel, dim3((unsigned)blocks), dim3(kSolveThreads), 0, s, nrows, rp, cols, vals, x,
                     b, y);
  return hipGetLastError();
}

hipError_t launch_galerkin(int64_t ncb, const int64_t* gptr, const int64_t* gidx, const int32_t* brow,
                           const int64_t* rp, const double* fvals, double* cvals, hipStream_t s) {
  if (ncb == 0) return hipSuccess;
  hipLaunchKernelGGL(galerkin_kernel, grid_for(16 * ncb), dim3(kSolveThreads), 0, s, ncb, gptr, gidx, brow, rp, fvals,
                     cvals);
  return hipGetLastError();
}

hipError_t launch_restrict(int64_t nc, const int64_t* aptr, const int32_t* amem, const double* r, double* bc,
                           hipStream_t s) {
  if (nc == 0) return hipSuccess;
  hipLaunchKernelGGL(restrict_kernel, grid_for(4 * nc), dim3(kSolveThreads), 0, s, nc, aptr, amem, r, bc);
  return hipGetLastError();
}

hipError_t launch_prolong(int64_t nf, const int32_t* agg, const double* xc, double* x, hipStream_t s) {
  if (nf == 0) return hipSuccess;
  hipLaunchKernelGGL(prolong_kernel, grid_for(4 * nf), dim3(kSolveThreads), 0, s, nf, agg, xc, x);
  return hipGetLastError();
}

hipError_t launch_cheb(int64_t nn, const double* binv, const double* r, double c1, double c2, double* d, double* x,
                       hipStream_t s) {
  if (nn == 0) return hipSuccess;
  hipLaunchKernelGGL(cheb_kernel, grid_for(nn), dim3(kSolveThreads), 0, s, nn, binv, r, c1, c2, d, x);
  return hipGetLastError();
}

hipError_t launch_dense_matvec(int n, const double* M, const double* x, double* y, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(dense_matvec_kernel, grid_for(n), dim3(kSolveThreads), 0, s, n, M, x, y);
  return hipGetLastError();
}

hipError_t launch_block_apply(int64_t nn, const double* binv, const double* x, double* y, hipStream_t s) {
  if (nn == 0) return hipSuccess;
  hipLaunchKernelGGL(block_apply_kernel, grid_for(nn), dim3(kSolveThreads), 0, s, nn, binv, x, y);
  return hipGetLastError();
}

hipError_t launch_multi_dot(int64_t n, int k, const double* V, int64_t ldv, const double* w, double* partial,
                            double* out, hipStream_t s) {
  if (k == 0) return hipSuccess;
  const int nb = multi_dot_blocks(n);
  hipLaunchKernelGGL(multi_dot_partial_kernel, dim3(nb), dim3(kSolveThreads), 0, s, n, k, V, ldv, w, partial);
  hipLaunchKernelGGL(multi_dot_final_kernel, dim3(k), dim3(kSolveThreads), 0, s, nb, partial, out);
  return hipGetLastError();
}

hipError_t launch_multi_axpy(int64_t n, int k, const double* V, int64_t ldv, const double* c, double sign, double* w,
                             hipStream_t s) {
  if (n == 0 || k == 0) return hipSuccess;
  hipLaunchKernelGGL(multi_axpy_kernel, grid_for(n), dim3(kSolveThreads), 0, s, n, k, V, ldv, c, sign, w);
  return hipGetLastError();
}

hipError_t launch_scale(int64_t n, double a, const double* x, double* y, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(scale_kernel, grid_for(n), dim3(kSolveThreads), 0, s, n, a, x, y);
  return hipGetLastError();
}

hipError_t launch_rsub(int64_t n, const double* b, double* y, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(rsub_kernel, grid_for(n), dim3(kSolveThreads), 0, s, n, b, y);
  return hipGetLastError();
}

}  // namespace goal
