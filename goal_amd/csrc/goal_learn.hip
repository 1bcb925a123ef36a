// Learned element error indicator: per-element features of a primal state
// and a 12 -> 64 -> 64 -> 1 ReLU network (BF16 weights, FP32 biases and
// accumulation) with its forward pass, mean-squared-error gradient and Adam
// update.  The network predicts log10(eta_E / E) where eta_E is the nested-
// adjoint indicator of Nested::set_error (src/goal_nested.cpp:395-412); it
// has no counterpart in the reference.
//
// Every reduction runs in a fixed order (per-block partial sums, then one
// ordered pass), so a training run is bitwise reproducible.

#include "goal_kernels.hpp"

namespace goal {
namespace {

constexpr int kIn = kMlpIn, kHid = kMlpHidden;
constexpr int oW1 = kMlpOffW1, oB1 = kMlpOffB1, oW2 = kMlpOffW2, oB2 = kMlpOffB2, oW3 = kMlpOffW3,
              oB3 = kMlpOffB3, kP = kMlpParams;

// ------------------------------------------------------------- features

constexpr int kFeatThreads = 256;

__device__ inline double tet_volume(const double* X, const int32_t* v) {
  const double* a = X + 3 * v[0];
  const double* b = X + 3 * v[1];
  const double* c = X + 3 * v[2];
  const double* d = X + 3 * v[3];
  const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double e3[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
  return (e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
          e1[2] * (e2[0] * e3[1] - e2[1] * e3[0])) /
         6.0;
}

// Reference-configuration gradient of the vertex-interpolated displacement:
// H[i][j] = sum_k (U[k+1][i] - U[0][i]) Ji[j][k], Ji the inverse of the
// matrix whose rows are the edges 01, 02, 03.
__device__ inline void linear_grad(const double X[4][3], const double U[4][3], double H[9]) {
  double J[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) J[r][c] = X[r + 1][c] - X[0][c];
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double Ji[3][3];
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += (U[k + 1][i] - U[0][i]) * Ji[j][k];
      H[3 * i + j] = s;
    }
}

__device__ inline void load_vertices(const FeatureArgs& a, int64_t e, double X[4][3], double U[4][3]) {
  const int32_t* v = a.conn + e * a.npe;
  for (int i = 0; i < 4; ++i)
    for (int c = 0; c < 3; ++c) {
      X[i][c] = a.coords[3 * (int64_t)v[i] + c];
      U[i][c] = a.x[4 * (int64_t)v[i] + c];
    }
}

// Per element: volume and grad u, into elem_scratch[e][0..9].
__global__ __launch_bounds__(kFeatThreads) void elem_geometry_kernel(FeatureArgs a) {
  const int64_t e = (int64_t)blockIdx.x * kFeatThreads + threadIdx.x;
  if (e >= a.num_elems) return;
  double X[4][3], U[4][3], H[9];
  load_vertices(a, e, X, U);
  linear_grad(X, U, H);
  double* o = a.elem_scratch + 10 * e;
  o[0] = tet_volume(a.coords, a.conn + e * a.npe);
  for (int k = 0; k < 9; ++k) o[1 + k] = H[k];
}

// Volume-weighted nodal means of the element von Mises stress and of grad u:
// node_scratch[n] = (sum int_e vm, sum |e| H_e) / sum |e| over the incident
// elements in ascending order.
__global__ __launch_bounds__(kFeatThreads) void recover_kernel(FeatureArgs a) {
  const int64_t n = (int64_t)blockIdx.x * kFeatThreads + threadIdx.x;
  if (n >= a.num_nodes) return;
  double s[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  double w = 0.0;
  for (int64_t k = a.inc_ptr[n]; k < a.inc_ptr[n + 1]; ++k) {
    const int64_t e = a.inc[k] / a.npe;
    const double* g = a.elem_scratch + 10 * e;
    s[0] += a.elem_vm[e];
    for (int c = 0; c < 9; ++c) s[1 + c] += g[0] * g[1 + c];
    w += g[0];
  }
  double* o = a.node_scratch + 10 * n;
  for (int c = 0; c < 10; ++c) o[c] = w > 0.0 ? s[c] / w : 0.0;
}

__device__ inline float flog10(double v) { return (float)log10(v); }

// One element per thread.  Features (FP32):
//  0 log10 h (RMS edge length)        6 log10(vm / E + 1e-12)
//  1-3 centroid x, y, z               7 log10(|vm - vm*| / E + 1e-12)
//  4 log10(|x_c - qoi_x| + h)         8 log10(|grad u|_F + 1e-12)
//  5 1 if x_c < qoi_x else 0          9 log10(|grad u - (grad u)*|_F + 1e-12)
//                                     10 nu
//                                     11 log10(mean point eqps + 1e-9)
// vm is the element mean von Mises stress; a starred value is the mean over
// the element's four vertices of the recovered nodal value.  grad u uses the
// vertex displacements (linear interpolation).
__global__ __launch_bounds__(kFeatThreads) void features_kernel(FeatureArgs a) {
  const int64_t e = (int64_t)blockIdx.x * kFeatThreads + threadIdx.x;
  if (e >= a.num_elems) return;
  const int32_t* v = a.conn + e * a.npe;
  double X[4][3];
  for (int i = 0; i < 4; ++i)
    for (int c = 0; c < 3; ++c) X[i][c] = a.coords[3 * (int64_t)v[i] + c];
  const double* ge = a.elem_scratch + 10 * e;
  double star[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    const double* nv = a.node_scratch + 10 * (int64_t)v[i];
    for (int c = 0; c < 10; ++c) star[c] += 0.25 * nv[c];
  }
  double h2 = 0.0, d2 = 0.0;
  for (int k = 0; k < 9; ++k) {
    h2 += ge[1 + k] * ge[1 + k];
    d2 += (ge[1 + k] - star[1 + k]) * (ge[1 + k] - star[1 + k]);
  }
  double l2 = 0.0;
  for (int k = 0; k < 6; ++k) {
    const int p = k < 3 ? 0 : (k < 5 ? 1 : 2);
    const int q = k < 3 ? k + 1 : (k < 5 ? k - 1 : 3);
    for (int c = 0; c < 3; ++c) l2 += (X[p][c] - X[q][c]) * (X[p][c] - X[q][c]);
  }
  const double h = sqrt(l2 / 6.0);
  double xc[3];
  for (int c = 0; c < 3; ++c) xc[c] = 0.25 * (X[0][c] + X[1][c] + X[2][c] + X[3][c]);
  const MatParams m = a.mats[a.elem_set[e]];
  const double E = 9.0 * m.kappa * m.mu / (3.0 * m.kappa + m.mu);
  const double nu = (3.0 * m.kappa - 2.0 * m.mu) / (2.0 * (3.0 * m.kappa + m.mu));
  const double vm = a.elem_vm[e] / ge[0];
  double eq = 0.0;
  if (a.eqps) {
    for (int q = 0; q < a.nq; ++q) eq += a.eqps[e * a.nq + q];
    eq /= a.nq;
  }
  float* f = a.feat + e * kIn;
  f[0] = flog10(h);
  f[1] = (float)xc[0];
  f[2] = (float)xc[1];
  f[3] = (float)xc[2];
  f[4] = flog10(fabs(xc[0] - a.qoi_x) + h);
  f[5] = xc[0] < a.qoi_x ? 1.0f : 0.0f;
  f[6] = flog10(fabs(vm) / E + 1e-12);
  f[7] = flog10(fabs(vm - star[0]) / E + 1e-12);
  f[8] = flog10(sqrt(h2) + 1e-12);
  f[9] = flog10(sqrt(d2) + 1e-12);
  f[10] = (float)nu;
  f[11] = flog10(eq + 1e-9);
}

// ---------------------------------------------------------------- network

__device__ inline float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Round to nearest, ties to even (finite inputs).
__device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Network parameters staged in LDS as FP32: weights from their BF16 copy,
// biases from the FP32 master.
struct SharedNet {
  float w1[kHid * kIn], b1[kHid], w2[kHid * kHid], b2[kHid], w3[kHid], b3;
  float mu[kIn], is[kIn], mu_y, s_y, c;
};

__device__ void stage_net(SharedNet& S, const MlpArgs& a) {
  for (int i = threadIdx.x; i < kHid * kIn; i += blockDim.x) S.w1[i] = bf16_to_f32(a.wq[oW1 + i]);
  for (int i = threadIdx.x; i < kHid * kHid; i += blockDim.x) S.w2[i] = bf16_to_f32(a.wq[oW2 + i]);
  for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
    S.w3[i] = bf16_to_f32(a.wq[oW3 + i]);
    S.b1[i] = a.params[oB1 + i];
    S.b2[i] = a.params[oB2 + i];
  }
  for (int i = threadIdx.x; i < kIn; i += blockDim.x) {
    S.mu[i] = a.norms[i];
    S.is[i] = a.norms[kIn + i];
  }
  if (threadIdx.x == 0) {
    S.b3 = a.params[oB3];
    S.mu_y = a.norms[2 * kIn];
    S.s_y = a.norms[2 * kIn + 1];
    S.c = a.norms[kMlpOffC];
  }
  __syncthreads();
}

__device__ inline float round_bf16(float f) { return bf16_to_f32(f32_to_bf16(f)); }

// ---------------------------------------------------------- forward (MFMA)
//
// Layers 1 and 2 run on the matrix cores as v_mfma_f32_16x16x16_bf16 with the
// samples as the N dimension.  For lane l, r = l % 16 and q = l / 16:
//   A operand  lane holds A[r][4q + i]      (weights: output row r of a block)
//   B operand  lane holds B[4q + i][r]      (inputs of sample r)
//   C/D        lane holds D[4q + i][r]      i = 0..3
// so the D of a layer-1 row block is, after bias, ReLU and rounding, exactly
// the B operand of layer 2 for the matching K block: the activations never
// leave the registers.  All weights sit in registers for the whole launch.
// The inputs of each product (normalized features, layer-1 activations) are
// rounded to BF16; accumulation is FP32.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kFwdThreads = 256;
constexpr int kFwdWaves = kFwdThreads / 64;
constexpr int kFwdUnroll = 4;

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// y[i] = s_y * net(bf16((x[i] - mu) * is)) + mu_y + c * x[i][0].  Rows with a
// feature outside the training box [lo, hi] are counted in *ood.
__global__ __launch_bounds__(kFwdThreads) void mlp_forward_kernel(MlpArgs a, const float* x, int64_t n, float* y,
                                                                  int* ood) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  // the BF16 weights, staged through LDS with coalesced loads; each wave
  // then reads its operand fragments from there
  __shared__ uint16_t sw[kP + 1];
  for (int i = threadIdx.x; i < kP; i += kFwdThreads) sw[i] = a.wq[i];
  __syncthreads();
  bf16x4 A1[4], A2[4][4];
  float b1[4][4], b2[4][4], w3[4][4];
  for (int mb = 0; mb < 4; ++mb) {
    for (int i = 0; i < 4; ++i) {
      const int k = 4 * q + i;
      A1[mb][i] = k < kIn ? (short)sw[oW1 + (16 * mb + r) * kIn + k] : (short)0;
      const int row = 16 * mb + 4 * q + i;   // D row held by this lane
      b1[mb][i] = a.params[oB1 + row];
      b2[mb][i] = a.params[oB2 + row];
      w3[mb][i] = bf16_to_f32(sw[oW3 + row]);
    }
    for (int kb = 0; kb < 4; ++kb)
      for (int i = 0; i < 4; ++i) A2[mb][kb][i] = (short)sw[oW2 + (16 * mb + r) * kHid + 16 * kb + 4 * q + i];
  }
  float mu[4], is[4], lo[4], hi[4];
  for (int i = 0; i < 4; ++i) {
    const int k = 4 * q + i;
    mu[i] = k < kIn ? a.norms[k] : 0.0f;
    is[i] = k < kIn ? a.norms[kIn + k] : 0.0f;
    lo[i] = k < kIn ? a.norms[kMlpOffLo + k] : 0.0f;
    hi[i] = k < kIn ? a.norms[kMlpOffHi + k] : 0.0f;
  }
  const float b3 = a.params[oB3], mu_y = a.norms[2 * kIn], s_y = a.norms[2 * kIn + 1], cx = a.norms[kMlpOffC];
  const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  const int64_t tiles = (n + 15) / 16;
  const int64_t wave = (int64_t)blockIdx.x * kFwdWaves + (threadIdx.x >> 6);
  const int64_t stride = (int64_t)gridDim.x * kFwdWaves;
  int outside = 0;  // rows of this lane outside the box (lanes q == 0 only)
  // kFwdUnroll consecutive tiles per iteration: their feature loads are all
  // issued before the first product, so each wave keeps several in flight.
  for (int64_t t0 = wave * kFwdUnroll; t0 < tiles; t0 += stride * kFwdUnroll) {   // uniform per wave
    float v[kFwdUnroll][4];
#pragma unroll
    for (int u = 0; u < kFwdUnroll; ++u) {
      const int64_t smp = (t0 + u) * 16 + r;
#pragma unroll
      for (int i = 0; i < 4; ++i) v[u][i] = (4 * q + i < kIn && smp < n) ? x[smp * kIn + 4 * q + i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kFwdUnroll; ++u) {
      const int64_t smp = (t0 + u) * 16 + r;
      const bool live = smp < n;
      bf16x4 B1;
      int out = 0;
      for (int i = 0; i < 4; ++i) {
        const int k = 4 * q + i;
        if (k < kIn && live) out |= !(v[u][i] >= lo[i] && v[u][i] <= hi[i]);
        B1[i] = (short)f32_to_bf16(k < kIn ? (v[u][i] - mu[i]) * is[i] : 0.0f);
      }
      bf16x4 B2[4];
      for (int mb = 0; mb < 4; ++mb) {
        const f32x4 z = mfma16(A1[mb], B1, zero);
        for (int i = 0; i < 4; ++i) B2[mb][i] = (short)f32_to_bf16(fmaxf(z[i] + b1[mb][i], 0.0f));
      }
      float o = 0.0f;
      for (int mb = 0; mb < 4; ++mb) {
        f32x4 z = zero;
        for (int kb = 0; kb < 4; ++kb) z = mfma16(A2[mb][kb], B2[kb], z);
        for (int i = 0; i < 4; ++i) o += w3[mb][i] * fmaxf(z[i] + b2[mb][i], 0.0f);
      }
      // sum the four row groups of sample r (lanes r, r+16, r+32, r+48)
      o += __shfl_xor(o, 16, 64);
      o += __shfl_xor(o, 32, 64);
      out |= __shfl_xor(out, 16, 64);
      out |= __shfl_xor(out, 32, 64);
      // feature 0 of sample r is v[u][0] of lane r (q == 0)
      if (q == 0 && live) {
        y[smp] = (o + b3) * s_y + mu_y + cx * v[u][0];
        outside += out ? 1 : 0;
      }
    }
  }
  // one atomic per wave: a row-by-row count on a single address serializes
  for (int off = 32; off > 0; off >>= 1) outside += __shfl_xor(outside, off, 64);
  if (lane == 0 && outside && ood) atomicAdd(ood, outside);
}

// Gradient of L = (1/B) sum_i (net(x_i) - t_i)^2, t_i = (y_i - mu_y - c x_i0)/s_y, over
// rows idx[0..B).  A block takes tiles of kTile rows: wave 0 runs the forward
// and back-propagation of its rows into LDS, then all threads accumulate the
// parameter gradients of the tile into registers.  Each block writes its sums
// to partial[block][kP] and its loss sum to partial_loss[block].
constexpr int kTile = 64, kGradThreads = 256, kPad = kHid + 1;

struct SharedTile {
  float xs[kTile][kIn + 1];
  float h1[kTile][kPad], h2[kTile][kPad], d2[kTile][kPad], d1[kTile][kPad];
  float d3[kTile];
};

__global__ __launch_bounds__(kGradThreads) void mlp_grad_kernel(MlpArgs a, const float* x, const float* y,
                                                                const int32_t* idx, int64_t B, float* partial,
                                                                float* partial_loss) {
  __shared__ SharedNet S;
  __shared__ SharedTile T;
  stage_net(S, a);
  const int tid = threadIdx.x;
  // parameter ownership: W2 entries tid + 256 r (r < 16), W1 entries tid + 256 r
  // (r < 3), and for tid < 193 one of W3 / b1 / b2 / b3
  float g2[16], g1[3], g3 = 0.0f;
  for (int r = 0; r < 16; ++r) g2[r] = 0.0f;
  for (int r = 0; r < 3; ++r) g1[r] = 0.0f;
  float loss = 0.0f;
  const float invB = 1.0f / (float)B;
  const int64_t tiles = (B + kTile - 1) / kTile;
  for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    if (tid < kTile) {
      const int64_t row = tile * kTile + tid;
      const bool live = row < B;
      float xs[kIn];
      float t = 0.0f;
      if (live) {
        const int64_t r = idx[row];
        for (int k = 0; k < kIn; ++k) xs[k] = round_bf16((x[r * kIn + k] - S.mu[k]) * S.is[k]);
        t = (y[r] - S.mu_y - S.c * x[r * kIn]) / S.s_y;
      } else {
        for (int k = 0; k < kIn; ++k) xs[k] = 0.0f;
      }
      for (int k = 0; k < kIn; ++k) T.xs[tid][k] = xs[k];
      // Each thread works on its own padded LDS rows, which keeps register
      // use low; the odd row stride makes the 64 rows conflict-free.
      for (int j = 0; j < kHid; ++j) {
        float s = S.b1[j];
#pragma unroll
        for (int k = 0; k < kIn; ++k) s += S.w1[j * kIn + k] * xs[k];
        T.h1[tid][j] = round_bf16(fmaxf(s, 0.0f));
      }
      float o = S.b3;
      for (int jj = 0; jj < kHid; ++jj) {
        float s = S.b2[jj];
        for (int j = 0; j < kHid; ++j) s += S.w2[jj * kHid + j] * T.h1[tid][j];
        const float h = fmaxf(s, 0.0f);
        T.h2[tid][jj] = h;
        o += S.w3[jj] * h;
      }
      const float r = live ? o - t : 0.0f;
      loss += r * r;
      const float d3 = 2.0f * r * invB;
      T.d3[tid] = d3;
      for (int jj = 0; jj < kHid; ++jj) T.d2[tid][jj] = T.h2[tid][jj] > 0.0f ? d3 * S.w3[jj] : 0.0f;
      for (int j = 0; j < kHid; ++j) {
        float s = 0.0f;
        for (int jj = 0; jj < kHid; ++jj) s += S.w2[jj * kHid + j] * T.d2[tid][jj];
        T.d1[tid][j] = T.h1[tid][j] > 0.0f ? s : 0.0f;
      }
    }
    __syncthreads();
    // W2 entry tid + 256 r is (row tid / 64 + 4 r, column tid % 64)
    {
      const int j = tid % kHid, jj0 = tid / kHid;
      int j1[3], k1[3];
      for (int r = 0; r < 3; ++r) {
        j1[r] = (tid + kGradThreads * r) / kIn;
        k1[r] = (tid + kGradThreads * r) % kIn;
      }
      const int sel = tid / kHid, col = tid % kHid;  // W3 / b1 / b2 slot for tid < 192
#pragma unroll 2
      for (int q = 0; q < kTile; ++q) {
        const float h = T.h1[q][j];
#pragma unroll
        for (int r = 0; r < 16; ++r) g2[r] += T.d2[q][jj0 + 4 * r] * h;
#pragma unroll
        for (int r = 0; r < 3; ++r) g1[r] += T.d1[q][j1[r]] * T.xs[q][k1[r]];
        if (sel == 0) g3 += T.d3[q] * T.h2[q][col];
        else if (sel == 1) g3 += T.d1[q][col];
        else if (sel == 2) g3 += T.d2[q][col];
        else if (tid == 3 * kHid) g3 += T.d3[q];
      }
    }
    __syncthreads();
  }
  float* out = partial + (int64_t)blockIdx.x * kP;
  for (int r = 0; r < 16; ++r) out[oW2 + tid + kGradThreads * r] = g2[r];
  for (int r = 0; r < 3; ++r) out[oW1 + tid + kGradThreads * r] = g1[r];
  if (tid < kHid) out[oW3 + tid] = g3;
  else if (tid < 2 * kHid) out[oB1 + tid - kHid] = g3;
  else if (tid < 3 * kHid) out[oB2 + tid - 2 * kHid] = g3;
  else if (tid == 3 * kHid) out[oB3] = g3;
  if (tid < kTile) {
    // wave 0 holds the tile losses: fixed-order wave sum
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    if (tid == 0) partial_loss[blockIdx.x] = loss;
  }
}

// grad[p] = sum_b partial[b][p]; loss = sum_b partial_loss[b] / B.
__global__ void mlp_grad_reduce_kernel(const float* partial, const float* partial_loss, int nb, int64_t B,
                                       float* grad, float* loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kP) {
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s += partial[(int64_t)b * kP + p];
    grad[p] = s;
  }
  if (p == 0) {
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s += partial_loss[b];
    *loss = s / (float)B;
  }
}

// Adam (Kingma & Ba 2015, algorithm 1) on the FP32 master; weight entries
// are re-rounded into the BF16 copy.
__global__ void mlp_adam_kernel(MlpArgs a, const float* grad, int step, float lr, float beta1, float beta2,
                                float eps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kP) return;
  const float g = grad[p];
  const float m = beta1 * a.adam_m[p] + (1.0f - beta1) * g;
  const float v = beta2 * a.adam_v[p] + (1.0f - beta2) * g * g;
  a.adam_m[p] = m;
  a.adam_v[p] = v;
  const float mh = m / (1.0f - powf(beta1, (float)step));
  const float vh = v / (1.0f - powf(beta2, (float)step));
  const float w = a.params[p] - lr * mh / (sqrtf(vh) + eps);
  a.params[p] = w;
  a.wq[p] = f32_to_bf16(w);
}

__global__ void mlp_round_kernel(MlpArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kP) a.wq[p] = f32_to_bf16(a.params[p]);
}

}  // namespace

int mlp_grad_blocks(int64_t B) {
  const int64_t tiles = (B + kTile - 1) / kTile;
  return (int)std::min<int64_t>(std::max<int64_t>(tiles, 1), 512);
}

hipError_t launch_features(const FeatureArgs& a, hipStream_t s) {
  if (a.num_elems == 0) return hipSuccess;
  const unsigned ge = (unsigned)((a.num_elems + kFeatThreads - 1) / kFeatThreads);
  const unsigned gn = (unsigned)((a.num_nodes + kFeatThreads - 1) / kFeatThreads);
  hipLaunchKernelGGL(elem_geometry_kernel, dim3(ge), dim3(kFeatThreads), 0, s, a);
  hipLaunchKernelGGL(recover_kernel, dim3(gn), dim3(kFeatThreads), 0, s, a);
  hipLaunchKernelGGL(features_kernel, dim3(ge), dim3(kFeatThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const MlpArgs& a, const float* x, int64_t n, float* y, int* ood, hipStream_t s) {
  if (n == 0) return hipSuccess;
  // persistent: exactly the workgroups that are resident at once; each wave
  // loads the weights once, then walks groups of 16-row tiles
  static const int resident = [] {
    int occ = 0, dev = 0, cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(&mlp_forward_kernel),
                                                     kFwdThreads, 0) != hipSuccess ||
        occ < 1) {
      occ = 1;
      (void)hipGetLastError();
    }
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return occ * cus;
  }();
  const int64_t tiles = (n + 15) / 16;
  const int64_t groups = (tiles + kFwdUnroll - 1) / kFwdUnroll;
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((groups + kFwdWaves - 1) / kFwdWaves, resident));
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(g), dim3(kFwdThreads), 0, s, a, x, n, y, ood);
  return hipGetLastError();
}

hipError_t launch_mlp_grad(const MlpArgs& a, const float* x, const float* y, const int32_t* idx, int64_t B,
                           float* partial, float* partial_loss, float* grad, float* loss, hipStream_t s) {
  if (B <= 0) return hipErrorInvalidValue;
  const int nb = mlp_grad_blocks(B);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(nb), dim3(kGradThreads), 0, s, a, x, y, idx, B, partial, partial_loss);
  hipLaunchKernelGGL(mlp_grad_reduce_kernel, dim3((kP + 255) / 256), dim3(256), 0, s, partial, partial_loss, nb, B,
                     grad, loss);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(const MlpArgs& a, const float* grad, int step, float lr, hipStream_t s) {
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kP + 255) / 256), dim3(256), 0, s, a, grad, step, lr, 0.9f, 0.999f,
                     1e-8f);
  return hipGetLastError();
}

hipError_t launch_mlp_round(const MlpArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(mlp_round_kernel, dim3((kP + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace goal
