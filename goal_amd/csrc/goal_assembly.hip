// Element assembly for goal's mixed u-p Total Lagrangian tets, in two passes.
//
// Pass 1 (element_kernel) runs, per element, the evaluator chain that
// goal::assemble() runs hook by hook (src/goal_assembly.cpp:65-88, chain order
// src/goal_mechanics.cpp:97-146): gather of u and p, interpolation at the
// quadrature points, kinematics F = I + grad u and J, the constitutive model,
// Mixed, MResidual, PResidual and Stabilization.  It writes the element
// residual r_e and, for the Jacobian, the element matrix k_e to element-major
// buffers with coalesced stores.
//
// Pass 2 (gather_kernel) is the scatter (Displacement/Pressure::scatter,
// src/goal_displacement.cpp:77-86,177-214, src/goal_pressure.cpp:76-83,170-203)
// turned around: one wavefront owns one mesh node, i.e. the 4 consecutive CSR
// rows of that node, sums the row blocks of its incident elements in LDS and
// writes the rows once.  No atomics, no zero fill, and the summation order is
// fixed, so results are bitwise reproducible.
//
// Pass 1 work layout: a workgroup takes G elements at a time (6 to 32, by
// element kind) and one lane per column of their element matrices.  The
// phases, separated by workgroup barriers, keep all element data in LDS and
// hand out their tasks over all lanes:
//   A  gather connectivity, vertex coordinates, DOFs        (element, DOF)
//   B  shape functions                                       (element, point, node)
//   C1 quadrature-point state: kinematics, model, states     (element, point)
//   C2 one column of the 13x13 point tangent (J2, elastic)   (element, point, direction)
//   D  residual rows                                         (element, row)
//   E  Jacobian: a lane owns one column (b,k), forms C = D B
//      and writes K(row, col) = B^T C for all rows           (element, column)
// For neo-Hookean, C2 is skipped: phase E evaluates C directly from the
// closed-form tangent (dev::nh_direction), which keeps the 13x13 tables out of
// LDS and lets three workgroups share a CU.

#include <algorithm>
#include <type_traits>

#include "goal_device.hpp"
#include "goal_kernels.hpp"

// Element-block store.  NT: a nontemporal (streaming) store.  Measured on the
// P2 64^3 cube, streaming the neo-Hookean FP32 blocks takes the chunked
// FP32-tangent assembly from 10.60 to 10.38 ms, while streaming the FP64
// blocks more than doubles their element pass (8.6 -> 20.3 ms), and streaming
// the gather's loads and CSR writes changes nothing: only FP32 blocks stream.
template <bool NT, typename T>
__device__ inline void put(T* p, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

namespace goal {
namespace {

using dev::Quad;

template <int MODEL>
struct ModelOf;
template <>
struct ModelOf<MODEL_ELASTIC> {
  using type = dev::Elastic;
};
template <>
struct ModelOf<MODEL_NEOHOOKEAN> {
  using type = dev::NeoHookean;
};
template <>
struct ModelOf<MODEL_J2> {
  using type = dev::J2Model;
};

// What phase C1 leaves in LDS for the tangent columns (phase C2, or phase E
// for neo-Hookean, kept in the element-matrix type S).
template <int MODEL, typename S = double>
struct QPStore {
  dev::QP qp;
  dev::Stab st;
  typename ModelOf<MODEL>::type m;
};
template <typename S>
struct QPStore<MODEL_NEOHOOKEAN, S> {
  dev::NHBaseT<S> b;
};

// Workgroup shape.  Phase E gives one lane to each column of an element
// matrix, so a workgroup takes G elements and G*ND lanes, rounded up to whole
// wavefronts: 6 ten-node tets on 256 lanes, 32 four-node one-point tets on
// 512, 4 four-node four-point tets on 64.  Four-wave ten-node workgroups let
// three workgroups share a CU at W = 3 waves per SIMD (the register budget);
// five-wave workgroups were only ever placed one per CU.
//
// Neo-Hookean ten-node tets with four points (MF) form their element matrices
// on the matrix cores instead, in tiles of 16 columns per wavefront: G = 16
// elements on four waves, their 48 tiles dealt round-robin, the 64 points of
// phase C1 filling wave 0.  The 77 KB of LDS this takes leave two workgroups
// per CU, 2 waves per SIMD (182 VGPRs for FP64 element matrices, 199 for
// FP32).  Measured on the P2 64^3 cube,
// FP64 element pass / FP32 one: G = 16 6.48 / 4.89 ms, 14 6.80 / 5.21,
// 12 6.73 / 4.79, 8 7.32 / 5.66 and 6 7.87 / 6.10 at 2 waves per SIMD; at 3
// waves per SIMD (168 registers) G = 6 6.89 / 5.22, 8 8.26 / 5.04 and
// 10 8.47 / 4.97 ms.  Larger groups share phase C1's one-wave latency and the
// barriers among more elements.
#ifndef GOAL_MF_G
#define GOAL_MF_G 16
#endif
// Four-node tets with one point: G = 32 (512 lanes, 60 KB of LDS, two
// workgroups per CU at 4 waves per SIMD).  Measured on the P1 J2 bar
// (point + element pass): 6.62 ms at G = 32, 6.76-6.80 at 16, 7.49 at 8 and
// 9.52 at 24 (three 6-wave workgroups per CU).
#ifndef GOAL_P1_G
#define GOAL_P1_G 32
#endif
#ifndef GOAL_MF_WAVES
#define GOAL_MF_WAVES 2
#endif
#ifndef GOAL_MF
#define GOAL_MF 1  // 0: the vector-ALU phase E for every kernel
#endif
template <int NN, int NQ, int MODEL>
struct Layout {
  static constexpr int ND = 4 * NN;
  static constexpr bool MF = GOAL_MF && MODEL == MODEL_NEOHOOKEAN && NN == 10 && NQ == 4;
  static constexpr int G = MF ? GOAL_MF_G : (NN == 10) ? 6 : (NQ == 1 ? GOAL_P1_G : 4);
  static constexpr int T = MF ? 256 : ((G * ND + 63) / 64) * 64;
  static constexpr int W = (MODEL == MODEL_J2) ? 1 : MF ? GOAL_MF_WAVES : GOAL_ELEM_WAVES;
};

template <int NN, int NQ, int G>
struct SBase {
  double x[G][4 * NN];
  double X[G][4][3];
  double sh[G][NQ][NN][kShapeRow];  // N_a, dN_a/dX, pad
  double wdv[G][NQ];
  double sig[G][NQ][13];
  int32_t conn[G][NN];
  int32_t rec[G][NN];
};
template <int NQ, int MODEL, int G, typename S>
struct SJac {
  double D[G][NQ][13][13];
  QPStore<MODEL> qs[G][NQ];
};
// neo-Hookean: phase E contracts the closed-form tangent itself
template <int NQ, int G, typename S>
struct SJac<NQ, MODEL_NEOHOOKEAN, G, S> {
  QPStore<MODEL_NEOHOOKEAN, S> qs[G][NQ];
};
template <int NN, int NQ, int G>
struct SLoc {
  double zq[G][NQ][4];      // z_diff (u,p) at qp
  double zgq[G][NQ][4][3];  // grad z_diff
  double zcq[G][NQ][4];     // z_coarse p and its gradient
  double zd[G][4 * NN];
  double zc[G][NN];
};
struct SNone {};

// neo-Hookean: the next group's phase-A data, fetched by wave 1 while wave 0
// computes the point states (phase C1)
template <int NN, int G>
struct SPrefetch {
  int32_t pconn[G][NN];
  int32_t prec[G][NN];
  double pX[G][4][3];
  double px[G][4 * NN];
};
struct SNoPrefetch {};
// (FP64 element matrices and residuals, and the matrix-core kernels of either
// precision; the vector-ALU FP32-tangent kernels measured slightly slower
// with it, 10.50 -> 10.55 ms, chunked beside the gather)
template <int NN, int NQ, int MODEL, int KIND, typename TE>
constexpr bool kPrefetch = MODEL == MODEL_NEOHOOKEAN && KIND != ASM_LOCALIZE && Layout<NN, NQ, MODEL>::T >= 128 &&
                           (std::is_same<TE, double>::value || Layout<NN, NQ, MODEL>::MF);

// J2 assemblies other than the localization take their point states from the
// thread-per-point pass (j2_point_kernel) instead of phases C1 and C2.
template <int MODEL, int KIND>
constexpr bool kPointPass = MODEL == MODEL_J2 && KIND != ASM_LOCALIZE;
template <int NQ, int G>
struct SJacD {
  double D[G][NQ][13][13];
};

template <int NN, int NQ, int MODEL, int KIND, int G, typename TE = double>
struct Smem : std::conditional<kPrefetch<NN, NQ, MODEL, KIND, TE>, SPrefetch<NN, G>, SNoPrefetch>::type {
  static constexpr bool JK = KIND == ASM_JACOBIAN || KIND == ASM_ADJOINT;
  using Ext = typename std::conditional<
      JK && kPointPass<MODEL, KIND>, SJacD<NQ, G>,
      typename std::conditional<
          JK, SJac<NQ, MODEL, G, TE>,
          typename std::conditional<KIND == ASM_LOCALIZE, SLoc<NN, NQ, G>, SNone>::type>::type>::type;
  SBase<NN, NQ, G> b;
  Ext ext;
};
// Waves per SIMD the element kernel is compiled for: the register budget
// (W, or GOAL_ELEM_WAVES for the J2 point-pass kinds), lowered to what the
// workgroup's LDS allows (160 KB per CU), so the compiler is not held to an
// occupancy the LDS rules out.
template <int NN, int NQ, int MODEL, int KIND, typename TE>
constexpr int elem_waves() {
  using Lay = Layout<NN, NQ, MODEL>;
  constexpr int want = kPointPass<MODEL, KIND> ? GOAL_ELEM_WAVES : Lay::W;
  constexpr int per_cu = 163840 / (int)sizeof(Smem<NN, NQ, MODEL, KIND, Lay::G, TE>);
  constexpr int lds = per_cu * (Lay::T / 64) / 4;
  return lds < 1 ? 1 : (lds < want ? lds : want);
}

// Element-group schedule.  Plain: group = block + it * gridDim.  XCD slabs:
// workgroups are dispatched round-robin over the 8 XCDs, so block b runs on
// XCD b % 8; XCD x is given the groups [x*per, (x+1)*per) and its blocks walk
// that slab together.
__device__ inline int64_t group_of(int64_t it, int64_t ngroups, int remap) {
  const int64_t nb = gridDim.x, b = blockIdx.x;
  if (!remap || (nb & 7)) {
    int64_t g = b + it * nb;
    return g < ngroups ? g : -1;
  }
  const int64_t x = b & 7, s = b >> 3, bx = nb >> 3;
  const int64_t per = (ngroups + 7) >> 3;
  int64_t g = s + it * bx;
  if (g >= per) return -1;
  g += x * per;
  return g < ngroups ? g : -1;
}

// Column col = (b, kc) of the element matrix, all rows at once:
//   C_q = D_q B_q[:, col]  (13 values),   K[(a,i)][col] = sum_q B_q[:, (a,i)]^T C_q.
// Quadrature points are the outer loop so each shape-function row of the LDS
// table is read once per column; the NN*4 sums stay in registers.
// Store column col = (b, kc) of an element matrix: into the row blocks of the
// element's nodes (rows (a, i) at record rec[a]), or, transposed, as row kc of
// node b's block.
template <int NN, bool TR, typename T, bool NTS = false>
__device__ inline void store_column(const T (*acc)[4], int col, const int32_t* rec, T* em) {
  constexpr int ND = 4 * NN;
  if constexpr (!TR) {
#pragma unroll
    for (int n = 0; n < NN; ++n) {
      T* o = em + (int64_t)rec[n] * (4 * ND) + col;
      for (int i = 0; i < 4; ++i) {
        if constexpr (NTS) put<true>(o + i * ND, acc[n][i]);
        else o[i * ND] = acc[n][i];
      }
    }
  } else {
    T* o = em + (int64_t)rec[col >> 2] * (4 * ND) + (col & 3) * ND;
#pragma unroll
    for (int n = 0; n < NN; ++n)
      for (int i = 0; i < 4; ++i) {
        if constexpr (NTS) put<true>(o + 4 * n + i, acc[n][i]);
        else o[4 * n + i] = acc[n][i];
      }
  }
}

// T: the element-matrix type (double; float for the FP32 tangent, which forms
// and accumulates K in FP32 from the FP64 point tangent).
template <int NN, int NQ, bool TR, typename T>
__device__ inline void emit_column(const double (*sh)[NN][kShapeRow], const double (*D)[13][13], int col,
                                   const int32_t* rec, T* em) {
  const int b = col >> 2, kc = col & 3;
  T acc[NN][4];
#pragma unroll
  for (int n = 0; n < NN; ++n)
    for (int i = 0; i < 4; ++i) acc[n][i] = 0.0;
#pragma unroll 1
  for (int q = 0; q < NQ; ++q) {
    const double* shb = sh[q][b];
    const double sb0 = shb[0], sb1 = shb[1], sb2 = shb[2], sb3 = shb[3];
    T C[13];
#pragma unroll
    for (int r = 0; r < 13; ++r) {
      const double* Dr = D[q][r];
      C[r] = (T)((kc < 3) ? Dr[3 * kc] * sb1 + Dr[3 * kc + 1] * sb2 + Dr[3 * kc + 2] * sb3
                          : Dr[9] * sb0 + Dr[10] * sb1 + Dr[11] * sb2 + Dr[12] * sb3);
    }
#pragma unroll
    for (int n = 0; n < NN; ++n) {
      const double* sa = sh[q][n];
      const T s0 = (T)sa[0], s1 = (T)sa[1], s2 = (T)sa[2], s3 = (T)sa[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) acc[n][i] += C[3 * i] * s1 + C[3 * i + 1] * s2 + C[3 * i + 2] * s3;
      acc[n][3] += C[9] * s0 + C[10] * s1 + C[11] * s2 + C[12] * s3;
    }
  }
  store_column<NN, TR, T>(acc, col, rec, em);
}

// neo-Hookean column (b, kc): C_q from the closed-form tangent contracted
// with node b's shape function (dev::nh_direction), then the rows as above.
template <int NN, int NQ, bool TR, typename T>
__device__ inline void emit_column_nh(const double (*sh)[NN][kShapeRow], const double* wdv,
                                      const QPStore<MODEL_NEOHOOKEAN, T>* qs, int col, const int32_t* rec,
                                      T* em) {
  const int b = col >> 2, kc = col & 3;
  T acc[NN][4];
#pragma unroll
  for (int n = 0; n < NN; ++n)
    for (int i = 0; i < 4; ++i) acc[n][i] = 0.0;
#pragma unroll 1
  for (int q = 0; q < NQ; ++q) {
    const double w = wdv[q];
    const double* shb = sh[q][b];
    const T G[3] = {(T)(w * shb[1]), (T)(w * shb[2]), (T)(w * shb[3])};
    T C[13];
    dev::nh_direction<T>(qs[q].b, kc, (T)(w * shb[0]), G, C);
#pragma unroll
    for (int n = 0; n < NN; ++n) {
      const double* sa = sh[q][n];
      const T s0 = (T)sa[0], s1 = (T)sa[1], s2 = (T)sa[2], s3 = (T)sa[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) acc[n][i] += C[3 * i] * s1 + C[3 * i + 1] * s2 + C[3 * i + 2] * s3;
      acc[n][3] += C[9] * s0 + C[10] * s1 + C[11] * s2 + C[12] * s3;
    }
  }
  store_column<NN, TR, T, std::is_same<T, float>::value>(acc, col, rec, em);
}

// Matrix-core form of the neo-Hookean element matrix (four points).  With the
// point index q as the inner dimension k of a 16x16x4 product,
//   K[(a,i), col] = sum_J sum_q G_q[a][J] C_q[3i+J][col]                      (i < 3)
//   K[(a,3), col] = sum_q N_q[a] C_q[9][col] + sum_J sum_q G_q[a][J] C_q[10+J][col]
// where G_q, N_q are the shape-function gradients and values at point q (rows
// a = element nodes, zero past NN) and C_q[:, col] = D_q B_q[:, col] is the
// point tangent applied to column col (dev::nh_direction, w dv folded in).
// A wavefront takes 16 columns at a time: lane l = 16 q + n evaluates C_q for
// column 16 t + n, which is exactly the B operand of the product (k = q, n),
// and holds G_q[n][J], N_q[n] as the A operands (m = n, k = q).  Thirteen
// products give the four 16x16 row blocks (i = 0..3) of the tile; lane l then
// holds four rows of column 16 t + (l % 16): a = 4 (l / 16) + r for the FP32
// product and a = 4 r + (l / 16) for the FP64 one, r = 0..3 (measured on
// gfx950 with a probe of known products).
// T: FP64 (v_mfma_f64_16x16x4f64) or FP32 (v_mfma_f32_16x16x4f32) element
// matrices, accumulated in their own precision.
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using acc = f64x4;
  __device__ static acc mac(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int g, int r) { return 4 * r + g; }
};
template <>
struct Mfma<float> {
  using acc = f32x4;
  __device__ static acc mac(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int g, int r) { return 4 * g + r; }
};

// Tiles of the element's columns, two ways.
//  - By kind (col): displacement columns of five nodes (15 lanes) per tile,
//    then the pressure columns (up to 16 nodes per tile), so every lane of a
//    tile takes the same branch of dev::nh_direction.  The pipelined FP32
//    kernel needs this: its branch-free columns fix the kind per tile.
//  - Contiguous (col_contig): columns 16 t .. 16 t + 15, so a tile stores
//    whole 128-byte runs of each block row and the lanes of a tile take both
//    branches.  The FP64 kernel uses these.  Measured on the P2 64^3 cube
//    (same GPU): FP64 element pass 7.33-7.40 -> 7.22-7.24 ms, its HBM writes
//    24.5 -> 21.6 GB (the blocks are 20.6 GB, residual rows and stresses 1 GB):
//    with tiles by kind every 128-byte line of a block row took pieces from
//    two tiles, and part of them reached HBM as partial writes.  Trading rows
//    between neighbouring lanes to store two columns at once (16-byte stores)
//    was slower: 7.75 against 7.14 ms.
template <int NN>
struct MfTiles {
  static constexpr int DT = (3 * NN + 14) / 15;
  static constexpr int PT = (NN + 15) / 16;
  static constexpr int NT = DT + PT;
  static_assert(NT == (4 * NN + 15) / 16, "as many contiguous tiles as tiles by kind");
  // element-matrix column of lane n in tile t, or -1
  __device__ static int col(int t, int n) {
    if (t < DT) {
      const int b = 5 * t + n / 3;
      return (n < 15 && b < NN) ? 4 * b + n % 3 : -1;
    }
    const int b = 16 * (t - DT) + n;
    return b < NN ? 4 * b + 3 : -1;
  }
  __device__ static int col_contig(int t, int n) {
    const int c = 16 * t + n;
    return c < 4 * NN ? c : -1;
  }
};

// The thirteen products of one tile and the stores of its values: lane l
// holds the A operands (aN, aG) of row l % 16 at point g = l / 16 and the
// point-tangent column C of column col (-1: a padding lane), and stores rows
// a = M::row(g, r) of column col.
template <int NN, bool TR, typename T>
__device__ inline void mf_product_store(T aN, T aG0, T aG1, T aG2, const T* C, int col, int g, const int32_t* rec,
                                        T* em) {
  constexpr int ND = 4 * NN;
  using M = Mfma<T>;
  using acc = typename M::acc;
  acc k0 = {0, 0, 0, 0}, k1 = {0, 0, 0, 0}, k2 = {0, 0, 0, 0}, k3 = {0, 0, 0, 0};
  k0 = M::mac(aG0, C[0], k0);
  k1 = M::mac(aG0, C[3], k1);
  k2 = M::mac(aG0, C[6], k2);
  k3 = M::mac(aN, C[9], k3);
  k0 = M::mac(aG1, C[1], k0);
  k1 = M::mac(aG1, C[4], k1);
  k2 = M::mac(aG1, C[7], k2);
  k3 = M::mac(aG0, C[10], k3);
  k0 = M::mac(aG2, C[2], k0);
  k1 = M::mac(aG2, C[5], k1);
  k2 = M::mac(aG2, C[8], k2);
  k3 = M::mac(aG1, C[11], k3);
  k3 = M::mac(aG2, C[12], k3);
  if (col < 0) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int a = M::row(g, r);
    if (a < NN) {
      if constexpr (!TR) {
        T* o = em + (int64_t)rec[a] * (4 * ND) + col;
        o[0] = k0[r];
        o[ND] = k1[r];
        o[2 * ND] = k2[r];
        o[3 * ND] = k3[r];
      } else {
        T* o = em + (int64_t)rec[col >> 2] * (4 * ND) + (col & 3) * ND + 4 * a;
        o[0] = k0[r];
        o[1] = k1[r];
        o[2] = k2[r];
        o[3] = k3[r];
      }
    }
  }
}

// One tile of contiguous columns (MfTiles::col_contig): each lane evaluates
// the column of its own kind (dev::nh_direction branches per lane).
template <int NN, bool TR, typename T>
__device__ inline void emit_tile_mfma(const double (*sh)[NN][kShapeRow], const double* wdv,
                                      const QPStore<MODEL_NEOHOOKEAN, T>* qs, const int32_t* rec, T* em, int lane,
                                      int t) {
  const int q = lane >> 4, n = lane & 15;
  T aN = (T)0, aG0 = (T)0, aG1 = (T)0, aG2 = (T)0;
  if (n < NN) {
    const double* sa = sh[q][n];
    aN = (T)sa[0];
    aG0 = (T)sa[1];
    aG1 = (T)sa[2];
    aG2 = (T)sa[3];
  }
  const int col = MfTiles<NN>::col_contig(t, n);
  T C[13];
  if (col >= 0) {
    const double w = wdv[q];
    const double* shb = sh[q][col >> 2];
    const T G[3] = {(T)(w * shb[1]), (T)(w * shb[2]), (T)(w * shb[3])};
    dev::nh_direction<T>(qs[q].b, col & 3, (T)(w * shb[0]), G, C);
  } else {
#pragma unroll
    for (int r = 0; r < 13; ++r) C[r] = (T)0;
  }
  mf_product_store<NN, TR, T>(aN, aG0, aG1, aG2, C, col, q, rec, em);
}

// The same tiles in software-pipelined form: a wave takes whole elements and,
// for tile t, forms the point-tangent columns C of the next tile on the
// vector ALUs while the 13 products of tile t run on the matrix cores.  The
// column and the products are branch-free, so both sit in one scheduling
// region: a padding lane evaluates a real column (or node row) instead, and
// since column n of B and row m of A reach only column n and row m of the
// product, which are not stored, what it holds is never seen.  The products
// and their summation order are those of emit_tile_mfma, and so are the
// bits of every stored value.
//
// Used for FP32 element matrices only.  Measured on the P2 64^3 cube (same
// GPU, same run): FP32 element pass 5.10 -> 4.62 ms, but FP64 7.32 -> 8.62 ms:
// on MI355X an FP64 16x16x4 product takes the FP64 vector rate (64 cycles),
// and overlapping the two in one wave only slows the FP64 kernel.  Explicit
// interleaving (sched_group_barrier: one product, then 1-2 LDS reads and
// 6-12 vector instructions) measured 1-2% slower than the compiler's order.
// The FP64 kernel keeps the round-robin deal of contiguous tiles, the fastest
// measured (with tiles by kind, 24.5 GB of HBM writes for 20.6 GB of
// blocks; the other deals below were measured with tiles by kind): dealing
// each wave whole elements (its three tiles in a row, unpipelined) took
// 7.34 -> 9.93 ms, 10.3 ms with the tile kinds staggered across the waves
// (HBM writes 24.5 -> 31.0 GB), waves 0-2 on one element's three tiles and
// wave 3 on a fourth element 10.5 ms (33.9 GB), and the pipelined loop over
// the round-robin tasks 7.32 -> 8.49 ms (FP32 4.64 -> 4.74 ms).
#ifndef GOAL_MF_PIPE
#define GOAL_MF_PIPE 1
#endif
template <int NN, int TILE, typename T>
__device__ inline void mf_tile_column(const double (*sh)[NN][kShapeRow], const double* wdv,
                                      const QPStore<MODEL_NEOHOOKEAN, T>* qs, int lane, T* C) {
  using Tiles = MfTiles<NN>;
  const int q = lane >> 4, n = lane & 15;
  const int col = Tiles::col(TILE, n);
  const bool valid = col >= 0;
  const int b = valid ? (col >> 2) : 0;
  const double w = wdv[q];
  const double* shb = sh[q][b];
  const T G[3] = {(T)(w * shb[1]), (T)(w * shb[2]), (T)(w * shb[3])};
  if constexpr (TILE < Tiles::DT) dev::nh_direction_u<T>(qs[q].b, valid ? (col & 3) : 0, G, C);
  else dev::nh_direction_p<T>(qs[q].b, (T)(w * shb[0]), G, C);
}

// A operands of an element's products: N_q[n], G_q[n][J] of lane l = 16 q + n
// (node 0's past the element's nodes: padding rows).
template <int NN, typename T>
__device__ inline void mf_rows(const double (*sh)[NN][kShapeRow], int lane, T* A) {
  const int q = lane >> 4, n = lane & 15;
  const double* sa = sh[q][n < NN ? n : 0];
#pragma unroll
  for (int j = 0; j < 4; ++j) A[j] = (T)sa[j];
}

template <int NN, int TILE, bool TR, typename T>
__device__ inline void mf_tile_product(const T* A, const T* C, const int32_t* rec, T* em, int lane) {
  mf_product_store<NN, TR, T>(A[0], A[1], A[2], A[3], C, MfTiles<NN>::col(TILE, lane & 15), lane >> 4, rec, em);
}

template <int NN, int NQ, int MODEL, int KIND, typename TE>
__global__ __launch_bounds__((Layout<NN, NQ, MODEL>::T), (elem_waves<NN, NQ, MODEL, KIND, TE>())) void element_kernel(
    AsmArgs a) {
  using Lay = Layout<NN, NQ, MODEL>;
  using M = typename ModelOf<MODEL>::type;
  constexpr int G = Lay::G, ND = Lay::ND, T = Lay::T;
  constexpr bool JAC = (KIND == ASM_JACOBIAN || KIND == ASM_ADJOINT);
  constexpr bool LOC = (KIND == ASM_LOCALIZE);
  __shared__ Smem<NN, NQ, MODEL, KIND, G, TE> sm;
  auto& s = sm.b;
  const int tid = threadIdx.x;
  const int64_t ne = a.num_elems;
  int64_t g_lo = 0, ngroups = (ne + G - 1) / G;
  if constexpr (MODEL == MODEL_NEOHOOKEAN) {
    if (a.group_hi > 0) {
      g_lo = a.group_lo;
      ngroups = a.group_hi - a.group_lo;
    }
  }
  const int64_t nqtot = ne * NQ;
  bool fetched = false;  // this group's phase-A data is in the prefetch buffers
  // J2 point-pass kinds: four-node tets load a group's point states into
  // registers one group ahead, so the loads are in flight during the previous
  // group's phases (ten-node tets have no registers to spare for it)
  constexpr bool PP = kPointPass<MODEL, KIND>;
  constexpr int PPS = PP ? (JAC ? kPointTangent : kPointStress) : 1;
  constexpr int NPF = PP ? (G * NQ * PPS + T - 1) / T : 1;
  constexpr bool PPF = PP && NN == 4 && NQ == 1;
  double pf[PPF ? NPF : 1];
  auto fetch_points = [&](int64_t it_) {
    const int64_t gq = group_of(it_, ngroups, a.xcd_remap);
    if (gq < 0) return;
    const int64_t e0q = (g_lo + gq) * G;
    const int npq = (int)min((int64_t)G, ne - e0q) * NQ;
#pragma unroll
    for (int i = 0; i < NPF; ++i) {
      const int t = tid + i * T;
      if (t < npq * PPS) pf[i] = a.pt[point_state_index(e0q * NQ + t % npq, t / npq, PPS)];
    }
  };
  if constexpr (PPF) fetch_points(0);
  if constexpr (kPointPass<MODEL, KIND> && JAC) {
    // the point tangent's structural zeros (s3 rows of the p column, P and r1
    // rows of the grad-p columns): never stored, never overwritten
    for (int t = tid; t < G * NQ * 33; t += T) {
      auto& D = sm.ext.D[t / (NQ * 33)][(t / 33) % NQ];
      const int z = t % 33;
      if (z < 3) D[10 + z][9] = 0.0;
      else D[(z - 3) % 10][10 + (z - 3) / 10] = 0.0;
    }
  }

  for (int64_t it = 0;; ++it) {
    const int64_t grp0 = group_of(it, ngroups, a.xcd_remap);
    if (grp0 < 0) break;
    const int64_t grp = g_lo + grp0;
    const int64_t e0 = grp * G;
    const int nv = (int)min((int64_t)G, ne - e0);  // elements of this group

    // ---- A: gather (Displacement/Pressure::gather, src/goal_displacement.cpp:139-156)
    constexpr bool PF = kPrefetch<NN, NQ, MODEL, KIND, TE>;
    if (PF && fetched) {
      // this group's data came in during the previous group's phase C1
      if constexpr (PF) {
        for (int t = tid; t < nv * NN; t += T) {
          s.conn[t / NN][t % NN] = sm.pconn[t / NN][t % NN];
          s.rec[t / NN][t % NN] = sm.prec[t / NN][t % NN];
        }
        for (int t = tid; t < nv * 12; t += T) s.X[t / 12][(t % 12) / 3][t % 3] = sm.pX[t / 12][(t % 12) / 3][t % 3];
        for (int t = tid; t < nv * ND; t += T) s.x[t / ND][t % ND] = sm.px[t / ND][t % ND];
      }
    } else {
    for (int t = tid; t < nv * NN; t += T) {
      s.conn[t / NN][t % NN] = a.conn[e0 * NN + t];
      s.rec[t / NN][t % NN] = a.rec[e0 * NN + t];
    }
    __syncthreads();
    for (int t = tid; t < nv * 12; t += T) {
      const int el = t / 12, c = t % 12;
      s.X[el][c / 3][c % 3] = a.coords[3 * (int64_t)s.conn[el][c / 3] + (c % 3)];
    }
    if constexpr (!kPointPass<MODEL, KIND>)  // the point pass read the DOFs
      for (int t = tid; t < nv * ND; t += T) {
        const int el = t / ND, k = t % ND;
        const int64_t dof = 4 * (int64_t)s.conn[el][k >> 2] + (k & 3);
        s.x[el][k] = a.x[dof];
        if constexpr (LOC) sm.ext.zd[el][k] = a.z_diff[dof];
      }
    if constexpr (LOC)
      for (int t = tid; t < nv * NN; t += T) sm.ext.zc[t / NN][t % NN] = a.z_coarse[4 * (int64_t)s.conn[t / NN][t % NN] + 3];
    }
    __syncthreads();

    // ---- B: shape functions at the quadrature points (u/uw/p/pw at_point)
    if constexpr (MODEL == MODEL_NEOHOOKEAN && (std::is_same<TE, double>::value || Lay::MF)) {
      // grad lambda and det once per element, staged in the stress slots of
      // the element (phase C1 writes them after the barrier below); each
      // (element, point, node) lane then reads the rows it needs by address,
      // not through per-lane selects over register arrays.  Same arithmetic
      // as dev::shape, so the same bits.  Measured on the P2 64^3 cube: FP64
      // element pass 8.55 -> 8.19 ms.  The vector-ALU FP32-tangent kernels
      // keep the register form: their chunked assembly shares the CUs with
      // the gather's LDS atomics and went from 10.51 to 10.93 ms with this
      // one.  The matrix-core FP32 kernels use it (element pass 6.60 ->
      // 6.05 ms with the prefetch below).
      for (int el = tid; el < nv; el += T) {
        double gl[4][3];
        double* o = &s.sig[el][0][0];
        o[12] = dev::tet_gradlambda(s.X[el], gl);
        for (int v = 0; v < 4; ++v)
          for (int c = 0; c < 3; ++c) o[3 * v + c] = gl[v][c];
      }
      __syncthreads();
      for (int t = tid; t < nv * NQ * NN; t += T) {
        const int el = t / (NQ * NN), r = t % (NQ * NN);
        const int q = r / NN, n = r % NN;
        const double* gl = &s.sig[el][0][0];
        double xi[3], w;
        Quad<NQ>::point(q, xi, w);
        const double lam0 = 1.0 - xi[0] - xi[1] - xi[2];
        auto lam = [&](int k) { return k == 0 ? lam0 : Quad<NQ>::coord(q, k - 1); };
        double N, dN[3];
        if (NN == 4 || n < 4) {
          const double la = lam(n);
          if (NN == 4) {
            N = la;
            for (int c = 0; c < 3; ++c) dN[c] = gl[3 * n + c];
          } else {
            N = la * (2.0 * la - 1.0);
            const double f = 4.0 * la - 1.0;
            for (int c = 0; c < 3; ++c) dN[c] = f * gl[3 * n + c];
          }
        } else {
          const int ed = n - 4;  // edges 01,12,20,03,13,23
          const int i = ed < 3 ? ed : ed - 3, j = ed < 3 ? (ed == 2 ? 0 : ed + 1) : 3;
          const double li = lam(i), lj = lam(j);
          N = 4.0 * li * lj;
          for (int c = 0; c < 3; ++c) dN[c] = 4.0 * (li * gl[3 * j + c] + lj * gl[3 * i + c]);
        }
        s.sh[el][q][n][0] = N;
        s.sh[el][q][n][1] = dN[0];
        s.sh[el][q][n][2] = dN[1];
        s.sh[el][q][n][3] = dN[2];
        if (n == 0) s.wdv[el][q] = w * gl[12];
      }
    } else {
      for (int t = tid; t < nv * NQ * NN; t += T) {
        const int el = t / (NQ * NN), r = t % (NQ * NN);
        const int q = r / NN, n = r % NN;
        double gl[4][3];
        const double det = dev::tet_gradlambda(s.X[el], gl);
        double xi[3], w;
        Quad<NQ>::point(q, xi, w);
        const double lam[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
        double N, dN[3];
        dev::shape<NN>(n, lam, gl, N, dN);
        s.sh[el][q][n][0] = N;
        s.sh[el][q][n][1] = dN[0];
        s.sh[el][q][n][2] = dN[1];
        s.sh[el][q][n][3] = dN[2];
        if (n == 0) s.wdv[el][q] = w * det;
      }
    }
    __syncthreads();

    // ---- C1: quadrature-point state (Kinematics, model, Mixed; states), one lane per point
    if constexpr (PF) {
      // the waves past the first idle in this phase (the points take fewer
      // lanes than one wave): they fetch the next group's connectivity,
      // incidence records, vertex coordinates and DOFs, one (element, node)
      // per lane
      static_assert(G * NQ <= 64 && G * NN <= T - 64, "phase C1 on wave 0, the next group's nodes on the others");
      const int64_t gn = group_of(it + 1, ngroups, a.xcd_remap);
      fetched = gn >= 0;
      const int l = tid - 64;
      if (fetched && l >= 0) {
        const int64_t en = (g_lo + gn) * G;
        if (l < (int)min((int64_t)G, ne - en) * NN) {
          const int el = l / NN, n = l % NN;
          const int32_t c = a.conn[en * NN + l];
          sm.pconn[el][n] = c;
          sm.prec[el][n] = a.rec[en * NN + l];
          if (n < 4)
            for (int k = 0; k < 3; ++k) sm.pX[el][n][k] = a.coords[3 * (int64_t)c + k];
          for (int k = 0; k < 4; ++k) sm.px[el][4 * n + k] = a.x[4 * (int64_t)c + k];
        }
      }
    }
    if constexpr (PP) {
      // the point pass left [P, r1, s3] and the tangent, times w dv, per
      // point (with PPF they arrived in pf during the previous group)
      constexpr int PS = PPS;
      const int npq = nv * NQ;
      auto put = [&](int t, double v) {
        const int r = t / npq, pq = t % npq;  // consecutive lanes: consecutive points
        if (r < 13) {
          s.sig[pq / NQ][pq % NQ][r] = v;
        } else if constexpr (JAC) {
          auto& D = sm.ext.D[pq / NQ][pq % NQ];
          if (r < 130) D[(r - 13) % 13][(r - 13) / 13] = v;
          else if (r < 140) D[r - 130][9] = v;
          else D[10 + (r - 140) % 3][10 + (r - 140) / 3] = v;
        }
      };
      if constexpr (PPF) {
#pragma unroll
        for (int i = 0; i < NPF; ++i)
          if (tid + i * T < npq * PS) put(tid + i * T, pf[i]);
      } else {
        for (int t = tid; t < npq * PS; t += T) put(t, a.pt[point_state_index(e0 * NQ + t % npq, t / npq, PS)]);
      }
      if constexpr (PPF) fetch_points(it + 1);
    }
    if constexpr (!kPointPass<MODEL, KIND>) {
      for (int t = tid; t < nv * NQ; t += T) {
        const int el = t / NQ, q = t % NQ;
        const int64_t e = e0 + el;
        const MatParams mp = a.mats[a.elem_set ? a.elem_set[e] : 0];
        dev::QP qp;
        for (int k = 0; k < 9; ++k) qp.H[k] = 0.0;
        qp.p = 0.0;
        qp.g[0] = qp.g[1] = qp.g[2] = 0.0;
  #pragma unroll 1
        for (int n = 0; n < NN; ++n) {
          const double N = s.sh[el][q][n][0];
          const double pn = s.x[el][4 * n + 3];
          qp.p += pn * N;
          for (int j = 0; j < 3; ++j) {
            const double dNj = s.sh[el][q][n][1 + j];
            qp.g[j] += pn * dNj;
            for (int i = 0; i < 3; ++i) qp.H[3 * i + j] += s.x[el][4 * n + i] * dNj;
          }
        }
        for (int k = 0; k < 9; ++k) qp.F[k] = qp.H[k] + ((k % 4 == 0) ? 1.0 : 0.0);
        qp.J = dev::det3(qp.F);
        dev::inv3(qp.F, qp.J, qp.Fi);
        dev::Stab st;
        st.init(qp, a.stabilization ? dev::stab_tau(s.X[el], mp) : 0.0);
        const double wdv = s.wdv[el][q];
        const int64_t sidx = e * NQ + q;
        M m;
        if constexpr (MODEL == MODEL_J2) {
          double Fpo[9];
          for (int k = 0; k < 9; ++k) Fpo[k] = a.Fp_old[k * nqtot + sidx];
          m.init(qp, mp, Fpo, a.eqps_old[sidx]);
          if (m.failed()) atomicOr(a.status, 1);
          if (a.save_state) {
            a.eqps[sidx] = m.alpha;
            if (m.plastic) {
              double Fpn[9];
              m.new_Fp(Fpo, Fpn);
              for (int k = 0; k < 9; ++k) a.Fp[k * nqtot + sidx] = Fpn[k];
            }
          }
        } else {
          m.init(qp, mp);
        }
        if (a.save_state)
          for (int k = 0; k < 9; ++k) a.sigma[k * nqtot + sidx] = m.sig[k];
        for (int k = 0; k < 9; ++k) s.sig[el][q][k] = m.P[k] * wdv;
        s.sig[el][q][9] = m.r1(qp) * wdv;
        double s3[3];
        st.value(s3);
        for (int j = 0; j < 3; ++j) s.sig[el][q][10 + j] = s3[j] * wdv;
        if constexpr (JAC && MODEL == MODEL_NEOHOOKEAN) {
          dev::nh_base(qp, m, st, sm.ext.qs[el][q].b);
        } else if constexpr (JAC) {
          auto& D = sm.ext.D[el][q];
          sm.ext.qs[el][q].qp = qp;
          sm.ext.qs[el][q].st = st;
          sm.ext.qs[el][q].m = m;
          // p column: dsigma/dp = I (Mixed); grad-p columns: stabilization only
          double dPp[9], dr1p;
          if constexpr (MODEL == MODEL_ELASTIC) {
            for (int k = 0; k < 9; ++k) dPp[k] = (k % 4 == 0) ? 1.0 : 0.0;
            dr1p = 1.0;
          } else {
            for (int i = 0; i < 3; ++i)
              for (int jj = 0; jj < 3; ++jj) dPp[3 * i + jj] = qp.J * qp.Fi[3 * jj + i];
            dr1p = 1.0 / mp.kappa;
          }
          for (int r = 0; r < 9; ++r) D[r][9] = dPp[r] * wdv;
          D[9][9] = dr1p * wdv;
          for (int r = 10; r < 13; ++r) D[r][9] = 0.0;
          for (int I = 0; I < 3; ++I) {
            double ds3[3];
            st.dir_g(qp, I, ds3);
            for (int r = 0; r < 10; ++r) D[r][10 + I] = 0.0;
            for (int j = 0; j < 3; ++j) D[10 + j][10 + I] = ds3[j] * wdv;
          }
        }
        if constexpr (LOC) {
          for (int i = 0; i < 4; ++i) {
            double z = 0.0, zg[3] = {0.0, 0.0, 0.0};
            for (int n = 0; n < NN; ++n) {
              const double zn = sm.ext.zd[el][4 * n + i];
              z += zn * s.sh[el][q][n][0];
              for (int j = 0; j < 3; ++j) zg[j] += zn * s.sh[el][q][n][1 + j];
            }
            sm.ext.zq[el][q][i] = z;
            for (int j = 0; j < 3; ++j) sm.ext.zgq[el][q][i][j] = zg[j];
          }
          double z = 0.0, zg[3] = {0.0, 0.0, 0.0};
          for (int n = 0; n < NN; ++n) {
            const double zn = sm.ext.zc[el][n];
            z += zn * s.sh[el][q][n][0];
            for (int j = 0; j < 3; ++j) zg[j] += zn * s.sh[el][q][n][1 + j];
          }
          sm.ext.zcq[el][q][0] = z;
          for (int j = 0; j < 3; ++j) sm.ext.zcq[el][q][1 + j] = zg[j];
        }
      }
    }
    __syncthreads();

    // ---- C2: tangent columns d sig / d F_kL, one (element, point, direction) per lane
    if constexpr (JAC && MODEL != MODEL_NEOHOOKEAN && !kPointPass<MODEL, KIND>) {
      for (int t = tid; t < nv * NQ * 9; t += T) {
        const int el = t / (NQ * 9), r = t % (NQ * 9);
        const int q = r / 9, c = r % 9;
        const double wdv = s.wdv[el][q];
        auto& qs = sm.ext.qs[el][q];
        dev::Dir d;
        qs.m.dir(qs.qp, c / 3, c % 3, d);
        double ds3[3];
        qs.st.dir(qs.qp, c / 3, c % 3, ds3);
        auto& D = sm.ext.D[el][q];
        for (int rr = 0; rr < 9; ++rr) D[rr][c] = d.dP[rr] * wdv;
        D[9][c] = d.dr1 * wdv;
        for (int j = 0; j < 3; ++j) D[10 + j][c] = ds3[j] * wdv;
      }
      __syncthreads();
    }

    // ---- D: residual rows (MResidual, PResidual, Stabilization), one (element, row) per lane
    for (int t = tid; t < nv * ND; t += T) {
      const int el = t / ND, k = t % ND;
      const int n = k >> 2, i = k & 3;
      double r = 0.0;
      for (int q = 0; q < NQ; ++q) {
        const double* sg = s.sig[el][q];
        const double* sh = s.sh[el][q][n];
        if constexpr (!LOC) {
          if (i < 3) r += sg[3 * i] * sh[1] + sg[3 * i + 1] * sh[2] + sg[3 * i + 2] * sh[3];
          else r += sg[9] * sh[0] + sg[10] * sh[1] + sg[11] * sh[2] + sg[12] * sh[3];
        } else {
          // DisplacementAdjoint / PressureAdjoint weights
          // (src/goal_displacement_adjoint.cpp:37-53, src/goal_pressure_adjoint.cpp:38-49)
          if (i < 3) {
            const double z = sm.ext.zq[el][q][i];
            const double* zg = sm.ext.zgq[el][q][i];
            for (int J = 0; J < 3; ++J) r += sg[3 * i + J] * (zg[J] * sh[0] + z * sh[1 + J]);
          } else {
            r += sg[9] * sm.ext.zq[el][q][3] * sh[0];
            const double zc = sm.ext.zcq[el][q][0];
            for (int J = 0; J < 3; ++J) r += sg[10 + J] * (sm.ext.zcq[el][q][1 + J] * sh[0] + zc * sh[1 + J]);
          }
        }
      }
      a.emr[(int64_t)s.rec[el][n] * 4 + i] = r;
    }

    // ---- E: Jacobian, one (element, column) per lane
    constexpr bool TR = (KIND == ASM_ADJOINT);
    TE* em;
    if constexpr (std::is_same<TE, float>::value) em = a.em32;
    else em = a.em;
    if constexpr (JAC && MODEL == MODEL_NEOHOOKEAN && Lay::MF) {
      if constexpr (GOAL_MF_PIPE && std::is_same<TE, float>::value) {
        // matrix cores, pipelined: wave w takes elements w, w + 4, ... and
        // forms each tile's columns during the previous tile's products
        static_assert(MfTiles<NN>::DT == 2 && MfTiles<NN>::NT == 3, "two displacement tiles and one pressure tile");
        constexpr int NW = T / 64;
        const int w = tid >> 6, lane = tid & 63;
        if (w < nv) {
          TE C0[13], C1[13], C2[13], A[4];
          mf_tile_column<NN, 0, TE>(s.sh[w], s.wdv[w], sm.ext.qs[w], lane, C0);
          for (int el = w; el < nv; el += NW) {
            const int en = min(el + NW, nv - 1);
            mf_rows<NN, TE>(s.sh[el], lane, A);
            mf_tile_column<NN, 1, TE>(s.sh[el], s.wdv[el], sm.ext.qs[el], lane, C1);
            mf_tile_product<NN, 0, TR, TE>(A, C0, s.rec[el], em, lane);
            mf_tile_column<NN, 2, TE>(s.sh[el], s.wdv[el], sm.ext.qs[el], lane, C2);
            mf_tile_product<NN, 1, TR, TE>(A, C1, s.rec[el], em, lane);
            mf_tile_column<NN, 0, TE>(s.sh[en], s.wdv[en], sm.ext.qs[en], lane, C0);
            mf_tile_product<NN, 2, TR, TE>(A, C2, s.rec[el], em, lane);
          }
        }
      } else {
        // matrix cores: the (element, tile) tasks dealt round-robin over the waves
        constexpr int NTL = MfTiles<NN>::NT;
        for (int task = tid >> 6; task < nv * NTL; task += T / 64) {
          const int el = task / NTL;
          emit_tile_mfma<NN, TR, TE>(s.sh[el], s.wdv[el], sm.ext.qs[el], s.rec[el], em, tid & 63, task % NTL);
        }
      }
    } else if constexpr (JAC && MODEL == MODEL_NEOHOOKEAN) {
      if constexpr (T == 256 && G * NN <= 64 && std::is_same<TE, double>::value) {
        // displacement columns on waves 0-2, pressure columns on wave 3: the
        // closed-form column takes a different (shorter) path for pressure,
        // and keeping the two apart leaves no wave running both (measured:
        // faster for FP64 columns, slower for FP32 ones)
        const int w = tid >> 6, l = tid & 63;
        if (w < 3) {
          for (int t = tid; t < 3 * nv * NN; t += 192) {
            const int el = t / (3 * NN), r = t % (3 * NN);
            emit_column_nh<NN, NQ, TR, TE>(s.sh[el], s.wdv[el], sm.ext.qs[el], 4 * (r / 3) + r % 3, s.rec[el], em);
          }
        } else if (l < nv * NN) {
          const int el = l / NN, b = l % NN;
          emit_column_nh<NN, NQ, TR, TE>(s.sh[el], s.wdv[el], sm.ext.qs[el], 4 * b + 3, s.rec[el], em);
        }
      } else {
        for (int t = tid; t < nv * ND; t += T) {
          const int el = t / ND, col = t % ND;
          emit_column_nh<NN, NQ, TR, TE>(s.sh[el], s.wdv[el], sm.ext.qs[el], col, s.rec[el], em);
        }
      }
    } else if constexpr (JAC) {
      for (int t = tid; t < nv * ND; t += T) {
        const int el = t / ND, col = t % ND;
        emit_column<NN, NQ, TR, TE>(s.sh[el], sm.ext.D[el], col, s.rec[el], em);
      }
    }
    __syncthreads();
  }
}

// Node-range gather.  Rows 4a..4a+3 of the CSR matrix are contiguous for
// every node a, and so are the element row blocks the element pass left for a
// (records inc_ptr[a] .. inc_ptr[a+1]); for a run of consecutive nodes both
// stay contiguous.  A workgroup takes one run (planned on the host so its CSR
// values fit the LDS budget `cap`), clears an LDS copy of those values, streams
// the run's blocks in column pairs, U pair loads per lane in flight and the
// next batch prefetched, adds each value at its slot with LDS atomics, then
// stores the rows once.  Wave i of the four handles row i (equation i) of
// every node, so no two waves touch the same slot and each slot is summed in
// one wave's fixed instruction and lane order: results are bitwise
// reproducible.  No atomics to HBM, no zero fill of the CSR values.
// Column-pair loads, U per lane per batch (measured on the P2 64^3 cube,
// same GPU, gather FP64 / FP32: single-column loads with U = 2 / 4 5.79 /
// 4.22 ms; pairs with U = 1 / 2 5.64 / 4.09; pairs U = 2 / 4 5.70 / 4.25;
// four columns per lane, U = 1 / 1 5.72 / 4.58).
#ifndef GOAL_GATHER_U64
#define GOAL_GATHER_U64 1
#endif
#ifndef GOAL_GATHER_U32
#define GOAL_GATHER_U32 2
#endif
// Four-node tets (16-column rows) keep single-column loads: measured on the
// P1 J2 bar (gather FP64 / FP32), single loads with U = 2 / 4 2.74-2.76 /
// 1.93-1.95 ms, U = 4 / 8 2.76 / 2.19; pairs with U = 1 / 2 3.16 / 2.01,
// U = 2 / 4 2.81 / 2.04.
#ifndef GOAL_GATHER_K_P1
#define GOAL_GATHER_K_P1 1
#endif
#ifndef GOAL_GATHER_U64_P1
#define GOAL_GATHER_U64_P1 2
#endif
#ifndef GOAL_GATHER_U32_P1
#define GOAL_GATHER_U32_P1 4
#endif
constexpr int kGatherThreads = 256;  // four waves: one per equation
constexpr int kGatherRecords = 256;  // incidence records staged per chunk
constexpr int kGatherNodes = 64;     // nodes per run, at most

template <int NN, typename TE>
__global__ __launch_bounds__(kGatherThreads) void gather_kernel(GatherArgs g) {
  constexpr int ND = 4 * NN, RB = 4 * ND, NT = kGatherThreads, RC = kGatherRecords, NC = kGatherNodes;
  constexpr bool P2 = NN == 10;
  constexpr int KV = P2 ? 2 : GOAL_GATHER_K_P1;  // adjacent columns per load
  constexpr int U = sizeof(TE) == 8 ? (P2 ? GOAL_GATHER_U64 : GOAL_GATHER_U64_P1)
                                    : (P2 ? GOAL_GATHER_U32 : GOAL_GATHER_U32_P1);  // loads per lane per batch
  constexpr int NP = ND / KV;                                                       // loads per row
  using V2 = TE __attribute__((ext_vector_type(KV)));
  extern __shared__ double acc[];
  __shared__ int64_t srow[4 * NC + 1];
  __shared__ int64_t sinc[NC + 1];
  __shared__ double racc[4 * NC];
  __shared__ uint32_t spos2[RC * NN / 2];  // posi rows of the chunk, two uint16 per word
  __shared__ uint8_t rnode[RC];            // run-local node of each record of the chunk
  const uint16_t* spos = reinterpret_cast<const uint16_t*>(spos2);
  const int tid = threadIdx.x;
  const int eq = tid >> 6, lane = tid & 63;  // this wave's row of every node
  for (int64_t w = blockIdx.x; w < g.num_ranges; w += gridDim.x) {
    const int64_t a0 = g.ranges[w];
    const int na = (int)(g.ranges[w + 1] - a0);
    for (int t = tid; t <= 4 * na; t += NT) srow[t] = g.row_ptr[4 * a0 + t];
    for (int t = tid; t <= na; t += NT) sinc[t] = g.inc_ptr[a0 + t];
    for (int t = tid; t < 4 * na; t += NT) racc[t] = 0.0;
    __syncthreads();
    const int64_t r0 = srow[0];
    const int nv = (int)(srow[4 * na] - r0);
    const int64_t k0 = sinc[0], k1 = sinc[na];
    if (g.vals)
      for (int t = tid; t < nv; t += NT) acc[t] = 0.0;
    for (int64_t kb = k0; kb < k1; kb += RC) {
      const int m = (int)min((int64_t)RC, k1 - kb);
      // this wave's values: row `eq` of each record, ND per record
      const TE* em = nullptr;
      if (g.vals) {
        if constexpr (std::is_same<TE, float>::value) em = g.em32 + kb * RB + eq * ND;
        else em = g.em + kb * RB + eq * ND;
      }
      // a lane loads KV adjacent columns of one row (the row and record
      // strides keep them aligned); q counts loads
      const int total = g.vals ? m * NP : 0;
      // loads are unconditional (index clamped into the chunk) so all U issue
      // back to back; out-of-range lanes are dropped at the add
      V2 src[U];
      if (total > 0)
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int q = min(u * 64 + lane, total - 1);
          src[u] = *reinterpret_cast<const V2*>(em + (q / NP) * RB + KV * (q % NP));
        }
      if (g.vals) {
        // kb * NN and m * NN are even (NN is 4 or 10): whole 32-bit words
        const uint32_t* pw = reinterpret_cast<const uint32_t*>(g.posi + kb * NN);
        const int nw = m * NN / 2;
        for (int t0 = 0; t0 < nw; t0 += 4 * NT) {
          uint32_t v4[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) v4[u] = pw[min(t0 + u * NT + tid, nw - 1)];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int t = t0 + u * NT + tid;
            if (t < nw) spos2[t] = v4[u];
          }
        }
      }
      for (int t = tid; t < na; t += NT) {
        const int64_t lo = max(sinc[t], kb), hi = min(sinc[t + 1], kb + m);
        for (int64_t k = lo; k < hi; ++k) rnode[k - kb] = (uint8_t)t;
      }
      double rv[RC / 64];
      if (g.R)
#pragma unroll
        for (int u = 0; u < RC / 64; ++u) rv[u] = g.emr[(kb + min(u * 64 + lane, m - 1)) * 4 + eq];
      __syncthreads();  // acc / racc cleared, spos / rnode staged
#pragma unroll
      for (int u = 0; u < RC / 64; ++u) {
        const int j = u * 64 + lane;
        if (g.R && j < m) atomicAdd(&racc[4 * rnode[j] + eq], rv[u]);
      }
      // double-buffered stream: batch A is summed while batch B is in flight
      // and the other way round (no register copies of loads in flight)
      auto add_batch = [&](const V2* buf, int qb) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int q = qb + u * 64 + lane;
          if (q < total) {
            const int j = q / NP, c = KV * (q % NP);  // the KV columns lie in node c / 4
            const int slot = (int)(srow[4 * rnode[j] + eq] - r0) + 4 * (int)spos[j * NN + (c >> 2)] + (c & 3);
#pragma unroll
            for (int k = 0; k < KV; ++k) atomicAdd(&acc[slot + k], (double)buf[u][k]);
          }
        }
      };
      auto load_batch = [&](V2* buf, int qb) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int q = min(qb + u * 64 + lane, total - 1);
          buf[u] = *reinterpret_cast<const V2*>(em + (q / NP) * RB + KV * (q % NP));
        }
      };
      for (int q0 = 0; q0 < total; q0 += 2 * U * 64) {
        V2 nxt[U];
        load_batch(nxt, q0 + U * 64);
        add_batch(src, q0);
        if (q0 + U * 64 >= total) break;
        load_batch(src, q0 + 2 * U * 64);
        add_batch(nxt, q0 + U * 64);
      }
      __syncthreads();  // spos / rnode are rewritten by the next chunk; acc complete
    }
    if (g.R)
      for (int t = tid; t < 4 * na; t += NT) {
        const int64_t row = 4 * a0 + t;
        g.R[row] = (g.overwrite_R ? 0.0 : g.R[row]) + racc[t];
      }
    if (g.vals) {
      double* out = g.vals + r0;
      if (g.overwrite_vals)
        for (int t = tid; t < nv; t += NT) out[t] = acc[t];
      else
        for (int t = tid; t < nv; t += NT) out[t] += acc[t];
    }
    __syncthreads();
  }
}

// J2 point pass, one thread per quadrature point (e, q): the work of phases
// C1 and C2 for J2 (interpolation, kinematics, the radial return of
// src/goal_J2.cpp:71-143 with its history stores, Mixed, the residual terms
// and, for the Jacobian kinds, the 13 x 13 point tangent), written to pt for
// the element pass.  The element kernel's phase C1 runs on one lane per point
// of a group (16 lanes of a workgroup for four-node tets) under the register
// budget of the return map; here every lane takes a point and no barrier
// holds it.  The arithmetic is that of phases B, C1 and C2, in the same order.
template <int NN, int NQ>
__global__ __launch_bounds__(256, 2) void j2_point_kernel(AsmArgs a, int jac) {
  const int64_t npts = a.num_elems * NQ, nqtot = npts;
  const int PS = jac ? kPointTangent : kPointStress;
  for (int64_t sidx = (int64_t)blockIdx.x * 256 + threadIdx.x; sidx < npts; sidx += (int64_t)gridDim.x * 256) {
    const int64_t e = sidx / NQ;
    const int q = (int)(sidx % NQ);
    const int32_t* ec = a.conn + e * NN;
    double X[4][3];
    for (int v = 0; v < 4; ++v)
      for (int c = 0; c < 3; ++c) X[v][c] = a.coords[3 * (int64_t)ec[v] + c];
    double gl[4][3];
    const double det = dev::tet_gradlambda(X, gl);
    double xi[3], w;
    Quad<NQ>::point(q, xi, w);
    const double lam[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
    const MatParams mp = a.mats[a.elem_set ? a.elem_set[e] : 0];
    dev::QP qp;
    for (int k = 0; k < 9; ++k) qp.H[k] = 0.0;
    qp.p = 0.0;
    qp.g[0] = qp.g[1] = qp.g[2] = 0.0;
#pragma unroll 1
    for (int n = 0; n < NN; ++n) {
      double N, dN[3];
      dev::shape<NN>(n, lam, gl, N, dN);
      const double* xn = a.x + 4 * (int64_t)ec[n];
      const double pn = xn[3];
      qp.p += pn * N;
      for (int j = 0; j < 3; ++j) {
        qp.g[j] += pn * dN[j];
        for (int i = 0; i < 3; ++i) qp.H[3 * i + j] += xn[i] * dN[j];
      }
    }
    for (int k = 0; k < 9; ++k) qp.F[k] = qp.H[k] + ((k % 4 == 0) ? 1.0 : 0.0);
    qp.J = dev::det3(qp.F);
    dev::inv3(qp.F, qp.J, qp.Fi);
    dev::Stab st;
    st.init(qp, a.stabilization ? dev::stab_tau(X, mp) : 0.0);
    const double wdv = w * det;
    double Fpo[9];
    for (int k = 0; k < 9; ++k) Fpo[k] = a.Fp_old[k * nqtot + sidx];
    dev::J2Model m;
    m.init(qp, mp, Fpo, a.eqps_old[sidx]);
    if (m.failed()) atomicOr(a.status, 1);
    if (a.save_state) {
      a.eqps[sidx] = m.alpha;
      if (m.plastic) {
        double Fpn[9];
        m.new_Fp(Fpo, Fpn);
        for (int k = 0; k < 9; ++k) a.Fp[k * nqtot + sidx] = Fpn[k];
      }
      for (int k = 0; k < 9; ++k) a.sigma[k * nqtot + sidx] = m.sig[k];
    }
    // value k at o[64 k]: a wave's 64 consecutive points store each value as
    // one contiguous run
    double* o = a.pt + point_state_index(sidx, 0, PS);
    for (int k = 0; k < 9; ++k) o[64 * k] = m.P[k] * wdv;
    o[64 * 9] = m.r1(qp) * wdv;
    double s3[3];
    st.value(s3);
    for (int j = 0; j < 3; ++j) o[64 * (10 + j)] = s3[j] * wdv;
    if (!jac) continue;
    // p column: dsigma/dp = I (Mixed); grad-p columns: stabilization only
    for (int i = 0; i < 3; ++i)
      for (int jj = 0; jj < 3; ++jj) o[64 * (130 + 3 * i + jj)] = (qp.J * qp.Fi[3 * jj + i]) * wdv;
    o[64 * 139] = (1.0 / mp.kappa) * wdv;
    for (int I = 0; I < 3; ++I) {
      double ds3[3];
      st.dir_g(qp, I, ds3);
      for (int j = 0; j < 3; ++j) o[64 * (140 + 3 * I + j)] = ds3[j] * wdv;
    }
    // F columns: directional derivatives through the return map, one
    // direction at a time (the derivatives pick the direction's rows and
    // columns by selects, so the loop stays rolled and the registers hold
    // one direction's temporaries)
#pragma unroll 1
    for (int c = 0; c < 9; ++c) {
      dev::Dir d;
      m.dir(qp, c / 3, c % 3, d);
      double ds3[3];
      st.dir(qp, c / 3, c % 3, ds3);
      double* oc = o + 64 * (13 + 13 * c);
      for (int rr = 0; rr < 9; ++rr) oc[64 * rr] = d.dP[rr] * wdv;
      oc[64 * 9] = d.dr1 * wdv;
      for (int j = 0; j < 3; ++j) oc[64 * (10 + j)] = ds3[j] * wdv;
    }
  }
}

template <int NN, int NQ, int MODEL, int KIND, typename TE>
int occupancy() {
  static const int occ = [] {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, reinterpret_cast<const void*>(&element_kernel<NN, NQ, MODEL, KIND, TE>), Layout<NN, NQ, MODEL>::T,
            0) != hipSuccess ||
        n < 1) {
      n = 1;
      (void)hipGetLastError();  // the fallback stands; keep the error off the next launch
    }
    return n;
  }();
  return occ;
}

// Persistent grid: the workgroups that fit on the device at once (a multiple
// of 8 so the XCD slab schedule applies), at most one per element group.
template <int NN, int NQ, int MODEL, int KIND, typename TE>
int resident_grid(int64_t ne) {
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int64_t groups = (ne + Layout<NN, NQ, MODEL>::G - 1) / Layout<NN, NQ, MODEL>::G;
  int64_t want = std::min<int64_t>((int64_t)cus * occupancy<NN, NQ, MODEL, KIND, TE>(), groups);
  return (int)std::max<int64_t>(8, (want + 7) / 8 * 8);
}

template <int NN, int NQ, int MODEL, int KIND, typename TE>
hipError_t launch4(const AsmArgs& a, int blocks, hipStream_t s) {
  if constexpr (kPointPass<MODEL, KIND>) {
    if (!a.pt) return hipErrorInvalidValue;
    const int64_t npts = a.num_elems * NQ;
    const int pb = (int)std::min<int64_t>((npts + 255) / 256, 65536);
    hipLaunchKernelGGL((j2_point_kernel<NN, NQ>), dim3(pb), dim3(256), 0, s, a,
                       (KIND == ASM_JACOBIAN || KIND == ASM_ADJOINT) ? 1 : 0);
  }
  if (blocks <= 0) blocks = resident_grid<NN, NQ, MODEL, KIND, TE>(a.num_elems);
  hipLaunchKernelGGL((element_kernel<NN, NQ, MODEL, KIND, TE>), dim3(blocks), dim3(Layout<NN, NQ, MODEL>::T), 0, s,
                     a);
  return hipGetLastError();
}

// a.em32 set: FP32 element matrices (Jacobian kinds only).
template <int NN, int NQ, int MODEL>
hipError_t launch3(int kind, const AsmArgs& a, int blocks, hipStream_t s) {
  const bool f32 = a.em32 != nullptr;
  switch (kind) {
    case ASM_RESIDUAL: return launch4<NN, NQ, MODEL, ASM_RESIDUAL, double>(a, blocks, s);
    case ASM_JACOBIAN:
      return f32 ? launch4<NN, NQ, MODEL, ASM_JACOBIAN, float>(a, blocks, s)
                 : launch4<NN, NQ, MODEL, ASM_JACOBIAN, double>(a, blocks, s);
    case ASM_ADJOINT:
      return f32 ? launch4<NN, NQ, MODEL, ASM_ADJOINT, float>(a, blocks, s)
                 : launch4<NN, NQ, MODEL, ASM_ADJOINT, double>(a, blocks, s);
    case ASM_LOCALIZE: return launch4<NN, NQ, MODEL, ASM_LOCALIZE, double>(a, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

template <int NN, int NQ>
hipError_t launch2(int kind, const AsmArgs& a, int blocks, hipStream_t s) {
  switch (a.model) {
    case MODEL_ELASTIC: return launch3<NN, NQ, MODEL_ELASTIC>(kind, a, blocks, s);
    case MODEL_NEOHOOKEAN: return launch3<NN, NQ, MODEL_NEOHOOKEAN>(kind, a, blocks, s);
    case MODEL_J2: return launch3<NN, NQ, MODEL_J2>(kind, a, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

// ------------------------------------------------------------------ QoI
// One thread per element.  AvgDisp: sum_q (sum_i u_i) w dv / 3
// (src/goal_avg_disp.cpp:16-21); AvgVM: sum_q vm(sigma) w dv over one element
// set (src/goal_avg_vm.cpp:51-61, src/goal_von_mises.cpp:6-17).  With dMdu set,
// the derivative of the element value w.r.t. the element DOFs is added to
// dMdu (QoI<FADT>::scatter, src/goal_qoi.cpp:63-76).
template <int NN, int NQ, int MODEL>
__global__ __launch_bounds__(256) void qoi_kernel(QoiArgs a) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.num_elems) return;
  const int32_t* ec = a.conn + e * NN;
  const int set = a.elem_set ? a.elem_set[e] : 0;
  double val = 0.0;
  if (a.kind == QOI_AVG_VM && a.qoi_set >= 0 && set != a.qoi_set) {
    a.elem_value[e] = 0.0;
    return;
  }
  double X[4][3];
  for (int v = 0; v < 4; ++v)
    for (int c = 0; c < 3; ++c) X[v][c] = a.coords[3 * (int64_t)ec[v] + c];
  double gl[4][3];
  const double det = dev::tet_gradlambda(X, gl);
  const MatParams mp = a.mats[set];
  const int64_t nqtot = a.num_elems * NQ;
  for (int q = 0; q < NQ; ++q) {
    double xi[3], w;
    Quad<NQ>::point(q, xi, w);
    const double lam[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
    const double wdv = w * det;
    if (a.kind == QOI_AVG_DISP) {
      double s = 0.0;
      for (int n = 0; n < NN; ++n) {
        double N, dN[3];
        dev::shape<NN>(n, lam, gl, N, dN);
        const int64_t base = 4 * (int64_t)ec[n];
        s += (a.x[base] + a.x[base + 1] + a.x[base + 2]) * N;
        if (a.dMdu)
          for (int i = 0; i < 3; ++i) unsafeAtomicAdd(a.dMdu + base + i, N * wdv / 3.0);
      }
      val += s * wdv / 3.0;
      continue;
    }
    dev::QP qp;
    for (int k = 0; k < 9; ++k) qp.H[k] = 0.0;
    qp.p = 0.0;
    qp.g[0] = qp.g[1] = qp.g[2] = 0.0;
    double Ns[NN], dNs[NN][3];
    for (int n = 0; n < NN; ++n) {
      dev::shape<NN>(n, lam, gl, Ns[n], dNs[n]);
      const int64_t base = 4 * (int64_t)ec[n];
      qp.p += a.x[base + 3] * Ns[n];
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) qp.H[3 * i + j] += a.x[base + i] * dNs[n][j];
    }
    for (int k = 0; k < 9; ++k) qp.F[k] = qp.H[k] + ((k % 4 == 0) ? 1.0 : 0.0);
    qp.J = dev::det3(qp.F);
    dev::inv3(qp.F, qp.J, qp.Fi);
    typename ModelOf<MODEL>::type m;
    const int64_t sidx = e * NQ + q;
    if constexpr (MODEL == MODEL_J2) {
      double Fpo[9];
      for (int k = 0; k < 9; ++k) Fpo[k] = a.Fp_old[k * nqtot + sidx];
      m.init(qp, mp, Fpo, a.eqps_old[sidx]);
      if (m.failed()) atomicOr(a.status, 1);
    } else {
      m.init(qp, mp);
    }
    const double* s = m.sig;
    const double s1 = (s[0] - s[4]) * (s[0] - s[4]);
    const double s2 = (s[4] - s[8]) * (s[4] - s[8]);
    const double s3 = (s[8] - s[0]) * (s[8] - s[0]);
    const double s4 = s[1] * s[1], s5 = s[5] * s[5], s6 = s[6] * s[6];
    const double vm = sqrt(0.5 * (s1 + s2 + s3 + 6.0 * (s4 + s5 + s6)));
    val += vm * wdv;
    if (a.dMdu) {
      // d vm = (1/(2 vm)) * 1/2 * d(s1+s2+s3+6(s4+s5+s6))
      auto dvm = [&](const double* ds) {
        const double t = 2.0 * (s[0] - s[4]) * (ds[0] - ds[4]) + 2.0 * (s[4] - s[8]) * (ds[4] - ds[8]) +
                         2.0 * (s[8] - s[0]) * (ds[8] - ds[0]) +
                         12.0 * (s[1] * ds[1] + s[5] * ds[5] + s[6] * ds[6]);
        return 0.25 * t / vm;
      };
      double gF[9];
      for (int c = 0; c < 9; ++c) {
        dev::Dir d;
        m.dir(qp, c / 3, c % 3, d);
        gF[c] = dvm(d.dsig);
      }
      double Ip[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      const double gp = dvm(Ip);
      for (int n = 0; n < NN; ++n) {
        const int64_t base = 4 * (int64_t)ec[n];
        for (int k = 0; k < 3; ++k)
          unsafeAtomicAdd(a.dMdu + base + k,
                          (gF[3 * k] * dNs[n][0] + gF[3 * k + 1] * dNs[n][1] + gF[3 * k + 2] * dNs[n][2]) * wdv);
        unsafeAtomicAdd(a.dMdu + base + 3, gp * Ns[n] * wdv);
      }
    }
  }
  a.elem_value[e] = val;
}

template <int NN, int NQ>
hipError_t launch_qoi2(const QoiArgs& a, hipStream_t s) {
  dim3 grid((unsigned)((a.num_elems + 255) / 256)), block(256);
  if (a.num_elems == 0) return hipSuccess;
  switch (a.model) {
    case MODEL_ELASTIC: hipLaunchKernelGGL((qoi_kernel<NN, NQ, MODEL_ELASTIC>), grid, block, 0, s, a); break;
    case MODEL_NEOHOOKEAN: hipLaunchKernelGGL((qoi_kernel<NN, NQ, MODEL_NEOHOOKEAN>), grid, block, 0, s, a); break;
    case MODEL_J2: hipLaunchKernelGGL((qoi_kernel<NN, NQ, MODEL_J2>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_element(int kind, int order, int quad_order, const AsmArgs& a, int blocks, hipStream_t s) {
  if (a.num_elems == 0) return hipSuccess;
  if (order == 1 && quad_order == 1) return launch2<4, 1>(kind, a, blocks, s);
  if (order == 1 && quad_order == 2) return launch2<4, 4>(kind, a, blocks, s);
  if (order == 2 && quad_order == 1) return launch2<10, 1>(kind, a, blocks, s);
  if (order == 2 && quad_order == 2) return launch2<10, 4>(kind, a, blocks, s);
  return hipErrorInvalidValue;
}

hipError_t launch_gather(int order, const GatherArgs& g, int blocks, hipStream_t s) {
  if (g.num_ranges == 0) return hipSuccess;
  if (blocks < 1 || g.cap < 1) return hipErrorInvalidValue;
  const size_t lds = g.vals ? sizeof(double) * (size_t)g.cap : 0;
  const bool f32 = g.em32 != nullptr;
  if (order == 1 && f32)
    hipLaunchKernelGGL((gather_kernel<4, float>), dim3(blocks), dim3(kGatherThreads), lds, s, g);
  else if (order == 1)
    hipLaunchKernelGGL((gather_kernel<4, double>), dim3(blocks), dim3(kGatherThreads), lds, s, g);
  else if (f32)
    hipLaunchKernelGGL((gather_kernel<10, float>), dim3(blocks), dim3(kGatherThreads), lds, s, g);
  else
    hipLaunchKernelGGL((gather_kernel<10, double>), dim3(blocks), dim3(kGatherThreads), lds, s, g);
  return hipGetLastError();
}

namespace {
template <int NN, int NQ>
int group_size_of(int model) {
  switch (model) {
    case MODEL_ELASTIC: return Layout<NN, NQ, MODEL_ELASTIC>::G;
    case MODEL_NEOHOOKEAN: return Layout<NN, NQ, MODEL_NEOHOOKEAN>::G;
    default: return Layout<NN, NQ, MODEL_J2>::G;
  }
}
}  // namespace

int point_stride(int kind) { return (kind == ASM_JACOBIAN || kind == ASM_ADJOINT) ? kPointTangent : kPointStress; }
bool uses_point_pass(int model, int kind) { return model == MODEL_J2 && kind != ASM_LOCALIZE; }

int element_group_size(int order, int quad_order, int model) {
  const int nq = quad_order == 1 ? 1 : 4;
  if (order == 2) return nq == 1 ? group_size_of<10, 1>(model) : group_size_of<10, 4>(model);
  return nq == 1 ? group_size_of<4, 1>(model) : group_size_of<4, 4>(model);
}

namespace {
template <int NN, int NQ, int MODEL>
int resident_of_kind(int kind, bool f32, int64_t ne) {
  const bool jac = kind == ASM_JACOBIAN || kind == ASM_ADJOINT;
  if (jac && f32)
    return kind == ASM_JACOBIAN ? resident_grid<NN, NQ, MODEL, ASM_JACOBIAN, float>(ne)
                                : resident_grid<NN, NQ, MODEL, ASM_ADJOINT, float>(ne);
  switch (kind) {
    case ASM_RESIDUAL: return resident_grid<NN, NQ, MODEL, ASM_RESIDUAL, double>(ne);
    case ASM_JACOBIAN: return resident_grid<NN, NQ, MODEL, ASM_JACOBIAN, double>(ne);
    case ASM_ADJOINT: return resident_grid<NN, NQ, MODEL, ASM_ADJOINT, double>(ne);
    default: return resident_grid<NN, NQ, MODEL, ASM_LOCALIZE, double>(ne);
  }
}
}  // namespace

int element_resident_grid(int kind, int order, int quad_order, int model, bool f32, int64_t ne) {
  if (model != MODEL_NEOHOOKEAN) return 8;  // chunked assembly is neo-Hookean only
  if (order == 2) return quad_order == 1 ? resident_of_kind<10, 1, MODEL_NEOHOOKEAN>(kind, f32, ne)
                                         : resident_of_kind<10, 4, MODEL_NEOHOOKEAN>(kind, f32, ne);
  return quad_order == 1 ? resident_of_kind<4, 1, MODEL_NEOHOOKEAN>(kind, f32, ne)
                         : resident_of_kind<4, 4, MODEL_NEOHOOKEAN>(kind, f32, ne);
}

int gather_nodes_per_range() { return kGatherNodes; }

int gather_grid(int order, bool f32, int cap, int64_t num_ranges) {
  int dev = 0, cus = 256, occ = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const void* f = order == 1 ? (f32 ? reinterpret_cast<const void*>(&gather_kernel<4, float>)
                                    : reinterpret_cast<const void*>(&gather_kernel<4, double>))
                             : (f32 ? reinterpret_cast<const void*>(&gather_kernel<10, float>)
                                    : reinterpret_cast<const void*>(&gather_kernel<10, double>));
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f, kGatherThreads, sizeof(double) * (size_t)cap) !=
          hipSuccess ||
      occ < 1) {
    occ = 1;
    (void)hipGetLastError();  // the fallback stands; keep the error off the next launch
  }
  return (int)std::max<int64_t>(1, std::min<int64_t>(num_ranges, (int64_t)cus * occ));
}

hipError_t launch_qoi(int order, int quad_order, const QoiArgs& a, hipStream_t s) {
  if (order == 1 && quad_order == 1) return launch_qoi2<4, 1>(a, s);
  if (order == 1 && quad_order == 2) return launch_qoi2<4, 4>(a, s);
  if (order == 2 && quad_order == 1) return launch_qoi2<10, 1>(a, s);
  if (order == 2 && quad_order == 2) return launch_qoi2<10, 4>(a, s);
  return hipErrorInvalidValue;
}

}  // namespace goal
