// Launch interface between the host library (goal_capi.cpp) and the CDNA4
// kernels.  Everything here is plain data: device pointers and sizes.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>

namespace goal {

enum Model : int { MODEL_ELASTIC = 0, MODEL_NEOHOOKEAN = 1, MODEL_J2 = 2 };

// Per element-set material constants, derived on the host the way the
// reference evaluators derive them in set_elem_set():
//   mu = E/(2(1+nu)), kappa = E/(3(1-2nu)), lambda = E nu/((1+nu)(1-2nu))
//   (src/goal_neohookean.cpp:44-52, src/goal_J2.cpp:54-64,
//    src/goal_presidual.cpp:43-51, src/goal_stabilization.cpp:45-52)
struct MatParams {
  double mu, kappa, lambda, K, Y, c0;
};

// What a fused element pass computes.
enum AsmKind : int {
  ASM_RESIDUAL = 0,  // ST chain: R += r_e                        (Primal::compute_resid)
  ASM_JACOBIAN = 1,  // FADT chain: R += r_e, dRdu += k_e         (Primal::compute_jacob)
  ASM_ADJOINT = 2,   // FADT chain, transposed scatter: dRdu^T    (NestedAdjoint::compute_adjoint)
  ASM_LOCALIZE = 3,  // ST chain, weights z_diff*N_a / z_coarse*N_a (NestedAdjoint::localize)
};

struct AsmArgs {
  int64_t num_elems;
  const int32_t* conn;      // [ne][npe]
  const double* coords;     // [num_nodes][3]
  const double* x;          // [num_dofs] u,p interleaved per node
  const int32_t* elem_set;  // [ne]
  const MatParams* mats;    // [num_sets]
  const int32_t* rec;       // [ne][npe] incidence record of (element, local node)
  // Element residual rows and matrix row blocks, node-major: the block of
  // (element e, local node a) is stored at record k = rec[e][a], so the blocks
  // a node sums are contiguous.  emr[k][4], em[k][4][4*npe] (for
  // ASM_ADJOINT the row blocks of the transposed element matrix).
  double* emr;
  double* em;               // null: residual only
  float* em32;              // FP32 element matrices instead of em (FP32 tangent)
  // per quadrature point history, structure of arrays: comp * (ne*nq) + e*nq + q
  const double* Fp_old;
  const double* eqps_old;
  double* Fp;
  double* eqps;
  double* sigma;
  // DWR weights (ASM_LOCALIZE): nodal fields laid out like x
  const double* z_diff;
  const double* z_coarse;
  int model;
  int stabilization;
  int save_state;
  int xcd_remap;            // 1: give each XCD a contiguous slab of elements
  int* status;              // bit 0: J2 return map failed
  // neo-Hookean only: this launch takes element groups [group_lo, group_hi)
  // (group_hi == 0: all of them), so an assembly can be split into chunks
  int64_t group_lo, group_hi;
  // J2 (all kinds but ASM_LOCALIZE): the point states, written by a
  // thread-per-point pass before the element pass reads them.  Per point
  // p = e * nq + q, PS values, all times w dv: the generalized stress
  // [P, r1, s3] (13) and, for the Jacobian kinds, the nonzero entries of the
  // point tangent D (13 x 13; rows P, r1, s3; columns F, p, grad p):
  //   13 + 13 c + r   D[r][c], c < 9 (the F directions, every row)
  //   130 + r         D[r][9], r < 10 (the p column; its s3 rows are zero)
  //   140 + 3 I + j   D[10 + j][10 + I] (grad p acts on s3 only)
  // Blocked by 64 points so that a wave's store of one value is one
  // contiguous 512-byte run: value k of point p sits at point_state_index.
  double* pt;
};

// Doubles per row of the element kernel's LDS shape table (N, dN/dX and one
// pad): with a 40-byte row stride the rows that one wave reads at once fall on
// different LDS banks (a 32-byte stride puts every eighth row on the same one).
constexpr int kShapeRow = 5;

// Doubles per point in AsmArgs::pt: residual kinds, Jacobian kinds.
constexpr int kPointStress = 13;
constexpr int kPointTangent = 13 + 117 + 10 + 9;
__host__ __device__ inline int64_t point_state_index(int64_t p, int k, int ps) {
  return ((p >> 6) * ps + k) * 64 + (p & 63);
}
// Length of AsmArgs::pt for npts points: whole blocks of 64.
inline int64_t point_state_doubles(int64_t npts, int ps) { return (npts + 63) / 64 * 64 * ps; }
int point_stride(int kind);
bool uses_point_pass(int model, int kind);

// Node-owner gather of element residuals / matrices into R and the CSR values.
struct GatherArgs {
  int64_t num_ranges;
  const int32_t* ranges;    // [num_ranges + 1] first node of each run of nodes
  int cap;                  // LDS budget: max CSR values of one run (>= 16 * max degree)
  const int64_t* inc_ptr;   // [num_nodes + 1] node -> its incidence records
  const uint16_t* posi;     // [records][npe] position of each element node in adj(node)
  const int64_t* row_ptr;   // [num_dofs + 1]
  const double* emr;        // node-major, see AsmArgs
  const double* em;         // null: residual only
  const float* em32;        // FP32 element matrices instead of em
  double* R;                // null: skip
  double* vals;             // null: skip
  int overwrite_R;          // 1: R = sum (R was pending zero), 0: R += sum
  int overwrite_vals;
};

// QoI selection (src/goal_mechanics.cpp:148-167)
enum QoiKind : int { QOI_AVG_DISP = 0, QOI_AVG_VM = 1 };

struct QoiArgs {
  int64_t num_elems;
  const int32_t* conn;
  const double* coords;
  const double* x;
  const int32_t* elem_set;
  const MatParams* mats;
  const double* Fp_old;
  const double* eqps_old;
  int model;
  int kind;
  int qoi_set;              // element set integrated over (AvgVM), -1: all
  double* elem_value;       // [ne] per-element QoI contribution
  double* dMdu;             // [num_dofs] or null: derivative scatter
  int* status;
};

// Launchers (goal_assembly.hip).  order: 1 or 2; quad_order: 1 or 2.
hipError_t launch_element(int kind, int order, int quad_order, const AsmArgs& a, int blocks, hipStream_t s);
hipError_t launch_gather(int order, const GatherArgs& g, int blocks, hipStream_t s);
// Nodes per run the gather kernel accepts (plan_gather_ranges must respect it).
int gather_nodes_per_range();
// Persistent grid for the gather: the workgroups resident at once, at most one per run.
int gather_grid(int order, bool f32, int cap, int64_t num_ranges);
hipError_t launch_qoi(int order, int quad_order, const QoiArgs& a, hipStream_t s);
// Elements per workgroup iteration of the element kernel, and its persistent
// grid for ne elements (the workgroups resident at once, at most one per group).
int element_group_size(int order, int quad_order, int model);
int element_resident_grid(int kind, int order, int quad_order, int model, bool f32, int64_t ne);

// Utility kernels (goal_util.hip)
// Device linear solve (goal_solve.hip).
hipError_t launch_spmv(int64_t nrows, const int64_t* rp, const int32_t* cols, const double* vals, const double* x,
                       double* y, hipStream_t s);
hipError_t launch_block_inverse(int64_t nn, const int64_t* rp, const int32_t* cols, const double* vals,
                                double* binv, int* status, hipStream_t s);
hipError_t launch_block_apply(int64_t nn, const double* binv, const double* x, double* y, hipStream_t s);
// aggregation multigrid: BSR (4x4 node blocks) levels, Galerkin sums, transfers, smoother
hipError_t launch_bsr_block_inverse(int64_t nb, const int64_t* bptr, const int32_t* bcol, const double* bval,
                                    double* binv, int* status, hipStream_t s);
hipError_t launch_bsr_spmv(int64_t nb, const int64_t* bptr, const int32_t* bcol, const double* bval, const double* x,
                           const double* b, double* y, hipStream_t s);  // y = A x, or b - A x with b
hipError_t launch_spmv_resid(int64_t nrows, const int64_t* rp, const int32_t* cols, const double* vals,
                             const double* x, const double* b, double* y, hipStream_t s);  // y = b - A x
hipError_t launch_galerkin(int64_t ncb, const int64_t* gptr, const int64_t* gidx, const int32_t* brow,
                           const int64_t* rp, const double* fvals, double* cvals, hipStream_t s);
hipError_t launch_restrict(int64_t nc, const int64_t* aptr, const int32_t* amem, const double* r, double* bc,
                           hipStream_t s);
hipError_t launch_prolong(int64_t nf, const int32_t* agg, const double* xc, double* x, hipStream_t s);
hipError_t launch_cheb(int64_t nn, const double* binv, const double* r, double c1, double c2, double* d, double* x,
                       hipStream_t s);
hipError_t launch_dense_matvec(int n, const double* M, const double* x, double* y, hipStream_t s);
// out[j] = V_j . w (j < k); partial holds k * multi_dot_blocks(n) doubles
int multi_dot_blocks(int64_t n);
hipError_t launch_multi_dot(int64_t n, int k, const double* V, int64_t ldv, const double* w, double* partial,
                            double* out, hipStream_t s);
hipError_t launch_multi_axpy(int64_t n, int k, const double* V, int64_t ldv, const double* c, double sign, double* w,
                             hipStream_t s);
hipError_t launch_scale(int64_t n, double a, const double* x, double* y, hipStream_t s);
hipError_t launch_rsub(int64_t n, const double* b, double* y, hipStream_t s);

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s);
hipError_t launch_copy(double* dst, const double* src, int64_t n, hipStream_t s);
hipError_t launch_axpy(double* y, const double* x, double a, int64_t n, hipStream_t s);
// Discretization graph on the device (goal_graph.hip): temp_bytes sizes the
// radix-sort scratch for all three steps; graph_adjacency leaves the unique
// node-pair keys (count in *num_unique, on the device) in `keys`.
hipError_t graph_temp_bytes(int64_t ne, int npe, int64_t nn, size_t* bytes);
hipError_t graph_adjacency(const int32_t* conn, int64_t ne, int npe, int64_t nn, void* temp, size_t temp_bytes,
                           uint64_t* keys, uint64_t* keys_sorted, int64_t* num_unique, hipStream_t s);
hipError_t graph_rows(const uint64_t* ukeys, int64_t nu, int64_t nn, int64_t* adj_ptr, int32_t* adj, hipStream_t s);
hipError_t graph_incidence(const int32_t* conn, int64_t ne, int npe, int64_t nn, void* temp, size_t temp_bytes,
                           int32_t* keys_out, int32_t* slots, int32_t* inc, int64_t* inc_ptr, hipStream_t s);
hipError_t graph_positions(const int32_t* conn, int64_t ne, int npe, const int64_t* adj_ptr, const int32_t* adj,
                           const int32_t* inc, uint16_t* pos, int32_t* rec, uint16_t* posi, hipStream_t s);
hipError_t launch_expand_csr(const int64_t* node_adj_ptr, const int32_t* node_adj, int64_t num_nodes,
                             int64_t* row_ptr, int32_t* cols, hipStream_t s);
// Deterministic two-pass reductions into out[0] (device).  scratch holds at
// least kReduceBlocks doubles.  op: 0 sum x, 1 sum x^2, 2 sum x*y, 3 sum |x|.
constexpr int kReduceBlocks = 512;
hipError_t launch_reduce(int op, const double* x, const double* y, int64_t n, double* scratch, double* out,
                         hipStream_t s);
// Dirichlet rows (src/goal_dbcs.cpp:39-99)
hipError_t launch_dbc(const int32_t* nodes, const double* values, int64_t n, int comp, const double* x,
                      double* R, double* dMdu, const int64_t* row_ptr, const int32_t* cols,
                      double* vals, hipStream_t s);
// Traction loads (src/goal_tbcs.cpp:30-66): R[rows[k]] -= values[k], rows distinct
hipError_t launch_sub_rows(const int64_t* rows, const double* values, int64_t n, double* R, hipStream_t s);
// out[k] = src[rows[k]]
hipError_t launch_take_rows(const int64_t* rows, const double* src, int64_t n, double* out, hipStream_t s);

// DWR (goal_dwr.hip)
hipError_t launch_set_coarse(double* z, const int32_t* edge_parents, int64_t num_old, int64_t num_new,
                             hipStream_t s);
// Uniform refinement (8 children per tet) from the base mesh's DOF CSR graph:
// ud[a] = number of neighbours of node a above a; with ebase = its exclusive
// prefix sum (ebase[nv] = number of edges), launch_refine writes the edges'
// parent pairs and midpoints (fX rows nv..) and the children (fconn [8 ne][4],
// positively oriented; fX rows 0..nv-1 must hold the base coordinates).
// status |= 1 if an element edge is missing from the graph.
hipError_t launch_upper_degree(const int64_t* row_ptr, const int32_t* cols, int64_t nv, int64_t* ud, hipStream_t s);
hipError_t launch_refine(const int32_t* conn, const double* X, const int64_t* row_ptr, const int32_t* cols,
                         const int64_t* ebase, int64_t ne, int64_t nv, int32_t* edge_parents, double* fX,
                         int32_t* fconn, int32_t* status, hipStream_t s);
hipError_t launch_subtract(const double* a, const double* b, double* out, int64_t n, hipStream_t s);
hipError_t launch_compute_error(int order, const int32_t* conn, int64_t ne, const double* e,
                                double* eta, hipStream_t s);
hipError_t launch_sum_contribs(const int32_t* vertex_nodes, int64_t nv, const double* e, double* scratch,
                               double* out, hipStream_t s);
hipError_t launch_set_error(const double* eta_child, int64_t num_base, double* eta_base, hipStream_t s);
// fine[k][child*nq + q] = base[k][(child/8)*nq + q] for ncomp components
hipError_t launch_transfer_state(const double* base, double* fine, int ncomp, int64_t num_fine, int nq,
                                 int64_t base_pts, hipStream_t s);


// Learned indicator (goal_learn.hip).  Parameter vector layout (FP32 master
// and its BF16 copy share it): W1[64][12], b1[64], W2[64][64], b2[64], W3[64], b3.
constexpr int kMlpIn = 12, kMlpHidden = 64;
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpHidden * kMlpIn;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHidden;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHidden;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHidden;
constexpr int kMlpParams = kMlpOffB3 + 1;
// Normalization vector: input shift mu[12], input scale 1/s[12], output
// shift and scale, the training feature box lo[12], hi[12], and c, the weight
// of raw feature 0 added to the output:
//   y = net((x - mu) * is) * s_y + mu_y + c * x[0].
constexpr int kMlpOffLo = 2 * kMlpIn + 2;
constexpr int kMlpOffHi = kMlpOffLo + kMlpIn;
constexpr int kMlpOffC = kMlpOffHi + kMlpIn;
constexpr int kMlpNorms = kMlpOffC + 1;

struct MlpArgs {
  float* params;       // [kMlpParams] FP32 master
  uint16_t* wq;        // [kMlpParams] BF16 copy (weight entries are the ones read)
  const float* norms;  // [kMlpNorms]
  float* adam_m;       // [kMlpParams]
  float* adam_v;       // [kMlpParams]
};

struct FeatureArgs {
  int64_t num_elems, num_nodes;
  int npe;
  const int32_t* conn;
  const double* coords;
  const double* x;
  const int32_t* elem_set;
  const MatParams* mats;
  const double* elem_vm;    // [ne] integral of the von Mises stress over each element
  const double* eqps;       // [ne*nq] or null
  int nq;
  const int64_t* inc_ptr;   // node -> incidence records
  const int32_t* inc;       // record -> e * npe + local node
  double qoi_x;             // QoI subdomain {x < qoi_x}
  double* elem_scratch;     // [ne][10] volume, grad u
  double* node_scratch;     // [num_nodes][10] recovered vm, grad u
  float* feat;              // [ne][kMlpIn]
};

hipError_t launch_features(const FeatureArgs& a, hipStream_t s);
hipError_t launch_mlp_forward(const MlpArgs& a, const float* x, int64_t n, float* y, int* ood, hipStream_t s);
// Blocks of the gradient kernel for a batch of B rows (partial holds that many
// kMlpParams rows and partial_loss that many floats).
int mlp_grad_blocks(int64_t B);
hipError_t launch_mlp_grad(const MlpArgs& a, const float* x, const float* y, const int32_t* idx, int64_t B,
                           float* partial, float* partial_loss, float* grad, float* loss, hipStream_t s);
hipError_t launch_mlp_adam(const MlpArgs& a, const float* grad, int step, float lr, hipStream_t s);
hipError_t launch_mlp_round(const MlpArgs& a, hipStream_t s);

}  // namespace goal
