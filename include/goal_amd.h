/*
 * goal_amd: MI355X (gfx950) element assembly and nested-adjoint DWR error
 * estimation for goal's mixed u-p Total Lagrangian tetrahedra.
 *
 * C ABI: plain pointers, sizes and opaque handles.  Every entry point returns
 * a status (GOAL_OK = 0) and never aborts; goal_last_error() describes the
 * last failure on the calling thread.  Handles own device memory; pointers
 * named *_dev are device pointers, all others are host pointers.  Work is
 * queued on one HIP stream per process; calls that hand data back to the host
 * synchronise it.
 *
 * Each declaration names the reference interface it replaces (file:line in
 * bgranzow/goal).
 */
#ifndef GOAL_AMD_H
#define GOAL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GOAL_OK = 0,
  GOAL_ERR_INVALID = 1,       /* bad argument / input (Teuchos validation)        */
  GOAL_ERR_HIP = 2,           /* HIP runtime failure                              */
  GOAL_ERR_J2_RETURN_MAP = 3, /* fail("J2: return mapping failed") src/goal_J2.cpp:119-120 */
  GOAL_ERR_NO_DEVICE = 4,     /* no HIP device visible                            */
  GOAL_ERR_ALLOC = 5          /* device allocation failed                         */
};

enum { GOAL_MODEL_ELASTIC = 0, GOAL_MODEL_NEOHOOKEAN = 1, GOAL_MODEL_J2 = 2 };
enum { GOAL_QOI_AVG_DISP = 0, GOAL_QOI_AVG_VM = 1 };

/* Message of the last failure on this thread (goal::fail, src/goal_control.cpp:91-99). */
const char* goal_last_error(void);
/* Library version string. */
const char* goal_version(void);

/* ---------------------------------------------------------------- device */
/* Device memory and stream helpers for callers without their own runtime. */
int goal_device_alloc(int64_t bytes, void** ptr_dev);
int goal_device_free(void* ptr_dev);
int goal_memcpy_h2d(void* dst_dev, const void* src, int64_t bytes);
int goal_memcpy_d2h(void* dst, const void* src_dev, int64_t bytes);
int goal_sync(void);
/* dst[i] = src[i] for i < n with the library's streaming copy kernel (async). */
int goal_device_copy(double* dst_dev, const double* src_dev, int64_t n);
/* p[i] = value for i < n (async; a fill kernel, so it may be graph-captured). */
int goal_device_fill(double* p_dev, int64_t n, double value);
/* y[i] = a * x[i] for i < n (async; y may be x). */
int goal_device_scale(double* y_dev, const double* x_dev, int64_t n, double a);
/* HIP events on the library stream: record into slot 0..255, elapsed a->b in ms. */
int goal_event_record(int slot);
int goal_event_elapsed(int slot_a, int slot_b, float* ms);
/* Device name and compute-unit count. */
int goal_device_info(char* name, int name_len, int* num_cus);

/* ------------------------------------------------------------ host mesh */
/* Host-only mesh layer (no device needed).  Call with NULL arrays to get
 * sizes[0..1] = num_nodes, num_elems, then again with arrays of that size. */
int goal_mesh_box(int nx, int ny, int nz, double lx, double ly, double lz, int order, double jitter, uint64_t seed,
                  int64_t* sizes, double* coords, int32_t* conn);
/* Host CSR graph of a mesh (the same builder goal_disc_create uses).  Call
 * with NULL row_ptr/cols to get *nnz, then with row_ptr [4*num_nodes+1] and
 * cols [nnz]. */
int goal_mesh_graph(int order, int64_t num_nodes, int64_t num_elems, const int32_t* conn, int64_t* nnz,
                    int64_t* row_ptr, int32_t* cols);
/* Host uniform refinement of a linear mesh (see goal_nested_create).  Call
 * with NULL outputs to get sizes[0..2] = fine nodes, fine elems, new vertices. */
int goal_mesh_refine(int64_t num_nodes, const double* coords, int64_t num_elems, const int32_t* conn,
                     int64_t* sizes, double* fine_coords, int32_t* fine_conn, int32_t* parent,
                     int32_t* edge_parents);

/* ------------------------------------------------------------ adaptation */
/* get_iso_target_size (src/goal_size_field.cpp:39-151) on a linear mesh:
 * per-element indicators eta [num_elems] -> per-vertex target size
 * vtx_size [num_nodes] for a target element count (p_order 1 as in the
 * reference's call). */
int goal_size_field(int64_t num_nodes, const double* coords, int64_t num_elems, const int32_t* conn,
                    const double* eta, int64_t target, int p_order, double* vtx_size);
/* The mesh step of Solver::adapt (src/main_nested.cpp:108-126, ma::adapt),
 * refinement only: edges longer than split_ratio x the mean target size of
 * their vertices are bisected (longest-edge propagation keeps the mesh
 * conforming) for up to max_passes passes.  Old vertices keep their ids; new
 * vertex num_old + k is the midpoint of new_vertex_parents[2k], [2k+1]. */
typedef struct goal_adapt_s* goal_adapt;
int goal_adapt_create(int64_t num_nodes, const double* coords, int64_t num_elems, const int32_t* conn,
                      const double* vtx_size, double split_ratio, int max_passes, goal_adapt* out);
int goal_adapt_destroy(goal_adapt a);
/* sizes[0..4] = nodes, elements, new vertices, edge cuts, passes run */
int goal_adapt_sizes(goal_adapt a, int64_t* sizes);
/* Any output may be NULL.  parent[e] = input element that contains e. */
int goal_adapt_get(goal_adapt a, double* coords, int32_t* conn, int32_t* parent, int32_t* new_vertex_parents);
/* A nodal field with ncomp values per vertex carried to the adapted mesh:
 * old vertices keep their values, new ones take the mean of their edge's. */
int goal_adapt_transfer(goal_adapt a, int ncomp, const double* old_field, double* new_field);

/* -------------------------------------------------------- discretization */
typedef struct goal_disc_s* goal_disc;

/* Disc(ParameterList) + Disc::build_data: mesh, node-major DOF numbering
 * lid = 4*node + eq and the CSR graph (src/goal_disc.cpp:111-120, 195-332).
 * order 1: conn is [num_elems][4]; order 2: [num_elems][10] (4 vertices, then
 * edges 01,12,20,03,13,23).  elem_set may be NULL (one element set). */
int goal_disc_create(int order, int64_t num_nodes, const double* coords, int64_t num_elems,
                     const int32_t* conn, const int32_t* elem_set, goal_disc* out);
/* Synthetic stand-in for loadMdsMesh + read_sets (src/goal_disc.cpp:91-99):
 * Kuhn box of nx*ny*nz hexahedra split into 6 tets each, interior vertices
 * jittered by +-jitter*h (seeded), node sets xmin..zmax on the closed faces. */
int goal_disc_create_box(int nx, int ny, int nz, double lx, double ly, double lz, int order, double jitter,
                         uint64_t seed, goal_disc* out);
/* goal's mesh inputs, read on the host: apf::loadMdsMesh(geom file, mesh
 * file) + read_sets(assoc file) (src/goal_disc.cpp:28-99), then the sets of
 * compute_elem_sets / compute_side_sets / compute_node_sets
 * (src/goal_disc.cpp:347-396).  mesh_file: a SCOREC MDS mesh of version 5
 * holding tetrahedra -- a one-part file (the reference's *-serial0.smb), or
 * "name.smb" for a mesh stored in parts name0.smb, name1.smb, ..., which are
 * merged through their shared vertices;
 * geom_file: its gmi .dmg model; assoc_file: "node set / side set / elem set"
 * lines with (model dim, model tag) entries.  Node sets hold the nodes
 * classified in the closure of their model entities; side sets the triangles
 * classified on their model faces, as (element, local face); element sets
 * are numbered in file order.  order 2 adds a node at each edge midpoint. */
typedef struct goal_mesh_file_s* goal_mesh_file;
int goal_mesh_file_read(const char* geom_file, const char* mesh_file, const char* assoc_file, int order,
                        goal_mesh_file* out);
int goal_mesh_file_destroy(goal_mesh_file f);
int goal_mesh_file_sizes(goal_mesh_file f, int64_t* num_nodes, int64_t* num_elems, int* num_elem_sets);
/* coords [num_nodes][3], conn [num_elems][4 or 10], elem_set [num_elems]; any may be NULL */
int goal_mesh_file_get_mesh(goal_mesh_file f, double* coords, int32_t* conn, int32_t* elem_set);
int goal_mesh_file_elem_set_name(goal_mesh_file f, int index, char* buf, int64_t len);
/* *n = the set's size; nodes / elems, faces (may be NULL) receive it */
int goal_mesh_file_node_set(goal_mesh_file f, const char* name, int64_t* n, int32_t* nodes);
int goal_mesh_file_side_set(goal_mesh_file f, const char* name, int64_t* n, int32_t* elems, int32_t* faces);
/* The Disc of a read mesh file, with its node, side and element sets. */
int goal_disc_create_from_file(goal_mesh_file f, goal_disc* out);
/* destroy_disc (src/goal_disc.cpp:452-454) */
int goal_disc_destroy(goal_disc d);
/* out[0..8] = num_nodes, num_elems, num_dofs, nnz, nodes_per_elem, num_vertices, order,
 *             num_elem_sets, max_node_degree */
int goal_disc_sizes(goal_disc d, int64_t* out);
int goal_disc_get_mesh(goal_disc d, double* coords, int32_t* conn, int32_t* elem_set);
/* Ghost graph after fillComplete: rows in lid order, sorted columns
 * (Disc::compute_graphs, src/goal_disc.cpp:307-332). row_ptr [num_dofs+1]. */
int goal_disc_get_graph(goal_disc d, int64_t* row_ptr, int32_t* cols);
/* Device pointers of the CSR graph (row_ptr [num_dofs+1] int64, cols [nnz] int32),
 * for device-resident consumers such as a GPU linear solver. */
int goal_disc_graph_dev(goal_disc d, int64_t** row_ptr_dev, int32_t** cols_dev);
/* Device pointers of the mesh: coords [num_nodes][3], conn [num_elems][npe] and
 * the element node-pair position table pos [num_elems][npe][npe] (uint16). */
int goal_disc_mesh_dev(goal_disc d, double** coords_dev, int32_t** conn_dev, uint16_t** pos_dev);
/* Disc::get_lids for every element, [num_elems][4*nodes_per_elem] (src/goal_disc.cpp:214-222) */
int goal_disc_get_lids(goal_disc d, int32_t* lids);
/* Node sets (Disc::get_nodes, src/goal_disc.cpp:179-182, compute_node_sets :378-396) */
int goal_disc_add_node_set(goal_disc d, const char* name, int64_t n, const int32_t* nodes);
int goal_disc_node_set_size(goal_disc d, const char* name, int64_t* n);
int goal_disc_get_node_set(goal_disc d, const char* name, int32_t* nodes);
/* Side sets (Disc::get_sides, src/goal_disc.cpp:174-177): element faces as
 * (element, local face) pairs, local face f of a tet being the vertices
 * {0,1,2}, {0,1,3}, {1,2,3}, {0,2,3} for f = 0..3 (apf's order).  A name with
 * no side set given but a node set of that name stands for the boundary faces
 * whose three vertices all lie in that node set (a box face's triangles for
 * the box meshes' face node sets), in ascending 4 * element + face order. */
int goal_disc_add_side_set(goal_disc d, const char* name, int64_t n, const int32_t* elems, const int32_t* faces);
int goal_disc_side_set_size(goal_disc d, const char* name, int64_t* n);
int goal_disc_get_side_set(goal_disc d, const char* name, int32_t* elems, int32_t* faces);
/* Element sets (Disc::get_elems / compute_elem_sets, src/goal_disc.cpp:169-172, 347-359) */
int goal_disc_set_elem_sets(goal_disc d, int num_sets, const int32_t* elem_set);
/* The u,p fields ("u", "p" created in src/goal_mechanics.cpp:71-85): x[4*node+eq]. */
int goal_disc_set_soln(goal_disc d, const double* x);
int goal_disc_get_soln(goal_disc d, double* x);
/* Disc::add_soln: u,p += du (src/goal_disc.cpp:398-422) */
int goal_disc_add_soln(goal_disc d, const double* du);
int goal_disc_add_soln_dev(goal_disc d, const double* du_dev);
double* goal_disc_soln_dev(goal_disc d);

/* ---------------------------------------------------------------- states */
typedef struct goal_states_s* goal_states;
/* Mechanics::make_states (src/goal_mechanics.cpp:87-95) with one history
 * point per quadrature point: sigma (all models), eqps and Fp with _old
 * copies (J2).  Initial values: sigma 0, eqps 0, Fp = I (src/goal_states.cpp:99-128). */
int goal_states_create(goal_disc d, int model, int quad_order, goal_states* out);
int goal_states_destroy(goal_states s);
/* States::update: X_old := X (src/goal_states.cpp:130-141) */
int goal_states_update(goal_states s);
/* name in {"sigma","eqps","Fp","eqps_old","Fp_old"}; layout [comp][num_elems*num_qp] */
int goal_states_get(goal_states s, const char* name, double* out);
int goal_states_set(goal_states s, const char* name, const double* in);

/* --------------------------------------------------------------- sol info */
typedef struct goal_solinfo_s* goal_solinfo;
/* SolInfo: R, dMdu (num_dofs) and dRdu values on the disc graph (src/goal_sol_info.cpp:6-22) */
int goal_solinfo_create(goal_disc d, goal_solinfo* out);
int goal_solinfo_destroy(goal_solinfo s);
/* zero_R / zero_dMdu / zero_dRdu / zero_all (src/goal_sol_info.cpp:51-70) */
int goal_solinfo_zero_R(goal_solinfo s);
int goal_solinfo_zero_dMdu(goal_solinfo s);
int goal_solinfo_zero_dRdu(goal_solinfo s);
int goal_solinfo_zero_all(goal_solinfo s);
/* download; any pointer may be NULL */
int goal_solinfo_get(goal_solinfo s, double* R, double* dMdu, double* dRdu_vals);
int goal_solinfo_set(goal_solinfo s, const double* R, const double* dMdu, const double* dRdu_vals);
/* device pointers of R, dMdu, dRdu values */
int goal_solinfo_dev(goal_solinfo s, double** R_dev, double** dMdu_dev, double** vals_dev);
/* R->norm2() (src/goal_primal.cpp:130) */
int goal_solinfo_norm_R(goal_solinfo s, double* norm);
/* -(R . z) (src/goal_nested_adjoint.cpp:213) with z a device vector */
int goal_solinfo_dot_R(goal_solinfo s, const double* z_dev, double* dot);
/* Linear solve dRdu x = b on the SolInfo's CSR values, standing in for
 * goal::solve (src/goal_linear_solve.cpp:22-90: Belos block GMRES, block
 * size 1, "krylov size" = restart, DGKS orthogonalization, left MueLu
 * preconditioner with "number of equations" 4).  Here: GMRES(restart)
 * preconditioned on the right by one V-cycle of aggregation multigrid over
 * the 4-equation node blocks (aggregates of nodes, Galerkin coarse matrices,
 * Chebyshev smoothing on the inverse diagonal blocks, a dense inverse on the
 * coarsest level), classical Gram-Schmidt with the DGKS test (a second pass
 * when the first removes more than 1 - 1/sqrt(2) of the norm).  b, x are
 * device vectors of num_dofs; x holds the initial guess on entry.  Stops when
 * ||b - A x|| <= rtol ||b - A x0|| or after max_iters inner iterations; like
 * the reference it returns normally when the limit is hit (*relres tells).
 * Outputs may be NULL. */
int goal_linear_solve(goal_solinfo s, const double* b_dev, double* x_dev, double rtol, int max_iters, int restart,
                      int* iters, double* relres);
/* The same with the preconditioner named: GOAL_PRECOND_AMG (goal_linear_solve's)
 * or GOAL_PRECOND_BLOCK_JACOBI (the inverse 4x4 diagonal block of every node). */
enum { GOAL_PRECOND_BLOCK_JACOBI = 0, GOAL_PRECOND_AMG = 1 };
int goal_linear_solve_pc(goal_solinfo s, const double* b_dev, double* x_dev, double rtol, int max_iters,
                         int restart, int precond, int* iters, double* relres);

/* ------------------------------------------------------------- mechanics */
typedef struct goal_material {
  int model;            /* GOAL_MODEL_*                                   */
  double E, nu, K, Y;   /* K, Y: J2 hardening modulus and yield stress    */
  double c0;            /* stabilization constant                          */
} goal_material;

typedef struct goal_mech_s* goal_mech;
/* Mechanics(ParameterList, Disc*) (src/goal_mechanics.cpp:41-63): one
 * material per element set (materials sublist), stabilization on/off,
 * quadrature order (1: the reference's one point, 2: four points).  All
 * element sets must use the same model. */
int goal_mech_create(goal_disc d, goal_states st, int num_sets, const goal_material* mats, int stabilization,
                     int quad_order, goal_mech* out);
int goal_mech_destroy(goal_mech m);
/* Options: "blocks" (workgroups per launch, 0 = auto), "xcd_remap" (0/1),
 * "check" (1: synchronise after each assembly and report J2 failures),
 * "stabilization" (0/1, overrides the value given at creation),
 * "tangent" (64: FP64 element matrices, the default; 32: element matrices
 * formed and stored in FP32 and summed into the FP64 CSR -- the residual stays
 * FP64, so Newton still converges to the FP64 tolerance, in more iterations
 * at worst),
 * "timing" (1: record device events around the element and gather kernels of
 * each assembly, read back with goal_mech_timing),
 * "overlap" (neo-Hookean only: 2..1024 runs the element pass in this many
 * chunks and gathers each finished chunk's node ranges on a second stream,
 * bitwise equal to one pass; 0 off; -1, the default, is one pass; not used
 * under stream capture),
 * "overlap_blocks" (element workgroups per chunk, 0 = 5/6 of the resident
 * grid). */
int goal_mech_set_option(goal_mech m, const char* key, int64_t value);
/* Device time of the last assembly's element pass and node gather, in ms
 * (waits for it).  Needs the "timing" option.  No counterpart in the
 * reference (goal::time() brackets whole phases on the host,
 * src/goal_primal.cpp:76-108); it is for profiling callers. */
int goal_mech_timing(goal_mech m, float* element_ms, float* gather_ms);
/* Accumulated kernel status bits since the last call (1: J2 return map failed). */
int goal_mech_status(goal_mech m, int* flags);

/* goal::assemble over the chain of Mechanics::build_resid<ST> with PRIMAL
 * scatter: R += r_e (src/goal_assembly.cpp:65-88, src/goal_primal.cpp:75-90).
 * save_states: the `save` flag of build_resid (src/goal_primal.cpp:49-50). */
int goal_asm_residual(goal_mech m, goal_solinfo s, int save_states);
/* goal::assemble over build_resid<FADT>: R += r_e and dRdu += k_e
 * (Displacement/Pressure<FADT>::scatter_primal, src/goal_displacement.cpp:177-194,
 * src/goal_pressure.cpp:170-185); adjoint != 0 scatters the transpose
 * (scatter_adjoint, src/goal_displacement.cpp:196-214, src/goal_pressure.cpp:187-203). */
int goal_asm_jacobian(goal_mech m, goal_solinfo s, int adjoint, int save_states);
/* goal_asm_jacobian without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The element
 * residuals and matrices between the two passes live in a 64 MiB array that the
 * library holds on each GPU until the process ends, so a mesh may have at most
 * 5,115 ten-node or 30,840 four-node tets (J2, whose point states take room
 * too: 3,748 ten-node four-point or 19,904 four-node one-point tets; larger
 * meshes: GOAL_ERR_INVALID, and nothing is launched).  Results are bitwise those of goal_asm_jacobian.  All
 * calls on one GPU share the array: calls that overlap on one GPU (other
 * streams or threads) race.  Under capture, set the mechanics' "check" option
 * to 0 (the status read-back synchronizes). */
int goal_asm_jacobian_noalloc(goal_mech m, goal_solinfo s, int adjoint, int save_states);
/* Functional::compute / QoI<ST>, QoI<FADT> (src/goal_functional.cpp:62-70,
 * src/goal_qoi.cpp:20-81, src/goal_avg_disp.cpp:16-21, src/goal_avg_vm.cpp:32-61).
 * kind GOAL_QOI_*; elem_set: AvgVM element set (-1 = all); value: J(u);
 * elem_values (host, may be NULL): per-element contributions; if s != NULL
 * the derivative dJ/du is added into s->dMdu. */
int goal_asm_qoi(goal_mech m, goal_solinfo s, int kind, int elem_set, double* value, double* elem_values);
/* NestedAdjoint::localize's assembly (src/goal_nested_adjoint.cpp:217-234):
 * residual of build_error (src/goal_mechanics.cpp:169-218) with weights
 * z_diff*N_a (uw, pw) and z_coarse*N_a (pwc), R += r_e. */
int goal_asm_localize(goal_mech m, goal_solinfo s, const double* z_diff_dev, const double* z_coarse_dev);

/* -------------------------------------------------- Dirichlet and traction */
/* goal::eval (src/goal_control.cpp:107-123) */
int goal_eval(const char* expr, double x, double y, double z, double t, double* out);
/* One entry [comp, node set, expr] of set_resid_dbcs (jacobian = 0) or
 * set_jac_dbcs (jacobian = 1) (src/goal_dbcs.cpp:39-99). */
int goal_dbc_apply(goal_disc d, goal_solinfo s, int comp, const char* node_set, const char* expr, double t,
                   int jacobian);
/* Same with explicit node ids and prescribed values. */
int goal_dbc_apply_values(goal_disc d, goal_solinfo s, int comp, int64_t n, const int32_t* nodes,
                          const double* values, int jacobian);
/* One entry [side set, T_x, T_y, T_z] of set_tbcs (src/goal_tbcs.cpp:30-66):
 * a dead traction load.  Per face, the reference's one-point rule (centroid,
 * w dv = face area): R[a, i] -= T_i(x_c, t) W_{a,i} |face| for every face node
 * a, with the "uw" weight W_{a,i} = N_a(centroid) when weight_dev is NULL
 * (primal and adjoint chains) or z_i(x_c) N_a(centroid) for a device DOF
 * vector z = weight_dev (the error chain's DisplacementAdjoint,
 * src/goal_displacement_adjoint.cpp:37-53).  Residual only; call it after the
 * assembly and before the Dirichlet rows, as Primal::compute_resid /
 * compute_jacob and NestedAdjoint::compute_adjoint / localize do
 * (src/goal_primal.cpp:75-109, src/goal_nested_adjoint.cpp:163-234). */
int goal_tbc_apply(goal_disc d, goal_solinfo s, const char* side_set, const char* tx, const char* ty,
                   const char* tz, double t, const double* weight_dev);

/* --------------------------------------------------------- batched assembly */
/* Many meshes assembled as one launch: the Jacobian-mode assemblies of n
 * (mechanics, sol info) pairs -- each what goal_asm_jacobian(m, s, 0, 1) does
 * after goal_solinfo_zero_all(s) -- captured into one HIP graph.  No
 * counterpart in the reference (one Mechanics per process there); it is for
 * callers that assemble many small meshes. */
typedef struct goal_batch_s* goal_batch;
int goal_batch_create(int n, const goal_mech* mechs, const goal_solinfo* infos, goal_batch* out);
int goal_batch_destroy(goal_batch b);
/* Autotuner: for each mesh, time `reps` assemblies with each candidate
 * persistent-grid size (the occupancy-sized grid and multiples of the CU
 * count) and keep the fastest as its "blocks" option.  Outputs (may be NULL):
 * the chosen grid (0 = occupancy-sized) and its median time in ms per mesh.
 * A captured graph is dropped (capture again). */
int goal_batch_tune(goal_batch b, int reps, int* blocks, float* ms);
/* Capture the batch (one eager assembly per mesh first sizes the buffers).
 * Every mesh's "timing" option must be off. */
int goal_batch_capture(goal_batch b);
/* Launch the captured graph on the library stream (async). */
int goal_batch_run(goal_batch b);

/* ----------------------------------------------------- learned indicator */
/* No counterpart in the reference: a 12 -> 64 -> 64 -> 1 ReLU network that
 * predicts log10(eta_E / E) for the nested-adjoint indicator eta_E of
 * Nested::set_error (src/goal_nested.cpp:395-412) from 12 per-element features
 * of the primal state, standing in for the nested adjoint solve.  Weights are
 * BF16, biases FP32, all accumulation FP32; training keeps an FP32 master
 * copy and runs Adam on the device. */
#define GOAL_MLP_IN 12
#define GOAL_MLP_HIDDEN 64
/* W1[64][12], b1[64], W2[64][64], b2[64], W3[64], b3 */
#define GOAL_MLP_PARAMS 5057
/* input shift[12], input scale[12] (x_hat = (x - shift) * scale), output
 * shift and scale, training feature box lo[12], hi[12], then c: the
 * prediction is y = net(x_hat) * scale + shift + c * x[0] */
#define GOAL_MLP_NORMS 51
/* Features [num_elems][12] FP32 on the device, for the current solution and
 * states of m's disc: log10 h, centroid x y z, log10(|x_c - qoi_x| + h),
 * [x_c < qoi_x], log10(vm/E), log10(|vm - vm*|/E), log10|grad u|,
 * log10|grad u - (grad u)*|, nu, log10(eqps + 1e-9).  vm is the element mean
 * von Mises stress (src/goal_von_mises.cpp:7-17); a starred value is the mean
 * over the element's vertices of the volume-weighted nodal average. */
int goal_learn_features(goal_mech m, double qoi_x, float* feat_dev);
typedef struct goal_mlp_s* goal_mlp;
/* Zero parameters, identity normalization, unbounded feature box. */
int goal_mlp_create(goal_mlp* out);
int goal_mlp_destroy(goal_mlp m);
/* params [GOAL_MLP_PARAMS] FP32 (weights are rounded to BF16 for use; Adam
 * state restarts), norms [GOAL_MLP_NORMS]; either may be NULL. */
int goal_mlp_set(goal_mlp m, const float* params, const float* norms);
/* FP32 master parameters, the BF16 weights in use, the normalization; any may be NULL. */
int goal_mlp_get(goal_mlp m, float* params, uint16_t* weights_bf16, float* norms);
/* y_dev[i] = prediction for row i of x_dev [n][12]; *num_outside (may be
 * NULL) = rows with a feature outside the training box. */
int goal_mlp_forward(goal_mlp m, const float* x_dev, int64_t n, float* y_dev, int64_t* num_outside);
/* Mean squared error over rows idx_dev[0..n) of (x_dev, y_dev) in the
 * normalized output space and its gradient (host [GOAL_MLP_PARAMS]) with
 * respect to the parameters as used (BF16 weights). */
int goal_mlp_gradient(goal_mlp m, const float* x_dev, const float* y_dev, const int32_t* idx_dev, int64_t n,
                      float* loss, float* grad);
/* steps Adam steps (beta 0.9, 0.999, eps 1e-8); step k uses rows
 * idx_dev[k*batch .. (k+1)*batch).  *last_loss = the last step's loss. */
int goal_mlp_train(goal_mlp m, const float* x_dev, const float* y_dev, const int32_t* idx_dev, int64_t batch,
                   int steps, float lr, float* last_loss);

/* ------------------------------------------------------------ nested DWR */
typedef struct goal_nested_s* goal_nested;
/* Nested(Disc*, UNIFORM) (src/goal_nested.cpp:13-28, 144-154): uniform
 * 1:8 refinement of a linear base mesh; the refined disc carries u,p
 * transferred from the base (new vertex = mean of its edge's vertices). */
int goal_nested_create(goal_disc base, goal_nested* out);
int goal_nested_destroy(goal_nested n);
/* The refined discretization, owned by n. */
goal_disc goal_nested_disc(goal_nested n);
/* parent [num_fine_elems]; edge_parents [num_new_vertices][2] (may be NULL) */
int goal_nested_get_maps(goal_nested n, int32_t* parent, int32_t* edge_parents, int64_t* num_new_vertices);
/* Copy base u,p to the old vertices of the nested solution and interpolate new ones. */
int goal_nested_transfer_soln(goal_nested n);
/* State transfer onto the nested mesh: each child element takes its parent's
 * current history (sigma, Fp, eqps), which become the nested "_old" fields as
 * in the reference (fields copied to the refined mesh, then States::update in
 * make_states, src/goal_mechanics.cpp:87-95, src/goal_states.cpp:130-141). */
int goal_nested_transfer_states(goal_nested n, goal_states base, goal_states fine);
/* Nested::set_coarse on a nested nodal field, in place (src/goal_nested.cpp:369-393) */
int goal_dwr_set_coarse(goal_nested n, double* field_dev);
/* NestedAdjoint::subtract (src/goal_nested_adjoint.cpp:182-197): out = fine - coarse */
int goal_dwr_subtract(const double* fine_dev, const double* coarse_dev, double* out_dev, int64_t n);
/* compute_error (src/goal_error.cpp:7-35): eta[e] = |sum_d e_u,d + e_p| at the centroid */
int goal_dwr_compute_error(goal_disc d, const double* e_dev, double* eta_dev);
/* sum_contribs (src/goal_error.cpp:37-56) */
int goal_dwr_sum_contribs(goal_disc d, const double* e_dev, double* bound);
/* Nested::set_error (src/goal_nested.cpp:395-412): eta_base[E] = sum of its children */
int goal_dwr_set_error(goal_nested n, const double* eta_child_dev, double* eta_base_dev);

#ifdef __cplusplus
}
#endif

#endif /* GOAL_AMD_H */
