// Host-side discretization: tetrahedral meshes (P1 / P2), node and element
// sets, node-major interleaved DOF numbering and the CSR sparsity graph.
//
// Reference behaviour followed here:
//   * DOF numbering lid = node * (dim + 1) + eq   (src/goal_disc.cpp:195-201)
//   * element-local DOF order n * (dim + 1) + eq  (src/goal_disc.cpp:214-222)
//   * graph = union of element cliques, rows in lid order, columns sorted
//     ascending (Tpetra fillComplete on one rank; src/goal_disc.cpp:307-332)
//   * P2 tet node order: 4 vertices, then edges {01,12,20,03,13,23}
//     (apf Lagrange quadratic tet, the shape the reference fields use via
//     apf::createFieldOn, src/goal_mechanics.cpp:71-85).
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace goal {

constexpr int kDim = 3;
constexpr int kNumEqs = kDim + 1;  // u_x, u_y, u_z, p on every node

// The six tet edges as (vertex, vertex) pairs in apf order.
constexpr int kTetEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

struct HostMesh {
  int order = 1;                 // 1: 4-node tets, 2: 10-node tets
  int nodes_per_elem = 4;
  int64_t num_nodes = 0;
  int64_t num_elems = 0;
  std::vector<double> coords;    // [num_nodes][3]
  std::vector<int32_t> conn;     // [num_elems][nodes_per_elem]
  std::vector<int32_t> elem_set; // [num_elems] element-set index (0 default)
  int num_elem_sets = 1;
  std::map<std::string, std::vector<int32_t>> node_sets;  // node ids, ascending
  // Vertex nodes: node id of each mesh vertex (identity for P1 meshes).
  std::vector<int32_t> vertex_nodes;
};

// CSR graph of the DOF-level Jacobian plus the node-level adjacency it is
// built from.  node_adj_ptr / node_adj give, for every node, the sorted list of
// nodes it shares an element with (itself included).  Because every node
// carries kNumEqs equations, the DOF-level row of (node a, eq i) is
//   { kNumEqs * b + j : b in adj(a), j = 0..kNumEqs-1 }  in that order,
// which is sorted ascending because adj(a) is.
struct HostGraph {
  std::vector<int64_t> node_adj_ptr;  // [num_nodes + 1]
  std::vector<int32_t> node_adj;      // [node_adj_ptr.back()]
  int max_degree = 0;
  int64_t num_dofs = 0;
  int64_t nnz = 0;
  // Per element, per ordered local node pair (a, b): position of node
  // conn[e][b] inside adj(conn[e][a]).  [num_elems][npe][npe]
  std::vector<uint16_t> pos;
  // Node -> incident elements, packed as e * npe + local node index, in
  // ascending element order.
  std::vector<int64_t> inc_ptr;  // [num_nodes + 1]
  std::vector<int32_t> inc;
  // Incidence record of (element e, local node a): inc[rec[e * npe + a]] ==
  // e * npe + a.  [num_elems][npe]
  std::vector<int32_t> rec;
  // pos rows in incidence order: posi[k][b] = pos[inc[k]][b].  [inc.size()][npe]
  std::vector<uint16_t> posi;
};

// Kuhn (Freudenthal) box mesh: nx*ny*nz hexahedra on [0,lx]x[0,ly]x[0,lz],
// each split into 6 positively oriented tets sharing the (0,0,0)-(1,1,1)
// diagonal.  Interior vertex coordinates are jittered by U(-jitter, jitter)*h
// per coordinate with a counter-based generator keyed by `seed`; vertices on
// the box boundary move only within their face (edge, corner) so the faces
// stay planar.  P2 edge nodes sit at the straight midpoints of their edges.
// Node sets "xmin", "xmax", "ymin", "ymax", "zmin", "zmax" hold every node on
// the closed face.
HostMesh make_box_mesh(int nx, int ny, int nz, double lx, double ly, double lz,
                       int order, double jitter, uint64_t seed);

// Validate a user-provided mesh and derive vertex_nodes.  Throws on error.
void finalize_mesh(HostMesh& m);

// Build the node adjacency, DOF counts and the per-element position table.
HostGraph build_graph(const HostMesh& m, int num_threads);

// Expand the node adjacency into the DOF-level CSR (row_ptr, cols) on the host.
void expand_csr(const HostMesh& m, const HostGraph& g, std::vector<int64_t>& row_ptr,
                std::vector<int32_t>& cols);

// Uniform refinement of a linear tet mesh: every tet is split into 8 children
// (4 corner tets + the inner octahedron cut along the diagonal joining the
// midpoints of edges 02 and 13).  Old vertices keep their ids, one new vertex
// per edge follows.  parent[child] = base element, children of base element E
// are 8E..8E+7.  edge_parents[k] = the two old vertices of new vertex
// (num_old_vertices + k).
struct Refinement {
  HostMesh fine;
  std::vector<int32_t> parent;        // [fine.num_elems]
  std::vector<int32_t> edge_parents;  // [num_new_vertices][2]
  int64_t num_old_vertices = 0;
};
Refinement refine_uniform(const HostMesh& base);

// Isotropic target size per vertex from per-element error indicators
// (get_iso_target_size, src/goal_size_field.cpp:39-151).  Linear meshes.
std::vector<double> iso_target_size(const HostMesh& m, const double* eta, int64_t target, int p_order);

// Refinement of a linear mesh toward a vertex size field by longest-edge
// bisection.  Old vertices keep their ids; new vertex num_old_vertices + k is
// the midpoint of new_vertex_parents[2k], [2k+1] (either may itself be a new
// vertex with a smaller k).  parent[e] = element of the input mesh that
// contains element e.
struct Adaptation {
  HostMesh mesh;
  std::vector<int32_t> parent;
  std::vector<int32_t> new_vertex_parents;
  int64_t num_old_vertices = 0;
  int64_t splits = 0;
  int passes = 0;
};
Adaptation adapt_refine(const HostMesh& m, const std::vector<double>& vtx_size, double split_ratio,
                        int max_passes);

// splitmix64 of (seed, key): the jitter source, reproducible on any host.
uint64_t splitmix64(uint64_t x);

}  // namespace goal
