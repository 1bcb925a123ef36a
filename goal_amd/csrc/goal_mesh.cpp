#include "goal_mesh.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace goal {

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

static double unit_uniform(uint64_t seed, uint64_t key) {
  uint64_t r = splitmix64(splitmix64(seed) ^ key);
  return (double)(r >> 11) * (1.0 / 9007199254740992.0);  // [0,1)
}

template <typename F>
static void parallel_for(int64_t n, int num_threads, F&& body) {
  if (num_threads <= 1 || n < 4096) {
    body(0, (int64_t)0, n);
    return;
  }
  std::vector<std::thread> pool;
  int64_t chunk = (n + num_threads - 1) / num_threads;
  for (int t = 0; t < num_threads; ++t) {
    int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([&, t, b, e] { body(t, b, e); });
  }
  for (auto& th : pool) th.join();
}

HostMesh make_box_mesh(int nx, int ny, int nz, double lx, double ly, double lz,
                       int order, double jitter, uint64_t seed) {
  if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("box: cell counts must be >= 1");
  if (order != 1 && order != 2) throw std::invalid_argument("box: order must be 1 or 2");
  if (!(jitter >= 0.0 && jitter < 0.2)) throw std::invalid_argument("box: jitter must be in [0, 0.2)");
  HostMesh m;
  m.order = order;
  m.nodes_per_elem = order == 1 ? 4 : 10;
  const double h[3] = {lx / nx, ly / ny, lz / nz};
  const int n[3] = {nx, ny, nz};
  const int64_t vx = nx + 1, vy = ny + 1, vz = nz + 1;
  const int64_t nverts = vx * vy * vz;

  // Vertex coordinates on the (n+1)^3 lattice.
  std::vector<double> vcoord(nverts * 3);
  for (int64_t k = 0; k < vz; ++k)
    for (int64_t j = 0; j < vy; ++j)
      for (int64_t i = 0; i < vx; ++i) {
        int64_t v = i + vx * (j + vy * k);
        const int64_t idx[3] = {i, j, k};
        for (int c = 0; c < 3; ++c) {
          double x = idx[c] * h[c];
          bool on_bnd = idx[c] == 0 || idx[c] == n[c];
          if (jitter > 0.0 && !on_bnd)
            x += (2.0 * unit_uniform(seed, (uint64_t)(3 * v + c)) - 1.0) * jitter * h[c];
          if (idx[c] == n[c]) x = (c == 0 ? lx : c == 1 ? ly : lz);
          vcoord[3 * v + c] = x;
        }
      }

  // Node lattice: the vertex lattice for P1, the doubled lattice for P2
  // (every doubled-lattice point is a vertex or the midpoint of a Kuhn edge).
  const int s = order;  // lattice stride of a vertex
  const int64_t Lx = (int64_t)s * nx + 1, Ly = (int64_t)s * ny + 1, Lz = (int64_t)s * nz + 1;
  m.num_nodes = Lx * Ly * Lz;
  m.coords.resize(m.num_nodes * 3);
  auto lattice_node = [&](int64_t I, int64_t J, int64_t K) { return I + Lx * (J + Ly * K); };
  for (int64_t K = 0; K < Lz; ++K)
    for (int64_t J = 0; J < Ly; ++J)
      for (int64_t I = 0; I < Lx; ++I) {
        int64_t node = lattice_node(I, J, K);
        if (s == 1) {
          for (int c = 0; c < 3; ++c) m.coords[3 * node + c] = vcoord[3 * node + c];
          continue;
        }
        const int64_t P[3] = {I, J, K};
        int64_t A[3], B[3];
        for (int c = 0; c < 3; ++c) {
          int64_t d = P[c] & 1;
          A[c] = (P[c] - d) / 2;
          B[c] = (P[c] + d) / 2;
        }
        int64_t va = A[0] + vx * (A[1] + vy * A[2]);
        int64_t vb = B[0] + vx * (B[1] + vy * B[2]);
        for (int c = 0; c < 3; ++c)
          m.coords[3 * node + c] = 0.5 * (vcoord[3 * va + c] + vcoord[3 * vb + c]);
      }

  // Six Kuhn tets per hexahedron, hexahedra in lexicographic (x fastest) order.
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  static const int perm_sign[6] = {+1, -1, -1, +1, +1, -1};
  m.num_elems = (int64_t)nx * ny * nz * 6;
  m.conn.resize(m.num_elems * m.nodes_per_elem);
  m.elem_set.assign(m.num_elems, 0);
  int64_t e = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        for (int t = 0; t < 6; ++t, ++e) {
          int64_t V[4][3];
          int64_t cur[3] = {i, j, k};
          for (int c = 0; c < 3; ++c) V[0][c] = cur[c];
          for (int st = 0; st < 3; ++st) {
            cur[perms[t][st]] += 1;
            for (int c = 0; c < 3; ++c) V[st + 1][c] = cur[c];
          }
          if (perm_sign[t] < 0)
            for (int c = 0; c < 3; ++c) std::swap(V[2][c], V[3][c]);
          int32_t* ec = &m.conn[e * m.nodes_per_elem];
          for (int a = 0; a < 4; ++a) ec[a] = (int32_t)lattice_node(s * V[a][0], s * V[a][1], s * V[a][2]);
          if (order == 2)
            for (int ed = 0; ed < 6; ++ed) {
              const int64_t* p = V[kTetEdge[ed][0]];
              const int64_t* q = V[kTetEdge[ed][1]];
              ec[4 + ed] = (int32_t)lattice_node(p[0] + q[0], p[1] + q[1], p[2] + q[2]);
            }
        }

  // Face node sets (closure of each face), ascending node id.
  const char* names[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
  const int64_t Lm[3] = {Lx - 1, Ly - 1, Lz - 1};
  for (int f = 0; f < 6; ++f) {
    int c = f / 2;
    int64_t target = (f % 2 == 0) ? 0 : Lm[c];
    std::vector<int32_t> set;
    for (int64_t K = 0; K < Lz; ++K)
      for (int64_t J = 0; J < Ly; ++J)
        for (int64_t I = 0; I < Lx; ++I) {
          const int64_t P[3] = {I, J, K};
          if (P[c] == target) set.push_back((int32_t)lattice_node(I, J, K));
        }
    m.node_sets[names[f]] = std::move(set);
  }
  finalize_mesh(m);
  return m;
}

void finalize_mesh(HostMesh& m) {
  if (m.order != 1 && m.order != 2) throw std::invalid_argument("mesh: order must be 1 or 2");
  m.nodes_per_elem = m.order == 1 ? 4 : 10;
  if ((int64_t)m.conn.size() != m.num_elems * m.nodes_per_elem)
    throw std::invalid_argument("mesh: connectivity size does not match num_elems * nodes_per_elem");
  if ((int64_t)m.coords.size() != m.num_nodes * 3)
    throw std::invalid_argument("mesh: coordinate array size does not match num_nodes * 3");
  if (m.num_nodes > (int64_t)INT32_MAX / kNumEqs)
    throw std::invalid_argument("mesh: too many nodes for 32-bit DOF ids");
  for (int32_t c : m.conn)
    if (c < 0 || c >= m.num_nodes) throw std::invalid_argument("mesh: connectivity entry out of range");
  if (m.elem_set.empty()) m.elem_set.assign(m.num_elems, 0);
  if ((int64_t)m.elem_set.size() != m.num_elems) throw std::invalid_argument("mesh: elem_set size mismatch");
  int mx = 0;
  for (int32_t s : m.elem_set) {
    if (s < 0) throw std::invalid_argument("mesh: negative element set id");
    mx = std::max(mx, (int)s);
  }
  m.num_elem_sets = std::max(m.num_elem_sets, mx + 1);
  std::vector<char> is_vertex(m.num_nodes, 0);
  for (int64_t e = 0; e < m.num_elems; ++e)
    for (int a = 0; a < 4; ++a) is_vertex[m.conn[e * m.nodes_per_elem + a]] = 1;
  m.vertex_nodes.clear();
  for (int64_t n = 0; n < m.num_nodes; ++n)
    if (is_vertex[n]) m.vertex_nodes.push_back((int32_t)n);
  for (auto& kv : m.node_sets)
    for (int32_t n : kv.second)
      if (n < 0 || n >= m.num_nodes) throw std::invalid_argument("mesh: node set entry out of range");
}

HostGraph build_graph(const HostMesh& m, int num_threads) {
  HostGraph g;
  const int npe = m.nodes_per_elem;
  const int64_t nn = m.num_nodes, ne = m.num_elems;
  if (num_threads < 1) num_threads = 1;

  // node -> element incidence
  std::vector<int64_t> inc_ptr(nn + 1, 0);
  for (int64_t i = 0; i < ne * npe; ++i) inc_ptr[m.conn[i] + 1]++;
  for (int64_t n = 0; n < nn; ++n) inc_ptr[n + 1] += inc_ptr[n];
  std::vector<int32_t> inc(inc_ptr[nn]);
  {
    std::vector<int64_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < npe; ++a) inc[fill[m.conn[e * npe + a]]++] = (int32_t)e;
  }

  // Adjacent node lists: two passes (degrees, then fill), each thread with a
  // private marker array.
  std::vector<int32_t> degree(nn, 0);
  auto collect = [&](std::vector<int32_t>& mark, std::vector<int32_t>& buf, int64_t a) {
    buf.clear();
    for (int64_t k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) {
      const int32_t* ec = &m.conn[(int64_t)inc[k] * npe];
      for (int b = 0; b < npe; ++b)
        if (mark[ec[b]] != (int32_t)a) {
          mark[ec[b]] = (int32_t)a;
          buf.push_back(ec[b]);
        }
    }
    std::sort(buf.begin(), buf.end());
  };
  parallel_for(nn, num_threads, [&](int, int64_t b, int64_t e) {
    std::vector<int32_t> mark(nn, -1), buf;
    for (int64_t a = b; a < e; ++a) {
      collect(mark, buf, a);
      degree[a] = (int32_t)buf.size();
    }
  });
  g.node_adj_ptr.assign(nn + 1, 0);
  for (int64_t a = 0; a < nn; ++a) {
    g.node_adj_ptr[a + 1] = g.node_adj_ptr[a] + degree[a];
    g.max_degree = std::max(g.max_degree, (int)degree[a]);
  }
  g.node_adj.resize(g.node_adj_ptr[nn]);
  parallel_for(nn, num_threads, [&](int, int64_t b, int64_t e) {
    std::vector<int32_t> mark(nn, -1), buf;
    for (int64_t a = b; a < e; ++a) {
      collect(mark, buf, a);
      std::copy(buf.begin(), buf.end(), g.node_adj.begin() + g.node_adj_ptr[a]);
    }
  });
  if (g.max_degree > 65535) throw std::runtime_error("graph: node degree exceeds 65535");
  if (ne * npe > (int64_t)INT32_MAX) throw std::runtime_error("graph: too many element-node incidences");
  g.inc_ptr = inc_ptr;
  g.inc.resize(inc.size());
  {
    std::vector<int64_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int a = 0; a < npe; ++a) g.inc[fill[m.conn[e * npe + a]]++] = (int32_t)(e * npe + a);
  }
  g.num_dofs = nn * kNumEqs;
  g.nnz = g.node_adj_ptr[nn] * kNumEqs * kNumEqs;

  // position table
  g.pos.resize(ne * npe * npe);
  parallel_for(ne, num_threads, [&](int, int64_t b, int64_t e) {
    for (int64_t el = b; el < e; ++el) {
      const int32_t* ec = &m.conn[el * npe];
      for (int a = 0; a < npe; ++a) {
        const int32_t* lo = &g.node_adj[g.node_adj_ptr[ec[a]]];
        const int32_t* hi = &g.node_adj[g.node_adj_ptr[ec[a] + 1]];
        for (int bb = 0; bb < npe; ++bb) {
          const int32_t* it = std::lower_bound(lo, hi, ec[bb]);
          g.pos[(el * npe + a) * npe + bb] = (uint16_t)(it - lo);
        }
      }
    }
  });
  const int64_t ninc = (int64_t)g.inc.size();
  g.rec.resize(ninc);
  for (int64_t k = 0; k < ninc; ++k) g.rec[g.inc[k]] = (int32_t)k;
  g.posi.resize(ninc * npe);
  parallel_for(ninc, num_threads, [&](int, int64_t b, int64_t e) {
    for (int64_t k = b; k < e; ++k)
      std::copy_n(&g.pos[(int64_t)g.inc[k] * npe], npe, &g.posi[k * npe]);
  });
  return g;
}

void expand_csr(const HostMesh& m, const HostGraph& g, std::vector<int64_t>& row_ptr,
                std::vector<int32_t>& cols) {
  const int64_t nn = m.num_nodes;
  row_ptr.resize(nn * kNumEqs + 1);
  cols.resize(g.nnz);
  for (int64_t a = 0; a < nn; ++a) {
    int64_t deg = g.node_adj_ptr[a + 1] - g.node_adj_ptr[a];
    int64_t base = g.node_adj_ptr[a] * kNumEqs * kNumEqs;
    for (int i = 0; i < kNumEqs; ++i) {
      int64_t r0 = base + i * kNumEqs * deg;
      row_ptr[a * kNumEqs + i] = r0;
      for (int64_t k = 0; k < deg; ++k)
        for (int j = 0; j < kNumEqs; ++j)
          cols[r0 + k * kNumEqs + j] = g.node_adj[g.node_adj_ptr[a] + k] * kNumEqs + j;
    }
  }
  row_ptr[nn * kNumEqs] = g.nnz;
}

static double tet_det(const double* x0, const double* x1, const double* x2, const double* x3) {
  double a[3], b[3], c[3];
  for (int i = 0; i < 3; ++i) {
    a[i] = x1[i] - x0[i];
    b[i] = x2[i] - x0[i];
    c[i] = x3[i] - x0[i];
  }
  return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
         a[2] * (b[0] * c[1] - b[1] * c[0]);
}

Refinement refine_uniform(const HostMesh& base) {
  if (base.order != 1) throw std::invalid_argument("refine: only linear meshes can be refined");
  Refinement r;
  const int64_t ne = base.num_elems, nv = base.num_nodes;
  r.num_old_vertices = nv;
  // unique edges
  struct EdgeRef {
    int32_t lo, hi;
    int64_t slot;  // e*6 + local edge
  };
  std::vector<EdgeRef> refs(ne * 6);
  for (int64_t e = 0; e < ne; ++e)
    for (int ed = 0; ed < 6; ++ed) {
      int32_t a = base.conn[e * 4 + kTetEdge[ed][0]], b = base.conn[e * 4 + kTetEdge[ed][1]];
      refs[e * 6 + ed] = {std::min(a, b), std::max(a, b), e * 6 + ed};
    }
  std::sort(refs.begin(), refs.end(), [](const EdgeRef& x, const EdgeRef& y) {
    return x.lo != y.lo ? x.lo < y.lo : x.hi < y.hi;
  });
  std::vector<int32_t> edge_vertex(ne * 6);
  int64_t nedges = 0;
  for (size_t k = 0; k < refs.size(); ++k) {
    if (k == 0 || refs[k].lo != refs[k - 1].lo || refs[k].hi != refs[k - 1].hi) {
      r.edge_parents.push_back(refs[k].lo);
      r.edge_parents.push_back(refs[k].hi);
      ++nedges;
    }
    edge_vertex[refs[k].slot] = (int32_t)(nv + nedges - 1);
  }
  HostMesh& f = r.fine;
  f.order = 1;
  f.nodes_per_elem = 4;
  f.num_nodes = nv + nedges;
  f.coords.resize(f.num_nodes * 3);
  std::copy(base.coords.begin(), base.coords.end(), f.coords.begin());
  for (int64_t k = 0; k < nedges; ++k)
    for (int c = 0; c < 3; ++c)
      f.coords[3 * (nv + k) + c] =
          0.5 * (base.coords[3 * r.edge_parents[2 * k] + c] + base.coords[3 * r.edge_parents[2 * k + 1] + c]);
  f.num_elems = ne * 8;
  f.conn.resize(f.num_elems * 4);
  f.elem_set.resize(f.num_elems);
  f.num_elem_sets = base.num_elem_sets;
  r.parent.resize(f.num_elems);
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t* v = &base.conn[e * 4];
    const int32_t* me = &edge_vertex[e * 6];
    // midpoints by vertex pair (edge order 01,12,20,03,13,23)
    int32_t m01 = me[0], m12 = me[1], m02 = me[2], m03 = me[3], m13 = me[4], m23 = me[5];
    const int32_t kids[8][4] = {{v[0], m01, m02, m03}, {m01, v[1], m12, m13}, {m02, m12, v[2], m23},
                                {m03, m13, m23, v[3]}, {m01, m02, m03, m13}, {m01, m02, m12, m13},
                                {m02, m03, m13, m23}, {m02, m12, m13, m23}};
    for (int c = 0; c < 8; ++c) {
      int64_t ce = e * 8 + c;
      int32_t* cc = &f.conn[ce * 4];
      for (int a = 0; a < 4; ++a) cc[a] = kids[c][a];
      double d = tet_det(&f.coords[3 * cc[0]], &f.coords[3 * cc[1]], &f.coords[3 * cc[2]], &f.coords[3 * cc[3]]);
      if (d < 0) std::swap(cc[2], cc[3]);
      r.parent[ce] = (int32_t)e;
      f.elem_set[ce] = base.elem_set.empty() ? 0 : base.elem_set[e];
    }
  }
  // node sets: old members plus new vertices whose two parents are members
  for (auto& kv : base.node_sets) {
    std::vector<char> in(nv, 0);
    for (int32_t n : kv.second) in[n] = 1;
    std::vector<int32_t> set(kv.second);
    for (int64_t k = 0; k < nedges; ++k)
      if (in[r.edge_parents[2 * k]] && in[r.edge_parents[2 * k + 1]]) set.push_back((int32_t)(nv + k));
    f.node_sets[kv.first] = std::move(set);
  }
  finalize_mesh(f);
  return r;
}

}  // namespace goal
