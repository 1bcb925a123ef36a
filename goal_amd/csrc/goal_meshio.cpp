// goal's mesh inputs on the host: a serial SCOREC MDS mesh file (.smb), its
// gmi model (.dmg) and goal's association file, as Disc's constructor reads
// them (apf::loadMdsMesh + read_sets, src/goal_disc.cpp:28-99) and as
// compute_elem_sets / compute_side_sets / compute_node_sets turn them into
// sets (src/goal_disc.cpp:347-396).
//
// The .smb layout read here (version 5; all integers are 32-bit unsigned and
// all reals 64-bit, both big-endian):
//   magic (0), version, dimension, part count
//   entity counts by type: vertex, edge, triangle, quad, hex, prism, pyramid, tet
//   downward adjacency: edges -> 2 vertices, triangles -> 3 edges, tets -> 4 triangles
//   vertex coordinates [num_vertices][3], vertex parametric coordinates [num_vertices][2]
//   remote copies of vertices: peer count P, the P peer parts, the P copy
//   counts, then per peer the local ids of the shared vertices (the k-th id
//   a part lists for peer p and the k-th id p lists for it are one vertex)
//   classification: (model tag, model dimension) per entity, by type
//   (tags and the coordinate field's shape follow and are not needed)
// A partitioned mesh "name.smb" is stored as name0.smb ... name<P-1>.smb (the
// part number goes before the extension, as apf writes and reads them); its
// parts are merged into one mesh through the vertex copies.
// The .dmg (gmi "mesh model") text: counts of regions, faces, edges and
// vertices; two bounding-box lines; vertices (tag x y z); edges (tag, its two
// vertex tags); faces (tag, loop count, then per loop its edge count and
// (edge tag, direction) lines); regions (tag, shell count, then per shell its
// face count and (face tag, direction) lines).
#include "goal_meshio.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <fstream>
#include <functional>
#include <set>
#include <sstream>
#include <stdexcept>

namespace goal {
namespace {

[[noreturn]] void fail(const std::string& what) { throw std::invalid_argument(what); }

std::vector<uint8_t> slurp(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) fail("cannot open " + path);
  return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

struct BigEndian {
  const std::vector<uint8_t>& b;
  size_t at = 0;
  std::string name;
  uint32_t u32() {
    if (at + 4 > b.size()) fail(name + ": file ends early");
    const uint32_t v = (uint32_t)b[at] << 24 | (uint32_t)b[at + 1] << 16 | (uint32_t)b[at + 2] << 8 | b[at + 3];
    at += 4;
    return v;
  }
  double f64() {
    if (at + 8 > b.size()) fail(name + ": file ends early");
    uint64_t v = 0;
    for (int k = 0; k < 8; ++k) v = v << 8 | b[at + k];
    at += 8;
    double d;
    std::memcpy(&d, &v, 8);
    return d;
  }
};

struct Smb {
  int64_t nv = 0, ne = 0, nt = 0, ntet = 0;
  std::vector<int32_t> edge_verts, tri_edges, tet_tris;  // downward adjacency
  std::vector<double> coords;
  std::vector<std::array<int32_t, 2>> cls[4];  // (dim, tag) of vertices, edges, triangles, tets
  uint32_t parts = 1;
  std::vector<std::pair<uint32_t, std::vector<int32_t>>> links;  // peer part, shared local vertex ids
};

bool exists(const std::string& path) { return std::ifstream(path).good(); }

Smb read_smb(const std::string& path, uint32_t part = 0) {
  const std::vector<uint8_t> bytes = slurp(path);
  BigEndian in{bytes, 0, path};
  if (in.u32() != 0) fail(path + ": not an MDS mesh file (bad magic number)");
  const uint32_t version = in.u32();
  if (version != 5) fail(path + ": MDS file version " + std::to_string(version) + " (only version 5 is read)");
  if (in.u32() != 3) fail(path + ": not a 3D mesh");
  const uint32_t parts = in.u32();
  if (parts < 1 || part >= parts) fail(path + ": bad part count");
  uint32_t n[8];
  for (auto& c : n) c = in.u32();
  if (n[3] || n[4] || n[5] || n[6]) fail(path + ": has non-tetrahedral elements");
  Smb s;
  s.parts = parts;
  s.nv = n[0];
  s.ne = n[1];
  s.nt = n[2];
  s.ntet = n[7];
  if (s.nv < 4 || s.ntet < 1) fail(path + ": no tetrahedra");
  // The adjacency, coordinates and parametric coordinates follow the counts;
  // a file too short for them is rejected before anything is sized by them.
  const uint64_t body = 4 * (2 * (uint64_t)s.ne + 3 * (uint64_t)s.nt + 4 * (uint64_t)s.ntet) + 40 * (uint64_t)s.nv;
  if (in.at + body > bytes.size()) fail(path + ": file ends early");
  auto down = [&](int64_t count, int per, int64_t bound, std::vector<int32_t>& out) {
    out.resize(count * per);
    for (auto& v : out) {
      const uint32_t x = in.u32();
      if ((int64_t)x >= bound) fail(path + ": adjacency entry out of range");
      v = (int32_t)x;
    }
  };
  down(s.ne, 2, s.nv, s.edge_verts);
  down(s.nt, 3, s.ne, s.tri_edges);
  down(s.ntet, 4, s.nt, s.tet_tris);
  s.coords.resize(3 * s.nv);
  for (auto& x : s.coords) x = in.f64();
  for (int64_t k = 0; k < 2 * s.nv; ++k) in.f64();  // parametric coordinates
  const uint32_t npeers = in.u32();
  if (npeers >= parts) fail(path + ": more peers than parts");
  s.links.resize(npeers);
  for (auto& l : s.links) {
    l.first = in.u32();
    if (l.first >= parts || l.first == part) fail(path + ": bad peer part");
  }
  std::vector<uint32_t> counts(npeers);
  for (auto& c : counts) {
    c = in.u32();
    if (c > s.nv) fail(path + ": more copies than vertices");
  }
  for (uint32_t k = 0; k < npeers; ++k) {
    s.links[k].second.resize(counts[k]);
    for (auto& v : s.links[k].second) {
      const uint32_t x = in.u32();
      if (x >= s.nv) fail(path + ": remote copy of a vertex out of range");
      v = (int32_t)x;
    }
  }
  const int64_t cnt[4] = {s.nv, s.ne, s.nt, s.ntet};
  for (int d = 0; d < 4; ++d) {
    s.cls[d].resize(cnt[d]);
    for (auto& c : s.cls[d]) {
      const uint32_t tag = in.u32(), dim = in.u32();
      if (dim > 3 || (int)dim < (d == 3 ? 3 : 0) || (int)dim < d) fail(path + ": bad classification");
      c = {(int32_t)dim, (int32_t)tag};
    }
  }
  return s;
}

// The parts of a partitioned mesh as one mesh: vertices joined through their
// remote copies (first copy's coordinates and classification), edges and
// triangles joined by their vertices, tets in part order.
Smb merge_parts(const std::vector<Smb>& P, const std::string& name) {
  const size_t np = P.size();
  std::vector<int64_t> base(np + 1, 0);
  for (size_t a = 0; a < np; ++a) base[a + 1] = base[a] + P[a].nv;
  std::vector<int64_t> up(base[np]);
  for (int64_t k = 0; k < base[np]; ++k) up[k] = k;
  std::function<int64_t(int64_t)> root = [&](int64_t k) {
    while (up[k] != k) k = up[k] = up[up[k]];
    return k;
  };
  for (size_t a = 0; a < np; ++a)
    for (auto const& l : P[a].links) {
      const Smb& q = P[l.first];
      const std::vector<int32_t>* back = nullptr;
      for (auto const& m : q.links)
        if (m.first == a) back = &m.second;
      if (!back || back->size() != l.second.size()) fail(name + ": remote copies of parts do not pair up");
      for (size_t k = 0; k < l.second.size(); ++k) {
        const int64_t x = root(base[a] + l.second[k]), y = root(base[l.first] + (*back)[k]);
        if (x != y) up[std::max(x, y)] = std::min(x, y);
      }
    }
  Smb m;
  std::vector<int32_t> gid(base[np], -1);
  for (size_t a = 0; a < np; ++a)
    for (int64_t v = 0; v < P[a].nv; ++v) {
      const int64_t r = root(base[a] + v);
      if (gid[r] < 0) {
        gid[r] = (int32_t)m.nv++;
        for (int c = 0; c < 3; ++c) m.coords.push_back(P[a].coords[3 * v + c]);
        m.cls[0].push_back(P[a].cls[0][v]);
      }
      gid[base[a] + v] = gid[r];
    }
  std::map<std::pair<int32_t, int32_t>, int32_t> edges;
  std::map<std::array<int32_t, 3>, int32_t> tris;
  for (size_t a = 0; a < np; ++a) {
    const Smb& q = P[a];
    std::vector<int32_t> eid(q.ne), tid(q.nt);
    for (int64_t e = 0; e < q.ne; ++e) {
      const int32_t x = gid[base[a] + q.edge_verts[2 * e]], y = gid[base[a] + q.edge_verts[2 * e + 1]];
      auto it = edges.emplace(std::make_pair(std::min(x, y), std::max(x, y)), (int32_t)m.ne);
      if (it.second) {
        m.edge_verts.push_back(x);
        m.edge_verts.push_back(y);
        m.cls[1].push_back(q.cls[1][e]);
        ++m.ne;
      }
      eid[e] = it.first->second;
    }
    for (int64_t t = 0; t < q.nt; ++t) {
      std::array<int32_t, 3> ev = {eid[q.tri_edges[3 * t]], eid[q.tri_edges[3 * t + 1]], eid[q.tri_edges[3 * t + 2]]};
      std::set<int32_t> vs;
      for (int32_t e : ev) {
        vs.insert(m.edge_verts[2 * e]);
        vs.insert(m.edge_verts[2 * e + 1]);
      }
      if (vs.size() != 3) fail(name + ": a triangle's edges do not close");
      std::array<int32_t, 3> key;
      std::copy(vs.begin(), vs.end(), key.begin());
      auto it = tris.emplace(key, (int32_t)m.nt);
      if (it.second) {
        for (int32_t e : ev) m.tri_edges.push_back(e);
        m.cls[2].push_back(q.cls[2][t]);
        ++m.nt;
      }
      tid[t] = it.first->second;
    }
    for (int64_t k = 0; k < q.ntet; ++k) {
      for (int f = 0; f < 4; ++f) m.tet_tris.push_back(tid[q.tet_tris[4 * k + f]]);
      m.cls[3].push_back(q.cls[3][k]);
      ++m.ntet;
    }
  }
  return m;
}

// One-part file as named, else the parts name0.smb ... of a partitioned mesh.
Smb read_mesh(const std::string& path) {
  if (exists(path)) {
    Smb s = read_smb(path);
    if (s.parts != 1) fail(path + ": one part of a partitioned mesh; name the mesh without its part number");
    return s;
  }
  const size_t dot = path.rfind(".smb");
  if (dot == std::string::npos || dot + 4 != path.size()) fail("cannot open " + path);
  auto part_file = [&](uint32_t k) { return path.substr(0, dot) + std::to_string(k) + ".smb"; };
  if (!exists(part_file(0))) fail("cannot open " + path + " or its part " + part_file(0));
  std::vector<Smb> P;
  P.push_back(read_smb(part_file(0), 0));
  for (uint32_t k = 1; k < P[0].parts; ++k) {
    P.push_back(read_smb(part_file(k), k));
    if (P.back().parts != P[0].parts) fail(part_file(k) + ": part counts differ");
  }
  return P.size() == 1 ? P[0] : merge_parts(P, path);
}

// Model topology: the entities in the closure of each model entity.
struct Model {
  std::map<std::array<int32_t, 2>, std::vector<std::array<int32_t, 2>>> down;  // (dim, tag) -> bounding entities
  std::set<std::array<int32_t, 2>> all;
  void closure(std::array<int32_t, 2> e, std::set<std::array<int32_t, 2>>& out) const {
    if (!out.insert(e).second) return;
    auto it = down.find(e);
    if (it != down.end())
      for (auto const& b : it->second) closure(b, out);
  }
};

Model read_dmg(const std::string& path) {
  std::ifstream f(path);
  if (!f) fail("cannot open " + path);
  int64_t nr, nf, ne, nvx;
  if (!(f >> nr >> nf >> ne >> nvx)) fail(path + ": bad header");
  double box[6];
  for (double& b : box)
    if (!(f >> b)) fail(path + ": bad bounding box");
  Model m;
  for (int64_t k = 0; k < nvx; ++k) {
    int32_t tag;
    double x, y, z;
    if (!(f >> tag >> x >> y >> z)) fail(path + ": bad vertex line");
    m.all.insert({0, tag});
  }
  for (int64_t k = 0; k < ne; ++k) {
    int32_t tag, a, b;
    if (!(f >> tag >> a >> b)) fail(path + ": bad edge line");
    m.all.insert({1, tag});
    m.down[{1, tag}] = {{0, a}, {0, b}};
  }
  auto bounded = [&](int dim, int64_t count) {
    for (int64_t k = 0; k < count; ++k) {
      int32_t tag, nloops;
      if (!(f >> tag >> nloops)) fail(path + ": bad face or region line");
      m.all.insert({dim, tag});
      auto& d = m.down[{dim, tag}];
      for (int32_t l = 0; l < nloops; ++l) {
        int32_t nuse;
        if (!(f >> nuse)) fail(path + ": bad loop or shell line");
        for (int32_t u = 0; u < nuse; ++u) {
          int32_t t, dir;
          if (!(f >> t >> dir)) fail(path + ": bad loop or shell entry");
          d.push_back({dim - 1, t});
        }
      }
    }
  };
  bounded(2, nf);
  bounded(3, nr);
  for (auto const& kv : m.down)
    for (auto const& b : kv.second)
      if (!m.all.count(b)) fail(path + ": a bounding entity is not in the model");
  return m;
}

struct AssocSet {
  int kind;  // 0 node set, 1 side set, 2 element set
  std::string name;
  std::vector<std::array<int32_t, 2>> ents;
};

// read_sets (src/goal_disc.cpp:28-67): "node set | side set | elem set <name>
// <count>" lines, each followed by <count> "<model dim> <model tag>" lines;
// reading stops at the first empty line.
std::vector<AssocSet> read_assoc(const std::string& path, const Model& model) {
  std::ifstream f(path);
  if (!f) fail("cannot open " + path);
  static const char* kinds[3] = {"node set", "side set", "elem set"};
  std::vector<AssocSet> out;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty()) break;
    int kind = -1;
    for (int k = 0; k < 3; ++k)
      if (line.compare(0, std::strlen(kinds[k]), kinds[k]) == 0) kind = k;
    if (kind < 0) fail(path + ": invalid association line: " + line);
    std::stringstream ss(line.substr(std::strlen(kinds[kind])));
    AssocSet s;
    s.kind = kind;
    int n;
    if (!(ss >> s.name >> n) || n < 0) fail(path + ": invalid association line: " + line);
    const int want = kind == 0 ? 0 : kind == 1 ? 2 : 3;
    for (int k = 0; k < n; ++k) {
      std::string el;
      if (!std::getline(f, el) || el.empty()) fail(path + ": association ends early after " + s.name);
      std::stringstream es(el);
      int32_t dim, tag;
      if (!(es >> dim >> tag)) fail(path + ": bad association entry: " + el);
      if (!model.all.count({dim, tag}))
        fail(path + ": no model entity with dim " + std::to_string(dim) + " and tag " + std::to_string(tag));
      if (kind > 0 && dim != want) fail(path + ": side sets need model faces and element sets model regions");
      s.ents.push_back({dim, tag});
    }
    out.push_back(std::move(s));
  }
  return out;
}

// Local vertices of tet face f (apf's order), the tet edges in the P2 node order.
constexpr int kFace[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};

}  // namespace

MeshFile read_mesh_files(const std::string& geom_file, const std::string& mesh_file, const std::string& assoc_file,
                         int order) {
  if (order != 1 && order != 2) fail("mesh order must be 1 or 2");
  const Smb s = read_mesh(mesh_file);
  const Model model = read_dmg(geom_file);
  const std::vector<AssocSet> sets = read_assoc(assoc_file, model);
  for (int d = 0; d < 4; ++d)
    for (auto const& c : s.cls[d])
      if (!model.all.count(c)) fail(mesh_file + ": classified on an entity the model lacks");

  MeshFile out;
  HostMesh& m = out.mesh;
  m.order = order;
  m.nodes_per_elem = order == 1 ? 4 : 10;
  m.num_elems = s.ntet;
  m.num_nodes = s.nv + (order == 2 ? s.ne : 0);
  m.coords = s.coords;
  if (order == 2) {
    m.coords.resize(3 * m.num_nodes);
    for (int64_t e = 0; e < s.ne; ++e)
      for (int c = 0; c < 3; ++c)
        m.coords[3 * (s.nv + e) + c] =
            0.5 * (s.coords[3 * s.edge_verts[2 * e] + c] + s.coords[3 * s.edge_verts[2 * e + 1] + c]);
  }
  // tet vertices: a triangle's three, then the fourth from a second triangle;
  // positively oriented
  std::map<std::pair<int32_t, int32_t>, int32_t> edge_of;
  for (int64_t e = 0; e < s.ne; ++e) {
    const int32_t a = s.edge_verts[2 * e], b = s.edge_verts[2 * e + 1];
    edge_of[{std::min(a, b), std::max(a, b)}] = (int32_t)e;
  }
  auto tri_verts = [&](int32_t t, int32_t* v) {
    const int32_t* e = &s.tri_edges[3 * t];
    v[0] = s.edge_verts[2 * e[0]];
    v[1] = s.edge_verts[2 * e[0] + 1];
    const int32_t c0 = s.edge_verts[2 * e[1]], c1 = s.edge_verts[2 * e[1] + 1];
    v[2] = (c0 != v[0] && c0 != v[1]) ? c0 : c1;
    if (v[2] == v[0] || v[2] == v[1] || v[0] == v[1]) fail(mesh_file + ": degenerate triangle");
  };
  m.conn.resize(m.num_elems * m.nodes_per_elem);
  std::vector<std::vector<int32_t>> tri_tets(s.nt);
  for (int64_t k = 0; k < s.ntet; ++k) {
    int32_t v[4], w[3];
    tri_verts(s.tet_tris[4 * k], v);
    tri_verts(s.tet_tris[4 * k + 1], w);
    v[3] = -1;
    for (int32_t x : w)
      if (x != v[0] && x != v[1] && x != v[2]) v[3] = x;
    if (v[3] < 0) fail(mesh_file + ": a tet's triangles do not span four vertices");
    const double* X[4];
    for (int a = 0; a < 4; ++a) X[a] = &s.coords[3 * (int64_t)v[a]];
    double d1[3], d2[3], d3[3];
    for (int c = 0; c < 3; ++c) {
      d1[c] = X[1][c] - X[0][c];
      d2[c] = X[2][c] - X[0][c];
      d3[c] = X[3][c] - X[0][c];
    }
    const double det = d1[0] * (d2[1] * d3[2] - d2[2] * d3[1]) - d1[1] * (d2[0] * d3[2] - d2[2] * d3[0]) +
                       d1[2] * (d2[0] * d3[1] - d2[1] * d3[0]);
    if (det == 0.0) fail(mesh_file + ": a tet has zero volume");
    if (det < 0) std::swap(v[1], v[2]);
    int32_t* c = &m.conn[k * m.nodes_per_elem];
    for (int a = 0; a < 4; ++a) c[a] = v[a];
    if (order == 2)
      for (int ed = 0; ed < 6; ++ed) {
        const int32_t a = v[kTetEdge[ed][0]], b = v[kTetEdge[ed][1]];
        auto it = edge_of.find({std::min(a, b), std::max(a, b)});
        if (it == edge_of.end()) fail(mesh_file + ": a tet edge is missing from the edge list");
        c[4 + ed] = (int32_t)(s.nv + it->second);
      }
    for (int f = 0; f < 4; ++f) tri_tets[s.tet_tris[4 * k + f]].push_back((int32_t)k);
  }
  // element sets: the assoc order gives the set index (compute_elem_sets)
  std::map<std::array<int32_t, 2>, int> region_set;
  for (auto const& st : sets)
    if (st.kind == 2) {
      const int idx = (int)out.elem_set_names.size();
      out.elem_set_names.push_back(st.name);
      for (auto const& e : st.ents) region_set[e] = idx;
    }
  m.elem_set.resize(m.num_elems);
  for (int64_t k = 0; k < s.ntet; ++k) {
    auto it = region_set.find(s.cls[3][k]);
    if (it == region_set.end()) fail(mesh_file + ": a tet's model region is in no element set");
    m.elem_set[k] = it->second;
  }
  m.num_elem_sets = std::max<int>(1, (int)out.elem_set_names.size());
  // node sets: nodes classified in the closure of a set's model entities
  // (apf::collectEntityModels walks up from the node's model entity)
  for (auto const& st : sets)
    if (st.kind == 0) {
      std::set<std::array<int32_t, 2>> cl;
      for (auto const& e : st.ents) model.closure(e, cl);
      std::vector<int32_t> nodes;
      for (int64_t v = 0; v < s.nv; ++v)
        if (cl.count(s.cls[0][v])) nodes.push_back((int32_t)v);
      if (order == 2)
        for (int64_t e = 0; e < s.ne; ++e)
          if (cl.count(s.cls[1][e])) nodes.push_back((int32_t)(s.nv + e));
      m.node_sets[st.name] = std::move(nodes);
    }
  // side sets: triangles classified on the set's model faces
  // (compute_side_sets), as (element, local face) of an adjacent tet
  for (auto const& st : sets)
    if (st.kind == 1) {
      std::set<std::array<int32_t, 2>> faces(st.ents.begin(), st.ents.end());
      std::vector<int32_t> sides;
      for (int64_t t = 0; t < s.nt; ++t) {
        if (!faces.count(s.cls[2][t])) continue;
        if (tri_tets[t].empty()) fail(mesh_file + ": a side-set triangle bounds no tet");
        const int32_t k = tri_tets[t][0];
        int32_t tv[3];
        tri_verts((int32_t)t, tv);
        std::sort(tv, tv + 3);
        const int32_t* c = &m.conn[(int64_t)k * m.nodes_per_elem];
        int lf = -1;
        for (int f = 0; f < 4; ++f) {
          int32_t fv[3] = {c[kFace[f][0]], c[kFace[f][1]], c[kFace[f][2]]};
          std::sort(fv, fv + 3);
          if (std::equal(fv, fv + 3, tv)) lf = f;
        }
        if (lf < 0) fail(mesh_file + ": a side-set triangle is not a face of its tet");
        sides.push_back(4 * k + lf);
      }
      std::sort(sides.begin(), sides.end());
      out.side_sets[st.name] = std::move(sides);
    }
  finalize_mesh(m);
  return out;
}

}  // namespace goal
