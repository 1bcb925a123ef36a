// Mesh adaptation driven by element error indicators, standing in for the
// reference's adapt step (src/main_nested.cpp:108-126):
//
//   * iso_target_size follows get_iso_target_size (src/goal_size_field.cpp:
//     39-151) term by term: G = sum_E |eta_E|^(2d/(2p+d)), size factor
//     (G/N)^(1/d), per-element h_new = factor * |eta_E|^(-2/(2p+d)) * h_E with
//     h_E the RMS edge length, clamped to [alpha h_E, beta h_E] (alpha 0.25,
//     beta 2), then averaged over the elements around each vertex.
//   * adapt_refine replaces ma::adapt with refinement only: every edge longer
//     than split_ratio times the mean target size of its two vertices is
//     bisected by longest-edge propagation (Rivara's LEPP): before an edge is
//     cut, every tetrahedron around it whose own longest edge is a different
//     one has that edge cut first.  Cutting an edge splits every tetrahedron
//     around it at the edge midpoint, so the mesh stays conforming; choosing
//     each element's longest edge keeps the shape of the children bounded.
//     The new vertex takes the mean target size of the edge's two vertices.
//     Passes repeat until no edge is too long or max_passes is reached.  No
//     coarsening and no shape fixing (ma's coarsen / fixShape) are done.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <unordered_map>
#include <utility>
#include <vector>

#include "goal_mesh.hpp"

namespace goal {

namespace {

inline uint64_t edge_key(int32_t a, int32_t b) {
  const uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
  return ((uint64_t)lo << 32) | hi;
}

inline double dist2(const double* x, int32_t a, int32_t b) {
  const double dx = x[3 * a] - x[3 * b], dy = x[3 * a + 1] - x[3 * b + 1], dz = x[3 * a + 2] - x[3 * b + 2];
  return dx * dx + dy * dy + dz * dz;
}

// Total order on edges: squared length, then key.  "Longer" is strict, so the
// longest-edge chains of the propagation always end.
struct EdgeLen {
  double l2;
  uint64_t key;
  bool operator<(const EdgeLen& o) const { return l2 != o.l2 ? l2 < o.l2 : key < o.key; }
};

class Bisector {
 public:
  Bisector(const HostMesh& m, const std::vector<double>& size)
      : coords_(m.coords), size_(size), tets_(m.conn), parent_(m.num_elems), num_old_(m.num_nodes) {
    for (int64_t e = 0; e < m.num_elems; ++e) parent_[e] = (int32_t)e;
    edges_.reserve((size_t)m.num_elems * 2);
    for (int64_t t = 0; t < m.num_elems; ++t) attach(t);
  }

  // One pass: cut every edge that is too long for the size field.
  int64_t pass(double ratio) {
    std::vector<std::pair<EdgeLen, uint64_t>> marked;
    for (const auto& kv : edges_) {
      const int32_t a = (int32_t)(kv.first >> 32), b = (int32_t)(kv.first & 0xffffffffu);
      const double l2 = dist2(coords_.data(), a, b);
      const double s = ratio * 0.5 * (size_[a] + size_[b]);
      if (l2 > s * s) marked.push_back({EdgeLen{l2, kv.first}, kv.first});
    }
    std::sort(marked.begin(), marked.end(),
              [](const std::pair<EdgeLen, uint64_t>& x, const std::pair<EdgeLen, uint64_t>& y) {
                return y.first < x.first;
              });
    const int64_t before = splits_;
    for (const auto& mk : marked)
      if (edges_.count(mk.second)) cut(mk.second);
    return splits_ - before;
  }

  void finish(Adaptation& out) {
    HostMesh& f = out.mesh;
    f.order = 1;
    f.nodes_per_elem = 4;
    f.num_nodes = (int64_t)coords_.size() / 3;
    f.num_elems = (int64_t)tets_.size() / 4;
    f.coords = std::move(coords_);
    f.conn = std::move(tets_);
    f.elem_set.assign(f.num_elems, 0);
    out.parent = std::move(parent_);
    out.new_vertex_parents = std::move(new_parents_);
    out.num_old_vertices = num_old_;
    out.splits = splits_;
    finalize_mesh(f);
  }

 private:
  std::vector<double> coords_;
  std::vector<double> size_;
  std::vector<int32_t> tets_;
  std::vector<int32_t> parent_;
  std::vector<int32_t> new_parents_;
  int64_t num_old_;
  int64_t splits_ = 0;
  std::unordered_map<uint64_t, std::vector<int32_t>> edges_;  // edge -> tetrahedra around it

  void attach(int64_t t) {
    const int32_t* v = &tets_[4 * t];
    for (int k = 0; k < 6; ++k) edges_[edge_key(v[kTetEdge[k][0]], v[kTetEdge[k][1]])].push_back((int32_t)t);
  }
  void detach(int64_t t) {
    const int32_t* v = &tets_[4 * t];
    for (int k = 0; k < 6; ++k) {
      auto it = edges_.find(edge_key(v[kTetEdge[k][0]], v[kTetEdge[k][1]]));
      auto& l = it->second;
      l.erase(std::find(l.begin(), l.end(), (int32_t)t));
      if (l.empty()) edges_.erase(it);
    }
  }
  EdgeLen longest(int64_t t) const {
    const int32_t* v = &tets_[4 * t];
    EdgeLen best{-1.0, 0};
    for (int k = 0; k < 6; ++k) {
      const int32_t a = v[kTetEdge[k][0]], b = v[kTetEdge[k][1]];
      const EdgeLen e{dist2(coords_.data(), a, b), edge_key(a, b)};
      if (best < e) best = e;
    }
    return best;
  }

  // Cut edge `key` after cutting, around it, every longer longest edge.
  void cut(uint64_t key) {
    std::vector<uint64_t> stack{key};
    while (!stack.empty()) {
      const uint64_t k = stack.back();
      auto it = edges_.find(k);
      if (it == edges_.end()) {  // already cut through an earlier chain
        stack.pop_back();
        continue;
      }
      uint64_t first = 0;
      bool blocked = false;
      for (int32_t t : it->second) {
        const EdgeLen le = longest(t);
        if (le.key != k) {
          first = le.key;
          blocked = true;
          break;
        }
      }
      if (blocked) {
        stack.push_back(first);
        continue;
      }
      bisect(k);
      stack.pop_back();
    }
  }

  void bisect(uint64_t key) {
    const int32_t a = (int32_t)(key >> 32), b = (int32_t)(key & 0xffffffffu);
    const int32_t m = (int32_t)(coords_.size() / 3);
    for (int c = 0; c < 3; ++c) coords_.push_back(0.5 * (coords_[3 * a + c] + coords_[3 * b + c]));
    size_.push_back(0.5 * (size_[a] + size_[b]));
    new_parents_.push_back(a);
    new_parents_.push_back(b);
    const std::vector<int32_t> around = edges_.at(key);
    for (int32_t t : around) {
      detach(t);
      const int64_t t2 = (int64_t)tets_.size() / 4;
      int32_t v[4];
      for (int c = 0; c < 4; ++c) v[c] = tets_[4 * t + c];
      tets_.insert(tets_.end(), v, v + 4);
      parent_.push_back(parent_[t]);
      // t keeps a and takes m for b; t2 keeps b and takes m for a.  Moving
      // one vertex along the edge towards the other keeps the orientation.
      for (int c = 0; c < 4; ++c) {
        if (tets_[4 * t + c] == b) tets_[4 * t + c] = m;
        if (tets_[4 * t2 + c] == a) tets_[4 * t2 + c] = m;
      }
      attach(t);
      attach(t2);
    }
    ++splits_;
  }
};

}  // namespace

std::vector<double> iso_target_size(const HostMesh& m, const double* eta, int64_t target, int p_order) {
  if (m.order != 1) throw std::invalid_argument("size field: linear meshes only");
  if (target <= 0) throw std::invalid_argument("size field: target must be positive");
  const double d = 3.0, p = p_order, alpha = 0.25, beta = 2.0;
  const int64_t ne = m.num_elems;
  double G = 0.0;
  for (int64_t e = 0; e < ne; ++e) G += std::pow(std::fabs(eta[e]), (2.0 * d) / (2.0 * p + d));
  const double factor = std::pow(G / (double)target, 1.0 / d);
  std::vector<double> hsum(m.num_nodes, 0.0), cnt(m.num_nodes, 0.0);
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t* v = &m.conn[4 * e];
    double h = 0.0;
    for (int k = 0; k < 6; ++k) h += dist2(m.coords.data(), v[kTetEdge[k][0]], v[kTetEdge[k][1]]);
    h = std::sqrt(h / 6.0);
    const double r = std::pow(std::fabs(eta[e]), -2.0 / (2.0 * p + d));
    double hn = factor * r * h;
    if (hn < alpha * h) hn = alpha * h;
    if (hn > beta * h) hn = beta * h;
    for (int a = 0; a < 4; ++a) {
      hsum[v[a]] += hn;
      cnt[v[a]] += 1.0;
    }
  }
  for (int64_t n = 0; n < m.num_nodes; ++n) hsum[n] = cnt[n] > 0 ? hsum[n] / cnt[n] : 0.0;
  return hsum;
}

Adaptation adapt_refine(const HostMesh& m, const std::vector<double>& vtx_size, double split_ratio,
                        int max_passes) {
  if (m.order != 1) throw std::invalid_argument("adapt: linear meshes only");
  if ((int64_t)vtx_size.size() != m.num_nodes) throw std::invalid_argument("adapt: size field length");
  if (!(split_ratio > 0.0)) throw std::invalid_argument("adapt: split ratio must be positive");
  Bisector bis(m, vtx_size);
  Adaptation out;
  int passes = 0;
  while (passes < max_passes) {
    ++passes;
    if (bis.pass(split_ratio) == 0) break;
  }
  bis.finish(out);
  out.passes = passes;
  return out;
}

}  // namespace goal
