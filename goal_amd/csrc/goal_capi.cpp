// C ABI of goal_amd (declared in include/goal_amd.h).  Host C++ that owns the
// device buffers, builds meshes and graphs, and queues the kernels of
// goal_assembly.hip / goal_util.hip / goal_dwr.hip on one HIP stream.
#include "goal_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "goal_amg.hpp"
#include "goal_expr.hpp"
#include "goal_kernels.hpp"
#include "goal_mesh.hpp"
#include "goal_meshio.hpp"

namespace {

thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

struct HipError : std::runtime_error {
  int code;
  HipError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    throw HipError(GOAL_ERR_ALLOC, std::string(what) + ": " + hipGetErrorString(e));
  throw HipError(GOAL_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
// A HIP call whose status is deliberately ignored (a free in a destructor, an
// occupancy query with a fallback) still leaves its error in the thread's
// last-error slot, where the hipGetLastError() after the next kernel launch
// would report it as that launch's failure.  Each checked call therefore
// first takes any such leftover out of the slot, naming it on stderr.
void clear_stale_error(const char* before) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    std::fprintf(stderr, "goal_amd: an unchecked HIP call failed earlier (%s); seen before %s\n",
                 hipGetErrorString(e), before);
}
#define HIPCHK(x)              \
  do {                         \
    clear_stale_error(#x);     \
    hip_check((x), #x);        \
  } while (0)

// Every entry point runs its body through guard(): exceptions become status
// codes plus a message, never an abort.
template <typename F>
int guard(F&& f) {
  try {
    return f();
  } catch (const HipError& e) {
    return set_error(e.code, "%s", e.what());
  } catch (const std::invalid_argument& e) {
    return set_error(GOAL_ERR_INVALID, "%s", e.what());
  } catch (const std::bad_alloc&) {
    return set_error(GOAL_ERR_ALLOC, "host allocation failed");
  } catch (const std::exception& e) {
    return set_error(GOAL_ERR_INVALID, "%s", e.what());
  }
}

void require(bool ok, const char* msg) {
  if (!ok) throw std::invalid_argument(msg);
}

// ------------------------------------------- allocation-free assembly scratch
// Element residuals and element matrices of goal_asm_jacobian_noalloc live in
// this file-scope device array instead of buffers the library allocates.  Each
// GPU holds its own copy at its own address until the process ends.  In host
// code `scratch` names a host-side stand-in, not device memory: host code uses
// only the address hipGetSymbolAddress returned for the current device.
constexpr int64_t kNoallocScratchDoubles = int64_t(1) << 23;  // 64 MiB
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

constexpr int kMaxDevices = 64;
double* g_noalloc_ptr[kMaxDevices] = {};

// Once per device, with that device current (called from the context set-up).
void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) throw HipError(GOAL_ERR_HIP, "device index beyond the scratch table");
  if (g_noalloc_ptr[device]) return;
  void* p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_noalloc_ptr[device] = static_cast<double*>(p);
}

// ---------------------------------------------------------------- context
struct Context {
  bool ready = false;
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;        // gathers of a chunked assembly (default priority)
  hipEvent_t chunk_ev[65] = {};      // chunk boundaries of a chunked assembly
  hipEvent_t events[256] = {};
  double* reduce = nullptr;  // reductions: kReduceBlocks partials + results
  std::mutex mu;
};
Context g_ctx;

Context& ctx() {
  if (g_ctx.ready) return g_ctx;
  std::lock_guard<std::mutex> lock(g_ctx.mu);
  if (g_ctx.ready) return g_ctx;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) throw HipError(GOAL_ERR_NO_DEVICE, "no HIP device available");
  HIPCHK(hipGetDevice(&g_ctx.device));
  HIPCHK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&g_ctx.side, hipStreamNonBlocking));
  for (auto& ev : g_ctx.events) HIPCHK(hipEventCreate(&ev));
  for (auto& ev : g_ctx.chunk_ev) HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  HIPCHK(hipMalloc(&g_ctx.reduce, sizeof(double) * (goal::kReduceBlocks + 16)));
  noalloc_setup(g_ctx.device);
  g_ctx.ready = true;
  return g_ctx;
}

hipStream_t stream() { return ctx().stream; }
double* reduce_scratch() { return ctx().reduce; }
double* result_slot(int k) { return ctx().reduce + goal::kReduceBlocks + k; }

template <typename T>
struct DevBuf {
  T* p = nullptr;
  int64_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(int64_t count) {
    release();
    if (count <= 0) return;
    HIPCHK(hipMalloc(&p, sizeof(T) * count));
    n = count;
  }
  void upload(const T* host, int64_t count) {
    alloc(count);
    if (count > 0) HIPCHK(hipMemcpyAsync(p, host, sizeof(T) * count, hipMemcpyHostToDevice, stream()));
  }
  void download(T* host, int64_t count) const {
    if (count > 0) {
      HIPCHK(hipMemcpyAsync(host, p, sizeof(T) * count, hipMemcpyDeviceToHost, stream()));
      HIPCHK(hipStreamSynchronize(stream()));
    }
  }
};

void sync() { HIPCHK(hipStreamSynchronize(stream())); }

int host_threads() {
  unsigned t = std::thread::hardware_concurrency();
  if (t < 1) t = 1;
  return (int)std::min<unsigned>(t, 32);
}

goal::MatParams derive(const goal_material& m) {
  goal::MatParams p;
  p.mu = m.E / (2.0 * (1.0 + m.nu));
  p.kappa = m.E / (3.0 * (1.0 - 2.0 * m.nu));
  p.lambda = m.E * m.nu / ((1.0 + m.nu) * (1.0 - 2.0 * m.nu));
  p.K = m.K;
  p.Y = m.Y;
  p.c0 = m.c0;
  return p;
}

}  // namespace

// ---------------------------------------------------------------- handles
// Aggregation multigrid over a disc's node-block matrices (the preconditioner
// of goal_linear_solve).  Structure once per disc; values for every solve.
struct AmgLevel {
  int64_t nn = 0;                         // nodes (4 equations each)
  DevBuf<int64_t> bptr;                   // coarse levels: node-block CSR
  DevBuf<int32_t> bcol;
  DevBuf<double> bval;
  std::vector<int64_t> bptr_h;            // the coarsest level's graph, for its dense inverse
  std::vector<int32_t> bcol_h;
  int64_t nc = 0, ncb = 0;                // transfer to the next level
  DevBuf<int32_t> agg, amem;
  DevBuf<int64_t> aptr, gptr, gidx;
  DevBuf<double> binv, b, x, r, d;        // smoother inverse blocks, work vectors
  double theta = 1.0, c2 = 0.0;           // smoother ellipse: centre, squared focal half-distance
};
struct AmgHierarchy {
  std::vector<std::unique_ptr<AmgLevel>> lv;
  DevBuf<int32_t> brow0;                  // row node of every finest-level block
  DevBuf<double> coarse_inv;              // dense inverse of the coarsest level
  DevBuf<double> arn;                     // Arnoldi basis of the smoother's spectrum estimate
  int coarse_n = 0;
};

struct goal_disc_s {
  goal::HostMesh mesh;
  std::vector<int64_t> node_adj_ptr;
  std::vector<int32_t> node_adj;          // host copy, made on demand (host_node_adj)
  DevBuf<int32_t> node_adj_dev;           // the columns as built on the device
  std::unique_ptr<AmgHierarchy> amg;
  int max_degree = 0;
  int64_t num_dofs = 0, nnz = 0;
  // gather plan: runs of consecutive nodes whose CSR values fit `gather_cap`
  int gather_cap = 0;
  int64_t num_ranges = 0;
  DevBuf<int32_t> ranges;
  std::vector<int32_t> ranges_host;      // the same node runs on the host
  std::vector<int64_t> range_last_elem;  // per run: largest element touching its nodes (lazy)
  DevBuf<double> coords, x;
  DevBuf<int32_t> conn, elem_set, cols, vertex_nodes, inc, rec;
  DevBuf<uint16_t> pos, posi;
  DevBuf<int64_t> row_ptr, inc_ptr;
  std::map<std::string, std::unique_ptr<DevBuf<int32_t>>> dev_node_sets;
  DevBuf<double> dbc_values;
  // side sets as 4 * element + local face: given ones, and ones derived from
  // the node set of the same name (dropped when that node set changes)
  std::map<std::string, std::vector<int32_t>> side_sets, derived_side_sets;
  DevBuf<int64_t> tbc_rows;
  DevBuf<double> tbc_values;
};

struct goal_states_s {
  goal_disc d = nullptr;
  int model = 0, nq = 1;
  int64_t npts = 0;
  DevBuf<double> sigma, eqps, eqps_old, Fp, Fp_old;
};

// zero_* only mark a buffer as pending zero; the next assembly then writes
// it instead of adding to it, and any other access fills it first.
struct goal_solinfo_s {
  goal_disc d = nullptr;
  DevBuf<double> R, dMdu, vals;
  bool zero_R = false, zero_dMdu = false, zero_vals = false;
  // linear-solve workspace, kept between solves
  DevBuf<double> V, t, z, binv, partial, h1, h2, y;
  DevBuf<int32_t> solve_status;
};

struct goal_mech_s {
  goal_disc d = nullptr;
  goal_states st = nullptr;
  int model = 0, stab = 1, quad_order = 1, num_sets = 1;
  int blocks = 0, xcd_remap = 1, check = 1;
  int tangent = 64;  // 32: FP32 element matrices for the Jacobian kinds
  // neo-Hookean assemblies in chunks, gathers on a second stream: > 1 chunks,
  // 0 off, -1 automatic (one pass: with matrix-core element matrices chunking
  // no longer pays)
  int overlap = -1;
  int overlap_blocks = 0;  // element workgroups per chunk (0: 5/6 of the resident grid)
  DevBuf<goal::MatParams> mats;
  DevBuf<int32_t> status;
  DevBuf<double> elem_value;
  DevBuf<double> emr, em;  // element residuals / matrices between the two passes
  DevBuf<double> pt;       // J2 point states between the point pass and the element pass
  DevBuf<double> feat_vol, feat_vm;  // learned-indicator feature scratch
  bool timing = false, timed = false;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  ~goal_mech_s() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

// Batched assembly: many (mechanics, sol info) pairs captured into one HIP
// graph, so a whole batch of meshes is one launch.
struct goal_batch_s {
  std::vector<goal_mech> mechs;
  std::vector<goal_solinfo> infos;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~goal_batch_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

struct goal_mlp_s {
  DevBuf<float> params, norms, adam_m, adam_v, grad, partial, partial_loss, loss;
  DevBuf<uint16_t> wq;
  DevBuf<int> ood;
  int step = 0;
  goal::MlpArgs args() {
    return goal::MlpArgs{params.p, wq.p, norms.p, adam_m.p, adam_v.p};
  }
};

struct goal_nested_s {
  goal_disc base = nullptr;
  goal_disc fine = nullptr;
  int64_t num_old = 0, num_new = 0;
  std::vector<int32_t> parent_host, edge_parents_host;
  DevBuf<int32_t> parent, edge_parents;
};

namespace {

// The discretization graph: on the device (radix sorts, goal_graph.hip) when
// the element-pair count fits the sorts' 32-bit item counts, else on the
// host (goal::build_graph); both give the same arrays bit for bit.  The host
// keeps the node adjacency's row pointers (gather plan); the column lists are
// copied to the host only when the multigrid set-up asks for them.
void disc_build(goal_disc d) {
  const goal::HostMesh& m = d->mesh;
  const int npe = m.nodes_per_elem;
  const int64_t nn = m.num_nodes, ne = m.num_elems, ns = ne * npe, np = ns * npe;
  d->amg.reset();
  d->coords.upload(m.coords.data(), (int64_t)m.coords.size());
  d->conn.upload(m.conn.data(), (int64_t)m.conn.size());
  d->elem_set.upload(m.elem_set.data(), (int64_t)m.elem_set.size());
  d->vertex_nodes.upload(m.vertex_nodes.data(), (int64_t)m.vertex_nodes.size());
  require(ns <= (int64_t)INT32_MAX, "graph: too many element-node incidences");
  DevBuf<int64_t> adj_ptr;
  DevBuf<int32_t> adj;
  d->node_adj.clear();
  if (np < (int64_t)INT32_MAX && nn < ((int64_t)1 << 30)) {
    size_t tb = 0;
    HIPCHK(goal::graph_temp_bytes(ne, npe, nn, &tb));
    DevBuf<uint8_t> temp;
    temp.alloc((int64_t)tb + 256);
    int64_t nu = 0;
    {
      DevBuf<uint64_t> keys, sorted;
      DevBuf<int64_t> count;
      keys.alloc(np);
      sorted.alloc(np);
      count.alloc(1);
      HIPCHK(goal::graph_adjacency(d->conn.p, ne, npe, nn, temp.p, tb, keys.p, sorted.p, count.p, stream()));
      count.download(&nu, 1);
      adj_ptr.alloc(nn + 1);
      adj.alloc(nu);
      HIPCHK(goal::graph_rows(keys.p, nu, nn, adj_ptr.p, adj.p, stream()));
      sync();  // the key buffers are released here
    }
    {
      DevBuf<int32_t> node_keys, slots;
      node_keys.alloc(ns);
      slots.alloc(ns);
      d->inc.alloc(ns);
      d->inc_ptr.alloc(nn + 1);
      HIPCHK(goal::graph_incidence(d->conn.p, ne, npe, nn, temp.p, tb, node_keys.p, slots.p, d->inc.p, d->inc_ptr.p,
                                   stream()));
      d->pos.alloc(np);
      d->rec.alloc(ns);
      d->posi.alloc(ns * npe);
      HIPCHK(goal::graph_positions(d->conn.p, ne, npe, adj_ptr.p, adj.p, d->inc.p, d->pos.p, d->rec.p, d->posi.p,
                                   stream()));
      sync();
    }
    d->node_adj_ptr.resize(nn + 1);
    adj_ptr.download(d->node_adj_ptr.data(), nn + 1);
    d->max_degree = 0;
    for (int64_t a = 0; a < nn; ++a)
      d->max_degree = std::max(d->max_degree, (int)(d->node_adj_ptr[a + 1] - d->node_adj_ptr[a]));
  } else {
    goal::HostGraph g = goal::build_graph(m, host_threads());
    d->node_adj_ptr = g.node_adj_ptr;
    d->node_adj = std::move(g.node_adj);
    d->max_degree = g.max_degree;
    d->pos.upload(g.pos.data(), (int64_t)g.pos.size());
    d->inc_ptr.upload(g.inc_ptr.data(), (int64_t)g.inc_ptr.size());
    d->inc.upload(g.inc.data(), (int64_t)g.inc.size());
    d->rec.upload(g.rec.data(), (int64_t)g.rec.size());
    d->posi.upload(g.posi.data(), (int64_t)g.posi.size());
    adj_ptr.upload(d->node_adj_ptr.data(), (int64_t)d->node_adj_ptr.size());
    adj.upload(d->node_adj.data(), (int64_t)d->node_adj.size());
  }
  require(d->max_degree <= 65535, "graph: node degree exceeds 65535");
  d->num_dofs = nn * goal::kNumEqs;
  d->nnz = d->node_adj_ptr[nn] * goal::kNumEqs * goal::kNumEqs;
  {
    // 16 * deg(a) CSR values per node; at least 2048 per run, and always one node
    d->gather_cap = std::max(2048, 16 * d->max_degree);
    const int nmax = goal::gather_nodes_per_range();
    std::vector<int32_t> r{0};
    int64_t cur = 0;
    for (int64_t a = 0; a < nn; ++a) {
      const int64_t v = 16 * (d->node_adj_ptr[a + 1] - d->node_adj_ptr[a]);
      if (a > r.back() && (cur + v > d->gather_cap || a - r.back() >= nmax)) {
        r.push_back((int32_t)a);
        cur = 0;
      }
      cur += v;
    }
    r.push_back((int32_t)nn);
    d->num_ranges = (int64_t)r.size() - 1;
    d->ranges.upload(r.data(), (int64_t)r.size());
    d->ranges_host = r;
  }
  d->row_ptr.alloc(d->num_dofs + 1);
  d->cols.alloc(d->nnz);
  HIPCHK(goal::launch_expand_csr(adj_ptr.p, adj.p, nn, d->row_ptr.p, d->cols.p, stream()));
  d->x.alloc(d->num_dofs);
  HIPCHK(goal::launch_fill(d->x.p, d->num_dofs, 0.0, stream()));
  if (d->node_adj.empty() && adj.n > 0) {  // keep the columns on the device for the multigrid set-up
    d->node_adj_dev.release();
    std::swap(d->node_adj_dev.p, adj.p);
    std::swap(d->node_adj_dev.n, adj.n);
  }
  sync();  // the graph buffers are released on return
}

// The node adjacency's column lists on the host (the multigrid aggregation
// runs there); copied from the device on first use.
const std::vector<int32_t>& host_node_adj(goal_disc d) {
  if (d->node_adj.empty() && d->node_adj_dev.n > 0) {
    d->node_adj.resize(d->node_adj_dev.n);
    d->node_adj_dev.download(d->node_adj.data(), d->node_adj_dev.n);
  }
  return d->node_adj;
}

goal_disc new_disc(goal::HostMesh&& m) {
  std::unique_ptr<goal_disc_s> d(new goal_disc_s);
  d->mesh = std::move(m);
  disc_build(d.get());
  return d.release();
}

const DevBuf<int32_t>& dev_node_set(goal_disc d, const std::string& name) {
  auto it = d->mesh.node_sets.find(name);
  if (it == d->mesh.node_sets.end()) throw std::invalid_argument("unknown node set: " + name);
  auto& slot = d->dev_node_sets[name];
  if (!slot) {
    slot.reset(new DevBuf<int32_t>);
    slot->upload(it->second.data(), (int64_t)it->second.size());
  }
  return *slot;
}

int nq_of(int quad_order) { return quad_order == 1 ? 1 : 4; }

int finish_assembly(goal_mech m) {
  if (!m->check) return GOAL_OK;
  int flags = 0;
  HIPCHK(hipMemcpyAsync(&flags, m->status.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
  sync();
  if (flags) {
    HIPCHK(hipMemsetAsync(m->status.p, 0, sizeof(int), stream()));
    if (flags & 1) return set_error(GOAL_ERR_J2_RETURN_MAP, "J2: return mapping failed");
  }
  return GOAL_OK;
}

// The kernels index the material table by element set: the disc's set count
// may have grown (goal_disc_set_elem_sets) since the mechanics was created.
void require_materials(goal_mech m) {
  require(m->num_sets >= m->d->mesh.num_elem_sets,
          "fewer materials than element sets (goal_disc_set_elem_sets after goal_mech_create)");
}

goal::AsmArgs asm_args(goal_mech m) {
  require_materials(m);
  goal::AsmArgs a{};
  goal_disc d = m->d;
  a.num_elems = d->mesh.num_elems;
  a.conn = d->conn.p;
  a.coords = d->coords.p;
  a.x = d->x.p;
  a.elem_set = d->elem_set.p;
  a.mats = m->mats.p;
  a.model = m->model;
  a.stabilization = m->stab;
  a.xcd_remap = m->xcd_remap;
  a.status = m->status.p;
  goal_states st = m->st;
  a.sigma = st->sigma.p;
  a.eqps = st->eqps.p;
  a.eqps_old = st->eqps_old.p;
  a.Fp = st->Fp.p;
  a.Fp_old = st->Fp_old.p;
  return a;
}

// 0 lets the launcher size a persistent grid from the kernel's occupancy.
int asm_blocks(goal_mech m) { return m->blocks; }


// Fill whatever zero_* left pending (before any access other than assembly).
void materialize(goal_solinfo s) {
  if (s->zero_R) HIPCHK(goal::launch_fill(s->R.p, s->R.n, 0.0, stream()));
  if (s->zero_dMdu) HIPCHK(goal::launch_fill(s->dMdu.p, s->dMdu.n, 0.0, stream()));
  if (s->zero_vals) HIPCHK(goal::launch_fill(s->vals.p, s->vals.n, 0.0, stream()));
  s->zero_R = s->zero_dMdu = s->zero_vals = false;
}

// Element pass + node gather.  want_vals: Jacobian assembly.
// Chunked assembly (neo-Hookean; option "overlap" = K chunks).  The element
// groups are split into K consecutive chunks on the library stream; after
// chunk c, the node runs whose incident elements all lie in chunks <= c are
// gathered on the side stream, while chunk c + 1 computes.  The element pass
// is issue-bound and the gather HBM-bound, so the two share the GPU; the
// element grid is held below the resident size to leave the gather room.
// Measured on the P2 64^3 cube with the vector-ALU element matrices: FP32-
// tangent Jacobian 11.97 -> 10.64 ms at 32 chunks, FP64 unchanged (its
// element pass already saturates HBM with the gather's share).  With the
// matrix-core element matrices one pass is faster for both (FP32 tangent
// 9.56 ms in one pass, 9.84 ms in 32 chunks), so the automatic setting is
// one pass.
// Same kernels and the same arithmetic as one pass: results are bitwise equal.
int chunks_of(goal_mech m, bool want_vals) {
  (void)want_vals;
  return m->overlap >= 0 ? m->overlap : 0;
}

bool chunked(goal_mech m, bool want_vals) {
  if (chunks_of(m, want_vals) < 2 || m->model != goal::MODEL_NEOHOOKEAN) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIPCHK(hipStreamIsCapturing(stream(), &cs));
  return cs == hipStreamCaptureStatusNone;  // graphs keep the one-stream order
}

void run_chunked(goal_mech m, goal_solinfo s, int kind, goal::AsmArgs a, bool want_vals) {
  goal_disc d = m->d;
  const int64_t ne = d->mesh.num_elems, nr = d->num_ranges;
  if (d->range_last_elem.empty()) {
    const int npe = d->mesh.nodes_per_elem;
    std::vector<int64_t> last(d->mesh.num_nodes, -1);
    for (int64_t e = 0; e < ne; ++e)
      for (int k = 0; k < npe; ++k) last[d->mesh.conn[e * npe + k]] = e;  // ascending e: the last write wins
    d->range_last_elem.assign(nr, -1);
    for (int64_t r = 0; r < nr; ++r)
      for (int32_t n = d->ranges_host[r]; n < d->ranges_host[r + 1]; ++n)
        d->range_last_elem[r] = std::max(d->range_last_elem[r], last[n]);
  }
  const int G = goal::element_group_size(d->mesh.order, m->quad_order, m->model);
  const int64_t ngroups = (ne + G - 1) / G;
  const int K = (int)std::min<int64_t>(chunks_of(m, want_vals), ngroups);
  int blocks = m->overlap_blocks;
  if (blocks <= 0) {
    const int full = goal::element_resident_grid(kind, d->mesh.order, m->quad_order, m->model, a.em32 != nullptr, ne);
    blocks = std::max(8, (full * 5 / 6) / 8 * 8);
  }
  // fills the gathers depend on go first, on the library stream
  if (s->zero_dMdu) {
    HIPCHK(goal::launch_fill(s->dMdu.p, s->dMdu.n, 0.0, stream()));
    s->zero_dMdu = false;
  }
  if (!want_vals && s->zero_vals) {
    HIPCHK(goal::launch_fill(s->vals.p, s->vals.n, 0.0, stream()));
    s->zero_vals = false;
  }
  goal::GatherArgs g{};
  g.cap = d->gather_cap;
  g.inc_ptr = d->inc_ptr.p;
  g.posi = d->posi.p;
  g.row_ptr = d->row_ptr.p;
  g.emr = a.emr;
  g.em = a.em;
  g.em32 = a.em32;
  g.R = s->R.p;
  g.vals = want_vals ? s->vals.p : nullptr;
  g.overwrite_R = s->zero_R ? 1 : 0;
  g.overwrite_vals = (want_vals && s->zero_vals) ? 1 : 0;
  int64_t r0 = 0;
  for (int c = 0; c < K; ++c) {
    a.group_lo = ngroups * c / K;
    a.group_hi = ngroups * (c + 1) / K;
    HIPCHK(goal::launch_element(kind, d->mesh.order, m->quad_order, a, blocks, stream()));
    const int64_t elem_end = std::min<int64_t>(ne, a.group_hi * G);
    int64_t r1 = r0;
    while (r1 < nr && d->range_last_elem[r1] < elem_end) ++r1;
    if (c == K - 1) r1 = nr;
    if (r1 > r0) {
      // the wait takes the event's latest record, so the events can cycle
      HIPCHK(hipEventRecord(ctx().chunk_ev[c % 64], stream()));
      HIPCHK(hipStreamWaitEvent(ctx().side, ctx().chunk_ev[c % 64], 0));
      g.num_ranges = r1 - r0;
      g.ranges = d->ranges.p + r0;
      HIPCHK(goal::launch_gather(d->mesh.order, g, goal::gather_grid(d->mesh.order, g.em32 != nullptr, d->gather_cap,
                                                                     g.num_ranges),
                                 ctx().side));
      r0 = r1;
    }
  }
  if (m->timing) HIPCHK(hipEventRecord(m->ev[1], stream()));
  HIPCHK(hipEventRecord(ctx().chunk_ev[64], ctx().side));
  HIPCHK(hipStreamWaitEvent(stream(), ctx().chunk_ev[64], 0));
  if (m->timing) {
    HIPCHK(hipEventRecord(m->ev[2], stream()));
    m->timed = true;
  }
  s->zero_R = false;
  if (want_vals) s->zero_vals = false;
}

// Element buffers between the two passes: the mechanics' own (allocated on
// first use), or, with emr_buf/em_buf given, the caller's.
// Doubles of J2 point states an assembly of this kind needs (0: none).
int64_t point_states(goal_mech m, int kind) {
  if (!goal::uses_point_pass(m->model, kind)) return 0;
  return goal::point_state_doubles(m->d->mesh.num_elems * m->st->nq, goal::point_stride(kind));
}

// With emr_buf given (and pt_buf for J2) the assembly also stays in one pass
// (the chunked path keeps a host-side table it builds on first use); the two
// give the same bits.
void run_assembly(goal_mech m, goal_solinfo s, int kind, goal::AsmArgs a, bool want_vals, double* emr_buf = nullptr,
                  double* em_buf = nullptr, double* pt_buf = nullptr) {
  const bool own_buffers = emr_buf == nullptr;
  goal_disc d = m->d;
  const int64_t nd = 4 * (int64_t)d->mesh.nodes_per_elem;
  if (!emr_buf) {
    if (m->emr.n < d->mesh.num_elems * nd) m->emr.alloc(d->mesh.num_elems * nd);
    if (want_vals && m->em.n < d->mesh.num_elems * nd * nd) m->em.alloc(d->mesh.num_elems * nd * nd);
    const int64_t np = point_states(m, kind);
    if (m->pt.n < np) m->pt.alloc(np);
    emr_buf = m->emr.p;
    em_buf = m->em.p;
    pt_buf = m->pt.p;
  }
  a.pt = pt_buf;
  a.rec = d->rec.p;
  a.emr = emr_buf;
  const bool f32 = want_vals && m->tangent == 32;
  a.em = (want_vals && !f32) ? em_buf : nullptr;
  a.em32 = f32 ? reinterpret_cast<float*>(em_buf) : nullptr;
  if (m->timing) HIPCHK(hipEventRecord(m->ev[0], stream()));
  if (own_buffers && chunked(m, want_vals)) {
    run_chunked(m, s, kind, a, want_vals);
    return;
  }
  HIPCHK(goal::launch_element(kind, d->mesh.order, m->quad_order, a, asm_blocks(m), stream()));
  if (m->timing) HIPCHK(hipEventRecord(m->ev[1], stream()));
  if (s->zero_dMdu) {
    HIPCHK(goal::launch_fill(s->dMdu.p, s->dMdu.n, 0.0, stream()));
    s->zero_dMdu = false;
  }
  if (!want_vals && s->zero_vals) {
    HIPCHK(goal::launch_fill(s->vals.p, s->vals.n, 0.0, stream()));
    s->zero_vals = false;
  }
  goal::GatherArgs g{};
  g.num_ranges = d->num_ranges;
  g.ranges = d->ranges.p;
  g.cap = d->gather_cap;
  g.inc_ptr = d->inc_ptr.p;
  g.posi = d->posi.p;
  g.row_ptr = d->row_ptr.p;
  g.emr = a.emr;
  g.em = a.em;
  g.em32 = a.em32;
  g.R = s->R.p;
  g.vals = want_vals ? s->vals.p : nullptr;
  g.overwrite_R = s->zero_R ? 1 : 0;
  g.overwrite_vals = (want_vals && s->zero_vals) ? 1 : 0;
  HIPCHK(goal::launch_gather(d->mesh.order, g,
                             goal::gather_grid(d->mesh.order, g.em32 != nullptr, d->gather_cap, d->num_ranges),
                             stream()));
  if (m->timing) {
    HIPCHK(hipEventRecord(m->ev[2], stream()));
    m->timed = true;
  }
  s->zero_R = false;
  if (want_vals) s->zero_vals = false;
}

// ------------------------------------------------- multigrid preconditioner
// Aggregation multigrid (one V-cycle per application) standing in for the
// MueLu preconditioner of goal's solve (src/goal_linear_solve.cpp:53-70).
//   - Levels: greedy aggregates of nodes (goal::amg_coarsen), the prolongator
//     P the aggregates' indicator on each of a node's 4 equations, coarse
//     matrices the Galerkin sums P^T A P, down to at most kAmgCoarseNodes nodes,
//     whose matrix is inverted densely.
//   - Smoother: Chebyshev of degree kAmgDegree on Binv A (Binv: inverse 4x4
//     diagonal blocks), before and after the coarse correction, on an ellipse
//     around the spectrum: real extent [hi / 10, hi], imaginary semi-axis
//     from the Ritz values of kAmgArnoldi Arnoldi steps on Binv A
//     (goal::chebyshev_ellipse).  The plastic J2 tangent makes Binv A
//     strongly non-normal (eigenvalues such as 0.46 +- 0.76i next to a
//     largest real one of 2.1), which Chebyshev on the real interval alone
//     amplifies.  A numpy model of this hierarchy: on a 446 K-tet graded
//     bar at a plastic stretch, GMRES stalled at a relative residual of 0.24
//     after 1,500 iterations with the interval and reached 1e-8 in 132 with
//     the ellipse; on the Newton-state 65 K-tet bar 75 and 80 iterations; on
//     the P2 cube's and plastic bar's Jacobians degree 4 on [lmax/10, ..]
//     took 57 and 69-92, block Jacobi alone 800+ without converging.
// Every kernel sums in a fixed order, so a solve is bitwise reproducible.
constexpr int kAmgCoarseNodes = 128;
constexpr int kAmgMaxLevels = 12;
constexpr int kAmgDegree = 4;
constexpr int kAmgArnoldi = 10;

void amg_build_structure(goal_disc d) {
  auto h = std::make_unique<AmgHierarchy>();
  const int64_t nn0 = d->mesh.num_nodes;
  std::vector<int64_t> fptr = d->node_adj_ptr;
  std::vector<int32_t> fcol = host_node_adj(d);
  int64_t nf = nn0;
  {  // row node of every finest-level block
    std::vector<int32_t> brow(fptr[nf]);
    for (int64_t a = 0; a < nf; ++a)
      for (int64_t k = fptr[a]; k < fptr[a + 1]; ++k) brow[k] = (int32_t)a;
    h->brow0.upload(brow.data(), (int64_t)brow.size());
    sync();
  }
  for (int l = 0;; ++l) {
    auto L = std::make_unique<AmgLevel>();
    L->nn = nf;
    if (l > 0) {
      L->bptr.upload(fptr.data(), (int64_t)fptr.size());
      L->bcol.upload(fcol.data(), (int64_t)fcol.size());
      L->bval.alloc(16 * fptr[nf]);
    }
    L->binv.alloc(16 * nf);
    L->b.alloc(4 * nf);
    L->x.alloc(4 * nf);
    L->r.alloc(4 * nf);
    L->d.alloc(4 * nf);
    const bool last = nf <= kAmgCoarseNodes || l + 1 >= kAmgMaxLevels;
    goal::AmgCoarsening c;
    if (!last) c = goal::amg_coarsen(nf, fptr.data(), fcol.data());
    if (last || c.nc >= nf) {  // coarsest: keep its graph for the dense inverse
      L->bptr_h = fptr;
      L->bcol_h = fcol;
      h->lv.push_back(std::move(L));
      sync();
      break;
    }
    L->nc = c.nc;
    L->ncb = (int64_t)c.ccol.size();
    L->agg.upload(c.agg.data(), (int64_t)c.agg.size());
    L->aptr.upload(c.aptr.data(), (int64_t)c.aptr.size());
    L->amem.upload(c.amem.data(), (int64_t)c.amem.size());
    L->gptr.upload(c.gptr.data(), (int64_t)c.gptr.size());
    L->gidx.upload(c.gidx.data(), (int64_t)c.gidx.size());
    sync();  // the host vectors go out of scope
    h->lv.push_back(std::move(L));
    fptr.swap(c.cptr);
    fcol.swap(c.ccol);
    nf = c.nc;
  }
  const int64_t nc = h->lv.back()->nn;
  h->coarse_n = (int)(4 * nc);
  h->arn.alloc((int64_t)(kAmgArnoldi + 1) * 4 * nn0);
  h->coarse_inv.alloc((int64_t)h->coarse_n * h->coarse_n);
  d->amg = std::move(h);
}

// y = b - A_l x, or y = A_l x without b (finest level: the solve's CSR).
void amg_residual(goal_solinfo s, AmgHierarchy& h, int l, const double* x, const double* b, double* y) {
  AmgLevel& L = *h.lv[l];
  goal_disc d = s->d;
  if (l == 0) {
    if (b) HIPCHK(goal::launch_spmv_resid(4 * L.nn, d->row_ptr.p, d->cols.p, s->vals.p, x, b, y, stream()));
    else HIPCHK(goal::launch_spmv(4 * L.nn, d->row_ptr.p, d->cols.p, s->vals.p, x, y, stream()));
  } else {
    HIPCHK(goal::launch_bsr_spmv(L.nn, L.bptr.p, L.bcol.p, L.bval.p, x, b, y, stream()));
  }
}

// Values of the hierarchy for the solve's current matrix.
void amg_setup(goal_solinfo s, int* status) {
  goal_disc d = s->d;
  if (!d->amg) amg_build_structure(d);
  AmgHierarchy& h = *d->amg;
  const int nl = (int)h.lv.size();
  for (int l = 0; l < nl; ++l) {
    AmgLevel& L = *h.lv[l];
    if (l == 0) {
      HIPCHK(goal::launch_block_inverse(L.nn, d->row_ptr.p, d->cols.p, s->vals.p, L.binv.p, status, stream()));
    } else {
      AmgLevel& F = *h.lv[l - 1];
      HIPCHK(goal::launch_galerkin(F.ncb, F.gptr.p, F.gidx.p, l == 1 ? h.brow0.p : nullptr, d->row_ptr.p,
                                   l == 1 ? s->vals.p : F.bval.p, L.bval.p, stream()));
      HIPCHK(goal::launch_bsr_block_inverse(L.nn, L.bptr.p, L.bcol.p, L.bval.p, L.binv.p, status, stream()));
    }
    if (l + 1 == nl) break;
    // Arnoldi on Binv A from a fixed start vector (classical Gram-Schmidt,
    // two passes); its Ritz values place the smoother's ellipse
    const int64_t nv = 4 * L.nn;
    double* Q = h.arn.p;
    std::vector<double> v0(nv);
    for (int64_t i = 0; i < nv; ++i)
      v0[i] = (double)(goal::splitmix64(0xA3C5u + (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    HIPCHK(hipMemcpyAsync(Q, v0.data(), sizeof(double) * nv, hipMemcpyHostToDevice, stream()));
    auto norm2 = [&](const double* v) {  // waits
      HIPCHK(goal::launch_reduce(1, v, nullptr, nv, reduce_scratch(), result_slot(3), stream()));
      double r2 = 0.0;
      HIPCHK(hipMemcpyAsync(&r2, result_slot(3), sizeof(double), hipMemcpyDeviceToHost, stream()));
      sync();
      return std::sqrt(r2);
    };
    const double q0 = norm2(Q);
    HIPCHK(goal::launch_scale(nv, 1.0 / q0, Q, Q, stream()));
    constexpr int M = kAmgArnoldi;
    std::vector<double> H((size_t)(M + 1) * M, 0.0), hc(M + 1);
    int steps = 0;
    for (int j = 0; j < M; ++j) {
      double* w = Q + (int64_t)(j + 1) * nv;
      amg_residual(s, h, l, Q + (int64_t)j * nv, nullptr, L.r.p);
      HIPCHK(goal::launch_block_apply(L.nn, L.binv.p, L.r.p, w, stream()));
      for (int pass = 0; pass < 2; ++pass) {
        HIPCHK(goal::launch_multi_dot(nv, j + 1, Q, nv, w, s->partial.p, s->h1.p, stream()));
        HIPCHK(goal::launch_multi_axpy(nv, j + 1, Q, nv, s->h1.p, -1.0, w, stream()));
        HIPCHK(hipMemcpyAsync(hc.data(), s->h1.p, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, stream()));
        sync();
        for (int i = 0; i <= j; ++i) H[(size_t)i * M + j] += hc[i];
      }
      const double hn = norm2(w);
      H[(size_t)(j + 1) * M + j] = hn;
      steps = j + 1;
      if (!(hn > 1e-12 * std::fabs(H[(size_t)j * M + j])) || !std::isfinite(hn)) break;  // invariant subspace
      HIPCHK(goal::launch_scale(nv, 1.0 / hn, w, w, stream()));
    }
    std::vector<double> Hs((size_t)steps * steps);
    for (int i = 0; i < steps; ++i)
      for (int j = 0; j < steps; ++j) Hs[(size_t)i * steps + j] = H[(size_t)i * M + j];
    goal::chebyshev_ellipse(goal::hessenberg_eigenvalues(steps, Hs), &L.theta, &L.c2);
    if (!(L.theta > 0.0)) {  // no Ritz value with a positive real part: plain block Jacobi steps
      L.theta = 1.0;
      L.c2 = 0.0;
    }
  }
  // the coarsest level's matrix, densely inverted on the host
  AmgLevel& C = *h.lv[nl - 1];
  const int n = h.coarse_n;
  std::vector<double> dense((size_t)n * n, 0.0);
  if (nl == 1) {  // the whole problem is the coarsest level: its CSR
    std::vector<double> vals(s->vals.n);
    std::vector<int64_t> rp(d->row_ptr.n);
    std::vector<int32_t> cols(d->cols.n);
    s->vals.download(vals.data(), s->vals.n);
    d->row_ptr.download(rp.data(), d->row_ptr.n);
    d->cols.download(cols.data(), d->cols.n);
    for (int r = 0; r < n; ++r)
      for (int64_t k = rp[r]; k < rp[r + 1]; ++k) dense[(size_t)r * n + cols[k]] = vals[k];
  } else {
    std::vector<double> bv(C.bval.n);
    C.bval.download(bv.data(), C.bval.n);
    for (int64_t a = 0; a < C.nn; ++a)
      for (int64_t k = C.bptr_h[a]; k < C.bptr_h[a + 1]; ++k)
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j)
            dense[(size_t)(4 * a + i) * n + 4 * (int64_t)C.bcol_h[k] + j] = bv[16 * k + 4 * i + j];
  }
  if (!goal::dense_inverse(n, dense)) {
    int one = 1;
    HIPCHK(hipMemcpyAsync(status, &one, sizeof(int), hipMemcpyHostToDevice, stream()));
  }
  HIPCHK(hipMemcpyAsync(h.coarse_inv.p, dense.data(), sizeof(double) * dense.size(), hipMemcpyHostToDevice, stream()));
  sync();
}

// x = Chebyshev smoothing of A_l x = b (from x = 0 when zero).
void amg_smooth(goal_solinfo s, AmgHierarchy& h, int l, const double* b, double* x, bool zero) {
  AmgLevel& L = *h.lv[l];
  // Chebyshev iteration for the ellipse with centre theta and foci
  // theta +- sqrt(c2): with kappa_0 = 1 / theta and kappa_{k+1} =
  // 1 / (2 theta - c2 kappa_k), d_0 = kappa_0 Binv r_0 and d_{k+1} =
  // c2 kappa_{k+1} kappa_k d_k + 2 kappa_{k+1} Binv r_{k+1}, all real also
  // when the foci are imaginary (c2 < 0)
  const double th = L.theta, c2 = L.c2;
  double k0 = 1.0 / th;
  const double* r = b;
  if (zero) {
    HIPCHK(goal::launch_fill(x, 4 * L.nn, 0.0, stream()));
  } else {
    amg_residual(s, h, l, x, b, L.r.p);
    r = L.r.p;
  }
  HIPCHK(goal::launch_cheb(L.nn, L.binv.p, r, 0.0, k0, L.d.p, x, stream()));
  for (int k = 1; k < kAmgDegree; ++k) {
    const double k1 = 1.0 / (2.0 * th - c2 * k0);
    amg_residual(s, h, l, x, b, L.r.p);
    HIPCHK(goal::launch_cheb(L.nn, L.binv.p, L.r.p, c2 * k1 * k0, 2.0 * k1, L.d.p, x, stream()));
    k0 = k1;
  }
}

// x = V-cycle applied to b on level l.
void amg_vcycle(goal_solinfo s, AmgHierarchy& h, int l, const double* b, double* x) {
  AmgLevel& L = *h.lv[l];
  if (l + 1 == (int)h.lv.size()) {
    HIPCHK(goal::launch_dense_matvec(h.coarse_n, h.coarse_inv.p, b, x, stream()));
    return;
  }
  AmgLevel& C = *h.lv[l + 1];
  amg_smooth(s, h, l, b, x, true);
  amg_residual(s, h, l, x, b, L.r.p);
  HIPCHK(goal::launch_restrict(L.nc, L.aptr.p, L.amem.p, L.r.p, C.b.p, stream()));
  amg_vcycle(s, h, l + 1, C.b.p, C.x.p);
  HIPCHK(goal::launch_prolong(L.nn, L.agg.p, C.x.p, x, stream()));
  amg_smooth(s, h, l, b, x, false);
}

}  // namespace

extern "C" {

const char* goal_last_error(void) { return g_err.c_str(); }
const char* goal_version(void) { return "goal_amd 0.1 (gfx950)"; }

int goal_device_alloc(int64_t bytes, void** ptr) {
  return guard([&]() -> int {
    require(ptr && bytes >= 0, "goal_device_alloc: bad arguments");
    ctx();
    *ptr = nullptr;
    if (bytes > 0) HIPCHK(hipMalloc(ptr, bytes));
    return GOAL_OK;
  });
}
int goal_device_free(void* p) {
  return guard([&]() -> int {
    if (p) HIPCHK(hipFree(p));
    return GOAL_OK;
  });
}
int goal_memcpy_h2d(void* dst, const void* src, int64_t bytes) {
  return guard([&]() -> int {
    if (bytes > 0) {
      HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
      sync();
    }
    return GOAL_OK;
  });
}
int goal_memcpy_d2h(void* dst, const void* src, int64_t bytes) {
  return guard([&]() -> int {
    if (bytes > 0) {
      HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
      sync();
    }
    return GOAL_OK;
  });
}
int goal_sync(void) {
  return guard([&]() -> int {
    sync();
    return GOAL_OK;
  });
}
int goal_event_record(int slot) {
  return guard([&]() -> int {
    require(slot >= 0 && slot < 256, "event slot out of range");
    HIPCHK(hipEventRecord(ctx().events[slot], stream()));
    return GOAL_OK;
  });
}
int goal_event_elapsed(int a, int b, float* ms) {
  return guard([&]() -> int {
    require(a >= 0 && a < 256 && b >= 0 && b < 256 && ms, "event slot out of range");
    HIPCHK(hipEventSynchronize(ctx().events[b]));
    HIPCHK(hipEventElapsedTime(ms, ctx().events[a], ctx().events[b]));
    return GOAL_OK;
  });
}
int goal_device_copy(double* dst, const double* src, int64_t n) {
  return guard([&]() -> int {
    require(dst && src && n >= 0, "goal_device_copy: bad arguments");
    HIPCHK(goal::launch_copy(dst, src, n, stream()));
    return GOAL_OK;
  });
}
int goal_device_scale(double* y, const double* x, int64_t n, double a) {
  return guard([&]() -> int {
    require(y && x && n >= 0, "goal_device_scale: bad arguments");
    HIPCHK(goal::launch_scale(n, a, x, y, stream()));
    return GOAL_OK;
  });
}
int goal_device_fill(double* p, int64_t n, double value) {
  return guard([&]() -> int {
    require(p && n >= 0, "goal_device_fill: bad arguments");
    HIPCHK(goal::launch_fill(p, n, value, stream()));
    return GOAL_OK;
  });
}
int goal_device_info(char* name, int len, int* num_cus) {
  return guard([&]() -> int {
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, ctx().device));
    if (name && len > 0) {
      std::snprintf(name, (size_t)len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (num_cus) *num_cus = p.multiProcessorCount;
    return GOAL_OK;
  });
}

// ------------------------------------------------------------- host mesh
int goal_mesh_box(int nx, int ny, int nz, double lx, double ly, double lz, int order, double jitter, uint64_t seed,
                  int64_t* sizes, double* coords, int32_t* conn) {
  return guard([&]() -> int {
    require(sizes != nullptr, "goal_mesh_box: null sizes");
    goal::HostMesh m = goal::make_box_mesh(nx, ny, nz, lx, ly, lz, order, jitter, seed);
    sizes[0] = m.num_nodes;
    sizes[1] = m.num_elems;
    if (coords) std::memcpy(coords, m.coords.data(), sizeof(double) * m.coords.size());
    if (conn) std::memcpy(conn, m.conn.data(), sizeof(int32_t) * m.conn.size());
    return GOAL_OK;
  });
}

int goal_mesh_graph(int order, int64_t num_nodes, int64_t num_elems, const int32_t* conn, int64_t* nnz,
                    int64_t* row_ptr, int32_t* cols) {
  return guard([&]() -> int {
    require(conn && nnz && num_nodes > 0 && num_elems > 0, "goal_mesh_graph: bad arguments");
    goal::HostMesh m;
    m.order = order;
    m.nodes_per_elem = order == 1 ? 4 : 10;
    m.num_nodes = num_nodes;
    m.num_elems = num_elems;
    m.coords.assign(num_nodes * 3, 0.0);
    m.conn.assign(conn, conn + num_elems * m.nodes_per_elem);
    goal::finalize_mesh(m);
    goal::HostGraph g = goal::build_graph(m, host_threads());
    *nnz = g.nnz;
    if (row_ptr && cols) {
      std::vector<int64_t> rp;
      std::vector<int32_t> cl;
      goal::expand_csr(m, g, rp, cl);
      std::memcpy(row_ptr, rp.data(), sizeof(int64_t) * rp.size());
      std::memcpy(cols, cl.data(), sizeof(int32_t) * cl.size());
    }
    return GOAL_OK;
  });
}

int goal_mesh_refine(int64_t num_nodes, const double* coords, int64_t num_elems, const int32_t* conn,
                     int64_t* sizes, double* fine_coords, int32_t* fine_conn, int32_t* parent,
                     int32_t* edge_parents) {
  return guard([&]() -> int {
    require(coords && conn && sizes && num_nodes > 0 && num_elems > 0, "goal_mesh_refine: bad arguments");
    goal::HostMesh m;
    m.order = 1;
    m.nodes_per_elem = 4;
    m.num_nodes = num_nodes;
    m.num_elems = num_elems;
    m.coords.assign(coords, coords + num_nodes * 3);
    m.conn.assign(conn, conn + num_elems * 4);
    goal::finalize_mesh(m);
    goal::Refinement r = goal::refine_uniform(m);
    sizes[0] = r.fine.num_nodes;
    sizes[1] = r.fine.num_elems;
    sizes[2] = (int64_t)r.edge_parents.size() / 2;
    if (fine_coords) std::memcpy(fine_coords, r.fine.coords.data(), sizeof(double) * r.fine.coords.size());
    if (fine_conn) std::memcpy(fine_conn, r.fine.conn.data(), sizeof(int32_t) * r.fine.conn.size());
    if (parent) std::memcpy(parent, r.parent.data(), sizeof(int32_t) * r.parent.size());
    if (edge_parents) std::memcpy(edge_parents, r.edge_parents.data(), sizeof(int32_t) * r.edge_parents.size());
    return GOAL_OK;
  });
}

// ----------------------------------------------------------------- adapt
struct goal_adapt_s {
  goal::Adaptation a;
};

extern "C++" {
namespace {
goal::HostMesh linear_mesh(int64_t num_nodes, const double* coords, int64_t num_elems, const int32_t* conn) {
  goal::HostMesh m;
  m.order = 1;
  m.nodes_per_elem = 4;
  m.num_nodes = num_nodes;
  m.num_elems = num_elems;
  m.coords.assign(coords, coords + num_nodes * 3);
  m.conn.assign(conn, conn + num_elems * 4);
  goal::finalize_mesh(m);
  return m;
}
}  // namespace
}  // extern "C++"

int goal_size_field(int64_t num_nodes, const double* coords, int64_t num_elems, const int32_t* conn,
                    const double* eta, int64_t target, int p_order, double* vtx_size) {
  return guard([&]() -> int {
    require(coords && conn && eta && vtx_size && num_nodes > 0 && num_elems > 0 && p_order >= 1,
            "goal_size_field: bad arguments");
    const goal::HostMesh m = linear_mesh(num_nodes, coords, num_elems, conn);
    const std::vector<double> s = goal::iso_target_size(m, eta, target, p_order);
    std::memcpy(vtx_size, s.data(), sizeof(double) * s.size());
    return GOAL_OK;
  });
}

int goal_adapt_create(int64_t num_nodes, const double* coords, int64_t num_elems, const int32_t* conn,
                      const double* vtx_size, double split_ratio, int max_passes, goal_adapt* out) {
  return guard([&]() -> int {
    require(out && coords && conn && vtx_size && num_nodes > 0 && num_elems > 0 && max_passes >= 0,
            "goal_adapt_create: bad arguments");
    const goal::HostMesh m = linear_mesh(num_nodes, coords, num_elems, conn);
    std::vector<double> size(vtx_size, vtx_size + num_nodes);
    for (double h : size) require(h > 0.0 && std::isfinite(h), "goal_adapt_create: sizes must be positive");
    auto* a = new goal_adapt_s;
    try {
      a->a = goal::adapt_refine(m, size, split_ratio, max_passes);
      require((int64_t)a->a.mesh.num_nodes <= (int64_t)INT32_MAX / 4, "goal_adapt_create: mesh too large");
    } catch (...) {
      delete a;
      throw;
    }
    *out = a;
    return GOAL_OK;
  });
}

int goal_adapt_destroy(goal_adapt a) {
  delete a;
  return GOAL_OK;
}

int goal_adapt_sizes(goal_adapt a, int64_t* sizes) {
  return guard([&]() -> int {
    require(a && sizes, "goal_adapt_sizes: bad arguments");
    sizes[0] = a->a.mesh.num_nodes;
    sizes[1] = a->a.mesh.num_elems;
    sizes[2] = (int64_t)a->a.new_vertex_parents.size() / 2;
    sizes[3] = a->a.splits;
    sizes[4] = a->a.passes;
    return GOAL_OK;
  });
}

int goal_adapt_get(goal_adapt a, double* coords, int32_t* conn, int32_t* parent, int32_t* new_vertex_parents) {
  return guard([&]() -> int {
    require(a != nullptr, "goal_adapt_get: bad arguments");
    const goal::Adaptation& r = a->a;
    if (coords) std::memcpy(coords, r.mesh.coords.data(), sizeof(double) * r.mesh.coords.size());
    if (conn) std::memcpy(conn, r.mesh.conn.data(), sizeof(int32_t) * r.mesh.conn.size());
    if (parent) std::memcpy(parent, r.parent.data(), sizeof(int32_t) * r.parent.size());
    if (new_vertex_parents)
      std::memcpy(new_vertex_parents, r.new_vertex_parents.data(), sizeof(int32_t) * r.new_vertex_parents.size());
    return GOAL_OK;
  });
}

int goal_adapt_transfer(goal_adapt a, int ncomp, const double* old_field, double* new_field) {
  return guard([&]() -> int {
    require(a && old_field && new_field && ncomp > 0, "goal_adapt_transfer: bad arguments");
    const goal::Adaptation& r = a->a;
    const int64_t nold = r.num_old_vertices, nnew = (int64_t)r.new_vertex_parents.size() / 2;
    std::memcpy(new_field, old_field, sizeof(double) * nold * ncomp);
    for (int64_t k = 0; k < nnew; ++k) {
      const int32_t p0 = r.new_vertex_parents[2 * k], p1 = r.new_vertex_parents[2 * k + 1];
      for (int c = 0; c < ncomp; ++c)
        new_field[(nold + k) * ncomp + c] = 0.5 * (new_field[(int64_t)p0 * ncomp + c] + new_field[(int64_t)p1 * ncomp + c]);
    }
    return GOAL_OK;
  });
}

// ------------------------------------------------------------------ disc
int goal_disc_create(int order, int64_t num_nodes, const double* coords, int64_t num_elems, const int32_t* conn,
                     const int32_t* elem_set, goal_disc* out) {
  return guard([&]() -> int {
    require(out && coords && conn && num_nodes > 0 && num_elems > 0, "goal_disc_create: bad arguments");
    ctx();
    goal::HostMesh m;
    m.order = order;
    m.nodes_per_elem = order == 1 ? 4 : 10;
    m.num_nodes = num_nodes;
    m.num_elems = num_elems;
    m.coords.assign(coords, coords + num_nodes * 3);
    m.conn.assign(conn, conn + num_elems * m.nodes_per_elem);
    if (elem_set) m.elem_set.assign(elem_set, elem_set + num_elems);
    goal::finalize_mesh(m);
    *out = new_disc(std::move(m));
    return GOAL_OK;
  });
}

int goal_disc_create_box(int nx, int ny, int nz, double lx, double ly, double lz, int order, double jitter,
                         uint64_t seed, goal_disc* out) {
  return guard([&]() -> int {
    require(out != nullptr, "goal_disc_create_box: null output");
    ctx();
    *out = new_disc(goal::make_box_mesh(nx, ny, nz, lx, ly, lz, order, jitter, seed));
    return GOAL_OK;
  });
}

// ------------------------------------------------------------ mesh files
struct goal_mesh_file_s {
  goal::MeshFile f;
};

int goal_mesh_file_read(const char* geom_file, const char* mesh_file, const char* assoc_file, int order,
                        goal_mesh_file* out) {
  return guard([&]() -> int {
    require(geom_file && mesh_file && assoc_file && out, "goal_mesh_file_read: bad arguments");
    std::unique_ptr<goal_mesh_file_s> f(new goal_mesh_file_s);
    f->f = goal::read_mesh_files(geom_file, mesh_file, assoc_file, order);
    *out = f.release();
    return GOAL_OK;
  });
}

int goal_mesh_file_destroy(goal_mesh_file f) {
  delete f;
  return GOAL_OK;
}

int goal_mesh_file_sizes(goal_mesh_file f, int64_t* num_nodes, int64_t* num_elems, int* num_elem_sets) {
  return guard([&]() -> int {
    require(f != nullptr, "null mesh file");
    if (num_nodes) *num_nodes = f->f.mesh.num_nodes;
    if (num_elems) *num_elems = f->f.mesh.num_elems;
    if (num_elem_sets) *num_elem_sets = (int)f->f.elem_set_names.size();
    return GOAL_OK;
  });
}

int goal_mesh_file_get_mesh(goal_mesh_file f, double* coords, int32_t* conn, int32_t* elem_set) {
  return guard([&]() -> int {
    require(f != nullptr, "null mesh file");
    const goal::HostMesh& m = f->f.mesh;
    if (coords) std::copy(m.coords.begin(), m.coords.end(), coords);
    if (conn) std::copy(m.conn.begin(), m.conn.end(), conn);
    if (elem_set) std::copy(m.elem_set.begin(), m.elem_set.end(), elem_set);
    return GOAL_OK;
  });
}

int goal_mesh_file_elem_set_name(goal_mesh_file f, int index, char* buf, int64_t len) {
  return guard([&]() -> int {
    require(f && buf && len > 0, "goal_mesh_file_elem_set_name: bad arguments");
    require(index >= 0 && index < (int)f->f.elem_set_names.size(), "element set index out of range");
    const std::string& n = f->f.elem_set_names[index];
    require((int64_t)n.size() < len, "name buffer too small");
    std::memcpy(buf, n.c_str(), n.size() + 1);
    return GOAL_OK;
  });
}

int goal_mesh_file_node_set(goal_mesh_file f, const char* name, int64_t* n, int32_t* nodes) {
  return guard([&]() -> int {
    require(f && name && n, "goal_mesh_file_node_set: bad arguments");
    auto it = f->f.mesh.node_sets.find(name);
    if (it == f->f.mesh.node_sets.end()) return set_error(GOAL_ERR_INVALID, "unknown node set: %s", name);
    *n = (int64_t)it->second.size();
    if (nodes) std::copy(it->second.begin(), it->second.end(), nodes);
    return GOAL_OK;
  });
}

int goal_mesh_file_side_set(goal_mesh_file f, const char* name, int64_t* n, int32_t* elems, int32_t* faces) {
  return guard([&]() -> int {
    require(f && name && n, "goal_mesh_file_side_set: bad arguments");
    auto it = f->f.side_sets.find(name);
    if (it == f->f.side_sets.end()) return set_error(GOAL_ERR_INVALID, "unknown side set: %s", name);
    *n = (int64_t)it->second.size();
    for (size_t k = 0; k < it->second.size(); ++k) {
      if (elems) elems[k] = it->second[k] / 4;
      if (faces) faces[k] = it->second[k] % 4;
    }
    return GOAL_OK;
  });
}

int goal_disc_create_from_file(goal_mesh_file f, goal_disc* out) {
  return guard([&]() -> int {
    require(f && out, "goal_disc_create_from_file: bad arguments");
    ctx();
    goal::HostMesh m = f->f.mesh;
    goal_disc d = new_disc(std::move(m));
    d->side_sets = f->f.side_sets;
    *out = d;
    return GOAL_OK;
  });
}

int goal_disc_destroy(goal_disc d) {
  return guard([&]() -> int {
    if (d) {
      sync();
      delete d;
    }
    return GOAL_OK;
  });
}

int goal_disc_sizes(goal_disc d, int64_t* out) {
  return guard([&]() -> int {
    require(d && out, "goal_disc_sizes: bad arguments");
    out[0] = d->mesh.num_nodes;
    out[1] = d->mesh.num_elems;
    out[2] = d->num_dofs;
    out[3] = d->nnz;
    out[4] = d->mesh.nodes_per_elem;
    out[5] = (int64_t)d->mesh.vertex_nodes.size();
    out[6] = d->mesh.order;
    out[7] = d->mesh.num_elem_sets;
    out[8] = d->max_degree;
    return GOAL_OK;
  });
}

int goal_disc_get_mesh(goal_disc d, double* coords, int32_t* conn, int32_t* elem_set) {
  return guard([&]() -> int {
    require(d != nullptr, "null disc");
    if (coords) std::memcpy(coords, d->mesh.coords.data(), sizeof(double) * d->mesh.coords.size());
    if (conn) std::memcpy(conn, d->mesh.conn.data(), sizeof(int32_t) * d->mesh.conn.size());
    if (elem_set) std::memcpy(elem_set, d->mesh.elem_set.data(), sizeof(int32_t) * d->mesh.elem_set.size());
    return GOAL_OK;
  });
}

int goal_disc_get_graph(goal_disc d, int64_t* row_ptr, int32_t* cols) {
  return guard([&]() -> int {
    require(d != nullptr, "null disc");
    if (row_ptr) d->row_ptr.download(row_ptr, d->num_dofs + 1);
    if (cols) d->cols.download(cols, d->nnz);
    return GOAL_OK;
  });
}

int goal_disc_graph_dev(goal_disc d, int64_t** row_ptr, int32_t** cols) {
  return guard([&]() -> int {
    require(d != nullptr, "null disc");
    if (row_ptr) *row_ptr = d->row_ptr.p;
    if (cols) *cols = d->cols.p;
    return GOAL_OK;
  });
}

int goal_disc_mesh_dev(goal_disc d, double** coords, int32_t** conn, uint16_t** pos) {
  return guard([&]() -> int {
    require(d != nullptr, "null disc");
    if (coords) *coords = d->coords.p;
    if (conn) *conn = d->conn.p;
    if (pos) *pos = d->pos.p;
    return GOAL_OK;
  });
}

int goal_disc_get_lids(goal_disc d, int32_t* lids) {
  return guard([&]() -> int {
    require(d && lids, "goal_disc_get_lids: bad arguments");
    const int npe = d->mesh.nodes_per_elem;
    for (int64_t e = 0; e < d->mesh.num_elems; ++e)
      for (int n = 0; n < npe; ++n)
        for (int eq = 0; eq < goal::kNumEqs; ++eq)
          lids[(e * npe + n) * goal::kNumEqs + eq] = d->mesh.conn[e * npe + n] * goal::kNumEqs + eq;
    return GOAL_OK;
  });
}

int goal_disc_add_node_set(goal_disc d, const char* name, int64_t n, const int32_t* nodes) {
  return guard([&]() -> int {
    require(d && name && n >= 0 && (n == 0 || nodes), "goal_disc_add_node_set: bad arguments");
    std::vector<int32_t> v(nodes, nodes + n);
    for (int32_t x : v) require(x >= 0 && x < d->mesh.num_nodes, "node set entry out of range");
    d->mesh.node_sets[name] = std::move(v);
    d->dev_node_sets.erase(name);
    d->derived_side_sets.erase(name);
    return GOAL_OK;
  });
}

int goal_disc_node_set_size(goal_disc d, const char* name, int64_t* n) {
  return guard([&]() -> int {
    require(d && name && n, "bad arguments");
    auto it = d->mesh.node_sets.find(name);
    if (it == d->mesh.node_sets.end()) return set_error(GOAL_ERR_INVALID, "unknown node set: %s", name);
    *n = (int64_t)it->second.size();
    return GOAL_OK;
  });
}

int goal_disc_get_node_set(goal_disc d, const char* name, int32_t* nodes) {
  return guard([&]() -> int {
    require(d && name && nodes, "bad arguments");
    auto it = d->mesh.node_sets.find(name);
    if (it == d->mesh.node_sets.end()) return set_error(GOAL_ERR_INVALID, "unknown node set: %s", name);
    std::copy(it->second.begin(), it->second.end(), nodes);
    return GOAL_OK;
  });
}

// ------------------------------------------------------------ side sets
extern "C++" {
namespace {

// Local vertices of tet face f, in apf's order.
constexpr int kTetFace[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};

int tet_edge_between(int a, int b) {
  for (int k = 0; k < 6; ++k)
    if ((goal::kTetEdge[k][0] == a && goal::kTetEdge[k][1] == b) ||
        (goal::kTetEdge[k][0] == b && goal::kTetEdge[k][1] == a))
      return k;
  return -1;
}

// The side set `name`: a given one, else the faces on the mesh boundary whose
// three vertices all lie in the node set of that name (for a box face node
// set: that face's triangles, the side set goal's assoc file names for it).
const std::vector<int32_t>& side_set(goal_disc d, const char* name) {
  auto given = d->side_sets.find(name);
  if (given != d->side_sets.end()) return given->second;
  auto derived = d->derived_side_sets.find(name);
  if (derived != d->derived_side_sets.end()) return derived->second;
  auto ns = d->mesh.node_sets.find(name);
  if (ns == d->mesh.node_sets.end())
    throw std::invalid_argument(std::string("unknown side set (and no node set of that name): ") + name);
  std::vector<char> in(d->mesh.num_nodes, 0);
  for (int32_t n : ns->second) in[n] = 1;
  const int npe = d->mesh.nodes_per_elem;
  std::map<std::array<int32_t, 3>, std::pair<int32_t, int>> seen;  // sorted vertices -> (side, count)
  for (int64_t e = 0; e < d->mesh.num_elems; ++e) {
    const int32_t* c = &d->mesh.conn[e * npe];
    for (int f = 0; f < 4; ++f) {
      std::array<int32_t, 3> v = {c[kTetFace[f][0]], c[kTetFace[f][1]], c[kTetFace[f][2]]};
      if (!in[v[0]] || !in[v[1]] || !in[v[2]]) continue;
      std::sort(v.begin(), v.end());
      auto it = seen.find(v);
      if (it == seen.end())
        seen.emplace(v, std::make_pair((int32_t)(4 * e + f), 1));
      else
        ++it->second.second;  // shared by two elements: an interior face
    }
  }
  std::vector<int32_t> sides;
  for (auto const& kv : seen)
    if (kv.second.second == 1) sides.push_back(kv.second.first);
  std::sort(sides.begin(), sides.end());
  return d->derived_side_sets[name] = std::move(sides);
}

}  // namespace
}  // extern "C++"

int goal_disc_add_side_set(goal_disc d, const char* name, int64_t n, const int32_t* elems, const int32_t* faces) {
  return guard([&]() -> int {
    require(d && name && n >= 0 && (n == 0 || (elems && faces)), "goal_disc_add_side_set: bad arguments");
    std::vector<int32_t> v((size_t)n);
    for (int64_t k = 0; k < n; ++k) {
      require(elems[k] >= 0 && elems[k] < d->mesh.num_elems, "side set element out of range");
      require(faces[k] >= 0 && faces[k] < 4, "side set local face must be 0..3");
      v[k] = 4 * elems[k] + faces[k];
    }
    d->side_sets[name] = std::move(v);
    return GOAL_OK;
  });
}

int goal_disc_side_set_size(goal_disc d, const char* name, int64_t* n) {
  return guard([&]() -> int {
    require(d && name && n, "goal_disc_side_set_size: bad arguments");
    *n = (int64_t)side_set(d, name).size();
    return GOAL_OK;
  });
}

int goal_disc_get_side_set(goal_disc d, const char* name, int32_t* elems, int32_t* faces) {
  return guard([&]() -> int {
    require(d && name && elems && faces, "goal_disc_get_side_set: bad arguments");
    const std::vector<int32_t>& v = side_set(d, name);
    for (size_t k = 0; k < v.size(); ++k) {
      elems[k] = v[k] / 4;
      faces[k] = v[k] % 4;
    }
    return GOAL_OK;
  });
}

int goal_disc_set_elem_sets(goal_disc d, int num_sets, const int32_t* elem_set) {
  return guard([&]() -> int {
    require(d && elem_set && num_sets >= 1, "goal_disc_set_elem_sets: bad arguments");
    for (int64_t e = 0; e < d->mesh.num_elems; ++e)
      require(elem_set[e] >= 0 && elem_set[e] < num_sets, "element set id out of range");
    d->mesh.elem_set.assign(elem_set, elem_set + d->mesh.num_elems);
    d->mesh.num_elem_sets = num_sets;
    d->elem_set.upload(d->mesh.elem_set.data(), d->mesh.num_elems);
    sync();
    return GOAL_OK;
  });
}

int goal_disc_set_soln(goal_disc d, const double* x) {
  return guard([&]() -> int {
    require(d && x, "bad arguments");
    HIPCHK(hipMemcpyAsync(d->x.p, x, sizeof(double) * d->num_dofs, hipMemcpyHostToDevice, stream()));
    sync();
    return GOAL_OK;
  });
}

int goal_disc_get_soln(goal_disc d, double* x) {
  return guard([&]() -> int {
    require(d && x, "bad arguments");
    d->x.download(x, d->num_dofs);
    return GOAL_OK;
  });
}

int goal_disc_add_soln(goal_disc d, const double* du) {
  return guard([&]() -> int {
    require(d && du, "bad arguments");
    DevBuf<double> tmp;
    tmp.upload(du, d->num_dofs);
    HIPCHK(goal::launch_axpy(d->x.p, tmp.p, 1.0, d->num_dofs, stream()));
    sync();
    return GOAL_OK;
  });
}

int goal_disc_add_soln_dev(goal_disc d, const double* du_dev) {
  return guard([&]() -> int {
    require(d && du_dev, "bad arguments");
    HIPCHK(goal::launch_axpy(d->x.p, du_dev, 1.0, d->num_dofs, stream()));
    return GOAL_OK;
  });
}

double* goal_disc_soln_dev(goal_disc d) { return d ? d->x.p : nullptr; }

// ---------------------------------------------------------------- states
int goal_states_create(goal_disc d, int model, int quad_order, goal_states* out) {
  return guard([&]() -> int {
    require(d && out, "goal_states_create: bad arguments");
    require(model >= 0 && model <= 2, "unknown model");
    require(quad_order == 1 || quad_order == 2, "quadrature order must be 1 or 2");
    std::unique_ptr<goal_states_s> s(new goal_states_s);
    s->d = d;
    s->model = model;
    s->nq = nq_of(quad_order);
    s->npts = d->mesh.num_elems * s->nq;
    s->sigma.alloc(9 * s->npts);
    HIPCHK(goal::launch_fill(s->sigma.p, 9 * s->npts, 0.0, stream()));
    if (model == GOAL_MODEL_J2) {
      s->eqps.alloc(s->npts);
      s->eqps_old.alloc(s->npts);
      s->Fp.alloc(9 * s->npts);
      s->Fp_old.alloc(9 * s->npts);
      HIPCHK(goal::launch_fill(s->eqps.p, s->npts, 0.0, stream()));
      HIPCHK(goal::launch_fill(s->eqps_old.p, s->npts, 0.0, stream()));
      for (int k = 0; k < 9; ++k) {
        const double v = (k % 4 == 0) ? 1.0 : 0.0;
        HIPCHK(goal::launch_fill(s->Fp.p + k * s->npts, s->npts, v, stream()));
        HIPCHK(goal::launch_fill(s->Fp_old.p + k * s->npts, s->npts, v, stream()));
      }
    }
    sync();
    *out = s.release();
    return GOAL_OK;
  });
}

int goal_states_destroy(goal_states s) {
  return guard([&]() -> int {
    if (s) {
      sync();
      delete s;
    }
    return GOAL_OK;
  });
}

int goal_states_update(goal_states s) {
  return guard([&]() -> int {
    require(s != nullptr, "null states");
    if (s->model == GOAL_MODEL_J2) {
      HIPCHK(goal::launch_copy(s->eqps_old.p, s->eqps.p, s->npts, stream()));
      HIPCHK(goal::launch_copy(s->Fp_old.p, s->Fp.p, 9 * s->npts, stream()));
    }
    return GOAL_OK;
  });
}

static DevBuf<double>* state_field(goal_states s, const char* name, int* ncomp) {
  std::string n(name);
  *ncomp = 9;
  if (n == "sigma") return &s->sigma;
  if (s->model == GOAL_MODEL_J2) {
    if (n == "Fp") return &s->Fp;
    if (n == "Fp_old") return &s->Fp_old;
    *ncomp = 1;
    if (n == "eqps") return &s->eqps;
    if (n == "eqps_old") return &s->eqps_old;
  }
  throw std::invalid_argument("unknown state field: " + n);
}

int goal_states_get(goal_states s, const char* name, double* out) {
  return guard([&]() -> int {
    require(s && name && out, "bad arguments");
    int nc;
    DevBuf<double>* f = state_field(s, name, &nc);
    f->download(out, nc * s->npts);
    return GOAL_OK;
  });
}

int goal_states_set(goal_states s, const char* name, const double* in) {
  return guard([&]() -> int {
    require(s && name && in, "bad arguments");
    int nc;
    DevBuf<double>* f = state_field(s, name, &nc);
    HIPCHK(hipMemcpyAsync(f->p, in, sizeof(double) * nc * s->npts, hipMemcpyHostToDevice, stream()));
    sync();
    return GOAL_OK;
  });
}

// ---------------------------------------------------------------- solinfo
int goal_solinfo_create(goal_disc d, goal_solinfo* out) {
  return guard([&]() -> int {
    require(d && out, "bad arguments");
    std::unique_ptr<goal_solinfo_s> s(new goal_solinfo_s);
    s->d = d;
    s->R.alloc(d->num_dofs);
    s->dMdu.alloc(d->num_dofs);
    s->vals.alloc(d->nnz);
    HIPCHK(goal::launch_fill(s->R.p, d->num_dofs, 0.0, stream()));
    HIPCHK(goal::launch_fill(s->dMdu.p, d->num_dofs, 0.0, stream()));
    HIPCHK(goal::launch_fill(s->vals.p, d->nnz, 0.0, stream()));
    sync();
    *out = s.release();
    return GOAL_OK;
  });
}

int goal_solinfo_destroy(goal_solinfo s) {
  return guard([&]() -> int {
    if (s) {
      sync();
      delete s;
    }
    return GOAL_OK;
  });
}

int goal_solinfo_zero_R(goal_solinfo s) {
  return guard([&]() -> int {
    require(s != nullptr, "null solinfo");
    s->zero_R = true;
    return GOAL_OK;
  });
}
int goal_solinfo_zero_dMdu(goal_solinfo s) {
  return guard([&]() -> int {
    require(s != nullptr, "null solinfo");
    s->zero_dMdu = true;
    return GOAL_OK;
  });
}
int goal_solinfo_zero_dRdu(goal_solinfo s) {
  return guard([&]() -> int {
    require(s != nullptr, "null solinfo");
    s->zero_vals = true;
    return GOAL_OK;
  });
}
int goal_solinfo_zero_all(goal_solinfo s) {
  return guard([&]() -> int {
    require(s != nullptr, "null solinfo");
    s->zero_R = s->zero_vals = s->zero_dMdu = true;
    return GOAL_OK;
  });
}

int goal_solinfo_get(goal_solinfo s, double* R, double* dMdu, double* vals) {
  return guard([&]() -> int {
    require(s != nullptr, "null solinfo");
    materialize(s);
    if (R) s->R.download(R, s->R.n);
    if (dMdu) s->dMdu.download(dMdu, s->dMdu.n);
    if (vals) s->vals.download(vals, s->vals.n);
    return GOAL_OK;
  });
}

int goal_solinfo_set(goal_solinfo s, const double* R, const double* dMdu, const double* vals) {
  return guard([&]() -> int {
    require(s != nullptr, "null solinfo");
    materialize(s);
    if (R) HIPCHK(hipMemcpyAsync(s->R.p, R, sizeof(double) * s->R.n, hipMemcpyHostToDevice, stream()));
    if (dMdu) HIPCHK(hipMemcpyAsync(s->dMdu.p, dMdu, sizeof(double) * s->dMdu.n, hipMemcpyHostToDevice, stream()));
    if (vals) HIPCHK(hipMemcpyAsync(s->vals.p, vals, sizeof(double) * s->vals.n, hipMemcpyHostToDevice, stream()));
    sync();
    return GOAL_OK;
  });
}

int goal_solinfo_dev(goal_solinfo s, double** R, double** dMdu, double** vals) {
  return guard([&]() -> int {
    require(s != nullptr, "null solinfo");
    materialize(s);
    if (R) *R = s->R.p;
    if (dMdu) *dMdu = s->dMdu.p;
    if (vals) *vals = s->vals.p;
    return GOAL_OK;
  });
}

int goal_linear_solve(goal_solinfo s, const double* b, double* x, double rtol, int max_iters, int restart,
                      int* iters_out, double* relres_out) {
  return goal_linear_solve_pc(s, b, x, rtol, max_iters, restart, GOAL_PRECOND_AMG, iters_out, relres_out);
}

int goal_linear_solve_pc(goal_solinfo s, const double* b, double* x, double rtol, int max_iters, int restart,
                         int precond, int* iters_out, double* relres_out) {
  return guard([&]() -> int {
    require(s && b && x && rtol > 0.0 && max_iters >= 0 && restart >= 1 && restart <= 1000,
            "goal_linear_solve: bad arguments");
    require(precond == GOAL_PRECOND_AMG || precond == GOAL_PRECOND_BLOCK_JACOBI,
            "goal_linear_solve: unknown preconditioner");
    materialize(s);
    goal_disc d = s->d;
    const int64_t n = d->num_dofs, nn = d->mesh.num_nodes;
    const int m = restart;
    if (s->V.n < (int64_t)(m + 1) * n) s->V.alloc((int64_t)(m + 1) * n);
    if (s->t.n < n) s->t.alloc(n);
    if (s->z.n < n) s->z.alloc(n);
    if (s->binv.n < 16 * nn) s->binv.alloc(16 * nn);
    const int nb = goal::multi_dot_blocks(n);
    const int mk = std::max(m, kAmgArnoldi);  // the multigrid set-up's Arnoldi shares these
    if (s->partial.n < (int64_t)(mk + 2) * nb) s->partial.alloc((int64_t)(mk + 2) * nb);
    if (s->h1.n < mk + 2) s->h1.alloc(mk + 2);
    if (s->h2.n < m + 1) s->h2.alloc(m + 1);
    if (s->y.n < m + 1) s->y.alloc(m + 1);
    if (s->solve_status.n < 1) s->solve_status.alloc(1);
    HIPCHK(hipMemsetAsync(s->solve_status.p, 0, sizeof(int32_t), stream()));
    const int64_t* rp = d->row_ptr.p;
    const int32_t* cols = d->cols.p;
    const double* A = s->vals.p;
    double* V = s->V.p;
    double* t = s->t.p;
    double* z = s->z.p;
    const bool amg = precond == GOAL_PRECOND_AMG;
    if (amg) amg_setup(s, s->solve_status.p);
    else HIPCHK(goal::launch_block_inverse(nn, rp, cols, A, s->binv.p, s->solve_status.p, stream()));
    auto prec = [&](const double* in, double* out) {  // out = M^{-1} in
      if (amg) amg_vcycle(s, *d->amg, 0, in, out);
      else HIPCHK(goal::launch_block_apply(nn, s->binv.p, in, out, stream()));
    };
    auto norm2 = [&](const double* v) {  // waits
      HIPCHK(goal::launch_reduce(1, v, nullptr, n, reduce_scratch(), result_slot(2), stream()));
      double r = 0.0;
      HIPCHK(hipMemcpyAsync(&r, result_slot(2), sizeof(double), hipMemcpyDeviceToHost, stream()));
      sync();
      return std::sqrt(r);
    };
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), hh1(m + 2), hh2(m + 1), yv(m + 1);
    auto Hm = [&](int i, int j) -> double& { return H[(size_t)j * (m + 1) + i]; };
    int total = 0;
    double beta0 = -1.0, relres = 0.0;
    for (;;) {
      // residual r = b - A x into V_0 (right preconditioning: the Krylov
      // space is that of A M^{-1}, and the residual GMRES minimizes is b - A x)
      HIPCHK(goal::launch_spmv(n, rp, cols, A, x, V, stream()));
      HIPCHK(goal::launch_rsub(n, b, V, stream()));
      const double beta = norm2(V);
      if (beta0 < 0.0) beta0 = beta;
      relres = beta0 > 0.0 ? beta / beta0 : 0.0;
      if (beta == 0.0 || relres <= rtol || total >= max_iters) break;
      HIPCHK(goal::launch_scale(n, 1.0 / beta, V, V, stream()));
      std::fill(g.begin(), g.end(), 0.0);
      g[0] = beta;
      int j = 0;
      bool done = false;
      while (j < m && total < max_iters) {
        double* w = V + (int64_t)(j + 1) * n;
        prec(V + (int64_t)j * n, t);
        HIPCHK(goal::launch_spmv(n, rp, cols, A, t, w, stream()));
        // classical Gram-Schmidt with the DGKS test (the reference's Belos
        // orthogonalization, src/goal_linear_solve.cpp:53-90): a second pass
        // only when the first removed more than 1 - 1/sqrt(2) of the norm
        // w is V_{j+1}, so one more dot in the same pass gives |w|^2
        HIPCHK(goal::launch_multi_dot(n, j + 2, V, n, w, s->partial.p, s->h1.p, stream()));
        HIPCHK(goal::launch_multi_axpy(n, j + 1, V, n, s->h1.p, -1.0, w, stream()));
        HIPCHK(hipMemcpyAsync(hh1.data(), s->h1.p, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, stream()));
        double hn = norm2(w);  // waits, so hh1 has arrived
        const double wn = std::sqrt(std::max(hh1[j + 1], 0.0));
        std::fill(hh2.begin(), hh2.begin() + j + 1, 0.0);
        if (hn < wn * 0.70710678118654752) {
          HIPCHK(goal::launch_multi_dot(n, j + 1, V, n, w, s->partial.p, s->h2.p, stream()));
          HIPCHK(goal::launch_multi_axpy(n, j + 1, V, n, s->h2.p, -1.0, w, stream()));
          HIPCHK(hipMemcpyAsync(hh2.data(), s->h2.p, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, stream()));
          hn = norm2(w);
        }
        for (int i = 0; i <= j; ++i) Hm(i, j) = hh1[i] + hh2[i];
        Hm(j + 1, j) = hn;
        if (hn > 0.0) HIPCHK(goal::launch_scale(n, 1.0 / hn, w, w, stream()));
        for (int i = 0; i < j; ++i) {
          const double a = Hm(i, j), c = Hm(i + 1, j);
          Hm(i, j) = cs[i] * a + sn[i] * c;
          Hm(i + 1, j) = -sn[i] * a + cs[i] * c;
        }
        const double a = Hm(j, j), c = Hm(j + 1, j), r = std::hypot(a, c);
        cs[j] = r > 0.0 ? a / r : 1.0;
        sn[j] = r > 0.0 ? c / r : 0.0;
        Hm(j, j) = r;
        Hm(j + 1, j) = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        ++j;
        ++total;
        relres = std::fabs(g[j]) / beta0;
        if (relres <= rtol || hn == 0.0) {
          done = true;
          break;
        }
      }
      // x += V y with H y = g (upper triangular, j x j)
      for (int i = j - 1; i >= 0; --i) {
        double acc = g[i];
        for (int k = i + 1; k < j; ++k) acc -= Hm(i, k) * yv[k];
        yv[i] = Hm(i, i) != 0.0 ? acc / Hm(i, i) : 0.0;
      }
      if (j > 0) {  // x += M^{-1} (V y)
        HIPCHK(hipMemcpyAsync(s->y.p, yv.data(), sizeof(double) * j, hipMemcpyHostToDevice, stream()));
        HIPCHK(goal::launch_fill(z, n, 0.0, stream()));
        HIPCHK(goal::launch_multi_axpy(n, j, V, n, s->y.p, 1.0, z, stream()));
        prec(z, t);
        HIPCHK(goal::launch_axpy(x, t, 1.0, n, stream()));
        sync();  // yv is reused by the next cycle
      }
      (void)done;  // the next pass recomputes the true residual and stops
    }
    int st = 0;
    HIPCHK(hipMemcpyAsync(&st, s->solve_status.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
    sync();
    if (iters_out) *iters_out = total;
    if (relres_out) *relres_out = relres;
    if (st) return set_error(GOAL_ERR_INVALID, "goal_linear_solve: singular diagonal block or coarsest matrix");
    return GOAL_OK;
  });
}

int goal_solinfo_norm_R(goal_solinfo s, double* norm) {
  return guard([&]() -> int {
    require(s && norm, "bad arguments");
    materialize(s);
    HIPCHK(goal::launch_reduce(1, s->R.p, nullptr, s->R.n, reduce_scratch(), result_slot(0), stream()));
    double v = 0.0;
    HIPCHK(hipMemcpyAsync(&v, result_slot(0), sizeof(double), hipMemcpyDeviceToHost, stream()));
    sync();
    *norm = std::sqrt(v);
    return GOAL_OK;
  });
}

int goal_solinfo_dot_R(goal_solinfo s, const double* z_dev, double* dot) {
  return guard([&]() -> int {
    require(s && z_dev && dot, "bad arguments");
    materialize(s);
    HIPCHK(goal::launch_reduce(2, s->R.p, z_dev, s->R.n, reduce_scratch(), result_slot(0), stream()));
    double v = 0.0;
    HIPCHK(hipMemcpyAsync(&v, result_slot(0), sizeof(double), hipMemcpyDeviceToHost, stream()));
    sync();
    *dot = -v;
    return GOAL_OK;
  });
}

// ------------------------------------------------------------- mechanics
int goal_mech_create(goal_disc d, goal_states st, int num_sets, const goal_material* mats, int stabilization,
                     int quad_order, goal_mech* out) {
  return guard([&]() -> int {
    require(d && st && mats && out && num_sets >= 1, "goal_mech_create: bad arguments");
    require(st->d == d, "states belong to another disc");
    require(quad_order == 1 || quad_order == 2, "quadrature order must be 1 or 2");
    require(nq_of(quad_order) == st->nq, "states were created for another quadrature order");
    require(num_sets >= d->mesh.num_elem_sets, "fewer materials than element sets");
    std::unique_ptr<goal_mech_s> m(new goal_mech_s);
    m->d = d;
    m->st = st;
    m->model = mats[0].model;
    m->stab = stabilization ? 1 : 0;
    m->quad_order = quad_order;
    m->num_sets = num_sets;
    std::vector<goal::MatParams> p(num_sets);
    for (int i = 0; i < num_sets; ++i) {
      require(mats[i].model == m->model, "all element sets must use the same model");
      require(mats[i].E > 0 && mats[i].nu > -1.0 && mats[i].nu < 0.5, "invalid E or nu");
      if (m->model == GOAL_MODEL_J2) require(mats[i].Y > 0, "J2 needs Y > 0");
      p[i] = derive(mats[i]);
    }
    require(m->model == st->model, "states were created for another model");
    m->mats.upload(p.data(), num_sets);
    m->status.alloc(1);
    HIPCHK(hipMemsetAsync(m->status.p, 0, sizeof(int), stream()));
    m->elem_value.alloc(d->mesh.num_elems);
    sync();
    *out = m.release();
    return GOAL_OK;
  });
}

int goal_mech_destroy(goal_mech m) {
  return guard([&]() -> int {
    if (m) {
      sync();
      delete m;
    }
    return GOAL_OK;
  });
}

int goal_mech_set_option(goal_mech m, const char* key, int64_t value) {
  return guard([&]() -> int {
    require(m && key, "bad arguments");
    std::string k(key);
    if (k == "blocks") {
      require(value >= 0 && value < (1 << 30), "blocks out of range");
      m->blocks = (int)value;
    } else if (k == "xcd_remap") {
      m->xcd_remap = value ? 1 : 0;
    } else if (k == "check") {
      m->check = value ? 1 : 0;
    } else if (k == "stabilization") {
      m->stab = value ? 1 : 0;
    } else if (k == "tangent") {
      require(value == 32 || value == 64, "tangent must be 32 or 64");
      m->tangent = (int)value;
    } else if (k == "overlap") {
      require(value >= -1 && value <= 1024, "overlap must be -1 (automatic) or 0..1024 chunks");
      m->overlap = (int)value;
    } else if (k == "overlap_blocks") {
      require(value >= 0 && value < (1 << 30), "overlap_blocks out of range");
      m->overlap_blocks = (int)value;
    } else if (k == "timing") {
      if (value)
        for (auto& e : m->ev)
          if (!e) HIPCHK(hipEventCreate(&e));
      m->timing = value != 0;
      m->timed = false;
    } else {
      return set_error(GOAL_ERR_INVALID, "unknown option: %s", key);
    }
    return GOAL_OK;
  });
}

int goal_mech_timing(goal_mech m, float* element_ms, float* gather_ms) {
  return guard([&]() -> int {
    require(m && element_ms && gather_ms, "bad arguments");
    require(m->timing && m->timed, "goal_mech_timing: no timed assembly (set the \"timing\" option first)");
    HIPCHK(hipEventSynchronize(m->ev[2]));
    HIPCHK(hipEventElapsedTime(element_ms, m->ev[0], m->ev[1]));
    HIPCHK(hipEventElapsedTime(gather_ms, m->ev[1], m->ev[2]));
    return GOAL_OK;
  });
}

int goal_mech_status(goal_mech m, int* flags) {
  return guard([&]() -> int {
    require(m && flags, "bad arguments");
    HIPCHK(hipMemcpyAsync(flags, m->status.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
    sync();
    HIPCHK(hipMemsetAsync(m->status.p, 0, sizeof(int), stream()));
    return GOAL_OK;
  });
}

int goal_asm_residual(goal_mech m, goal_solinfo s, int save_states) {
  return guard([&]() -> int {
    require(m && s && s->d == m->d, "goal_asm_residual: bad arguments");
    goal::AsmArgs a = asm_args(m);
    a.save_state = save_states ? 1 : 0;
    run_assembly(m, s, goal::ASM_RESIDUAL, a, false);
    return finish_assembly(m);
  });
}

int goal_asm_jacobian(goal_mech m, goal_solinfo s, int adjoint, int save_states) {
  return guard([&]() -> int {
    require(m && s && s->d == m->d, "goal_asm_jacobian: bad arguments");
    goal::AsmArgs a = asm_args(m);
    a.save_state = save_states ? 1 : 0;
    run_assembly(m, s, adjoint ? goal::ASM_ADJOINT : goal::ASM_JACOBIAN, a, true);
    return finish_assembly(m);
  });
}

int goal_asm_jacobian_noalloc(goal_mech m, goal_solinfo s, int adjoint, int save_states) {
  return guard([&]() -> int {
    require(m && s && s->d == m->d, "goal_asm_jacobian_noalloc: bad arguments");
    goal::AsmArgs a = asm_args(m);  // ctx() is up: the mechanics was created through it
    const int64_t ne = m->d->mesh.num_elems, nd = 4 * (int64_t)m->d->mesh.nodes_per_elem;
    // element matrices (FP64; FP32 ones take half), element residuals, then
    // (J2) the point states
    const int kind = adjoint ? goal::ASM_ADJOINT : goal::ASM_JACOBIAN;
    const bool pts = goal::uses_point_pass(m->model, kind);
    auto need_of = [&](int64_t n) {
      return n * (nd * nd + nd) + (pts ? goal::point_state_doubles(n * m->st->nq, goal::point_stride(kind)) : 0);
    };
    const int64_t need = need_of(ne);
    if (need > kNoallocScratchDoubles) {
      int64_t most = kNoallocScratchDoubles / (nd * nd + nd + (pts ? m->st->nq * goal::point_stride(kind) : 0));
      while (most > 0 && need_of(most) > kNoallocScratchDoubles) --most;
      return set_error(GOAL_ERR_INVALID,
                       "goal_asm_jacobian_noalloc: %lld elements need %lld doubles of scratch, more than its "
                       "%lld (at most %lld elements of this mesh and model)",
                       (long long)ne, (long long)need, (long long)kNoallocScratchDoubles, (long long)most);
    }
    double* base = g_noalloc_ptr[ctx().device];
    require(base != nullptr, "goal_asm_jacobian_noalloc: scratch address not set up for this device");
    a.save_state = save_states ? 1 : 0;
    // every value the gather reads is written by this call's element pass, and
    // every point state the element pass reads by its point pass, so nothing
    // left from an earlier call needs clearing
    run_assembly(m, s, kind, a, true, base + ne * nd * nd, base, base + ne * (nd * nd + nd));
    return finish_assembly(m);
  });
}

int goal_batch_create(int n, const goal_mech* mechs, const goal_solinfo* infos, goal_batch* out) {
  return guard([&]() -> int {
    require(n >= 1 && mechs && infos && out, "goal_batch_create: bad arguments");
    auto b = std::make_unique<goal_batch_s>();
    for (int i = 0; i < n; ++i) {
      require(mechs[i] && infos[i] && infos[i]->d == mechs[i]->d, "goal_batch_create: mismatched pair");
      b->mechs.push_back(mechs[i]);
      b->infos.push_back(infos[i]);
    }
    *out = b.release();
    return GOAL_OK;
  });
}

int goal_batch_destroy(goal_batch b) {
  return guard([&]() -> int {
    if (b) {
      sync();
      delete b;
    }
    return GOAL_OK;
  });
}

int goal_batch_tune(goal_batch b, int reps, int* blocks_out, float* ms_out) {
  return guard([&]() -> int {
    require(b && reps >= 1, "goal_batch_tune: bad arguments");
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx().device);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto time_one = [&](goal_mech m, goal_solinfo s) {
      std::vector<float> t;
      for (int r = 0; r < reps + 1; ++r) {
        s->zero_R = s->zero_dMdu = s->zero_vals = true;
        goal::AsmArgs a = asm_args(m);
        a.save_state = 1;
        HIPCHK(hipEventRecord(e0, stream()));
        run_assembly(m, s, goal::ASM_JACOBIAN, a, true);
        HIPCHK(hipEventRecord(e1, stream()));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (r > 0) t.push_back(ms);  // the first run allocates and warms up
      }
      std::sort(t.begin(), t.end());
      return t[t.size() / 2];
    };
    for (size_t i = 0; i < b->mechs.size(); ++i) {
      goal_mech m = b->mechs[i];
      // time the path the graph captures: one pass, never chunked
      const int overlap = m->overlap;
      m->overlap = 0;
      // candidate persistent grids: the occupancy-sized one (0) and fixed
      // multiples of the CU count; more workgroups than element groups are
      // never useful
      const int64_t ne = m->d->mesh.num_elems;
      std::vector<int> cand{0};
      for (int k : {1, 2, 4, 8, 16})
        if ((int64_t)k * cus <= std::max<int64_t>(ne / 4, 8)) cand.push_back(k * cus);
      int best = 0;
      float best_ms = 1e30f;
      for (int c : cand) {
        m->blocks = c;
        const float ms = time_one(m, b->infos[i]);
        if (ms < best_ms) {
          best_ms = ms;
          best = c;
        }
      }
      m->overlap = overlap;
      m->blocks = best;
      if (blocks_out) blocks_out[i] = best;
      if (ms_out) ms_out[i] = best_ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (b->exec) {  // launch shapes changed: capture again before the next run
      (void)hipGraphExecDestroy(b->exec);
      b->exec = nullptr;
    }
    if (b->graph) {
      (void)hipGraphDestroy(b->graph);
      b->graph = nullptr;
    }
    return GOAL_OK;
  });
}

int goal_batch_capture(goal_batch b) {
  return guard([&]() -> int {
    require(b, "goal_batch_capture: bad arguments");
    for (size_t i = 0; i < b->mechs.size(); ++i) {
      goal_mech m = b->mechs[i];
      require(!m->timing, "goal_batch_capture: turn the \"timing\" option off first");
      // one eager assembly sizes the element buffers (no allocation inside a capture)
      b->infos[i]->zero_R = b->infos[i]->zero_dMdu = b->infos[i]->zero_vals = true;
      goal::AsmArgs a = asm_args(m);
      a.save_state = 1;
      run_assembly(m, b->infos[i], goal::ASM_JACOBIAN, a, true);
    }
    sync();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    HIPCHK(hipStreamBeginCapture(stream(), hipStreamCaptureModeRelaxed));
    try {
      for (size_t i = 0; i < b->mechs.size(); ++i) {
        // pending zero: the captured kernels overwrite R and the CSR values,
        // so every replay assembles from scratch
        b->infos[i]->zero_R = b->infos[i]->zero_dMdu = b->infos[i]->zero_vals = true;
        goal::AsmArgs a = asm_args(b->mechs[i]);
        a.save_state = 1;
        run_assembly(b->mechs[i], b->infos[i], goal::ASM_JACOBIAN, a, true);
      }
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream(), &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(stream(), &b->graph));
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return GOAL_OK;
  });
}

int goal_batch_run(goal_batch b) {
  return guard([&]() -> int {
    require(b && b->exec, "goal_batch_run: capture the batch first");
    HIPCHK(hipGraphLaunch(b->exec, stream()));
    // the replay writes R, dMdu and the CSR values of every pair from scratch,
    // so a zero_* still pending from before the launch is satisfied
    for (goal_solinfo s : b->infos) s->zero_R = s->zero_dMdu = s->zero_vals = false;
    return GOAL_OK;
  });
}

int goal_asm_localize(goal_mech m, goal_solinfo s, const double* z_diff, const double* z_coarse) {
  return guard([&]() -> int {
    require(m && s && s->d == m->d && z_diff && z_coarse, "goal_asm_localize: bad arguments");
    goal::AsmArgs a = asm_args(m);
    a.save_state = 0;
    a.z_diff = z_diff;
    a.z_coarse = z_coarse;
    run_assembly(m, s, goal::ASM_LOCALIZE, a, false);
    return finish_assembly(m);
  });
}

int goal_asm_qoi(goal_mech m, goal_solinfo s, int kind, int elem_set, double* value, double* elem_values) {
  return guard([&]() -> int {
    require(m && value, "goal_asm_qoi: bad arguments");
    require(kind == GOAL_QOI_AVG_DISP || kind == GOAL_QOI_AVG_VM, "unknown QoI kind");
    require(!s || s->d == m->d, "solinfo belongs to another disc");
    require_materials(m);
    goal::QoiArgs q{};
    goal_disc d = m->d;
    q.num_elems = d->mesh.num_elems;
    q.conn = d->conn.p;
    q.coords = d->coords.p;
    q.x = d->x.p;
    q.elem_set = d->elem_set.p;
    q.mats = m->mats.p;
    q.Fp_old = m->st->Fp_old.p;
    q.eqps_old = m->st->eqps_old.p;
    q.model = m->model;
    q.kind = kind;
    q.qoi_set = elem_set;
    q.elem_value = m->elem_value.p;
    if (s) materialize(s);
    q.dMdu = s ? s->dMdu.p : nullptr;
    q.status = m->status.p;
    HIPCHK(goal::launch_qoi(d->mesh.order, m->quad_order, q, stream()));
    HIPCHK(goal::launch_reduce(0, m->elem_value.p, nullptr, d->mesh.num_elems, reduce_scratch(), result_slot(1), stream()));
    HIPCHK(hipMemcpyAsync(value, result_slot(1), sizeof(double), hipMemcpyDeviceToHost, stream()));
    sync();
    if (elem_values) m->elem_value.download(elem_values, d->mesh.num_elems);
    int st = finish_assembly(m);
    return st;
  });
}

// ---------------------------------------------------- learned indicator
int goal_learn_features(goal_mech m, double qoi_x, float* feat_dev) {
  return guard([&]() -> int {
    require(m && feat_dev, "goal_learn_features: bad arguments");
    require_materials(m);
    goal_disc d = m->d;
    const int64_t ne = d->mesh.num_elems, nn = d->mesh.num_nodes;
    // per-element integral of the von Mises stress (the AvgVM integrand over all elements)
    goal::QoiArgs q{};
    q.num_elems = ne;
    q.conn = d->conn.p;
    q.coords = d->coords.p;
    q.x = d->x.p;
    q.elem_set = d->elem_set.p;
    q.mats = m->mats.p;
    q.Fp_old = m->st->Fp_old.p;
    q.eqps_old = m->st->eqps_old.p;
    q.model = m->model;
    q.kind = GOAL_QOI_AVG_VM;
    q.qoi_set = -1;
    q.elem_value = m->elem_value.p;
    q.dMdu = nullptr;
    q.status = m->status.p;
    HIPCHK(goal::launch_qoi(d->mesh.order, m->quad_order, q, stream()));
    require(d->mesh.num_elems > 0, "goal_learn_features: empty mesh");
    if (m->feat_vol.n < 10 * ne) m->feat_vol.alloc(10 * ne);
    if (m->feat_vm.n < 10 * nn) m->feat_vm.alloc(10 * nn);
    goal::FeatureArgs f{};
    f.num_elems = ne;
    f.num_nodes = nn;
    f.npe = d->mesh.nodes_per_elem;
    f.conn = d->conn.p;
    f.coords = d->coords.p;
    f.x = d->x.p;
    f.elem_set = d->elem_set.p;
    f.mats = m->mats.p;
    f.elem_vm = m->elem_value.p;
    f.eqps = m->model == goal::MODEL_J2 ? m->st->eqps.p : nullptr;
    f.nq = m->st->nq;
    f.inc_ptr = d->inc_ptr.p;
    f.inc = d->inc.p;
    f.qoi_x = qoi_x;
    f.elem_scratch = m->feat_vol.p;
    f.node_scratch = m->feat_vm.p;
    f.feat = feat_dev;
    HIPCHK(goal::launch_features(f, stream()));
    return finish_assembly(m);
  });
}

int goal_mlp_create(goal_mlp* out) {
  return guard([&]() -> int {
    require(out != nullptr, "goal_mlp_create: bad arguments");
    ctx();
    auto* m = new goal_mlp_s;
    try {
      m->params.alloc(goal::kMlpParams);
      m->wq.alloc(goal::kMlpParams);
      m->norms.alloc(goal::kMlpNorms);
      m->adam_m.alloc(goal::kMlpParams);
      m->adam_v.alloc(goal::kMlpParams);
      m->grad.alloc(goal::kMlpParams);
      m->loss.alloc(1);
      m->ood.alloc(1);
      std::vector<float> zero(goal::kMlpParams, 0.0f), norms(goal::kMlpNorms, 0.0f);
      for (int k = 0; k < goal::kMlpIn; ++k) {
        norms[goal::kMlpIn + k] = 1.0f;
        norms[goal::kMlpOffLo + k] = -INFINITY;
        norms[goal::kMlpOffHi + k] = INFINITY;
      }
      norms[2 * goal::kMlpIn + 1] = 1.0f;
      HIPCHK(hipMemcpyAsync(m->params.p, zero.data(), 4 * zero.size(), hipMemcpyHostToDevice, stream()));
      HIPCHK(hipMemcpyAsync(m->adam_m.p, zero.data(), 4 * zero.size(), hipMemcpyHostToDevice, stream()));
      HIPCHK(hipMemcpyAsync(m->adam_v.p, zero.data(), 4 * zero.size(), hipMemcpyHostToDevice, stream()));
      HIPCHK(hipMemcpyAsync(m->norms.p, norms.data(), 4 * norms.size(), hipMemcpyHostToDevice, stream()));
      HIPCHK(goal::launch_mlp_round(m->args(), stream()));
      sync();
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
    return GOAL_OK;
  });
}

int goal_mlp_destroy(goal_mlp m) {
  delete m;
  return GOAL_OK;
}

int goal_mlp_set(goal_mlp m, const float* params, const float* norms) {
  return guard([&]() -> int {
    require(m != nullptr, "goal_mlp_set: bad arguments");
    if (params) {
      for (int p = 0; p < goal::kMlpParams; ++p) require(std::isfinite(params[p]), "goal_mlp_set: non-finite parameter");
      std::vector<float> zero(goal::kMlpParams, 0.0f);
      HIPCHK(hipMemcpyAsync(m->params.p, params, 4 * goal::kMlpParams, hipMemcpyHostToDevice, stream()));
      HIPCHK(hipMemcpyAsync(m->adam_m.p, zero.data(), 4 * zero.size(), hipMemcpyHostToDevice, stream()));
      HIPCHK(hipMemcpyAsync(m->adam_v.p, zero.data(), 4 * zero.size(), hipMemcpyHostToDevice, stream()));
      HIPCHK(goal::launch_mlp_round(m->args(), stream()));
      m->step = 0;
    }
    if (norms) {
      for (int k = 0; k < goal::kMlpIn; ++k) require(norms[goal::kMlpIn + k] > 0.0f, "goal_mlp_set: input scale must be positive");
      require(norms[2 * goal::kMlpIn + 1] > 0.0f, "goal_mlp_set: output scale must be positive");
      HIPCHK(hipMemcpyAsync(m->norms.p, norms, 4 * goal::kMlpNorms, hipMemcpyHostToDevice, stream()));
    }
    sync();
    return GOAL_OK;
  });
}

int goal_mlp_get(goal_mlp m, float* params, uint16_t* weights_bf16, float* norms) {
  return guard([&]() -> int {
    require(m != nullptr, "goal_mlp_get: bad arguments");
    if (params) m->params.download(params, goal::kMlpParams);
    if (weights_bf16) m->wq.download(weights_bf16, goal::kMlpParams);
    if (norms) m->norms.download(norms, goal::kMlpNorms);
    return GOAL_OK;
  });
}

int goal_mlp_forward(goal_mlp m, const float* x_dev, int64_t n, float* y_dev, int64_t* num_outside) {
  return guard([&]() -> int {
    require(m && n >= 0 && (n == 0 || (x_dev && y_dev)), "goal_mlp_forward: bad arguments");
    HIPCHK(hipMemsetAsync(m->ood.p, 0, sizeof(int), stream()));
    HIPCHK(goal::launch_mlp_forward(m->args(), x_dev, n, y_dev, m->ood.p, stream()));
    if (num_outside) {
      int c = 0;
      HIPCHK(hipMemcpyAsync(&c, m->ood.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
      sync();
      *num_outside = c;
    }
    return GOAL_OK;
  });
}

namespace {
void mlp_workspace(goal_mlp m, int64_t batch) {
  const int nb = goal::mlp_grad_blocks(batch);
  if (m->partial.n < (int64_t)nb * goal::kMlpParams) m->partial.alloc((int64_t)nb * goal::kMlpParams);
  if (m->partial_loss.n < nb) m->partial_loss.alloc(nb);
}
}  // namespace

int goal_mlp_gradient(goal_mlp m, const float* x_dev, const float* y_dev, const int32_t* idx_dev, int64_t n,
                      float* loss, float* grad) {
  return guard([&]() -> int {
    require(m && x_dev && y_dev && idx_dev && n > 0, "goal_mlp_gradient: bad arguments");
    mlp_workspace(m, n);
    HIPCHK(goal::launch_mlp_grad(m->args(), x_dev, y_dev, idx_dev, n, m->partial.p, m->partial_loss.p, m->grad.p,
                                 m->loss.p, stream()));
    if (grad) m->grad.download(grad, goal::kMlpParams);
    if (loss) m->loss.download(loss, 1);
    sync();
    return GOAL_OK;
  });
}

int goal_mlp_train(goal_mlp m, const float* x_dev, const float* y_dev, const int32_t* idx_dev, int64_t batch,
                   int steps, float lr, float* last_loss) {
  return guard([&]() -> int {
    require(m && x_dev && y_dev && idx_dev && batch > 0 && steps >= 0 && lr > 0.0f, "goal_mlp_train: bad arguments");
    mlp_workspace(m, batch);
    for (int k = 0; k < steps; ++k) {
      HIPCHK(goal::launch_mlp_grad(m->args(), x_dev, y_dev, idx_dev + (int64_t)k * batch, batch, m->partial.p,
                                   m->partial_loss.p, m->grad.p, m->loss.p, stream()));
      HIPCHK(goal::launch_mlp_adam(m->args(), m->grad.p, ++m->step, lr, stream()));
    }
    if (last_loss) m->loss.download(last_loss, 1);
    sync();
    return GOAL_OK;
  });
}

// ------------------------------------------------------------- Dirichlet
int goal_eval(const char* expr, double x, double y, double z, double t, double* out) {
  return guard([&]() -> int {
    require(expr && out, "goal_eval: bad arguments");
    *out = goal::eval(expr, x, y, z, t);
    return GOAL_OK;
  });
}

static int dbc_apply_dev(goal_disc d, goal_solinfo s, int comp, const int32_t* nodes_dev, int64_t n,
                         const std::vector<double>& values, int jacobian) {
  materialize(s);
  d->dbc_values.upload(values.data(), n);
  HIPCHK(goal::launch_dbc(nodes_dev, d->dbc_values.p, n, comp, d->x.p, s->R.p, jacobian ? s->dMdu.p : nullptr,
                          d->row_ptr.p, d->cols.p, jacobian ? s->vals.p : nullptr, stream()));
  sync();
  return GOAL_OK;
}

int goal_dbc_apply(goal_disc d, goal_solinfo s, int comp, const char* node_set, const char* expr, double t,
                   int jacobian) {
  return guard([&]() -> int {
    require(d && s && s->d == d && node_set && expr, "goal_dbc_apply: bad arguments");
    require(comp >= 0 && comp < goal::kDim, "DBC component must be 0..2");
    const DevBuf<int32_t>& nodes = dev_node_set(d, node_set);
    const std::vector<int32_t>& hn = d->mesh.node_sets[node_set];
    goal::Expr f(expr);
    std::vector<double> vals(hn.size());
    for (size_t k = 0; k < hn.size(); ++k) {
      const double* X = &d->mesh.coords[3 * (int64_t)hn[k]];
      vals[k] = f(X[0], X[1], X[2], t);
    }
    return dbc_apply_dev(d, s, comp, nodes.p, (int64_t)hn.size(), vals, jacobian);
  });
}

int goal_dbc_apply_values(goal_disc d, goal_solinfo s, int comp, int64_t n, const int32_t* nodes,
                          const double* values, int jacobian) {
  return guard([&]() -> int {
    require(d && s && s->d == d && n >= 0 && (n == 0 || (nodes && values)), "goal_dbc_apply_values: bad arguments");
    require(comp >= 0 && comp < goal::kDim, "DBC component must be 0..2");
    for (int64_t k = 0; k < n; ++k) require(nodes[k] >= 0 && nodes[k] < d->mesh.num_nodes, "DBC node out of range");
    DevBuf<int32_t> dn;
    dn.upload(nodes, n);
    std::vector<double> v(values, values + n);
    return dbc_apply_dev(d, s, comp, dn.p, n, v, jacobian);
  });
}

// One traction entry (src/goal_tbcs.cpp:30-66): per face of the side set, the
// reference's one-point rule -- the centroid, weight 1/2, dv = twice the area
// (apf getIntPoint / getIntWeight / getDV on a triangle) -- so each face node
// a takes R[a, i] -= T_i(x_c, t) W_{a,i} |face|.  The weight is the "uw"
// evaluator's value at the centroid: N_a (VectorWeight) when weight_dev is
// null, z_i(x_c) N_a (DisplacementAdjoint over the DOF vector weight_dev,
// src/goal_displacement_adjoint.cpp:37-53) otherwise.  N_a(centroid) is 1/3
// on a linear face, -1/9 (vertices) and 4/9 (edge nodes) on a quadratic one.
// Shares of a row are summed on the host in side-set order, then one kernel
// subtracts them.  A dead load: the Jacobian is untouched.
int goal_tbc_apply(goal_disc d, goal_solinfo s, const char* side_set_name, const char* tx, const char* ty,
                   const char* tz, double t, const double* weight_dev) {
  return guard([&]() -> int {
    require(d && s && s->d == d && side_set_name && tx && ty && tz, "goal_tbc_apply: bad arguments");
    const std::vector<int32_t>& sides = side_set(d, side_set_name);
    const goal::Expr T[3] = {goal::Expr(tx), goal::Expr(ty), goal::Expr(tz)};
    const int npe = d->mesh.nodes_per_elem;
    const bool p2 = d->mesh.order == 2;
    const int nfn = p2 ? 6 : 3;
    struct Face {
      int32_t nodes[6];
      double N[6], T[3], w_dv;
    };
    std::vector<Face> faces(sides.size());
    for (size_t k = 0; k < sides.size(); ++k) {
      Face& F = faces[k];
      const int64_t e = sides[k] / 4;
      const int f = sides[k] % 4;
      const int32_t* c = &d->mesh.conn[e * npe];
      const int lv[3] = {kTetFace[f][0], kTetFace[f][1], kTetFace[f][2]};
      const double* X[3];
      for (int j = 0; j < 3; ++j) X[j] = &d->mesh.coords[3 * (int64_t)c[lv[j]]];
      double a[3], b[3], xc[3];
      for (int i = 0; i < 3; ++i) {
        a[i] = X[1][i] - X[0][i];
        b[i] = X[2][i] - X[0][i];
        xc[i] = (X[0][i] + X[1][i] + X[2][i]) / 3.0;
      }
      const double nrm[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
      F.w_dv = 0.5 * std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
      for (int i = 0; i < 3; ++i) F.T[i] = T[i](xc[0], xc[1], xc[2], t);
      for (int j = 0; j < 3; ++j) {
        F.nodes[j] = c[lv[j]];
        F.N[j] = p2 ? -1.0 / 9.0 : 1.0 / 3.0;
      }
      if (p2)
        for (int j = 0; j < 3; ++j) {
          F.nodes[3 + j] = c[4 + tet_edge_between(lv[j], lv[(j + 1) % 3])];
          F.N[3 + j] = 4.0 / 9.0;
        }
    }
    // the displacement rows of every face node: where the loads go, and
    // where the weight field is read
    std::vector<int64_t> rows;
    for (const Face& F : faces)
      for (int j = 0; j < nfn; ++j)
        for (int i = 0; i < goal::kDim; ++i) rows.push_back(goal::kNumEqs * (int64_t)F.nodes[j] + i);
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    const int64_t nr = (int64_t)rows.size();
    auto index = [&](int64_t row) { return std::lower_bound(rows.begin(), rows.end(), row) - rows.begin(); };
    materialize(s);
    d->tbc_rows.upload(rows.data(), nr);
    std::vector<double> z;
    if (weight_dev) {
      d->tbc_values.alloc(nr);
      HIPCHK(goal::launch_take_rows(d->tbc_rows.p, weight_dev, nr, d->tbc_values.p, stream()));
      z.resize(nr);
      d->tbc_values.download(z.data(), nr);
    }
    std::vector<double> load(nr, 0.0);
    for (const Face& F : faces) {
      double zc[3] = {1.0, 1.0, 1.0};
      if (weight_dev)
        for (int i = 0; i < goal::kDim; ++i) {
          zc[i] = 0.0;
          for (int j = 0; j < nfn; ++j) zc[i] += z[index(goal::kNumEqs * (int64_t)F.nodes[j] + i)] * F.N[j];
        }
      for (int j = 0; j < nfn; ++j)
        for (int i = 0; i < goal::kDim; ++i)
          load[index(goal::kNumEqs * (int64_t)F.nodes[j] + i)] += F.T[i] * zc[i] * F.N[j] * F.w_dv;
    }
    d->tbc_values.upload(load.data(), nr);
    HIPCHK(goal::launch_sub_rows(d->tbc_rows.p, d->tbc_values.p, nr, s->R.p, stream()));
    sync();
    return GOAL_OK;
  });
}

// ------------------------------------------------------------ nested DWR
// Nested::Nested's uniform refinement (src/goal_nested.cpp:13-28, 144-154) on
// the device: the children, the new vertices' parent pairs and coordinates
// come from kernels over the base Disc's device mesh and CSR graph, numbered
// as goal::refine_uniform numbers them (edges in (lo, hi) order), so the same
// fine mesh results; the host then adds the node sets and builds the fine
// Disc's graph.
extern "C++" {
namespace {
goal::Refinement refine_on_device(goal_disc base) {
  const goal::HostMesh& m = base->mesh;
  const int64_t nv = m.num_nodes, ne = m.num_elems;
  DevBuf<int64_t> ud;
  ud.alloc(nv);
  HIPCHK(goal::launch_upper_degree(base->row_ptr.p, base->cols.p, nv, ud.p, stream()));
  std::vector<int64_t> eb(nv + 1, 0);
  ud.download(eb.data() + 1, nv);
  for (int64_t a = 0; a < nv; ++a) eb[a + 1] += eb[a];
  const int64_t nedges = eb[nv];
  require(nv + nedges <= (int64_t)INT32_MAX / goal::kNumEqs, "nested refinement: mesh too large");
  DevBuf<int64_t> ebase;
  ebase.upload(eb.data(), nv + 1);
  DevBuf<double> fX;
  fX.alloc(3 * (nv + nedges));
  HIPCHK(goal::launch_copy(fX.p, base->coords.p, 3 * nv, stream()));
  DevBuf<int32_t> ep, fconn, status;
  ep.alloc(2 * nedges);
  fconn.alloc(32 * ne);
  const int32_t zero = 0;
  status.upload(&zero, 1);
  HIPCHK(goal::launch_refine(base->conn.p, base->coords.p, base->row_ptr.p, base->cols.p, ebase.p, ne, nv, ep.p, fX.p,
                             fconn.p, status.p, stream()));
  int32_t st = 0;
  status.download(&st, 1);
  require(st == 0, "nested refinement: an element edge is missing from the base graph");
  goal::Refinement r;
  r.num_old_vertices = nv;
  r.edge_parents.resize(2 * nedges);
  ep.download(r.edge_parents.data(), 2 * nedges);
  r.parent.resize(8 * ne);
  for (int64_t c = 0; c < 8 * ne; ++c) r.parent[c] = (int32_t)(c / 8);
  goal::HostMesh& f = r.fine;
  f.order = 1;
  f.nodes_per_elem = 4;
  f.num_nodes = nv + nedges;
  f.num_elems = 8 * ne;
  f.coords.resize(3 * f.num_nodes);
  fX.download(f.coords.data(), 3 * f.num_nodes);
  f.conn.resize(32 * ne);
  fconn.download(f.conn.data(), 32 * ne);
  f.elem_set.resize(8 * ne);
  for (int64_t c = 0; c < 8 * ne; ++c) f.elem_set[c] = m.elem_set.empty() ? 0 : m.elem_set[c / 8];
  f.num_elem_sets = m.num_elem_sets;
  // node sets: old members plus new vertices whose two parents are members
  for (auto const& kv : m.node_sets) {
    std::vector<char> in(nv, 0);
    for (int32_t n : kv.second) in[n] = 1;
    std::vector<int32_t> set(kv.second);
    for (int64_t k = 0; k < nedges; ++k)
      if (in[r.edge_parents[2 * k]] && in[r.edge_parents[2 * k + 1]]) set.push_back((int32_t)(nv + k));
    f.node_sets[kv.first] = std::move(set);
  }
  goal::finalize_mesh(f);
  return r;
}
}  // namespace
}  // extern "C++"

int goal_nested_create(goal_disc base, goal_nested* out) {
  return guard([&]() -> int {
    require(base && out, "goal_nested_create: bad arguments");
    require(base->mesh.order == 1, "nested refinement needs a linear base mesh");
    goal::Refinement r = refine_on_device(base);
    std::unique_ptr<goal_nested_s> n(new goal_nested_s);
    n->base = base;
    n->num_old = r.num_old_vertices;
    n->num_new = (int64_t)r.edge_parents.size() / 2;
    n->parent_host = r.parent;
    n->edge_parents_host = r.edge_parents;
    n->parent.upload(r.parent.data(), (int64_t)r.parent.size());
    n->edge_parents.upload(r.edge_parents.data(), (int64_t)r.edge_parents.size());
    // given side sets: each base face splits into four faces of its children
    std::map<std::string, std::vector<int32_t>> fine_sides;
    if (!base->side_sets.empty()) {
      std::map<std::pair<int32_t, int32_t>, int32_t> mid;
      for (int64_t k = 0; k < n->num_new; ++k) {
        const int32_t a = r.edge_parents[2 * k], b = r.edge_parents[2 * k + 1];
        mid[{std::min(a, b), std::max(a, b)}] = (int32_t)(n->num_old + k);
      }
      auto m = [&](int32_t a, int32_t b) { return mid.at({std::min(a, b), std::max(a, b)}); };
      for (auto const& kv : base->side_sets) {
        std::vector<int32_t>& out = fine_sides[kv.first];
        for (int32_t side : kv.second) {
          const int64_t e = side / 4;
          const int32_t* c = &base->mesh.conn[e * 4];
          const int32_t a = c[kTetFace[side % 4][0]], b = c[kTetFace[side % 4][1]], d = c[kTetFace[side % 4][2]];
          std::array<std::array<int32_t, 3>, 4> sub = {{{a, m(a, b), m(d, a)}, {b, m(a, b), m(b, d)},
                                                        {d, m(b, d), m(d, a)}, {m(a, b), m(b, d), m(d, a)}}};
          for (auto& t : sub) std::sort(t.begin(), t.end());
          for (int64_t ch = 8 * e; ch < 8 * e + 8; ++ch) {
            const int32_t* cc = &r.fine.conn[ch * 4];
            for (int f = 0; f < 4; ++f) {
              std::array<int32_t, 3> t = {cc[kTetFace[f][0]], cc[kTetFace[f][1]], cc[kTetFace[f][2]]};
              std::sort(t.begin(), t.end());
              if (std::find(sub.begin(), sub.end(), t) != sub.end()) out.push_back((int32_t)(4 * ch + f));
            }
          }
        }
        std::sort(out.begin(), out.end());
        require(out.size() == 4 * kv.second.size(), "nested refinement: a side did not split into four");
      }
    }
    n->fine = new_disc(std::move(r.fine));
    n->fine->side_sets = std::move(fine_sides);
    *out = n.release();
    return GOAL_OK;
  });
}

int goal_nested_destroy(goal_nested n) {
  return guard([&]() -> int {
    if (n) {
      sync();
      delete n->fine;
      delete n;
    }
    return GOAL_OK;
  });
}

goal_disc goal_nested_disc(goal_nested n) { return n ? n->fine : nullptr; }

int goal_nested_get_maps(goal_nested n, int32_t* parent, int32_t* edge_parents, int64_t* num_new) {
  return guard([&]() -> int {
    require(n != nullptr, "null nested");
    if (parent) std::copy(n->parent_host.begin(), n->parent_host.end(), parent);
    if (edge_parents) std::copy(n->edge_parents_host.begin(), n->edge_parents_host.end(), edge_parents);
    if (num_new) *num_new = n->num_new;
    return GOAL_OK;
  });
}

int goal_nested_transfer_soln(goal_nested n) {
  return guard([&]() -> int {
    require(n != nullptr, "null nested");
    HIPCHK(goal::launch_copy(n->fine->x.p, n->base->x.p, 4 * n->num_old, stream()));
    HIPCHK(goal::launch_set_coarse(n->fine->x.p, n->edge_parents.p, n->num_old, n->num_new, stream()));
    return GOAL_OK;
  });
}

int goal_nested_transfer_states(goal_nested n, goal_states base, goal_states fine) {
  return guard([&]() -> int {
    require(n && base && fine, "bad arguments");
    require(base->d == n->base && fine->d == n->fine, "states belong to other discs");
    require(base->model == fine->model && base->nq == fine->nq, "states differ in model or quadrature");
    const int64_t nf = n->fine->mesh.num_elems;
    HIPCHK(goal::launch_transfer_state(base->sigma.p, fine->sigma.p, 9, nf, fine->nq, base->npts, stream()));
    if (base->model == GOAL_MODEL_J2) {
      HIPCHK(goal::launch_transfer_state(base->Fp.p, fine->Fp_old.p, 9, nf, fine->nq, base->npts, stream()));
      HIPCHK(goal::launch_transfer_state(base->Fp.p, fine->Fp.p, 9, nf, fine->nq, base->npts, stream()));
      HIPCHK(goal::launch_transfer_state(base->eqps.p, fine->eqps_old.p, 1, nf, fine->nq, base->npts, stream()));
      HIPCHK(goal::launch_transfer_state(base->eqps.p, fine->eqps.p, 1, nf, fine->nq, base->npts, stream()));
    }
    return GOAL_OK;
  });
}

int goal_dwr_set_coarse(goal_nested n, double* field) {
  return guard([&]() -> int {
    require(n && field, "bad arguments");
    HIPCHK(goal::launch_set_coarse(field, n->edge_parents.p, n->num_old, n->num_new, stream()));
    return GOAL_OK;
  });
}

int goal_dwr_subtract(const double* fine, const double* coarse, double* out, int64_t n) {
  return guard([&]() -> int {
    require(fine && coarse && out && n >= 0, "bad arguments");
    HIPCHK(goal::launch_subtract(fine, coarse, out, n, stream()));
    return GOAL_OK;
  });
}

int goal_dwr_compute_error(goal_disc d, const double* e, double* eta) {
  return guard([&]() -> int {
    require(d && e && eta, "bad arguments");
    HIPCHK(goal::launch_compute_error(d->mesh.order, d->conn.p, d->mesh.num_elems, e, eta, stream()));
    return GOAL_OK;
  });
}

int goal_dwr_sum_contribs(goal_disc d, const double* e, double* bound) {
  return guard([&]() -> int {
    require(d && e && bound, "bad arguments");
    HIPCHK(goal::launch_sum_contribs(d->vertex_nodes.p, d->vertex_nodes.n, e, reduce_scratch(), result_slot(2), stream()));
    HIPCHK(hipMemcpyAsync(bound, result_slot(2), sizeof(double), hipMemcpyDeviceToHost, stream()));
    sync();
    return GOAL_OK;
  });
}

int goal_dwr_set_error(goal_nested n, const double* eta_child, double* eta_base) {
  return guard([&]() -> int {
    require(n && eta_child && eta_base, "bad arguments");
    HIPCHK(goal::launch_set_error(eta_child, n->base->mesh.num_elems, eta_base, stream()));
    return GOAL_OK;
  });
}

}  // extern "C"
