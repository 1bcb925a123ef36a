// goal_amd C++ surface: the reference's evaluator / assembly / driver classes
// re-expressed over the C ABI of goal_amd.h, so that code written against
// goal::Integrator, goal::assemble(), goal::Mechanics, goal::Primal,
// goal::Functional, goal::NestedAdjoint and goal::compute_error() keeps its
// shape while every element loop runs as fused gfx950 kernels.
//
// Header-only; link with libgoal_amd.so.  Errors throw goal::Error (the
// reference calls abort() via goal::fail, src/goal_control.cpp:91-99).
//
// What changes relative to the reference:
//   * The per-element hooks (gather, in_elem, at_point, out_elem, scatter) of
//     the built-in evaluators do no host work: assemble() reads the chain the
//     evaluators describe and runs it on the device in one launch pair.  The
//     per-element hooks of user-written integrators are not called.
//   * FADT is a tag type: Jacobian chains use the hand-derived analytic
//     tangent, not dual numbers.
//   * ParameterList input is replaced by plain structs (MechanicsParams,
//     PrimalParams, DirichletBC, TractionBC).
#pragma once

#include <cmath>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "goal_amd.h"

namespace goal {

// ---------------------------------------------------------------- basics
using ST = double;  // src/goal_scalar_types.hpp:8
struct FADT {};     // Jacobian chain tag (src/goal_scalar_types.hpp:9)

enum EvalModes { NONE, PRIMAL, ADJOINT };  // src/goal_eval_modes.hpp:6

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int status) {
  if (status != GOAL_OK) throw Error(status, goal_last_error());
}

// goal::eval (src/goal_control.cpp:107-123)
inline double eval(std::string const& v, double x, double y, double z, double t) {
  double out = 0.0;
  check(goal_eval(v.c_str(), x, y, z, t, &out));
  return out;
}

// ------------------------------------------------------------ disc / data
// Disc (src/goal_disc.hpp:27-93): mesh, node-major numbering, CSR graph.
class Disc {
 public:
  Disc(int order, int64_t num_nodes, const double* coords, int64_t num_elems, const int32_t* conn,
       const int32_t* elem_set = nullptr) {
    check(goal_disc_create(order, num_nodes, coords, num_elems, conn, elem_set, &h_));
    read_sizes();
  }
  static std::unique_ptr<Disc> box(int nx, int ny, int nz, double lx, double ly, double lz, int order,
                                   double jitter = 0.0, uint64_t seed = 0) {
    goal_disc h = nullptr;
    check(goal_disc_create_box(nx, ny, nz, lx, ly, lz, order, jitter, seed, &h));
    return std::unique_ptr<Disc>(new Disc(h, true));
  }
  // Disc(ParameterList) with "geom file", "mesh file", "assoc file"
  // (src/goal_disc.cpp:101-120): goal's mesh inputs with their node, side
  // and element sets; order 2 adds edge-midpoint nodes.
  static std::unique_ptr<Disc> load(std::string const& geom_file, std::string const& mesh_file,
                                    std::string const& assoc_file, int order = 1) {
    goal_mesh_file f = nullptr;
    check(goal_mesh_file_read(geom_file.c_str(), mesh_file.c_str(), assoc_file.c_str(), order, &f));
    goal_disc h = nullptr;
    const int st = goal_disc_create_from_file(f, &h);
    goal_mesh_file_destroy(f);
    check(st);
    return std::unique_ptr<Disc>(new Disc(h, true));
  }
  ~Disc() {
    if (own_ && h_) goal_disc_destroy(h_);
  }
  Disc(const Disc&) = delete;
  Disc& operator=(const Disc&) = delete;

  goal_disc handle() const { return h_; }
  int get_num_dims() const { return 3; }
  int get_num_eqs() const { return 4; }
  int64_t get_num_nodes() const { return sizes_[0]; }
  int64_t get_num_elems() const { return sizes_[1]; }
  int64_t get_num_dofs() const { return sizes_[2]; }
  int64_t get_nnz() const { return sizes_[3]; }
  int get_num_elem_sets() const { return (int)sizes_[7]; }
  int get_order() const { return (int)sizes_[6]; }
  // get_lid(node, eq) (src/goal_disc.cpp:209-212)
  static int get_lid(int node, int eq) { return node * 4 + eq; }
  std::vector<int32_t> get_nodes(std::string const& set) const {
    int64_t n = 0;
    check(goal_disc_node_set_size(h_, set.c_str(), &n));
    std::vector<int32_t> out((size_t)n);
    if (n) check(goal_disc_get_node_set(h_, set.c_str(), out.data()));
    return out;
  }
  void set_soln(std::vector<double> const& x) { check(goal_disc_set_soln(h_, x.data())); }
  std::vector<double> get_soln() const {
    std::vector<double> x((size_t)get_num_dofs());
    check(goal_disc_get_soln(h_, x.data()));
    return x;
  }
  // Disc::add_soln (src/goal_disc.cpp:398-422)
  void add_soln(std::vector<double> const& du) { check(goal_disc_add_soln(h_, du.data())); }
  void add_soln_dev(const double* du_dev) { check(goal_disc_add_soln_dev(h_, du_dev)); }

 private:
  friend class Nested;
  Disc(goal_disc h, bool own) : h_(h), own_(own) { read_sizes(); }
  void read_sizes() { check(goal_disc_sizes(h_, sizes_)); }
  goal_disc h_ = nullptr;
  bool own_ = true;
  int64_t sizes_[9] = {};
};

// SolInfo (src/goal_sol_info.hpp:12-39).  On one GPU owned == ghost, so the
// gather_* (Export ADD) calls and fill states are no-ops.
class SolInfo {
 public:
  explicit SolInfo(Disc* d) : disc_(d) { check(goal_solinfo_create(d->handle(), &h_)); }
  ~SolInfo() {
    if (h_) goal_solinfo_destroy(h_);
  }
  SolInfo(const SolInfo&) = delete;
  SolInfo& operator=(const SolInfo&) = delete;
  Disc* get_disc() { return disc_; }
  goal_solinfo handle() const { return h_; }
  void zero_R() { check(goal_solinfo_zero_R(h_)); }
  void zero_dMdu() { check(goal_solinfo_zero_dMdu(h_)); }
  void zero_dRdu() { check(goal_solinfo_zero_dRdu(h_)); }
  void zero_all() { check(goal_solinfo_zero_all(h_)); }
  void gather_R() {}
  void gather_dMdu() {}
  void gather_dRdu() {}
  void gather_all() {}
  void resume_fill() {}
  void complete_fill() {}
  double norm_R() {
    double n = 0.0;
    check(goal_solinfo_norm_R(h_, &n));
    return n;
  }
  std::vector<double> R() { return get(0); }
  std::vector<double> dMdu() { return get(1); }
  std::vector<double> dRdu_values() { return get(2); }
  void dev(double** R, double** dMdu, double** vals) { check(goal_solinfo_dev(h_, R, dMdu, vals)); }

 private:
  std::vector<double> get(int which) {
    std::vector<double> v((size_t)(which == 2 ? disc_->get_nnz() : disc_->get_num_dofs()));
    check(goal_solinfo_get(h_, which == 0 ? v.data() : nullptr, which == 1 ? v.data() : nullptr,
                           which == 2 ? v.data() : nullptr));
    return v;
  }
  Disc* disc_;
  goal_solinfo h_ = nullptr;
};

// ---------------------------------------------------------- integrators
// Integrator (src/goal_integrator.hpp:17-33).  Element and point arguments are
// opaque here: the device runs the element loop.
struct MeshElement;
struct Vector3 {
  double x[3];
};

class Integrator {
 public:
  Integrator() {}
  virtual ~Integrator() {}
  std::string const& get_name() { return name; }
  virtual void set_time(double, double) {}
  virtual void pre_process(SolInfo*) {}
  virtual void set_elem_set(int) {}
  virtual void gather(MeshElement*) {}
  virtual void in_elem(MeshElement*) {}
  virtual void at_point(Vector3 const&, double, double) {}
  virtual void out_elem() {}
  virtual void scatter(SolInfo*) {}
  virtual void post_process(SolInfo*) {}

 protected:
  std::string name;
};

using Evaluators = std::vector<std::shared_ptr<Integrator>>;

// find_evaluator / set_time (src/goal_assembly.cpp:12-22)
inline std::shared_ptr<Integrator> find_evaluator(std::string const& n, Evaluators const& E) {
  for (auto const& e : E)
    if (e->get_name() == n) return e;
  return nullptr;
}
inline void set_time(Evaluators& E, double t_now, double t_old) {
  for (auto& e : E) e->set_time(t_now, t_old);
}

class Mechanics;

// Built-in evaluators: descriptors of one fused chain.
namespace chain {
enum Role { FIELD_U, FIELD_P, WEIGHT_U, WEIGHT_P, WEIGHT_PC, KINEMATICS, MODEL, MIXED, MRESIDUAL, PRESIDUAL, STAB, QOI };
class Node : public Integrator {
 public:
  Node(Role r, std::string n) : role(r) { name = std::move(n); }
  Role role;
  bool fad = false;           // FADT chain
  int mode = PRIMAL;          // Displacement / Pressure scatter mode
  bool save = false;          // Model save_state
  Mechanics* mech = nullptr;  // Model
  const double* z_dev = nullptr;  // adjoint weights: nodal field (u,p per node)
  int qoi_kind = -1, qoi_set = -1;
  double qoi_value = 0.0;
};
}  // namespace chain

// Displacement<T>, Pressure<T> (src/goal_displacement.cpp:11-23,98-112,
// src/goal_pressure.cpp:11-24,95-110)
template <typename T>
std::shared_ptr<Integrator> Displacement(int mode) {
  auto n = std::make_shared<chain::Node>(chain::FIELD_U, "u");
  n->fad = std::is_same<T, FADT>::value;
  n->mode = mode;
  return n;
}
template <typename T>
std::shared_ptr<Integrator> Pressure(int mode) {
  auto n = std::make_shared<chain::Node>(chain::FIELD_P, "p");
  n->fad = std::is_same<T, FADT>::value;
  n->mode = mode;
  return n;
}
// VectorWeight / ScalarWeight (src/goal_vector_weight.cpp:6-11, src/goal_scalar_weight.cpp:6-11)
inline std::shared_ptr<Integrator> VectorWeight() { return std::make_shared<chain::Node>(chain::WEIGHT_U, "uw"); }
inline std::shared_ptr<Integrator> ScalarWeight() { return std::make_shared<chain::Node>(chain::WEIGHT_P, "pw"); }
// DisplacementAdjoint / PressureAdjoint (src/goal_displacement_adjoint.cpp:9-16,
// src/goal_pressure_adjoint.cpp:9-17): weights z N_a from a nodal field z
// (device pointer, 4 values per node; the u part or the p part is used).
inline std::shared_ptr<Integrator> DisplacementAdjoint(const double* z_dev) {
  auto n = std::make_shared<chain::Node>(chain::WEIGHT_U, "uw");
  n->z_dev = z_dev;
  return n;
}
inline std::shared_ptr<Integrator> PressureAdjoint(const double* z_dev, std::string const& nm) {
  auto n = std::make_shared<chain::Node>(nm == "pwc" ? chain::WEIGHT_PC : chain::WEIGHT_P, nm);
  n->z_dev = z_dev;
  return n;
}
// AvgDisp / AvgVM (src/goal_avg_disp.cpp:9-14, src/goal_avg_vm.cpp:22-30)
class QoI : public chain::Node {
 public:
  QoI(int kind, int set, std::string const& nm) : chain::Node(chain::QOI, nm) {
    qoi_kind = kind;
    qoi_set = set;
  }
  ST const& get_qoi_value() const { return qoi_value; }
};

// ------------------------------------------------------------ mechanics
struct MechanicsParams {
  std::string model = "neohookean";   // "elastic" | "neohookean" | "J2"
  bool stabilization = true;          // src/goal_mechanics.cpp:33
  std::vector<goal_material> materials;  // one per element set
  int quad_order = 1;                 // 1: the reference's one point; 2: four points
  int tangent = 64;                   // 32: FP32 element matrices (FP64 residual)
};

inline int model_id(std::string const& m) {
  if (m == "elastic") return GOAL_MODEL_ELASTIC;
  if (m == "neohookean") return GOAL_MODEL_NEOHOOKEAN;
  if (m == "J2") return GOAL_MODEL_J2;
  throw Error(GOAL_ERR_INVALID, "unknown model: " + m);
}

// Mechanics (src/goal_mechanics.cpp:41-95): owns the states and the device
// material table; builds the evaluator chains.
class Mechanics {
 public:
  Mechanics(MechanicsParams const& p, Disc* d) : params_(p), disc_(d) {
    const int model = model_id(p.model);
    std::vector<goal_material> mats = p.materials;
    for (auto& m : mats) m.model = model;
    if (mats.empty()) throw Error(GOAL_ERR_INVALID, "mechanics: no materials");
    check(goal_states_create(d->handle(), model, p.quad_order, &states_));
    check(goal_mech_create(d->handle(), states_, (int)mats.size(), mats.data(), p.stabilization ? 1 : 0,
                           p.quad_order, &mech_));
    check(goal_mech_set_option(mech_, "tangent", p.tangent));
  }
  ~Mechanics() {
    if (mech_) goal_mech_destroy(mech_);
    if (states_) goal_states_destroy(states_);
  }
  Mechanics(const Mechanics&) = delete;
  Mechanics& operator=(const Mechanics&) = delete;
  Disc* get_disc() { return disc_; }
  goal_mech handle() const { return mech_; }
  goal_states states() const { return states_; }
  MechanicsParams const& params() const { return params_; }
  // States::update (src/goal_states.cpp:130-141)
  void update_states() { check(goal_states_update(states_)); }

  // build_resid<T> (src/goal_mechanics.cpp:97-146)
  template <typename T>
  void build_resid(Evaluators& E, bool save) {
    const bool fad = std::is_same<T, FADT>::value;
    auto add = [&](chain::Role r, const char* n) {
      auto node = std::make_shared<chain::Node>(r, n);
      node->fad = fad;
      node->mech = this;
      node->save = save;
      E.push_back(node);
    };
    add(chain::KINEMATICS, "kinematics");
    add(chain::MODEL, "model");
    add(chain::MIXED, "mixed");
    add(chain::MRESIDUAL, "mresidual");
    add(chain::PRESIDUAL, "pressure");
    if (params_.stabilization) add(chain::STAB, "stabilization");
  }
  // build_functional<T> (src/goal_mechanics.cpp:148-167): "avg disp" or
  // "avg vm" over element set `set`.
  template <typename T>
  void build_functional(std::string const& type, int set, Evaluators& E) {
    int kind = type == "avg disp" ? GOAL_QOI_AVG_DISP : type == "avg vm" ? GOAL_QOI_AVG_VM : -1;
    if (kind < 0) throw Error(GOAL_ERR_INVALID, "unknown functional type: " + type);
    auto q = std::make_shared<QoI>(kind, set, type);
    q->fad = std::is_same<T, FADT>::value;
    q->mech = this;
    E.push_back(q);
  }
  // build_error (src/goal_mechanics.cpp:169-218)
  void build_error(Evaluators& E) { build_resid<ST>(E, false); }

 private:
  MechanicsParams params_;
  Disc* disc_;
  goal_states states_ = nullptr;
  goal_mech mech_ = nullptr;
};

// assemble (src/goal_assembly.cpp:65-88): pre_process and set_elem_set hooks,
// then the chain the built-in evaluators describe, then post_process.
inline void assemble(Evaluators const& E, SolInfo* s) {
  for (auto const& e : E) e->pre_process(s);
  const int nsets = s->get_disc()->get_num_elem_sets();
  for (int es = 0; es < nsets; ++es)
    for (auto const& e : E) e->set_elem_set(es);
  auto node = [&](chain::Role r) -> chain::Node* {
    for (auto const& e : E) {
      auto* n = dynamic_cast<chain::Node*>(e.get());
      if (n && n->role == r) return n;
    }
    return nullptr;
  };
  chain::Node* u = node(chain::FIELD_U);
  chain::Node* model = node(chain::MODEL);
  chain::Node* mres = node(chain::MRESIDUAL);
  chain::Node* uw = node(chain::WEIGHT_U);
  chain::Node* pw = node(chain::WEIGHT_P);
  chain::Node* pwc = node(chain::WEIGHT_PC);
  chain::Node* qoi = node(chain::QOI);
  if (model && mres && u) {
    goal_mech m = model->mech->handle();
    check(goal_mech_set_option(m, "stabilization", node(chain::STAB) ? 1 : 0));
    if (uw && uw->z_dev) {
      // error chain: weights z_diff (uw, pw) and z_coarse (pwc)
      if (!pw || !pwc || !pw->z_dev || !pwc->z_dev)
        throw Error(GOAL_ERR_INVALID, "assemble: adjoint weights need uw, pw and pwc fields");
      if (uw->z_dev != pw->z_dev) throw Error(GOAL_ERR_INVALID, "assemble: uw and pw must share one field");
      check(goal_asm_localize(m, s->handle(), uw->z_dev, pwc->z_dev));
    } else if (u->mode != NONE) {
      if (model->fad)
        check(goal_asm_jacobian(m, s->handle(), u->mode == ADJOINT ? 1 : 0, model->save ? 1 : 0));
      else
        check(goal_asm_residual(m, s->handle(), model->save ? 1 : 0));
    }
    check(goal_mech_set_option(m, "stabilization", model->mech->params().stabilization ? 1 : 0));
  }
  if (qoi) {
    check(goal_asm_qoi(qoi->mech->handle(), qoi->fad ? s->handle() : nullptr, qoi->qoi_kind, qoi->qoi_set,
                       &qoi->qoi_value, nullptr));
  }
  for (auto const& e : E) e->post_process(s);
}

// ------------------------------------------------------------ Dirichlet
// One entry [component, node set, expression] (src/goal_dbcs.cpp:39-99).
struct DirichletBC {
  int comp;
  std::string set;
  std::string value;
};
inline void set_resid_dbcs(std::vector<DirichletBC> const& bcs, SolInfo* s, double t) {
  for (auto const& b : bcs)
    check(goal_dbc_apply(s->get_disc()->handle(), s->handle(), b.comp, b.set.c_str(), b.value.c_str(), t, 0));
}
inline void set_jac_dbcs(std::vector<DirichletBC> const& bcs, SolInfo* s, double t) {
  for (auto const& b : bcs)
    check(goal_dbc_apply(s->get_disc()->handle(), s->handle(), b.comp, b.set.c_str(), b.value.c_str(), t, 1));
}

// ------------------------------------------------------------ tractions
// One entry [side set, T_x, T_y, T_z] of the "traction bcs" list
// (src/goal_tbcs.cpp:30-66).  A side set name with no side set given to the
// disc stands for the boundary faces of the node set of that name.
struct TractionBC {
  std::string set;
  std::string value[3];
};
// set_tbcs (src/goal_tbcs.cpp:68-79).  weight_dev: null for the primal and
// adjoint chains (weight N_a), the z_diff DOF vector for the error chain.
inline void set_tbcs(std::vector<TractionBC> const& bcs, SolInfo* s, double t, const double* weight_dev = nullptr) {
  for (auto const& b : bcs)
    check(goal_tbc_apply(s->get_disc()->handle(), s->handle(), b.set.c_str(), b.value[0].c_str(),
                         b.value[1].c_str(), b.value[2].c_str(), t, weight_dev));
}

// Linear solve callback standing in for goal::solve (Belos GMRES + MueLu,
// src/goal_linear_solve.cpp:72-90): solve A x = b with A the device CSR of
// s (row_ptr / cols from the disc, values from s), b and x device vectors.
// Return 0 on success.
using LinearSolve = std::function<int(Disc* d, SolInfo* s, const double* b_dev, double* x_dev)>;

// The library's device solve as a LinearSolve: GMRES(restart) preconditioned
// by aggregation multigrid over the 4-equation node blocks (as goal's MueLu
// with "number of equations" 4; GOAL_PRECOND_BLOCK_JACOBI for the inverse
// diagonal blocks alone) from a zero initial guess, the parameters of goal's
// "linear solve" list ("tolerance", "max iters", "krylov size";
// src/goal_linear_solve.cpp:22-50).  Like the reference it does not fail when
// the iteration limit is reached.
struct GmresSolve {
  double tolerance = 1e-10;
  int max_iters = 1000;
  int krylov_size = 100;
  int preconditioner = GOAL_PRECOND_AMG;
  int operator()(Disc* d, SolInfo* s, const double* b_dev, double* x_dev) const {
    int st = goal_device_fill(x_dev, d->get_num_dofs(), 0.0);
    if (st != GOAL_OK) return st;
    return goal_linear_solve_pc(s->handle(), b_dev, x_dev, tolerance, max_iters, krylov_size, preconditioner,
                                nullptr, nullptr);
  }
};

struct PrimalParams {
  std::vector<DirichletBC> dirichlet;
  std::vector<TractionBC> traction;  // "traction bcs"
  int max_iters = 5;       // "nonlinear max iters"
  double tolerance = 1e-8; // "nonlinear tolerance"
};

// Primal (src/goal_primal.cpp:43-137)
class Primal {
 public:
  Primal(PrimalParams const& p, Mechanics* m, LinearSolve solve) : params_(p), mech_(m), solve_(std::move(solve)) {
    residual_.push_back(Displacement<ST>(PRIMAL));
    residual_.push_back(VectorWeight());
    residual_.push_back(Pressure<ST>(PRIMAL));
    residual_.push_back(ScalarWeight());
    jacobian_.push_back(Displacement<FADT>(PRIMAL));
    jacobian_.push_back(VectorWeight());
    jacobian_.push_back(Pressure<FADT>(PRIMAL));
    jacobian_.push_back(ScalarWeight());
    mech_->build_resid<ST>(residual_, true);
    mech_->build_resid<FADT>(jacobian_, true);
  }
  ~Primal() { destroy_data(); }
  void build_data() { sol_info_.reset(new SolInfo(mech_->get_disc())); }
  void destroy_data() { sol_info_.reset(); }
  SolInfo* get_sol_info() { return sol_info_.get(); }
  Mechanics* get_mech() { return mech_; }
  void compute_resid(double t_now, double t_old) {
    sol_info_->zero_R();
    set_time(residual_, t_now, t_old);
    assemble(residual_, sol_info_.get());
    set_tbcs(params_.traction, sol_info_.get(), t_now);
    sol_info_->gather_R();
    set_resid_dbcs(params_.dirichlet, sol_info_.get(), t_now);
  }
  void compute_jacob(double t_now, double t_old) {
    sol_info_->resume_fill();
    sol_info_->zero_all();
    set_time(jacobian_, t_now, t_old);
    assemble(jacobian_, sol_info_.get());
    set_tbcs(params_.traction, sol_info_.get(), t_now);
    sol_info_->gather_all();
    set_jac_dbcs(params_.dirichlet, sol_info_.get(), t_now);
    sol_info_->complete_fill();
  }
  // Newton: returns the iteration count; throws if not converged.
  int solve(double t_now, double t_old) {
    const int64_t n = mech_->get_disc()->get_num_dofs();
    double *du = nullptr, *b = nullptr;
    check(goal_device_alloc(8 * n, (void**)&du));
    check(goal_device_alloc(8 * n, (void**)&b));
    int iter = 1;
    bool converged = false;
    try {
      while (iter <= params_.max_iters && !converged) {
        compute_jacob(t_now, t_old);
        double* R = nullptr;
        sol_info_->dev(&R, nullptr, nullptr);
        check(goal_device_scale(b, R, n, -1.0));  // b = -R on the device
        if (solve_(mech_->get_disc(), sol_info_.get(), b, du) != 0)
          throw Error(GOAL_ERR_INVALID, "linear solve failed");
        mech_->get_disc()->add_soln_dev(du);
        compute_resid(t_now, t_old);
        last_norm_ = sol_info_->norm_R();
        if (last_norm_ < params_.tolerance) converged = true;
        ++iter;
      }
    } catch (...) {
      goal_device_free(du);
      goal_device_free(b);
      throw;
    }
    goal_device_free(du);
    goal_device_free(b);
    if (!converged)
      throw Error(GOAL_ERR_INVALID, "newton's method failed in " + std::to_string(params_.max_iters) + " iterations");
    return iter - 1;
  }
  double last_norm() const { return last_norm_; }

 private:
  PrimalParams params_;
  Mechanics* mech_;
  LinearSolve solve_;
  Evaluators residual_, jacobian_;
  std::unique_ptr<SolInfo> sol_info_;
  double last_norm_ = 0.0;
};

// Functional (src/goal_functional.cpp:34-70): the ST chain with NONE
// scatter plus the QoI.
class Functional {
 public:
  Functional(std::string const& type, int set, Primal* pr) : primal_(pr) {
    E_.push_back(Displacement<ST>(NONE));
    E_.push_back(VectorWeight());
    E_.push_back(Pressure<ST>(NONE));
    E_.push_back(ScalarWeight());
    pr->get_mech()->build_resid<ST>(E_, false);
    pr->get_mech()->build_functional<ST>(type, set, E_);
    qoi_ = std::static_pointer_cast<QoI>(E_.back());
  }
  void compute(double t_now, double t_old) {
    set_time(E_, t_now, t_old);
    assemble(E_, primal_->get_sol_info());
  }
  double get_value() const { return qoi_->get_qoi_value(); }

 private:
  Primal* primal_;
  Evaluators E_;
  std::shared_ptr<QoI> qoi_;
};

// compute_error / sum_contribs (src/goal_error.hpp:10-11) on a nodal field e
// (device, 4 values per node: u_error then p_error).
inline std::vector<double> compute_error(Disc* d, const double* e_dev) {
  double* eta = nullptr;
  check(goal_device_alloc(8 * d->get_num_elems(), (void**)&eta));
  std::vector<double> out((size_t)d->get_num_elems());
  int st = goal_dwr_compute_error(d->handle(), e_dev, eta);
  if (st == GOAL_OK) st = goal_memcpy_d2h(out.data(), eta, 8 * d->get_num_elems());
  goal_device_free(eta);
  check(st);
  return out;
}
inline double sum_contribs(Disc* d, const double* e_dev) {
  double b = 0.0;
  check(goal_dwr_sum_contribs(d->handle(), e_dev, &b));
  return b;
}

// Nested (src/goal_nested.hpp:17-42), UNIFORM mode.
class Nested {
 public:
  explicit Nested(Disc* base) {
    check(goal_nested_create(base->handle(), &h_));
    fine_.reset(new Disc(goal_nested_disc(h_), false));
  }
  ~Nested() {
    fine_.reset();
    if (h_) goal_nested_destroy(h_);
  }
  Disc* disc() { return fine_.get(); }
  goal_nested handle() const { return h_; }
  void transfer_soln() { check(goal_nested_transfer_soln(h_)); }
  void set_coarse(double* field_dev) { check(goal_dwr_set_coarse(h_, field_dev)); }

 private:
  goal_nested h_ = nullptr;
  std::unique_ptr<Disc> fine_;
};

// NestedAdjoint::run (src/goal_nested_adjoint.cpp:163-257): nested mesh,
// adjoint solve, z - I_H z, localized residual, indicators summed onto the
// base elements.  Returns eta per base element; bound and z.R are kept.
class NestedAdjoint {
 public:
  NestedAdjoint(Mechanics* base_mech, std::string const& qoi, int qoi_set, std::vector<DirichletBC> bcs,
                LinearSolve solve, std::vector<TractionBC> tbcs = {})
      : mp_(base_mech->params()), base_mech_(base_mech), qoi_(qoi), qoi_set_(qoi_set), bcs_(std::move(bcs)),
        tbcs_(std::move(tbcs)), base_(base_mech->get_disc()), solve_(std::move(solve)) {}

  std::vector<double> run(double t_now, double t_old) {
    Nested nested(base_);
    Disc* fine = nested.disc();
    nested.transfer_soln();
    Mechanics mech(mp_, fine);
    check(goal_nested_transfer_states(nested.handle(), base_mech_->states(), mech.states()));
    SolInfo si(fine);
    const int64_t n = fine->get_num_dofs();
    std::vector<double*> bufs;
    auto dalloc = [&](int64_t count) {
      double* p = nullptr;
      check(goal_device_alloc(8 * count, (void**)&p));
      bufs.push_back(p);
      return p;
    };
    try {
      double *z = dalloc(n), *zc = dalloc(n), *zd = dalloc(n), *eta_f = dalloc(fine->get_num_elems()),
             *eta_b = dalloc(base_->get_num_elems());
      // compute_adjoint: transposed Jacobian, dJ/du, Dirichlet rows
      Evaluators adj;
      adj.push_back(Displacement<FADT>(ADJOINT));
      adj.push_back(VectorWeight());
      adj.push_back(Pressure<FADT>(ADJOINT));
      adj.push_back(ScalarWeight());
      mech.build_resid<FADT>(adj, false);
      mech.build_functional<FADT>(qoi_, qoi_set_, adj);
      si.zero_all();
      set_time(adj, t_now, t_old);
      assemble(adj, &si);
      set_tbcs(tbcs_, &si, t_now);
      set_jac_dbcs(bcs_, &si, t_now);
      double *R = nullptr, *dMdu = nullptr;
      si.dev(&R, &dMdu, nullptr);
      if (solve_(fine, &si, dMdu, z) != 0) throw Error(GOAL_ERR_INVALID, "adjoint solve failed");
      check(goal_solinfo_dot_R(si.handle(), z, &zR_));
      check(goal_device_copy(zc, z, n));
      nested.set_coarse(zc);
      check(goal_dwr_subtract(z, zc, zd, n));
      // localize
      Evaluators err;
      err.push_back(Displacement<ST>(PRIMAL));
      err.push_back(DisplacementAdjoint(zd));
      err.push_back(Pressure<ST>(PRIMAL));
      err.push_back(PressureAdjoint(zd, "pw"));
      err.push_back(PressureAdjoint(zc, "pwc"));
      mech.build_error(err);
      si.zero_R();
      set_time(err, t_now, t_old);
      assemble(err, &si);
      set_tbcs(tbcs_, &si, t_now, zd);
      set_resid_dbcs(bcs_, &si, t_now);
      si.dev(&R, nullptr, nullptr);
      bound_ = sum_contribs(fine, R);
      check(goal_dwr_compute_error(fine->handle(), R, eta_f));
      check(goal_dwr_set_error(nested.handle(), eta_f, eta_b));
      std::vector<double> out((size_t)base_->get_num_elems());
      check(goal_memcpy_d2h(out.data(), eta_b, 8 * base_->get_num_elems()));
      for (double* p : bufs) goal_device_free(p);
      return out;
    } catch (...) {
      for (double* p : bufs) goal_device_free(p);
      throw;
    }
  }
  double bound() const { return bound_; }
  double z_dot_R() const { return zR_; }

 private:
  MechanicsParams mp_;
  Mechanics* base_mech_;
  std::string qoi_;
  int qoi_set_;
  std::vector<DirichletBC> bcs_;
  std::vector<TractionBC> tbcs_;
  Disc* base_;
  LinearSolve solve_;
  double bound_ = 0.0, zR_ = 0.0;
};

// get_iso_target_size (src/goal_size_field.cpp:142-151) on a linear disc:
// per-element indicators -> per-vertex target sizes for `target` elements.
inline std::vector<double> get_iso_target_size(Disc* d, std::vector<double> const& eta, int64_t target) {
  if (d->get_order() != 1) throw Error(GOAL_ERR_INVALID, "size field: linear meshes only");
  const int64_t nn = d->get_num_nodes(), ne = d->get_num_elems();
  if ((int64_t)eta.size() != ne) throw Error(GOAL_ERR_INVALID, "size field: one indicator per element");
  std::vector<double> coords((size_t)(3 * nn)), size((size_t)nn);
  std::vector<int32_t> conn((size_t)(4 * ne));
  check(goal_disc_get_mesh(d->handle(), coords.data(), conn.data(), nullptr));
  check(goal_size_field(nn, coords.data(), ne, conn.data(), eta.data(), target, 1, size.data()));
  return size;
}

// The mesh step of Solver::adapt (src/main_nested.cpp:108-126) without
// coarsening: a new disc refined toward `size` by longest-edge bisection.
// Children keep their parent's element set; the solution is carried over
// (new vertices take their edge's mean).  Node sets are not carried over.
struct Adapted {
  std::unique_ptr<Disc> disc;
  std::vector<int32_t> parent;  // element -> element of the input disc
  int64_t splits = 0;
};
inline Adapted adapt(Disc* d, std::vector<double> const& size, double split_ratio = 1.5, int max_passes = 8) {
  const int64_t nn = d->get_num_nodes(), ne = d->get_num_elems();
  std::vector<double> coords((size_t)(3 * nn));
  std::vector<int32_t> conn((size_t)(4 * ne)), es((size_t)ne);
  check(goal_disc_get_mesh(d->handle(), coords.data(), conn.data(), es.data()));
  goal_adapt a = nullptr;
  check(goal_adapt_create(nn, coords.data(), ne, conn.data(), size.data(), split_ratio, max_passes, &a));
  Adapted out;
  try {
    int64_t sz[5];
    check(goal_adapt_sizes(a, sz));
    std::vector<double> c2((size_t)(3 * sz[0]));
    std::vector<int32_t> t2((size_t)(4 * sz[1]));
    out.parent.resize((size_t)sz[1]);
    check(goal_adapt_get(a, c2.data(), t2.data(), out.parent.data(), nullptr));
    std::vector<int32_t> es2((size_t)sz[1]);
    for (int64_t e = 0; e < sz[1]; ++e) es2[e] = es[out.parent[e]];
    out.disc.reset(new Disc(1, sz[0], c2.data(), sz[1], t2.data(), es2.data()));
    std::vector<double> x2((size_t)(4 * sz[0]));
    const std::vector<double> x = d->get_soln();
    check(goal_adapt_transfer(a, 4, x.data(), x2.data()));
    out.disc->set_soln(x2);
    out.splits = sz[3];
  } catch (...) {
    goal_adapt_destroy(a);
    throw;
  }
  goal_adapt_destroy(a);
  return out;
}

// Learned stand-in for NestedAdjoint::run (no counterpart in the reference):
// per-element indicators eta_E = E * 10^net(features) from a trained network,
// or the nested adjoint's when more than ood_limit of the elements have
// features outside the training box.
class LearnedIndicator {
 public:
  LearnedIndicator(std::vector<float> const& params, std::vector<float> const& norms, double qoi_x,
                   double ood_limit = 0.01)
      : qoi_x_(qoi_x), ood_limit_(ood_limit) {
    if (params.size() != GOAL_MLP_PARAMS || norms.size() != GOAL_MLP_NORMS)
      throw Error(GOAL_ERR_INVALID, "learned indicator: parameter or normalization size");
    check(goal_mlp_create(&mlp_));
    check(goal_mlp_set(mlp_, params.data(), norms.data()));
  }
  ~LearnedIndicator() {
    if (mlp_) goal_mlp_destroy(mlp_);
  }
  LearnedIndicator(const LearnedIndicator&) = delete;
  LearnedIndicator& operator=(const LearnedIndicator&) = delete;

  std::vector<double> run(Mechanics* mech, NestedAdjoint* fallback, double t_now, double t_old) {
    Disc* d = mech->get_disc();
    const int64_t ne = d->get_num_elems();
    float *feat = nullptr, *y = nullptr;
    check(goal_device_alloc(4 * ne * GOAL_MLP_IN, (void**)&feat));
    if (goal_device_alloc(4 * ne, (void**)&y) != GOAL_OK) {
      goal_device_free(feat);
      throw Error(GOAL_ERR_ALLOC, "learned indicator: device allocation");
    }
    std::vector<float> yh((size_t)ne);
    int64_t outside = 0;
    int st = goal_learn_features(mech->handle(), qoi_x_, feat);
    if (st == GOAL_OK) st = goal_mlp_forward(mlp_, feat, ne, y, &outside);
    if (st == GOAL_OK) st = goal_memcpy_d2h(yh.data(), y, 4 * ne);
    goal_device_free(feat);
    goal_device_free(y);
    check(st);
    fell_back_ = outside > ood_limit_ * (double)ne;
    if (fell_back_) {
      if (!fallback) throw Error(GOAL_ERR_INVALID, "learned indicator: features out of range and no fallback");
      return fallback->run(t_now, t_old);
    }
    std::vector<int32_t> es((size_t)ne);
    std::vector<double> coords((size_t)(3 * d->get_num_nodes()));
    std::vector<int32_t> conn((size_t)(ne * (d->get_order() == 1 ? 4 : 10)));
    check(goal_disc_get_mesh(d->handle(), coords.data(), conn.data(), es.data()));
    std::vector<double> eta((size_t)ne);
    auto const& mats = mech->params().materials;
    for (int64_t e = 0; e < ne; ++e) eta[e] = mats.at((size_t)es[e]).E * std::pow(10.0, (double)yh[e]);
    return eta;
  }
  bool fell_back() const { return fell_back_; }

 private:
  goal_mlp mlp_ = nullptr;
  double qoi_x_, ood_limit_;
  bool fell_back_ = false;
};

}  // namespace goal
