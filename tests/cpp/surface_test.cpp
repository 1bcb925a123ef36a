// Drives goal_amd through its C++ surface (include/goal_amd.hpp) the way the
// reference's GoalPrimal / GoalNested drivers drive goal (src/main_primal.cpp,
// src/main_nested.cpp): Newton load steps, the functional, one nested-adjoint
// error estimate.  The linear solves use a dense LU, adequate for the small
// meshes this program is run on.  Prints one JSON object.
//
// usage: surface_test <order> <quad_order> <model> <n> <steps> [tangent bits: 64 | 32] [solver: lu | gmres]
//                     [extra: adapt | learned:<file> | known:<case>]
//   adapt            after the load steps (P1), adapt once with the nested indicators
//                    (get_iso_target_size, adapt) and solve the last step again there
//   learned:<file>   the learned indicator on the final state; <file> holds the
//                    network as raw float32: GOAL_MLP_PARAMS parameters, then
//                    GOAL_MLP_NORMS normalization values
//   known:<case>     the settings of one of the reference's regression inputs
//                    (unit cube, rollers on xmin/ymin/zmin, Newton tolerance
//                    1e-8 in at most 5 iterations): nh_uniaxial
//                    (example/primal/neohookean_uniaxial_3D.yaml), j2_uniaxial
//                    (J2_uniaxial_3D.yaml) or j2_traction (J2_traction_3D.yaml,
//                    a traction 1.0*t in y on ymax), in place of the default
//                    zmin/zmax compression; known:<case>:<dir> runs it on the
//                    mesh files <dir>/cube.dmg, <dir>/cube.smb, <dir>/cube.txt
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "goal_amd.hpp"

namespace {

// Dense LU with partial pivoting of the device CSR system A x = b.
int dense_solve(goal::Disc* d, goal::SolInfo* s, const double* b_dev, double* x_dev) {
  const int64_t n = d->get_num_dofs();
  std::vector<int64_t> rp(n + 1);
  std::vector<int32_t> cols(d->get_nnz());
  goal::check(goal_disc_get_graph(d->handle(), rp.data(), cols.data()));
  std::vector<double> vals = s->dRdu_values();
  std::vector<double> A((size_t)(n * n), 0.0), b(n);
  for (int64_t r = 0; r < n; ++r)
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) A[r * n + cols[k]] = vals[k];
  goal::check(goal_memcpy_d2h(b.data(), b_dev, 8 * n));
  std::vector<int64_t> piv(n);
  for (int64_t k = 0; k < n; ++k) {
    int64_t p = k;
    for (int64_t i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > std::fabs(A[p * n + k])) p = i;
    if (A[p * n + k] == 0.0) return 1;
    if (p != k) {
      for (int64_t j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
      std::swap(b[k], b[p]);
    }
    for (int64_t i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] / A[k * n + k];
      if (f == 0.0) continue;
      for (int64_t j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      b[i] -= f * b[k];
    }
  }
  std::vector<double> x(n);
  for (int64_t i = n - 1; i >= 0; --i) {
    double acc = b[i];
    for (int64_t j = i + 1; j < n; ++j) acc -= A[i * n + j] * x[j];
    x[i] = acc / A[i * n + i];
  }
  goal::check(goal_memcpy_h2d(x_dev, x.data(), 8 * n));
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s order quad model n steps\n", argv[0]);
    return 2;
  }
  const int order = std::atoi(argv[1]), quad = std::atoi(argv[2]), model = std::atoi(argv[3]);
  const int n = std::atoi(argv[4]), steps = std::atoi(argv[5]);
  const int tangent = argc > 6 ? std::atoi(argv[6]) : 64;
  const bool gmres = argc > 7 && std::string(argv[7]) == "gmres";
  std::string extra = argc > 8 ? argv[8] : "";
  try {
    std::unique_ptr<goal::Disc> disc;
    const size_t files = extra.rfind("known:", 0) == 0 ? extra.find(':', 6) : std::string::npos;
    if (files != std::string::npos) {
      const std::string dir = extra.substr(files + 1);
      disc = goal::Disc::load(dir + "/cube.dmg", dir + "/cube.smb", dir + "/cube.txt", order);
      extra = extra.substr(0, files);
    } else {
      disc = goal::Disc::box(n, n, n, 1.0, 1.0, 1.0, order, 0.1, 11);
    }
    goal::MechanicsParams mp;
    mp.model = model == 2 ? "J2" : model == 0 ? "elastic" : "neohookean";
    mp.quad_order = quad;
    mp.tangent = tangent;
    goal_material mat{};
    if (model == 2) mat = goal_material{GOAL_MODEL_J2, 200.0, 0.3, 2.0, 0.25, 1.0};
    else mat = goal_material{GOAL_MODEL_NEOHOOKEAN, 1000.0, 0.25, 0.0, 0.0, 1.0};
    goal::PrimalParams pp;
    pp.dirichlet = {{0, "zmin", "0.0"}, {1, "zmin", "0.0"}, {2, "zmin", "0.0"}, {2, "zmax", "0.01*t"}};
    pp.max_iters = 10;
    pp.tolerance = 1e-10;
    if (extra.rfind("known:", 0) == 0) {
      const std::string kc = extra.substr(6);
      if (kc != "nh_uniaxial" && kc != "j2_uniaxial" && kc != "j2_traction") {
        std::fprintf(stderr, "unknown case %s\n", kc.c_str());
        return 2;
      }
      if (kc != "nh_uniaxial") mat = goal_material{GOAL_MODEL_J2, 1000.0, 0.25, 100.0, 10.0, 1.0};
      mp.model = kc == "nh_uniaxial" ? "neohookean" : "J2";
      pp.dirichlet = {{0, "xmin", "0.0"}, {1, "ymin", "0.0"}, {2, "zmin", "0.0"}};
      if (kc == "j2_traction")
        pp.traction = {{"ymax", {"0.0", "1.0*t", "0.0"}}};
      else
        pp.dirichlet.push_back({0, "xmax", "0.01*t"});
      pp.max_iters = 5;
      pp.tolerance = 1e-8;
    }
    mp.materials = {mat};
    goal::Mechanics mech(mp, disc.get());
    goal::GmresSolve gm;
    gm.tolerance = 1e-12;
    gm.max_iters = 5000;
    const goal::LinearSolve solve = gmres ? goal::LinearSolve(gm) : goal::LinearSolve(dense_solve);
    goal::Primal primal(pp, &mech, solve);
    primal.build_data();
    goal::Functional functional("avg disp", -1, &primal);
    std::printf("{\"iters\": [");
    double t_old = 0.0;
    double J = 0.0;
    for (int step = 1; step <= steps; ++step) {
      const double t = step;
      const int it = primal.solve(t, t_old);
      functional.compute(t, t_old);
      J = functional.get_value();
      std::printf("%s%d", step > 1 ? ", " : "", it);
      if (step < steps) mech.update_states();
      t_old = t;
    }
    std::printf("], \"J\": %.17g, \"norm\": %.6e", J, primal.last_norm());
    if (order == 1) {
      goal::NestedAdjoint adj(&mech, "avg disp", -1, pp.dirichlet, solve, pp.traction);
      std::vector<double> eta = adj.run(t_old, t_old - 1.0);
      double s = 0.0;
      for (double e : eta) s += e;
      std::printf(", \"eta_sum\": %.17g, \"bound\": %.17g, \"zR\": %.17g", s, adj.bound(), adj.z_dot_R());
      if (extra.rfind("learned:", 0) == 0) {
        std::vector<float> buf(GOAL_MLP_PARAMS + GOAL_MLP_NORMS);
        FILE* f = std::fopen(extra.substr(8).c_str(), "rb");
        if (!f || std::fread(buf.data(), sizeof(float), buf.size(), f) != buf.size()) {
          std::fprintf(stderr, "cannot read %s\n", extra.substr(8).c_str());
          return 2;
        }
        std::fclose(f);
        goal::LearnedIndicator learned(std::vector<float>(buf.begin(), buf.begin() + GOAL_MLP_PARAMS),
                                       std::vector<float>(buf.begin() + GOAL_MLP_PARAMS, buf.end()), 0.25);
        std::vector<double> le = learned.run(&mech, &adj, t_old, t_old - 1.0);
        double ls = 0.0;
        for (double e : le) ls += e;
        std::printf(", \"learned_eta_sum\": %.17g, \"learned_fell_back\": %d", ls, learned.fell_back() ? 1 : 0);
      }
      if (extra == "adapt" && model != 2) {
        // Solver::adapt: size field for 3x the elements, refine, solve again
        const int64_t ne = disc->get_num_elems();
        std::vector<double> size = goal::get_iso_target_size(disc.get(), eta, 3 * ne);
        goal::Adapted next = goal::adapt(disc.get(), size);
        goal::Disc* d2 = next.disc.get();
        std::vector<double> c2((size_t)(3 * d2->get_num_nodes()));
        std::vector<int32_t> t2((size_t)(4 * d2->get_num_elems()));
        goal::check(goal_disc_get_mesh(d2->handle(), c2.data(), t2.data(), nullptr));
        for (const char* face : {"zmin", "zmax"}) {
          const double z = std::string(face) == "zmin" ? 0.0 : 1.0;
          std::vector<int32_t> nodes;
          for (int64_t k = 0; k < d2->get_num_nodes(); ++k)
            if (c2[3 * k + 2] == z) nodes.push_back((int32_t)k);
          goal::check(goal_disc_add_node_set(d2->handle(), face, (int64_t)nodes.size(), nodes.data()));
        }
        goal::Mechanics mech2(mp, d2);
        goal::Primal primal2(pp, &mech2, solve);
        primal2.build_data();
        const int it2 = primal2.solve(t_old, t_old - 1.0);
        goal::Functional f2("avg disp", -1, &primal2);
        f2.compute(t_old, t_old - 1.0);
        std::printf(", \"adapt\": {\"elems\": %lld, \"splits\": %lld, \"iters\": %d, \"J\": %.17g, \"norm\": %.6e}",
                    (long long)d2->get_num_elems(), (long long)next.splits, it2, f2.get_value(), primal2.last_norm());
      }
    }
    mech.update_states();
    std::printf("}\n");
  } catch (goal::Error const& e) {
    std::fprintf(stderr, "goal error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
