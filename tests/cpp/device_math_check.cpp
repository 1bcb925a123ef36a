// Host build of the element kernels' point math (goal_amd/csrc/goal_device.hpp:
// the J2 radial return and its directional derivatives, neo-Hookean, the
// stabilization flux) and of the element kernel's LDS index arithmetic, for
// checks that need no GPU:
//   * every directional derivative (J2 plastic and elastic branches,
//     neo-Hookean, stabilization) against central finite differences of the
//     point values it differentiates;
//   * J2 invariants: det Fp = 1 after the exponential update, eqps = eqps_old +
//     sqrt(2/3) dgam, consistency |dev sigma| J = sqrt(2/3) (Y + K eqps);
//   * every LDS index the phases A-E form for every thread of a workgroup,
//     for every element count 1..G of a group, inside its array.
// Built under AddressSanitizer and UndefinedBehaviorSanitizer, twice: with
// stack variables filled with a pattern and with zeros before first use.  The
// two builds print every computed value in hex; equal output means no value
// depends on a variable read before it is written.  Exit status 0: all
// checks passed.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "goal_device.hpp"

using namespace goal;
using namespace goal::dev;

namespace {

int g_fail = 0;
void check(bool ok, const char* what, double a, double b) {
  if (!ok) {
    if (g_fail < 20) std::fprintf(stderr, "FAIL %s: %.17g vs %.17g\n", what, a, b);
    ++g_fail;
  }
}

uint64_t g_rng = 0x2545F4914F6CDD1Dull;
double urand() {  // xorshift64*, uniform in [-1, 1)
  g_rng ^= g_rng >> 12;
  g_rng ^= g_rng << 25;
  g_rng ^= g_rng >> 27;
  return (double)((g_rng * 2685821657736338717ull) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

void print_hex(const double* v, int n) {
  for (int i = 0; i < n; ++i) {
    uint64_t u;
    std::memcpy(&u, &v[i], 8);
    std::printf("%016llx ", (unsigned long long)u);
  }
  std::printf("\n");
}

void make_qp(const double* F, double p, const double* g, QP& q) {
  for (int k = 0; k < 9; ++k) {
    q.F[k] = F[k];
    q.H[k] = F[k] - ((k % 4 == 0) ? 1.0 : 0.0);
  }
  q.J = det3(q.F);
  inv3(q.F, q.J, q.Fi);
  q.p = p;
  for (int j = 0; j < 3; ++j) q.g[j] = g[j];
}

MatParams mat(double E, double nu, double K, double Y) {
  MatParams m{};
  m.mu = E / (2.0 * (1.0 + nu));
  m.kappa = E / (3.0 * (1.0 - 2.0 * nu));
  m.lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  m.K = K;
  m.Y = Y;
  m.c0 = 1.0;
  return m;
}

// point values the tangent differentiates: P (9), r1, s3 (3)
template <typename Model>
void values(const Model& m, const Stab& st, const QP& q, double* out) {
  for (int k = 0; k < 9; ++k) out[k] = m.P[k];
  out[9] = m.r1(q);
  st.value(out + 10);
}

// Fp_old = exp(sym traceless A), det 1
void random_Fp(double amp, double* Fp) {
  double A[9];
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) A[3 * i + j] = A[3 * j + i] = amp * urand();
  const double tr = (A[0] + A[4] + A[8]) / 3.0;
  for (int i = 0; i < 3; ++i) A[4 * i] -= tr;
  expm3(A, Fp);
}

void check_j2(int trials, bool plastic_wanted) {
  const MatParams mp = mat(200.0, 0.3, 2.0, 0.25);
  const double sq23 = std::sqrt(2.0 / 3.0);
  int plastic_seen = 0;
  for (int t = 0; t < trials; ++t) {
    const double amp = plastic_wanted ? 0.02 : 2e-4;
    double F[9], g[3], Fp[9];
    for (int k = 0; k < 9; ++k) F[k] = ((k % 4 == 0) ? 1.0 : 0.0) + amp * urand();
    for (int j = 0; j < 3; ++j) g[j] = urand();
    random_Fp(plastic_wanted ? 0.01 : 0.0, Fp);
    const double eqps_old = plastic_wanted ? 0.01 * std::fabs(urand()) : 0.0;
    const double p = 0.3 * urand();
    QP q;
    make_qp(F, p, g, q);
    J2Model m;
    m.init(q, mp, Fp, eqps_old);
    Stab st;
    st.init(q, 0.01);
    plastic_seen += m.plastic ? 1 : 0;
    check(!m.failed(), "J2 return map converged", m.iters, 30);
    double v0[13];
    values(m, st, q, v0);
    print_hex(v0, 13);
    if (m.plastic) {
      double Fpn[9];
      m.new_Fp(Fp, Fpn);
      print_hex(Fpn, 9);
      check(std::fabs(det3(Fpn) - 1.0) < 1e-12, "det Fp = 1", det3(Fpn), 1.0);
      check(std::fabs(m.alpha - (eqps_old + sq23 * m.dgam)) <= 1e-15 * (1.0 + m.alpha), "eqps update", m.alpha,
            eqps_old + sq23 * m.dgam);
      double sn = 0.0;  // consistency: |s| = sqrt(2/3) (Y + K eqps) after the return
      for (int k = 0; k < 9; ++k) sn += m.s[k] * m.s[k];
      check(std::fabs(std::sqrt(sn) - sq23 * (mp.Y + mp.K * m.alpha)) < 1e-10, "consistency", std::sqrt(sn),
            sq23 * (mp.Y + mp.K * m.alpha));
    }
    // directional derivatives against central differences of the values
    for (int c = 0; c < 9; ++c) {
      const int k = c / 3, L = c % 3;
      Dir d;
      m.dir(q, k, L, d);
      double ds3[3];
      st.dir(q, k, L, ds3);
      double an[13];
      for (int r = 0; r < 9; ++r) an[r] = d.dP[r];
      an[9] = d.dr1;
      for (int j = 0; j < 3; ++j) an[10 + j] = ds3[j];
      print_hex(an, 13);
      const double h = 1e-6;
      double vp[13], vm[13];
      for (int s = -1; s <= 1; s += 2) {
        double Fh[9];
        for (int kk = 0; kk < 9; ++kk) Fh[kk] = F[kk];
        Fh[3 * k + L] += s * h;
        QP qh;
        make_qp(Fh, p, g, qh);
        J2Model mh;
        mh.init(qh, mp, Fp, eqps_old);
        Stab sh;
        sh.init(qh, 0.01);
        values(mh, sh, qh, s < 0 ? vm : vp);
        check(mh.plastic == m.plastic, "branch unchanged under the FD step", mh.plastic, m.plastic);
      }
      double nrm = 0.0, err = 0.0;
      for (int r = 0; r < 13; ++r) {
        const double fd = (vp[r] - vm[r]) / (2.0 * h);
        nrm = std::fmax(nrm, std::fabs(fd));
        err = std::fmax(err, std::fabs(fd - an[r]));
      }
      check(err <= 2e-6 * (1.0 + nrm), m.plastic ? "J2 plastic tangent vs FD" : "J2 elastic tangent vs FD", err,
            nrm);
    }
  }
  if (plastic_wanted) check(plastic_seen > trials / 2, "plastic branch exercised", plastic_seen, trials);
  else check(plastic_seen == 0, "elastic branch exercised", plastic_seen, 0);
}

void check_nh(int trials) {
  const MatParams mp = mat(1000.0, 0.25, 0.0, 0.0);
  for (int t = 0; t < trials; ++t) {
    double F[9], g[3];
    for (int k = 0; k < 9; ++k) F[k] = ((k % 4 == 0) ? 1.0 : 0.0) + 0.05 * urand();
    for (int j = 0; j < 3; ++j) g[j] = urand();
    const double p = urand();
    QP q;
    make_qp(F, p, g, q);
    NeoHookean m;
    m.init(q, mp);
    Stab st;
    st.init(q, 0.02);
    NHBase b;
    nh_base(q, m, st, b);
    for (int c = 0; c < 9; ++c) {
      const int k = c / 3, L = c % 3;
      Dir d;
      m.dir(q, k, L, d);
      // the closed-form column with G = e_L (and N = 0) is the same direction
      double G[3] = {0.0, 0.0, 0.0}, C[13];
      G[L] = 1.0;
      nh_direction<double>(b, k, 0.0, G, C);
      double ds3[3];
      st.dir(q, k, L, ds3);
      print_hex(C, 13);
      for (int r = 0; r < 9; ++r) check(std::fabs(C[r] - d.dP[r]) <= 1e-12 * (1.0 + std::fabs(d.dP[r])),
                                        "closed-form NH column = directional derivative", C[r], d.dP[r]);
      check(std::fabs(C[9] - d.dr1) <= 1e-12 * (1.0 + std::fabs(d.dr1)), "NH dr1", C[9], d.dr1);
      for (int j = 0; j < 3; ++j)
        check(std::fabs(C[10 + j] - ds3[j]) <= 1e-12 * (1.0 + std::fabs(ds3[j])), "NH ds3", C[10 + j], ds3[j]);
      const double h = 1e-6;
      double vp[13], vm[13];
      for (int s = -1; s <= 1; s += 2) {
        double Fh[9];
        for (int kk = 0; kk < 9; ++kk) Fh[kk] = F[kk];
        Fh[3 * k + L] += s * h;
        QP qh;
        make_qp(Fh, p, g, qh);
        NeoHookean mh;
        mh.init(qh, mp);
        Stab sh;
        sh.init(qh, 0.02);
        values(mh, sh, qh, s < 0 ? vm : vp);
      }
      double nrm = 0.0, err = 0.0;
      for (int r = 0; r < 13; ++r) {
        const double fd = (vp[r] - vm[r]) / (2.0 * h);
        nrm = std::fmax(nrm, std::fabs(fd));
        err = std::fmax(err, std::fabs(fd - C[r]));
      }
      check(err <= 1e-6 * (1.0 + nrm), "NH tangent vs FD", err, nrm);
    }
  }
}

// The element kernel's LDS indexing (goal_assembly.hip, element_kernel):
// for every thread t < T and every element count nv <= G of a group, the
// indices each phase forms stay inside the arrays they address.
void check_indices(int NN, int NQ, int G, int T, bool mf) {
  const int ND = 4 * NN;
  auto in = [&](long idx, long extent, const char* what) { check(idx >= 0 && idx < extent, what, idx, extent); };
  for (int nv = 1; nv <= G; ++nv)
    for (int tid = 0; tid < T; ++tid) {
      for (int t = tid; t < nv * NN; t += T) in(t / NN, G, "A conn element");
      for (int t = tid; t < nv * 12; t += T) in((t / 12) * 12 + t % 12, G * 12, "A coords");
      for (int t = tid; t < nv * ND; t += T) in((t / ND) * ND + t % ND, G * ND, "A dofs");
      for (int t = tid; t < nv * NQ * NN; t += T) {
        const int el = t / (NQ * NN), r = t % (NQ * NN);
        in(((el * NQ + r / NN) * NN + r % NN) * kShapeRow + 3, G * NQ * NN * kShapeRow, "B shape table");
      }
      for (int t = tid; t < nv * NQ; t += T) in((t / NQ) * NQ + t % NQ, G * NQ, "C1 point");
      for (int t = tid; t < nv * NQ * 9; t += T) {  // C2: D[el][q][rr][c], rr < 13, c < 9
        const int el = t / (NQ * 9), r = t % (NQ * 9), q = r / 9, c = r % 9;
        in(((el * NQ + q) * 13 + 12) * 13 + c, G * NQ * 169, "C2 tangent table");
      }
      for (int t = tid; t < nv * ND; t += T) {  // D: sig[el][q][..], sh[el][q][n][..]
        const int el = t / ND, k = t % ND, n = k >> 2;
        in((el * NQ + NQ - 1) * 13 + 12, G * NQ * 13, "D stress");
        in(((el * NQ + NQ - 1) * NN + n) * kShapeRow + 3, G * NQ * NN * kShapeRow, "D shape table");
      }
      if (!mf) {
        for (int t = tid; t < nv * ND; t += T) {  // E: one column per lane
          const int el = t / ND, col = t % ND, b = col >> 2;
          in(el, G, "E element");
          in(b, NN, "E column node");
        }
      } else {
        const int DT = (3 * NN + 14) / 15, NT = DT + (NN + 15) / 16;
        for (int task = tid >> 6; task < nv * NT; task += T / 64) {
          const int el = task / NT, tt = task % NT, n = tid & 63 & 15;
          in(el, G, "E tile element");
          int col;
          if (tt < DT) {
            const int b = 5 * tt + n / 3;
            col = (n < 15 && b < NN) ? 4 * b + n % 3 : -1;
          } else {
            const int b = 16 * (tt - DT) + n;
            col = b < NN ? 4 * b + 3 : -1;
          }
          if (col >= 0) in(col, ND, "E tile column");
        }
      }
    }
}

// The J2 point states: the element kernel's read of value r (r >= 13) and its
// fill of the tangent's structural zeros together cover every entry of the
// 13 x 13 table exactly once, and the blocked index of every value of every
// point stays inside the buffer point_state_doubles sizes.
void check_point_states() {
  int hits[13][13] = {};
  for (int r = 13; r < kPointTangent; ++r) {
    int row, col;
    if (r < 130) row = (r - 13) % 13, col = (r - 13) / 13;
    else if (r < 140) row = r - 130, col = 9;
    else row = 10 + (r - 140) % 3, col = 10 + (r - 140) / 3;
    check(row >= 0 && row < 13 && col >= 0 && col < 13, "point tangent position", row, col);
    if (row >= 0 && row < 13 && col >= 0 && col < 13) ++hits[row][col];
  }
  for (int z = 0; z < 33; ++z) {
    const int row = z < 3 ? 10 + z : (z - 3) % 10, col = z < 3 ? 9 : 10 + (z - 3) / 10;
    ++hits[row][col];
  }
  for (int i = 0; i < 13; ++i)
    for (int j = 0; j < 13; ++j) check(hits[i][j] == 1, "point tangent entry covered once", i * 13 + j, hits[i][j]);
  for (long npts : {1L, 63L, 64L, 65L, 1000L}) {
    for (int ps : {kPointStress, kPointTangent}) {
      const long n = (long)point_state_doubles(npts, ps);
      for (long p = 0; p < npts; ++p)
        for (int k = 0; k < ps; ++k) {
          const long idx = (long)point_state_index(p, k, ps);
          if (idx < 0 || idx >= n) check(false, "point state index", idx, n);
        }
    }
  }
}

}  // namespace

int main() {
  check_point_states();
  check_j2(200, true);
  check_j2(50, false);
  check_nh(100);
  // the kernels' shapes: (nodes, points, elements per group, threads, matrix cores)
  check_indices(4, 1, 16, 256, false);
  check_indices(4, 4, 4, 64, false);
  check_indices(10, 1, 6, 256, false);
  check_indices(10, 4, 6, 256, false);
  check_indices(10, 4, 6, 256, true);
  if (g_fail) {
    std::fprintf(stderr, "%d checks failed\n", g_fail);
    return 1;
  }
  std::fprintf(stderr, "all checks passed\n");
  return 0;
}
