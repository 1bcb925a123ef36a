// Fused element assembly for goal's mixed u-p Total Lagrangian tets.
//
// One kernel performs, per element, the whole evaluator chain that
// goal::assemble() runs hook by hook (src/goal_assembly.cpp:65-88, chain order
// src/goal_mechanics.cpp:97-146): gather of u and p, interpolation, kinematics
// F = I + grad u, J, the constitutive model, Mixed, MResidual, PResidual,
// Stabilization, and the scatter into the FP64 residual and CSR Jacobian.
//
// Work layout (CDNA4, wave64): a workgroup is one wavefront and processes a
// group of EPW elements at a time (EPW = 4 for 4-node tets: 16 lanes per
// element; EPW = 1 for 10-node tets).  Element data (connectivity, vertex
// coordinates, DOFs, CSR row starts, node-pair positions) is staged in LDS.
// Lanes then cooperate in four phases: shape functions per (qp, node);
// constitutive update per (qp, direction) - each lane evaluates one column of
// the 13x13 generalized tangent; residual rows; and Jacobian columns, where a
// lane holds D*B for its column in registers and walks the element rows,
// adding each entry into the CSR with a hardware FP64 atomic (4 adjacent lanes
// hit 4 adjacent columns of one row).  Workgroups can be mapped so that each
// XCD sweeps one contiguous slab of the element range (xcd_remap), keeping the
// CSR rows an XCD accumulates into within its own L2.

#include "goal_device.hpp"
#include "goal_kernels.hpp"

namespace goal {
namespace {

using dev::Quad;

template <int NN, int NQ>
struct Smem {
  static constexpr int EPW = (NN == 4) ? 4 : 1;
  static constexpr int ND = 4 * NN;
  double x[EPW][ND];
  double X[EPW][4][3];
  double sh[EPW][NQ][NN][4];  // N_a, dN_a/dX
  double wdv[EPW][NQ];
  double sig[EPW][NQ][13];
  double D[EPW][NQ][13][13];
  double zq[EPW][NQ][4];      // z_diff (u,p) at qp            (localize)
  double zgq[EPW][NQ][4][3];  // grad z_diff                   (localize)
  double zcq[EPW][NQ][4];     // z_coarse p and its gradient   (localize)
  double zd[EPW][ND];
  double zc[EPW][NN];
  int64_t rowbase[EPW][ND];
  int32_t conn[EPW][NN];
  uint16_t pos[EPW][NN * NN];
};

// Element-group schedule.  Plain: group = block + it * gridDim.  XCD slabs:
// workgroups are dispatched round-robin over the 8 XCDs, so block b runs on
// XCD b % 8; XCD x is given the groups [x*per, (x+1)*per) and its blocks walk
// that slab together.
__device__ inline int64_t group_of(int64_t it, int64_t ngroups, int remap) {
  const int64_t nb = gridDim.x, b = blockIdx.x;
  if (!remap || (nb & 7)) {
    int64_t g = b + it * nb;
    return g < ngroups ? g : -1;
  }
  const int64_t x = b & 7, s = b >> 3, bx = nb >> 3;
  const int64_t per = (ngroups + 7) >> 3;
  int64_t g = s + it * bx;
  if (g >= per) return -1;
  g += x * per;
  return g < ngroups ? g : -1;
}

__device__ inline void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

template <int MODEL>
struct ModelOf;
template <>
struct ModelOf<MODEL_ELASTIC> {
  using type = dev::Elastic;
};
template <>
struct ModelOf<MODEL_NEOHOOKEAN> {
  using type = dev::NeoHookean;
};
template <>
struct ModelOf<MODEL_J2> {
  using type = dev::J2Model;
};

template <int NN, int NQ, int MODEL, int KIND>
__global__ __launch_bounds__(64) void assembly_kernel(AsmArgs a) {
  using S = Smem<NN, NQ>;
  constexpr int EPW = S::EPW, ND = S::ND, LPE = 64 / EPW;
  constexpr bool JAC = (KIND == ASM_JACOBIAN || KIND == ASM_ADJOINT);
  constexpr bool LOC = (KIND == ASM_LOCALIZE);
  __shared__ S sm;
  const int lane = threadIdx.x;
  const int el = lane / LPE;
  const int l = lane % LPE;
  const int64_t ne = a.num_elems;
  const int64_t ngroups = (ne + EPW - 1) / EPW;
  const int64_t nqtot = ne * NQ;

  for (int64_t it = 0;; ++it) {
    const int64_t grp = group_of(it, ngroups, a.xcd_remap);
    if (grp < 0) break;
    const int64_t e = grp * EPW + el;
    const bool valid = e < ne;

    // ---- gather (Displacement/Pressure::gather, src/goal_displacement.cpp:139-156)
    if (valid) {
      const int32_t* ec = a.conn + e * NN;
      if (l < NN) sm.conn[el][l] = ec[l];
      if (l < 12) sm.X[el][l / 3][l % 3] = a.coords[3 * (int64_t)ec[l / 3] + (l % 3)];
      for (int k = l; k < ND; k += LPE) {
        const int64_t dof = 4 * (int64_t)ec[k >> 2] + (k & 3);
        sm.x[el][k] = a.x[dof];
        if (JAC) sm.rowbase[el][k] = a.row_ptr[dof];
        if (LOC) sm.zd[el][k] = a.z_diff[dof];
      }
      if (LOC)
        for (int k = l; k < NN; k += LPE) sm.zc[el][k] = a.z_coarse[4 * (int64_t)ec[k] + 3];
      if (JAC)
        for (int k = l; k < NN * NN; k += LPE) sm.pos[el][k] = a.pos[e * (NN * NN) + k];
    }
    __syncthreads();

    // ---- shape functions at the quadrature points (u/uw/p/pw at_point)
    if (valid)
      for (int t = l; t < NQ * NN; t += LPE) {
        const int q = t / NN, n = t % NN;
        double gl[4][3];
        const double det = dev::tet_gradlambda(sm.X[el], gl);
        double xi[3], w;
        Quad<NQ>::point(q, xi, w);
        const double lam[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
        double N, dN[3];
        dev::shape<NN>(n, lam, gl, N, dN);
        sm.sh[el][q][n][0] = N;
        sm.sh[el][q][n][1] = dN[0];
        sm.sh[el][q][n][2] = dN[1];
        sm.sh[el][q][n][3] = dN[2];
        if (n == 0) sm.wdv[el][q] = w * det;
      }
    __syncthreads();

    // ---- constitutive update: lane (qp, c), c < 9 -> tangent column F_kL,
    //      c == 9 -> value, p and grad-p columns, state writes
    if (valid)
      for (int t = l; t < NQ * 10; t += LPE) {
        const int q = t / 10, c = t % 10;
        const MatParams mp = a.mats[a.elem_set ? a.elem_set[e] : 0];
        dev::QP qp;
        for (int k = 0; k < 9; ++k) qp.H[k] = 0.0;
        qp.p = 0.0;
        qp.g[0] = qp.g[1] = qp.g[2] = 0.0;
        for (int n = 0; n < NN; ++n) {
          const double N = sm.sh[el][q][n][0];
          const double pn = sm.x[el][4 * n + 3];
          qp.p += pn * N;
          for (int j = 0; j < 3; ++j) {
            const double dNj = sm.sh[el][q][n][1 + j];
            qp.g[j] += pn * dNj;
            for (int i = 0; i < 3; ++i) qp.H[3 * i + j] += sm.x[el][4 * n + i] * dNj;
          }
        }
        for (int k = 0; k < 9; ++k) qp.F[k] = qp.H[k] + ((k % 4 == 0) ? 1.0 : 0.0);
        qp.J = dev::det3(qp.F);
        dev::inv3(qp.F, qp.J, qp.Fi);
        dev::Stab st;
        const double tau = a.stabilization ? dev::stab_tau(sm.X[el], mp) : 0.0;
        st.init(qp, tau);
        const double wdv = sm.wdv[el][q];
        const int64_t sidx = e * NQ + q;

        typename ModelOf<MODEL>::type m;
        if constexpr (MODEL == MODEL_J2) {
          double Fpo[9];
          for (int k = 0; k < 9; ++k) Fpo[k] = a.Fp_old[k * nqtot + sidx];
          m.init(qp, mp, Fpo, a.eqps_old[sidx]);
          if (c == 9) {
            if (m.failed()) atomicOr(a.status, 1);
            if (a.save_state) {
              a.eqps[sidx] = m.alpha;
              if (m.plastic) {
                double Fpn[9];
                m.new_Fp(Fpo, Fpn);
                for (int k = 0; k < 9; ++k) a.Fp[k * nqtot + sidx] = Fpn[k];
              }
            }
          }
        } else {
          m.init(qp, mp);
        }
        if (c == 9) {
          if (a.save_state)
            for (int k = 0; k < 9; ++k) a.sigma[k * nqtot + sidx] = m.sig[k];
          for (int k = 0; k < 9; ++k) sm.sig[el][q][k] = m.P[k] * wdv;
          sm.sig[el][q][9] = m.r1(qp) * wdv;
          double s3[3];
          st.value(s3);
          for (int j = 0; j < 3; ++j) sm.sig[el][q][10 + j] = s3[j] * wdv;
          if (LOC) {
            for (int i = 0; i < 4; ++i) {
              double z = 0.0, zg[3] = {0.0, 0.0, 0.0};
              for (int n = 0; n < NN; ++n) {
                const double zn = sm.zd[el][4 * n + i];
                z += zn * sm.sh[el][q][n][0];
                for (int j = 0; j < 3; ++j) zg[j] += zn * sm.sh[el][q][n][1 + j];
              }
              sm.zq[el][q][i] = z;
              for (int j = 0; j < 3; ++j) sm.zgq[el][q][i][j] = zg[j];
            }
            double z = 0.0, zg[3] = {0.0, 0.0, 0.0};
            for (int n = 0; n < NN; ++n) {
              const double zn = sm.zc[el][n];
              z += zn * sm.sh[el][q][n][0];
              for (int j = 0; j < 3; ++j) zg[j] += zn * sm.sh[el][q][n][1 + j];
            }
            sm.zcq[el][q][0] = z;
            for (int j = 0; j < 3; ++j) sm.zcq[el][q][1 + j] = zg[j];
          }
          if (JAC) {
            // p column: dsigma = I (Mixed), dP = d(P)/dp, dr1 = dr1/dp
            double dPp[9], dr1p;
            if constexpr (MODEL == MODEL_ELASTIC) {
              for (int k = 0; k < 9; ++k) dPp[k] = (k % 4 == 0) ? 1.0 : 0.0;
              dr1p = 1.0;
            } else {
              for (int i = 0; i < 3; ++i)
                for (int jj = 0; jj < 3; ++jj) dPp[3 * i + jj] = qp.J * qp.Fi[3 * jj + i];
              dr1p = 1.0 / mp.kappa;
            }
            for (int r = 0; r < 9; ++r) sm.D[el][q][r][9] = dPp[r] * wdv;
            sm.D[el][q][9][9] = dr1p * wdv;
            for (int r = 10; r < 13; ++r) sm.D[el][q][r][9] = 0.0;
            // grad-p columns: only the stabilization flux depends on grad p
            for (int I = 0; I < 3; ++I) {
              double ds3[3];
              st.dir_g(qp, I, ds3);
              for (int r = 0; r < 10; ++r) sm.D[el][q][r][10 + I] = 0.0;
              for (int j = 0; j < 3; ++j) sm.D[el][q][10 + j][10 + I] = ds3[j] * wdv;
            }
          }
        } else if (JAC) {
          const int kk = c / 3, LL = c % 3;
          dev::Dir d;
          m.dir(qp, kk, LL, d);
          double ds3[3];
          st.dir(qp, kk, LL, ds3);
          for (int r = 0; r < 9; ++r) sm.D[el][q][r][c] = d.dP[r] * wdv;
          sm.D[el][q][9][c] = d.dr1 * wdv;
          for (int j = 0; j < 3; ++j) sm.D[el][q][10 + j][c] = ds3[j] * wdv;
        }
      }
    __syncthreads();

    // ---- residual rows (MResidual, PResidual, Stabilization + scatter)
    if (valid && a.R)
      for (int k = l; k < ND; k += LPE) {
        const int n = k >> 2, i = k & 3;
        double r = 0.0;
        for (int q = 0; q < NQ; ++q) {
          const double* sg = sm.sig[el][q];
          const double* sh = sm.sh[el][q][n];
          if (!LOC) {
            if (i < 3) r += sg[3 * i] * sh[1] + sg[3 * i + 1] * sh[2] + sg[3 * i + 2] * sh[3];
            else r += sg[9] * sh[0] + sg[10] * sh[1] + sg[11] * sh[2] + sg[12] * sh[3];
          } else {
            // DisplacementAdjoint / PressureAdjoint weights
            // (src/goal_displacement_adjoint.cpp:37-53, src/goal_pressure_adjoint.cpp:38-49)
            if (i < 3) {
              const double z = sm.zq[el][q][i];
              const double* zg = sm.zgq[el][q][i];
              for (int J = 0; J < 3; ++J) r += sg[3 * i + J] * (zg[J] * sh[0] + z * sh[1 + J]);
            } else {
              r += sg[9] * sm.zq[el][q][3] * sh[0];
              const double zc = sm.zcq[el][q][0];
              for (int J = 0; J < 3; ++J) r += sg[10 + J] * (sm.zcq[el][q][1 + J] * sh[0] + zc * sh[1 + J]);
            }
          }
        }
        atomic_add(a.R + 4 * (int64_t)sm.conn[el][n] + i, r);
      }

    // ---- Jacobian columns: C = D B for column (b,k), then K(row, col) = B^T C
    if constexpr (JAC) {
      if (valid && a.vals)
        for (int col = l; col < ND; col += LPE) {
          const int b = col >> 2, kc = col & 3;
          double C[NQ][13];
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const double* shb = sm.sh[el][q][b];
            for (int r = 0; r < 13; ++r) {
              const double* Dr = sm.D[el][q][r];
              if (kc < 3)
                C[q][r] = Dr[3 * kc] * shb[1] + Dr[3 * kc + 1] * shb[2] + Dr[3 * kc + 2] * shb[3];
              else
                C[q][r] = Dr[9] * shb[0] + Dr[10] * shb[1] + Dr[11] * shb[2] + Dr[12] * shb[3];
            }
          }
          for (int n = 0; n < NN; ++n) {
            double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
              const double* sh = sm.sh[el][q][n];
              for (int i = 0; i < 3; ++i) v[i] += C[q][3 * i] * sh[1] + C[q][3 * i + 1] * sh[2] + C[q][3 * i + 2] * sh[3];
              v[3] += C[q][9] * sh[0] + C[q][10] * sh[1] + C[q][11] * sh[2] + C[q][12] * sh[3];
            }
            if (KIND == ASM_JACOBIAN) {
              const int64_t off = 4 * (int64_t)sm.pos[el][n * NN + b] + kc;
              for (int i = 0; i < 4; ++i) atomic_add(a.vals + sm.rowbase[el][4 * n + i] + off, v[i]);
            } else {
              // transposed scatter (Displacement<FADT>::scatter_adjoint,
              // src/goal_displacement.cpp:196-214): entry (row, col) -> (col, row)
              const int64_t base = sm.rowbase[el][col] + 4 * (int64_t)sm.pos[el][b * NN + n];
              for (int i = 0; i < 4; ++i) atomic_add(a.vals + base + i, v[i]);
            }
          }
        }
    }
    __syncthreads();
  }
}

template <int NN, int NQ, int MODEL>
hipError_t launch3(int kind, const AsmArgs& a, int blocks, hipStream_t s) {
  dim3 grid(blocks), block(64);
  switch (kind) {
    case ASM_RESIDUAL: hipLaunchKernelGGL((assembly_kernel<NN, NQ, MODEL, ASM_RESIDUAL>), grid, block, 0, s, a); break;
    case ASM_JACOBIAN: hipLaunchKernelGGL((assembly_kernel<NN, NQ, MODEL, ASM_JACOBIAN>), grid, block, 0, s, a); break;
    case ASM_ADJOINT: hipLaunchKernelGGL((assembly_kernel<NN, NQ, MODEL, ASM_ADJOINT>), grid, block, 0, s, a); break;
    case ASM_LOCALIZE: hipLaunchKernelGGL((assembly_kernel<NN, NQ, MODEL, ASM_LOCALIZE>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int NN, int NQ>
hipError_t launch2(int kind, const AsmArgs& a, int blocks, hipStream_t s) {
  switch (a.model) {
    case MODEL_ELASTIC: return launch3<NN, NQ, MODEL_ELASTIC>(kind, a, blocks, s);
    case MODEL_NEOHOOKEAN: return launch3<NN, NQ, MODEL_NEOHOOKEAN>(kind, a, blocks, s);
    case MODEL_J2: return launch3<NN, NQ, MODEL_J2>(kind, a, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

// ------------------------------------------------------------------ QoI
// One thread per element.  AvgDisp: sum_q (sum_i u_i) w dv / 3
// (src/goal_avg_disp.cpp:16-21); AvgVM: sum_q vm(sigma) w dv over one element
// set (src/goal_avg_vm.cpp:51-61, src/goal_von_mises.cpp:6-17).  With dMdu set,
// the derivative of the element value w.r.t. the element DOFs is added to
// dMdu (QoI<FADT>::scatter, src/goal_qoi.cpp:63-76).
template <int NN, int NQ, int MODEL>
__global__ __launch_bounds__(256) void qoi_kernel(QoiArgs a) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.num_elems) return;
  const int32_t* ec = a.conn + e * NN;
  const int set = a.elem_set ? a.elem_set[e] : 0;
  double val = 0.0;
  if (a.kind == QOI_AVG_VM && a.qoi_set >= 0 && set != a.qoi_set) {
    a.elem_value[e] = 0.0;
    return;
  }
  double X[4][3];
  for (int v = 0; v < 4; ++v)
    for (int c = 0; c < 3; ++c) X[v][c] = a.coords[3 * (int64_t)ec[v] + c];
  double gl[4][3];
  const double det = dev::tet_gradlambda(X, gl);
  const MatParams mp = a.mats[set];
  const int64_t nqtot = a.num_elems * NQ;
  for (int q = 0; q < NQ; ++q) {
    double xi[3], w;
    Quad<NQ>::point(q, xi, w);
    const double lam[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
    const double wdv = w * det;
    if (a.kind == QOI_AVG_DISP) {
      double s = 0.0;
      for (int n = 0; n < NN; ++n) {
        double N, dN[3];
        dev::shape<NN>(n, lam, gl, N, dN);
        const int64_t base = 4 * (int64_t)ec[n];
        s += (a.x[base] + a.x[base + 1] + a.x[base + 2]) * N;
        if (a.dMdu)
          for (int i = 0; i < 3; ++i) atomic_add(a.dMdu + base + i, N * wdv / 3.0);
      }
      val += s * wdv / 3.0;
      continue;
    }
    dev::QP qp;
    for (int k = 0; k < 9; ++k) qp.H[k] = 0.0;
    qp.p = 0.0;
    qp.g[0] = qp.g[1] = qp.g[2] = 0.0;
    double Ns[NN], dNs[NN][3];
    for (int n = 0; n < NN; ++n) {
      dev::shape<NN>(n, lam, gl, Ns[n], dNs[n]);
      const int64_t base = 4 * (int64_t)ec[n];
      qp.p += a.x[base + 3] * Ns[n];
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) qp.H[3 * i + j] += a.x[base + i] * dNs[n][j];
    }
    for (int k = 0; k < 9; ++k) qp.F[k] = qp.H[k] + ((k % 4 == 0) ? 1.0 : 0.0);
    qp.J = dev::det3(qp.F);
    dev::inv3(qp.F, qp.J, qp.Fi);
    typename ModelOf<MODEL>::type m;
    const int64_t sidx = e * NQ + q;
    if constexpr (MODEL == MODEL_J2) {
      double Fpo[9];
      for (int k = 0; k < 9; ++k) Fpo[k] = a.Fp_old[k * nqtot + sidx];
      m.init(qp, mp, Fpo, a.eqps_old[sidx]);
      if (m.failed()) atomicOr(a.status, 1);
    } else {
      m.init(qp, mp);
    }
    const double* s = m.sig;
    const double s1 = (s[0] - s[4]) * (s[0] - s[4]);
    const double s2 = (s[4] - s[8]) * (s[4] - s[8]);
    const double s3 = (s[8] - s[0]) * (s[8] - s[0]);
    const double s4 = s[1] * s[1], s5 = s[5] * s[5], s6 = s[6] * s[6];
    const double vm = sqrt(0.5 * (s1 + s2 + s3 + 6.0 * (s4 + s5 + s6)));
    val += vm * wdv;
    if (a.dMdu) {
      // d vm = (1/(2 vm)) * 1/2 * d(s1+s2+s3+6(s4+s5+s6))
      auto dvm = [&](const double* ds) {
        const double t = 2.0 * (s[0] - s[4]) * (ds[0] - ds[4]) + 2.0 * (s[4] - s[8]) * (ds[4] - ds[8]) +
                         2.0 * (s[8] - s[0]) * (ds[8] - ds[0]) +
                         12.0 * (s[1] * ds[1] + s[5] * ds[5] + s[6] * ds[6]);
        return 0.25 * t / vm;
      };
      double gF[9];
      for (int c = 0; c < 9; ++c) {
        dev::Dir d;
        m.dir(qp, c / 3, c % 3, d);
        gF[c] = dvm(d.dsig);
      }
      double Ip[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      const double gp = dvm(Ip);
      for (int n = 0; n < NN; ++n) {
        const int64_t base = 4 * (int64_t)ec[n];
        for (int k = 0; k < 3; ++k)
          atomic_add(a.dMdu + base + k,
                     (gF[3 * k] * dNs[n][0] + gF[3 * k + 1] * dNs[n][1] + gF[3 * k + 2] * dNs[n][2]) * wdv);
        atomic_add(a.dMdu + base + 3, gp * Ns[n] * wdv);
      }
    }
  }
  a.elem_value[e] = val;
}

template <int NN, int NQ>
hipError_t launch_qoi2(const QoiArgs& a, hipStream_t s) {
  dim3 grid((unsigned)((a.num_elems + 255) / 256)), block(256);
  if (a.num_elems == 0) return hipSuccess;
  switch (a.model) {
    case MODEL_ELASTIC: hipLaunchKernelGGL((qoi_kernel<NN, NQ, MODEL_ELASTIC>), grid, block, 0, s, a); break;
    case MODEL_NEOHOOKEAN: hipLaunchKernelGGL((qoi_kernel<NN, NQ, MODEL_NEOHOOKEAN>), grid, block, 0, s, a); break;
    case MODEL_J2: hipLaunchKernelGGL((qoi_kernel<NN, NQ, MODEL_J2>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_assembly(int kind, int order, int quad_order, const AsmArgs& a, int blocks, hipStream_t s) {
  if (a.num_elems == 0) return hipSuccess;
  if (blocks < 1) return hipErrorInvalidValue;
  if (order == 1 && quad_order == 1) return launch2<4, 1>(kind, a, blocks, s);
  if (order == 1 && quad_order == 2) return launch2<4, 4>(kind, a, blocks, s);
  if (order == 2 && quad_order == 1) return launch2<10, 1>(kind, a, blocks, s);
  if (order == 2 && quad_order == 2) return launch2<10, 4>(kind, a, blocks, s);
  return hipErrorInvalidValue;
}

hipError_t launch_qoi(int order, int quad_order, const QoiArgs& a, hipStream_t s) {
  if (order == 1 && quad_order == 1) return launch_qoi2<4, 1>(a, s);
  if (order == 1 && quad_order == 2) return launch_qoi2<4, 4>(a, s);
  if (order == 2 && quad_order == 1) return launch_qoi2<10, 1>(a, s);
  if (order == 2 && quad_order == 2) return launch_qoi2<10, 4>(a, s);
  return hipErrorInvalidValue;
}

}  // namespace goal
