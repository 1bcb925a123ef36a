"""Independent numpy references used to pin the oracle (CPU tests only).

Nothing here restates the oracle's code path: these are the textbook objects
the reference formulas must agree with (hyperelastic energies whose gradient
is the residual, closed-form radial return, brute-force graphs, a Kuhn mesh
generator written from its definition)."""
from __future__ import annotations

import itertools

import numpy as np

A4 = 0.5854101966249685
B4 = 0.1381966011250105


def quad(nq):
    if nq == 1:
        return [np.array([0.25, 0.25, 0.25])], [1.0 / 6.0]
    pts = [np.array([B4, B4, B4]), np.array([A4, B4, B4]), np.array([B4, A4, B4]), np.array([B4, B4, A4])]
    return pts, [1.0 / 24.0] * 4


def shapes(order, X, xi):
    """P1/P2 tet shape values and X-gradients at parent point xi (numpy)."""
    lam = np.array([1.0 - xi.sum(), xi[0], xi[1], xi[2]])
    J = np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]]).T  # dX/dxi
    gl = np.zeros((4, 3))
    Ji = np.linalg.inv(J)
    gl[1:] = Ji
    gl[0] = -Ji.sum(axis=0)
    if order == 1:
        return lam, gl, np.linalg.det(J)
    edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    N = np.zeros(10)
    dN = np.zeros((10, 3))
    for a in range(4):
        N[a] = lam[a] * (2 * lam[a] - 1)
        dN[a] = (4 * lam[a] - 1) * gl[a]
    for k, (i, j) in enumerate(edges):
        N[4 + k] = 4 * lam[i] * lam[j]
        dN[4 + k] = 4 * (lam[i] * gl[j] + lam[j] * gl[i])
    return N, dN, np.linalg.det(J)


def qp_fields(order, coords, conn, x, e, nq):
    """Per quadrature point: (w*dv, N, dN, u, H, p, g) for element e."""
    ec = conn[e]
    X = coords[ec[:4]]
    un = np.stack([x[4 * ec + i] for i in range(3)], axis=1)
    pn = x[4 * ec + 3]
    out = []
    pts, ws = quad(nq)
    for xi, w in zip(pts, ws):
        N, dN, dv = shapes(order, X, xi)
        u = N @ un
        H = un.T @ dN
        p = N @ pn
        g = dN.T @ pn
        out.append((w * dv, N, dN, u, H, p, g))
    return out


def nh_energy(order, coords, conn, x, mu, nq):
    """Pi(u, p) = sum_q [mu/2 (J^{-2/3} tr(F^T F) - 3) + p (J - 1)] w dv.
    Its u-gradient is the momentum residual of the mixed neo-Hookean chain."""
    tot = 0.0
    for e in range(conn.shape[0]):
        for wdv, N, dN, u, H, p, g in qp_fields(order, coords, conn, x, e, nq):
            F = np.eye(3) + H
            J = np.linalg.det(F)
            tot += (0.5 * mu * (J ** (-2.0 / 3.0) * np.sum(F * F) - 3.0) + p * (J - 1.0)) * wdv
    return tot


def pressure_potential(order, coords, conn, x, mu, kappa, c0, nq, stab=True):
    """Pi_p(u, p) = sum_q [p^2/(2 kappa) - 1/2 (J - 1/J) p + 1/2 tau J g.C^{-1}g] w dv.
    Its p-gradient (u fixed) is the pressure residual with stabilization."""
    tot = 0.0
    for e in range(conn.shape[0]):
        ec = conn[e]
        X = coords[ec[:4]]
        h2 = np.mean([np.sum((X[j] - X[i]) ** 2) for i, j in [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]])
        tau = 0.5 * c0 * h2 / mu
        for wdv, N, dN, u, H, p, g in qp_fields(order, coords, conn, x, e, nq):
            F = np.eye(3) + H
            J = np.linalg.det(F)
            Ci = np.linalg.inv(F.T @ F)
            val = p * p / (2.0 * kappa) - 0.5 * (J - 1.0 / J) * p
            if stab:
                val += 0.5 * tau * J * g @ Ci @ g
            tot += val * wdv
    return tot


def j2_radial_return(F, Fp_old, eqps_old, mu, K, Y):
    """Closed-form return of linear-hardening finite-strain J2:
    dgam = f / (2 mubar + 2K/3) when f > 0; returns (plastic, dgam, eqps_new,
    trial deviator s_tr, mubar, J)."""
    J = np.linalg.det(F)
    Cpi = np.linalg.inv(Fp_old @ Fp_old.T)
    be = J ** (-2.0 / 3.0) * F @ Cpi @ F.T
    s = mu * (be - np.trace(be) / 3 * np.eye(3))
    mubar = mu * np.trace(be) / 3
    f = np.linalg.norm(s) - np.sqrt(2.0 / 3.0) * (Y + K * eqps_old)
    if f <= 1e-12:
        return False, 0.0, eqps_old, s, mubar, J
    dgam = f / (2 * mubar + 2.0 * K / 3.0)
    return True, dgam, eqps_old + np.sqrt(2.0 / 3.0) * dgam, s, mubar, J


def graph_bruteforce(npe, num_nodes, conn):
    """Clique-union CSR with sorted columns via numpy set operations."""
    pairs = set()
    for ec in conn:
        for a, b in itertools.product(ec, ec):
            pairs.add((int(a), int(b)))
    adj = [[] for _ in range(num_nodes)]
    for a, b in pairs:
        adj[a].append(b)
    rp = [0]
    cols = []
    for a in range(num_nodes):
        nb = sorted(adj[a])
        for i in range(4):
            cols += [4 * b + j for b in nb for j in range(4)]
            rp.append(len(cols))
    return np.array(rp, dtype=np.int64), np.array(cols, dtype=np.int32)


def splitmix64(x):
    m = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & m
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
    return x ^ (x >> 31)


def kuhn_box(nx, ny, nz, lx, ly, lz, order, jitter, seed):
    """Kuhn triangulation written from its definition: every hexahedron is cut
    into the 6 simplices of monotone lattice paths (0,0,0)->(1,1,1)."""
    n = (nx, ny, nz)
    L = (lx, ly, lz)
    h = [L[c] / n[c] for c in range(3)]
    vx, vy, vz = nx + 1, ny + 1, nz + 1
    vc = np.zeros((vx * vy * vz, 3))
    for k in range(vz):
        for j in range(vy):
            for i in range(vx):
                v = i + vx * (j + vy * k)
                for c, idx in enumerate((i, j, k)):
                    xx = idx * h[c]
                    if jitter > 0 and 0 < idx < n[c]:
                        r = splitmix64(splitmix64(seed) ^ (3 * v + c))
                        u = (r >> 11) * (1.0 / 9007199254740992.0)
                        xx += (2.0 * u - 1.0) * jitter * h[c]
                    if idx == n[c]:
                        xx = L[c]
                    vc[v, c] = xx
    s = order
    Lx, Ly, Lz = s * nx + 1, s * ny + 1, s * nz + 1

    def node(I, J, K):
        return I + Lx * (J + Ly * K)

    coords = np.zeros((Lx * Ly * Lz, 3))
    for K in range(Lz):
        for J in range(Ly):
            for I in range(Lx):
                if s == 1:
                    coords[node(I, J, K)] = vc[node(I, J, K)]
                else:
                    P = (I, J, K)
                    A = [(p - (p & 1)) // 2 for p in P]
                    B = [(p + (p & 1)) // 2 for p in P]
                    va = A[0] + vx * (A[1] + vy * A[2])
                    vb = B[0] + vx * (B[1] + vy * B[2])
                    coords[node(I, J, K)] = 0.5 * (vc[va] + vc[vb])
    conn = []
    perms = list(itertools.permutations(range(3)))
    edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                for perm in perms:
                    cur = [i, j, k]
                    V = [tuple(cur)]
                    for ax in perm:
                        cur[ax] += 1
                        V.append(tuple(cur))
                    d = np.linalg.det(np.array([np.subtract(V[m], V[0]) for m in (1, 2, 3)], dtype=float))
                    if d < 0:
                        V[2], V[3] = V[3], V[2]
                    row = [node(*(s * np.array(v))) for v in V]
                    if order == 2:
                        row += [node(*(np.array(V[a]) + np.array(V[b]))) for a, b in edges]
                    conn.append(row)
    return coords, np.array(conn, dtype=np.int32)
