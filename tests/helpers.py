"""Shared set-up for the parity tests: the same seeded inputs for the HIP
path (through the C ABI) and for the oracle."""
from __future__ import annotations

import numpy as np

import oracle as O

NH = dict(model=1, E=1000.0, nu=0.25, K=0.0, Y=0.0, c0=1.0)          # cube configs
J2 = dict(model=2, E=200.0, nu=0.3, K=2.0, Y=0.25, c0=1.0)           # bar config
ELASTIC = dict(model=0, E=1000.0, nu=0.25, K=0.0, Y=0.0, c0=1.0)
MATS = {0: ELASTIC, 1: NH, 2: J2}

# Normwise relative tolerance for the plain FP64 path against the oracle.
RTOL = 1e-11


def mat_row(m):
    return [m["E"], m["nu"], m["K"], m["Y"], m["c0"]]


def rel(a, b):
    """Normwise relative difference ||a - b|| / ||b||."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    if nb == 0.0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)


def random_state(num_nodes, seed, scale_u=0.02, scale_p=1.0):
    """A non-converged iterate: random displacements and pressures."""
    rng = np.random.default_rng(seed)
    x = np.zeros(4 * num_nodes)
    x[0::4] = scale_u * rng.standard_normal(num_nodes)
    x[1::4] = scale_u * rng.standard_normal(num_nodes)
    x[2::4] = scale_u * rng.standard_normal(num_nodes)
    x[3::4] = scale_p * rng.standard_normal(num_nodes)
    return x


def random_history(n_pts, seed, amp=0.02):
    """Plastic history: Fp_old = exp(sym traceless A) (det 1), eqps_old >= 0."""
    rng = np.random.default_rng(seed)
    Fp = np.zeros((9, n_pts))
    for k in range(n_pts):
        A = amp * rng.standard_normal((3, 3))
        A = 0.5 * (A + A.T)
        A -= np.trace(A) / 3 * np.eye(3)
        w, V = np.linalg.eigh(A)
        Fp[:, k] = (V @ np.diag(np.exp(w)) @ V.T).ravel()
    eqps = np.abs(0.01 * rng.standard_normal(n_pts))
    return Fp, eqps


def oracle_mesh_from(disc):
    coords, conn, es = disc.mesh()
    return O.Mesh(disc.order, coords, conn, es)


def fan_mesh(order=1, subdiv=1, jitter=0.0, seed=0):
    """Tets fanned around one centre node: an icosphere's triangles joined to the
    origin.  subdiv=1 gives 80 tets sharing node 0 (more than 64 incident
    elements).  order 2 adds straight edge midpoints in apf edge order."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, dtype=float) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        mid = {}
        F2 = []

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = V[a] + V[b]
                V.append(p / np.linalg.norm(p))
                mid[key] = len(V) - 1
            return mid[key]

        for a, b, c in F:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            F2 += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = F2
    rng = np.random.default_rng(seed)
    coords = np.vstack([np.zeros(3)] + [v * (1.0 + jitter * rng.uniform(-1, 1)) for v in V])
    conn = []
    for a, b, c in F:
        tet = [0, a + 1, b + 1, c + 1]
        X = coords[tet]
        if np.linalg.det(np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) < 0:
            tet[2], tet[3] = tet[3], tet[2]
        conn.append(tet)
    conn = np.array(conn, dtype=np.int32)
    if order == 1:
        return coords, conn
    edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    emap = {}
    pts = list(coords)
    conn2 = []
    for tet in conn:
        row = list(tet)
        for a, b in edges:
            key = (min(tet[a], tet[b]), max(tet[a], tet[b]))
            if key not in emap:
                emap[key] = len(pts)
                pts.append(0.5 * (coords[key[0]] + coords[key[1]]))
            row.append(emap[key])
        conn2.append(row)
    return np.array(pts), np.array(conn2, dtype=np.int32)
