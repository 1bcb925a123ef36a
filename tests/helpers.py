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
