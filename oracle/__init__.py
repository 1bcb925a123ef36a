"""ORACLE -- test infrastructure only (tests/, smoke check, bench cpu_baseline).

ctypes wrapper of liboracle.so, the C++ restatement of goal's assembly and DWR
path (oracle/oracle.cpp).  Built by `make -C oracle` (also run by
__graft_entry__.build()).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

K_RESIDUAL, K_JACOBIAN, K_ADJOINT, K_LOCALIZE = 0, 1, 2, 3


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def _load():
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    return C.CDLL(str(LIB_PATH))


lib = _load()
_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


_sig("oracle_graph", _i64, _i32, _i64, _i64, _p, _p, _p, _i64)
_sig("oracle_assemble", _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p,
     _p, _p, _p, _p, _p, _p, _p, _i32)
_sig("oracle_qoi", _i32, _i32, _i32, _i32, _i32, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _i32,
     C.POINTER(C.c_double), _p, _p)
_sig("oracle_dbc", None, _i64, _p, _p, _i32, _p, _p, _p, _p, _p, _p)
_sig("oracle_traction", None, _i32, _i64, _p, _p, _p, _p, _p)
_sig("oracle_set_coarse", None, _p, _p, _i64, _i64)
_sig("oracle_parts_create", _p, _i32, _i64, _i64, _p, _p, _i32)
_sig("oracle_parts_destroy", None, _p)
_sig("oracle_parts_rows", None, _p, _p)
_sig("oracle_assemble_parts", _i32, _p, _i32, _i32, _i32, _i32, _i32, _i32, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
     _p, _p, _p, _p)
_sig("oracle_compute_error", None, _i32, _i64, _p, _p, _p)
_sig("oracle_sum_contribs", C.c_double, _i64, _p, _p)
_sig("oracle_set_error", None, _i64, _p, _p, _i64, _p)


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    return a.ctypes.data_as(_p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32a(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def graph(order, num_nodes, conn):
    """Brute-force clique-union CSR: rows in lid order, sorted columns."""
    conn = _i32a(conn)
    npe = 4 if order == 1 else 10
    rp = np.empty(4 * num_nodes + 1, dtype=np.int64)
    nnz = lib.oracle_graph(npe, num_nodes, conn.shape[0], _ptr(conn), _ptr(rp), None, 0)
    cols = np.empty(nnz, dtype=np.int32)
    lib.oracle_graph(npe, num_nodes, conn.shape[0], _ptr(conn), _ptr(rp), _ptr(cols), nnz)
    return rp, cols


class Mesh:
    """Arrays describing one discretization for the oracle."""

    def __init__(self, order, coords, conn, elem_set=None):
        self.order = order
        self.coords = _f64(coords)
        self.conn = _i32a(conn)
        self.num_nodes = self.coords.shape[0]
        self.num_elems = self.conn.shape[0]
        self.elem_set = _i32a(elem_set) if elem_set is not None else np.zeros(self.num_elems, dtype=np.int32)
        self.row_ptr, self.cols = graph(order, self.num_nodes, self.conn)
        self.num_dofs = 4 * self.num_nodes
        self.nnz = int(self.row_ptr[-1])


class States:
    def __init__(self, mesh: Mesh, model: int, quad_order: int):
        nq = 1 if quad_order == 1 else 4
        n = mesh.num_elems * nq
        self.sigma = np.zeros((9, n))
        self.eqps = np.zeros(n)
        self.eqps_old = np.zeros(n)
        eye = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float64)[:, None]
        self.Fp = np.repeat(eye, n, axis=1).copy()
        self.Fp_old = self.Fp.copy()

    def update(self):
        self.eqps_old[:] = self.eqps
        self.Fp_old[:] = self.Fp


class OracleError(RuntimeError):
    pass


def assemble(kind, mesh: Mesh, model, mats, x, states: States, quad_order=1, stabilization=True, save=True,
             R=None, vals=None, z_diff=None, z_coarse=None, threads=1):
    """One oracle assembly pass; returns (R, vals) (accumulating into given arrays)."""
    mats = np.ascontiguousarray(np.asarray(mats, dtype=np.float64).reshape(-1, 5))
    x = _f64(x)
    if R is None:
        R = np.zeros(mesh.num_dofs)
    if vals is None and kind in (K_JACOBIAN, K_ADJOINT):
        vals = np.zeros(mesh.nnz)
    st = lib.oracle_assemble(kind, mesh.order, quad_order, model, int(bool(stabilization)), int(bool(save)),
                             mesh.num_nodes, mesh.num_elems, _ptr(mesh.coords), _ptr(mesh.conn),
                             _ptr(mesh.elem_set), _ptr(mats), _ptr(x), _ptr(mesh.row_ptr), _ptr(mesh.cols),
                             _ptr(R), _ptr(vals), _ptr(states.Fp_old), _ptr(states.eqps_old), _ptr(states.Fp),
                             _ptr(states.eqps), _ptr(states.sigma), _ptr(_f64(z_diff)), _ptr(_f64(z_coarse)),
                             threads)
    if st == 3:
        raise OracleError("J2: return mapping failed")
    if st != 0:
        raise OracleError(f"oracle_assemble failed with status {st}")
    return R, vals


def qoi(kind, mesh: Mesh, model, mats, x, states: States, quad_order=1, qoi_set=-1, derivative=False):
    mats = np.ascontiguousarray(np.asarray(mats, dtype=np.float64).reshape(-1, 5))
    x = _f64(x)
    v = C.c_double(0.0)
    ev = np.zeros(mesh.num_elems)
    dm = np.zeros(mesh.num_dofs) if derivative else None
    st = lib.oracle_qoi(kind, mesh.order, quad_order, model, mesh.num_nodes, mesh.num_elems, _ptr(mesh.coords),
                        _ptr(mesh.conn), _ptr(mesh.elem_set), _ptr(mats), _ptr(x), _ptr(states.Fp_old),
                        _ptr(states.eqps_old), qoi_set, C.byref(v), _ptr(ev), _ptr(dm))
    if st != 0:
        raise OracleError(f"oracle_qoi failed with status {st}")
    return v.value, ev, dm


def dbc(mesh: Mesh, nodes, values, comp, x, R, dMdu=None, vals=None):
    nodes = _i32a(nodes)
    values = _f64(values)
    lib.oracle_dbc(nodes.size, _ptr(nodes), _ptr(values), comp, _ptr(_f64(x)), _ptr(R), _ptr(dMdu),
                   _ptr(mesh.row_ptr), _ptr(mesh.cols), _ptr(vals))


# local vertices of tet face f (apf's order) and the tet edges in the P2 node
# order {01, 12, 20, 03, 13, 23}
_TET_FACES = ((0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3))
_TET_EDGES = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))


def boundary_faces(mesh: Mesh, nodes):
    """Faces on the mesh boundary with all three vertices in `nodes`, as node
    lists (3 vertices, then for P2 the nodes of edges v0v1, v1v2, v2v0), in
    ascending (element, local face) order: the side set of a box face."""
    inset = np.zeros(mesh.num_nodes, dtype=bool)
    inset[np.asarray(nodes)] = True
    conn = mesh.conn.reshape(mesh.num_elems, -1)
    cand = []
    for f, (a, b, c) in enumerate(_TET_FACES):
        ok = np.nonzero(inset[conn[:, a]] & inset[conn[:, b]] & inset[conn[:, c]])[0]
        cand += [(int(e), f) for e in ok]
    keys = [tuple(sorted(conn[e, list(_TET_FACES[f])])) for e, f in cand]
    count = {}
    for k in keys:
        count[k] = count.get(k, 0) + 1
    out = []
    for (e, f), k in sorted(zip(cand, keys)):
        if count[k] != 1:
            continue
        lv = _TET_FACES[f]
        row = [int(conn[e, v]) for v in lv]
        if mesh.order == 2:
            for j in range(3):
                pair = {lv[j], lv[(j + 1) % 3]}
                row.append(int(conn[e, 4 + next(i for i, ed in enumerate(_TET_EDGES) if set(ed) == pair)]))
        out.append(row)
    return np.asarray(out, dtype=np.int32).reshape(-1, 3 if mesh.order == 1 else 6)


def traction(mesh: Mesh, faces, T, t, R, z=None):
    """One traction entry (src/goal_tbcs.cpp:30-66): T(x, y, z, t) -> 3
    components, evaluated at each face centroid; R -= T W |face| per face node
    (W = N_a, or z_i N_a with a DOF vector z)."""
    faces = _i32a(faces)
    nfn = 3 if mesh.order == 1 else 6
    c = mesh.coords.reshape(-1, 3)
    xc = c[faces.reshape(-1, nfn)[:, :3]].mean(axis=1)
    Tv = np.ascontiguousarray([T(x[0], x[1], x[2], t) for x in xc], dtype=np.float64).reshape(-1, 3)
    lib.oracle_traction(mesh.order, faces.size // nfn, _ptr(faces), _ptr(mesh.coords), _ptr(Tv),
                        _ptr(_f64(z)), _ptr(R))


def set_coarse(z, edge_parents, num_old):
    ep = _i32a(edge_parents)
    lib.oracle_set_coarse(_ptr(z), _ptr(ep), num_old, ep.size // 2)


def compute_error(mesh: Mesh, e):
    eta = np.zeros(mesh.num_elems)
    lib.oracle_compute_error(mesh.order, mesh.num_elems, _ptr(mesh.conn), _ptr(_f64(e)), _ptr(eta))
    return eta


def sum_contribs(vertex_nodes, e):
    vn = _i32a(vertex_nodes)
    return lib.oracle_sum_contribs(vn.size, _ptr(vn), _ptr(_f64(e)))


def set_error(parent, child, num_base):
    parent = _i32a(parent)
    base = np.zeros(num_base)
    lib.oracle_set_error(parent.size, _ptr(parent), _ptr(_f64(child)), num_base, _ptr(base))
    return base


class Parts:
    """The rank-per-part assembly plan of a mesh (one part per thread: a
    contiguous element range and only the rows its elements touch)."""

    def __init__(self, mesh: Mesh, threads: int):
        self.mesh = mesh
        self.threads = threads
        self.h = lib.oracle_parts_create(mesh.order, mesh.num_nodes, mesh.num_elems, _ptr(mesh.conn),
                                         _ptr(mesh.row_ptr), threads)

    def rows(self):
        """(rows held over all parts, rows of the mesh)."""
        out = np.zeros(2, dtype=np.int64)
        lib.oracle_parts_rows(self.h, _ptr(out))
        return int(out[0]), int(out[1])

    def assemble(self, kind, model, mats, x, states: States, quad_order=1, stabilization=True, save=True,
                 R=None, vals=None):
        """One assembly over the parts, exported into R and vals (added to
        what they hold).  Returns (R, vals, seconds) with seconds = the longest
        part's (zero, element assembly, export) times."""
        m = self.mesh
        mats = np.ascontiguousarray(np.asarray(mats, dtype=np.float64).reshape(-1, 5))
        x = _f64(x)
        if R is None:
            R = np.zeros(m.num_dofs)
        if vals is None and kind in (K_JACOBIAN, K_ADJOINT):
            vals = np.zeros(m.nnz)
        sec = np.zeros(3)
        st = lib.oracle_assemble_parts(self.h, kind, m.order, quad_order, model, int(bool(stabilization)),
                                       int(bool(save)), _ptr(m.coords), _ptr(m.conn), _ptr(m.elem_set), _ptr(mats),
                                       _ptr(x), _ptr(m.cols), _ptr(R), _ptr(vals), _ptr(states.Fp_old),
                                       _ptr(states.eqps_old), _ptr(states.Fp), _ptr(states.eqps),
                                       _ptr(states.sigma), _ptr(sec))
        if st != 0:
            raise RuntimeError(f"oracle_assemble_parts failed: {st}")
        return R, vals, sec

    def __del__(self):
        if getattr(self, "h", None):
            lib.oracle_parts_destroy(self.h)
            self.h = None
