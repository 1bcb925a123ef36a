"""Python bindings of goal_amd's C ABI (include/goal_amd.h) over ctypes.

The product is the shared library libgoal_amd.so (C++ host code + gfx950 HIP
kernels).  This module only marshals numpy arrays to it; it performs no
computation of its own and has no fallback: if the library is missing, import
fails with an error that says how to build it.

Objects mirror the reference's classes: Disc (src/goal_disc.hpp:27-93),
States (src/goal_states.hpp), SolInfo (src/goal_sol_info.hpp:12-39),
Mechanics (src/goal_mechanics.hpp:21-48), Nested (src/goal_nested.hpp:17-42).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
# GOAL_AMD_LIB selects another build of the same library (e.g. a tuning variant).
LIB_PATH = Path(os.environ["GOAL_AMD_LIB"]) if os.environ.get("GOAL_AMD_LIB") else _HERE / "libgoal_amd.so"

OK = 0
ERR_INVALID = 1
ERR_HIP = 2
ERR_J2_RETURN_MAP = 3
ERR_NO_DEVICE = 4
ERR_ALLOC = 5

MODEL_ELASTIC = 0
MODEL_NEOHOOKEAN = 1
MODEL_J2 = 2
QOI_AVG_DISP = 0
QOI_AVG_VM = 1


class GoalError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"goal_amd error {code}: {msg}")
        self.code = code


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python goal_amd/build.py` "
            "(or __graft_entry__.build()); there is no non-HIP fallback"
        )
    return C.CDLL(str(LIB_PATH))


lib = _load()

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_d = C.c_double
_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class Material(C.Structure):
    _fields_ = [("model", C.c_int), ("E", C.c_double), ("nu", C.c_double), ("K", C.c_double),
                ("Y", C.c_double), ("c0", C.c_double)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("goal_last_error", C.c_char_p)
_sig("goal_version", C.c_char_p)
_sig("goal_device_alloc", _i32, _i64, C.POINTER(_p))
_sig("goal_device_free", _i32, _p)
_sig("goal_memcpy_h2d", _i32, _p, _p, _i64)
_sig("goal_memcpy_d2h", _i32, _p, _p, _i64)
_sig("goal_sync", _i32)
_sig("goal_event_record", _i32, _i32)
_sig("goal_event_elapsed", _i32, _i32, _i32, C.POINTER(C.c_float))
_sig("goal_device_copy", _i32, _p, _p, _i64)
_sig("goal_device_info", _i32, C.c_char_p, _i32, C.POINTER(C.c_int))
_sig("goal_mesh_box", _i32, _i32, _i32, _i32, _d, _d, _d, _i32, _d, C.c_uint64, _p, _p, _p)
_sig("goal_mesh_graph", _i32, _i32, _i64, _i64, _p, _i64p, _p, _p)
_sig("goal_mesh_refine", _i32, _i64, _p, _i64, _p, _p, _p, _p, _p, _p)
_sig("goal_disc_create", _i32, _i32, _i64, _p, _i64, _p, _p, C.POINTER(_p))
_sig("goal_size_field", _i32, _i64, _p, _i64, _p, _p, _i64, _i32, _p)
_sig("goal_adapt_create", _i32, _i64, _p, _i64, _p, _p, _d, _i32, C.POINTER(_p))
_sig("goal_adapt_destroy", _i32, _p)
_sig("goal_adapt_sizes", _i32, _p, _p)
_sig("goal_adapt_get", _i32, _p, _p, _p, _p, _p)
_sig("goal_adapt_transfer", _i32, _p, _i32, _p, _p)
_sig("goal_disc_create_box", _i32, _i32, _i32, _i32, _d, _d, _d, _i32, _d, C.c_uint64, C.POINTER(_p))
_sig("goal_disc_destroy", _i32, _p)
_sig("goal_disc_sizes", _i32, _p, _p)
_sig("goal_disc_get_mesh", _i32, _p, _p, _p, _p)
_sig("goal_disc_get_graph", _i32, _p, _p, _p)
_sig("goal_disc_graph_dev", _i32, _p, C.POINTER(_p), C.POINTER(_p))
_sig("goal_disc_mesh_dev", _i32, _p, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p))
_sig("goal_disc_get_lids", _i32, _p, _p)
_sig("goal_disc_add_node_set", _i32, _p, C.c_char_p, _i64, _p)
_sig("goal_disc_node_set_size", _i32, _p, C.c_char_p, _i64p)
_sig("goal_disc_get_node_set", _i32, _p, C.c_char_p, _p)
_sig("goal_disc_set_elem_sets", _i32, _p, _i32, _p)
_sig("goal_disc_set_soln", _i32, _p, _p)
_sig("goal_disc_get_soln", _i32, _p, _p)
_sig("goal_disc_add_soln", _i32, _p, _p)
_sig("goal_disc_add_soln_dev", _i32, _p, _p)
_sig("goal_disc_soln_dev", _p, _p)
_sig("goal_states_create", _i32, _p, _i32, _i32, C.POINTER(_p))
_sig("goal_states_destroy", _i32, _p)
_sig("goal_states_update", _i32, _p)
_sig("goal_states_get", _i32, _p, C.c_char_p, _p)
_sig("goal_states_set", _i32, _p, C.c_char_p, _p)
_sig("goal_solinfo_create", _i32, _p, C.POINTER(_p))
_sig("goal_solinfo_destroy", _i32, _p)
_sig("goal_solinfo_zero_R", _i32, _p)
_sig("goal_solinfo_zero_dMdu", _i32, _p)
_sig("goal_solinfo_zero_dRdu", _i32, _p)
_sig("goal_solinfo_zero_all", _i32, _p)
_sig("goal_solinfo_get", _i32, _p, _p, _p, _p)
_sig("goal_solinfo_set", _i32, _p, _p, _p, _p)
_sig("goal_solinfo_dev", _i32, _p, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p))
_sig("goal_solinfo_norm_R", _i32, _p, _dp)
_sig("goal_solinfo_dot_R", _i32, _p, _p, _dp)
_sig("goal_linear_solve", _i32, _p, _p, _p, C.c_double, _i32, _i32, C.POINTER(_i32), _dp)
_sig("goal_linear_solve_pc", _i32, _p, _p, _p, C.c_double, _i32, _i32, _i32, C.POINTER(_i32), _dp)
PRECOND_BLOCK_JACOBI, PRECOND_AMG = 0, 1
_sig("goal_device_fill", _i32, _p, _i64, C.c_double)
_sig("goal_device_scale", _i32, _p, _p, _i64, C.c_double)
_sig("goal_mech_create", _i32, _p, _p, _i32, _p, _i32, _i32, C.POINTER(_p))
_sig("goal_mech_destroy", _i32, _p)
_sig("goal_mech_set_option", _i32, _p, C.c_char_p, _i64)
_sig("goal_mech_timing", _i32, _p, C.POINTER(C.c_float), C.POINTER(C.c_float))
_sig("goal_mech_status", _i32, _p, C.POINTER(C.c_int))
_sig("goal_asm_residual", _i32, _p, _p, _i32)
_sig("goal_asm_jacobian", _i32, _p, _p, _i32, _i32)
_sig("goal_asm_qoi", _i32, _p, _p, _i32, _i32, _dp, _p)
_sig("goal_asm_localize", _i32, _p, _p, _p, _p)
_sig("goal_eval", _i32, C.c_char_p, _d, _d, _d, _d, _dp)
_sig("goal_dbc_apply", _i32, _p, _p, _i32, C.c_char_p, C.c_char_p, _d, _i32)
_sig("goal_dbc_apply_values", _i32, _p, _p, _i32, _i64, _p, _p, _i32)
_sig("goal_tbc_apply", _i32, _p, _p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, _d, _p)
_sig("goal_disc_add_side_set", _i32, _p, C.c_char_p, _i64, _p, _p)
_sig("goal_mesh_file_read", _i32, C.c_char_p, C.c_char_p, C.c_char_p, _i32, C.POINTER(_p))
_sig("goal_mesh_file_destroy", _i32, _p)
_sig("goal_mesh_file_sizes", _i32, _p, _i64p, _i64p, C.POINTER(C.c_int))
_sig("goal_mesh_file_get_mesh", _i32, _p, _p, _p, _p)
_sig("goal_mesh_file_elem_set_name", _i32, _p, _i32, C.c_char_p, _i64)
_sig("goal_mesh_file_node_set", _i32, _p, C.c_char_p, _i64p, _p)
_sig("goal_mesh_file_side_set", _i32, _p, C.c_char_p, _i64p, _p, _p)
_sig("goal_disc_create_from_file", _i32, _p, C.POINTER(_p))
_sig("goal_disc_side_set_size", _i32, _p, C.c_char_p, _i64p)
_sig("goal_disc_get_side_set", _i32, _p, C.c_char_p, _p, _p)
_sig("goal_batch_create", _i32, _i32, _p, _p, C.POINTER(_p))
_sig("goal_batch_destroy", _i32, _p)
_sig("goal_batch_tune", _i32, _p, _i32, _p, _p)
_sig("goal_batch_capture", _i32, _p)
_sig("goal_batch_run", _i32, _p)
_sig("goal_learn_features", _i32, _p, _d, _p)
_sig("goal_mlp_create", _i32, C.POINTER(_p))
_sig("goal_mlp_destroy", _i32, _p)
_sig("goal_mlp_set", _i32, _p, _p, _p)
_sig("goal_mlp_get", _i32, _p, _p, _p, _p)
_sig("goal_mlp_forward", _i32, _p, _p, _i64, _p, _i64p)
_sig("goal_mlp_gradient", _i32, _p, _p, _p, _p, _i64, C.POINTER(C.c_float), _p)
_sig("goal_mlp_train", _i32, _p, _p, _p, _p, _i64, _i32, C.c_float, C.POINTER(C.c_float))
_sig("goal_nested_create", _i32, _p, C.POINTER(_p))
_sig("goal_nested_destroy", _i32, _p)
_sig("goal_nested_disc", _p, _p)
_sig("goal_nested_get_maps", _i32, _p, _p, _p, _i64p)
_sig("goal_nested_transfer_soln", _i32, _p)
_sig("goal_nested_transfer_states", _i32, _p, _p, _p)
_sig("goal_dwr_set_coarse", _i32, _p, _p)
_sig("goal_dwr_subtract", _i32, _p, _p, _p, _i64)
_sig("goal_dwr_compute_error", _i32, _p, _p, _p)
_sig("goal_dwr_sum_contribs", _i32, _p, _p, _dp)
_sig("goal_dwr_set_error", _i32, _p, _p, _p)


def check(status: int) -> None:
    if status != OK:
        raise GoalError(status, lib.goal_last_error().decode())


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_p)


def version() -> str:
    return lib.goal_version().decode()


def device_info() -> tuple[str, int]:
    buf = C.create_string_buffer(256)
    cus = C.c_int(0)
    check(lib.goal_device_info(buf, 256, C.byref(cus)))
    return buf.value.decode(), cus.value


def sync() -> None:
    check(lib.goal_sync())


def event_record(slot: int) -> None:
    check(lib.goal_event_record(slot))


def event_elapsed(a: int, b: int) -> float:
    ms = C.c_float(0.0)
    check(lib.goal_event_elapsed(a, b, C.byref(ms)))
    return float(ms.value)


def device_copy(dst: "DeviceArray", src: "DeviceArray", n: int | None = None) -> None:
    check(lib.goal_device_copy(dst.ptr, src.ptr, dst.n if n is None else n))


def eval_expr(expr: str, x: float, y: float, z: float, t: float) -> float:
    out = C.c_double(0.0)
    check(lib.goal_eval(expr.encode(), x, y, z, t, C.byref(out)))
    return out.value


def mesh_box(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, order=1, jitter=0.0, seed=0):
    """Host-only Kuhn box mesh: (coords [nn,3], conn [ne, 4|10])."""
    sizes = np.zeros(2, dtype=np.int64)
    check(lib.goal_mesh_box(nx, ny, nz, lx, ly, lz, order, jitter, seed, _ptr(sizes), None, None))
    coords = np.empty((sizes[0], 3))
    conn = np.empty((sizes[1], 4 if order == 1 else 10), dtype=np.int32)
    check(lib.goal_mesh_box(nx, ny, nz, lx, ly, lz, order, jitter, seed, _ptr(sizes), _ptr(coords), _ptr(conn)))
    return coords, conn


def mesh_graph(order, num_nodes, conn):
    """Host CSR graph (row_ptr int64, cols int32) built by the library."""
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    nnz = C.c_int64(0)
    check(lib.goal_mesh_graph(order, num_nodes, conn.shape[0], _ptr(conn), C.byref(nnz), None, None))
    rp = np.empty(4 * num_nodes + 1, dtype=np.int64)
    cols = np.empty(nnz.value, dtype=np.int32)
    check(lib.goal_mesh_graph(order, num_nodes, conn.shape[0], _ptr(conn), C.byref(nnz), _ptr(rp), _ptr(cols)))
    return rp, cols


def mesh_refine(coords, conn):
    """Host 1:8 refinement: (fine_coords, fine_conn, parent, edge_parents)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    sizes = np.zeros(3, dtype=np.int64)
    check(lib.goal_mesh_refine(coords.shape[0], _ptr(coords), conn.shape[0], _ptr(conn), _ptr(sizes),
                               None, None, None, None))
    fc = np.empty((sizes[0], 3))
    fconn = np.empty((sizes[1], 4), dtype=np.int32)
    parent = np.empty(sizes[1], dtype=np.int32)
    ep = np.empty((sizes[2], 2), dtype=np.int32)
    check(lib.goal_mesh_refine(coords.shape[0], _ptr(coords), conn.shape[0], _ptr(conn), _ptr(sizes),
                               _ptr(fc), _ptr(fconn), _ptr(parent), _ptr(ep)))
    return fc, fconn, parent, ep


def size_field(coords, conn, eta, target: int, p_order: int = 1) -> np.ndarray:
    """Per-vertex target size from per-element indicators (linear mesh)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    eta = np.ascontiguousarray(eta, dtype=np.float64)
    out = np.empty(coords.shape[0])
    check(lib.goal_size_field(coords.shape[0], _ptr(coords), conn.shape[0], _ptr(conn), _ptr(eta), int(target),
                              p_order, _ptr(out)))
    return out


class Adaptation:
    """Refinement of a linear mesh toward a vertex size field (host)."""

    def __init__(self, coords, conn, vtx_size, split_ratio: float = 1.5, max_passes: int = 8):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        vtx_size = np.ascontiguousarray(vtx_size, dtype=np.float64)
        self.h = _p()
        check(lib.goal_adapt_create(coords.shape[0], _ptr(coords), conn.shape[0], _ptr(conn), _ptr(vtx_size),
                                    float(split_ratio), int(max_passes), C.byref(self.h)))
        sz = np.zeros(5, dtype=np.int64)
        check(lib.goal_adapt_sizes(self.h, _ptr(sz)))
        self.num_nodes, self.num_elems, self.num_new, self.splits, self.passes = (int(v) for v in sz)
        self.num_old = self.num_nodes - self.num_new
        self.coords = np.empty((self.num_nodes, 3))
        self.conn = np.empty((self.num_elems, 4), dtype=np.int32)
        self.parent = np.empty(self.num_elems, dtype=np.int32)
        self.new_vertex_parents = np.empty((self.num_new, 2), dtype=np.int32)
        check(lib.goal_adapt_get(self.h, _ptr(self.coords), _ptr(self.conn), _ptr(self.parent),
                                 _ptr(self.new_vertex_parents)))

    def transfer(self, field) -> np.ndarray:
        """Nodal field [num_old, k] -> [num_nodes, k] (new vertices: edge means)."""
        f = np.ascontiguousarray(field, dtype=np.float64)
        k = 1 if f.ndim == 1 else f.shape[1]
        out = np.empty((self.num_nodes, k) if f.ndim > 1 else self.num_nodes)
        check(lib.goal_adapt_transfer(self.h, k, _ptr(f), _ptr(out)))
        return out

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.goal_adapt_destroy(self.h)
            self.h = _p()


class DeviceArray:
    """A device buffer of float64 owned by this object."""

    def __init__(self, n: int):
        self.n = int(n)
        self.ptr = _p()
        check(lib.goal_device_alloc(8 * self.n, C.byref(self.ptr)))

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.size)
        check(lib.goal_memcpy_h2d(d.ptr, _ptr(a), 8 * a.size))
        return d

    def fill(self, value: float) -> None:
        check(lib.goal_device_fill(self.ptr, self.n, float(value)))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float64)
        check(lib.goal_memcpy_d2h(_ptr(out), self.ptr, 8 * self.n))
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib.goal_device_free(self.ptr)
            self.ptr = _p()


class DeviceBuffer:
    """A device buffer of n items of a numpy dtype, owned by this object."""

    def __init__(self, n: int, dtype=np.float32):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.ptr = _p()
        check(lib.goal_device_alloc(max(1, self.n) * self.dtype.itemsize, C.byref(self.ptr)))

    @classmethod
    def from_host(cls, a: np.ndarray, dtype=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size > self.n:
            raise ValueError("DeviceBuffer.upload: array larger than the buffer")
        check(lib.goal_memcpy_h2d(self.ptr, _ptr(a), a.nbytes))

    def numpy(self, n: int | None = None) -> np.ndarray:
        out = np.empty(self.n if n is None else int(n), dtype=self.dtype)
        check(lib.goal_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib.goal_device_free(self.ptr)
            self.ptr = _p()


MLP_IN, MLP_HIDDEN, MLP_PARAMS, MLP_NORMS = 12, 64, 5057, 51


def learn_features(mech: "Mechanics", qoi_x: float, out: DeviceBuffer | None = None) -> DeviceBuffer:
    """Per-element features [num_elems][12] (FP32, device) of mech's disc."""
    n = mech.disc.num_elems * MLP_IN
    if out is None or out.n < n:
        out = DeviceBuffer(n, np.float32)
    check(lib.goal_learn_features(mech.h, float(qoi_x), out.ptr))
    return out


class MLP:
    """12 -> 64 -> 64 -> 1 ReLU network on the device (BF16 weights)."""

    def __init__(self):
        self.h = _p()
        check(lib.goal_mlp_create(C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.goal_mlp_destroy(self.h)
            self.h = _p()

    def set(self, params=None, norms=None) -> None:
        pa = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        na = None if norms is None else np.ascontiguousarray(norms, dtype=np.float32)
        if pa is not None and pa.size != MLP_PARAMS:
            raise ValueError("MLP.set: params must have %d entries" % MLP_PARAMS)
        if na is not None and na.size != MLP_NORMS:
            raise ValueError("MLP.set: norms must have %d entries" % MLP_NORMS)
        check(lib.goal_mlp_set(self.h, _ptr(pa), _ptr(na)))

    def get(self):
        """(FP32 master parameters, BF16 weights as uint16, normalization)."""
        pa = np.empty(MLP_PARAMS, dtype=np.float32)
        wq = np.empty(MLP_PARAMS, dtype=np.uint16)
        na = np.empty(MLP_NORMS, dtype=np.float32)
        check(lib.goal_mlp_get(self.h, _ptr(pa), _ptr(wq), _ptr(na)))
        return pa, wq, na

    def forward(self, x: DeviceBuffer, n: int, y: DeviceBuffer, count: bool = True) -> int | None:
        """y[:n] = predictions for rows x[:n]; returns rows outside the
        training box (count=False: asynchronous, returns None)."""
        if x.n < n * MLP_IN or y.n < n:
            raise ValueError("MLP.forward: buffers too small")
        out = C.c_int64(0)
        check(lib.goal_mlp_forward(self.h, x.ptr, int(n), y.ptr, C.byref(out) if count else None))
        return out.value if count else None

    def gradient(self, x: DeviceBuffer, y: DeviceBuffer, idx: DeviceBuffer, n: int):
        loss = C.c_float(0.0)
        grad = np.empty(MLP_PARAMS, dtype=np.float32)
        check(lib.goal_mlp_gradient(self.h, x.ptr, y.ptr, idx.ptr, int(n), C.byref(loss), _ptr(grad)))
        return loss.value, grad

    def train(self, x: DeviceBuffer, y: DeviceBuffer, idx: DeviceBuffer, batch: int, steps: int, lr: float) -> float:
        if idx.n < batch * steps:
            raise ValueError("MLP.train: index buffer shorter than batch * steps")
        loss = C.c_float(0.0)
        check(lib.goal_mlp_train(self.h, x.ptr, y.ptr, idx.ptr, int(batch), int(steps), float(lr), C.byref(loss)))
        return loss.value


class Disc:
    """Mesh + node-major DOF numbering + CSR graph on the device."""

    def __init__(self, handle, owner=None):
        self.h = handle
        self._owner = owner  # nested discs are owned by their Nested
        s = np.zeros(9, dtype=np.int64)
        check(lib.goal_disc_sizes(self.h, _ptr(s)))
        (self.num_nodes, self.num_elems, self.num_dofs, self.nnz, self.nodes_per_elem,
         self.num_vertices, self.order, self.num_elem_sets, self.max_degree) = (int(v) for v in s)

    @classmethod
    def box(cls, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, order=1, jitter=0.0, seed=0):
        h = _p()
        check(lib.goal_disc_create_box(nx, ny, nz, lx, ly, lz, order, jitter, seed, C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, order, coords, conn, elem_set=None):
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        es = None if elem_set is None else np.ascontiguousarray(elem_set, dtype=np.int32)
        h = _p()
        check(lib.goal_disc_create(order, coords.shape[0], _ptr(coords), conn.shape[0], _ptr(conn), _ptr(es),
                                   C.byref(h)))
        return cls(h)

    def __del__(self):
        if self._owner is None and getattr(self, "h", None):
            lib.goal_disc_destroy(self.h)
            self.h = None

    def mesh(self):
        coords = np.empty((self.num_nodes, 3))
        conn = np.empty((self.num_elems, self.nodes_per_elem), dtype=np.int32)
        es = np.empty(self.num_elems, dtype=np.int32)
        check(lib.goal_disc_get_mesh(self.h, _ptr(coords), _ptr(conn), _ptr(es)))
        return coords, conn, es

    def graph(self):
        rp = np.empty(self.num_dofs + 1, dtype=np.int64)
        cols = np.empty(self.nnz, dtype=np.int32)
        check(lib.goal_disc_get_graph(self.h, _ptr(rp), _ptr(cols)))
        return rp, cols

    def graph_dev(self):
        rp, cols = _p(), _p()
        check(lib.goal_disc_graph_dev(self.h, C.byref(rp), C.byref(cols)))
        return rp.value, cols.value

    def mesh_dev(self):
        c, cn, ps = _p(), _p(), _p()
        check(lib.goal_disc_mesh_dev(self.h, C.byref(c), C.byref(cn), C.byref(ps)))
        return c.value, cn.value, ps.value

    def lids(self):
        out = np.empty((self.num_elems, 4 * self.nodes_per_elem), dtype=np.int32)
        check(lib.goal_disc_get_lids(self.h, _ptr(out)))
        return out

    def node_set(self, name: str) -> np.ndarray:
        n = C.c_int64(0)
        check(lib.goal_disc_node_set_size(self.h, name.encode(), C.byref(n)))
        out = np.empty(n.value, dtype=np.int32)
        if n.value:
            check(lib.goal_disc_get_node_set(self.h, name.encode(), _ptr(out)))
        return out

    def add_node_set(self, name: str, nodes) -> None:
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        check(lib.goal_disc_add_node_set(self.h, name.encode(), nodes.size, _ptr(nodes)))

    def side_set(self, name: str):
        """(elements, local faces) of a side set; a node set's name gives its
        boundary faces."""
        n = C.c_int64(0)
        check(lib.goal_disc_side_set_size(self.h, name.encode(), C.byref(n)))
        elems = np.empty(n.value, dtype=np.int32)
        faces = np.empty(n.value, dtype=np.int32)
        if n.value:
            check(lib.goal_disc_get_side_set(self.h, name.encode(), _ptr(elems), _ptr(faces)))
        return elems, faces

    def add_side_set(self, name: str, elems, faces) -> None:
        elems = np.ascontiguousarray(elems, dtype=np.int32)
        faces = np.ascontiguousarray(faces, dtype=np.int32)
        if elems.shape != faces.shape:
            raise ValueError("one local face per element")
        check(lib.goal_disc_add_side_set(self.h, name.encode(), elems.size, _ptr(elems), _ptr(faces)))

    def set_elem_sets(self, num_sets: int, elem_set) -> None:
        es = np.ascontiguousarray(elem_set, dtype=np.int32)
        check(lib.goal_disc_set_elem_sets(self.h, num_sets, _ptr(es)))
        self.num_elem_sets = num_sets

    def set_soln(self, x) -> None:
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == self.num_dofs
        check(lib.goal_disc_set_soln(self.h, _ptr(x)))

    def soln(self) -> np.ndarray:
        out = np.empty(self.num_dofs)
        check(lib.goal_disc_get_soln(self.h, _ptr(out)))
        return out

    def add_soln(self, du) -> None:
        du = np.ascontiguousarray(du, dtype=np.float64)
        check(lib.goal_disc_add_soln(self.h, _ptr(du)))

    def add_soln_dev(self, du: "DeviceArray") -> None:
        check(lib.goal_disc_add_soln_dev(self.h, du.ptr))

    def soln_dev(self):
        return lib.goal_disc_soln_dev(self.h)


class States:
    def __init__(self, disc: Disc, model: int, quad_order: int):
        self.disc = disc
        self.h = _p()
        check(lib.goal_states_create(disc.h, model, quad_order, C.byref(self.h)))
        self.nq = 1 if quad_order == 1 else 4
        self.model = model

    def __del__(self):
        if getattr(self, "h", None):
            lib.goal_states_destroy(self.h)
            self.h = None

    def update(self) -> None:
        check(lib.goal_states_update(self.h))

    def get(self, name: str) -> np.ndarray:
        nc = 1 if name.startswith("eqps") else 9
        out = np.empty((nc, self.disc.num_elems * self.nq))
        check(lib.goal_states_get(self.h, name.encode(), _ptr(out)))
        return out

    def set(self, name: str, a) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(lib.goal_states_set(self.h, name.encode(), _ptr(a)))


class SolInfo:
    def __init__(self, disc: Disc):
        self.disc = disc
        self.h = _p()
        check(lib.goal_solinfo_create(disc.h, C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None):
            lib.goal_solinfo_destroy(self.h)
            self.h = None

    def zero_R(self):
        check(lib.goal_solinfo_zero_R(self.h))

    def zero_dMdu(self):
        check(lib.goal_solinfo_zero_dMdu(self.h))

    def zero_dRdu(self):
        check(lib.goal_solinfo_zero_dRdu(self.h))

    def zero_all(self):
        check(lib.goal_solinfo_zero_all(self.h))

    def R(self) -> np.ndarray:
        out = np.empty(self.disc.num_dofs)
        check(lib.goal_solinfo_get(self.h, _ptr(out), None, None))
        return out

    def dMdu(self) -> np.ndarray:
        out = np.empty(self.disc.num_dofs)
        check(lib.goal_solinfo_get(self.h, None, _ptr(out), None))
        return out

    def vals(self) -> np.ndarray:
        out = np.empty(self.disc.nnz)
        check(lib.goal_solinfo_get(self.h, None, None, _ptr(out)))
        return out

    def set(self, R=None, dMdu=None, vals=None) -> None:
        conv = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (R, dMdu, vals)]
        check(lib.goal_solinfo_set(self.h, *[_ptr(a) for a in conv]))

    def dev(self):
        r, m, v = _p(), _p(), _p()
        check(lib.goal_solinfo_dev(self.h, C.byref(r), C.byref(m), C.byref(v)))
        return r.value, m.value, v.value

    def solve(self, b, x, rtol=1e-10, max_iters=2000, restart=100, precond=PRECOND_AMG):
        """dRdu x = b on the device (GMRES preconditioned by aggregation
        multigrid, or node block-Jacobi); x holds the initial guess.  b, x:
        DeviceArray or raw device pointers.  Returns (inner iterations,
        relative residual ||b - Ax|| / ||b - Ax0||)."""
        it = C.c_int(0)
        rr = C.c_double(0.0)
        bp = b.ptr if isinstance(b, DeviceArray) else b
        xp = x.ptr if isinstance(x, DeviceArray) else x
        check(lib.goal_linear_solve_pc(self.h, bp, xp, rtol, max_iters, restart, precond, C.byref(it),
                                       C.byref(rr)))
        return it.value, rr.value

    def norm_R(self) -> float:
        out = C.c_double(0.0)
        check(lib.goal_solinfo_norm_R(self.h, C.byref(out)))
        return out.value

    def dot_R(self, z_dev_ptr) -> float:
        out = C.c_double(0.0)
        check(lib.goal_solinfo_dot_R(self.h, z_dev_ptr, C.byref(out)))
        return out.value


@dataclass
class MaterialSpec:
    model: int
    E: float
    nu: float
    K: float = 0.0
    Y: float = 0.0
    c0: float = 1.0

    def as_struct(self) -> Material:
        return Material(self.model, self.E, self.nu, self.K, self.Y, self.c0)

    def as_row(self):
        return [self.E, self.nu, self.K, self.Y, self.c0]


class Mechanics:
    def __init__(self, disc: Disc, states: States, mats, stabilization=True, quad_order=1):
        self.disc, self.states = disc, states
        self.mats = list(mats)
        arr = (Material * len(self.mats))(*[m.as_struct() for m in self.mats])
        self.h = _p()
        check(lib.goal_mech_create(disc.h, states.h, len(self.mats), C.cast(arr, _p), int(bool(stabilization)),
                                   quad_order, C.byref(self.h)))
        self.quad_order = quad_order
        self.stabilization = bool(stabilization)

    def __del__(self):
        if getattr(self, "h", None):
            lib.goal_mech_destroy(self.h)
            self.h = None

    def set_option(self, key: str, value: int) -> None:
        check(lib.goal_mech_set_option(self.h, key.encode(), int(value)))

    def timing(self) -> tuple[float, float]:
        """(element pass ms, gather ms) of the last assembly; needs set_option("timing", 1)."""
        a, b = C.c_float(0.0), C.c_float(0.0)
        check(lib.goal_mech_timing(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def status(self) -> int:
        f = C.c_int(0)
        check(lib.goal_mech_status(self.h, C.byref(f)))
        return f.value

    def residual(self, s: SolInfo, save_states=True) -> None:
        check(lib.goal_asm_residual(self.h, s.h, int(bool(save_states))))

    def jacobian(self, s: SolInfo, adjoint=False, save_states=True) -> None:
        check(lib.goal_asm_jacobian(self.h, s.h, int(bool(adjoint)), int(bool(save_states))))

    def qoi(self, kind: int, s: SolInfo | None = None, elem_set: int = -1, elem_values=False):
        v = C.c_double(0.0)
        ev = np.empty(self.disc.num_elems) if elem_values else None
        check(lib.goal_asm_qoi(self.h, s.h if s is not None else None, kind, elem_set, C.byref(v), _ptr(ev)))
        return (v.value, ev) if elem_values else v.value

    def localize(self, s: SolInfo, z_diff_dev, z_coarse_dev) -> None:
        check(lib.goal_asm_localize(self.h, s.h, z_diff_dev, z_coarse_dev))


class Batch:
    """Jacobian-mode assemblies of many (Mechanics, SolInfo) pairs as one
    captured HIP graph (goal_batch_*)."""

    def __init__(self, mechs, infos):
        self.mechs, self.infos = list(mechs), list(infos)
        self.h = _p()
        mh = (_p * len(self.mechs))(*[m.h.value for m in self.mechs])
        sh = (_p * len(self.infos))(*[s.h.value for s in self.infos])
        check(lib.goal_batch_create(len(self.mechs), mh, sh, C.byref(self.h)))

    def tune(self, reps: int = 3):
        """Autotune each mesh's grid; returns (blocks, median ms) per mesh."""
        blocks = np.zeros(len(self.mechs), dtype=np.int32)
        ms = np.zeros(len(self.mechs), dtype=np.float32)
        check(lib.goal_batch_tune(self.h, reps, _ptr(blocks), _ptr(ms)))
        return blocks, ms

    def capture(self) -> None:
        check(lib.goal_batch_capture(self.h))

    def run(self) -> None:
        check(lib.goal_batch_run(self.h))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.goal_batch_destroy(self.h)
            self.h = _p()


def dbc_apply(disc: Disc, s: SolInfo, comp: int, node_set: str, expr: str, t: float, jacobian: bool) -> None:
    check(lib.goal_dbc_apply(disc.h, s.h, comp, node_set.encode(), expr.encode(), t, int(bool(jacobian))))


def dbc_apply_values(disc: Disc, s: SolInfo, comp: int, nodes, values, jacobian: bool) -> None:
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64)
    check(lib.goal_dbc_apply_values(disc.h, s.h, comp, nodes.size, _ptr(nodes), _ptr(values), int(bool(jacobian))))


class MeshFile:
    """goal's mesh inputs (serial SCOREC .smb, gmi .dmg, association file)
    read on the host (src/goal_disc.cpp:28-99, 347-396); needs no GPU."""

    def __init__(self, geom_file, mesh_file, assoc_file, order=1):
        self.h = _p()
        check(lib.goal_mesh_file_read(str(geom_file).encode(), str(mesh_file).encode(), str(assoc_file).encode(),
                                      order, C.byref(self.h)))
        nn, ne, ns = C.c_int64(0), C.c_int64(0), C.c_int(0)
        check(lib.goal_mesh_file_sizes(self.h, C.byref(nn), C.byref(ne), C.byref(ns)))
        self.num_nodes, self.num_elems, self.num_elem_sets = nn.value, ne.value, ns.value
        self.order = order
        self.elem_set_names = []
        for i in range(self.num_elem_sets):
            buf = C.create_string_buffer(256)
            check(lib.goal_mesh_file_elem_set_name(self.h, i, buf, 256))
            self.elem_set_names.append(buf.value.decode())

    def mesh(self):
        coords = np.empty((self.num_nodes, 3))
        conn = np.empty((self.num_elems, 4 if self.order == 1 else 10), dtype=np.int32)
        es = np.empty(self.num_elems, dtype=np.int32)
        check(lib.goal_mesh_file_get_mesh(self.h, _ptr(coords), _ptr(conn), _ptr(es)))
        return coords, conn, es

    def node_set(self, name: str) -> np.ndarray:
        n = C.c_int64(0)
        check(lib.goal_mesh_file_node_set(self.h, name.encode(), C.byref(n), None))
        out = np.empty(n.value, dtype=np.int32)
        check(lib.goal_mesh_file_node_set(self.h, name.encode(), C.byref(n), _ptr(out)))
        return out

    def side_set(self, name: str):
        n = C.c_int64(0)
        check(lib.goal_mesh_file_side_set(self.h, name.encode(), C.byref(n), None, None))
        elems = np.empty(n.value, dtype=np.int32)
        faces = np.empty(n.value, dtype=np.int32)
        check(lib.goal_mesh_file_side_set(self.h, name.encode(), C.byref(n), _ptr(elems), _ptr(faces)))
        return elems, faces

    def disc(self) -> "Disc":
        h = _p()
        check(lib.goal_disc_create_from_file(self.h, C.byref(h)))
        return Disc(h)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.goal_mesh_file_destroy(self.h)
            self.h = _p()


def tbc_apply(disc: Disc, s: SolInfo, side_set: str, exprs, t: float, weight=None) -> None:
    """One traction entry [side set, T_x, T_y, T_z] (src/goal_tbcs.cpp:30-66);
    weight: None (weight N_a) or a device DOF vector z (weight z_i N_a)."""
    tx, ty, tz = (str(e).encode() for e in exprs)
    w = None if weight is None else (weight.ptr if hasattr(weight, "ptr") else _p(int(weight)))
    check(lib.goal_tbc_apply(disc.h, s.h, side_set.encode(), tx, ty, tz, t, w))


class Nested:
    def __init__(self, base: Disc):
        self.base = base
        self.h = _p()
        check(lib.goal_nested_create(base.h, C.byref(self.h)))
        self.disc = Disc(lib.goal_nested_disc(self.h), owner=self)
        n = C.c_int64(0)
        check(lib.goal_nested_get_maps(self.h, None, None, C.byref(n)))
        self.num_new = n.value
        self.num_old = base.num_nodes

    def __del__(self):
        if getattr(self, "h", None):
            lib.goal_nested_destroy(self.h)
            self.h = None

    def maps(self):
        parent = np.empty(self.disc.num_elems, dtype=np.int32)
        ep = np.empty((self.num_new, 2), dtype=np.int32)
        check(lib.goal_nested_get_maps(self.h, _ptr(parent), _ptr(ep), None))
        return parent, ep

    def transfer_soln(self) -> None:
        check(lib.goal_nested_transfer_soln(self.h))

    def transfer_states(self, base_states: "States", fine_states: "States") -> None:
        check(lib.goal_nested_transfer_states(self.h, base_states.h, fine_states.h))

    def set_coarse(self, field: DeviceArray) -> None:
        check(lib.goal_dwr_set_coarse(self.h, field.ptr))

    def set_error(self, eta_child: DeviceArray, eta_base: DeviceArray) -> None:
        check(lib.goal_dwr_set_error(self.h, eta_child.ptr, eta_base.ptr))


def dwr_subtract(fine: DeviceArray, coarse: DeviceArray, out: DeviceArray) -> None:
    check(lib.goal_dwr_subtract(fine.ptr, coarse.ptr, out.ptr, out.n))


def dwr_compute_error(disc: Disc, e_dev_ptr, eta: DeviceArray) -> None:
    check(lib.goal_dwr_compute_error(disc.h, e_dev_ptr, eta.ptr))


def dwr_sum_contribs(disc: Disc, e_dev_ptr) -> float:
    out = C.c_double(0.0)
    check(lib.goal_dwr_sum_contribs(disc.h, e_dev_ptr, C.byref(out)))
    return out.value
