"""Host mesh layer, expression evaluator and the C ABI surface (CPU only)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from tests import npref

g = pytest.importorskip("goal_amd")

ROOT = Path(__file__).resolve().parents[1]


def test_every_declared_symbol_is_exported():
    header = (ROOT / "include" / "goal_amd.h").read_text()
    names = re.findall(r"^\s*(?:int|double\*|goal_disc|const char\*)\s+(goal_\w+)\s*\(", header, re.M)
    assert len(names) > 50
    lib = C.CDLL(str(g.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_eval_known_answer():
    # test/control.cpp:12-15 of the reference
    v = g.eval_expr("sin(x)*cos(y) + 2*exp(t) - z", 1.0, 2.0, 3.0, 4.0)
    assert abs(v - (np.sin(1.0) * np.cos(2.0) + 2.0 * np.exp(4.0) - 3.0)) < 1e-15
    assert g.eval_expr("0.01*t", 0, 0, 0, 3.0) == 0.01 * 3.0
    assert g.eval_expr("-x^2 + pow(y, 3) / 2", 2.0, 2.0, 0, 0) == -4.0 + 4.0
    assert g.eval_expr("max(x, y) - min(x, y)", 1.0, 5.0, 0, 0) == 4.0
    with pytest.raises(g.GoalError) as ei:
        g.eval_expr("sin(x", 0, 0, 0, 0)
    assert ei.value.code == g.ERR_INVALID


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("jitter", [0.0, 0.1])
def test_kuhn_box_matches_definition(order, jitter):
    c_lib, n_lib = g.mesh_box(3, 2, 2, 1.0, 2.0, 1.5, order=order, jitter=jitter, seed=0x5EED)
    c_ref, n_ref = npref.kuhn_box(3, 2, 2, 1.0, 2.0, 1.5, order, jitter, 0x5EED)
    assert np.array_equal(n_lib, n_ref)
    assert np.array_equal(c_lib, c_ref)


@pytest.mark.parametrize("order", [1, 2])
def test_kuhn_box_properties(order):
    n = 4
    coords, conn = g.mesh_box(n, n, n, order=order, jitter=0.1, seed=1)
    assert coords.shape[0] == (order * n + 1) ** 3
    assert conn.shape == (6 * n ** 3, 4 if order == 1 else 10)
    X = coords[conn[:, :4]]
    vol = np.einsum("ei,ei->e", np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), X[:, 3] - X[:, 0]) / 6
    assert vol.min() > 0
    assert abs(vol.sum() - 1.0) < 1e-13
    # conforming: every interior face shared by exactly two tets, boundary faces by one
    faces = {}
    for t in conn[:, :4]:
        for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            key = tuple(sorted(t[list(f)]))
            faces[key] = faces.get(key, 0) + 1
    assert set(faces.values()) <= {1, 2}
    nb = sum(1 for v in faces.values() if v == 1)
    assert nb == 6 * 2 * n * n
    if order == 2:
        # edge nodes sit at edge midpoints
        for k, (a, b) in enumerate([(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]):
            assert np.allclose(coords[conn[:, 4 + k]], 0.5 * (coords[conn[:, a]] + coords[conn[:, b]]), atol=1e-15)


def test_jitter_keeps_faces_planar():
    coords, conn = g.mesh_box(5, 4, 3, order=2, jitter=0.15, seed=9)
    for c in range(3):
        lo = np.isclose(coords[:, c], 0.0)
        hi = np.isclose(coords[:, c], 1.0)
        assert lo.sum() == hi.sum() > 0


def test_refinement_properties():
    coords, conn = g.mesh_box(2, 2, 1, order=1, jitter=0.1, seed=4)
    fc, fconn, parent, ep = g.mesh_refine(coords, conn)
    ne = conn.shape[0]
    assert fconn.shape == (8 * ne, 4)
    assert np.array_equal(parent, np.repeat(np.arange(ne), 8))
    assert np.array_equal(fc[: coords.shape[0]], coords)
    assert np.allclose(fc[coords.shape[0]:], 0.5 * (coords[ep[:, 0]] + coords[ep[:, 1]]), atol=0)

    def vols(c, t):
        X = c[t]
        return np.einsum("ei,ei->e", np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), X[:, 3] - X[:, 0]) / 6

    v0 = vols(coords, conn)
    v1 = vols(fc, fconn)
    assert v1.min() > 0
    assert np.allclose(v1, np.repeat(v0, 8) / 8, rtol=1e-12)
    # every edge of the base mesh gets exactly one new vertex
    edges = set()
    for t in conn:
        for a in range(4):
            for b in range(a + 1, 4):
                edges.add((min(t[a], t[b]), max(t[a], t[b])))
    assert len(edges) == ep.shape[0]
    assert set(map(tuple, np.sort(ep, axis=1))) == edges


def test_device_entry_points_fail_cleanly_without_gpu():
    """On a host without a GPU every device entry point returns an error code
    instead of aborting."""
    h = C.c_void_p()
    st = g.lib.goal_disc_create_box(2, 2, 2, 1.0, 1.0, 1.0, 1, 0.0, 0, C.byref(h))
    if st == g.OK:  # a GPU is present: nothing to check here
        g.lib.goal_disc_destroy(h)
        pytest.skip("GPU present")
    assert st in (g.ERR_NO_DEVICE, g.ERR_HIP)
    assert "device" in g.lib.goal_last_error().decode().lower() or st == g.ERR_HIP


def test_invalid_inputs_are_rejected():
    with pytest.raises(g.GoalError) as ei:
        g.mesh_box(0, 1, 1)
    assert ei.value.code == g.ERR_INVALID
    with pytest.raises(g.GoalError):
        g.mesh_box(2, 2, 2, order=3)
    with pytest.raises(g.GoalError):
        g.mesh_box(2, 2, 2, jitter=0.5)
    bad = np.array([[0, 1, 2, 99]], dtype=np.int32)
    with pytest.raises(g.GoalError) as ei:
        g.mesh_graph(1, 4, bad)
    assert "out of range" in str(ei.value)
