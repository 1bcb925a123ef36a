"""The reference's own known answers on this path, reproduced by the oracle.

goal's regression inputs example/primal/{neohookean_uniaxial,J2_uniaxial,
J2_traction}_3D.yaml check the final avg-disp QoI against a stored value
(src/goal_regression.cpp:14-23, absolute tolerance 1e-8).  All three load a
unit cube on rollers (xmin u_x, ymin u_y, zmin u_z = 0) either by u_x = 0.01 t
on xmax or by a traction (0, t, 0) on ymax.  The exact solutions are homogeneous
(uniaxial stress), and a homogeneous deformation satisfies the discrete
equations of any tet mesh exactly:
  * interior and free-face nodes see P : sum_e grad N_a |e| = 0;
  * the pressure is constant, so the stabilization term vanishes;
  * avg disp integrates a linear field, which the one-point rule does exactly.
So the reference's values hold on goal_amd's own jittered box meshes, not only
on its SCOREC .smb files.  The Newton iterates are homogeneous too.  Stopping
at ||R|| < 1e-8 after the same iteration as the reference therefore gives the
same value to rounding, and the values match to ~1e-17, far inside the
reference's 1e-8.

The neo-Hookean value is also pinned by a closed form.  It solves
sigma_22 = 0 for the lateral stretch of the mixed, stabilized model
(src/goal_neohookean.cpp:59-72, src/goal_mixed.cpp:33-46,
src/goal_presidual.cpp:53-77).
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy.optimize import brentq

from tests import drivers

ABS_TOL = 1e-15   # the reference's own check uses 1e-8


def test_nh_uniaxial_value_is_the_uniaxial_stress_closed_form():
    E, nu = 1000.0, 0.25
    mu, kappa = E / (2 * (1 + nu)), E / (3 * (1 - 2 * nu))
    l1 = 1.0 + 0.01 * 3   # u_x = 0.01 t on xmax at t = 3

    def sigma22(l2):
        J = l1 * l2 * l2
        p = 0.5 * kappa * (J - 1.0 / J)   # pressure residual p / kappa - (J - 1/J) / 2 = 0
        return mu * J ** (-5.0 / 3.0) * (l2 * l2 - (l1 * l1 + 2 * l2 * l2) / 3.0) + p

    l2 = brentq(sigma22, 0.9, 1.0, xtol=1e-16, rtol=4 * np.finfo(float).eps)
    avg = ((l1 - 1.0) + 2.0 * (l2 - 1.0)) / 6.0   # mean of (u_x + u_y + u_z) / 3 over the unit cube
    assert abs(avg - drivers.KNOWN["nh_uniaxial"]["value"]) < 1e-16


@pytest.mark.parametrize("case", sorted(drivers.KNOWN))
@pytest.mark.parametrize("n,seed", [(2, 11), (3, 5)])
def test_oracle_reproduces_reference_known_answer(case, n, seed):
    out = drivers.known_newton(case, order=1, quad=1, n=n, seed=seed)
    ref = drivers.KNOWN[case]["value"]
    print(case, n, out["iters"], out["J"], out["J"] - ref)
    assert abs(out["J"] - ref) < ABS_TOL, (out["J"], ref)


@pytest.mark.parametrize("case", ["nh_uniaxial", "j2_uniaxial"])
def test_oracle_p2_four_points_reproduces_reference_known_answer(case):
    """The homogeneous solution is exact for P2 with four points as well."""
    out = drivers.known_newton(case, order=2, quad=2, n=2)
    assert abs(out["J"] - drivers.KNOWN[case]["value"]) < ABS_TOL


def test_traction_total_force_and_weighting():
    """The traction loads on a face sum to T |face| (P1: 1/3 per node and
    face; P2: -1/9 vertices, 4/9 edge nodes).  With a weight vector z, every
    share scales by z_i at the face centroid."""
    import goal_amd as g
    import oracle as O
    for order in (1, 2):
        coords, conn = g.mesh_box(2, 3, 2, ly=2.0, order=order, jitter=0.1, seed=3)
        m = O.Mesh(order, coords, conn)
        faces = O.boundary_faces(m, drivers.face_nodes(coords, 1, 2.0))
        assert faces.shape == (2 * 2 * 2, 3 if order == 1 else 6)   # 2 x 2 quads, 2 triangles each
        R = np.zeros(m.num_dofs)
        O.traction(m, faces, lambda x, y, z, t: (0.5 * t, -1.0, 2.0), 2.0, R)
        f = R.reshape(-1, 4)
        assert np.allclose(f[:, :3].sum(axis=0), [-1.0, 1.0, -2.0], rtol=1e-14)   # T = (1, -1, 2), area 1
        assert np.all(f[:, 3] == 0.0)
        # a constant weight z_i = c_i scales component i by c_i (the face
        # shape functions sum to one at the centroid)
        z = np.tile([2.0, -1.0, 0.5, 7.0], m.num_nodes)
        Rz = np.zeros(m.num_dofs)
        O.traction(m, faces, lambda x, y, z_, t: (0.5 * t, -1.0, 2.0), 2.0, Rz, z=z)
        assert np.allclose(Rz.reshape(-1, 4)[:, :3], f[:, :3] * [2.0, -1.0, 0.5], rtol=1e-13, atol=1e-16)


def test_boundary_faces_of_box_face():
    """The side set of a box face: 2 triangles per boundary quad, each face
    once, all nodes on the face, and no face of the interior plane y = 1."""
    import goal_amd as g
    import oracle as O
    coords, conn = g.mesh_box(3, 2, 2, ly=2.0, order=2, jitter=0.1, seed=4)
    m = O.Mesh(2, coords, conn)
    for y, expect in ((0.0, 2 * 3 * 2), (2.0, 2 * 3 * 2)):
        f = O.boundary_faces(m, drivers.face_nodes(coords, 1, y))
        assert f.shape == (expect, 6)
        assert np.all(coords[f.ravel(), 1] == y)
        assert len({tuple(sorted(r[:3])) for r in f.tolist()}) == expect
    # the unjittered mesh has faces on the interior plane y = 1, each shared by
    # two tets: none of them is a boundary face
    coords, conn = g.mesh_box(3, 2, 2, ly=2.0, order=2, jitter=0.0, seed=4)
    m = O.Mesh(2, coords, conn)
    mid = np.nonzero(coords[:, 1] == 1.0)[0]
    conn4 = m.conn.reshape(m.num_elems, -1)[:, :4]
    on_plane = sum(int(np.all(coords[conn4[e, list(fv)], 1] == 1.0))
                   for e in range(m.num_elems) for fv in O._TET_FACES)
    assert on_plane == 2 * (2 * 3 * 2)   # 12 faces on the plane, each seen from both sides
    assert O.boundary_faces(m, mid).shape[0] == 0


# --------------------------------------------------- the reference's own mesh
GOLDEN_SMB = __import__("pathlib").Path(__file__).resolve().parent / "golden" / "smb"


def _cube_file(order=1, mesh="cube-serial0.smb"):
    """The reference's cube: "cube-serial0.smb" (one part) or "cube.smb" (its
    four parts cube0.smb ... cube3.smb, as the example inputs name it)."""
    import goal_amd as g
    return g.MeshFile(GOLDEN_SMB / "cube.dmg", GOLDEN_SMB / mesh, GOLDEN_SMB / "cube.txt", order=order)


def faces_as_nodes(conn, order, elems, faces):
    """(element, local face) pairs as face node lists (3 vertices, then for
    P2 the nodes of edges v0v1, v1v2, v2v0), in apf's tet face order."""
    import oracle as O
    out = []
    for e, f in zip(elems, faces):
        lv = O._TET_FACES[f]
        row = [int(conn[e, v]) for v in lv]
        if order == 2:
            for j in range(3):
                pair = {lv[j], lv[(j + 1) % 3]}
                row.append(int(conn[e, 4 + next(i for i, ed in enumerate(O._TET_EDGES) if set(ed) == pair)]))
        out.append(row)
    return np.asarray(out, dtype=np.int32)


@pytest.mark.parametrize("order", [1, 2])
def test_mesh_file_reader_on_reference_cube(order):
    """The reference's test mesh (test/mesh/cube: cube-serial0.smb, cube.dmg,
    cube.txt, committed as fixtures) read by the library on the host: 51
    vertices, 230 edges, 132 positively oriented tets filling the unit cube;
    the node sets (closure of their model faces) are exactly the nodes on
    those faces; the side sets are the face triangles, area 1 each; one
    element set 'cube'."""
    f = _cube_file(order)
    coords, conn, es = f.mesh()
    assert (f.num_nodes, f.num_elems) == ((51, 132) if order == 1 else (51 + 230, 132))
    assert f.elem_set_names == ["cube"] and np.all(es == 0)
    X = coords[conn[:, :4]]
    vol = np.einsum("ij,ij->i", X[:, 1] - X[:, 0], np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0])) / 6.0
    assert vol.min() > 0 and abs(vol.sum() - 1.0) < 1e-14
    if order == 2:   # edge nodes at the midpoints of their tet edges
        import oracle as O
        for ed, (a, b) in enumerate(O._TET_EDGES):
            assert np.allclose(coords[conn[:, 4 + ed]], 0.5 * (coords[conn[:, a]] + coords[conn[:, b]]), atol=1e-15)
    for name, comp, val in (("xmin", 0, 0.0), ("xmax", 0, 1.0), ("ymin", 1, 0.0), ("zmin", 2, 0.0),
                            ("zmax", 2, 1.0)):
        assert np.array_equal(np.sort(f.node_set(name)), np.nonzero(coords[:, comp] == val)[0]), name
    for name, comp, val in (("ymax", 1, 1.0), ("xmax", 0, 1.0)):
        elems, faces = f.side_set(name)
        fn = faces_as_nodes(conn, order, elems, faces)
        assert np.all(coords[fn.ravel(), comp] == val)
        a = coords[fn[:, 1]] - coords[fn[:, 0]]
        b = coords[fn[:, 2]] - coords[fn[:, 0]]
        assert abs(0.5 * np.linalg.norm(np.cross(a, b), axis=1).sum() - 1.0) < 1e-14, name


def test_mesh_file_reader_rejects_bad_input(tmp_path):
    import goal_amd as g
    raw = (GOLDEN_SMB / "cube-serial0.smb").read_bytes()
    bad_version = raw[:4] + (6).to_bytes(4, "big") + raw[8:]
    (tmp_path / "v6.smb").write_bytes(bad_version)
    (tmp_path / "short.smb").write_bytes(raw[:5000])
    # a tet count of 4e9 in the header: rejected before anything is sized by it
    (tmp_path / "huge.smb").write_bytes(raw[:44] + (0xFFFFFFF0).to_bytes(4, "big") + raw[48:])
    for name in ("v6.smb", "short.smb", "huge.smb"):
        with pytest.raises(g.GoalError):
            g.MeshFile(GOLDEN_SMB / "cube.dmg", tmp_path / name, GOLDEN_SMB / "cube.txt")
    with pytest.raises(g.GoalError, match="ends early"):
        g.MeshFile(GOLDEN_SMB / "cube.dmg", tmp_path / "huge.smb", GOLDEN_SMB / "cube.txt")
    (tmp_path / "bad.txt").write_text("node set xmin 1\n2 99\n")
    with pytest.raises(g.GoalError, match="no model entity"):
        g.MeshFile(GOLDEN_SMB / "cube.dmg", GOLDEN_SMB / "cube-serial0.smb", tmp_path / "bad.txt")


def test_partitioned_mesh_merges_to_the_serial_one():
    """cube.smb as four parts (cube0..3.smb, joined through their shared
    vertices) is the one-part cube-serial0.smb: the same vertices, tets,
    node sets and side-set sizes; P2 edge nodes are not duplicated."""
    def key(c):
        return sorted(map(tuple, np.round(c, 12).tolist()))

    s, p = _cube_file(1), _cube_file(1, "cube.smb")
    cs, ks, _ = s.mesh()
    cp, kp, _ = p.mesh()
    assert (p.num_nodes, p.num_elems) == (51, 132) and key(cs) == key(cp)
    assert sorted(tuple(key(cs[t])) for t in ks) == sorted(tuple(key(cp[t])) for t in kp)
    for name, comp, val in (("xmin", 0, 0.0), ("xmax", 0, 1.0), ("ymin", 1, 0.0), ("zmin", 2, 0.0)):
        assert np.array_equal(np.sort(p.node_set(name)), np.nonzero(cp[:, comp] == val)[0])
    assert len(p.side_set("ymax")[0]) == len(s.side_set("ymax")[0]) == 16
    assert _cube_file(2, "cube.smb").num_nodes == 51 + 230


@pytest.mark.parametrize("mesh", ["cube-serial0.smb", "cube.smb"])
@pytest.mark.parametrize("case", sorted(drivers.KNOWN))
def test_oracle_reproduces_known_answer_on_reference_mesh(case, mesh):
    """The regression inputs as the reference runs them: its cube mesh (one
    part, or the four parts the example inputs name), model and association
    file (read by the library), its node sets and side set, P1 with one
    point.  The oracle's Newton gives the stored values."""
    f = _cube_file(1, mesh)
    coords, conn, _ = f.mesh()
    sets = {k: f.node_set(k) for k in ("xmin", "ymin", "zmin", "xmax")}
    ymax = faces_as_nodes(conn, 1, *f.side_set("ymax"))
    out = drivers.known_newton(case, mesh=(coords, conn, sets, ymax))
    print(case, out["iters"], out["J"] - drivers.KNOWN[case]["value"])
    assert abs(out["J"] - drivers.KNOWN[case]["value"]) < ABS_TOL
