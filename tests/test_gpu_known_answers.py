"""The reference's known answers and its traction loads on the GPU, through
the C ABI and the C++ surface.

* The three 3D regression inputs (example/primal/{neohookean_uniaxial,
  J2_uniaxial,J2_traction}_3D.yaml) run with device assembly, Dirichlet rows,
  traction loads and the device GMRES.  The final avg disp must equal the
  reference's stored value to 1e-14 absolute, where goal's own check allows
  1e-8 (the homogeneous solutions make the value mesh-independent; see
  tests/test_known_answers.py).  The Newton iteration counts must equal the
  oracle's.
* Side sets derived from box-face node sets must be the oracle's boundary
  faces.  Traction loads (plain and z-weighted, P1 and P2) must match the
  oracle's to 1e-14 relative.  Only the summation order differs.
* The C++ surface (goal::Primal with TractionBC, goal::NestedAdjoint with
  the z_diff-weighted traction in its localized residual) must match the
  oracle's Newton and nested-DWR drivers.
"""
from __future__ import annotations

import numpy as np
import pytest

import goal_amd as g
import oracle as O
from tests import drivers
from tests.test_gpu_drivers import run_surface

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case,order,quad,n", [(c, 1, 1, 3) for c in sorted(drivers.KNOWN)] + [
    ("nh_uniaxial", 2, 2, 2), ("j2_uniaxial", 2, 2, 2)])
def test_gpu_reproduces_reference_known_answer(case, order, quad, n):
    gpu = drivers.gpu_known_newton(case, order=order, quad=quad, n=n)
    ref = drivers.KNOWN[case]["value"]
    print(case, order, gpu["iters"], repr(gpu["J"]), gpu["J"] - ref)
    assert abs(gpu["J"] - ref) < 1e-14, (gpu["J"], ref)
    assert gpu["iters"] == drivers.known_newton(case, order=order, quad=quad, n=n)["iters"]


def test_gpu_p2_traction_matches_oracle():
    """On P2 faces the reference's one-point rule loads the vertices with
    -1/9 and the edge nodes with 4/9 of the face's force.  That is not the
    consistent load of a constant traction, so the P2 solution is not the
    homogeneous one and its value is not the P1 regression value.  The GPU
    path must still be the oracle's (to 1e-12 relative)."""
    gpu = drivers.gpu_known_newton("j2_traction", order=2, quad=2, n=2)
    ref = drivers.known_newton("j2_traction", order=2, quad=2, n=2)
    print(gpu, ref)
    assert gpu["iters"] == ref["iters"]
    assert abs(gpu["J"] - ref["J"]) <= 1e-12 * abs(ref["J"])
    assert abs(ref["J"] - drivers.KNOWN["j2_traction"]["value"]) > 1e-7   # the face rule shows


def _faces_as_nodes(disc, elems, faces):
    conn = disc.mesh()[1]
    out = []
    for e, f in zip(elems, faces):
        lv = O._TET_FACES[f]
        row = [int(conn[e, v]) for v in lv]
        if disc.order == 2:
            for j in range(3):
                pair = {lv[j], lv[(j + 1) % 3]}
                row.append(int(conn[e, 4 + next(i for i, ed in enumerate(O._TET_EDGES) if set(ed) == pair)]))
        out.append(row)
    return np.asarray(out, dtype=np.int32)


@pytest.mark.parametrize("order", [1, 2])
def test_side_sets_and_traction_loads_match_oracle(order):
    disc = g.Disc.box(3, 2, 2, ly=2.0, order=order, jitter=0.1, seed=7)
    coords, conn, _ = disc.mesh()
    m = O.Mesh(order, coords, conn)
    si = g.SolInfo(disc)
    rng = np.random.default_rng(order)
    z = rng.standard_normal(disc.num_dofs)
    zd = g.DeviceArray.from_host(z)
    T = ("0.5*t + x", "-1.0 + y*z", "2.0*t")

    def Tf(x, y, z_, t):
        return (0.5 * t + x, -1.0 + y * z_, 2.0 * t)

    for name, comp, val in (("ymax", 1, 2.0), ("xmin", 0, 0.0), ("zmax", 2, 1.0)):
        elems, faces = disc.side_set(name)
        ref_faces = O.boundary_faces(m, drivers.face_nodes(coords, comp, val))
        assert np.array_equal(_faces_as_nodes(disc, elems, faces), ref_faces), name
        for weight in (None, zd):
            si.zero_all()
            si.set(R=np.zeros(disc.num_dofs))
            g.tbc_apply(disc, si, name, T, 1.5, weight=weight)
            ref = np.zeros(disc.num_dofs)
            O.traction(m, ref_faces, Tf, 1.5, ref, z=None if weight is None else z)
            got = si.R()
            assert np.linalg.norm(got - ref) <= 1e-14 * np.linalg.norm(ref), (name, weight is not None)
    # an explicit side set: the loads of the faces given
    elems, faces = disc.side_set("ymax")
    disc.add_side_set("half", elems[::2], faces[::2])
    si.set(R=np.zeros(disc.num_dofs))
    g.tbc_apply(disc, si, "half", T, 1.5)
    ref = np.zeros(disc.num_dofs)
    O.traction(m, _faces_as_nodes(disc, elems[::2], faces[::2]), Tf, 1.5, ref)
    assert np.linalg.norm(si.R() - ref) <= 1e-14 * np.linalg.norm(ref)
    with pytest.raises(g.GoalError):
        g.tbc_apply(disc, si, "no such set", T, 1.0)


@pytest.mark.parametrize("case", sorted(drivers.KNOWN))
def test_cpp_surface_known_answer_and_dwr_with_traction(case):
    """goal::Primal / goal::NestedAdjoint (C++ surface) on the regression
    inputs: the known value, and the nested-DWR quantities against the
    oracle's drivers.

    The exact solutions are discrete solutions on the nested mesh too, so
    most estimates vanish up to the Newton tolerance.  Two do not.  Under the
    traction, eta and the bound come from the z_diff-weighted traction share,
    whose one-point face rule does not integrate the quadratic weight
    exactly.  For J2 uniaxial, z.R comes from the children restarting the
    return map from the parent's converged state.  These are compared at 1e-8
    relative.  The vanishing ones (below 1e-12 in the oracle) must stay below
    1e-11."""
    model = drivers.KNOWN[case]["model"]
    gpu = run_surface(1, 1, model, 3, drivers.KNOWN_STEPS, extra=f"known:{case}")
    ref = drivers.known_newton(case, order=1, quad=1, n=3, dwr=True)
    print(case, gpu, ref)
    assert abs(gpu["J"] - drivers.KNOWN[case]["value"]) < 1e-14
    assert gpu["iters"] == ref["iters"]
    for k in ("eta_sum", "bound", "zR"):
        if abs(ref[k]) > 1e-12:
            assert abs(gpu[k] - ref[k]) <= 1e-8 * abs(ref[k]), k
        else:
            assert abs(gpu[k]) < 1e-11, k


@pytest.mark.parametrize("mesh", ["cube-serial0.smb", "cube.smb"])
@pytest.mark.parametrize("case", sorted(drivers.KNOWN))
def test_gpu_known_answer_on_reference_mesh_files(case, mesh):
    """The regression inputs end to end as goal runs them: the reference's
    cube mesh, model and association file read by the library, a Disc with
    the file's node sets and side set, device Newton.  The final avg disp is
    the stored value."""
    from tests.test_known_answers import _cube_file
    disc = _cube_file(1, mesh).disc()
    gpu = drivers.gpu_known_newton(case, disc=disc)
    print(case, gpu["iters"], gpu["J"] - drivers.KNOWN[case]["value"])
    assert abs(gpu["J"] - drivers.KNOWN[case]["value"]) < 1e-14


@pytest.mark.parametrize("case", sorted(drivers.KNOWN))
def test_cpp_surface_known_answer_on_reference_mesh_files(case):
    """goal::Disc::load (C++ surface) on the reference's four-part cube, then
    goal::Primal with its Dirichlet and traction entries: the stored value.
    goal::NestedAdjoint then refines the file's side set with the mesh, and
    its estimates match the oracle's on the same mesh (as in the box test
    above)."""
    from tests.test_known_answers import GOLDEN_SMB, _cube_file, faces_as_nodes
    gpu = run_surface(1, 1, drivers.KNOWN[case]["model"], 3, drivers.KNOWN_STEPS, extra=f"known:{case}:{GOLDEN_SMB}")
    f = _cube_file(1, "cube.smb")
    coords, conn, _ = f.mesh()
    sets = {k: f.node_set(k) for k in ("xmin", "ymin", "zmin", "xmax")}
    ref = drivers.known_newton(case, mesh=(coords, conn, sets, faces_as_nodes(conn, 1, *f.side_set("ymax"))),
                               dwr=True)
    print(case, gpu, ref)
    assert abs(gpu["J"] - drivers.KNOWN[case]["value"]) < 1e-14
    assert gpu["iters"] == ref["iters"]
    for k in ("eta_sum", "bound", "zR"):
        if abs(ref[k]) > 1e-12:
            assert abs(gpu[k] - ref[k]) <= 1e-8 * abs(ref[k]), k
        else:
            assert abs(gpu[k]) < 1e-11, k
