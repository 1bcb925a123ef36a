"""Oracle-side drivers (CPU, test infrastructure): Newton load steps as in
Primal::solve (src/goal_primal.cpp:111-137), the functional, and one
nested-adjoint estimate as in NestedAdjoint::run
(src/goal_nested_adjoint.cpp:163-257), all on the oracle's assembly with a
direct sparse solve standing in for Belos/MueLu."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import goal_amd as g
import oracle as O

NH_ROW = [1000.0, 0.25, 0.0, 0.0, 1.0]
J2_ROW = [200.0, 0.3, 2.0, 0.25, 1.0]


def face_nodes(coords, comp, value):
    return np.nonzero(coords[:, comp] == value)[0].astype(np.int32)


def dbcs(coords, t):
    zmin = face_nodes(coords, 2, 0.0)
    zmax = face_nodes(coords, 2, 1.0)
    return [(0, zmin, np.zeros(zmin.size)), (1, zmin, np.zeros(zmin.size)), (2, zmin, np.zeros(zmin.size)),
            (2, zmax, np.full(zmax.size, 0.01 * t))]


# The bar of the J2 tension workload: [0,8]x[0,1]^2, rollers on xmin/ymin/zmin
# and u_x = BAR_RATE t on xmax (J2_uniaxial_3D.yaml's boundary conditions,
# example/primal/J2_uniaxial_3D.yaml:9-13, scaled to 1% strain at t = 20).
BAR_LENGTHS = (8.0, 1.0, 1.0)
BAR_RATE = 0.004
BAR_BCS = [(0, "xmin", "0.0"), (1, "ymin", "0.0"), (2, "zmin", "0.0"), (0, "xmax", "0.004*t")]
CUBE_BCS = [(0, "zmin", "0.0"), (1, "zmin", "0.0"), (2, "zmin", "0.0"), (2, "zmax", "0.01*t")]


def bar_dbcs(coords, t):
    xmin = face_nodes(coords, 0, 0.0)
    ymin = face_nodes(coords, 1, 0.0)
    zmin = face_nodes(coords, 2, 0.0)
    xmax = face_nodes(coords, 0, BAR_LENGTHS[0])
    return [(0, xmin, np.zeros(xmin.size)), (1, ymin, np.zeros(ymin.size)), (2, zmin, np.zeros(zmin.size)),
            (0, xmax, np.full(xmax.size, BAR_RATE * t))]


def problem_mesh(problem, n, order, jitter, seed):
    """(coords, conn, dims, lengths) of the cube (n^3) or bar (n = (nx, ny, nz)) problem."""
    if problem == "bar":
        dims = tuple(n)
        coords, conn = g.mesh_box(*dims, *BAR_LENGTHS, order=order, jitter=jitter, seed=seed)
        return coords, conn, dims, BAR_LENGTHS
    coords, conn = g.mesh_box(n, n, n, order=order, jitter=jitter, seed=seed)
    return coords, conn, (n, n, n), (1.0, 1.0, 1.0)


def csr(m, vals):
    return sp.csr_matrix((vals, m.cols, m.row_ptr), shape=(m.num_dofs, m.num_dofs))


def newton_steps(order, quad, model, n, steps, tol=1e-10, max_iters=10, jitter=0.1, seed=11, dwr=True,
                 threads=1, problem="cube"):
    """Oracle Newton load steps (and, for P1, one nested-adjoint estimate) of
    the cube (n hexahedra a side) or the bar (n = (nx, ny, nz))."""
    coords, conn, _, _ = problem_mesh(problem, n, order, jitter, seed)
    bcs_of = bar_dbcs if problem == "bar" else dbcs
    m = O.Mesh(order, coords, conn)
    mats = [NH_ROW if model == 1 else J2_ROW]
    st = O.States(m, model, quad)
    x = np.zeros(m.num_dofs)
    iters = []
    Js = []
    t_old = 0.0
    J = 0.0
    norm = 0.0
    for step in range(1, steps + 1):
        t = float(step)
        it = 1
        converged = False
        while it <= max_iters and not converged:
            R, V = O.assemble(O.K_JACOBIAN, m, model, mats, x, st, quad_order=quad, threads=threads)
            dM = np.zeros(m.num_dofs)
            for comp, nodes, vals in bcs_of(coords, t):
                O.dbc(m, nodes, vals, comp, x, R, dM, V)
            du = spla.spsolve(csr(m, V).tocsc(), -R)
            x = x + du
            R, _ = O.assemble(O.K_RESIDUAL, m, model, mats, x, st, quad_order=quad, threads=threads)
            for comp, nodes, vals in bcs_of(coords, t):
                O.dbc(m, nodes, vals, comp, x, R)
            norm = np.linalg.norm(R)
            converged = norm < tol
            it += 1
        if not converged:
            raise RuntimeError("newton failed")
        iters.append(it - 1)
        J, _, _ = O.qoi(0, m, model, mats, x, st, quad_order=quad)
        Js.append(J)
        if step < steps:
            st.update()
        t_old = t
    out = {"iters": iters, "J": J, "J_steps": Js, "norm": norm}
    if order == 1 and dwr:
        out.update(nested_estimate(coords, conn, x, st, model, mats, quad, t_old, bcs=bcs_of))
    return out


def nested_estimate(coords, conn, xb, st_base, model, mats, quad, t, bcs=dbcs, tbc=None):
    """tbc: None or (node set of the loaded face as a function of the
    coordinates, T(x, y, z, t)) -- a traction on that face's boundary faces."""
    fc, fconn, parent, ep = g.mesh_refine(coords, conn)
    m = O.Mesh(1, fc, fconn)
    nq = 1 if quad == 1 else 4
    x = np.zeros(m.num_dofs)
    x[: xb.size] = xb
    O.set_coarse(x, ep, coords.shape[0])
    st = O.States(m, model, quad)
    # child takes the parent's current state as its "_old" state
    pts = np.arange(fconn.shape[0] * nq)
    src = (pts // nq // 8) * nq + pts % nq
    st.sigma[:] = st_base.sigma[:, src]
    st.Fp_old[:] = st_base.Fp[:, src]
    st.Fp[:] = st.Fp_old
    st.eqps_old[:] = st_base.eqps[src]
    st.eqps[:] = st.eqps_old
    faces = O.boundary_faces(m, tbc[0](fc)) if tbc else None
    R, VT = O.assemble(O.K_ADJOINT, m, model, mats, x, st, quad_order=quad, save=False)
    if tbc:
        O.traction(m, faces, tbc[1], t, R)
    _, _, dM = O.qoi(0, m, model, mats, x, st, quad_order=quad, derivative=True)
    for comp, nodes, vals in bcs(fc, t):
        O.dbc(m, nodes, vals, comp, x, R, dM, VT)
    z = spla.spsolve(csr(m, VT).tocsc(), dM)
    zR = -(R @ z)
    zc = z.copy()
    O.set_coarse(zc, ep, coords.shape[0])
    zd = z - zc
    e, _ = O.assemble(O.K_LOCALIZE, m, model, mats, x, st, quad_order=quad, save=False, z_diff=zd, z_coarse=zc)
    if tbc:
        O.traction(m, faces, tbc[1], t, e, z=zd)
    for comp, nodes, vals in bcs(fc, t):
        O.dbc(m, nodes, vals, comp, x, e)
    bound = O.sum_contribs(np.arange(m.num_nodes), e)
    eta = O.compute_error(m, e)
    base = O.set_error(parent, eta, conn.shape[0])
    return {"eta_sum": float(base.sum()), "bound": float(bound), "zR": float(zR)}


# The reference's 3D regression inputs on this path (unit cube, rollers on
# xmin / ymin / zmin, 3 load steps of dt = 1, Newton tolerance 1e-8 in at most
# 5 iterations) and their known final avg-disp values.
KNOWN = {
    "nh_uniaxial": dict(model=1, mat=[1000.0, 0.25, 0.0, 0.0, 1.0], traction=False, value=2.541285341193943e-03,
                        src="example/primal/neohookean_uniaxial_3D.yaml:1-36"),
    "j2_uniaxial": dict(model=2, mat=[1000.0, 0.25, 100.0, 10.0, 1.0], traction=False, value=1.073955612775838e-03,
                        src="example/primal/J2_uniaxial_3D.yaml:1-38"),
    "j2_traction": dict(model=2, mat=[1000.0, 0.25, 100.0, 10.0, 1.0], traction=True, value=2.512233163668167e-04,
                        src="example/primal/J2_traction_3D.yaml:1-39"),
}
KNOWN_STEPS, KNOWN_TOL, KNOWN_MAX_ITERS = 3, 1e-8, 5


def known_traction(x, y, z, t):
    return (0.0, 1.0 * t, 0.0)


def known_bcs(case, sets=None):
    """The regression inputs' Dirichlet entries; sets: node sets by name (the
    mesh file's), else the box faces found from the coordinates."""
    def bcs(coords, t):
        if sets is None:
            xmin, ymin, zmin = face_nodes(coords, 0, 0.0), face_nodes(coords, 1, 0.0), face_nodes(coords, 2, 0.0)
            xmax = face_nodes(coords, 0, 1.0)
        else:
            xmin, ymin, zmin, xmax = sets["xmin"], sets["ymin"], sets["zmin"], sets["xmax"]
        out = [(0, xmin, np.zeros(xmin.size)), (1, ymin, np.zeros(ymin.size)), (2, zmin, np.zeros(zmin.size))]
        if not KNOWN[case]["traction"]:
            out.append((0, xmax, np.full(xmax.size, 0.01 * t)))
        return out
    return bcs


def known_newton(case, order=1, quad=1, n=2, jitter=0.1, seed=11, dwr=False, mesh=None):
    """Oracle Newton on one of the reference's regression inputs.  The mesh is
    a jittered n^3 box (the solutions are homogeneous, so any mesh of the unit
    cube reproduces the reference's value), or mesh = (coords, conn, node sets
    by name, ymax faces as node lists) from a mesh file."""
    k = KNOWN[case]
    if mesh is None:
        coords, conn = g.mesh_box(n, n, n, order=order, jitter=jitter, seed=seed)
        sets = None
    else:
        coords, conn, sets, ymax_faces = mesh
    m = O.Mesh(order, coords, conn)
    model, mats = k["model"], [k["mat"]]
    bcs = known_bcs(case, sets)
    faces = None
    if k["traction"]:
        faces = O.boundary_faces(m, face_nodes(coords, 1, 1.0)) if mesh is None else ymax_faces
    st = O.States(m, model, quad)
    x = np.zeros(m.num_dofs)
    iters, J, norm = [], 0.0, 0.0
    for step in range(1, KNOWN_STEPS + 1):
        t = float(step)
        for it in range(1, KNOWN_MAX_ITERS + 1):
            R, V = O.assemble(O.K_JACOBIAN, m, model, mats, x, st, quad_order=quad)
            if faces is not None:
                O.traction(m, faces, known_traction, t, R)
            for comp, nodes, vals in bcs(coords, t):
                O.dbc(m, nodes, vals, comp, x, R, np.zeros(m.num_dofs), V)
            x = x + spla.spsolve(csr(m, V).tocsc(), -R)
            R, _ = O.assemble(O.K_RESIDUAL, m, model, mats, x, st, quad_order=quad)
            if faces is not None:
                O.traction(m, faces, known_traction, t, R)
            for comp, nodes, vals in bcs(coords, t):
                O.dbc(m, nodes, vals, comp, x, R)
            norm = np.linalg.norm(R)
            if norm < KNOWN_TOL:
                break
        else:
            raise RuntimeError("newton failed")
        iters.append(it)
        J, _, _ = O.qoi(0, m, model, mats, x, st, quad_order=quad)
        if step < KNOWN_STEPS:
            st.update()
    out = {"iters": iters, "J": J, "norm": norm}
    if dwr and order == 1:
        # on the nested mesh the sets are the unit cube's faces, found from the
        # coordinates (a mesh file's sets are exactly those faces)
        tbc = (lambda c: face_nodes(c, 1, 1.0), known_traction) if k["traction"] else None
        out.update(nested_estimate(coords, conn, x, st, model, mats, quad, float(KNOWN_STEPS), bcs=known_bcs(case),
                                   tbc=tbc))
    return out


def gpu_known_newton(case, order=1, quad=1, n=2, jitter=0.1, seed=11, lin_rtol=1e-13, disc=None):
    """The same on the GPU through the C ABI: device assembly, Dirichlet rows,
    traction loads and norms, the library's device GMRES.  disc: a Disc with
    node sets xmin, ymin, zmin, xmax and side set ymax (e.g. read from the
    reference's mesh files), else a jittered n^3 box."""
    k = KNOWN[case]
    if disc is None:
        disc = g.Disc.box(n, n, n, order=order, jitter=jitter, seed=seed)
    model = k["model"]
    st = g.States(disc, model, quad)
    mech = g.Mechanics(disc, st, [g.MaterialSpec(model=model, **dict(zip(("E", "nu", "K", "Y", "c0"), k["mat"])))],
                       stabilization=True, quad_order=quad)
    si = g.SolInfo(disc)
    bcs = [(0, "xmin", "0.0"), (1, "ymin", "0.0"), (2, "zmin", "0.0")]
    if not k["traction"]:
        bcs.append((0, "xmax", "0.01*t"))
    b, du = g.DeviceArray(disc.num_dofs), g.DeviceArray(disc.num_dofs)
    iters, norm = [], 0.0
    for step in range(1, KNOWN_STEPS + 1):
        t = float(step)
        for it in range(1, KNOWN_MAX_ITERS + 1):
            si.zero_all()
            mech.jacobian(si, adjoint=False, save_states=True)
            if k["traction"]:
                g.tbc_apply(disc, si, "ymax", ("0.0", "1.0*t", "0.0"), t)
            for comp, ns, expr in bcs:
                g.dbc_apply(disc, si, comp, ns, expr, t, True)
            g.check(g.lib.goal_memcpy_h2d(b.ptr, g._ptr(np.ascontiguousarray(-si.R())), 8 * disc.num_dofs))
            du.fill(0.0)
            si.solve(b, du, rtol=lin_rtol, max_iters=5000, restart=100)
            disc.add_soln_dev(du)
            si.zero_R()
            mech.residual(si, save_states=True)
            if k["traction"]:
                g.tbc_apply(disc, si, "ymax", ("0.0", "1.0*t", "0.0"), t)
            for comp, ns, expr in bcs:
                g.dbc_apply(disc, si, comp, ns, expr, t, False)
            norm = si.norm_R()
            if norm < KNOWN_TOL:
                break
        else:
            raise RuntimeError("newton failed")
        iters.append(it)
        if step < KNOWN_STEPS:
            st.update()
    return {"iters": iters, "J": mech.qoi(0), "norm": norm}


def gpu_newton_steps(order, quad, model, n, steps, tol=1e-10, max_iters=10, jitter=0.1, seed=11, tangent=64,
                     solver="device", lin_rtol=1e-12, problem="cube"):
    """The same load steps with the assembly, Dirichlet rows and norms on the
    GPU (through the C ABI); the linear solve is the library's device GMRES
    (solver="device") or scipy's direct solve on the host (solver="scipy")."""
    if problem == "bar":
        disc = g.Disc.box(*n, *BAR_LENGTHS, order=order, jitter=jitter, seed=seed)
    else:
        disc = g.Disc.box(n, n, n, order=order, jitter=jitter, seed=seed)
    mat = dict(zip(("E", "nu", "K", "Y", "c0"), NH_ROW if model == 1 else J2_ROW))
    st = g.States(disc, model, quad)
    mech = g.Mechanics(disc, st, [g.MaterialSpec(model=model, **mat)], stabilization=True, quad_order=quad)
    mech.set_option("tangent", tangent)
    rp, cols = disc.graph()
    si = g.SolInfo(disc)
    bcs = BAR_BCS if problem == "bar" else CUBE_BCS
    iters, Js, lin = [], [], []
    norm = 0.0
    b, du = g.DeviceArray(disc.num_dofs), g.DeviceArray(disc.num_dofs)
    for step in range(1, steps + 1):
        t = float(step)
        it, converged = 1, False
        while it <= max_iters and not converged:
            si.zero_all()
            mech.jacobian(si, adjoint=False, save_states=True)
            for comp, ns, expr in bcs:
                g.dbc_apply(disc, si, comp, ns, expr, t, True)
            if solver == "device":
                g.check(g.lib.goal_memcpy_h2d(b.ptr, g._ptr(np.ascontiguousarray(-si.R())), 8 * disc.num_dofs))
                du.fill(0.0)
                k, rr = si.solve(b, du, rtol=lin_rtol, max_iters=5000, restart=100)
                lin.append(k)
                disc.add_soln_dev(du)
            else:
                A = sp.csr_matrix((si.vals(), cols, rp), shape=(disc.num_dofs, disc.num_dofs))
                disc.add_soln(spla.spsolve(A.tocsc(), -si.R()))
            si.zero_R()
            mech.residual(si, save_states=True)
            for comp, ns, expr in bcs:
                g.dbc_apply(disc, si, comp, ns, expr, t, False)
            norm = si.norm_R()
            converged = norm < tol
            it += 1
        if not converged:
            raise RuntimeError("newton failed")
        iters.append(it - 1)
        Js.append(mech.qoi(0))
        if step < steps:
            st.update()
    return {"iters": iters, "J": Js[-1], "J_steps": Js, "norm": norm, "linear_iters": lin}
