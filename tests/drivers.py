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


def csr(m, vals):
    return sp.csr_matrix((vals, m.cols, m.row_ptr), shape=(m.num_dofs, m.num_dofs))


def newton_steps(order, quad, model, n, steps, tol=1e-10, max_iters=10):
    coords, conn = g.mesh_box(n, n, n, order=order, jitter=0.1, seed=11)
    m = O.Mesh(order, coords, conn)
    mats = [NH_ROW if model == 1 else J2_ROW]
    st = O.States(m, model, quad)
    x = np.zeros(m.num_dofs)
    iters = []
    t_old = 0.0
    J = 0.0
    norm = 0.0
    for step in range(1, steps + 1):
        t = float(step)
        it = 1
        converged = False
        while it <= max_iters and not converged:
            R, V = O.assemble(O.K_JACOBIAN, m, model, mats, x, st, quad_order=quad)
            dM = np.zeros(m.num_dofs)
            for comp, nodes, vals in dbcs(coords, t):
                O.dbc(m, nodes, vals, comp, x, R, dM, V)
            du = spla.spsolve(csr(m, V).tocsc(), -R)
            x = x + du
            R, _ = O.assemble(O.K_RESIDUAL, m, model, mats, x, st, quad_order=quad)
            for comp, nodes, vals in dbcs(coords, t):
                O.dbc(m, nodes, vals, comp, x, R)
            norm = np.linalg.norm(R)
            converged = norm < tol
            it += 1
        if not converged:
            raise RuntimeError("newton failed")
        iters.append(it - 1)
        J, _, _ = O.qoi(0, m, model, mats, x, st, quad_order=quad)
        if step < steps:
            st.update()
        t_old = t
    out = {"iters": iters, "J": J, "norm": norm}
    if order == 1:
        out.update(nested_estimate(coords, conn, x, st, model, mats, quad, t_old))
    return out


def nested_estimate(coords, conn, xb, st_base, model, mats, quad, t):
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
    R, VT = O.assemble(O.K_ADJOINT, m, model, mats, x, st, quad_order=quad, save=False)
    _, _, dM = O.qoi(0, m, model, mats, x, st, quad_order=quad, derivative=True)
    for comp, nodes, vals in dbcs(fc, t):
        O.dbc(m, nodes, vals, comp, x, R, dM, VT)
    z = spla.spsolve(csr(m, VT).tocsc(), dM)
    zR = -(R @ z)
    zc = z.copy()
    O.set_coarse(zc, ep, coords.shape[0])
    zd = z - zc
    e, _ = O.assemble(O.K_LOCALIZE, m, model, mats, x, st, quad_order=quad, save=False, z_diff=zd, z_coarse=zc)
    for comp, nodes, vals in dbcs(fc, t):
        O.dbc(m, nodes, vals, comp, x, e)
    bound = O.sum_contribs(np.arange(m.num_nodes), e)
    eta = O.compute_error(m, e)
    base = O.set_error(parent, eta, conn.shape[0])
    return {"eta_sum": float(base.sum()), "bound": float(bound), "zR": float(zR)}
