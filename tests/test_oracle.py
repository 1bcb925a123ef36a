"""Pins for the oracle (CPU).  Each test is chosen so that a plausible mistake
in the oracle's restatement (a dropped term, a wrong sign or index, a
transposed operand) fails it:

  * Jacobian = central finite differences of the residual (all models,
    P1/P2, 1/4 quadrature points, stabilization on) -- pins the dual numbers
    and the consistency of the chain;
  * momentum residual = gradient of the neo-Hookean energy
    mu/2 (J^{-2/3} tr C - 3) + p (J - 1) (independent numpy energy);
  * pressure residual = p-gradient of p^2/(2 kappa) - 1/2 (J - 1/J) p +
    1/2 tau J g.C^{-1} g (independent numpy potential);
  * J2 return map against the closed form dgam = f/(2 mubar + 2K/3) and the
    consistency condition |dev sigma| J = sqrt(2/3) (Y + K eqps);
  * J2 elastic branch = neo-Hookean; rigid rotation -> zero stress;
  * sum over nodes of the momentum residual = 0 (partition of unity of grad N);
  * patch test: affine u with matching constant p -> interior residual 0;
  * transposed scatter = transpose of the primal Jacobian;
  * QoIs: closed forms (constant displacement, uniaxial stretch) and FD of dMdu;
  * DWR: partition-of-unity identity of the localized residual, set_error
    conservation, compute_error / sum_contribs by brute force;
  * set_coarse = the base mesh's P1 interpolant at every nested vertex
    (barycentric evaluation in the parent tet), affine fields unchanged.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from tests import npref
from tests.helpers import J2, NH, ELASTIC, mat_row, random_history, random_state

g = pytest.importorskip("goal_amd")


def box(n=2, order=1, jitter=0.1, seed=7):
    coords, conn = g.mesh_box(n, n, n, order=order, jitter=jitter, seed=seed)
    return O.Mesh(order, coords, conn)


def dense(m, vals):
    return sp.csr_matrix((vals, m.cols, m.row_ptr), shape=(m.num_dofs, m.num_dofs)).toarray()


MODELS = {0: ELASTIC, 1: NH, 2: J2}


@pytest.mark.parametrize("order,quad", [(1, 1), (1, 2), (2, 2)])
@pytest.mark.parametrize("model", [0, 1, 2])
def test_jacobian_matches_finite_differences(order, quad, model):
    m = box(2, order)
    mats = [mat_row(MODELS[model])]
    x = random_state(m.num_nodes, 3)
    st = O.States(m, model, quad)
    if model == 2:
        Fp, eq = random_history(m.num_elems * st.eqps.size // m.num_elems, 4)
        st.Fp_old[:] = Fp
        st.eqps_old[:] = eq
    R, V = O.assemble(O.K_JACOBIAN, m, model, mats, x, st, quad_order=quad, save=False)
    R0, _ = O.assemble(O.K_RESIDUAL, m, model, mats, x, st, quad_order=quad, save=False)
    assert np.array_equal(R, R0)
    A = dense(m, V)
    rng = np.random.default_rng(0)
    for j in rng.choice(m.num_dofs, 10, replace=False):
        h = 1e-6
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        Rp, _ = O.assemble(O.K_RESIDUAL, m, model, mats, xp, st, quad_order=quad, save=False)
        Rm, _ = O.assemble(O.K_RESIDUAL, m, model, mats, xm, st, quad_order=quad, save=False)
        fd = (Rp - Rm) / (2 * h)
        assert np.linalg.norm(fd - A[:, j]) <= 1e-6 * max(np.linalg.norm(A[:, j]), 1.0)


@pytest.mark.parametrize("order,quad", [(1, 1), (2, 2)])
def test_adjoint_is_transpose(order, quad):
    m = box(2, order)
    for model in (1, 2):
        mats = [mat_row(MODELS[model])]
        x = random_state(m.num_nodes, 5)
        st = O.States(m, model, quad)
        _, V = O.assemble(O.K_JACOBIAN, m, model, mats, x, st, quad_order=quad, save=False)
        _, VT = O.assemble(O.K_ADJOINT, m, model, mats, x, st, quad_order=quad, save=False)
        assert np.array_equal(dense(m, VT), dense(m, V).T)


@pytest.mark.parametrize("order,quad", [(1, 1), (2, 2)])
def test_momentum_residual_is_energy_gradient(order, quad):
    m = box(1, order, jitter=0.1)
    mu = NH["E"] / (2 * (1 + NH["nu"]))
    x = random_state(m.num_nodes, 8, scale_u=0.03, scale_p=5.0)
    st = O.States(m, 1, quad)
    R, _ = O.assemble(O.K_RESIDUAL, m, 1, [mat_row(NH)], x, st, quad_order=quad, save=False)
    rng = np.random.default_rng(1)
    for node in rng.choice(m.num_nodes, 6, replace=False):
        for i in range(3):
            j = 4 * node + i
            h = 1e-6
            xp, xm = x.copy(), x.copy()
            xp[j] += h
            xm[j] -= h
            d = (npref.nh_energy(order, m.coords, m.conn, xp, mu, 1 if quad == 1 else 4)
                 - npref.nh_energy(order, m.coords, m.conn, xm, mu, 1 if quad == 1 else 4)) / (2 * h)
            assert abs(d - R[j]) <= 1e-7 * max(1.0, abs(R[j]))


@pytest.mark.parametrize("order,quad", [(1, 1), (2, 2)])
def test_pressure_residual_is_potential_gradient(order, quad):
    m = box(1, order, jitter=0.1)
    mu = NH["E"] / (2 * (1 + NH["nu"]))
    kappa = NH["E"] / (3 * (1 - 2 * NH["nu"]))
    x = random_state(m.num_nodes, 9, scale_u=0.03, scale_p=5.0)
    st = O.States(m, 1, quad)
    nq = 1 if quad == 1 else 4
    R, _ = O.assemble(O.K_RESIDUAL, m, 1, [mat_row(NH)], x, st, quad_order=quad, save=False)
    for node in range(0, m.num_nodes, 3):
        j = 4 * node + 3
        h = 1e-5
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        d = (npref.pressure_potential(order, m.coords, m.conn, xp, mu, kappa, NH["c0"], nq)
             - npref.pressure_potential(order, m.coords, m.conn, xm, mu, kappa, NH["c0"], nq)) / (2 * h)
        assert abs(d - R[j]) <= 1e-7 * max(1.0, abs(R[j]))


def test_momentum_sums_to_zero():
    for order, quad in ((1, 1), (2, 2)):
        m = box(2, order)
        for model in (0, 1, 2):
            x = random_state(m.num_nodes, 2)
            st = O.States(m, model, quad)
            R, _ = O.assemble(O.K_RESIDUAL, m, model, [mat_row(MODELS[model])], x, st, quad_order=quad, save=False)
            for i in range(3):
                assert abs(R[i::4].sum()) <= 1e-10 * np.abs(R[i::4]).sum()


def test_rigid_rotation_is_stress_free():
    m = box(2, 2)
    th = 0.7
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    Q = Q @ np.array([[1, 0, 0], [0, np.cos(0.3), -np.sin(0.3)], [0, np.sin(0.3), np.cos(0.3)]])
    u = m.coords @ (Q - np.eye(3)).T
    x = np.zeros(m.num_dofs)
    for i in range(3):
        x[i::4] = u[:, i]
    for model in (1, 2):
        st = O.States(m, model, 2)
        R, _ = O.assemble(O.K_RESIDUAL, m, model, [mat_row(MODELS[model])], x, st, quad_order=2)
        assert np.abs(R).max() < 1e-11
        assert np.abs(st.sigma).max() < 1e-10


def test_patch_test_affine_field():
    """u = (F0 - I) X and p = 1/2 kappa (J0 - 1/J0): constant P, so every node
    off the boundary has zero residual, and the pressure equation vanishes."""
    for order, quad in ((1, 1), (2, 2)):
        m = box(3, order, jitter=0.1)
        F0 = np.array([[1.05, 0.02, -0.01], [0.01, 0.97, 0.03], [0.0, -0.02, 1.02]])
        J0 = np.linalg.det(F0)
        kappa = NH["E"] / (3 * (1 - 2 * NH["nu"]))
        x = np.zeros(m.num_dofs)
        u = m.coords @ (F0 - np.eye(3)).T
        for i in range(3):
            x[i::4] = u[:, i]
        x[3::4] = 0.5 * kappa * (J0 - 1 / J0)
        st = O.States(m, 1, quad)
        R, _ = O.assemble(O.K_RESIDUAL, m, 1, [mat_row(NH)], x, st, quad_order=quad)
        interior = np.all((m.coords > 1e-12) & (m.coords < 1 - 1e-12), axis=1)
        Ru = np.stack([R[i::4] for i in range(3)], axis=1)
        assert np.abs(Ru[interior]).max() < 1e-11
        assert np.abs(R[3::4]).max() < 1e-11
        # the stress is the closed-form neo-Hookean Cauchy stress
        mu = NH["E"] / (2 * (1 + NH["nu"]))
        b = F0 @ F0.T
        sig = mu * J0 ** (-5 / 3) * (b - np.trace(b) / 3 * np.eye(3)) + 0.5 * kappa * (J0 - 1 / J0) * np.eye(3)
        assert np.allclose(st.sigma.T, sig.ravel()[None, :], rtol=0, atol=1e-10)


def test_elastic_is_linearized_neohookean():
    """Without stabilization the Elastic model (src/goal_elastic.cpp:61-70 with
    the small-strain PResidual) is the neo-Hookean chain linearized at u = 0,
    p = 0: momentum rows agree, pressure rows are kappa times the neo-Hookean
    ones (p - (2mu/3 + lambda) tr H = kappa (p/kappa - tr H)).  Exact for the
    Jacobian at the reference state, first order for the residual."""
    m = box(2, 2)
    kappa = NH["E"] / (3 * (1 - 2 * NH["nu"]))
    rows = np.repeat(np.arange(m.num_dofs), np.diff(m.row_ptr))
    scale = np.where(rows % 4 == 3, kappa, 1.0)
    for quad in (1, 2):
        x0 = np.zeros(m.num_dofs)
        _, VE = O.assemble(O.K_JACOBIAN, m, 0, [mat_row(ELASTIC)], x0, O.States(m, 0, quad), quad_order=quad,
                           stabilization=False)
        _, VN = O.assemble(O.K_JACOBIAN, m, 1, [mat_row(NH)], x0, O.States(m, 1, quad), quad_order=quad,
                           stabilization=False)
        assert np.abs(VE - scale * VN).max() <= 1e-12 * np.abs(VE).max()
        x = random_state(m.num_nodes, 5, scale_u=1e-6, scale_p=1e-3)
        RE, _ = O.assemble(O.K_RESIDUAL, m, 0, [mat_row(ELASTIC)], x, O.States(m, 0, quad), quad_order=quad,
                           stabilization=False)
        RN, _ = O.assemble(O.K_RESIDUAL, m, 1, [mat_row(NH)], x, O.States(m, 1, quad), quad_order=quad,
                           stabilization=False)
        for i in range(3):
            assert np.linalg.norm(RE[i::4] - RN[i::4]) <= 1e-4 * np.linalg.norm(RN[i::4])
        assert np.linalg.norm(RE[3::4] - kappa * RN[3::4]) <= 1e-4 * np.linalg.norm(RE[3::4])


def test_j2_elastic_branch_equals_neohookean():
    m = box(2, 1)
    x = random_state(m.num_nodes, 4)
    big_y = dict(J2, Y=1e9, E=NH["E"], nu=NH["nu"])
    st2 = O.States(m, 2, 1)
    st1 = O.States(m, 1, 1)
    R2, V2 = O.assemble(O.K_JACOBIAN, m, 2, [mat_row(big_y)], x, st2)
    R1, V1 = O.assemble(O.K_JACOBIAN, m, 1, [mat_row(NH)], x, st1)
    assert np.abs(R2 - R1).max() <= 1e-12 * np.abs(R1).max()
    assert np.abs(V2 - V1).max() <= 1e-12 * np.abs(V1).max()
    assert np.all(st2.eqps == 0.0)


def test_j2_return_map_closed_form():
    m = box(2, 1)
    x = random_state(m.num_nodes, 6, scale_u=0.02)
    st = O.States(m, 2, 1)
    Fp, eq = random_history(m.num_elems, 8)
    st.Fp_old[:] = Fp
    st.eqps_old[:] = eq
    st.Fp[:] = Fp
    st.eqps[:] = eq
    O.assemble(O.K_RESIDUAL, m, 2, [mat_row(J2)], x, st, quad_order=1)
    mu = J2["E"] / (2 * (1 + J2["nu"]))
    n_plastic = 0
    for e in range(m.num_elems):
        (wdv, N, dN, u, H, p, gg), = npref.qp_fields(1, m.coords, m.conn, x, e, 1)
        F = np.eye(3) + H
        plastic, dgam, eq_new, s_tr, mubar, J = npref.j2_radial_return(F, Fp[:, e].reshape(3, 3), eq[e], mu,
                                                                        J2["K"], J2["Y"])
        assert abs(st.eqps[e] - eq_new) <= 1e-12 * max(1.0, eq_new)
        sig = st.sigma[:, e].reshape(3, 3)
        dev = sig - np.trace(sig) / 3 * np.eye(3)
        if plastic:
            n_plastic += 1
            # consistency: |s| = sqrt(2/3)(Y + K eqps_new), s = J dev(sigma)
            assert abs(np.linalg.norm(dev) * J - np.sqrt(2 / 3) * (J2["Y"] + J2["K"] * eq_new)) < 1e-10
            # Fp = exp(dgam N) Fp_old with N = s_tr/|s_tr|
            Nn = s_tr / np.linalg.norm(s_tr)
            w, Vv = np.linalg.eigh(dgam * Nn)
            Fp_new = Vv @ np.diag(np.exp(w)) @ Vv.T @ Fp[:, e].reshape(3, 3)
            assert np.allclose(st.Fp[:, e].reshape(3, 3), Fp_new, atol=1e-13)
        else:
            assert np.array_equal(st.Fp[:, e], Fp[:, e])
    assert n_plastic > 0


def test_qoi_closed_forms_and_derivative():
    m = box(2, 2, jitter=0.1)
    # constant displacement c: AvgDisp = V (c0 + c1 + c2) / 3
    c = np.array([0.1, -0.2, 0.4])
    x = np.zeros(m.num_dofs)
    for i in range(3):
        x[i::4] = c[i]
    st = O.States(m, 1, 2)
    v, _, _ = O.qoi(0, m, 1, [mat_row(NH)], x, st, quad_order=2)
    assert abs(v - c.sum() / 3) < 1e-14
    # uniaxial stretch lam along x with p = bulk pressure: vm = mu lam^{-5/3} (lam^2 - 1) (volume 1)
    lam = 1.03
    kappa = NH["E"] / (3 * (1 - 2 * NH["nu"]))
    mu = NH["E"] / (2 * (1 + NH["nu"]))
    x = np.zeros(m.num_dofs)
    x[0::4] = (lam - 1) * m.coords[:, 0]
    x[3::4] = 0.5 * kappa * (lam - 1 / lam)
    v, _, _ = O.qoi(1, m, 1, [mat_row(NH)], x, st, quad_order=2)
    assert abs(v - mu * lam ** (-5 / 3) * (lam ** 2 - 1)) < 1e-12
    # derivative by central differences
    x = random_state(m.num_nodes, 12)
    for kind in (0, 1):
        v, _, dm = O.qoi(kind, m, 1, [mat_row(NH)], x, st, quad_order=2, derivative=True)
        for j in np.random.default_rng(kind).choice(m.num_dofs, 8, replace=False):
            h = 1e-6
            xp, xm = x.copy(), x.copy()
            xp[j] += h
            xm[j] -= h
            d = (O.qoi(kind, m, 1, [mat_row(NH)], xp, st, quad_order=2)[0]
                 - O.qoi(kind, m, 1, [mat_row(NH)], xm, st, quad_order=2)[0]) / (2 * h)
            assert abs(d - dm[j]) <= 1e-6 * max(1.0, abs(dm[j]))


def test_dirichlet_rows():
    m = box(2, 1)
    x = random_state(m.num_nodes, 1)
    st = O.States(m, 1, 1)
    R, V = O.assemble(O.K_JACOBIAN, m, 1, [mat_row(NH)], x, st)
    dM = np.ones(m.num_dofs)
    nodes = np.array([0, 5, 7], dtype=np.int32)
    vals = np.array([0.1, 0.2, 0.3])
    O.dbc(m, nodes, vals, 2, x, R, dM, V)
    A = dense(m, V)
    for k, nd in enumerate(nodes):
        row = 4 * nd + 2
        assert R[row] == x[row] - vals[k]
        assert dM[row] == 0.0
        e = np.zeros(m.num_dofs)
        e[row] = 1.0
        assert np.array_equal(A[row], e)


def test_dwr_partition_of_unity():
    """sum_a e_u(a) = z_diff . R^u and sum_a e_p(a) = z_diff_p . R^p_bulk +
    z_coarse_p . R^p_stab (the stabilization uses the coarse adjoint)."""
    coords, conn = g.mesh_box(2, 2, 2, order=1, jitter=0.1, seed=2)
    fc, fconn, parent, ep = g.mesh_refine(coords, conn)
    m = O.Mesh(1, fc, fconn)
    x = random_state(m.num_nodes, 3)
    rng = np.random.default_rng(4)
    zf = rng.standard_normal(m.num_dofs)
    zc = zf.copy()
    O.set_coarse(zc, ep, coords.shape[0])
    zd = zf - zc
    for model in (1, 2):
        st = O.States(m, model, 1)
        mats = [mat_row(MODELS[model])]
        e, _ = O.assemble(O.K_LOCALIZE, m, model, mats, x, st, save=False, z_diff=zd, z_coarse=zc)
        R, _ = O.assemble(O.K_RESIDUAL, m, model, mats, x, st, save=False)
        Rn, _ = O.assemble(O.K_RESIDUAL, m, model, mats, x, st, save=False, stabilization=False)
        su = sum(e[i::4].sum() for i in range(3))
        assert abs(su - sum(zd[i::4] @ R[i::4] for i in range(3))) < 1e-10 * np.abs(R).sum()
        sp_ = e[3::4].sum()
        want = zd[3::4] @ Rn[3::4] + zc[3::4] @ (R[3::4] - Rn[3::4])
        assert abs(sp_ - want) < 1e-10 * np.abs(R).sum()


def _p1_interpolant_at_fine_vertices(coords, conn, fc, fconn, parent, zb):
    """The base mesh's P1 interpolant of the nodal field zb [nv, 4], evaluated
    at every fine vertex from its definition: barycentric coordinates of the
    vertex in the base tet that holds its child element (parent map), then
    sum_k lambda_k zb[conn[E][k]].  Independent of edge_parents."""
    out = np.full((fc.shape[0], 4), np.nan)
    for c, E in enumerate(parent):
        X = coords[conn[E]]                      # 4 x 3 base vertices
        A = np.vstack([np.ones(4), X.T])         # [1; x; y; z] per vertex
        for v in fconn[c]:
            lam = np.linalg.solve(A, np.concatenate([[1.0], fc[v]]))
            assert lam.min() > -1e-12            # the vertex lies in its parent tet
            out[v] = lam @ zb[conn[E]]
    return out


@pytest.mark.parametrize("seed", [2, 9])
def test_set_coarse_is_the_base_p1_interpolant(seed):
    """Nested::set_coarse (src/goal_nested.cpp:369-393) pinned from the
    definition of the coarse field: on the nested mesh it is the base mesh's
    P1 interpolant, so every new vertex gets the base field evaluated there.
    Checked for all four components (u and p) on a jittered base mesh with a
    random field, with the new vertices' values scrambled before the call.
    A dropped component, wrong parents or wrong weights fail it."""
    coords, conn = g.mesh_box(2, 2, 3, order=1, jitter=0.15, seed=seed)
    fc, fconn, parent, ep = g.mesh_refine(coords, conn)
    nv = coords.shape[0]
    rng = np.random.default_rng(seed)
    zb = rng.standard_normal((nv, 4))
    z = rng.standard_normal((fc.shape[0], 4)) * 100.0   # garbage on new vertices
    z[:nv] = zb
    zf = z.ravel().copy()
    O.set_coarse(zf, ep, nv)
    want = _p1_interpolant_at_fine_vertices(coords, conn, fc, fconn, parent, zb)
    assert not np.isnan(want).any()
    got = zf.reshape(-1, 4)
    assert np.array_equal(got[:nv], zb)                 # old vertices untouched
    assert np.abs(got - want).max() < 1e-13 * np.abs(zb).max()


def test_set_coarse_keeps_affine_fields():
    """An affine field (u and p) is its own P1 interpolant: after set_coarse
    every nested vertex carries the affine field's value at its position."""
    coords, conn = g.mesh_box(3, 2, 2, order=1, jitter=0.1, seed=4)
    fc, fconn, parent, ep = g.mesh_refine(coords, conn)
    A = np.array([[0.3, -1.2, 0.5], [2.0, 0.1, -0.7], [-0.4, 0.9, 1.1], [1.5, -0.5, 0.25]])
    b = np.array([0.1, -0.2, 0.3, 4.0])
    exact = fc @ A.T + b
    z = exact.copy()
    z[coords.shape[0]:] = -7.0
    zf = z.ravel().copy()
    O.set_coarse(zf, ep, coords.shape[0])
    assert np.abs(zf.reshape(-1, 4) - exact).max() < 1e-13 * np.abs(exact).max()


def test_error_indicator_helpers():
    coords, conn = g.mesh_box(2, 1, 1, order=1, jitter=0.0)
    fc, fconn, parent, ep = g.mesh_refine(coords, conn)
    m = O.Mesh(1, fc, fconn)
    rng = np.random.default_rng(5)
    e = rng.standard_normal(m.num_dofs)
    eta = O.compute_error(m, e)
    brute = np.abs(np.array([sum(e[4 * n + c] for n in el for c in range(4)) / 4.0 for el in fconn]))
    assert np.allclose(eta, brute, rtol=1e-13, atol=1e-14)
    s = O.sum_contribs(np.arange(m.num_nodes), e)
    assert abs(s - np.abs(e.reshape(-1, 4).sum(axis=1)).sum()) < 1e-12
    base = O.set_error(parent, eta, conn.shape[0])
    assert abs(base.sum() - eta.sum()) < 1e-12
    assert np.allclose(base, np.bincount(parent, weights=eta))
    # P2 centroid weights: vertices -1/8, edges 1/4 (sum 1)
    coords2, conn2 = g.mesh_box(1, 1, 1, order=2)
    m2 = O.Mesh(2, coords2, conn2)
    e2 = np.zeros(m2.num_dofs)
    e2[0::4] = 1.0
    assert np.allclose(O.compute_error(m2, e2), 1.0)


def test_graph_three_ways():
    for order in (1, 2):
        coords, conn = g.mesh_box(2, 1, 2, order=order, jitter=0.05, seed=3)
        rp_o, cols_o = O.graph(order, coords.shape[0], conn)
        rp_l, cols_l = g.mesh_graph(order, coords.shape[0], conn)
        rp_n, cols_n = npref.graph_bruteforce(conn.shape[1], coords.shape[0], conn)
        assert np.array_equal(rp_o, rp_n) and np.array_equal(cols_o, cols_n)
        assert np.array_equal(rp_l, rp_n) and np.array_equal(cols_l, cols_n)


@pytest.mark.parametrize("order,quad,model", [(2, 2, 1), (1, 1, 2)])
def test_rank_per_part_assembly_equals_serial(order, quad, model):
    """The CPU baseline's rank-per-part assembly (each part holds only the
    rows its elements touch, then exports ghost rows to their owners) gives
    the serial oracle's R and CSR values up to summation order, for 1..7
    parts, residual, Jacobian and transposed kinds; each part holds fewer rows
    than the mesh, and together at least all of them."""
    coords, conn = g.mesh_box(6, 5, 4, order=order, jitter=0.1, seed=3)
    m = O.Mesh(order, coords, conn)
    x = random_state(m.num_nodes, 4, scale_u=0.005)
    mats = [mat_row(MODELS[model])]
    for kind in (O.K_JACOBIAN, O.K_ADJOINT, O.K_RESIDUAL):
        st = O.States(m, model, quad)
        R1, V1 = O.assemble(kind, m, model, mats, x, st, quad_order=quad)
        for T in (1, 2, 7):
            parts = O.Parts(m, T)
            st2 = O.States(m, model, quad)
            R2, V2, sec = parts.assemble(kind, model, mats, x, st2, quad_order=quad)
            assert np.linalg.norm(R2 - R1) <= 1e-14 * np.linalg.norm(R1)
            if V1 is not None:
                assert np.linalg.norm(V2 - V1) <= 1e-14 * np.linalg.norm(V1)
            assert np.array_equal(st.sigma, st2.sigma)
            held, total = parts.rows()
            assert total == m.num_dofs and (held == total if T == 1 else held > total)
            assert (sec >= 0).all()
