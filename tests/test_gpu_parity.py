"""HIP path (through the C ABI) against the oracle on the same seeded inputs.

Bar: CSR structure and index maps bit-exact; R, dRdu, dMdu, states, QoI and
indicators within RTOL = 1e-11 normwise relative (FP64 arithmetic on both
sides, different operation order and atomic accumulation order)."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import MATS, RTOL, mat_row, oracle_mesh_from, random_history, random_state, rel

pytestmark = pytest.mark.gpu

g = pytest.importorskip("goal_amd")

CASES = [
    # (order, quad_order, model)
    (1, 1, 1), (1, 1, 2), (1, 1, 0),
    (1, 2, 1), (1, 2, 2),
    (2, 2, 1), (2, 2, 2), (2, 2, 0),
]


def make(order, quad, model, n=3, seed=11, jitter=0.1, history=True):
    disc = g.Disc.box(n, n, n, order=order, jitter=jitter, seed=seed)
    m = MATS[model]
    st = g.States(disc, model, quad)
    mech = g.Mechanics(disc, st, [g.MaterialSpec(**m)], stabilization=True, quad_order=quad)
    x = random_state(disc.num_nodes, seed)
    disc.set_soln(x)
    om = oracle_mesh_from(disc)
    ost = O.States(om, model, quad)
    if model == 2 and history:
        Fp, eq = random_history(disc.num_elems * st.nq, seed + 1)
        st.set("Fp_old", Fp)
        st.set("eqps_old", eq)
        st.set("Fp", Fp)
        st.set("eqps", eq)
        ost.Fp_old[:] = Fp
        ost.eqps_old[:] = eq
        ost.Fp[:] = Fp
        ost.eqps[:] = eq
    return disc, st, mech, x, om, ost, [mat_row(m)]


@pytest.mark.parametrize("order,quad,model", CASES)
def test_graph_bit_exact(order, quad, model):
    disc = g.Disc.box(3, 2, 2, order=order, jitter=0.1, seed=3)
    rp, cols = disc.graph()
    om = oracle_mesh_from(disc)
    assert np.array_equal(rp, om.row_ptr)
    assert np.array_equal(cols, om.cols)
    coords, conn, _ = disc.mesh()
    lids = disc.lids()
    ref = (4 * conn[:, :, None] + np.arange(4)[None, None, :]).reshape(conn.shape[0], -1)
    assert np.array_equal(lids, ref)


@pytest.mark.parametrize("order,quad,model", CASES)
def test_residual(order, quad, model):
    disc, st, mech, x, om, ost, mats = make(order, quad, model)
    si = g.SolInfo(disc)
    si.zero_R()
    mech.residual(si, save_states=True)
    Rg = si.R()
    Ro, _ = O.assemble(O.K_RESIDUAL, om, model, mats, x, ost, quad_order=quad)
    assert rel(Rg, Ro) < RTOL
    assert rel(st.get("sigma"), ost.sigma) < RTOL
    if model == 2:
        assert rel(st.get("eqps"), ost.eqps) < RTOL
        assert rel(st.get("Fp"), ost.Fp) < RTOL


@pytest.mark.parametrize("order,quad,model", CASES)
def test_jacobian(order, quad, model):
    disc, st, mech, x, om, ost, mats = make(order, quad, model)
    si = g.SolInfo(disc)
    si.zero_all()
    mech.jacobian(si, adjoint=False, save_states=True)
    Rg, Vg = si.R(), si.vals()
    Ro, Vo = O.assemble(O.K_JACOBIAN, om, model, mats, x, ost, quad_order=quad)
    assert rel(Rg, Ro) < RTOL
    assert rel(Vg, Vo) < RTOL
    if model == 2:
        assert rel(st.get("Fp"), ost.Fp) < RTOL


@pytest.mark.parametrize("order,quad,model", [(1, 1, 1), (1, 1, 2), (2, 2, 1)])
def test_jacobian_adjoint(order, quad, model):
    disc, st, mech, x, om, ost, mats = make(order, quad, model)
    si = g.SolInfo(disc)
    si.zero_all()
    mech.jacobian(si, adjoint=True, save_states=False)
    Ro, Vo = O.assemble(O.K_ADJOINT, om, model, mats, x, ost, quad_order=quad, save=False)
    assert rel(si.R(), Ro) < RTOL
    assert rel(si.vals(), Vo) < RTOL


# FP32 tangent: element matrices formed and stored in FP32, summed into the
# FP64 CSR.  Bar for dRdu: 2e-6 normwise relative (FP32 rounding, ~1e-7, over
# a few tens of accumulated terms); R and the states stay FP64 (RTOL).
F32_TOL = 2e-6


@pytest.mark.parametrize("order,quad,model", CASES)
@pytest.mark.parametrize("adjoint", [False, True])
def test_jacobian_fp32_tangent(order, quad, model, adjoint):
    disc, st, mech, x, om, ost, mats = make(order, quad, model)
    mech.set_option("tangent", 32)
    si = g.SolInfo(disc)
    si.zero_all()
    mech.jacobian(si, adjoint=adjoint, save_states=not adjoint)
    kind = O.K_ADJOINT if adjoint else O.K_JACOBIAN
    Ro, Vo = O.assemble(kind, om, model, mats, x, ost, quad_order=quad, save=not adjoint)
    assert rel(si.R(), Ro) < RTOL
    e = rel(si.vals(), Vo)
    assert 0.0 < e < F32_TOL  # rounded to FP32, and no further


@pytest.mark.parametrize("order,quad,model", [(1, 1, 1), (2, 2, 1), (1, 1, 2), (2, 2, 2)])
def test_qoi(order, quad, model):
    disc, st, mech, x, om, ost, mats = make(order, quad, model)
    si = g.SolInfo(disc)
    for kind in (g.QOI_AVG_DISP, g.QOI_AVG_VM):
        si.zero_dMdu()
        val, ev = mech.qoi(kind, si, elem_set=-1, elem_values=True)
        vo, evo, dmo = O.qoi(kind, om, model, mats, x, ost, quad_order=quad, derivative=True)
        assert abs(val - vo) <= RTOL * abs(vo)
        assert rel(ev, evo) < RTOL
        assert rel(si.dMdu(), dmo) < RTOL


@pytest.mark.parametrize("order", [1, 2])
def test_dirichlet_rows(order):
    quad = 1 if order == 1 else 2
    disc, st, mech, x, om, ost, mats = make(order, quad, 1)
    si = g.SolInfo(disc)
    si.zero_all()
    mech.jacobian(si)
    R, V = si.R(), si.vals()
    dM = np.linspace(0.0, 1.0, disc.num_dofs)
    si_dm = dM.copy()
    import ctypes as C
    g.check(g.lib.goal_solinfo_set(si.h, None, si_dm.ctypes.data_as(C.c_void_p), None))
    for comp in range(3):
        g.dbc_apply(disc, si, comp, "zmin", "0.0", 1.0, jacobian=True)
    g.dbc_apply(disc, si, 2, "zmax", "0.01*t", 3.0, jacobian=True)
    Ro, Vo, dMo = R.copy(), V.copy(), dM.copy()
    coords = om.coords
    zmin, zmax = disc.node_set("zmin"), disc.node_set("zmax")
    for comp in range(3):
        O.dbc(om, zmin, np.zeros(zmin.size), comp, x, Ro, dMo, Vo)
    O.dbc(om, zmax, 0.01 * 3.0 * np.ones(zmax.size), 2, x, Ro, dMo, Vo)
    assert np.array_equal(si.R(), Ro)
    assert np.array_equal(si.vals(), Vo)
    assert np.array_equal(si.dMdu(), dMo)
    assert np.allclose(coords[zmax, 2], 1.0) and np.allclose(coords[zmin, 2], 0.0)


@pytest.mark.parametrize("dims,seed", [((2, 2, 2), 5), ((7, 3, 5), 11), ((12, 12, 12), 2)])
def test_nested_refinement_on_device_equals_host(dims, seed):
    """The nested Disc's uniform refinement runs on the device (children,
    parent pairs, midpoints from the base Disc's CSR graph).  It is the host
    refinement bit for bit: coordinates, children and their orientation,
    parent maps, node sets."""
    coords, conn = g.mesh_box(*dims, order=1, jitter=0.1, seed=seed)
    base = g.Disc.from_arrays(1, coords, conn)
    for name in ("zmin", "xmax"):
        comp, val = (2, 0.0) if name == "zmin" else (0, 1.0)
        base.add_node_set(name, np.nonzero(coords[:, comp] == val)[0])
    nested = g.Nested(base)
    hc, hconn, hparent, hep = g.mesh_refine(coords, conn)
    fc, fconn, _ = nested.disc.mesh()
    parent, ep = nested.maps()
    assert np.array_equal(fc, hc) and np.array_equal(fconn, hconn)
    assert np.array_equal(parent, hparent) and np.array_equal(np.ravel(ep), np.ravel(hep))
    nv = coords.shape[0]
    pairs = np.reshape(hep, (-1, 2))
    for name in ("zmin", "xmax"):
        b = set(base.node_set(name).tolist())
        want = sorted(b | {nv + k for k, (i, j) in enumerate(pairs.tolist()) if i in b and j in b})
        assert sorted(nested.disc.node_set(name).tolist()) == want, name


@pytest.mark.parametrize("model", [1, 2])
def test_localize_and_indicators(model):
    base = g.Disc.box(2, 2, 2, order=1, jitter=0.1, seed=5)
    nested = g.Nested(base)
    fine = nested.disc
    quad = 1
    m = MATS[model]
    st = g.States(fine, model, quad)
    mech = g.Mechanics(fine, st, [g.MaterialSpec(**m)], quad_order=quad)
    xb = random_state(base.num_nodes, 5)
    base.set_soln(xb)
    nested.transfer_soln()
    x = fine.soln()
    parent, ep = nested.maps()
    # transfer: old vertices copied, new ones averaged
    xo = np.zeros(fine.num_dofs)
    xo[: xb.size] = xb
    O.set_coarse(xo, ep, base.num_nodes)
    assert np.array_equal(x, xo)
    om = oracle_mesh_from(fine)
    ost = O.States(om, model, quad)
    rng = np.random.default_rng(9)
    z_fine = rng.standard_normal(fine.num_dofs)
    zf = g.DeviceArray.from_host(z_fine)
    zc = g.DeviceArray.from_host(z_fine)
    nested.set_coarse(zc)
    zco = z_fine.copy()
    O.set_coarse(zco, ep, base.num_nodes)
    assert np.array_equal(zc.numpy(), zco)
    zd = g.DeviceArray(fine.num_dofs)
    g.dwr_subtract(zf, zc, zd)
    zdo = z_fine - zco
    assert np.array_equal(zd.numpy(), zdo)
    si = g.SolInfo(fine)
    si.zero_R()
    mech.localize(si, zd.ptr, zc.ptr)
    Ro, _ = O.assemble(O.K_LOCALIZE, om, model, [mat_row(m)], x, ost, quad_order=quad, save=False,
                       z_diff=zdo, z_coarse=zco)
    assert rel(si.R(), Ro) < RTOL
    R_dev = si.dev()[0]
    eta = g.DeviceArray(fine.num_elems)
    g.dwr_compute_error(fine, R_dev, eta)
    eta_o = O.compute_error(om, Ro)
    assert rel(eta.numpy(), eta_o) < RTOL
    bound = g.dwr_sum_contribs(fine, R_dev)
    bound_o = O.sum_contribs(np.arange(fine.num_nodes), Ro)
    assert abs(bound - bound_o) <= RTOL * abs(bound_o)
    eb = g.DeviceArray(base.num_elems)
    nested.set_error(eta, eb)
    eb_o = O.set_error(parent, eta.numpy(), base.num_elems)
    assert rel(eb.numpy(), eb_o) < 1e-15


def test_norm_and_dot():
    disc = g.Disc.box(2, 2, 2, order=1)
    si = g.SolInfo(disc)
    rng = np.random.default_rng(1)
    R = rng.standard_normal(disc.num_dofs)
    import ctypes as C
    g.check(g.lib.goal_solinfo_set(si.h, R.ctypes.data_as(C.c_void_p), None, None))
    assert abs(si.norm_R() - np.linalg.norm(R)) < 1e-13 * np.linalg.norm(R)
    z = g.DeviceArray.from_host(rng.standard_normal(disc.num_dofs))
    assert abs(si.dot_R(z.ptr) + R @ z.numpy()) < 1e-12 * np.abs(R).sum()


@pytest.mark.parametrize("order", [1, 2])
def test_high_incidence_node(order):
    """Node 0 belongs to 80 elements: the gather handles it in two chunks."""
    from tests.helpers import fan_mesh
    quad = 1 if order == 1 else 2
    coords, conn = fan_mesh(order, subdiv=1, jitter=0.05, seed=3)
    disc = g.Disc.from_arrays(order, coords, conn)
    om = O.Mesh(order, coords, conn)
    assert (conn == 0).sum() == 80
    for model in (1, 2):
        m = MATS[model]
        st = g.States(disc, model, quad)
        mech = g.Mechanics(disc, st, [g.MaterialSpec(**m)], quad_order=quad)
        x = random_state(disc.num_nodes, 7)
        disc.set_soln(x)
        si = g.SolInfo(disc)
        for adjoint in (False, True):
            si.zero_all()
            mech.jacobian(si, adjoint=adjoint, save_states=False)
            ost = O.States(om, model, quad)
            kind = O.K_ADJOINT if adjoint else O.K_JACOBIAN
            Ro, Vo = O.assemble(kind, om, model, [mat_row(m)], x, ost, quad_order=quad, save=False)
            assert rel(si.R(), Ro) < RTOL
            assert rel(si.vals(), Vo) < RTOL


@pytest.mark.parametrize("order", [1, 2])
def test_assembly_bitwise_deterministic(order):
    quad = 1 if order == 1 else 2
    disc, st, mech, x, om, ost, mats = make(order, quad, 1, n=4)
    si = g.SolInfo(disc)
    out = []
    for _ in range(2):
        si.zero_all()
        mech.jacobian(si, save_states=False)
        out.append((si.R(), si.vals()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_accumulate_without_zero():
    """Assembly adds into R / dRdu when they were not zeroed (SolInfo semantics)."""
    disc, st, mech, x, om, ost, mats = make(1, 1, 1)
    si = g.SolInfo(disc)
    si.zero_all()
    mech.jacobian(si, save_states=False)
    R1, V1 = si.R(), si.vals()
    mech.jacobian(si, save_states=False)
    assert rel(si.R(), 2 * R1) < 1e-15
    assert rel(si.vals(), 2 * V1) < 1e-15


@pytest.mark.parametrize("precond", [g.PRECOND_AMG, g.PRECOND_BLOCK_JACOBI])
@pytest.mark.parametrize("order", [1, 2])
def test_linear_solve(order, precond):
    """Device GMRES (right preconditioned by aggregation multigrid or node
    block-Jacobi) on an assembled Jacobian with Dirichlet rows, against scipy's
    direct solve."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    disc, st, mech, x, om, ost, mats = make(order, 2 if order == 2 else 1, 1)
    # a physical iterate near the reference configuration (the random one can
    # invert elements, which no Newton step of the mechanics problem solves at)
    disc.set_soln(1e-4 * random_state(disc.num_nodes, 5))
    si = g.SolInfo(disc)
    si.zero_all()
    mech.jacobian(si)
    for comp in range(3):
        g.dbc_apply(disc, si, comp, "zmin", "0.0", 0.0, True)
    rp, cols = disc.graph()
    A = sp.csr_matrix((si.vals(), cols, rp), shape=(disc.num_dofs, disc.num_dofs))
    rhs = np.random.default_rng(3).standard_normal(disc.num_dofs)
    ref = spla.spsolve(A.tocsc(), rhs)
    b, y = g.DeviceArray.from_host(rhs), g.DeviceArray(disc.num_dofs)
    y.fill(0.0)
    it, rr = si.solve(b, y, rtol=1e-12, max_iters=5000, restart=100, precond=precond)
    out = y.numpy()
    assert rr <= 1e-12 and it > 0
    assert np.linalg.norm(A @ out - rhs) <= 1e-9 * np.linalg.norm(rhs)
    assert rel(out, ref) < 1e-8
    # restarting from the solution keeps it (the stopping test is relative to
    # the initial residual, so this runs to the iteration limit)
    si.solve(b, y, rtol=1e-10, max_iters=60, restart=30, precond=precond)
    assert rel(y.numpy(), ref) < 1e-8


def _cube_newton_system(n):
    """The P2 neo-Hookean cube's load-step-1 Jacobian with its Dirichlet rows
    (zmin held, u_z = 0.01 on zmax) at u = 0, on an n^3 jittered box."""
    disc = g.Disc.box(n, n, n, order=2, jitter=0.1, seed=0x5EED)
    st = g.States(disc, 1, 2)
    mech = g.Mechanics(disc, st, [g.MaterialSpec(**MATS[1])], stabilization=True, quad_order=2)
    si = g.SolInfo(disc)
    si.zero_all()
    mech.jacobian(si)
    for comp, ns, expr in ((0, "zmin", "0.0"), (1, "zmin", "0.0"), (2, "zmin", "0.0"), (2, "zmax", "0.01")):
        g.dbc_apply(disc, si, comp, ns, expr, 1.0, True)
    return disc, si


def test_multigrid_preconditioner_cuts_iterations():
    """On the 12^3 P2 cube system (15,625 nodes, three multigrid levels) the
    multigrid-preconditioned GMRES reaches 1e-8 in a fifth of the block-Jacobi
    iterations or fewer, the true residual agrees, and two solves of the same
    system are bitwise equal (fixed-order reductions throughout)."""
    import scipy.sparse as sp

    disc, si = _cube_newton_system(12)
    rp, cols = disc.graph()
    A = sp.csr_matrix((si.vals(), cols, rp), shape=(disc.num_dofs, disc.num_dofs))
    rhs = -si.R()
    b, y = g.DeviceArray.from_host(rhs), g.DeviceArray(disc.num_dofs)
    its, sols = {}, []
    for pc in (g.PRECOND_AMG, g.PRECOND_AMG, g.PRECOND_BLOCK_JACOBI):
        y.fill(0.0)
        it, rr = si.solve(b, y, rtol=1e-8, max_iters=6000, restart=100, precond=pc)
        out = y.numpy()
        assert rr <= 1e-8
        assert np.linalg.norm(A @ out - rhs) <= 2e-8 * np.linalg.norm(rhs)
        its.setdefault(pc, []).append(it)
        sols.append(out)
    assert np.array_equal(sols[0], sols[1]) and its[g.PRECOND_AMG][0] == its[g.PRECOND_AMG][1]
    print("iterations: multigrid %d, block Jacobi %d" % (its[g.PRECOND_AMG][0], its[g.PRECOND_BLOCK_JACOBI][0]))
    assert 5 * its[g.PRECOND_AMG][0] <= its[g.PRECOND_BLOCK_JACOBI][0]


def _graded_plastic_bar(target=60000, cycles=4, strain=0.01):
    """A J2 bar refined toward its clamp by the library's adaptation (the
    meshes the adaptive runs produce), at a prescribed 1% stretch that is
    plastic everywhere: the Jacobian whose block-Jacobi-preconditioned
    spectrum has eigenvalues far off the real axis."""
    from goal_amd import learned as LP

    rng = np.random.default_rng(2024)
    LP.sample_problem(rng, "cube", 8)
    p = LP.sample_problem(rng, "bar", 4)
    coords, conn = LP.base_mesh(p)
    for c in range(cycles):
        cen = coords[conn].mean(axis=1)
        eta = np.exp(-3.0 * cen[:, 0]) + 0.02 * np.exp(-3.0 * (8.0 - cen[:, 0]))
        tgt = int(conn.shape[0] * (target / conn.shape[0]) ** (1.0 / (cycles - c)))
        ad = g.Adaptation(coords, conn, g.size_field(coords, conn, eta, tgt), 1.5, 8)
        coords, conn = ad.coords, ad.conn
    m = LP.Model(p, coords, conn)
    X, Y, Z = coords[:, 0], coords[:, 1], coords[:, 2]
    x = np.zeros(m.disc.num_dofs)
    x[0::4] = strain * X
    x[1::4] = -0.3 * strain * Y + 0.1 * strain * np.sin(np.pi * X / 8) * Z
    x[2::4] = -0.3 * strain * Z + 0.1 * strain * np.cos(np.pi * X / 8) * Y
    x[3::4] = 0.5 * strain * p.E / 3.0
    m.disc.set_soln(x)
    m.si.zero_all()
    m.mech.jacobian(m.si, adjoint=False, save_states=True)
    for comp, ns, expr in m.bcs:
        g.dbc_apply(m.disc, m.si, comp, ns, expr, 1.0, True)
    return m


def test_multigrid_smoother_handles_plastic_bar():
    """On a 65 K-tet clamp-graded J2 bar at a plastic stretch, the multigrid
    GMRES reaches 1e-8 (a Chebyshev smoother on the real interval alone
    stalls near 0.3 on this system in the numpy model of the hierarchy), and
    the true residual agrees."""
    import scipy.sparse as sp

    m = _graded_plastic_bar()
    assert (m.states.get("eqps") > 0).all()  # every point on the plastic branch
    n = m.disc.num_dofs
    rp, cols = m.disc.graph()
    A = sp.csr_matrix((m.si.vals(), cols, rp), shape=(n, n))
    rhs = -m.si.R()
    b, y = g.DeviceArray.from_host(rhs), g.DeviceArray(n)
    y.fill(0.0)  # the initial guess (the relative residual is against b - A x0)
    it, rr = m.si.solve(b, y, rtol=1e-8, max_iters=1500, restart=100, precond=g.PRECOND_AMG)
    print("graded plastic bar: %d tets, %d GMRES iterations, relres %.2e" % (m.ne, it, rr))
    assert rr <= 1e-8 and it <= 600
    assert np.linalg.norm(A @ y.numpy() - rhs) <= 2e-8 * np.linalg.norm(rhs)


def make_box(order, quad, model, dims, lengths, seed, jitter=0.1):
    """Like make() on an nx x ny x nz box of the given side lengths."""
    disc = g.Disc.box(*dims, lx=lengths[0], ly=lengths[1], lz=lengths[2], order=order, jitter=jitter,
                      seed=seed)
    m = MATS[model]
    st = g.States(disc, model, quad)
    mech = g.Mechanics(disc, st, [g.MaterialSpec(**m)], stabilization=True, quad_order=quad)
    x = random_state(disc.num_nodes, seed, scale_u=0.01)
    disc.set_soln(x)
    om = oracle_mesh_from(disc)
    ost = O.States(om, model, quad)
    if model == 2:
        Fp, eq = random_history(disc.num_elems * st.nq, seed + 1)
        for k, v in (("Fp_old", Fp), ("eqps_old", eq), ("Fp", Fp), ("eqps", eq)):
            st.set(k, v)
        ost.Fp_old[:] = Fp
        ost.eqps_old[:] = eq
        ost.Fp[:] = Fp
        ost.eqps[:] = eq
    return disc, st, mech, x, om, ost, [mat_row(m)]


# The batched workload in miniature: meshes of uneven size and aspect ratio,
# P2 neo-Hookean with 4 points and P1 J2 with 1 point and plastic history.
BATCH = [
    # (order, quad, model, (nx, ny, nz), (lx, ly, lz))
    (2, 2, 1, (3, 3, 3), (1.0, 1.0, 1.0)),
    (1, 1, 2, (8, 2, 2), (4.0, 1.0, 1.0)),
    (2, 2, 1, (6, 2, 1), (6.0, 2.0, 1.0)),
    (1, 1, 2, (5, 5, 4), (1.0, 1.0, 0.8)),
    (2, 2, 1, (2, 2, 2), (1.0, 1.0, 1.0)),
    (1, 1, 2, (12, 6, 6), (8.0, 2.0, 2.0)),
    (2, 2, 1, (4, 3, 2), (2.0, 1.5, 1.0)),
    (1, 1, 2, (3, 3, 3), (1.0, 1.0, 1.0)),
    (2, 2, 2, (3, 2, 2), (3.0, 2.0, 2.0)),
]


def test_batch_graph_matches_oracle_and_eager():
    """Nine meshes of uneven size (48 to 2,592 tets) and aspect ratio (1 to
    6), P2 neo-Hookean 4-point, P1 J2 1-point with history and one P2 J2,
    tuned and captured into one graph.  Every replay (with eager library work
    and synchronizations in between, and zero_all() pending before the launch)
    gives R, the CSR values and the updated states of every mesh within RTOL of
    the oracle, and bit for bit the eager assemblies."""
    items = []
    for k, (order, quad, model, dims, lengths) in enumerate(BATCH):
        disc, st, mech, x, om, ost, mats = make_box(order, quad, model, dims, lengths, seed=40 + k)
        mech.set_option("check", 0)
        items.append(dict(disc=disc, st=st, mech=mech, si=g.SolInfo(disc), om=om, ost=ost, mats=mats,
                          model=model, quad=quad, x=x))
    sizes = [it["disc"].num_elems for it in items]
    assert len(items) >= 8 and max(sizes) >= 20 * min(sizes)
    want = []
    for it in items:
        R, V = O.assemble(O.K_JACOBIAN, it["om"], it["model"], it["mats"], it["x"], it["ost"],
                          quad_order=it["quad"])
        want.append((R, V, it["ost"].sigma.copy()))
    eager = []
    for it in items:
        it["si"].zero_all()
        it["mech"].jacobian(it["si"], save_states=True)
        eager.append((it["si"].R(), it["si"].vals()))
    for it, (R, V, sig), (Re, Ve) in zip(items, want, eager):
        assert rel(Re, R) < RTOL and rel(Ve, V) < RTOL
    b = g.Batch([it["mech"] for it in items], [it["si"] for it in items])
    blocks, ms = b.tune(reps=2)
    assert (ms > 0).all()
    b.capture()
    other = g.SolInfo(items[0]["disc"])
    for rep in range(3):
        for it in items:  # scribble over the outputs, then leave a zero pending
            it["si"].set(R=np.full(it["disc"].num_dofs, 7.0), vals=np.full(it["om"].cols.size, 5.0))
            it["si"].zero_all()
        other.zero_all()
        items[0]["mech"].residual(other, save_states=False)
        other.norm_R()
        b.run()
        g.sync()
        for it, (R, V, sig), (Re, Ve) in zip(items, want, eager):
            Rg, Vg = it["si"].R(), it["si"].vals()
            assert np.array_equal(Rg, Re) and np.array_equal(Vg, Ve), rep
            assert rel(Rg, R) < RTOL and rel(Vg, V) < RTOL, rep
            assert rel(it["st"].get("sigma"), sig) < RTOL, rep
            assert np.array_equal(it["si"].dMdu(), np.zeros(it["disc"].num_dofs))


def test_assembly_refuses_more_element_sets_than_materials():
    """goal_disc_set_elem_sets after goal_mech_create may raise the set count
    past the material table: assembly, QoI and features refuse to launch."""
    disc, st, mech, x, om, ost, mats = make(1, 1, 1, n=2)
    si = g.SolInfo(disc)
    es = np.zeros(disc.num_elems, dtype=np.int32)
    es[::2] = 1
    disc.set_elem_sets(2, es)
    for call in (lambda: mech.jacobian(si), lambda: mech.residual(si), lambda: mech.qoi(g.QOI_AVG_DISP, si),
                 lambda: g.learn_features(mech, 0.25)):
        with pytest.raises(g.GoalError) as err:
            call()
        assert err.value.code == g.ERR_INVALID


@pytest.mark.parametrize("order,quad,n,tangent", [(2, 2, 4, 64), (2, 2, 4, 32), (1, 1, 7, 64), (1, 2, 5, 64)])
def test_chunked_assembly_is_bitwise_equal(order, quad, n, tangent):
    """The neo-Hookean assembly split into chunks, with the gathers of ready
    node runs on a second stream while the next chunk computes, gives the
    one-pass results bit for bit (the same kernels over disjoint work): R,
    the CSR values, the residual-only pass and the adjoint blocks."""
    disc, st, mech, x, om, ost, mats = make(order, quad, 1, n=n)
    mech.set_option("tangent", tangent)
    mech.set_option("overlap", 0)   # one pass
    si = g.SolInfo(disc)
    ref = {}
    for adjoint in (False, True):
        si.zero_all()
        mech.jacobian(si, adjoint=adjoint, save_states=True)
        ref[adjoint] = (si.R(), si.vals())
    si.zero_R()
    mech.residual(si, save_states=True)
    ref_res = si.R()
    for chunks, blocks in ((2, 0), (5, 0), (8, 16), (-1, 0)):   # -1: the automatic choice
        mech.set_option("overlap", chunks)
        mech.set_option("overlap_blocks", blocks)
        for adjoint in (False, True):
            si.set(R=np.full(disc.num_dofs, 3.0))   # stale values the assembly must overwrite
            si.zero_all()
            mech.jacobian(si, adjoint=adjoint, save_states=True)
            assert np.array_equal(si.R(), ref[adjoint][0]), (chunks, adjoint)
            assert np.array_equal(si.vals(), ref[adjoint][1]), (chunks, adjoint)
        si.zero_R()
        mech.residual(si, save_states=True)
        assert np.array_equal(si.R(), ref_res), chunks
    mech.set_option("overlap", 0)


# Parity at scale: meshes with thousands of element groups, so that every
# workgroup of the resident grid walks several groups, the XCD slab schedule
# splits the groups over all eight XCDs, the phase-A prefetch and the J2
# point-state prefetch carry data across group iterations, and the gather
# covers thousands of node runs.
SCALE = [
    # (order, quad, model, dims, lengths, tangent)
    (2, 2, 1, (16, 16, 16), (1.0, 1.0, 1.0), 64),   # 24,576 tets, 4,096 groups of 6
    (2, 2, 1, (16, 16, 16), (1.0, 1.0, 1.0), 32),
    (1, 1, 2, (48, 8, 8), (8.0, 1.0, 1.0), 64),     # 18,432 tets, 1,152 groups of 16
    (2, 2, 2, (8, 8, 6), (1.0, 1.0, 0.75), 64),     # 2,304 tets, 384 groups of 6, 4 points
]


@pytest.mark.parametrize("order,quad,model,dims,lengths,tangent", SCALE)
def test_jacobian_at_scale(order, quad, model, dims, lengths, tangent):
    """R, the CSR values and the stresses of a smooth iterate with small
    random noise (neo-Hookean) or a 1% plastic stretch over random plastic
    history (J2) against the oracle: RTOL for FP64, F32_TOL for FP32 element
    matrices."""
    disc = g.Disc.box(*dims, lx=lengths[0], ly=lengths[1], lz=lengths[2], order=order, jitter=0.1, seed=77)
    m = MATS[model]
    st = g.States(disc, model, quad)
    mech = g.Mechanics(disc, st, [g.MaterialSpec(**m)], stabilization=True, quad_order=quad)
    mech.set_option("tangent", tangent)
    coords, _, _ = disc.mesh()
    X, Y, Z = coords[:, 0], coords[:, 1], coords[:, 2]
    rng = np.random.default_rng(5)
    h = min(lengths[i] / dims[i] for i in range(3))
    x = np.zeros(disc.num_dofs)
    if model == 2:
        x[0::4] = 0.01 * X
        x[1::4] = -0.003 * Y + 1e-3 * np.sin(np.pi * X / lengths[0]) * Z
        x[2::4] = -0.003 * Z + 1e-3 * np.cos(np.pi * X / lengths[0]) * Y
    else:
        x[0::4] = 0.02 * np.sin(np.pi * X) * np.sin(np.pi * Y) * Z
        x[1::4] = 0.01 * X * Y * Z
        x[2::4] = 0.03 * Z * (1.0 + 0.5 * np.sin(np.pi * Y))
    for c in range(3):
        x[c::4] += 0.01 * h * rng.standard_normal(disc.num_nodes)
    x[3::4] = rng.standard_normal(disc.num_nodes)
    disc.set_soln(x)
    om = oracle_mesh_from(disc)
    ost = O.States(om, model, quad)
    if model == 2:
        Fp, eq = random_history(disc.num_elems * st.nq, 6)
        for k, v in (("Fp_old", Fp), ("eqps_old", eq), ("Fp", Fp), ("eqps", eq)):
            st.set(k, v)
        ost.Fp_old[:] = Fp
        ost.eqps_old[:] = eq
        ost.Fp[:] = Fp
        ost.eqps[:] = eq
    si = g.SolInfo(disc)
    si.zero_all()
    mech.jacobian(si, adjoint=False, save_states=True)
    assert mech.status() == 0
    Ro, Vo = O.assemble(O.K_JACOBIAN, om, model, [mat_row(m)], x, ost, quad_order=quad)
    assert np.isfinite(Ro).all() and np.isfinite(Vo).all()
    assert rel(si.R(), Ro) < RTOL
    assert rel(st.get("sigma"), ost.sigma) < RTOL
    if model == 2:
        assert (ost.eqps > eq).any()  # the stretch is past yield somewhere
        assert rel(st.get("Fp"), ost.Fp) < RTOL
    e = rel(si.vals(), Vo)
    assert e < (RTOL if tangent == 64 else F32_TOL)
    if tangent == 32:
        assert e > 0.0
