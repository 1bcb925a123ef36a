"""CPU checks of the learned-indicator oracle (oracle/learned.py) and of the
host adaptation layer (goal_size_field / goal_adapt_*), no GPU needed.

The network and its gradient have no reference counterpart; the oracle is
pinned by central finite differences and by torch autograd (float64), and its
BF16 rounding by known values and torch's own conversion.  The size field is
pinned by the closed form it reduces to on a mesh of congruent elements.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import goal_amd as g
from oracle import learned as L


def test_bf16_rounding_known_values():
    one = 1.0
    vals = np.array([one, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -(1 + 3 * 2 ** -8), 3.0e-5, 0.0],
                    dtype=np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -(1 + 2 ** -6)], dtype=np.float32)
    got = L.bf16(vals)
    assert np.array_equal(got[:5], want)  # ties go to the even mantissa
    rng = np.random.default_rng(0)
    r = (rng.standard_normal(20000) * 10.0 ** rng.uniform(-6, 6, 20000)).astype(np.float32)
    ref = torch.from_numpy(r).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(L.bf16(r), ref)


def _random_net(rng, scale=0.3):
    p = (rng.standard_normal(L.NPARAMS) * scale).astype(np.float32)
    norms = np.zeros(L.NNORMS, dtype=np.float32)
    norms[:12] = rng.standard_normal(12) * 0.1
    norms[12:24] = rng.uniform(0.5, 2.0, 12)
    norms[24], norms[25] = 0.3, 1.7
    norms[L.OFF_LO:L.OFF_HI] = -np.inf
    norms[L.OFF_HI:L.OFF_C] = np.inf
    norms[L.OFF_C] = 5.0
    return p, norms


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p, norms = _random_net(rng)
    X = rng.standard_normal((37, 12))
    Y = rng.standard_normal(37)
    loss, grad = L.mlp_loss_grad(p, norms, X, Y, rounded=False)
    pd = p.astype(np.float64)
    # central differences in float64 on the unrounded network (ReLU kinks are
    # avoided: a step of 1e-6 does not cross one for these inputs)
    h = 1e-6
    idx = np.concatenate([rng.choice(L.OFF_B1, 40, replace=False),
                          L.OFF_B1 + rng.choice(64, 10, replace=False),
                          L.OFF_W2 + rng.choice(4096, 60, replace=False),
                          L.OFF_B2 + rng.choice(64, 10, replace=False),
                          L.OFF_W3 + rng.choice(64, 10, replace=False), [L.OFF_B3]])

    def f(q):
        return _loss64(q, norms, X, Y)

    for i in idx:
        qp, qm = pd.copy(), pd.copy()
        qp[i] += h
        qm[i] -= h
        fd = (f(qp) - f(qm)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(grad[i])), (i, fd, grad[i])
    assert abs(loss - f(pd)) <= 1e-12 * loss


def _loss64(q, norms, X, Y):
    """The loss with float64 parameters (no float32 cast), for differencing."""
    W1 = q[L.OFF_W1:L.OFF_B1].reshape(64, 12)
    b1 = q[L.OFF_B1:L.OFF_W2]
    W2 = q[L.OFF_W2:L.OFF_B2].reshape(64, 64)
    b2 = q[L.OFF_B2:L.OFF_W3]
    W3 = q[L.OFF_W3:L.OFF_B3]
    b3 = q[L.OFF_B3]
    nrm = norms.astype(np.float64)
    xs = (X - nrm[:12]) * nrm[12:24]
    o = np.maximum(np.maximum(xs @ W1.T + b1, 0) @ W2.T + b2, 0) @ W3 + b3
    t = (Y - nrm[24] - nrm[L.OFF_C] * X[:, 0]) / nrm[25]
    return float(np.mean((o - t) ** 2))


def test_mlp_forward_and_gradient_match_torch_autograd():
    rng = np.random.default_rng(2)
    p, norms = _random_net(rng)
    X = rng.standard_normal((64, 12))
    Y = rng.standard_normal(64)
    W1, b1, W2, b2, W3, b3 = (torch.tensor(np.asarray(a, dtype=np.float64), requires_grad=True)
                              for a in L.unpack(p, rounded=True))
    nrm = torch.tensor(norms.astype(np.float64))
    x = torch.tensor(X)
    # the device rounds the matrix-product inputs to BF16 (torch's own
    # conversion here); gradients pass straight through the rounding
    xs32 = torch.tensor(X, dtype=torch.float32)
    n32 = torch.tensor(norms[:24], dtype=torch.float32)
    xs = ((xs32 - n32[:12]) * n32[12:24]).to(torch.bfloat16).to(torch.float64)
    h1 = torch.relu(xs @ W1.T + b1)
    h1 = h1 + (h1.to(torch.float32).to(torch.bfloat16).to(torch.float64) - h1).detach()
    o = torch.relu(h1 @ W2.T + b2) @ W3 + b3
    y = o * nrm[25] + nrm[24] + nrm[L.OFF_C] * x[:, 0]
    assert np.allclose(L.mlp_forward(p, norms, X), y.detach().numpy(), rtol=1e-13, atol=1e-13)
    t = (torch.tensor(Y) - nrm[24] - nrm[L.OFF_C] * x[:, 0]) / nrm[25]
    loss = torch.mean((o - t) ** 2)
    loss.backward()
    lo, grad = L.mlp_loss_grad(p, norms, X, Y)
    assert abs(lo - loss.item()) <= 1e-13 * lo
    tg = np.concatenate([W1.grad.numpy().ravel(), b1.grad.numpy(), W2.grad.numpy().ravel(), b2.grad.numpy(),
                         W3.grad.numpy(), [b3.grad.item()]])
    assert np.allclose(grad, tg, rtol=1e-12, atol=1e-14)


def test_adam_matches_torch():
    rng = np.random.default_rng(3)
    w = rng.standard_normal(50)
    tw = torch.tensor(w.copy(), requires_grad=True)
    opt = torch.optim.Adam([tw], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros(50)
    v = np.zeros(50)
    for step in range(1, 6):
        gr = rng.standard_normal(50)
        tw.grad = torch.tensor(gr)
        opt.step()
        w, m, v = L.adam(w, m, v, gr, step, 1e-2)
    assert np.allclose(w, tw.detach().numpy(), rtol=1e-13, atol=1e-15)


def test_size_field_matches_oracle_and_closed_form():
    coords, conn = g.mesh_box(4, 4, 4)
    ne = conn.shape[0]
    # Kuhn tets of one cube are congruent: h is the same everywhere, and with a
    # uniform indicator h_new = h (ne / N)^(1/3) wherever it is not clamped
    h = np.sqrt(sum(((coords[conn[:, a]] - coords[conn[:, b]]) ** 2).sum(1) for a, b in L.TET_EDGES) / 6)
    for N, ratio in ((8 * ne, 0.5), (ne, 1.0), (ne // 8, 2.0), (1000 * ne, 0.25)):
        s = g.size_field(coords, conn, np.full(ne, 3e-4), N)
        assert np.allclose(s, h[0] * ratio, rtol=1e-12)
    rng = np.random.default_rng(4)
    c2, t2 = g.mesh_box(5, 3, 4, 2.0, 1.0, 1.5, jitter=0.15, seed=5)
    eta = 10.0 ** rng.uniform(-8, -3, t2.shape[0])
    eta[::17] = 0.0   # |eta|^(-2/5) = inf: clamped to 2h, as in the reference
    got = g.size_field(c2, t2, eta, 5000)
    assert np.allclose(got, L.size_field(c2, t2, eta, 5000), rtol=1e-12)


def _faces(conn):
    f = np.concatenate([conn[:, [0, 1, 2]], conn[:, [0, 1, 3]], conn[:, [0, 2, 3]], conn[:, [1, 2, 3]]])
    return np.unique(np.sort(f, axis=1), axis=0, return_counts=True)


@pytest.mark.parametrize("shape,jitter", [((6, 6, 6), 0.0), ((4, 3, 5), 0.15)])
def test_adaptation_invariants(shape, jitter):
    coords, conn = g.mesh_box(*shape, 1.0, 1.0, 1.0, jitter=jitter, seed=3)
    xc = coords[conn].mean(axis=1)
    eta = np.exp(-np.linalg.norm(xc - 0.2, axis=1) / 0.15) + 1e-4
    size = g.size_field(coords, conn, eta, 8 * conn.shape[0])
    A = g.Adaptation(coords, conn, size, 1.5, 20)
    v0 = L.tet_volumes(coords, conn)
    v1 = L.tet_volumes(A.coords, A.conn)
    assert A.num_elems > conn.shape[0] and A.passes < 20
    assert (v1 > 0).all()                                   # orientation kept
    assert np.allclose(np.bincount(A.parent, v1, conn.shape[0]), v0, rtol=1e-12)  # children tile parents
    _, cnt = _faces(A.conn)
    assert cnt.max() == 2                                   # conforming: no hanging faces
    u0, c0 = _faces(conn)
    u1, c1 = _faces(A.conn)
    area = lambda c, f: 0.5 * np.linalg.norm(np.cross(c[f[:, 1]] - c[f[:, 0]], c[f[:, 2]] - c[f[:, 0]]), axis=1)
    assert np.isclose(area(A.coords, u1[c1 == 1]).sum(), area(coords, u0[c0 == 1]).sum(), rtol=1e-12)
    # old vertices keep ids; new ones are midpoints of their recorded edges
    assert np.array_equal(A.coords[:coords.shape[0]], coords)
    ep = A.new_vertex_parents
    assert np.allclose(A.coords[A.num_old:], 0.5 * (A.coords[ep[:, 0]] + A.coords[ep[:, 1]]), rtol=0, atol=1e-15)
    assert (ep.max(axis=1) < A.num_old + np.arange(A.num_new)).all()
    # a linear field is carried exactly
    lin = coords @ np.array([[1.0, 2.0], [-0.5, 0.25], [3.0, 1.0]]) + 0.7
    assert np.allclose(A.transfer(lin), A.coords @ np.array([[1.0, 2.0], [-0.5, 0.25], [3.0, 1.0]]) + 0.7,
                       rtol=1e-13, atol=1e-13)
    # no edge is left longer than the criterion (interpolated sizes at new vertices)
    s = A.transfer(size)
    for a, b in L.TET_EDGES:
        l = np.linalg.norm(A.coords[A.conn[:, a]] - A.coords[A.conn[:, b]], axis=1)
        assert (l <= 1.5 * 0.5 * (s[A.conn[:, a]] + s[A.conn[:, b]]) * (1 + 1e-12)).all()
    # longest-edge bisection keeps the shape measure of these meshes bounded
    l2 = sum(((A.coords[A.conn[:, a]] - A.coords[A.conn[:, b]]) ** 2).sum(1) for a, b in L.TET_EDGES) / 6
    l2b = sum(((coords[conn[:, a]] - coords[conn[:, b]]) ** 2).sum(1) for a, b in L.TET_EDGES) / 6
    q1 = (6 * np.sqrt(2) * v1 / l2 ** 1.5).min()
    q0 = (6 * np.sqrt(2) * v0 / l2b ** 1.5).min()
    assert q1 >= 0.5 * q0


def test_adaptation_rejects_bad_sizes():
    coords, conn = g.mesh_box(2, 2, 2)
    with pytest.raises(g.GoalError):
        g.Adaptation(coords, conn, np.zeros(coords.shape[0]), 1.5, 4)


def test_features_oracle_closed_form():
    """A linear displacement u = A X and a constant von Mises stress: grad u = A
    exactly, the recovered stress and gradient equal the element ones, and the
    geometric features follow from the mesh."""
    coords, conn = g.mesh_box(3, 2, 2, 1.0, 0.5, 0.5, jitter=0.15, seed=9)
    A = np.array([[0.01, 0.002, 0.0], [0.0, -0.003, 0.001], [0.004, 0.0, 0.02]])
    x = np.zeros((coords.shape[0], 4))
    x[:, :3] = coords @ A.T
    x[:, 3] = 2.5
    vol = L.tet_volumes(coords, conn)
    E, nu = 1000.0, 0.25
    f = L.features(coords, conn, x, 7.0 * vol, E, nu, None, 0.25)
    assert np.allclose(f[:, 8], np.log10(np.linalg.norm(A) + 1e-12), atol=1e-12)
    assert np.allclose(f[:, 6], np.log10(7.0 / E + 1e-12), atol=1e-12)
    assert np.allclose(f[:, 7], -12.0, atol=1e-9)
    assert np.allclose(f[:, 9], -12.0, atol=1e-6)   # the recovered gradient is the constant one
    assert np.all(f[:, 10] == nu) and np.allclose(f[:, 11], -9.0)
    xc = coords[conn].mean(axis=1)
    assert np.allclose(f[:, 1:4], xc) and np.array_equal(f[:, 5], (xc[:, 0] < 0.25).astype(float))
