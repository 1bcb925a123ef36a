"""ORACLE -- test infrastructure only (tests/, smoke check, bench cpu_baseline).

numpy restatement, in float64, of the learned-indicator pieces and of the
adaptation size field:

* size_field: get_iso_target_size (src/goal_size_field.cpp:39-151).
* features: the 12 per-element features the device computes from a primal
  state (no reference counterpart; von Mises as src/goal_von_mises.cpp:7-17
  enters through the per-element integral the QoI kernel returns).
* bf16, mlp_forward, mlp_loss_grad, adam: the 12 -> 64 -> 64 -> 1 ReLU network
  with BF16 weights, BF16-rounded inputs to its two matrix products and FP32
  biases (no reference counterpart).  The gradient is back-propagation written
  out by hand; tests pin it with central finite differences and torch autograd.
"""
from __future__ import annotations

import numpy as np

IN, HID = 12, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
NPARAMS = OFF_B3 + 1
OFF_LO = 2 * IN + 2
OFF_HI = OFF_LO + IN
OFF_C = OFF_HI + IN
NNORMS = OFF_C + 1

TET_EDGES = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))


def size_field(coords, conn, eta, target, p_order=1):
    """Vertex target sizes (src/goal_size_field.cpp:39-151): G = sum |eta|^(2d/(2p+d))
    (:39-54), factor (G/N)^(1/d) (:56-61), h = RMS edge length (:63-70),
    h_new = factor |eta|^(-2/(2p+d)) h clamped to [h/4, 2h] (:72-83), then the
    plain mean over the elements around each vertex (:97-108)."""
    d, p = 3.0, float(p_order)
    eta = np.abs(np.asarray(eta, dtype=np.float64))
    G = np.sum(eta ** ((2.0 * d) / (2.0 * p + d)))
    factor = (G / target) ** (1.0 / d)
    h2 = sum(((coords[conn[:, a]] - coords[conn[:, b]]) ** 2).sum(1) for a, b in TET_EDGES)
    h = np.sqrt(h2 / 6.0)
    with np.errstate(divide="ignore"):
        r = eta ** (-2.0 / (2.0 * p + d))
    hn = np.clip(factor * r * h, 0.25 * h, 2.0 * h)
    s = np.zeros(coords.shape[0])
    c = np.zeros(coords.shape[0])
    for a in range(4):
        np.add.at(s, conn[:, a], hn)
        np.add.at(c, conn[:, a], 1.0)
    return np.where(c > 0, s / np.maximum(c, 1), 0.0)


def tet_volumes(coords, conn):
    a, b, c, d = (coords[conn[:, i]] for i in range(4))
    return np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a) / 6.0


def features(coords, conn, x, elem_vm, E, nu, eqps, qoi_x):
    """Per-element features (see goal_learn_features).  conn: vertex columns of
    the connectivity; x: nodal solution [num_nodes][4]; elem_vm: integral of
    the von Mises stress over each element; E, nu: per element; eqps: mean
    point eqps per element or None."""
    conn = np.asarray(conn)[:, :4]
    vol = tet_volumes(coords, conn)
    X = coords[conn]                        # [ne][4][3]
    J = X[:, 1:, :] - X[:, :1, :]           # rows: edges 01, 02, 03
    Ji = np.linalg.inv(J)                   # [ne][3][3]
    U = x[conn][:, :, :3]
    dU = U[:, 1:, :] - U[:, :1, :]          # [ne][k][i]
    H = np.einsum("eki,ejk->eij", dU, Ji).reshape(-1, 9)
    nn = coords.shape[0]
    s = np.zeros((nn, 10))
    w = np.zeros(nn)
    for a in range(4):
        np.add.at(s[:, 0], conn[:, a], elem_vm)
        for c in range(9):
            np.add.at(s[:, 1 + c], conn[:, a], vol * H[:, c])
        np.add.at(w, conn[:, a], vol)
    star = np.where(w[:, None] > 0, s / np.where(w > 0, w, 1.0)[:, None], 0.0)
    se = star[conn].mean(axis=1)            # [ne][10]
    h = np.sqrt(sum(((coords[conn[:, a]] - coords[conn[:, b]]) ** 2).sum(1) for a, b in TET_EDGES) / 6.0)
    xc = X.mean(axis=1)
    vm = elem_vm / vol
    eq = np.zeros(conn.shape[0]) if eqps is None else np.asarray(eqps)
    f = np.empty((conn.shape[0], IN))
    f[:, 0] = np.log10(h)
    f[:, 1:4] = xc
    f[:, 4] = np.log10(np.abs(xc[:, 0] - qoi_x) + h)
    f[:, 5] = (xc[:, 0] < qoi_x).astype(float)
    f[:, 6] = np.log10(np.abs(vm) / E + 1e-12)
    f[:, 7] = np.log10(np.abs(vm - se[:, 0]) / E + 1e-12)
    f[:, 8] = np.log10(np.sqrt((H ** 2).sum(axis=1)) + 1e-12)
    f[:, 9] = np.log10(np.sqrt(((H - se[:, 1:]) ** 2).sum(axis=1)) + 1e-12)
    f[:, 10] = nu
    f[:, 11] = np.log10(eq + 1e-9)
    return f


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as
    float32.  Finite inputs."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(params, rounded=True):
    """(W1, b1, W2, b2, W3, b3) in float64; weights BF16-rounded as the device uses them."""
    p = np.asarray(params, dtype=np.float32)
    wr = bf16 if rounded else (lambda v: v)
    W1 = wr(p[OFF_W1:OFF_B1]).astype(np.float64).reshape(HID, IN)
    W2 = wr(p[OFF_W2:OFF_B2]).astype(np.float64).reshape(HID, HID)
    W3 = wr(p[OFF_W3:OFF_B3]).astype(np.float64)
    b1 = p[OFF_B1:OFF_W2].astype(np.float64)
    b2 = p[OFF_B2:OFF_W3].astype(np.float64)
    b3 = float(p[OFF_B3])
    return W1, b1, W2, b2, W3, b3


def _layers(net, norms, X, rounded=True):
    """Forward pass.  With `rounded`, the inputs of the two matrix products
    are BF16-rounded as on the device: the normalized features (computed in
    float32) and the layer-1 activations."""
    W1, b1, W2, b2, W3, b3 = net
    if rounded:
        n32 = np.asarray(norms, dtype=np.float32)
        xs = bf16((np.asarray(X, dtype=np.float32) - n32[:IN]) * n32[IN:2 * IN]).astype(np.float64)
    else:
        nrm = np.asarray(norms, dtype=np.float64)
        xs = (np.asarray(X, dtype=np.float64) - nrm[:IN]) * nrm[IN:2 * IN]
    z1 = xs @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    if rounded:
        h1 = bf16(h1.astype(np.float32)).astype(np.float64)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    o = h2 @ W3 + b3
    return xs, z1, h1, h2, o


def mlp_forward(params, norms, X, rounded=True):
    """y = net((x - mu) * is) * s_y + mu_y + c * x[0]."""
    nrm = np.asarray(norms, dtype=np.float64)
    o = _layers(unpack(params, rounded), norms, X, rounded)[-1]
    return o * nrm[2 * IN + 1] + nrm[2 * IN] + nrm[OFF_C] * np.asarray(X, dtype=np.float64)[:, 0]


def mlp_loss_grad(params, norms, X, Y, rounded=True):
    """L = mean_i (net(x_i) - t_i)^2 with t_i = (y_i - mu_y - c x_i0)/s_y, and
    dL/dparams (with respect to the weights as used; the BF16 rounding of the
    activations passes gradients straight through)."""
    net = unpack(params, rounded)
    W1, b1, W2, b2, W3, b3 = net
    nrm = np.asarray(norms, dtype=np.float64)
    xs, z1, h1, h2, o = _layers(net, norms, X, rounded)
    t = (np.asarray(Y, dtype=np.float64) - nrm[2 * IN] - nrm[OFF_C] * np.asarray(X, dtype=np.float64)[:, 0]) / nrm[
        2 * IN + 1]
    B = xs.shape[0]
    r = o - t
    loss = float(np.mean(r * r))
    d3 = 2.0 * r / B                       # dL/do
    d2 = np.outer(d3, W3) * (h2 > 0)       # dL/dz2
    d1 = (d2 @ W2) * (z1 > 0)              # dL/dz1
    g = np.zeros(NPARAMS)
    g[OFF_W1:OFF_B1] = (d1.T @ xs).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W3:OFF_B3] = d3 @ h2
    g[OFF_B3] = d3.sum()
    return loss, g


def adam(params, m, v, grad, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, algorithm 1); returns (params, m, v)."""
    m = beta1 * m + (1.0 - beta1) * grad
    v = beta2 * v + (1.0 - beta2) * grad * grad
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
