"""Goal-oriented adaptation with nested-adjoint DWR and with a learned indicator.

The adapt loop follows Solver::solve / Solver::adapt (src/main_nested.cpp:
90-140): on each cycle the primal problem is solved, element indicators are
computed, a size field is built from them (get_iso_target_size,
src/goal_size_field.cpp:39-151) and the mesh is adapted toward it.  Two
sources of indicators:

* nested_dwr -- NestedAdjoint::run (src/goal_nested_adjoint.cpp:163-257):
  uniform refinement, adjoint solve on the refined mesh, PU-localized
  residual, per-base-element sums.
* learned_eta -- no counterpart in the reference: 12 features per element
  from the primal state and the BF16 network of goal_mlp_*, which predicts
  log10(eta / E).  When more than `ood_limit` of the elements fall outside the
  training feature box it falls back to nested_dwr.

Problems are the box configurations of the assembly workloads with one load
step: the unit cube clamped at z = 0 and displaced at z = 1 (neo-Hookean), and
the [0,8]x[0,1]^2 bar clamped at x = 0 and displaced at x = 8 (J2).  The QoI
is AvgVM (src/goal_avg_vm.cpp:32-61) over the elements with centroid
x < 0.25.  All compute runs on the device through the C ABI; this module only
sequences it.
"""
from __future__ import annotations

import json
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

import goal_amd as g

# weights trained on the device from nested-DWR indicators of 260 seeded runs
DEFAULT_MODEL = Path(__file__).resolve().parent / "data" / "learned_avgvm.npz"
QOI_X = 0.25
QOI_SET = 1
AVG_VM = 1
# the network predicts log10(eta / E) as net(...) + H_EXPONENT * log10(h): the
# P1 indicator of a smooth solution scales like h^(2p+d) = h^5
H_EXPONENT = 5.0
# features whose training range bounds the physics of a prediction (the
# geometric ones 0-4 are left unbounded so finer meshes stay in range)
BOXED_FEATURES = (6, 7, 8, 9, 10, 11)


@dataclass
class Problem:
    geometry: str          # "cube" (neo-Hookean) or "bar" (J2)
    E: float
    nu: float
    K: float = 0.0
    Y: float = 0.0
    load: tuple = (0.0, 0.0, 0.01)
    n: int = 4             # hexahedra per unit length

    @property
    def model(self) -> int:
        return 2 if self.geometry == "bar" else 1

    @property
    def lengths(self):
        return (8.0, 1.0, 1.0) if self.geometry == "bar" else (1.0, 1.0, 1.0)

    @property
    def divisions(self):
        return (8 * self.n, self.n, self.n) if self.geometry == "bar" else (self.n,) * 3

    @property
    def faces(self):
        """(axis, fixed coordinate, loaded coordinate)"""
        return (0, 0.0, 8.0) if self.geometry == "bar" else (2, 0.0, 1.0)

    def material(self) -> g.MaterialSpec:
        return g.MaterialSpec(model=self.model, E=self.E, nu=self.nu, K=self.K, Y=self.Y, c0=1.0)


def sample_problem(rng: np.random.Generator, geometry: str, n: int) -> Problem:
    """Random material and load for one training or evaluation run.  n must
    be a multiple of 4 so that x = 0.25 is a plane of the mesh: the QoI
    element set is then the same region on every adapted mesh."""
    if n % 4:
        raise ValueError("sample_problem: n must be a multiple of 4")
    if geometry == "cube":
        E, nu = rng.uniform(500.0, 2000.0), rng.uniform(0.2, 0.35)
        load = (rng.uniform(-0.02, 0.02), rng.uniform(-0.02, 0.02), rng.choice([-1.0, 1.0]) * rng.uniform(0.01, 0.03))
        return Problem("cube", E, nu, load=tuple(float(v) for v in load), n=n)
    E, nu = rng.uniform(150.0, 250.0), rng.uniform(0.25, 0.35)
    K, Y = rng.uniform(1.0, 3.0), rng.uniform(0.2, 0.35)
    load = (rng.uniform(0.004, 0.016), rng.uniform(-0.03, 0.03), rng.uniform(-0.03, 0.03))
    return Problem("bar", E, nu, K=K, Y=Y, load=tuple(float(v) for v in load), n=n)


def base_mesh(p: Problem):
    return g.mesh_box(*p.divisions, *p.lengths, order=1, jitter=0.0)


class Model:
    """One linear mesh of a problem: disc, states, mechanics, sol info."""

    def __init__(self, p: Problem, coords, conn):
        self.p = p
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        es = (self.coords[self.conn].mean(axis=1)[:, 0] < QOI_X).astype(np.int32)
        self.disc = g.Disc.from_arrays(1, self.coords, self.conn, es)
        self.disc.set_elem_sets(2, es)
        axis, c0, c1 = p.faces
        self.disc.add_node_set("fixed", np.nonzero(self.coords[:, axis] == c0)[0])
        self.disc.add_node_set("loaded", np.nonzero(self.coords[:, axis] == c1)[0])
        self.bcs = [(0, "fixed", "0.0"), (1, "fixed", "0.0"), (2, "fixed", "0.0")] + [
            (k, "loaded", "%r*t" % p.load[k]) for k in range(3)]
        self.states = g.States(self.disc, p.model, 1)
        self.mats = [p.material(), p.material()]
        self.mech = g.Mechanics(self.disc, self.states, self.mats, stabilization=True, quad_order=1)
        self.si = g.SolInfo(self.disc)
        self.ne = self.conn.shape[0]
        self.newton_info = {}

    def newton(self, t=1.0, tol=1e-10, max_iters=30, lin_rtol=1e-8, lin_max=20000, restart=100, log=None,
               max_halvings=6) -> dict:
        """Primal::solve (src/goal_primal.cpp:111-137) with the device GMRES and
        one addition: a backtracking line search on ||R||.  goal takes every
        full Newton step; here a step that does not lower ||R|| by 1e-4 of its
        length is halved, up to max_halvings times (the last trial is kept).
        Full steps that lower ||R|| are taken unchanged, so converging solves
        are goal's iteration for iteration; on J2 bars refined hard at the
        clamp, full steps stall.  log, if given, is called after every
        iteration."""
        n = self.disc.num_dofs
        b, du = g.DeviceArray(n), g.DeviceArray(n)
        lin, it, alphas = [], 0, []
        t0 = time.perf_counter()

        def residual_norm():
            self.si.zero_R()
            self.mech.residual(self.si, save_states=True)
            for comp, ns, expr in self.bcs:
                g.dbc_apply(self.disc, self.si, comp, ns, expr, t, False)
            return self.si.norm_R()

        norm = residual_norm()  # the line search's reference; goal always takes a first step
        while it < max_iters:
            self.si.zero_all()
            self.mech.jacobian(self.si, adjoint=False, save_states=True)
            for comp, ns, expr in self.bcs:
                g.dbc_apply(self.disc, self.si, comp, ns, expr, t, True)
            R, _, _ = self.si.dev()
            g.check(g.lib.goal_device_scale(b.ptr, R, n, -1.0))
            du.fill(0.0)
            k, _ = self.si.solve(b, du, rtol=lin_rtol, max_iters=lin_max, restart=restart)
            lin.append(k)
            self.disc.add_soln_dev(du)            # x0 + du
            alpha, new = 1.0, residual_norm()
            for h in range(max_halvings):
                if new < (1.0 - 1e-4 * alpha) * norm:
                    break
                # x0 + alpha du -> x0 + alpha/2 du: add -alpha/2 du
                g.check(g.lib.goal_device_scale(du.ptr, du.ptr, n, -0.5 if h == 0 else 0.5))
                self.disc.add_soln_dev(du)
                alpha *= 0.5
                new = residual_norm()
            norm = new
            alphas.append(alpha)
            it += 1
            if log:
                log(f"  newton {it}: {k} GMRES iterations, " + (f"step {alpha:g}, " if alpha < 1.0 else "")
                    + f"||R|| = {norm:.3e} ({time.perf_counter() - t0:.1f} s)")
            if norm < tol:
                break
        g.sync()
        self.newton_info = {"iters": it, "linear_iters": lin, "norm": norm, "converged": bool(norm < tol),
                            "seconds": time.perf_counter() - t0, "steps": alphas}
        return self.newton_info

    def qoi(self) -> float:
        return self.mech.qoi(AVG_VM, None, elem_set=QOI_SET)

    def E(self) -> float:
        return self.p.E


def nested_dwr(m: Model, t=1.0, lin_rtol=1e-8, lin_max=40000, restart=100) -> dict:
    """NestedAdjoint::run (src/goal_nested_adjoint.cpp:163-257) on the device:
    returns eta per base element, the bound, z.R and the time taken."""
    t0 = time.perf_counter()
    nested = g.Nested(m.disc)
    fine = nested.disc
    nested.transfer_soln()
    st = g.States(fine, m.p.model, 1)
    mech = g.Mechanics(fine, st, m.mats, stabilization=True, quad_order=1)
    nested.transfer_states(m.states, st)
    si = g.SolInfo(fine)
    n = fine.num_dofs
    g.sync()
    t1 = time.perf_counter()
    si.zero_all()
    mech.jacobian(si, adjoint=True, save_states=False)
    mech.qoi(AVG_VM, si, elem_set=QOI_SET)
    for comp, ns, expr in m.bcs:
        g.dbc_apply(fine, si, comp, ns, expr, t, True)
    R, dMdu, _ = si.dev()
    z = g.DeviceArray(n)
    z.fill(0.0)
    g.sync()
    t2 = time.perf_counter()
    k, rr = si.solve(dMdu, z, rtol=lin_rtol, max_iters=lin_max, restart=restart)
    t3 = time.perf_counter()
    zR = si.dot_R(z.ptr.value)
    zc = g.DeviceArray(n)
    g.device_copy(zc, z, n)
    nested.set_coarse(zc)
    zd = g.DeviceArray(n)
    g.dwr_subtract(z, zc, zd)
    si.zero_R()
    mech.localize(si, zd.ptr.value, zc.ptr.value)
    for comp, ns, expr in m.bcs:
        g.dbc_apply(fine, si, comp, ns, expr, t, False)
    R, _, _ = si.dev()
    bound = g.dwr_sum_contribs(fine, R)
    eta_f = g.DeviceArray(fine.num_elems)
    g.dwr_compute_error(fine, R, eta_f)
    eta_b = g.DeviceArray(m.ne)
    nested.set_error(eta_f, eta_b)
    eta = eta_b.numpy()
    g.sync()
    t4 = time.perf_counter()
    return {"eta": eta, "bound": bound, "zR": zR, "adjoint_iters": k, "adjoint_relres": rr,
            "fine_elems": fine.num_elems, "seconds": t4 - t0,
            "split": {"refine_and_setup": t1 - t0, "adjoint_assembly": t2 - t1, "adjoint_solve": t3 - t2,
                      "localize": t4 - t3}}


@dataclass
class Learned:
    """A trained network and the normalization it was trained with."""
    mlp: g.MLP
    norms: np.ndarray
    ood_limit: float = 0.01
    history: list = field(default_factory=list)
    meta: dict = field(default_factory=dict)


def features(m: Model, buf: g.DeviceBuffer | None = None) -> g.DeviceBuffer:
    return g.learn_features(m.mech, QOI_X, buf)


def learned_eta(m: Model, net: Learned, fallback: bool = True) -> dict:
    """Indicators from the network; nested DWR when too many elements are out of range."""
    t0 = time.perf_counter()
    feat = features(m)
    y = g.DeviceBuffer(m.ne, np.float32)
    outside = net.mlp.forward(feat, m.ne, y)
    if fallback and outside > net.ood_limit * m.ne:
        out = nested_dwr(m)
        out.update(fallback=True, outside=outside)
        out["seconds"] = time.perf_counter() - t0
        return out
    eta = m.E() * np.power(10.0, y.numpy().astype(np.float64))
    g.sync()
    return {"eta": eta, "fallback": False, "outside": outside, "seconds": time.perf_counter() - t0}


def adapt(m: Model, eta, target: int, split_ratio: float = 1.5, max_passes: int = 8) -> Model:
    """Size field + refinement; the new model starts from the transferred solution."""
    if not np.all(np.isfinite(eta)):
        raise ValueError("adapt: non-finite indicators")
    size = g.size_field(m.coords, m.conn, eta, int(target))
    A = g.Adaptation(m.coords, m.conn, size, split_ratio, max_passes)
    new = Model(m.p, A.coords, A.conn)
    new.disc.set_soln(A.transfer(m.disc.soln().reshape(-1, 4)).ravel())
    return new


# ------------------------------------------------------------------ training

def make_dataset(num_runs: int, seed: int = 0, cube_cycles: int = 4, bar_cycles: int = 2, growth: float = 3.0,
                 cap: int = 1200, lin_rtol: float = 1e-6, checkpoint=None, geometries=("cube", "bar"), log=print):
    """Training samples from nested DWR on seeded cube and bar problems:
    features X [N][12], targets log10(eta / E) [N] and the mesh index of each
    sample.  Run r (geometries[r % len(geometries)]) draws its problem from a
    generator seeded with (seed, r) and adapts `cube_cycles` / `bar_cycles`
    times with the nested indicators, so every mesh of the run contributes; at
    most `cap` random elements of a mesh are kept.  The linear solves stop at
    `lin_rtol` (the indicators need a few digits, not ten).  With `checkpoint`
    (a path) the arrays so far are saved every 25 runs."""
    Xs, Ys, Gs, runs = [], [], [], []

    def arrays():
        return (np.concatenate(Xs).astype(np.float32), np.concatenate(Ys).astype(np.float32),
                np.concatenate(Gs), runs)

    skipped = []
    for r in range(num_runs):
        rng = np.random.default_rng([seed, r])
        geometry = geometries[r % len(geometries)]
        n = int(rng.choice([4, 8])) if geometry == "cube" else 4
        cycles = cube_cycles if geometry == "cube" else bar_cycles
        p = sample_problem(rng, geometry, n)
        first = len(runs)
        try:
            _run_into(p, r, cycles, growth, cap, lin_rtol, rng, Xs, Ys, Gs, runs)
        except (g.GoalError, ValueError) as e:   # a failed solve ends the run; its earlier meshes stay
            skipped.append(r)
            if log:
                log(f"[learned] run {r} stopped: {e}")
        for rec in runs[first:]:
            rec["seed"] = seed
        if log:
            last = runs[-1]["elems"] if len(runs) > first else 0
            log(f"[learned] dataset: {r + 1}/{num_runs} runs, {sum(len(y) for y in Ys)} samples "
                f"(run {r}: {len(runs) - first} meshes, up to {last} tets)")
        if checkpoint is not None and ((r + 1) % 25 == 0 or r + 1 == num_runs):
            X_, Y_, G_, _ = arrays()
            np.savez(checkpoint, X=X_, Y=Y_, G=G_, runs=np.frombuffer(json.dumps(runs).encode(), dtype=np.uint8))
    return arrays()


def _run_into(p, r, cycles, growth, cap, lin_rtol, rng, Xs, Ys, Gs, runs):
    """One training run: solve, nested DWR, adapt, `cycles` times."""
    m = Model(p, *base_mesh(p))
    info = m.newton(lin_rtol=lin_rtol)
    for c in range(cycles + 1):
        if not info["converged"]:
            raise ValueError(f"Newton did not converge on {m.ne} tets (|R| = {info['norm']:.3e})")
        d = nested_dwr(m, lin_rtol=lin_rtol)
        if not np.all(np.isfinite(d["eta"])):
            raise ValueError(f"non-finite nested indicators on {m.ne} tets")
        X = features(m).numpy(m.ne * g.MLP_IN).reshape(m.ne, g.MLP_IN)
        keep = np.nonzero(d["eta"] > 0)[0]
        if keep.size > cap:
            keep = np.sort(rng.choice(keep, cap, replace=False))
        Xs.append(X[keep])
        Ys.append(np.log10(d["eta"][keep] / p.E))
        Gs.append(np.full(keep.size, len(runs), dtype=np.int32))
        runs.append({"run": r, "cycle": c, "geometry": p.geometry, "elems": m.ne, "newton": info["iters"],
                     "converged": info["converged"], "eta_sum": float(d["eta"].sum()), "zR": d["zR"],
                     "nested_seconds": round(d["seconds"], 4)})
        if c < cycles:
            m = adapt(m, d["eta"], int(m.ne * growth))
            info = m.newton(lin_rtol=lin_rtol)


def fit_norms(X: np.ndarray, Y: np.ndarray, margin: float = 0.05) -> np.ndarray:
    norms = np.zeros(g.MLP_NORMS, dtype=np.float32)
    mu = X.mean(axis=0)
    sd = X.std(axis=0)
    sd[sd < 1e-6] = 1.0
    norms[:12] = mu
    norms[12:24] = 1.0 / sd
    norms[50] = H_EXPONENT
    r = Y - H_EXPONENT * X[:, 0]
    norms[24] = r.mean()
    norms[25] = max(float(r.std()), 1e-6)
    lo, hi = X.min(axis=0), X.max(axis=0)
    pad = margin * (hi - lo) + 1e-6
    norms[26:38] = -np.inf
    norms[38:50] = np.inf
    for k in BOXED_FEATURES:
        norms[26 + k] = lo[k] - pad[k]
        norms[38 + k] = hi[k] + pad[k]
    return norms


def train(X: np.ndarray, Y: np.ndarray, epochs: int = 60, batch: int = 4096, lr: float = 2e-3, seed: int = 0,
          val_frac: float = 0.1, groups: np.ndarray | None = None, log=print) -> tuple[Learned, dict]:
    """Adam on the device (goal_mlp_train) over shuffled minibatches.  With
    `groups` (one id per mesh) the output shift is then calibrated so that the
    predicted indicator sum of a mesh matches the nested one in the median:
    a least-squares fit in log space underestimates sums of exponentials."""
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    perm = rng.permutation(n)
    nval = int(val_frac * n)
    val, tr = perm[:nval], perm[nval:]
    norms = fit_norms(X[tr], Y[tr])
    mlp = g.MLP()
    mlp.set(he_init(rng), norms)
    xd = g.DeviceBuffer.from_host(X.ravel(), np.float32)
    yd = g.DeviceBuffer.from_host(Y, np.float32)
    steps = max(1, len(tr) // batch)
    hist = []
    idx = g.DeviceBuffer(steps * batch, np.int32)
    t0 = time.perf_counter()
    for ep in range(epochs):
        sched = lr * 0.5 * (1.0 + np.cos(np.pi * ep / epochs))
        idx.upload(rng.permutation(tr)[: steps * batch].astype(np.int32))
        loss = mlp.train(xd, yd, idx, batch, steps, sched)
        hist.append(loss)
    g.sync()
    seconds = time.perf_counter() - t0
    net = Learned(mlp, norms, history=hist)
    stats = {"train_loss": hist[-1], "epochs": epochs, "steps": epochs * steps, "batch": batch,
             "train_seconds": seconds, "samples": len(tr)}
    if groups is not None:
        yp = g.DeviceBuffer(n, np.float32)
        mlp.forward(xd, n, yp)
        pred = yp.numpy().astype(np.float64)
        trm = np.zeros(n, dtype=bool)
        trm[tr] = True
        ratios = []
        for gid in np.unique(groups[trm]):
            sel = trm & (groups == gid)
            ratios.append(np.log10(np.sum(10.0 ** Y[sel].astype(np.float64)) / np.sum(10.0 ** pred[sel])))
        shift = float(np.median(ratios))
        norms = norms.copy()
        norms[24] += shift
        mlp.set(None, norms)
        net.norms = norms
        stats["calibration_log10"] = shift
    if nval:
        vi = g.DeviceBuffer.from_host(val.astype(np.int32), np.int32)
        vloss, _ = mlp.gradient(xd, yd, vi, nval)
        stats["val_loss"] = vloss
        xv = g.DeviceBuffer.from_host(X[val].ravel(), np.float32)
        yv = g.DeviceBuffer(nval, np.float32)
        mlp.forward(xv, nval, yv)
        err = yv.numpy() - Y[val]
        stats["val_rmse_log10"] = float(np.sqrt(np.mean(err ** 2)))
        stats["val_base_rmse_log10"] = float(np.std(Y[val] - H_EXPONENT * X[val, 0]))
    if log:
        log(f"[learned] trained {epochs} epochs x {steps} steps in {seconds:.2f} s: " +
            ", ".join(f"{k} {v:.4g}" for k, v in stats.items() if isinstance(v, float)))
    return net, stats


def save(net: Learned, path, meta: dict | None = None) -> None:
    import json
    params, _, norms = net.mlp.get()
    np.savez(path, params=params, norms=norms, meta=np.frombuffer(json.dumps(meta or {}).encode(), dtype=np.uint8))


def load(path) -> Learned:
    """A network saved by save(); numeric arrays only (no pickles)."""
    import json
    with np.load(path, allow_pickle=False) as f:
        params, norms, meta = f["params"], f["norms"], f["meta"]
    mlp = g.MLP()
    mlp.set(params, norms)
    net = Learned(mlp, norms)
    net.meta = json.loads(bytes(meta).decode())
    return net


def he_init(rng: np.random.Generator) -> np.ndarray:
    p = np.zeros(g.MLP_PARAMS, dtype=np.float32)
    h, i = g.MLP_HIDDEN, g.MLP_IN
    o = 0
    p[o:o + h * i] = rng.uniform(-1, 1, h * i) * np.sqrt(6.0 / i)
    o += h * i + h
    p[o:o + h * h] = rng.uniform(-1, 1, h * h) * np.sqrt(6.0 / h)
    o += h * h + h
    p[o:o + h] = rng.uniform(-1, 1, h) * np.sqrt(6.0 / h) * 0.1
    return p


# ---------------------------------------------------------------- evaluation

def run_cycles(p: Problem, method: str, cycles: int, final_elems: int, net: Learned | None = None,
               log=print):
    """The adapt loop with one indicator source ("nested", "learned" or
    "uniform": every element equal, the method without indicators).  The
    target of cycle c is ne0 * growth^(c+1) with growth = (final/ne0)^(1/cycles).
    Returns the per-cycle records and the last model."""
    m = Model(p, *base_mesh(p))
    ne0 = m.ne
    growth = (final_elems / ne0) ** (1.0 / cycles)
    out = []
    for c in range(cycles + 1):
        info = m.newton(log=log)
        J = m.qoi()
        rec = {"cycle": c, "elems": m.ne, "J": J, "newton": info["iters"], "converged": info["converged"],
               "newton_seconds": round(info["seconds"], 4), "linear_iters": info["linear_iters"],
               "damped_steps": sum(1 for a in info["steps"] if a < 1.0)}
        if c < cycles:
            if method == "nested":
                d = nested_dwr(m)
                rec.update(zR=d["zR"], bound=d["bound"], adjoint_iters=d["adjoint_iters"],
                           nested_split={k: round(v, 4) for k, v in d["split"].items()})
                if net is not None:   # the network on the same mesh, for a like-for-like comparison
                    l = learned_eta(m, net, fallback=False)
                    rec.update(learned_eta_sum=float(np.sum(l["eta"])), learned_seconds=l["seconds"],
                               outside=int(l["outside"]), spearman=_rank_corr(d["eta"], l["eta"]))
            elif method == "learned":
                d = learned_eta(m, net)
                rec.update(outside=int(d["outside"]))
            else:
                d = {"eta": np.ones(m.ne), "seconds": 0.0}
            rec.update(estimate_seconds=d["seconds"], eta_sum=float(np.sum(d["eta"])),
                       fallback=bool(d.get("fallback", False)))
            t0 = time.perf_counter()
            m = adapt(m, d["eta"], int(round(ne0 * growth ** (c + 1))))
            rec["adapt_seconds"] = round(time.perf_counter() - t0, 3)
        out.append(rec)
        if log:
            log(f"[learned] {method} cycle {c}: {rec['elems']} tets, J {J:.10g}, newton {info['iters']} "
                f"({info['seconds']:.2f} s)" +
                (f", estimate {rec['estimate_seconds'] * 1e3:.2f} ms" if "estimate_seconds" in rec else ""))
    return out, m


def _rank_corr(a, b) -> float:
    ra = np.argsort(np.argsort(a)).astype(np.float64)
    rb = np.argsort(np.argsort(b)).astype(np.float64)
    return float(np.corrcoef(ra, rb)[0, 1])


def overkill_qoi(m: Model, log=print) -> dict:
    """The QoI on the uniform 1:8 refinement of m's mesh, solved from the
    interpolated solution: the reference value that effectivities and QoI
    errors are measured against."""
    fc, fconn, _, ep = g.mesh_refine(m.coords, m.conn)
    x = m.disc.soln().reshape(-1, 4)
    xf = np.concatenate([x, 0.5 * (x[ep[:, 0]] + x[ep[:, 1]])])
    del ep
    ref = Model(m.p, fc, fconn)
    ref.disc.set_soln(xf.ravel())
    info = ref.newton(log=log)
    J = ref.qoi()
    if log:
        log(f"[learned] overkill: {ref.ne} tets, J {J:.10g}, newton {info['iters']} ({info['seconds']:.1f} s), "
            f"converged {info['converged']}")
    return {"J": J, "elems": ref.ne, "newton": info["iters"], "converged": info["converged"],
            "seconds": info["seconds"]}
