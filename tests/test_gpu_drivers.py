"""The C++ surface (include/goal_amd.hpp) driving Newton load steps, the
functional and a nested-adjoint estimate on the GPU, against the same drivers
run on the oracle (tests/drivers.py).  Bar: identical Newton iteration counts
per load step, final ||R|| < 1e-10 on both, QoI and DWR quantities within 1e-9
relative (both sides converge to 1e-10 in ||R||, so iterates differ by
roughly the solver tolerance)."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests import drivers

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "cpp" / "surface_test"


def run_surface(order, quad, model, n, steps, tangent=64, solver="lu", extra=None):
    if not EXE.exists():
        pytest.fail(f"{EXE} is not built: run __graft_entry__.build()")
    args = [str(EXE), str(order), str(quad), str(model), str(n), str(steps), str(tangent), solver]
    r = subprocess.run(args + ([extra] if extra else []), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


@pytest.mark.parametrize("order,quad,model,n,steps", [
    (1, 1, 1, 3, 3),   # reference parity mode: P1, one point, neo-Hookean
    (1, 1, 2, 3, 4),   # J2 (plastic after the first steps)
    (2, 2, 1, 2, 2),   # P2 with four points
])
def test_newton_functional_and_dwr(order, quad, model, n, steps):
    gpu = run_surface(order, quad, model, n, steps)
    ref = drivers.newton_steps(order, quad, model, n, steps)
    assert gpu["iters"] == ref["iters"]
    assert gpu["norm"] < 1e-10 and ref["norm"] < 1e-10
    assert abs(gpu["J"] - ref["J"]) <= 1e-9 * abs(ref["J"])
    if order == 1:
        for k in ("eta_sum", "bound", "zR"):
            assert abs(gpu[k] - ref[k]) <= 1e-8 * abs(ref[k]), k


@pytest.mark.parametrize("order,quad,model,n,steps", [
    (2, 2, 1, 2, 3),   # P2, four points, neo-Hookean
    (1, 1, 2, 3, 4),   # J2
])
def test_fp32_tangent_newton(order, quad, model, n, steps):
    """FP32 element matrices, FP64 residual: Newton still reaches ||R|| < 1e-10
    with at most one extra iteration per load step, and the converged
    functional and DWR estimate match the FP64 path (the adjoint system itself
    is FP32-rounded, so its quantities agree to ~1e-6)."""
    f64 = run_surface(order, quad, model, n, steps)
    f32 = run_surface(order, quad, model, n, steps, tangent=32)
    assert f32["norm"] < 1e-10
    assert all(b - a in (0, 1) for a, b in zip(f64["iters"], f32["iters"])), (f64["iters"], f32["iters"])
    assert abs(f32["J"] - f64["J"]) <= 1e-9 * abs(f64["J"])
    if order == 1:
        for k in ("eta_sum", "bound"):
            assert abs(f32[k] - f64[k]) <= 1e-5 * abs(f64[k]), k
        # z.R sums products of opposite sign (J2 case: 4e-7 out of terms ~1e-3),
        # which amplifies the adjoint's FP32 rounding: 1e-4 relative
        assert abs(f32["zR"] - f64["zR"]) <= 1e-4 * abs(f64["zR"])


GOLDEN = Path(__file__).resolve().parent / "golden"


def test_cube8_p2_newton_matches_oracle():
    """The CPU-runnable workload end to end: 8^3 hexahedra -> 3,072 P2 tets,
    neo-Hookean, 4-point quadrature, ten load steps to u_z = 0.1 on z = 1,
    Newton to ||R|| < 1e-10 with the device GMRES.  Same iteration count per
    load step as the oracle with a direct solve (tests/golden/cube8_p2_newton.json),
    the average displacement within 1e-9 relative at every step."""
    ref = json.loads((GOLDEN / "cube8_p2_newton.json").read_text())
    c = ref["config"]
    gpu = drivers.gpu_newton_steps(c["order"], c["quad"], c["model"], c["n"], c["steps"], jitter=c["jitter"],
                                   seed=c["seed"])
    print("newton iterations", gpu["iters"], "linear iterations", gpu["linear_iters"])
    assert gpu["iters"] == ref["iters"]
    assert gpu["norm"] < 1e-10
    for a, b in zip(gpu["J_steps"], ref["J_steps"]):
        assert abs(a - b) <= 1e-9 * abs(b)


def test_bar48_j2_tension_matches_oracle():
    """The J2 tension bar's load path at reduced size: 48 x 6 x 6 hexahedra ->
    10,368 P1 tets on [0,8]x[0,1]^2, J2 with linear hardening and per-point
    history, rollers on xmin/ymin/zmin, u_x = 0.004 t on xmax for 20 load
    steps (1% strain), Newton to ||R|| < 1e-10 with the device GMRES and its
    multigrid preconditioner.  Same iteration count per load step as the
    oracle with a direct solve (tests/golden/bar48_j2_newton.json), the
    average displacement within 1e-9 relative at every step."""
    ref = json.loads((GOLDEN / "bar48_j2_newton.json").read_text())
    c = ref["config"]
    gpu = drivers.gpu_newton_steps(c["order"], c["quad"], c["model"], tuple(c["n"]), c["steps"],
                                   jitter=c["jitter"], seed=c["seed"], problem=c["problem"])
    print("newton iterations", gpu["iters"], "linear iterations", gpu["linear_iters"])
    assert gpu["iters"] == ref["iters"]
    assert gpu["norm"] < 1e-10
    for a, b in zip(gpu["J_steps"], ref["J_steps"]):
        assert abs(a - b) <= 1e-9 * abs(b)


@pytest.mark.parametrize("order,quad,model,n,steps", [(1, 1, 1, 3, 3), (2, 2, 1, 2, 2)])
def test_cpp_surface_with_device_gmres(order, quad, model, n, steps):
    """goal::GmresSolve as the LinearSolve of Primal and NestedAdjoint: the
    same Newton iteration counts and converged quantities as the dense LU."""
    lu = run_surface(order, quad, model, n, steps)
    gm = run_surface(order, quad, model, n, steps, solver="gmres")
    assert gm["iters"] == lu["iters"]
    assert gm["norm"] < 1e-10
    assert abs(gm["J"] - lu["J"]) <= 1e-9 * abs(lu["J"])
    if order == 1:
        for k in ("eta_sum", "bound", "zR"):
            assert abs(gm[k] - lu[k]) <= 1e-7 * abs(lu[k]), k


def test_cpp_surface_adapt_and_learned_indicator(tmp_path):
    """The adapt step and the learned indicator through the C++ surface
    (goal::get_iso_target_size, goal::adapt, goal::LearnedIndicator): the
    adapted problem solves, and the learned indicators equal the Python
    binding's on the same state."""
    import numpy as np

    import goal_amd as g
    from goal_amd import learned as LP

    with np.load(LP.DEFAULT_MODEL, allow_pickle=False) as f:
        params, norms = f["params"].astype(np.float32), f["norms"].astype(np.float32)
    net_file = tmp_path / "net.f32"
    np.concatenate([params, norms]).tofile(net_file)
    n = 4
    out = subprocess.run([str(EXE), "1", "1", "1", str(n), "1", "64", "gmres", f"learned:{net_file}"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    cpp = json.loads(out.stdout)
    ada = run_surface(1, 1, 1, n, 1, solver="gmres", extra="adapt")
    print("learned", cpp.get("learned_eta_sum"), "nested", cpp["eta_sum"], "adapt", ada["adapt"], "J", ada["J"])
    base = 6 * n ** 3
    assert 1.5 * base < ada["adapt"]["elems"] < 4 * base
    assert ada["adapt"]["norm"] < 1e-10
    # the same state in Python: box, material, load step of the driver
    disc = g.Disc.box(n, n, n, order=1, jitter=0.1, seed=11)
    st = g.States(disc, 1, 1)
    mech = g.Mechanics(disc, st, [g.MaterialSpec(model=1, E=1000.0, nu=0.25, c0=1.0)], stabilization=True)
    si = g.SolInfo(disc)
    bcs = [(0, "zmin", "0.0"), (1, "zmin", "0.0"), (2, "zmin", "0.0"), (2, "zmax", "0.01*t")]
    b, du = g.DeviceArray(disc.num_dofs), g.DeviceArray(disc.num_dofs)
    for _ in range(10):
        si.zero_all()
        mech.jacobian(si, adjoint=False, save_states=True)
        for comp, ns, expr in bcs:
            g.dbc_apply(disc, si, comp, ns, expr, 1.0, True)
        R, _, _ = si.dev()
        g.check(g.lib.goal_device_scale(b.ptr, R, disc.num_dofs, -1.0))
        du.fill(0.0)
        si.solve(b, du, rtol=1e-12, max_iters=5000, restart=100)
        disc.add_soln_dev(du)
        si.zero_R()
        mech.residual(si, save_states=True)
        for comp, ns, expr in bcs:
            g.dbc_apply(disc, si, comp, ns, expr, 1.0, False)
        if si.norm_R() < 1e-10:
            break
    net = LP.load(LP.DEFAULT_MODEL)
    feat = g.learn_features(mech, LP.QOI_X)
    y = g.DeviceBuffer(disc.num_elems, np.float32)
    outside = net.mlp.forward(feat, disc.num_elems, y)
    py = float(np.sum(1000.0 * np.power(10.0, y.numpy().astype(np.float64))))
    if cpp["learned_fell_back"]:
        assert outside > 0.01 * disc.num_elems
        assert abs(cpp["learned_eta_sum"] - cpp["eta_sum"]) <= 1e-9 * cpp["eta_sum"]
    else:
        assert abs(cpp["learned_eta_sum"] - py) <= 1e-5 * py
