"""The C++ surface (include/goal_amd.hpp) driving Newton load steps, the
functional and a nested-adjoint estimate on the GPU, against the same drivers
run on the oracle (tests/drivers.py).  Bar: identical Newton iteration counts
per load step, final ||R|| < 1e-10 on both, QoI and DWR quantities within 1e-9
relative (both sides converge to 1e-10 in ||R||, so iterates differ by
roughly the solver tolerance)."""
import json
import subprocess
from pathlib import Path

import pytest

from tests import drivers

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "cpp" / "surface_test"


def run_surface(order, quad, model, n, steps):
    if not EXE.exists():
        pytest.fail(f"{EXE} is not built: run __graft_entry__.build()")
    r = subprocess.run([str(EXE), str(order), str(quad), str(model), str(n), str(steps)], capture_output=True,
                       text=True, timeout=300)
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
