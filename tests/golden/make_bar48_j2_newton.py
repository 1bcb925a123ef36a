"""Writes tests/golden/bar48_j2_newton.json: the oracle's Newton iteration counts
and average-displacement QoI per load step for a reduced J2 tension bar -- the
bar workload's geometry and material at 48 x 6 x 6 hexahedra x 6 Kuhn tets
(10,368 P1 tets): [0,8]x[0,1]^2, jitter 10% h (seed 11), J2 E=200 nu=0.3 Y=0.25
K=2 c0=1 with stabilization, 1-point quadrature, rollers on xmin/ymin/zmin and
u_x = 0.004 t on xmax for t = 1..20 (1% strain), Newton on ||R||_2 < 1e-10.
Run from the repository root: python tests/golden/make_bar48_j2_newton.py"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests import drivers  # noqa: E402

CFG = dict(problem="bar", order=1, quad=1, model=2, n=[48, 6, 6], steps=20, jitter=0.1, seed=11)


def main():
    t0 = time.time()
    out = drivers.newton_steps(CFG["order"], CFG["quad"], CFG["model"], CFG["n"], CFG["steps"],
                               jitter=CFG["jitter"], seed=CFG["seed"], dwr=False, threads=os.cpu_count() or 1,
                               problem=CFG["problem"])
    out["config"] = CFG
    out["norm"] = float(out["norm"])
    with open(os.path.join(ROOT, "tests", "golden", "bar48_j2_newton.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out["iters"], out["J"], f"{time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
