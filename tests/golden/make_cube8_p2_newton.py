"""Writes tests/golden/cube8_p2_newton.json: the oracle's Newton iteration counts
and average-displacement QoI per load step for the CPU-runnable workload -- unit
cube, 8^3 hexahedra x 6 Kuhn tets (3,072 P2 tets), no jitter, neo-Hookean
E=1000 nu=0.25 c0=1 with stabilization, 4-point quadrature, z=0 clamped, u_z =
0.01 t on z=1 for t = 1..10, Newton on ||R||_2 < 1e-10.  Run from the
repository root: python tests/golden/make_cube8_p2_newton.py"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests import drivers  # noqa: E402

CFG = dict(order=2, quad=2, model=1, n=8, steps=10, jitter=0.0, seed=0)


def main():
    t0 = time.time()
    out = drivers.newton_steps(CFG["order"], CFG["quad"], CFG["model"], CFG["n"], CFG["steps"],
                               jitter=CFG["jitter"], seed=CFG["seed"], dwr=False, threads=os.cpu_count() or 1)
    out["config"] = CFG
    out["norm"] = float(out["norm"])
    with open(os.path.join(ROOT, "tests", "golden", "cube8_p2_newton.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out["iters"], out["J"], f"{time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
