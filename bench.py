#!/usr/bin/env python3
"""Benchmark: P2-tet residual + Jacobian assembly throughput on one MI355X.

Workload (default): unit cube, 64^3 hexahedra split into 6 Kuhn tets each
(1,572,864 P2 tets, 2,146,689 nodes, 8,586,756 DOFs), vertex positions
jittered +-10% h (seed 0x5EED), compressible neo-Hookean E=1000 nu=0.25 c0=1
with stabilization, 4-point quadrature, FP64 end to end.  One step = one
Jacobian-mode assembly pass as in Primal::compute_jacob: zero R and dRdu,
then the fused residual + Jacobian kernel (it also writes the stress state).
The input iterate is load step 1 of the cube problem (zmin held, u_z = 0.01 on
zmax), solved from u = 0 by Newton on the device before timing starts (about
1.5 minutes, untimed; --iterate synthetic uses a smooth synthetic field
instead).  The neo-Hookean element work does not depend on the iterate: both
inputs time the same to 0.5%.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import sys
import time
from pathlib import Path

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_ELEM_P2 = 13576  # algorithmic bytes per P2/P2 element, residual + Jacobian pass
B_ELEM_P1 = 2416
NCOPY = 1 << 28  # doubles per buffer of the in-run copy kernel (2 GiB read + 2 GiB written)
ASSEMBLY_KERNELS = ("j2_point_kernel", "element_kernel", "gather_kernel")  # the point pass runs for J2 only
PROFILES = os.path.join(ROOT, "profiles")


NH_MAT = dict(E=1000.0, nu=0.25, K=0.0, Y=0.0, c0=1.0)
J2_MAT = dict(E=200.0, nu=0.3, K=2.0, Y=0.25, c0=1.0)
B_STATE_J2 = 232  # J2 history per stored point: read Fp_old + eqps_old, write Fp + eqps + sigma


def workload(name: str, n: int, order: int):
    """Mesh, material and iterate of a benchmark workload."""
    if name == "cube":
        return dict(name=name, dims=(n, n, n), lengths=(1.0, 1.0, 1.0), order=order, quad=2, model=1, mat=NH_MAT,
                    b_elem=B_ELEM_P2 if order == 2 else B_ELEM_P1, npts=4,
                    metric=f"P{order}-tet resid+Jac assembled elems/s",
                    desc=f"cube {n}^3 hex, jitter 10% h, neo-Hookean E=1000 nu=0.25 c0=1, 4-pt quadrature")
    if name == "bar":
        return dict(name=name, dims=(384, 48, 48), lengths=(8.0, 1.0, 1.0), order=1, quad=1, model=2, mat=J2_MAT,
                    b_elem=B_ELEM_P1 + B_STATE_J2, npts=1, metric="P1-tet J2 resid+Jac assembled elems/s",
                    desc="bar [0,8]x[0,1]^2 as 384x48x48 hex, jitter 10% h, J2 E=200 nu=0.3 Y=0.25 K=2 c0=1, "
                         "1-pt quadrature, FP64 per-point history, 1% axial strain (plastic)")
    raise ValueError(name)


def profile_tag(n: int, order: int, tangent: int = 64, name: str = "cube") -> str:
    base = f"assembly_p{order}_nh_cube{n}" if name == "cube" else "assembly_p1_j2_bar384"
    return base + ("_f32tangent" if tangent == 32 else "")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bar_state(coords):
    """Uniaxial stretch of 1% along x (yield strain 1.25e-3), a smooth transverse
    perturbation, and the matching bulk pressure."""
    import numpy as np

    X, Y, Z = coords[:, 0], coords[:, 1], coords[:, 2]
    x = np.zeros(4 * coords.shape[0])
    x[0::4] = 0.01 * X
    x[1::4] = -0.003 * Y + 1e-3 * np.sin(np.pi * X / 8) * Z
    x[2::4] = -0.003 * Z + 1e-3 * np.cos(np.pi * X / 8) * Y
    x[3::4] = 0.5
    return x


def newton_iterate(g, disc, mech, si, t=1.0, tol=1e-10, max_newton=4, lin_rtol=1e-8, lin_max=4000):
    """The cube's load step 1 from u = 0 (zmin held, u_z = 0.01 t on zmax, the
    small-cube problem's boundary conditions), solved by Newton with the
    library's device assembly, Dirichlet rows and GMRES, so that the timed
    assemblies read a real Newton iterate.  Returns what it did."""
    import numpy as np

    n = disc.num_dofs
    disc.set_soln(np.zeros(n))
    bcs = [(0, "zmin", "0.0"), (1, "zmin", "0.0"), (2, "zmin", "0.0"), (2, "zmax", "0.01*t")]
    b, du = g.DeviceArray(n), g.DeviceArray(n)
    t0 = time.perf_counter()
    iters, lin, norms = 0, [], []
    while iters < max_newton:
        iters += 1
        si.zero_all()
        mech.jacobian(si, adjoint=False, save_states=True)
        for comp, ns, expr in bcs:
            g.dbc_apply(disc, si, comp, ns, expr, t, True)
        g.check(g.lib.goal_memcpy_h2d(b.ptr, g._ptr(np.ascontiguousarray(-si.R())), 8 * n))
        du.fill(0.0)
        k, _ = si.solve(b, du, rtol=lin_rtol, max_iters=lin_max, restart=100)
        lin.append(k)
        disc.add_soln_dev(du)
        si.zero_R()
        mech.residual(si, save_states=True)
        for comp, ns, expr in bcs:
            g.dbc_apply(disc, si, comp, ns, expr, t, False)
        norms.append(si.norm_R())
        log(f"newton {iters}: {k} GMRES iterations, ||R|| = {norms[-1]:.3e} ({time.perf_counter() - t0:.1f} s)")
        if norms[-1] < tol:
            break
    return {"kind": "newton", "load_step": 1, "newton_iters": iters, "gmres_iters": lin,
            "norms": [float(f"{v:.4g}") for v in norms], "seconds": round(time.perf_counter() - t0, 1)}


BAR_BCS = [(0, "xmin", "0.0"), (1, "ymin", "0.0"), (2, "zmin", "0.0"), (0, "xmax", "0.004*t")]


def bar_load_path(g, disc, mech, si, st, steps=20, tol=1e-10, max_newton=8, lin_rtol=1e-8, lin_max=4000):
    """The bar workload's tension load path: rollers on xmin/ymin/zmin, u_x =
    0.004 t on xmax for t = 1..steps (1% strain at t = 20), Newton on every
    step with the library's device assembly, Dirichlet rows and multigrid
    GMRES, the history updated between steps (States::update).  The timed
    assemblies then read the last step's plastic state."""
    import numpy as np

    n = disc.num_dofs
    disc.set_soln(np.zeros(n))
    b, du = g.DeviceArray(n), g.DeviceArray(n)
    t0 = time.perf_counter()
    per_step = []
    for step in range(1, steps + 1):
        t = float(step)
        its, lin, norm = 0, [], None
        while its < max_newton:
            its += 1
            si.zero_all()
            mech.jacobian(si, adjoint=False, save_states=True)
            for comp, ns, expr in BAR_BCS:
                g.dbc_apply(disc, si, comp, ns, expr, t, True)
            R, _, _ = si.dev()
            g.check(g.lib.goal_device_scale(b.ptr, R, n, -1.0))
            du.fill(0.0)
            k, _ = si.solve(b, du, rtol=lin_rtol, max_iters=lin_max, restart=100)
            lin.append(k)
            disc.add_soln_dev(du)
            si.zero_R()
            mech.residual(si, save_states=True)
            for comp, ns, expr in BAR_BCS:
                g.dbc_apply(disc, si, comp, ns, expr, t, False)
            norm = si.norm_R()
            if norm < tol:
                break
        eq = st.get("eqps")
        per_step.append({"t": step, "newton_iters": its, "gmres_iters": lin, "norm": float(f"{norm:.4g}"),
                         "plastic_points": int((eq > 0).sum())})
        log(f"load step {step}: {its} Newton iterations, GMRES {lin}, ||R|| = {norm:.3e}, "
            f"{per_step[-1]['plastic_points']} plastic points ({time.perf_counter() - t0:.1f} s)")
        if norm >= tol:
            raise RuntimeError(f"bar load step {step}: Newton did not converge (||R|| = {norm:.3e})")
        if step < steps:
            st.update()
    return {"kind": "load_path", "steps": steps, "per_step": per_step, "seconds": round(time.perf_counter() - t0, 1),
            "qoi_avg_disp": mech.qoi(0)}


def smooth_state(coords, amp=0.02):
    import numpy as np

    X, Y, Z = coords[:, 0], coords[:, 1], coords[:, 2]
    x = np.zeros(4 * coords.shape[0])
    x[0::4] = amp * np.sin(np.pi * X) * np.cos(np.pi * Y) * Z
    x[1::4] = amp * np.cos(np.pi * X) * np.sin(np.pi * Z) * Y
    x[2::4] = amp * np.sin(np.pi * Y) * X * Z + 0.05 * Z
    x[3::4] = 5.0 * np.cos(np.pi * X) * np.cos(np.pi * Y)
    return x


def host_cores():
    """Cores this process may run on, and what the host reports: the CPU
    baseline uses the cores allotted to the process (its affinity set, capped
    by OMP_NUM_THREADS where the host sets it)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    use = min(avail, int(omp)) if omp.isdigit() and int(omp) > 0 else avail
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return use, {"nproc": os.cpu_count(), "affinity": avail, "omp_num_threads": omp or None, "cpu_model": model}


def sample_of(disc, wl, layers, st=None):
    """The first `layers` hexahedron layers (lowest z) of the workload's own
    mesh as a mesh of their own, with the workload's iterate on its nodes.
    Elements are numbered layer by layer (6 tets per hexahedron, x fastest),
    so the sample is an element prefix; its nodes keep their relative order
    (vertices, then edge nodes)."""
    import numpy as np

    nx, ny, _ = wl["dims"]
    coords, conn, _ = disc.mesh()
    ne = min(disc.num_elems, layers * 6 * nx * ny)
    used = np.unique(conn[:ne])
    remap = np.full(disc.num_nodes, -1, dtype=np.int32)
    remap[used] = np.arange(used.size, dtype=np.int32)
    x = disc.soln().reshape(-1, 4)[used].ravel()
    hist = None
    if st is not None and wl["model"] == 2:  # the J2 history of the sample's points (an element prefix)
        npts = ne * wl["npts"]
        hist = (np.ascontiguousarray(st.get("Fp_old")[:, :npts]), np.ascontiguousarray(st.get("eqps_old")[:, :npts]))
    return coords[used], remap[conn[:ne]], x, hist


def cpu_baseline(sample, threads: int, wl, reps: int = 5):
    """The oracle (C++ restatement: dual numbers, column-search CSR scatter)
    assembling `sample` (coords, conn, x) with one part per core, each part
    holding only its own and ghost rows, then the export of ghost rows into
    their owners -- zero_all + assemble + gather_all of Primal::compute_jacob.
    Median of `reps` after one warm-up; the phase split is of the median run."""
    import numpy as np

    import oracle as O

    coords, conn, x, hist = sample
    order, quad = wl["order"], wl["quad"]
    m = O.Mesh(order, coords, conn)
    parts = O.Parts(m, threads)
    st = O.States(m, wl["model"], quad)
    if hist is not None:
        st.Fp_old[:] = hist[0]
        st.eqps_old[:] = hist[1].ravel()
    mt = wl["mat"]
    mats = [[mt["E"], mt["nu"], mt["K"], mt["Y"], mt["c0"]]]
    R = np.zeros(m.num_dofs)
    V = np.zeros(m.nnz)
    runs = []
    for k in range(reps + 1):
        R[:] = 0.0
        V[:] = 0.0
        t0 = time.perf_counter()
        _, _, sec = parts.assemble(O.K_JACOBIAN, wl["model"], mats, x, st, quad_order=quad, R=R, vals=V)
        dt = time.perf_counter() - t0
        if k > 0:
            runs.append((dt, sec))
    runs.sort(key=lambda r: r[0])
    med, sec = runs[len(runs) // 2]
    held, total = parts.rows()
    split = {"zero_s": round(float(sec[0]), 4), "elements_s": round(float(sec[1]), 4),
             "export_s": round(float(sec[2]), 4), "rows_held_over_rows": round(held / total, 3)}
    return m.num_elems / med, m.num_elems, med, split, (coords, conn, x, hist, R, V)


def accuracy(g, wl, sample):
    """The GPU assembly of the CPU-baseline sample against the oracle's R and
    dR/du of the same run (normwise relative), FP64 and FP32-tangent
    element matrices."""
    import numpy as np

    coords, conn, x, hist, R, V = sample
    disc = g.Disc.from_arrays(wl["order"], coords, conn)
    st = g.States(disc, wl["model"], wl["quad"])
    if hist is not None:
        st.set("Fp_old", hist[0])
        st.set("eqps_old", hist[1])
    mech = g.Mechanics(disc, st, [g.MaterialSpec(model=wl["model"], **wl["mat"])], stabilization=True,
                       quad_order=wl["quad"])
    disc.set_soln(x)
    si = g.SolInfo(disc)
    out = {}
    for tangent in (64, 32):
        mech.set_option("tangent", tangent)
        si.zero_all()
        mech.jacobian(si, adjoint=False, save_states=False)
        r, v = si.R(), si.vals()
        tag = "" if tangent == 64 else "_fp32_tangent"
        out["R_rel" + tag] = float(np.linalg.norm(r - R) / np.linalg.norm(R))
        out["dRdu_rel" + tag] = float(np.linalg.norm(v - V) / np.linalg.norm(V))
    return out


def _read_csv(path):
    import csv

    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _find(dirpath, suffix):
    for root, _, files in os.walk(dirpath):
        for fn in files:
            if fn.endswith(suffix):
                return os.path.join(root, fn)
    raise FileNotFoundError(f"no *{suffix} under {dirpath}")


def _short(name: str, tangent: int = 64):
    """Short name of a kernel, or None for an assembly kernel instantiated for
    the other element-matrix precision."""
    for k in ASSEMBLY_KERNELS + ("copy_kernel", "fill_kernel", "expand_csr_kernel"):
        if k in name:
            if k in ASSEMBLY_KERNELS:
                want, other = ("double", "float") if tangent == 64 else ("float", "double")
                if other in name and want not in name:
                    return None
            return k
    return name


def summarize(kt_dir, fetch_dir, write_dir, tag, commands, tangent=64):
    """Turn rocprofv3 CSV output (kernel trace + stats, FETCH_SIZE pass,
    WRITE_SIZE pass) into the text summaries kept under profiles/."""
    stats = _read_csv(_find(kt_dir, "kernel_stats.csv"))
    per = {}
    for which, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        for r in _read_csv(_find(d, "counter_collection.csv")):
            if r["Counter_Name"] != which:
                continue
            k = _short(r["Kernel_Name"], tangent)
            if k is None:
                continue
            per.setdefault(k, {}).setdefault(which, []).append(float(r["Counter_Value"]))
    mean = {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in per.items()}
    # calibration: the copy kernel reads and writes 8 * NCOPY bytes per launch
    cbytes = 8.0 * NCOPY
    cal_fetch = cbytes / (mean["copy_kernel"]["FETCH_SIZE"] * 1024.0)
    cal_write = cbytes / (mean["copy_kernel"]["WRITE_SIZE"] * 1024.0)
    kern = {}
    for r in stats:
        k = _short(r["Name"], tangent)
        if k in ASSEMBLY_KERNELS or k == "copy_kernel":
            kern[k] = {"calls": int(r["Calls"]), "avg_ns": float(r["AverageNs"])}
    out = {"tag": tag, "commands": commands, "calibration": {"fetch": cal_fetch, "write": cal_write},
           "kernels": {}}
    present = [k for k in ASSEMBLY_KERNELS if k in kern]
    for k in present:
        fk, wk = mean[k]["FETCH_SIZE"], mean[k]["WRITE_SIZE"]
        out["kernels"][k] = {"avg_ms": kern[k]["avg_ns"] / 1e6, "fetch_kib": fk, "write_kib": wk,
                             "read_bytes": fk * 1024.0 * cal_fetch, "write_bytes": wk * 1024.0 * cal_write}
    out["traffic_bytes"] = sum(v["read_bytes"] + v["write_bytes"] for v in out["kernels"].values())
    with open(os.path.join(PROFILES, f"{tag}_traffic.json"), "w") as f:
        json.dump(out, f, indent=1)
    lines = [f"# {tag}: kernel trace and HBM traffic", "", "Commands (repository root on the GPU host):", ""]
    lines += [f"    {c}" for c in commands]
    lines += ["", "## rocprofv3 --kernel-trace --stats", "", "| kernel | calls | average ms | total ms | % |",
              "|---|---|---|---|---|"]
    for r in stats:
        lines.append(f"| `{r['Name'][:90]}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.4f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.2f} |")
    inst = "double" if tangent == 64 else "float"
    lines += ["", f"The benchmark also times the other element-matrix precision in the same run; the traffic "
              f"and the per-kernel times below count only the `{inst}` instantiations of the assembly kernels.",
              "", "## FETCH_SIZE / WRITE_SIZE (KiB per dispatch, mean)", "",
              f"Calibration: the copy kernel moves {cbytes / 2**30:.0f} GiB each way per launch; the counters read "
              f"{mean['copy_kernel']['FETCH_SIZE'] / 2**20:.3f} GiB fetched and "
              f"{mean['copy_kernel']['WRITE_SIZE'] / 2**20:.3f} GiB written, so FETCH_SIZE is scaled by "
              f"{cal_fetch:.4f} and WRITE_SIZE by {cal_write:.4f}.", "",
              "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | read GB (scaled) | written GB (scaled) |",
              "|---|---|---|---|---|"]
    for k in present + ["copy_kernel"]:
        fk, wk = mean[k]["FETCH_SIZE"], mean[k]["WRITE_SIZE"]
        lines.append(f"| {k} | {fk:.0f} | {wk:.0f} | {fk * 1024 * cal_fetch / 1e9:.3f} | "
                     f"{wk * 1024 * cal_write / 1e9:.3f} |")
    lines += ["", f"HBM traffic per assembly ({' + '.join(present)}): {out['traffic_bytes'] / 1e9:.3f} GB", ""]
    with open(os.path.join(PROFILES, f"{tag}.md"), "w") as f:
        f.write("\n".join(lines))
    print(json.dumps({"tag": tag, "traffic_GB": out["traffic_bytes"] / 1e9,
                      "avg_ms": {k: v["avg_ms"] for k, v in out["kernels"].items()}}))
    return 0


def summarize_batch(kt_dir, fetch_dir, write_dir, commands, replays):
    """profiles/assembly_batch64.md and its _traffic.json from rocprofv3 CSV
    output of `bench.py --workload batch`: the last `replays` graph replays of
    the kernel trace (each one dispatch per kernel node of the graph), with
    HBM traffic from the FETCH_SIZE / WRITE_SIZE runs over the same dispatches."""
    specs = batch_meshes()
    nodes = 3 * len(specs) + sum(1 for sp_ in specs if sp_["model"] == 2)  # fill, element, gather; J2 point pass
    trace = sorted(_read_csv(_find(kt_dir, "kernel_trace.csv")), key=lambda r: int(r["Start_Timestamp"]))
    timed = trace[-replays * nodes:]
    spans = []
    for k in range(replays):
        rr = timed[k * nodes:(k + 1) * nodes]
        spans.append((max(int(r["End_Timestamp"]) for r in rr) - min(int(r["Start_Timestamp"]) for r in rr)) / 1e6)

    def short(name):
        base = name.replace("void ", "").replace("goal::(anonymous namespace)::", "")
        return base.split("(")[0]

    per = {}
    for r in timed:
        k = short(r["Kernel_Name"])
        d = per.setdefault(k, {"n": 0, "ns": 0.0})
        d["n"] += 1
        d["ns"] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    # counters: the last replays' dispatches of each counter run, per kernel name
    counters = {}
    for which, dd in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = [r for r in _read_csv(_find(dd, "counter_collection.csv")) if r["Counter_Name"] == which]
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        for r in rows[-replays * nodes:]:
            k = short(r["Kernel_Name"])
            counters.setdefault(k, {}).setdefault(which, 0.0)
            counters[k][which] += float(r["Counter_Value"])
    cal = {"fetch": 2.0, "write": 1.0}  # the copy-kernel calibration of the cube profiles
    kernels = {}
    for k, d in per.items():
        c = counters.get(k, {})
        kernels[k] = {"dispatches_per_replay": d["n"] / replays, "ms_per_replay": d["ns"] / 1e6 / replays,
                      "read_bytes": c.get("FETCH_SIZE", 0.0) * 1024 * cal["fetch"] / replays,
                      "write_bytes": c.get("WRITE_SIZE", 0.0) * 1024 * cal["write"] / replays}
    traffic = sum(v["read_bytes"] + v["write_bytes"] for v in kernels.values())
    ne = [6 * sp_["dims"][0] * sp_["dims"][1] * sp_["dims"][2] for sp_ in specs]
    alg = sum((B_ELEM_P2 if sp_["order"] == 2 else B_ELEM_P1 + B_STATE_J2) * n for sp_, n in zip(specs, ne))
    out = {"tag": "assembly_batch64", "commands": commands, "calibration": cal,
           "replay_ms": statistics.mean(spans), "kernel_ms_per_replay": sum(v["ms_per_replay"] for v in kernels.values()),
           "kernels": kernels, "traffic_bytes": traffic, "algorithmic_bytes": alg}
    with open(os.path.join(PROFILES, "assembly_batch64_traffic.json"), "w") as f:
        json.dump(out, f, indent=1)
    stats = _read_csv(_find(kt_dir, "kernel_stats.csv"))
    lines = ["# assembly_batch64: kernel trace and HBM traffic of one graph replay", "",
             f"The batched workload: {len(specs)} seeded box meshes, {sum(ne):,} tets in all "
             f"({min(ne):,} to {max(ne):,} per mesh). Even meshes are P2 neo-Hookean with 4 points, odd ones P1 J2 "
             f"with 1 point and per-point history.",
             f"One captured HIP graph holds every mesh's residual + Jacobian assembly: per mesh one fill, one element "
             f"kernel and one gather, and for the J2 meshes the point pass before the element kernel: {nodes} "
             f"dispatches per replay.", "", "Commands (repository root on the GPU host):", ""]
    lines += [f"    {c}" for c in commands]
    lines += ["", "## rocprofv3 --kernel-trace --stats (whole run: autotuning, capture, warm-up and timed replays)",
              "", "| kernel | calls | average ms | total ms | % |", "|---|---|---|---|---|"]
    for r in stats:
        lines.append(f"| `{r['Name'][:90]}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.4f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.2f} |")
    lines += ["", f"## The {replays} timed replays (the last {replays * nodes} dispatches of the trace)", "",
              "Replay spans, first dispatch start to last dispatch end: "
              + ", ".join(f"{x:.3f}" for x in spans) + " ms.", "",
              "Traffic is from the two counter runs over the same dispatches, FETCH_SIZE ×2 and WRITE_SIZE ×1 "
              "(the copy-kernel calibration of `assembly_p2_nh_cube64.md`).", "",
              "| kernel | dispatches per replay | ms per replay | read GB per replay | written GB per replay |",
              "|---|---|---|---|---|"]
    for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_replay"]):
        lines.append(f"| `{k}` | {v['dispatches_per_replay']:.0f} | {v['ms_per_replay']:.3f} | "
                     f"{v['read_bytes'] / 1e9:.3f} | {v['write_bytes'] / 1e9:.3f} |")
    lines += ["", f"Per replay: {out['kernel_ms_per_replay']:.3f} ms of kernel time and {traffic / 1e9:.2f} GB of HBM "
              f"traffic; algorithmic bytes {alg / 1e9:.2f} GB (13,576 B per P2 tet, 2,648 B per P1 J2 tet), "
              f"ratio {traffic / alg:.2f}.", ""]
    with open(os.path.join(PROFILES, "assembly_batch64.md"), "w") as f:
        f.write("\n".join(lines))
    print(json.dumps({"tag": "assembly_batch64", "replay_ms": out["replay_ms"], "traffic_GB": traffic / 1e9}))
    return 0


def load_traffic(tag, element_ms, gather_ms):
    """HBM bytes per assembly from the committed profile of this workload, if
    its kernel times match this run's within 15% (otherwise it is stale)."""
    path = os.path.join(PROFILES, f"{tag}_traffic.json")
    if not os.path.exists(path) or element_ms is None:
        return None, "no profile"
    with open(path) as f:
        prof = json.load(f)
    # the library's element-pass events bracket the point pass too
    ke = sum(prof["kernels"][k]["avg_ms"] for k in ("j2_point_kernel", "element_kernel") if k in prof["kernels"])
    kg = prof["kernels"]["gather_kernel"]["avg_ms"]
    if abs(ke - element_ms) > 0.15 * element_ms or abs(kg - gather_ms) > 0.15 * gather_ms:
        return None, f"stale profile (element {ke:.2f} ms, gather {kg:.2f} ms)"
    return prof["traffic_bytes"], os.path.relpath(path, ROOT)


def batch_meshes(count=64, seed=0xBA7C):
    """The batched workload: tet counts log-uniform in [2k, 400k], box aspect
    ratios uniform in [1, 8], even meshes P2 neo-Hookean (4 points), odd ones P1
    J2 (1 point, per-point history); each box is nx x ny x ny hexahedra x 6."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        target = float(np.exp(rng.uniform(np.log(2000.0), np.log(400000.0))))
        ar = float(rng.uniform(1.0, 8.0))
        ny = max(1, int(round((target / (6.0 * ar)) ** (1.0 / 3.0))))
        nx = max(1, int(round(ar * ny)))
        out.append(dict(dims=(nx, ny, ny), lengths=(nx / ny, 1.0, 1.0), order=2 if i % 2 == 0 else 1,
                        quad=2 if i % 2 == 0 else 1, model=1 if i % 2 == 0 else 2,
                        mat=NH_MAT if i % 2 == 0 else J2_MAT))
    return out


def bench_batch(args) -> int:
    import numpy as np

    import goal_amd as g

    name, _ = g.device_info()
    t0 = time.perf_counter()
    items = []
    for spec in batch_meshes():
        disc = g.Disc.box(*spec["dims"], *spec["lengths"], order=spec["order"], jitter=0.1, seed=0x5EED)
        st = g.States(disc, spec["model"], spec["quad"])
        mech = g.Mechanics(disc, st, [g.MaterialSpec(model=spec["model"], **spec["mat"])], stabilization=True,
                           quad_order=spec["quad"])
        mech.set_option("check", 0)
        coords, _, _ = disc.mesh()
        disc.set_soln(smooth_state(coords) if spec["model"] == 1 else bar_state(coords))
        items.append((disc, st, mech, g.SolInfo(disc)))
    ne = np.array([it[0].num_elems for it in items])
    log(f"{len(items)} meshes, {ne.sum()} tets ({ne.min()}..{ne.max()} per mesh), built in "
        f"{time.perf_counter() - t0:.1f} s")
    batch = g.Batch([it[2] for it in items], [it[3] for it in items])
    t0 = time.perf_counter()
    blocks, ms = batch.tune(reps=3)
    log(f"autotuned grids in {time.perf_counter() - t0:.1f} s: "
        f"{int((blocks == 0).sum())} occupancy-sized, others {sorted(set(int(b) for b in blocks if b))}")
    batch.capture()
    for _ in range(args.warmup):
        batch.run()
    g.sync()
    K = args.steps
    g.event_record(0)
    t0 = time.perf_counter()
    for _ in range(K):
        batch.run()
    g.event_record(1)
    g.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3 / K
    graph_ms = g.event_elapsed(0, 1) / K
    total = int(ne.sum())
    elems_per_s = total / (wall_ms * 1e-3)
    per_mesh = np.sort(ms.astype(np.float64))
    p50, p99 = float(np.percentile(per_mesh, 50)), float(np.percentile(per_mesh, 99))
    log(f"{K} graph launches: {wall_ms:.3f} ms per batch ({graph_ms:.3f} ms by events), "
        f"{elems_per_s / 1e6:.2f} M elem/s; per mesh p50 {p50:.3f} ms, p99 {p99:.3f} ms; "
        f"sum of per-mesh times {per_mesh.sum():.3f} ms")
    b_elem = np.array([B_ELEM_P2 if it[0].order == 2 else B_ELEM_P1 + B_STATE_J2 for it in items])
    achieved = float((b_elem * ne).sum()) / (graph_ms * 1e-3) / 1e9
    # HBM bytes per replay from the committed profile of this workload, used
    # only while its replay time matches this run's within 15%
    traffic, traffic_src = None, "no profile"
    prof_path = os.path.join(PROFILES, "assembly_batch64_traffic.json")
    if os.path.exists(prof_path):
        with open(prof_path) as f:
            prof = json.load(f)
        if abs(prof["replay_ms"] - graph_ms) <= 0.15 * graph_ms:
            traffic, traffic_src = prof["traffic_bytes"], os.path.relpath(prof_path, ROOT)
        else:
            traffic_src = f"stale profile (replay {prof['replay_ms']:.2f} ms)"
    out = {
        "metric": "batched resid+Jac assembled elems/s (64 meshes, one graph launch)",
        "value": round(elems_per_s, 1),
        "unit": "elem/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (64 seeded Kuhn-tet boxes, smooth synthetic iterates)",
        "config": {"workload": f"64 meshes, {total} tets, tet counts log-uniform 2k-400k, aspect ratios U[1,8], "
                               f"even meshes P2 neo-Hookean 4-pt, odd meshes P1 J2 1-pt; one captured HIP graph of "
                               f"all residual+Jacobian assemblies; per-mesh grid sizes autotuned",
                   "elements": total, "meshes": len(items)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic)},
        "cpu_baseline": None,
        "detail": {"graph_ms": round(graph_ms, 4), "traffic_source": traffic_src,
                   "roofline_unit": "one graph replay = 64 element + 64 gather kernels; "
                                    "B_elem x elements of every mesh / replay duration by HIP events",
                   "per_mesh_ms_p50": round(p50, 4),
                   "per_mesh_ms_p99": round(p99, 4), "sum_per_mesh_ms": round(float(per_mesh.sum()), 4),
                   "tuned_blocks": [int(b) for b in blocks], "elements_per_mesh": [int(x) for x in ne],
                   "device": name},
    }
    print(json.dumps(out), flush=True)
    return 0


def time_learned_kernels(m, net, reps=20):
    """Device time of the two learned-indicator launches on the final mesh,
    by HIP events on the library stream."""
    import numpy as np

    import goal_amd as g
    from goal_amd import learned as LP

    feat = LP.features(m)
    y = g.DeviceBuffer(m.ne, np.float32)
    net.mlp.forward(feat, m.ne, y)
    g.sync()
    g.event_record(10)
    for _ in range(reps):
        LP.features(m, feat)
    g.event_record(11)
    for _ in range(reps):
        net.mlp.forward(feat, m.ne, y, count=False)
    g.event_record(12)
    g.sync()
    f_ms, m_ms = g.event_elapsed(10, 11) / reps, g.event_elapsed(11, 12) / reps
    n = m.ne
    fwd_bytes = n * (4 * 12 + 4)             # feature rows in, predictions out
    fwd_flops = n * 2 * (64 * 16 + 64 * 64)  # the two products as issued (K of layer 1 padded to 16)
    log(f"learned kernels on {n} tets: features {f_ms:.3f} ms (with the AvgVM element pass), "
        f"network {m_ms:.3f} ms = {fwd_bytes / (m_ms * 1e-3) / 1e9:.0f} GB/s, "
        f"{fwd_flops / (m_ms * 1e-3) / 1e12:.1f} TFLOP/s bf16")
    return {"elements": n, "features_ms": round(f_ms, 4), "forward_ms": round(m_ms, 4),
            "forward_bytes": fwd_bytes, "forward_flops": fwd_flops,
            "forward_GBps": round(fwd_bytes / (m_ms * 1e-3) / 1e9, 1),
            "forward_TFLOPs": round(fwd_flops / (m_ms * 1e-3) / 1e12, 2)}


def bench_mlp(args) -> int:
    """The learned indicator's network alone: forward passes over 2 M rows of
    synthetic features drawn around the packaged network's normalization."""
    import numpy as np

    import goal_amd as g
    from goal_amd import learned as LP

    net = LP.load(Path(args.model) if args.model else LP.DEFAULT_MODEL)
    n = args.rows
    rng = np.random.default_rng(7)
    mu, inv = net.norms[:12], net.norms[12:24]
    X = (mu + rng.standard_normal((n, 12)) / inv).astype(np.float32)
    xd = g.DeviceBuffer.from_host(X.ravel(), np.float32)
    yd = g.DeviceBuffer(n, np.float32)
    for _ in range(args.warmup):
        net.mlp.forward(xd, n, yd, count=False)
    g.sync()
    K = args.steps
    g.event_record(0)
    t0 = time.perf_counter()
    for _ in range(K):
        net.mlp.forward(xd, n, yd, count=False)
    g.event_record(1)
    g.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3 / K
    ev_ms = g.event_elapsed(0, 1) / K
    nbytes = n * (4 * 12 + 4)
    flops = n * 2 * (64 * 16 + 64 * 64)
    gbs = nbytes / (ev_ms * 1e-3) / 1e9
    tfs = flops / (ev_ms * 1e-3) / 1e12
    log(f"{n} rows: {ev_ms * 1e3:.1f} us per forward pass (events), {gbs:.0f} GB/s, {tfs:.1f} TFLOP/s bf16")
    out = {
        "metric": "learned-indicator network forward rows/s",
        "value": round(n / (wall_ms * 1e-3), 1),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 weights and product inputs, f32 accumulation",
        "data": "synthetic (Gaussian features around the packaged normalization)",
        "config": {"workload": f"12-64-64-1 network forward, {n} rows, v_mfma_f32_16x16x16_bf16"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "mlp_forward_kernel",
                     "bf16_TFLOPs": round(tfs, 2), "bf16_mfma_frac": round(tfs / 2500.0, 5)},
        "cpu_baseline": None,
        "detail": {"event_ms": round(ev_ms, 5), "bytes_per_pass": nbytes, "flops_per_pass": flops},
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_learned(args) -> int:
    """Goal-oriented adaptation: learned indicator vs plain nested DWR."""
    import numpy as np

    import goal_amd as g
    from goal_amd import learned as LP

    name, _ = g.device_info()
    path = Path(args.model) if args.model else LP.DEFAULT_MODEL
    train_info = None
    if args.retrain or not path.exists():
        t0 = time.perf_counter()
        if args.dataset:   # datasets saved by earlier runs' checkpoints, concatenated
            parts = []
            for path in args.dataset.split(","):
                with np.load(path, allow_pickle=False) as f:
                    parts.append((f["X"], f["Y"], f["G"], json.loads(bytes(f["runs"]).decode())))
            X = np.concatenate([q[0] for q in parts])
            Y = np.concatenate([q[1] for q in parts])
            offs = np.cumsum([0] + [len(q[3]) for q in parts])
            G = np.concatenate([q[2] + o for q, o in zip(parts, offs)]).astype(np.int32)
            runs = [r for q in parts for r in q[3]]
        else:
            ckpt = Path(args.dataset_out) if args.dataset_out else None
            geos = tuple(args.train_geometries.split(","))
            X, Y, G, runs = LP.make_dataset(args.train_runs, seed=args.train_seed, cube_cycles=args.train_cube_cycles,
                                            bar_cycles=args.train_bar_cycles, checkpoint=ckpt, geometries=geos, log=log)
        data_s = time.perf_counter() - t0
        log(f"dataset: {X.shape[0]} samples from {args.train_runs} runs ({len(runs)} meshes) in {data_s:.1f} s")
        net, stats = LP.train(X, Y, epochs=40, batch=4096, lr=2e-3, seed=0, groups=G, log=log)
        train_info = dict(stats, dataset_seconds=round(data_s, 1), runs=len({(r.get("seed", 0), r["run"]) for r in runs}),
                          meshes=len(runs), samples=int(X.shape[0]),
                          max_train_elems={geo: max(r["elems"] for r in runs if r["geometry"] == geo)
                                           for geo in ("cube", "bar") if any(r["geometry"] == geo for r in runs)})
        out_path = Path(args.save_model) if args.save_model else Path("learned_avgvm.npz")
        out_path.parent.mkdir(parents=True, exist_ok=True)
        LP.save(net, out_path, meta=train_info)
        with open(out_path.with_suffix(".runs.json"), "w") as f:
            json.dump(runs, f)
        log(f"saved {out_path}")
        if args.train_only:
            print(json.dumps({"training": train_info, "model": str(out_path)}), flush=True)
            return 0
    else:
        net = LP.load(path)
        train_info = dict(net.meta, model=str(path.name))
    rng = np.random.default_rng(args.eval_seed)
    problems = [(LP.sample_problem(rng, "cube", 8), args.final_elems),
                (LP.sample_problem(rng, "bar", 4), args.final_elems_bar)]
    problems = [pf for pf in problems if pf[0].geometry in args.problems.split(",")]
    report = []
    kernels = None
    methods = ("nested", "learned", "uniform")
    for p, final in problems:
        log(f"problem: {p}, final {final} tets")
        res = {}
        for method in methods:
            recs, last = LP.run_cycles(p, method, args.cycles, final, net=net, log=log)
            res[method] = recs
            if method == "nested":
                ref = LP.overkill_qoi(last, log=log)
            if method == "learned" and p.geometry == "cube":
                kernels = time_learned_kernels(last, net)
            del last
        for method in methods:
            for r in res[method]:
                err = abs(ref["J"] - r["J"])
                r["qoi_error"] = err
                if "eta_sum" in r and method != "uniform":
                    r["effectivity"] = r["eta_sum"] / err if err > 0 else None
                if "learned_eta_sum" in r:
                    r["learned_effectivity_same_mesh"] = r["learned_eta_sum"] / err if err > 0 else None
        report.append({"problem": p.__dict__, "final_elems": final, "reference": ref, **res})
    est_n = sum(r["estimate_seconds"] for rep in report for r in rep["nested"] if "estimate_seconds" in r)
    est_l = sum(r["estimate_seconds"] for rep in report for r in rep["learned"] if "estimate_seconds" in r)
    # effectivities of the indicator sums on the same (nested-adapted) meshes
    same = [(r["effectivity"], r["learned_effectivity_same_mesh"]) for rep in report for r in rep["nested"]
            if r.get("effectivity") and r.get("learned_effectivity_same_mesh")]
    gaps = [abs(a - b) for a, b in same]
    rel = [abs(b / a - 1.0) for a, b in same]
    fallbacks = sum(1 for rep in report for r in rep["learned"] if r.get("fallback"))
    final_err = {rep["problem"]["geometry"]: {m: rep[m][-1]["qoi_error"] for m in methods} for rep in report}
    out = {
        "metric": "learned vs plain nested-DWR error estimation speed-up (AvgVM on x<0.25, 5 adapt cycles)",
        "value": round(est_n / est_l, 1),
        "unit": "x",
        "n_gpus": 1,
        "steps": args.cycles,
        "warmup": 0,
        "ms_per_step": round(est_l * 1e3 / max(1, args.cycles * len(problems)), 4),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64 assembly and solves; bf16 weights, f32 accumulation in the network",
        "data": "synthetic (seeded cube and bar problems, random material and load)",
        "config": {"workload": f"cube (neo-Hookean, 8^3 hex base, to {args.final_elems} tets) and bar (J2, "
                               f"32x4x4 hex base, to {args.final_elems_bar} tets), one load step, AvgVM over "
                               f"elements with centroid x<0.25, {args.cycles} adapt cycles; reference J from the "
                               f"1:8 refinement of the final nested-adapted mesh",
                   "final_elems": args.final_elems, "final_elems_bar": args.final_elems_bar,
                   "cycles": args.cycles},
        "roofline": None if kernels is None else {
            "bound": "hbm", "achieved": kernels["forward_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(kernels["forward_GBps"] / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "mlp_forward_kernel", "bf16_TFLOPs": kernels["forward_TFLOPs"],
            "bf16_mfma_frac": round(kernels["forward_TFLOPs"] / 2500.0, 5)},
        "cpu_baseline": None,
        "detail": {"estimate_seconds_nested": round(est_n, 4), "estimate_seconds_learned": round(est_l, 5),
                   "learned_kernels": kernels,
                   "effectivity_pairs_same_mesh": [[round(a, 4), round(b, 4)] for a, b in same],
                   "mean_abs_effectivity_gap": round(float(np.mean(gaps)), 4) if gaps else None,
                   "max_abs_effectivity_gap": round(float(np.max(gaps)), 4) if gaps else None,
                   "mean_rel_effectivity_gap": round(float(np.mean(rel)), 4) if rel else None,
                   "final_qoi_error": final_err,
                   "learned_fallbacks": fallbacks, "training": train_info, "device": name,
                   "problems": report},
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64, help="hexahedra per cube side")
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--workload", default="cube", choices=("cube", "bar", "batch", "learned", "mlp"),
                    help="cube: P2 neo-Hookean cube (the headline); bar: P1 J2 bar, 5.3 M tets; "
                         "batch: 64 meshes of uneven size in one graph launch; learned: adaptation with the "
                         "learned indicator against plain nested DWR; mlp: its network's forward pass alone")
    ap.add_argument("--rows", type=int, default=2_000_000, help="mlp: rows per forward pass")
    ap.add_argument("--model", default="", help="learned: trained network (.npz); default the packaged one")
    ap.add_argument("--retrain", action="store_true", help="learned: build the dataset and train first")
    ap.add_argument("--save-model", default="", help="learned: where a retrained network is written")
    ap.add_argument("--train-runs", type=int, default=200)
    ap.add_argument("--train-only", action="store_true", help="learned: stop after training")
    ap.add_argument("--dataset", default="", help="learned: train from saved datasets (.npz, comma-separated)")
    ap.add_argument("--train-geometries", default="cube,bar", help="learned: problems of the training runs, in turn")
    ap.add_argument("--train-cube-cycles", type=int, default=4, help="learned: adapt cycles of a cube training run")
    ap.add_argument("--train-bar-cycles", type=int, default=2, help="learned: adapt cycles of a bar training run")
    ap.add_argument("--train-seed", type=int, default=0)
    ap.add_argument("--dataset-out", default="", help="learned: save the generated dataset here (every 25 runs)")
    ap.add_argument("--cycles", type=int, default=5)
    ap.add_argument("--final-elems", type=int, default=2_000_000, help="learned: final size of the cube runs")
    ap.add_argument("--final-elems-bar", type=int, default=2_000_000, help="learned: final size of the bar runs")
    ap.add_argument("--eval-seed", type=int, default=2024)
    ap.add_argument("--problems", default="cube,bar", help="learned: which evaluation problems to run")
    ap.add_argument("--blocks", type=int, default=0)
    ap.add_argument("--iterate", default="newton", choices=("synthetic", "newton"),
                    help="the assembled state -- a smooth synthetic field, or solved by Newton on the device "
                         "(untimed): the cube's load step 1, the bar's whole tension load path")
    ap.add_argument("--load-steps", type=int, default=20, help="bar: tension load steps before timing")
    ap.add_argument("--overlap", type=int, default=-1,
                    help="neo-Hookean: assemble in this many chunks, gathering finished chunks on a second stream "
                         "(-1: the library's automatic choice, 0: off)")
    ap.add_argument("--overlap-blocks", type=int, default=0, help="element workgroups per chunk (0: 5/6 resident)")
    ap.add_argument("--xcd-remap", type=int, default=1)
    ap.add_argument("--cpu-layers", type=int, default=8,
                    help="hexahedron layers (lowest z) of the workload's mesh in the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline parts (0: the cores this process has)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tangent", type=int, default=64, choices=(64, 32),
                    help="element-matrix precision of the measured path (the residual is FP64 either way)")
    ap.add_argument("--summarize", nargs=3, metavar=("KT_DIR", "FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/ summaries from rocprofv3 CSV output of this benchmark and exit")
    args = ap.parse_args()

    if args.summarize and args.workload == "batch":
        base = f"python3 bench.py --workload batch --steps {args.steps} --warmup 1"
        cmds = [f"rocprofv3 --kernel-trace --stats --output-format csv -d {args.summarize[0]} -o run -- {base}",
                f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d {args.summarize[1]} -o run -- {base}",
                f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d {args.summarize[2]} -o run -- {base}"]
        return summarize_batch(*args.summarize, cmds, args.steps)
    if args.summarize:
        base = (f"python3 bench.py --workload {args.workload} --iterate synthetic"
                + (f" --overlap {args.overlap}" if args.overlap >= 0 else "")
                + f" --n {args.n} --order {args.order} --tangent {args.tangent}")
        cmds = [f"rocprofv3 --kernel-trace --stats --output-format csv -d {args.summarize[0]} -o run -- "
                f"{base} --steps 5 --warmup 1 --no-cpu-baseline",
                f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d {args.summarize[1]} -o run -- "
                f"{base} --steps 2 --warmup 1 --no-cpu-baseline",
                f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d {args.summarize[2]} -o run -- "
                f"{base} --steps 2 --warmup 1 --no-cpu-baseline"]
        return summarize(*args.summarize, profile_tag(args.n, args.order, args.tangent, args.workload), cmds,
                         args.tangent)

    if args.gpus != 1:
        print(
            "bench.py measures one GPU: goal_amd has no multi-GPU path (the assembly would shard by "
            "mesh partition, one part per GPU, with an exchange of interface rows). Use --gpus 1.",
            file=sys.stderr,
        )
        return 2

    import numpy as np

    import goal_amd as g

    name, cus = g.device_info()
    log(f"device: {name}, {cus} CUs")
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "mlp":
        return bench_mlp(args)
    wl = workload(args.workload, args.n, args.order)
    order, quad = wl["order"], wl["quad"]
    n = args.n
    t0 = time.perf_counter()
    disc = g.Disc.box(*wl["dims"], *wl["lengths"], order=order, jitter=0.1, seed=0x5EED)
    log(f"mesh+graph: {disc.num_elems} tets, {disc.num_nodes} nodes, {disc.num_dofs} dofs, nnz {disc.nnz}, "
        f"max degree {disc.max_degree} ({time.perf_counter() - t0:.1f} s)")
    st = g.States(disc, wl["model"], quad)
    mech = g.Mechanics(disc, st, [g.MaterialSpec(model=wl["model"], **wl["mat"])], stabilization=True,
                       quad_order=quad)
    mech.set_option("check", 0)
    if args.blocks:
        mech.set_option("blocks", args.blocks)
    mech.set_option("xcd_remap", args.xcd_remap)
    mech.set_option("tangent", args.tangent)
    if args.overlap >= 0:
        mech.set_option("overlap", args.overlap)
    mech.set_option("overlap_blocks", args.overlap_blocks)
    coords, _, _ = disc.mesh()
    si = g.SolInfo(disc)
    iterate = {"kind": "synthetic"}
    if args.iterate == "newton" and wl["name"] == "cube":
        iterate = newton_iterate(g, disc, mech, si)
    elif args.iterate == "newton":
        iterate = bar_load_path(g, disc, mech, si, st, steps=args.load_steps)
    else:
        disc.set_soln(smooth_state(coords) if wl["name"] == "cube" else bar_state(coords))

    def step(slot=None):
        si.zero_all()
        if slot is not None:
            g.event_record(slot)
        mech.jacobian(si, adjoint=False, save_states=True)
        if slot is not None:
            g.event_record(slot + 1)

    for _ in range(args.warmup):
        step()
    g.sync()
    assert mech.status() == 0
    K = args.steps
    use_events = 2 * K <= 250
    g.sync()
    t0 = time.perf_counter()
    for k in range(K):
        step(2 * k if use_events else None)
    g.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / K
    kernel_ms = statistics.mean(g.event_elapsed(2 * k, 2 * k + 1) for k in range(K)) if use_events else None
    elems_per_s = disc.num_elems / (ms_per_step / 1e3)
    log(f"{K} steps: {ms_per_step:.3f} ms/step, assembly kernel {kernel_ms:.3f} ms, {elems_per_s / 1e6:.2f} M elem/s")

    # element pass / gather split, from events the library records around each kernel
    mech.set_option("timing", 1)
    splits = []
    for _ in range(5):
        step()
        splits.append(mech.timing())
    mech.set_option("timing", 0)
    element_ms = statistics.mean(e for e, _ in splits)
    gather_ms = statistics.mean(gg for _, gg in splits)
    log(f"element pass {element_ms:.3f} ms, gather {gather_ms:.3f} ms")

    def chunks(tangent):   # the library's choice for the "overlap" option (automatic: one pass)
        return max(args.overlap, 0)

    # the other element-matrix precision, same mesh and iterate, for comparison
    other = 32 if args.tangent == 64 else 64
    mech.set_option("tangent", other)
    for _ in range(args.warmup):
        step()
    g.sync()
    t0 = time.perf_counter()
    for _ in range(K):
        step()
    g.sync()
    other_ms = (time.perf_counter() - t0) * 1e3 / K
    mech.set_option("timing", 1)
    osplit = []
    for _ in range(5):
        step()
        osplit.append(mech.timing())
    mech.set_option("timing", 0)
    mech.set_option("tangent", args.tangent)
    log(f"tangent {other}: {other_ms:.3f} ms/step (element {statistics.mean(e for e, _ in osplit):.3f} ms, "
        f"gather {statistics.mean(gg for _, gg in osplit):.3f} ms)")
    step()
    g.sync()

    # sanity: the measured state is a real assembly (finite, nonzero)
    R = si.R()
    assert np.isfinite(R).all() and np.abs(R).max() > 0

    # in-run STREAM-style copy bandwidth (read + write bytes / time)
    ncopy = NCOPY
    a = g.DeviceArray(ncopy)
    b = g.DeviceArray(ncopy)
    g.device_copy(b, a)
    g.sync()
    g.event_record(252)
    for _ in range(5):
        g.device_copy(b, a)
    g.event_record(253)
    copy_ms = g.event_elapsed(252, 253) / 5
    copy_gbs = 2 * 8 * ncopy / (copy_ms * 1e-3) / 1e9
    del a, b
    log(f"copy kernel: {copy_gbs:.0f} GB/s")

    b_elem = wl["b_elem"]
    achieved = b_elem * disc.num_elems / (kernel_ms * 1e-3) / 1e9 if kernel_ms else None
    traffic, traffic_src = load_traffic(profile_tag(n, order, args.tangent, wl["name"]), element_ms, gather_ms)
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
        "traffic": round(traffic) if traffic else None,
    }

    cpu = None
    acc = None
    if not args.no_cpu_baseline:
        threads, host = host_cores()
        if args.cpu_threads:
            threads = args.cpu_threads
        layers = args.cpu_layers
        log(f"cpu baseline: oracle on the first {layers} hex layers of this mesh at this iterate, {threads} parts "
            f"({host})")
        sample = sample_of(disc, wl, layers, st)
        v, ne_s, med, split, sample = cpu_baseline(sample, threads, wl)
        acc = accuracy(g, wl, sample)
        acc["bar_fp64"] = 1e-11
        acc["sample"] = "the cpu_baseline sample, GPU against the oracle's output of the same run"
        log(f"accuracy on the sample: R {acc['R_rel']:.2e}, dR/du {acc['dRdu_rel']:.2e} (FP64); "
            f"dR/du {acc['dRdu_rel_fp32_tangent']:.2e} (FP32 tangent)")
        cpu = {
            "value": round(v, 1),
            "unit": "elem/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{ne_s} P{order} tets: the first {layers} of the {wl['dims'][2]} hex layers of this "
                      f"mesh at this iterate ({iterate['kind']}); zero_all + residual/Jacobian with "
                      f"FAD-{16 if order == 1 else 40} and column-search CSR scatter + export of ghost rows, "
                      f"one part per core holding its own and ghost rows; median of 5 after 1 warm-up "
                      f"({med:.3f} s each); host {host['cpu_model']}, nproc {host['nproc']}, "
                      f"{host['affinity']} cores in this process's affinity set, OMP_NUM_THREADS "
                      f"{host['omp_num_threads']}",
            "host": host,
            "split": split,
        }
        log(f"cpu baseline: {v:.0f} elem/s ({elems_per_s / v:.0f}x), split {split}")

    out = {
        "metric": wl["metric"],
        "value": round(elems_per_s, 1),
        "unit": "elem/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.tangent == 64 else "f64 residual, f32 element matrices",
        "data": ("synthetic (Kuhn-tet box mesh; the iterate is load step 1 solved by device Newton)"
                 if iterate["kind"] == "newton" else
                 "synthetic (Kuhn-tet box mesh; the iterate is the last of 20 tension load steps solved by "
                 "device Newton, with its plastic history)" if iterate["kind"] == "load_path" else
                 "synthetic (Kuhn-tet box mesh, smooth synthetic displacement/pressure iterate)"),
        "config": {
            "workload": f"{wl['desc']} -> {disc.num_elems} P{order} tets, {disc.num_dofs} dofs, nnz {disc.nnz}; "
                        f"residual + Jacobian (zero_all + assembly)",
            "elements": disc.num_elems,
            "order": order,
            "quad_points": wl["npts"],
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": {
            "kernel_ms": round(kernel_ms, 4) if kernel_ms else None,
            "element_kernel_ms": round(element_ms, 4),
            "gather_kernel_ms": round(gather_ms, 4),
            "roofline_unit": "one assembly launch = element_kernel + gather_kernel on one stream, "
                             "HIP events around both; B_elem x elements / its duration",
            "traffic_source": traffic_src,
            "tangent_bits": args.tangent,
            "iterate": iterate,
            "accuracy": acc,
            "chunks": chunks(args.tangent),
            "chunked_split": "with chunks > 1 the element time runs from the first chunk to the last and "
                             "overlaps the side-stream gathers of finished node runs; the gather time is "
                             "the tail after the last chunk",
            f"tangent{other}": {"ms_per_step": round(other_ms, 4),
                                "chunks": chunks(other),
                                "elem_per_s": round(disc.num_elems / (other_ms * 1e-3), 1),
                                "element_kernel_ms": round(statistics.mean(e for e, _ in osplit), 4),
                                "gather_kernel_ms": round(statistics.mean(gg for _, gg in osplit), 4)},
            "algorithmic_bytes_per_elem": b_elem,
            "copy_kernel_GBs": round(copy_gbs, 1),
            "device": name,
        },
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
