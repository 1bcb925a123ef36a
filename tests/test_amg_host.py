"""Host checks of the multigrid set-up's dense helpers (goal_amd/csrc/goal_amg.cpp),
built with g++ under AddressSanitizer and UBSan (tests/cpp/amg_host_check.cpp):

- goal::hessenberg_eigenvalues against numpy.linalg.eigvals on seeded random
  upper Hessenberg matrices (the Arnoldi matrices of the smoother's spectrum
  estimate are 10 x 10), and against closed forms: a triangular matrix (its
  diagonal), rotation-scaling blocks a +- bi, the companion matrix of a
  polynomial with known roots, a zero subdiagonal that splits the matrix;
- goal::chebyshev_ellipse against its definition: real Ritz values give a real
  interval [hi/10, hi] (c2 = a^2), complex ones an ellipse through them.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def _exe():
    exe = ROOT / "tests" / "cpp" / "amg_host_check"
    if not exe.exists():
        import importlib.util
        spec = importlib.util.spec_from_file_location("graft_entry", ROOT / "__graft_entry__.py")
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        try:
            mod.build_amg_host_check()
        except (OSError, subprocess.CalledProcessError) as e:
            pytest.skip(f"host build of the multigrid helpers needs g++: {e}")
    return exe


def _run(mats):
    text = [str(len(mats))]
    for m in mats:
        n = m.shape[0]
        text.append(str(n))
        text.append(" ".join(repr(float(v)) for v in m.ravel()))
    r = subprocess.run([str(_exe())], input="\n".join(text) + "\n", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
    rows = [list(map(float, line.split())) for line in r.stdout.strip().splitlines()]
    out, k = [], 0
    for m in mats:
        n = m.shape[0]
        ev = np.array([complex(a, b) for a, b in rows[k:k + n]])
        theta, c2 = rows[k + n]
        out.append((ev, theta, c2))
        k += n + 1
    return out


def _match(got, want, tol):
    """Every wanted eigenvalue has a distinct computed one within tol."""
    left = list(got)
    for w in want:
        d = [abs(g - w) for g in left]
        i = int(np.argmin(d))
        assert d[i] <= tol * max(1.0, abs(w)), (w, left)
        left.pop(i)


def test_hessenberg_eigenvalues_random_against_numpy():
    rng = np.random.default_rng(12)
    mats = []
    for n in (1, 2, 3, 5, 10, 10, 10, 16):
        h = np.triu(rng.standard_normal((n, n)), -1)
        mats.append(h)
    for h, (ev, _, _) in zip(mats, _run(mats)):
        _match(ev, np.linalg.eigvals(h), 1e-9)


def test_hessenberg_eigenvalues_closed_forms():
    tri = np.triu(np.arange(1.0, 26.0).reshape(5, 5))
    rot = np.zeros((4, 4))
    rot[:2, :2] = [[0.4, -0.7], [0.7, 0.4]]
    rot[2:, 2:] = [[2.0, -0.1], [0.1, 2.0]]
    rot[1, 2] = 0.3
    rot[2, 1] = 1e-3
    roots = np.array([0.5, 1.0, 1.5, 2.0, 3.0])
    c = np.poly(roots)                     # monic, highest power first
    comp = np.zeros((5, 5))
    comp[0, :] = -c[1:]
    comp[1:, :-1] = np.eye(4)              # upper Hessenberg companion
    split = np.triu(np.ones((4, 4)), -1)
    split[2, 1] = 0.0
    res = _run([tri, rot, comp, split])
    _match(res[0][0], np.diag(tri), 1e-12)
    _match(res[1][0], np.linalg.eigvals(rot), 1e-10)
    assert np.abs(res[1][0].imag).max() > 0.6
    _match(res[2][0], roots, 1e-8)
    _match(res[3][0], np.linalg.eigvals(split), 1e-10)


def test_chebyshev_ellipse_definition():
    real = np.diag([0.5, 1.0, 2.0])
    (_, theta, c2), = _run([real])
    hi = 1.1 * 2.0
    assert theta == pytest.approx(0.55 * hi, rel=1e-14)
    assert c2 == pytest.approx((0.45 * hi) ** 2, rel=1e-14)
    # a complex pair inside the real extent sets the imaginary semi-axis
    blk = np.array([[1.0, -0.8, 0.0], [0.8, 1.0, 0.0], [0.0, 0.0, 2.0]])
    (ev, theta, c2), = _run([blk])
    hi = 1.1 * max(2.0, abs(complex(1.0, 0.8)))
    d, a = 0.55 * hi, 0.45 * hi
    u = (1.0 - d) / a
    b = 1.1 * 0.8 / np.sqrt(1.0 - u * u)
    assert theta == pytest.approx(d, rel=1e-12)
    assert c2 == pytest.approx(a * a - b * b, rel=1e-9)
    assert c2 < a * a
    # the fitted ellipse encloses the pair
    assert ((1.0 - d) / a) ** 2 + (0.8 / b) ** 2 <= 1.0
