"""The element kernels' point math (goal_amd/csrc/goal_device.hpp) and LDS index
arithmetic built for the host under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/cpp/device_math_check.cpp): J2 radial return
and tangent (plastic and elastic branches), neo-Hookean closed-form columns and
the stabilization flux against central finite differences, the J2 invariants,
and every phase's LDS indices for every workgroup shape.  Two builds differ
only in how stack variables are filled before first use (a pattern, zeros);
their printed values must be equal bit for bit, so no value depends on a
variable read before it is written."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _exes():
    exes = [ROOT / "tests" / "cpp" / f"device_math_check_{f}" for f in ("pattern", "zero")]
    if not all(e.exists() for e in exes):
        import importlib.util
        spec = importlib.util.spec_from_file_location("graft_entry", ROOT / "__graft_entry__.py")
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        try:
            mod.build_device_math_check()
        except (OSError, subprocess.CalledProcessError) as e:
            pytest.skip(f"host build of the device math needs hipcc: {e}")
    return exes


def test_device_math_on_host_under_sanitizers():
    outs = []
    for exe in _exes():
        r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-3000:]
        assert "all checks passed" in r.stderr
        outs.append(r.stdout)
    assert len(outs[0].splitlines()) > 3000
    assert outs[0] == outs[1], "a computed value depends on an uninitialized variable"
