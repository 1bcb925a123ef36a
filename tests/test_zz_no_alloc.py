"""goal_asm_jacobian_noalloc through the C ABI against goal_asm_jacobian: the
allocation-free entry point keeps its element buffers in the library's
static device array and must give the same bits, call after call, with a
smaller mesh on the second call (the array keeps the first call's contents)."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import MATS, random_history, random_state

pytestmark = pytest.mark.gpu

g = pytest.importorskip("goal_amd")

lib = g.lib
lib.goal_asm_jacobian_noalloc.restype = C.c_int
lib.goal_asm_jacobian_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]


def setup(order, quad, model, dims, seed):
    disc = g.Disc.box(*dims, order=order, jitter=0.1, seed=seed)
    st = g.States(disc, model, quad)
    mech = g.Mechanics(disc, st, [g.MaterialSpec(**MATS[model])], stabilization=True, quad_order=quad)
    # small enough that no element inverts on the finest box (J2 has no
    # finite stress at J <= 0)
    disc.set_soln(random_state(disc.num_nodes, seed, scale_u=0.002))
    if model == 2:
        Fp, eq = random_history(disc.num_elems * st.nq, seed + 1)
        for k, v in (("Fp_old", Fp), ("eqps_old", eq), ("Fp", Fp), ("eqps", eq)):
            st.set(k, v)
    return disc, st, mech


def noalloc(mech, si, adjoint=False, save=True):
    g.check(lib.goal_asm_jacobian_noalloc(mech.h, si.h, int(adjoint), int(save)))


# (order, quad, model, box dims, tangent): the first call is the larger mesh
PAIRS = [
    ((2, 2, 1, (8, 8, 10), 64), (2, 2, 1, (3, 2, 2), 64)),   # 3,840 then 72 P2 tets
    ((1, 1, 2, (16, 15, 12), 64), (1, 1, 2, (4, 3, 2), 64)),  # 17,280 then 144 P1 J2 tets
    ((2, 2, 1, (6, 6, 6), 32), (2, 2, 1, (2, 2, 3), 32)),    # FP32 element matrices
]


@pytest.mark.parametrize("pair", PAIRS)
def test_noalloc_matches_jacobian_twice(pair):
    for k, (order, quad, model, dims, tangent) in enumerate(pair):
        disc, st, mech = setup(order, quad, model, dims, seed=60 + k)
        mech.set_option("tangent", tangent)
        si = g.SolInfo(disc)
        for adjoint in (False, True):
            si.zero_all()
            mech.jacobian(si, adjoint=adjoint, save_states=True)
            R, V, sig = si.R(), si.vals(), st.get("sigma")
            assert np.isfinite(R).all() and np.isfinite(V).all()
            si.set(R=np.full(disc.num_dofs, 3.0), vals=np.full(disc.nnz, -2.0))   # stale outputs
            st.set("sigma", np.zeros_like(sig))
            si.zero_all()
            noalloc(mech, si, adjoint=adjoint, save=True)
            assert np.array_equal(si.R(), R), (k, adjoint)
            assert np.array_equal(si.vals(), V), (k, adjoint)
            assert np.array_equal(st.get("sigma"), sig), (k, adjoint)
            # accumulating semantics as goal_asm_jacobian: without a zero, it adds
            noalloc(mech, si, adjoint=adjoint, save=False)
            mech.jacobian(si, adjoint=adjoint, save_states=False)
            R3, V3 = si.R(), si.vals()
            si.zero_all()
            mech.jacobian(si, adjoint=adjoint, save_states=False)
            mech.jacobian(si, adjoint=adjoint, save_states=False)
            mech.jacobian(si, adjoint=adjoint, save_states=False)
            assert np.array_equal(R3, si.R()) and np.array_equal(V3, si.vals()), (k, adjoint)


def test_noalloc_rejects_meshes_beyond_its_scratch():
    """5,115 ten-node neo-Hookean tets fit the 64 MiB array; 6,000 do not,
    and the call refuses before launching anything (the outputs keep their
    values)."""
    disc, st, mech = setup(2, 2, 1, (10, 10, 10), seed=5)
    si = g.SolInfo(disc)
    si.set(R=np.full(disc.num_dofs, 1.5))
    with pytest.raises(g.GoalError) as err:
        noalloc(mech, si)
    assert err.value.code == g.ERR_INVALID
    assert "5115" in str(err.value)
    assert np.array_equal(si.R(), np.full(disc.num_dofs, 1.5))
