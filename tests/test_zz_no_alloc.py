"""psr_solve_steps_noalloc (steps staged through a fixed per-GPU device
array, no allocation) against psr_solve_steps on the same steps: bit for
bit, on two calls in a row, the second smaller than the first, with H-only
isothermal and H + He temperature physics."""
import ctypes as C

import numpy as np
import pytest

import psphray_amd as psr
from psphray_amd import _abi

from conftest import RATES, print_case

pytestmark = pytest.mark.gpu


def _steps(golden, key, n, seed):
    c = np.load(f"{golden}/chem_lsoda.npz")
    base = c[f"{key}_in"]
    rng = np.random.default_rng(seed)
    s = np.ascontiguousarray(base[rng.integers(0, len(base), n)], psr.STEP_DTYPE).copy()
    s["time"] *= rng.uniform(0.5, 1.5, n)
    return s


def _noalloc(ctx, steps, rtol=0.0):
    out = np.ascontiguousarray(steps, psr.STEP_DTYPE).copy()
    rc = ctx._lib.psr_solve_steps_noalloc(ctx.h, out.ctypes.data, len(out), rtol)
    return rc, out


@pytest.mark.parametrize("key,hmf,iso", [("h_iso", 1.0, True), ("he", 0.76, False)])
def test_noalloc_steps_equal_the_allocating_entry_point(golden, key, hmf, iso):
    with psr.Context(0) as ctx:
        ctx.set_physics(psr.Physics(RATES, psr.default_units(hmf=hmf),
                                    psr.make_flags(isothermal=iso, h_only=iso)))
        for n, seed in ((6000, 1), (700, 2)):
            print_case("noalloc", key, n)
            s = _steps(golden, key, n, seed)
            ref = ctx.solve_steps(s)
            rc, out = _noalloc(ctx, s)
            assert rc == 0, ctx._lib.psr_last_error(ctx.h)
            assert out.tobytes() == ref.tobytes()


def test_noalloc_refuses_more_than_its_array_holds(golden):
    with psr.Context(0) as ctx:
        ctx.set_physics(psr.Physics(RATES))
        n = _abi.PSR_NOALLOC_MAX_STEPS + 1
        s = _steps(golden, "h_iso", n, 3)
        rc, out = _noalloc(ctx, s)
        assert rc == _abi.PSR_ERR_CAPACITY
        assert out.tobytes() == s.tobytes()
        # the array itself still works afterwards, at its full size
        s = s[: _abi.PSR_NOALLOC_MAX_STEPS]
        rc, out = _noalloc(ctx, s)
        assert rc == 0
        assert out.tobytes() == ctx.solve_steps(s).tobytes()
