"""fsa_scene_get_statistics_noalloc against fsa_scene_get_statistics, through the C ABI.

The no-allocation entry point keeps its partials in a device array shared by every
call on the GPU; two calls in a row, the second on a smaller scene, must each equal
the allocating entry point bit for bit (the second call reads nothing the first left)."""
import ctypes as C

import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd._native import lib
from fluidsolver_amd.solver import Scene

pytestmark = pytest.mark.gpu


def _stats(fn, sc, zero):
    out = abi.SceneStatistics()
    rc = fn(sc.handle, float(zero), C.byref(out))
    assert rc == abi.FSA_OK, lib().fsa_last_error().decode()
    return bytes(out)


def test_statistics_without_allocation_equals_allocating(gpu_context):
    L = lib()
    big = scenes.dam_scene((24, 20, 16), (60, 30, 30), 3, seed=41)
    small = scenes.dam_scene((5, 4, 3), (12, 8, 8), 3, seed=42)
    print("case no-alloc statistics", big.size(), small.size(), flush=True)
    results = []
    for coll, steps, zero in ((big, 3, 0.0), (small, 2, -1.5)):
        sc = Scene(gpu_context, abi.default_desc(3, abi.SOLVER_IISPH), coll.copy())
        sc.step(steps)
        want = _stats(L.fsa_scene_get_statistics, sc, zero)
        got = _stats(L.fsa_scene_get_statistics_noalloc, sc, zero)
        assert got == want
        results.append(abi.SceneStatistics.from_buffer_copy(got))
        sc.close()
    assert results[0].normal_particles == 24 * 20 * 16 and results[1].normal_particles == 60
    assert results[1].kinetic_energy > 0.0
