"""1000-step runs: the GPU against the oracle.

* against the FP32-arithmetic / FP64-reduction oracle (the device's exact
  semantics) every step is bit-identical: iteration count, dt, FP64 average
  density error and the final state;
* against the pure reference-order oracle (FP32 serial convergence sum as in
  IISPHFluidSolver3D.hpp:467-509) iteration counts agree within +-1 per step and
  the average density error and kinetic energy within 1%.
"""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.solver import Scene

pytestmark = pytest.mark.gpu

STEPS = 1000


def kinetic_energy(c):
    fl = c.type == abi.TYPE_NORMAL
    v = c.velocity[fl].astype(np.float64)
    return 0.5 * float(np.sum(c.mass[fl].astype(np.float64) * np.sum(v * v, 1)))


@pytest.mark.timeout(600)
def test_thousand_steps_against_oracle(gpu_context, oracle_lib):
    """A collapsing column with dt = 0.02 and a 1e-4 density tolerance: the
    Jacobi loop runs 2-30 iterations per step (about 12 on average), so the
    convergence decision is exercised on every step."""
    O = oracle_lib
    coll = scenes.dam_scene((6, 14, 6), (18, 18, 10), 3, seed=3)
    desc = abi.default_desc(3, abi.SOLVER_IISPH, timestep=0.02, iisph_max_density_error_allowed=1e-4)
    print(f"case 1000 steps n={coll.size()}", flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    exact = O.OracleScene(desc, coll.copy(), mode=O.MODE_FP32_R64)
    ref = O.OracleScene(desc, coll.copy(), mode=O.MODE_FP32)
    g_it, e_it, r_it, g_err, r_err = [], [], [], [], []
    for step in range(STEPS):
        g = sc.step(1)
        e = exact.step(1)
        r = ref.step(1)
        g_it.append(g.iterations)
        e_it.append(e.iterations)
        r_it.append(r.iterations)
        g_err.append(g.average_density_error)
        r_err.append(r.average_density_error)
        assert g.iterations == e.iterations and g.actual_time_step == e.actual_time_step, step
        assert g.average_density_error == pytest.approx(e.average_density_error, rel=1e-12, abs=1e-300), step
        if step % 100 == 0:
            print(f"step {step}: iterations {g.iterations} avg err {g.average_density_error:.3e}", flush=True)
    gs = sc.download(coll.copy())
    es = exact.download()
    rs = ref.download()
    np.testing.assert_array_equal(gs.movement, es.movement)
    np.testing.assert_array_equal(gs.particle_data, es.particle_data)
    assert np.mean(g_it) > 5
    assert np.max(np.abs(np.array(g_it) - np.array(r_it))) <= 1
    assert np.mean(g_err) == pytest.approx(np.mean(r_err), rel=0.01)
    assert kinetic_energy(gs) == pytest.approx(kinetic_energy(rs), rel=0.01)
    assert kinetic_energy(gs) > 0.0
