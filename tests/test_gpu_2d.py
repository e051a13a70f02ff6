"""2D path on the GPU against the oracle: HashedNeighborhoodSearch and
QuadraticNeighborhoodSearchDynamicAllocated neighbor lists (order included),
IISPHFluidSolver and SESPHFluidSolver steps bit for bit, and the 2D breaking
dam (64x64 fluid in a 200x100 box) for 500 steps.

The oracle runs with FP64 convergence sums (the device's reduction); the
loop condition then tests the FP64 average rounded to float, as the device
does."""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.collection import ParticleCollection
from fluidsolver_amd.solver import Scene

pytestmark = pytest.mark.gpu


def _rows(off, cols, rows):
    return {int(r): cols[off[r] : off[r + 1]].tolist() for r in rows}


def _oracle(O, desc, coll, quadratic):
    search = O.SEARCH_QUADRATIC2D if quadratic else O.SEARCH_HASHED2D
    return O.OracleScene(desc, coll.copy(), search=search, mode=O.MODE_FP32_R64)


def _shuffled(coll, seed):
    """Rows in random order: grid cells then hold members whose insertion order
    (by entry, then index) differs from position order, so list order matters."""
    perm = np.random.default_rng(seed).permutation(coll.size())
    c = coll.copy()
    c.movement = c.movement[perm].copy()
    c.particle_data = c.particle_data[perm].copy()
    c.info = c.info[perm].copy()
    c.external = c.external[perm].copy()
    return c


def _desc(solver, quadratic, **kw):
    d = abi.default_desc(2, solver, **kw)
    d.search = abi.SEARCH_QUADRATIC_2D if quadratic else abi.SEARCH_HASHED_2D
    if solver == abi.SOLVER_SESPH:
        d.sesph_stiffness_k = 1000.0
    return d


def _compare(sc, orc, coll, steps, label, iisph):
    fluid = np.nonzero(coll.type == abi.TYPE_NORMAL)[0]
    for step in range(steps):
        sc.find_neighbors()
        orc.find_neighbors()
        goff, gcols = sc.export_neighbors()
        ooff, ocols = orc.export_neighbors()
        assert _rows(goff, gcols, fluid) == _rows(ooff, ocols, fluid), (label, step)
        st = sc.simulate()
        ost = orc.simulate()
        assert st.iterations == ost.iterations, (label, step)
        assert st.actual_time_step == ost.actual_time_step, (label, step)
        assert st.average_density_error == ost.average_density_error, (label, step)
        g = sc.download(coll.copy())
        r = orc.download()
        if iisph:
            np.testing.assert_array_equal(sc.download_iisph()[fluid], orc.download_iisph()[fluid],
                                          err_msg=f"{label} iisph step {step}")
        np.testing.assert_array_equal(g.movement, r.movement, err_msg=f"{label} movement step {step}")
        np.testing.assert_array_equal(g.particle_data, r.particle_data, err_msg=f"{label} pdata step {step}")
        np.testing.assert_array_equal(g.external, r.external, err_msg=f"{label} external step {step}")


@pytest.mark.parametrize("quadratic", [False, True])
@pytest.mark.parametrize("solver", [abi.SOLVER_IISPH, abi.SOLVER_SESPH])
def test_2d_steps_bitwise(gpu_context, oracle_lib, solver, quadratic):
    """A falling block with dt = 0.01 crosses cells every few steps (about 70 cell
    changes in 25 steps); with shuffled rows the hashed grid's insertion order
    then departs from index order.  List order is checked through the sums."""
    coll = _shuffled(scenes.dam_scene((12, 16), (40, 24), 3, seed=11), 5)
    desc = _desc(solver, quadratic, timestep=0.01)
    label = f"2d solver={solver} quadratic={quadratic}"
    print("case", label, coll.size(), flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    orc = _oracle(oracle_lib, desc, coll, quadratic)
    _compare(sc, orc, coll, 25, label, solver == abi.SOLVER_IISPH)
    sc.close()


def test_2d_inactive_and_external(gpu_context, oracle_lib):
    """Inactive particles stay in the hashed grid and in the lists (the solver
    skips them); external accelerations enter the predicted velocity."""
    coll = _shuffled(scenes.dam_scene((10, 10), (30, 20), 3, seed=12), 6)
    t = coll.info["type"]
    fluid = np.nonzero(t == abi.TYPE_NORMAL)[0]
    t[fluid[::7]] = abi.TYPE_INACTIVE
    coll.external[fluid[1::5]] = np.array([0.5, -0.25], np.float32)
    desc = _desc(abi.SOLVER_IISPH, False, timestep=0.005)
    print("case 2d inactive+external", coll.size(), flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    orc = _oracle(oracle_lib, desc, coll, False)
    _compare(sc, orc, coll, 10, "2d inactive", True)
    sc.close()


def test_2d_grid_count_truncation(gpu_context, oracle_lib):
    """Above 65535 particles the grid update visits only (count mod 65536)
    particles (uint16 particleAmount_t); the rest are never neighbors."""
    f = scenes.fluid_block((30, 30), 13)
    pad = scenes.lattice((317, 205), origin=(1000, 1000)).astype(np.float32)
    pos = np.concatenate([f, pad]).astype(np.float32)
    types = np.concatenate([np.zeros(len(f), np.uint8), np.ones(len(pad), np.uint8)])
    coll = ParticleCollection.from_arrays(2, pos, types)
    assert 65536 < coll.size() < 65536 + len(f)
    desc = _desc(abi.SOLVER_IISPH, False)
    print("case 2d uint16 count", coll.size(), flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    orc = _oracle(oracle_lib, desc, coll, False)
    _compare(sc, orc, coll, 2, "2d truncation", True)
    sc.close()


def test_2d_dam_500_steps(gpu_context, oracle_lib):
    """64x64 fluid in a 200x100 box with a 3-layer boundary, dt 0.001, 500 steps
    (integer-abs loop condition): every step bit-identical to the oracle; against
    the FP32-sum oracle the iteration counts agree within one."""
    O = oracle_lib
    coll = scenes.config1_dam2d(seed=1)
    desc = _desc(abi.SOLVER_IISPH, False)
    print("case 2d dam 500 steps", coll.size(), flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    exact = _oracle(O, desc, coll, False)
    ref = O.OracleScene(desc, coll.copy(), search=O.SEARCH_HASHED2D, mode=O.MODE_FP32)
    for step in range(500):
        g = sc.step(1)
        e = exact.step(1)
        r = ref.step(1)
        assert g.iterations == e.iterations and g.average_density_error == e.average_density_error, step
        assert abs(int(g.iterations) - int(r.iterations)) <= 1, step
        if step % 100 == 0:
            print(f"step {step}: iterations {g.iterations} err {g.average_density_error:.3e}", flush=True)
    np.testing.assert_array_equal(sc.download(coll.copy()).movement, exact.download().movement)
    sc.close()
