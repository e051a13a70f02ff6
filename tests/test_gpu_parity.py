"""GPU parity: the device path against the CPU oracle, through the C ABI.

The device keeps the reference's particle order, neighbor order and FP32
evaluation order, so against the oracle in MODE_FP32_R64 (reference FP32
arithmetic, FP64 convergence sums like the device) every state array must be
bit-identical, step after step.  Neighbor sets are compared by caller index.
"""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.solver import Scene

pytestmark = pytest.mark.gpu


def _fluid_rows(coll):
    return np.nonzero(coll.type == abi.TYPE_NORMAL)[0]


def _csr_rows(off, cols, rows):
    return {int(r): cols[off[r] : off[r + 1]].tolist() for r in rows}


@pytest.mark.parametrize("dims,interior,seed", [((6, 6, 6), (18, 9, 9), 1), ((12, 10, 8), (30, 16, 14), 2)])
def test_neighbor_sets_match_oracle(gpu_context, oracle_lib, dims, interior, seed):
    O = oracle_lib
    coll = scenes.dam_scene(dims, interior, 3, seed=seed)
    desc = abi.default_desc(3, abi.SOLVER_IISPH)
    print(f"case dims={dims} n={coll.size()}", flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    sc.find_neighbors()
    goff, gcols = sc.export_neighbors()
    orc = O.OracleScene(desc, coll.copy(), search=O.SEARCH_QUADRATIC3D)
    orc.find_neighbors()
    ooff, ocols = orc.export_neighbors()
    rows = _fluid_rows(coll)
    g = _csr_rows(goff, gcols, rows)
    o = _csr_rows(ooff, ocols, rows)
    assert g == o


@pytest.mark.parametrize("solver", [abi.SOLVER_IISPH, abi.SOLVER_SESPH])
def test_steps_bitwise_equal_oracle(gpu_context, oracle_lib, solver):
    O = oracle_lib
    coll = scenes.dam_scene((10, 10, 10), (30, 15, 15), 3, seed=3)
    desc = abi.default_desc(3, solver)
    if solver == abi.SOLVER_SESPH:
        desc.sesph_stiffness_k = 1000.0
    print(f"case solver={solver} n={coll.size()}", flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    orc = O.OracleScene(desc, coll.copy(), mode=O.MODE_FP32_R64)
    fluid = coll.type == abi.TYPE_NORMAL
    for step in range(5):
        st = sc.step(1)
        ost = orc.step(1)
        g = sc.download(coll.copy())
        r = orc.download()
        if solver == abi.SOLVER_IISPH:
            gi = sc.download_iisph()[fluid]
            oi = orc.download_iisph()[fluid]
            names = ["vstar_x", "vstar_y", "vstar_z", "anp_x", "anp_y", "anp_z", "source", "diag", "err"]
            for c, name in enumerate(names):
                bad = np.nonzero(gi[:, c] != oi[:, c])[0]
                assert bad.size == 0, (
                    f"step {step} {name}: {bad.size} differ, e.g. gpu {gi[bad[:3], c]} oracle {oi[bad[:3], c]}"
                )
        for c, name in enumerate(["density"]):
            bad = np.nonzero(g.density != r.density)[0]
            assert bad.size == 0, f"step {step} density: {bad.size} differ, e.g. {g.density[bad[:3]]} vs {r.density[bad[:3]]}"
        assert st.iterations == ost.iterations, step
        assert st.actual_time_step == ost.actual_time_step, step
        assert st.average_density_error == ost.average_density_error, step
        np.testing.assert_array_equal(g.movement, r.movement, err_msg=f"movement step {step}")
        np.testing.assert_array_equal(g.particle_data, r.particle_data, err_msg=f"particle data step {step}")
