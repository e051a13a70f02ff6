"""The reference's own scenario files run on the GPU, bit for bit against the
oracle: JSON version 2 read by the library, particles from the LZ4-framed
.data files, the solver / search / settings the file names; the dynamic CFL
timestep generator on the device (single scenes, batches, 2D); snapshots of
the device state written as .data files and read back.

The oracle runs with FP64 convergence sums (the device's reduction)."""
import os

import numpy as np
import pytest

from fluidsolver_amd import abi
from fluidsolver_amd import scenario as S
from fluidsolver_amd.solver import Scene

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenarios")


def oracle_search(O, search):
    return {abi.SEARCH_COMPRESSED_3D: O.SEARCH_COMPRESSED3D, abi.SEARCH_HASHED_3D: O.SEARCH_HASHED3D,
            abi.SEARCH_QUADRATIC_3D: O.SEARCH_QUADRATIC3D, abi.SEARCH_HASHED_2D: O.SEARCH_HASHED2D,
            abi.SEARCH_QUADRATIC_2D: O.SEARCH_QUADRATIC2D}[search]


def compare(sc, orc, coll, steps, label):
    iisph = sc.desc.solver == abi.SOLVER_IISPH
    dts = []
    for step in range(steps):
        assert sc.next_timestep() == orc.next_timestep(), (label, step)
        st = sc.step(1)
        ost = orc.step(1)
        assert st.desired_time_step == ost.desired_time_step, (label, step)
        assert st.actual_time_step == ost.actual_time_step, (label, step)
        assert st.iterations == ost.iterations, (label, step)
        assert st.average_density_error == ost.average_density_error, (label, step)
        dts.append(st.desired_time_step)
        if step % 5 == 4 or step == steps - 1:
            g = sc.download(coll.copy())
            r = orc.download()
            np.testing.assert_array_equal(g.movement, r.movement, err_msg=f"{label} movement {step}")
            np.testing.assert_array_equal(g.particle_data, r.particle_data, err_msg=f"{label} pdata {step}")
            if iisph:
                fl = coll.type == abi.TYPE_NORMAL
                np.testing.assert_array_equal(sc.download_iisph()[fl], orc.download_iisph()[fl], err_msg=label)
    return dts


SCENARIOS = [
    ("3d-small-breaking-dam.json", 40),           # SESPH-3D, hashed-3d, k 1e5
    ("filled-container.json", 30),                # IISPH-3D, hashed-3d, single-layer boundary
    ("boundary-test-multi-layer-3d.json", 60),    # SESPH-3D, hashed-3d, dt 0.01
    ("breaking-dam-1-layer-boundary.json", 30),   # SESPH-3D, hashed-3d, single-layer gamma 2 / 3
    ("box.json", 60),                             # SESPH 2D, quadratic-dynamic-allocated
    ("boundary-test.json", 60),                   # SESPH 2D, quadratic-dynamic-allocated
]


@pytest.mark.parametrize("name,steps", SCENARIOS)
def test_reference_scenarios_bitwise(gpu_context, oracle_lib, name, steps):
    O = oracle_lib
    scn = S.load_scenario(os.path.join(GOLD, name))
    coll = scn.collection
    print("case scenario", name, coll.size(), steps, flush=True)
    sc = Scene(gpu_context, scn.desc, coll.copy())
    orc = O.OracleScene(scn.desc, coll.copy(), search=oracle_search(O, scn.desc.search), mode=O.MODE_FP32_R64)
    compare(sc, orc, coll, steps, name)
    sc.close()


@pytest.mark.parametrize("name,steps,iisph", [("3d-small-breaking-dam.json", 40, False),
                                              ("3d-small-breaking-dam.json", 30, True),
                                              ("box.json", 60, False)])
def test_dynamic_cfl_bitwise(gpu_context, oracle_lib, name, steps, iisph):
    """DynamicCflTimestepGenerator: the desired step follows the state (for
    IISPH the stored acceleration is the pressure acceleration); the IISPH
    in-step correction falls back to max_timestep."""
    O = oracle_lib
    scn = S.load_scenario(os.path.join(GOLD, name))
    d = scn.desc
    if iisph:
        d.solver = abi.SOLVER_IISPH  # IISPHSettings3D defaults from the scenario's desc
    d.timestep_generator = abi.TIMESTEP_DYNAMIC_CFL
    # the reference's default upper bound.  From rest, IISPH's stored (pressure)
    # acceleration stays at or below FLT_EPSILON at tiny steps, which keeps the
    # generator at min_timestep: a larger minimum lets the dam build pressure.
    d.max_timestep, d.min_timestep = 0.02, (0.001 if iisph else 0.00001)
    coll = scn.collection
    print("case dynamic cfl", name, "iisph" if iisph else "", flush=True)
    sc = Scene(gpu_context, d, coll.copy())
    orc = O.OracleScene(d, coll.copy(), search=oracle_search(O, d.search), mode=O.MODE_FP32_R64)
    dts = compare(sc, orc, coll, steps, "dynamic " + name)
    if iisph:
        assert len(set(dts)) > 1, dts
    else:
        inner = [t for t in dts if t not in (np.float32(0.02), np.float32(0.00001))]
        assert len(set(inner)) > 3, dts  # the step follows the flow, not only the clamps
    sc.close()


def test_dynamic_cfl_batched(gpu_context, oracle_lib):
    """Scenes with their own generators advanced in one batch: each scene's
    desired step comes from its own state."""
    O = oracle_lib
    scn = S.load_scenario(os.path.join(GOLD, "filled-container.json"))
    coll = scn.collection
    descs = []
    for k, (gen, mx) in enumerate([(abi.TIMESTEP_DYNAMIC_CFL, 0.004), (abi.TIMESTEP_CONSTANT, 0.0),
                                   (abi.TIMESTEP_DYNAMIC_CFL, 0.002)]):
        d = abi.SceneDesc.from_buffer_copy(bytes(scn.desc))
        d.timestep_generator = gen
        if mx:
            d.max_timestep, d.min_timestep = mx, 0.0001
        d.iisph_gamma = 0.7 + 0.1 * k
        descs.append(d)
    print("case dynamic cfl batched", flush=True)
    scs = [Scene(gpu_context, d, coll.copy()) for d in descs]
    orcs = [O.OracleScene(d, coll.copy(), search=O.SEARCH_HASHED3D, mode=O.MODE_FP32_R64) for d in descs]
    for step in range(20):
        sts = gpu_context.batch_step(scs, 1)
        for q, (st, orc) in enumerate(zip(sts, orcs)):
            ost = orc.step(1)
            assert (st.desired_time_step, st.actual_time_step, st.iterations) == (
                ost.desired_time_step, ost.actual_time_step, ost.iterations), (q, step)
    for q, (sc, orc) in enumerate(zip(scs, orcs)):
        np.testing.assert_array_equal(sc.download(coll.copy()).movement, orc.download().movement, err_msg=str(q))
        sc.close()


def test_snapshot_dump_round_trip(gpu_context, oracle_lib, tmp_path):
    """Snapshots of the device state as .data files (FluidConsole --dump):
    the written file reads back to the downloaded state, which equals the
    oracle's, and a scenario saved with it reloads and continues identically."""
    O = oracle_lib
    scn = S.load_scenario(os.path.join(GOLD, "3d-small-breaking-dam.json"))
    coll = scn.collection
    print("case snapshot", flush=True)
    sc = Scene(gpu_context, scn.desc, coll.copy())
    orc = O.OracleScene(scn.desc, coll.copy(), search=O.SEARCH_HASHED3D, mode=O.MODE_FP32_R64)
    sc.step(10)
    orc.step(10)
    snap = sc.download(coll.copy())
    S.save_scenario(tmp_path / "snap.json", scn, collection=snap, particles="snap.data")
    back = S.load_scenario(tmp_path / "snap.json")
    assert S.records_equal(back.collection, snap)
    assert S.records_equal(back.collection, orc.download())
    # continuing from the snapshot: a fresh scene from the file follows the running one
    sc2 = Scene(gpu_context, back.desc, back.collection.copy())
    orc2 = O.OracleScene(back.desc, back.collection.copy(), search=O.SEARCH_HASHED3D, mode=O.MODE_FP32_R64)
    sc2.step(5)
    orc2.step(5)
    np.testing.assert_array_equal(sc2.download(coll.copy()).movement, orc2.download().movement)
    sc.close()
    sc2.close()
