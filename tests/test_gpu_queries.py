"""Position queries on the GPU against the oracle, bit for bit:
NeighborhoodInterface::get_neighbors(position) in each 3D search's order
(after a step: the cells of the last search with the moved positions),
SensorPlane values of each type, and the raytracer's volume probe.

Inactive particles are not resident on the device, so they are never
reported; the reference's quadratic position iterator lists them (the oracle
follows it), so for that search the oracle's rows are filtered before the
comparison, and SensorPlane, which weighs every listed particle, is compared
on scenes without inactive particles there."""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.solver import FsaError, Scene

pytestmark = pytest.mark.gpu

f32 = np.float32
SEARCH = {"compressed": (abi.SEARCH_COMPRESSED_3D, "SEARCH_COMPRESSED3D"),
          "hashed": (abi.SEARCH_HASHED_3D, "SEARCH_HASHED3D"),
          "quadratic": (abi.SEARCH_QUADRATIC_3D, "SEARCH_QUADRATIC3D")}


def scene(seed, inactive_every=0):
    """A dam break with every cell coordinate positive (cell (-1,-1,-1) shares
    the compressed search's inactive sentinel key, a reference defect the device
    does not reproduce)."""
    c = scenes.dam_scene((6, 7, 6), (16, 11, 11), 2, seed=seed)
    c.position[:] += f32(10.0)
    if inactive_every:
        c.info["type"][np.nonzero(c.type == abi.TYPE_NORMAL)[0][::inactive_every]] = abi.TYPE_INACTIVE
    return c


def pair(ctx, O, coll, search, solver=abi.SOLVER_IISPH):
    d = abi.default_desc(3, solver, timestep=0.004, sesph_stiffness_k=1000.0)
    d.search = SEARCH[search][0]
    sc = Scene(ctx, d, coll.copy())
    orc = O.OracleScene(d, coll.copy(), search=getattr(O, SEARCH[search][1]), mode=O.MODE_FP32_R64)
    return sc, orc


def queries(coll, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = coll.position.min(0) - 3.0, coll.position.max(0) + 3.0
    q = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    q[:16] = coll.position[rng.choice(coll.size(), 16, replace=False)]
    q[16] = [-1e6, 5.0, 5.0]
    q[17] = [1e9, -1e9, 1e9]
    return q


def oracle_rows(orc, q, drop_inactive):
    off, rows = orc.query_neighbors(q)
    if not drop_inactive:
        return off, rows
    keep = orc.download().type[rows.astype(np.int64)] != abi.TYPE_INACTIVE
    counts = [int(keep[off[k]:off[k + 1]].sum()) for k in range(len(q))]
    return np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64), rows[keep]


def volume_array(samples):
    return np.array([[s.density, s.boundary_only_density, s.fluid_only_density, *s.fluid_only_normal,
                      *s.boundary_only_normal] for s in samples], dtype=np.float32)


@pytest.mark.parametrize("search", ["compressed", "hashed", "quadratic"])
@pytest.mark.parametrize("inactive", [0, 9])
def test_queries_bitwise(gpu_context, oracle_lib, search, inactive):
    coll = scene(81 + inactive, inactive_every=inactive)
    print("case queries", search, "inactive", inactive, coll.size(), flush=True)
    sc, orc = pair(gpu_context, oracle_lib, coll, search)
    for step in range(6):
        st, ost = sc.step(1), orc.step(1)
        assert (st.actual_time_step, st.iterations) == (ost.actual_time_step, ost.iterations), step
    got = sc.download(coll.copy())
    np.testing.assert_array_equal(got.movement, orc.download().movement)
    q = queries(got, 600, 9)
    off, rows = sc.query_neighbors(q)
    roff, rrows = oracle_rows(orc, q, drop_inactive=(search == "quadratic" and inactive > 0))
    np.testing.assert_array_equal(off, roff)
    np.testing.assert_array_equal(rows, rrows)
    assert rows.size > 0 and np.any(np.diff(off) == 0)
    np.testing.assert_array_equal(volume_array(sc.evaluate_volume(q)), volume_array(orc.evaluate_volume(q)))
    if search == "quadratic" and inactive:
        return  # SensorPlane weighs the inactive particles the reference's quadratic iterator lists
    for kind in (abi.SENSOR_PLANE_DENSITY, abi.SENSOR_PLANE_VELOCITY, abi.SENSOR_PLANE_PRESSURE):
        plane = abi.sensor_plane(origin=(9.0, 9.5, 12.3), span_x=(1.0, 0.0, 0.2), span_y=(0.0, 2.0, 0.0),
                                 width=16.0, height=11.0, samples_x=41, samples_y=23, sensor_type=kind)
        g, r = sc.sensor_plane(plane), orc.sensor_plane(plane)
        np.testing.assert_array_equal(g, r, err_msg=f"sensor plane {kind}")
        if kind != abi.SENSOR_PLANE_PRESSURE:  # pressures may all be clamped to 0 in a holed fluid
            assert np.any(g != 0.0), kind
    sc.close()


def test_queries_after_search_and_in_batches(gpu_context, oracle_lib):
    """Queries after an explicit search (the raytracer's state) and on scenes
    stepped in one batch."""
    O = oracle_lib
    coll = scene(90)
    print("case queries batched", flush=True)
    pairs = [pair(gpu_context, O, coll, "compressed"), pair(gpu_context, O, coll, "hashed", abi.SOLVER_SESPH)]
    scs = [p[0] for p in pairs]
    for step in range(5):
        gpu_context.batch_step(scs, 1)
        for _, orc in pairs:
            orc.step(1)
    q = queries(coll, 300, 11)
    for k, (sc, orc) in enumerate(pairs):
        for phase in ("after step", "after search"):
            off, rows = sc.query_neighbors(q)
            roff, rrows = orc.query_neighbors(q)
            np.testing.assert_array_equal(off, roff, err_msg=f"{k} {phase}")
            np.testing.assert_array_equal(rows, rrows, err_msg=f"{k} {phase}")
            np.testing.assert_array_equal(volume_array(sc.evaluate_volume(q)), volume_array(orc.evaluate_volume(q)))
            sc.find_neighbors()
            orc.find_neighbors()
        sc.close()


def test_query_state_and_argument_errors(gpu_context):
    coll = scene(91)
    print("case query errors", flush=True)
    d = abi.default_desc(3, abi.SOLVER_IISPH, timestep=0.004)
    sc = Scene(gpu_context, d, coll.copy())
    q = np.zeros((1, 3), dtype=np.float32)
    with pytest.raises(FsaError) as e:
        sc.query_neighbors(q)  # no search yet
    assert e.value.code == abi.FSA_E_STATE
    sc.find_neighbors()
    off, rows = sc.query_neighbors(np.zeros((0, 3), dtype=np.float32))
    assert off.tolist() == [0] and rows.size == 0
    with pytest.raises(FsaError) as e:
        sc.query_neighbors(np.array([[np.nan, 0.0, 0.0]], dtype=np.float32))
    assert e.value.code == abi.FSA_E_INVALID
    with pytest.raises(FsaError) as e:
        sc.sensor_plane(abi.sensor_plane(samples_x=0))
    assert e.value.code == abi.FSA_E_INVALID
    sc.upload(coll.copy())
    with pytest.raises(FsaError) as e:
        sc.evaluate_volume(q)  # the layout changed since the search
    assert e.value.code == abi.FSA_E_STATE
    sc.close()
    c2 = scenes.dam_scene((5, 5), (12, 8), 2, seed=2)
    s2 = Scene(gpu_context, abi.default_desc(2, abi.SOLVER_SESPH), c2)
    s2.find_neighbors()
    with pytest.raises(FsaError) as e:
        s2.query_neighbors(q)
    assert e.value.code == abi.FSA_E_INVALID
    s2.close()
