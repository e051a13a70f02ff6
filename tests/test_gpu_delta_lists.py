"""Delta-coded neighbor lists (FSA_LIST_DELTA) on the GPU.

The solver walks of a DELTA scene read each list as 16-bit deltas with an
escape stream for the first entry and deltas beyond 32767; the lists and every
result must be the same as with the 32-bit lists (FSA_LIST_COMPRESSED), bit for
bit, and the same as the oracle.  The cases cover lists with no escape beyond
the first entry, lists with several (a block around the origin, whose 27
neighbor cells fall into the 8 sign octants of the Morton order), and a scene
whose lists need more escapes than the escape rows hold, where the walks fall
back to the 32-bit lists for that search.
"""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.collection import ParticleCollection
from fluidsolver_amd.solver import Scene

pytestmark = pytest.mark.gpu


def _desc(solver, lists):
    d = abi.default_desc(3, solver, neighbor_lists=lists)
    if solver == abi.SOLVER_SESPH:
        d.sesph_stiffness_k = 1000.0
    return d


def _same_state(a, b, label):
    np.testing.assert_array_equal(a.movement, b.movement, err_msg=f"movement {label}")
    np.testing.assert_array_equal(a.particle_data, b.particle_data, err_msg=f"particle data {label}")


def _run_pair(ctx, coll, solver, steps, label):
    """Steps the same collection with 32-bit and with delta-coded lists; every step's stats and state must agree."""
    plain = Scene(ctx, _desc(solver, abi.LIST_COMPRESSED), coll.copy())
    delta = Scene(ctx, _desc(solver, abi.LIST_DELTA), coll.copy())
    for step in range(steps):
        sp = plain.step(1)
        sd = delta.step(1)
        assert sp.status == sd.status == abi.FSA_OK, (label, step, sp.status, sd.status)
        assert sp.iterations == sd.iterations, (label, step)
        assert sp.average_density_error == sd.average_density_error, (label, step)
        assert sp.actual_time_step == sd.actual_time_step, (label, step)
        _same_state(plain.download(coll.copy()), delta.download(coll.copy()), f"{label} step {step}")
    return plain, delta


@pytest.mark.parametrize("solver", [abi.SOLVER_IISPH, abi.SOLVER_SESPH])
def test_delta_steps_equal_oracle(gpu_context, oracle_lib, solver):
    O = oracle_lib
    coll = scenes.dam_scene((10, 10, 10), (30, 15, 15), 3, seed=3)
    desc = _desc(solver, abi.LIST_DELTA)
    print(f"case solver={solver} n={coll.size()}", flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    orc = O.OracleScene(desc, coll.copy(), mode=O.MODE_FP32_R64)
    for step in range(5):
        st = sc.step(1)
        ost = orc.step(1)
        assert st.iterations == ost.iterations, step
        assert st.average_density_error == ost.average_density_error, step
        _same_state(sc.download(coll.copy()), orc.download(), f"step {step}")
    stats = sc.neighbor_stats()
    # codes are 2 bytes per entry (rounded to pairs) plus 4 per escape: well under 4 bytes per entry
    assert stats["device_list_bytes"] < 3 * stats["total_neighbors"], stats
    sc.close()


@pytest.mark.parametrize("solver", [abi.SOLVER_IISPH, abi.SOLVER_SESPH])
def test_delta_lists_with_many_escapes(gpu_context, solver):
    """A fluid block around the origin: the lists of particles near the origin
    jump between the sign octants of the Morton order (more than 32767 rows apart)."""
    side = 70
    coll = scenes.dam_scene((side, side, side), (side + 10, side + 10, side + 10), 3, seed=5)
    coll.position[:] -= np.float32(side / 2.0)
    print(f"case solver={solver} n={coll.size()}", flush=True)
    plain, delta = _run_pair(gpu_context, coll, solver, 3, "origin block")
    off, cols = delta.export_neighbors()
    # some list does jump by more than the 16-bit codes hold, several times
    rows = np.repeat(np.arange(coll.size()), np.diff(off).astype(np.int64))
    plain_off, plain_cols = plain.export_neighbors()
    np.testing.assert_array_equal(off, plain_off)
    np.testing.assert_array_equal(cols, plain_cols)
    del rows
    sd = delta.neighbor_stats()
    sp = plain.neighbor_stats()
    assert sd["total_neighbors"] == sp["total_neighbors"]
    assert sd["reference_codec_bytes"] == sp["reference_codec_bytes"]
    assert sd["device_list_bytes"] < sp["device_list_bytes"]
    plain.close()
    delta.close()


def _interleaved_cell(n_fluid, gap, seed):
    """n_fluid fluid particles near one corner of the cell [0, 2)^3 and, between
    consecutive ones in collection order, `gap` boundary particles clustered at
    the opposite corner (more than the search radius away): every fluid list
    holds all n_fluid fluid particles, `gap` + 1 rows apart."""
    rng = np.random.default_rng(seed)
    fl = rng.uniform(0.05, 0.55, size=(n_fluid, 3)).astype(np.float32)
    parts_pos, parts_type = [], []
    for q in range(n_fluid):
        parts_pos.append(fl[q : q + 1])
        parts_type.append(np.array([abi.TYPE_NORMAL], np.uint8))
        if q + 1 < n_fluid:
            parts_pos.append(rng.uniform(1.95, 1.99, size=(gap, 3)).astype(np.float32))
            parts_type.append(np.full(gap, abi.TYPE_BOUNDARY, np.uint8))
    return ParticleCollection.from_arrays(3, np.concatenate(parts_pos), np.concatenate(parts_type))


@pytest.mark.parametrize("n_fluid", [5, 20])
def test_delta_escape_queue_and_fallback(gpu_context, n_fluid):
    """5 fluid particles: 5 escapes per list (loaded as the walk reaches them);
    20: more escapes than the escape rows hold, so that search's walks read the
    32-bit lists.  Both must equal the 32-bit walk bit for bit."""
    coll = _interleaved_cell(n_fluid, 33000, seed=n_fluid)
    print(f"case n_fluid={n_fluid} n={coll.size()}", flush=True)
    fresh = Scene(gpu_context, _desc(abi.SOLVER_IISPH, abi.LIST_DELTA), coll.copy())
    fresh.find_neighbors()
    fluid = np.nonzero(coll.type == abi.TYPE_NORMAL)[0]
    for r in fluid[:3]:
        assert fresh.get_neighbors(int(r)) == [int(x) for x in fluid], r
    sd = fresh.neighbor_stats()
    assert sd["total_neighbors"] == n_fluid * n_fluid, sd
    if n_fluid <= 15:
        # every escape stored (the first entry and n_fluid - 1 deltas of 33001):
        # 4 bytes per code pair plus 4 per escape
        assert sd["device_list_bytes"] == n_fluid * (4 * ((n_fluid + 1) // 2) + 4 * n_fluid), sd
    else:
        # the walks read the 32-bit lists for this search
        assert sd["device_list_bytes"] == 4 * sd["total_neighbors"], sd
    fresh.close()
    plain, delta = _run_pair(gpu_context, coll, abi.SOLVER_IISPH, 3, f"interleaved {n_fluid}")
    plain.close()
    delta.close()


def test_delta_batch_equals_alone(gpu_context):
    """A batch of delta scenes walks delta lists; a batch mixing delta and 32-bit
    scenes walks the 32-bit lists; both equal each scene stepped alone."""
    colls = [scenes.dam_scene(d, i, 3, seed=s) for d, i, s in
             [((8, 8, 8), (24, 12, 12), 11), ((12, 9, 7), (30, 16, 14), 12), ((6, 6, 6), (18, 9, 9), 13)]]
    for mix in (False, True):
        lists = [abi.LIST_DELTA, abi.LIST_COMPRESSED if mix else abi.LIST_DELTA, abi.LIST_DELTA]
        batch = [Scene(gpu_context, _desc(abi.SOLVER_IISPH, l), c.copy()) for l, c in zip(lists, colls)]
        alone = [Scene(gpu_context, _desc(abi.SOLVER_IISPH, abi.LIST_COMPRESSED), c.copy()) for c in colls]
        print(f"case mix={mix}", flush=True)
        for step in range(3):
            st = gpu_context.batch_step(batch, 1)
            for q, (sb, sa, c) in enumerate(zip(batch, alone, colls)):
                ss = sa.step(1)
                assert st[q].status == abi.FSA_OK and ss.status == abi.FSA_OK
                assert st[q].iterations == ss.iterations, (mix, step, q)
                _same_state(sb.download(c.copy()), sa.download(c.copy()), f"mix={mix} step {step} scene {q}")
        for s in batch + alone:
            s.close()


def test_delta_mode_cannot_change(gpu_context):
    coll = scenes.dam_scene((6, 6, 6), (18, 9, 9), 3, seed=1)
    sc = Scene(gpu_context, _desc(abi.SOLVER_IISPH, abi.LIST_DELTA), coll.copy())
    with pytest.raises(Exception):
        sc.set_desc(_desc(abi.SOLVER_IISPH, abi.LIST_COMPRESSED))
    sc.set_desc(_desc(abi.SOLVER_SESPH, abi.LIST_DELTA))  # other settings may change
    plain = Scene(gpu_context, _desc(abi.SOLVER_IISPH, abi.LIST_PLAIN), coll.copy())
    with pytest.raises(Exception):
        plain.set_desc(_desc(abi.SOLVER_IISPH, abi.LIST_DELTA))
    sc.close()
    plain.close()
