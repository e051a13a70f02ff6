"""The compressed search's sort paths on the GPU, against the oracle's stable sort.

A scene whose grid has few enough cells sorts each particle by its cell's rank
with a counting sort: buckets of up to 12 particles are ordered per lane, up to
64 by a wave, and a bucket above 64 particles makes the scene redo its step
with the radix sort.  Cells of 40 and of 80 particles (a packed cluster of
boundary particles away from the fluid) take the two larger paths; steps must
stay bit-identical to the oracle, alone and in a batch with scenes that do not
retry.
"""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.collection import ParticleCollection
from fluidsolver_amd.solver import Scene

pytestmark = pytest.mark.gpu


def _with_cluster(seed, count):
    """A small dam plus `count` boundary particles packed inside one cell of the
    box's free corner (all within the search radius of each other)."""
    dam = scenes.dam_scene((8, 8, 8), (30, 15, 15), 3, seed=seed)
    rng = np.random.default_rng(seed)
    pts = rng.uniform(24.2, 25.6, size=(count, 3)).astype(np.float32)
    pts[:, 1] = rng.uniform(10.2, 11.6, size=count).astype(np.float32)
    cl = ParticleCollection.from_arrays(3, pts, np.full(count, abi.TYPE_BOUNDARY, np.uint8))
    return ParticleCollection.concatenate([dam, cl])


def _same(sc, orc, coll, steps, label):
    for step in range(steps):
        st = sc.step(1)
        ost = orc.step(1)
        assert st.status == abi.FSA_OK, (label, step, st.status)
        assert st.iterations == ost.iterations, (label, step)
        g = sc.download(coll.copy())
        r = orc.download()
        np.testing.assert_array_equal(g.movement, r.movement, err_msg=f"{label} step {step}")
        np.testing.assert_array_equal(g.particle_data, r.particle_data, err_msg=f"{label} step {step}")


@pytest.mark.parametrize("count", [40, 80])
def test_dense_cell_sort_paths_match_oracle(gpu_context, oracle_lib, count):
    O = oracle_lib
    coll = _with_cluster(7, count)
    desc = abi.default_desc(3, abi.SOLVER_IISPH)
    print(f"case dense cell {count} n={coll.size()}", flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    orc = O.OracleScene(desc, coll.copy(), mode=O.MODE_FP32_R64)
    sc.find_neighbors()
    orc.find_neighbors()
    go, gc = sc.export_neighbors()
    oo, oc = orc.export_neighbors()
    fl = np.nonzero(coll.type == abi.TYPE_NORMAL)[0]
    for r in fl[:: max(1, len(fl) // 50)]:
        assert gc[go[r]:go[r + 1]].tolist() == oc[oo[r]:oo[r + 1]].tolist(), r
    _same(sc, orc, coll, 4, f"cluster {count}")
    sc.close()


def test_retry_in_a_batch(gpu_context, oracle_lib):
    """A scene that retries with the radix sort shares a batch with scenes that
    do not: every scene equals the oracle."""
    O = oracle_lib
    colls = [_with_cluster(11, 80), scenes.dam_scene((9, 7, 8), (27, 14, 14), 3, seed=12), _with_cluster(13, 30)]
    desc = abi.default_desc(3, abi.SOLVER_IISPH)
    print("case retry in a batch", flush=True)
    batch = [Scene(gpu_context, desc, c.copy()) for c in colls]
    orcs = [O.OracleScene(desc, c.copy(), mode=O.MODE_FP32_R64) for c in colls]
    for step in range(4):
        st = gpu_context.batch_step(batch, 1)
        for q, (sc, orc, c) in enumerate(zip(batch, orcs, colls)):
            ost = orc.step(1)
            assert st[q].status == abi.FSA_OK and st[q].iterations == ost.iterations, (step, q)
            g = sc.download(c.copy())
            r = orc.download()
            np.testing.assert_array_equal(g.movement, r.movement, err_msg=f"step {step} scene {q}")
            np.testing.assert_array_equal(g.particle_data, r.particle_data, err_msg=f"step {step} scene {q}")
    for sc in batch:
        sc.close()
