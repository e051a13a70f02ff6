"""3D hashed and quadratic neighbor orders on the GPU against the oracle, and
re-uploads that keep the order the reference collection holds.

HashedNeighborhoodSearch3D walks the 27 cells dx-outer with each cell's
members ascending and re-sorts the collection with the unstable quick sort on
its 101st call; QuadraticNeighborhoodSearch3D lists ascending indices.  The
rows are shuffled so that index order and position order differ, which makes
list order (and so every FP32 sum) depend on the search's ordering rules."""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.solver import Scene

pytestmark = pytest.mark.gpu


def _shuffled(coll, seed):
    perm = np.random.default_rng(seed).permutation(coll.size())
    c = coll.copy()
    c.movement = c.movement[perm].copy()
    c.particle_data = c.particle_data[perm].copy()
    c.info = c.info[perm].copy()
    c.external = c.external[perm].copy()
    return c


def _compare_step(sc, orc, coll, label, step, iisph):
    fl = coll.type == abi.TYPE_NORMAL
    st = sc.step(1)
    ost = orc.step(1)
    assert st.iterations == ost.iterations, (label, step)
    assert st.actual_time_step == ost.actual_time_step, (label, step)
    assert st.average_density_error == ost.average_density_error, (label, step)
    g = sc.download(coll.copy())
    r = orc.download()
    np.testing.assert_array_equal(g.movement, r.movement, err_msg=f"{label} movement step {step}")
    np.testing.assert_array_equal(g.particle_data, r.particle_data, err_msg=f"{label} pdata step {step}")
    if iisph:
        np.testing.assert_array_equal(sc.download_iisph()[fl], orc.download_iisph()[fl], err_msg=f"{label} {step}")


@pytest.mark.parametrize("search", ["hashed", "quadratic"])
@pytest.mark.parametrize("solver", [abi.SOLVER_IISPH, abi.SOLVER_SESPH])
def test_3d_search_modes_bitwise(gpu_context, oracle_lib, search, solver):
    O = oracle_lib
    coll = _shuffled(scenes.dam_scene((5, 7, 5), (15, 10, 10), 2, seed=61), 8)
    t = coll.info["type"]
    t[np.nonzero(t == abi.TYPE_NORMAL)[0][::11]] = abi.TYPE_INACTIVE  # inactive rows take part in the re-sort
    d = abi.default_desc(3, solver, timestep=0.004)
    d.search = abi.SEARCH_HASHED_3D if search == "hashed" else abi.SEARCH_QUADRATIC_3D
    if solver == abi.SOLVER_SESPH:
        d.sesph_stiffness_k = 1000.0
    osearch = O.SEARCH_HASHED3D if search == "hashed" else O.SEARCH_QUADRATIC3D
    label = f"3d {search} solver={solver}"
    steps = 106 if search == "hashed" else 12  # hashed: across the re-sort on the 101st search
    print("case", label, coll.size(), steps, flush=True)
    sc = Scene(gpu_context, d, coll.copy())
    orc = O.OracleScene(d, coll.copy(), search=osearch, mode=O.MODE_FP32_R64)
    for step in range(steps):
        _compare_step(sc, orc, coll, label, step, solver == abi.SOLVER_IISPH)
    if search == "hashed":
        assert not np.array_equal(orc.internal_ids(), np.arange(coll.size()))  # the re-sort happened
    sc.close()


@pytest.mark.parametrize("search", ["compressed", "hashed"])
def test_reupload_keeps_collection_order(gpu_context, oracle_lib, search):
    """Rewriting particle data mid-run (an entity changing velocities) keeps
    the order the reference collection has by then, so later sorts start from
    it: results stay bit-identical."""
    O = oracle_lib
    coll = _shuffled(scenes.dam_scene((6, 6, 6), (18, 10, 10), 2, seed=62), 9)
    d = abi.default_desc(3, abi.SOLVER_IISPH, timestep=0.004)
    d.search = abi.SEARCH_HASHED_3D if search == "hashed" else abi.SEARCH_COMPRESSED_3D
    osearch = O.SEARCH_HASHED3D if search == "hashed" else O.SEARCH_COMPRESSED3D
    print("case reupload", search, flush=True)
    sc = Scene(gpu_context, d, coll.copy())
    orc = O.OracleScene(d, coll.copy(), search=osearch, mode=O.MODE_FP32_R64)
    total = 0
    for phase in range(3):
        for _ in range(40 if search == "hashed" else 4):
            _compare_step(sc, orc, coll, f"reupload {search}", total, True)
            total += 1
        g = sc.download(coll.copy())
        fl = np.nonzero(g.type == abi.TYPE_NORMAL)[0]
        g.velocity[fl[::3]] += np.float32([0.05, 0.0, -0.02])
        sc.upload(g)
        orc.upload(g)
    sc.close()


@pytest.mark.parametrize("search", ["hashed", "quadratic"])
@pytest.mark.parametrize("spacing", [0.75, 1.0])
def test_3d_search_mode_lists_dense_and_sparse(gpu_context, oracle_lib, search, spacing):
    """Neighbor lists of the hashed and quadratic orders, entry for entry.  At
    spacing 0.75 a cell holds up to 27 particles, so a quadratic cell's 27
    neighbor cells hold more candidates than the wave keeps at once and the
    cell's particles take the serial path; at spacing 1 they do not."""
    O = oracle_lib
    from fluidsolver_amd.collection import ParticleCollection

    rng = np.random.default_rng(63)
    n = 14
    f = (scenes.lattice((n, n, n)) * np.float32(spacing) + rng.uniform(-0.01, 0.01, (n**3, 3))).astype(np.float32)
    b = (scenes.box_shell((n, n, n), 2) * np.float32(spacing)).astype(np.float32)
    coll = _shuffled(ParticleCollection.from_arrays(
        3, np.concatenate([f, b]) + np.float32(10.0),
        np.concatenate([np.zeros(len(f), np.uint8), np.ones(len(b), np.uint8)])), 10)
    d = abi.default_desc(3, abi.SOLVER_IISPH)
    d.search = abi.SEARCH_HASHED_3D if search == "hashed" else abi.SEARCH_QUADRATIC_3D
    print("case dense lists", search, spacing, coll.size(), flush=True)
    sc = Scene(gpu_context, d, coll.copy())
    sc.find_neighbors()
    goff, gcols = sc.export_neighbors()
    orc = O.OracleScene(d, coll.copy(), search=O.SEARCH_HASHED3D if search == "hashed" else O.SEARCH_QUADRATIC3D)
    orc.find_neighbors()
    ooff, ocols = orc.export_neighbors()
    fl = np.nonzero(coll.type == abi.TYPE_NORMAL)[0]
    for r in fl:
        np.testing.assert_array_equal(gcols[goff[r]:goff[r + 1]], ocols[ooff[r]:ooff[r + 1]], err_msg=str(r))
    longest = max(int(goff[r + 1] - goff[r]) for r in fl)
    assert longest > (70 if spacing < 1.0 else 30), longest
    sc.close()


def test_hashed_resort_of_a_large_collection(gpu_context, oracle_lib):
    """The 101st hashed search re-sorts a collection large enough for the
    quick sort's parts to be partitioned on several threads (parts of 10000
    and more): the order, and so every later step, equals the oracle's."""
    O = oracle_lib
    coll = _shuffled(scenes.dam_scene((24, 24, 24), (40, 30, 30), 2, seed=64), 11)
    assert coll.size() > 25000
    d = abi.default_desc(3, abi.SOLVER_SESPH, timestep=0.004, sesph_stiffness_k=1000.0)
    d.search = abi.SEARCH_HASHED_3D
    print("case hashed resort large", coll.size(), flush=True)
    sc = Scene(gpu_context, d, coll.copy())
    orc = O.OracleScene(d, coll.copy(), search=O.SEARCH_HASHED3D, mode=O.MODE_FP32_R64)
    for step in range(104):
        st, ost = sc.step(1), orc.step(1)
        assert st.actual_time_step == ost.actual_time_step, step
    assert not np.array_equal(orc.internal_ids(), np.arange(coll.size()))  # the re-sort happened
    g, r = sc.download(coll.copy()), orc.download()
    np.testing.assert_array_equal(g.movement, r.movement)
    np.testing.assert_array_equal(g.particle_data, r.particle_data)
    sc.close()


def _fluid_lists(off, cols, fl):
    """The entries of rows fl of a CSR list array, concatenated, and their lengths."""
    off = np.asarray(off, np.int64)
    starts, lens = off[fl], off[fl + 1] - off[fl]
    idx = np.repeat(starts - np.cumsum(lens) + lens, lens) + np.arange(int(lens.sum()))
    return np.asarray(cols)[idx], lens


@pytest.mark.parametrize("inactive", [False, True])
def test_hashed_resort_on_the_device_large(gpu_context, oracle_lib, inactive):
    """The 101st hashed search re-sorts a collection of about 250k rows on the
    device (parts split level by level in chunks, the small parts sorted per
    block): every fluid particle's list afterwards, entry for entry, equals the
    oracle's, whose quick sort replays Sort::quick_sort.  Inactive rows take
    part in the sort (their key is the never-updated cell's)."""
    O = oracle_lib
    coll = _shuffled(scenes.dam_scene((56, 56, 56), (80, 60, 60), 2, seed=65), 12)
    if inactive:
        t = coll.info["type"]
        t[np.nonzero(t == abi.TYPE_NORMAL)[0][::13]] = abi.TYPE_INACTIVE
    assert coll.size() > 200000
    d = abi.default_desc(3, abi.SOLVER_IISPH)
    d.search = abi.SEARCH_HASHED_3D
    print("case hashed device resort", coll.size(), "inactive" if inactive else "all active", flush=True)
    sc = Scene(gpu_context, d, coll.copy())
    orc = O.OracleScene(d, coll.copy(), search=O.SEARCH_HASHED3D)
    for _ in range(101):
        sc.find_neighbors()
        orc.find_neighbors()
    assert not np.array_equal(orc.internal_ids(), np.arange(coll.size()))  # the re-sort happened
    fl = np.nonzero(coll.type == abi.TYPE_NORMAL)[0]
    gcols, glen = _fluid_lists(*sc.export_neighbors(), fl)
    ocols, olen = _fluid_lists(*orc.export_neighbors(), fl)
    np.testing.assert_array_equal(glen, olen)
    np.testing.assert_array_equal(gcols, ocols)
    sc.close()
