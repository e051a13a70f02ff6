"""Pins the oracle's neighborhood searches.

2D: the reference's known answers (NeighborhoodSearchTests.cpp:98-366) for
HashedNeighborhoodSearch and QuadraticNeighborhoodSearchDynamicAllocated.
3D: the compressed search (Morton sort + run table + 27-cell scan) against the
brute-force QuadraticNeighborhoodSearch3D on random and lattice inputs,
including exact-distance ties, cells on both sides of zero and inactive
particles.
"""
import numpy as np
import pytest

from fluidsolver_amd import abi
from fluidsolver_amd.collection import ParticleCollection


def sample_grid(start, step, stop):
    """GetSampleGrid (NeighborhoodSearchTests.cpp:56-67): float loops, y outer."""
    pts = []
    y = np.float32(start)
    while y <= np.float32(stop):
        x = np.float32(start)
        while x <= np.float32(stop):
            pts.append((x, y))
            x = np.float32(x + np.float32(step))
        y = np.float32(y + np.float32(step))
    return np.array(pts, dtype=np.float32)


def scene2d(O, pts, radius, search):
    c = ParticleCollection.from_arrays(2, pts, np.zeros(len(pts), np.uint8), tag=np.arange(len(pts)))
    d = abi.default_desc(2, abi.SOLVER_SESPH)
    d.particle_size = np.float32(radius) / np.float32(2.0)
    return O.OracleScene(d, c, search=search)


def neighbors(off, cols, i):
    return sorted(cols[off[i] : off[i + 1]].tolist())


SEARCHES_2D = ["hashed", "quadratic"]


def _search_id(O, name):
    return O.SEARCH_HASHED2D if name == "hashed" else O.SEARCH_QUADRATIC2D


@pytest.mark.parametrize("name", SEARCHES_2D)
def test_single_particle(oracle_lib, name):
    O = oracle_lib
    s = scene2d(O, np.array([[234.56341, 6578.45]], np.float32), 2.0, _search_id(O, name))
    s.find_neighbors()
    off, cols = s.export_neighbors()
    assert neighbors(off, cols, 0) == [0]


P0 = [0, 1, 2, 11, 12, 22]
P38 = [16, 26, 27, 28, 36, 37, 38, 39, 40, 48, 49, 50, 60]


@pytest.mark.parametrize("name", SEARCHES_2D)
@pytest.mark.parametrize("scale", [1.0, 2.0])
def test_uniform_sampled(oracle_lib, name, scale):
    O = oracle_lib
    s = scene2d(O, sample_grid(-5 * scale, 1 * scale, 5 * scale), 2.0 * scale, _search_id(O, name))
    s.find_neighbors()
    off, cols = s.export_neighbors()
    assert neighbors(off, cols, 0) == P0
    assert neighbors(off, cols, 38) == P38


@pytest.mark.parametrize("name", SEARCHES_2D)
def test_particles_at_origin(oracle_lib, name):
    O = oracle_lib
    s = scene2d(O, sample_grid(-1, 1, 1), 2.0, _search_id(O, name))
    s.find_neighbors()
    off, cols = s.export_neighbors()
    expect = {
        0: [0, 1, 2, 3, 4, 6],
        1: [0, 1, 2, 3, 4, 5, 7],
        2: [0, 1, 2, 4, 5, 8],
        3: [0, 1, 3, 4, 5, 6, 7],
        4: [0, 1, 2, 3, 4, 5, 6, 7, 8],
        5: [1, 2, 3, 4, 5, 7, 8],
        6: [0, 3, 4, 6, 7, 8],
        7: [1, 3, 4, 5, 6, 7, 8],
        8: [2, 4, 5, 6, 7, 8],
    }
    for i, e in expect.items():
        assert neighbors(off, cols, i) == e, i


@pytest.mark.parametrize("name", SEARCHES_2D)
def test_uniform_sampled_moving(oracle_lib, name):
    """1024 translations by (0.34, -0.5) with r = 2.01; the hashed search keeps
    its grid across calls and moves particles between cells incrementally."""
    O = oracle_lib
    s = scene2d(O, sample_grid(-5, 1, 5), 2.01, _search_id(O, name))
    steps = 1024 if name == "hashed" else 64
    for _ in range(steps):
        s.translate((0.34, -0.5))
        s.find_neighbors()
        off, cols = s.export_neighbors()
        assert neighbors(off, cols, 0) == P0
        assert neighbors(off, cols, 38) == P38


def brute_force_sets(pos, types, r):
    """QuadraticNeighborhoodSearch3D restated in numpy: dot(x_i - x_j) <= r^2
    with ((dx*dx + dy*dy) + dz*dz) in float32."""
    n = pos.shape[0]
    r2 = np.float32(r) * np.float32(r)
    out = {}
    act = np.nonzero(types != abi.TYPE_INACTIVE)[0]
    for i in act:
        d = (pos[i] - pos[act]).astype(np.float32)
        dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        out[int(i)] = sorted(act[dd <= r2].tolist())
    return out


def scene3d(O, pos, types, search, ps=1.0):
    c = ParticleCollection.from_arrays(3, pos.astype(np.float32), types.astype(np.uint8))
    d = abi.default_desc(3, abi.SOLVER_IISPH)
    d.particle_size = ps
    return O.OracleScene(d, c, search=search), c


@pytest.mark.parametrize("seed,inactive", [(0, True), (1, True), (2, False), (3, False)])
def test_compressed_equals_brute_force_random(oracle_lib, seed, inactive):
    O = oracle_lib
    rng = np.random.default_rng(seed)
    n = 1500
    pos = rng.uniform(-6.0, 6.0, size=(n, 3)).astype(np.float32)
    pos[:50] = np.round(pos[:50])  # lattice points: exact-distance ties at r = 2
    pos[50:60] = pos[0]  # coincident particles
    types = rng.choice([0, 0, 0, 1, 2] if inactive else [0, 0, 0, 1], size=n).astype(np.uint8)
    if inactive:
        # keep cell (-1,-1,-1) empty: its Morton key equals the inactive sentinel
        # (see test_inactive_sentinel_collides_with_cell_minus_one)
        in_cell = np.all((pos >= -2.0) & (pos < 0.0), axis=1)
        pos[in_cell] += np.float32(3.0)
    s, _ = scene3d(O, pos, types, O.SEARCH_COMPRESSED3D)
    s.find_neighbors()
    off, cols = s.export_neighbors()
    bf = brute_force_sets(pos, types, 2.0)
    for i in range(n):
        if types[i] == abi.TYPE_INACTIVE:
            assert neighbors(off, cols, i) == []
        else:
            assert neighbors(off, cols, i) == bf[i], i


def test_lattice_counts(oracle_lib):
    O = oracle_lib
    g = np.stack(np.meshgrid(*[np.arange(9, dtype=np.float32)] * 3, indexing="ij"), -1).reshape(-1, 3)
    s, _ = scene3d(O, g, np.zeros(len(g), np.uint8), O.SEARCH_COMPRESSED3D)
    s.find_neighbors()
    off, cols = s.export_neighbors()
    counts = np.diff(off.astype(np.int64))
    interior = np.all((g >= 2) & (g <= 6), axis=1)
    assert np.all(counts[interior] == 33)
    assert counts.min() == 1 + 3 + 3 + 1 + 3  # corner: itself, 3 at 1, 3 at sqrt2, 1 at sqrt3, 3 at 2


def test_compressed_codec_size_model(oracle_lib):
    """Encoded list size of the reference codec on a jittered block (~26 B/list)."""
    O = oracle_lib
    from fluidsolver_amd import scenes

    c = scenes.dam_scene((20, 20, 20), (60, 30, 30), 3, seed=2)
    d = abi.default_desc(3, abi.SOLVER_IISPH)
    s = O.OracleScene(d, c)
    s.find_neighbors()
    total, nbytes = s.codec_stats()
    n = c.size()
    assert 20 < total / n < 34
    assert 15 < nbytes / n < 40


def test_inactive_sentinel_collides_with_cell_minus_one(oracle_lib):
    """Reference quirk: libFluid keys inactive particles with cell_index = -1
    (CompressedNeighbors.cpp:177-181), and cell (-1,-1,-1) has the Morton key
    morton3D_64(2^31 - 1, ...) = 2^64 - 1 as well.  After the stable sort an
    inactive particle uploaded first precedes the active particles of that cell,
    the run table stops at it (:207-210), and the cell is never found: its
    particle does not even list itself.  The device path sorts active particles
    only and finds the cell."""
    O = oracle_lib
    assert O.cell_key(-1, -1, -1) == 2**64 - 1
    pos = np.array([[5.0, 5.0, 5.0], [-0.5, -0.5, -0.5], [0.5, -0.5, -0.5]], np.float32)
    types = np.array([abi.TYPE_INACTIVE, abi.TYPE_NORMAL, abi.TYPE_NORMAL], np.uint8)
    s, _ = scene3d(O, pos, types, O.SEARCH_COMPRESSED3D)
    s.find_neighbors()
    off, cols = s.export_neighbors()
    assert neighbors(off, cols, 1) == [2]  # particle 1 misses itself
    assert neighbors(off, cols, 2) == [2]  # and particle 2 misses particle 1
    assert brute_force_sets(pos, types, 2.0)[1] == [1, 2]


# ---- HashedNeighborhoodSearch3D (HashedNeighborhoodSearch3D.cpp:69-247) ----------------------


@pytest.mark.parametrize("seed", [4, 5])
def test_hashed3d_sets_and_stencil_order(oracle_lib, seed):
    """Same sets as brute force; each list walks the 27 cells dx-outer, dz-inner
    with each cell's members ascending (std::set), inactive particles excluded."""
    O = oracle_lib
    rng = np.random.default_rng(seed)
    n = 1200
    pos = rng.uniform(-5.0, 5.0, size=(n, 3)).astype(np.float32)
    pos[:40] = np.round(pos[:40])
    types = rng.choice([0, 0, 0, 1, 2], size=n).astype(np.uint8)
    s, _ = scene3d(O, pos, types, O.SEARCH_HASHED3D)
    s.find_neighbors()
    off, cols = s.export_neighbors()
    bf = brute_force_sets(pos, types, 2.0)
    for i in range(n):
        assert neighbors(off, cols, i) == (bf[i] if types[i] != abi.TYPE_INACTIVE else []), i
    roff, rcols = s.export_raw()
    ids = s.internal_ids().astype(np.int64)
    assert np.array_equal(ids, np.arange(n))  # no re-sort before the 101st call
    cell = np.floor(pos / np.float32(2.0)).astype(np.int64)
    for i in range(n):
        lst = rcols[roff[i] : roff[i + 1]].astype(np.int64)
        rel = cell[lst] - cell[i] + 1  # offsets in {0,1,2}
        order = rel[:, 0] * 9 + rel[:, 1] * 3 + rel[:, 2]
        key = order * n + lst
        assert np.all(np.diff(key) > 0), i


def test_hashed3d_resorts_on_call_101(oracle_lib):
    """improve_cache_efficiency: calls 1-100 keep the order; call 101 applies the
    unstable quick sort by the Morton key of each particle's cell from the grid
    update of call 100 (inactive particles keep the INT_MIN default cell)."""
    O = oracle_lib
    rng = np.random.default_rng(7)
    n = 600
    pos = rng.uniform(0.0, 12.0, size=(n, 3)).astype(np.float32)
    types = rng.choice([0, 0, 1, 2], size=n).astype(np.uint8)
    s, c = scene3d(O, pos, types, O.SEARCH_HASHED3D)
    for call in range(1, 101):
        s.find_neighbors()
        assert np.array_equal(s.internal_ids(), np.arange(n)), call
        if call == 100:
            p100 = s.download().position.copy()  # positions of the grid update of call 100
        s.translate((0.013, -0.007, 0.011))
    cells = np.floor(p100 / np.float32(2.0)).astype(np.int64)
    imin = np.int64(-(2 ** 31))
    cells[types == abi.TYPE_INACTIVE] = imin
    keys = np.array([O.morton3(*[int(v - imin) for v in cells[i]]) for i in range(n)], np.uint64)
    s.find_neighbors()
    np.testing.assert_array_equal(s.internal_ids(), O.quick_sort(keys))
