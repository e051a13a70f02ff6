"""GPU parity: the device path against the CPU oracle, through the C ABI.

The device keeps the reference's particle order, neighbor order and FP32
evaluation order, so against the oracle in MODE_FP32_R64 (reference FP32
arithmetic, FP64 convergence sums like the device) every state array must be
bit-identical, step after step.  Neighbor sets are compared by caller index.
"""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.collection import ParticleCollection
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


def _compare_steps(sc, orc, coll, steps, label, check_iisph=True):
    fluid = coll.type == abi.TYPE_NORMAL
    for step in range(steps):
        st = sc.step(1)
        ost = orc.step(1)
        g = sc.download(coll.copy())
        r = orc.download()
        assert st.iterations == ost.iterations, (label, step)
        assert st.actual_time_step == ost.actual_time_step, (label, step)
        assert st.average_density_error == pytest.approx(ost.average_density_error, rel=1e-12, abs=1e-300), (label, step)
        if check_iisph:
            np.testing.assert_array_equal(sc.download_iisph()[fluid], orc.download_iisph()[fluid], err_msg=f"{label} {step}")
        np.testing.assert_array_equal(g.movement, r.movement, err_msg=f"{label} movement step {step}")
        np.testing.assert_array_equal(g.particle_data, r.particle_data, err_msg=f"{label} pdata step {step}")
        np.testing.assert_array_equal(g.external, r.external, err_msg=f"{label} external step {step}")


def test_dense_lattice_lists_and_steps(gpu_context, oracle_lib):
    """Spacing 0.75: about 80 neighbors per particle and neighbor cells holding up
    to 27 particles, so the 27-cell candidate sets exceed one LDS chunk."""
    O = oracle_lib
    rng = np.random.default_rng(21)
    f = (scenes.lattice((16, 16, 16)) * np.float32(0.75) + rng.uniform(-0.01, 0.01, (4096, 3))).astype(np.float32)
    b = (scenes.box_shell((16, 16, 16), 2) * np.float32(0.75)).astype(np.float32)
    from fluidsolver_amd.collection import ParticleCollection

    coll = ParticleCollection.from_arrays(
        3, np.concatenate([f, b]), np.concatenate([np.zeros(len(f), np.uint8), np.ones(len(b), np.uint8)])
    )
    desc = abi.default_desc(3, abi.SOLVER_IISPH)
    print(f"case dense lattice n={coll.size()}", flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    sc.find_neighbors()
    goff, gcols = sc.export_neighbors()
    orc = O.OracleScene(desc, coll.copy(), search=O.SEARCH_QUADRATIC3D)
    orc.find_neighbors()
    ooff, ocols = orc.export_neighbors()
    rows = _fluid_rows(coll)
    assert _csr_rows(goff, gcols, rows) == _csr_rows(ooff, ocols, rows)
    assert max(len(v) for v in _csr_rows(goff, gcols, rows).values()) > 70
    # stepping a lattice this dense compresses it past the 97-neighbor capacity;
    # step a lattice of spacing 0.85 (about 55 neighbors) instead
    f2 = (scenes.lattice((14, 14, 14)) * np.float32(0.85) + rng.uniform(-0.01, 0.01, (2744, 3))).astype(np.float32)
    b2 = (scenes.box_shell((14, 14, 14), 2) * np.float32(0.85)).astype(np.float32)
    coll2 = ParticleCollection.from_arrays(
        3, np.concatenate([f2, b2]), np.concatenate([np.zeros(len(f2), np.uint8), np.ones(len(b2), np.uint8)])
    )
    sc2 = Scene(gpu_context, desc, coll2.copy())
    orc2 = O.OracleScene(desc, coll2.copy(), mode=O.MODE_FP32_R64)
    _compare_steps(sc2, orc2, coll2, 3, "dense")


def test_capacity_overflow_is_reported(gpu_context):
    """More than 97 neighbors (NeighborStorage capacity) raises FSA_E_CAPACITY."""
    from fluidsolver_amd.collection import ParticleCollection
    from fluidsolver_amd.solver import FsaError

    pos = (scenes.lattice((10, 10, 10)) * np.float32(0.45)).astype(np.float32)
    coll = ParticleCollection.from_arrays(3, pos, np.zeros(len(pos), np.uint8))
    sc = Scene(gpu_context, abi.default_desc(3, abi.SOLVER_IISPH), coll)
    with pytest.raises(FsaError) as e:
        sc.find_neighbors()
    assert e.value.code == abi.FSA_E_CAPACITY


@pytest.mark.parametrize("name", ["small_breaking_dam_3d", "filled_container_single_layer_3d",
                                  "boundary_test_multi_layer_3d"])
def test_reference_fixtures_bitwise(gpu_context, oracle_lib, name):
    """The reference's own scenarios with their JSON settings (SESPH, single-layer
    IISPH), GPU against the oracle, step by step."""
    from fixtures import load

    O = oracle_lib
    coll, desc, _ = load(name)
    print(f"case fixture {name} n={coll.size()}", flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    orc = O.OracleScene(desc, coll.copy(), mode=O.MODE_FP32_R64)
    _compare_steps(sc, orc, coll, 4, name, check_iisph=desc.solver == abi.SOLVER_IISPH)


@pytest.mark.parametrize("solver", [abi.SOLVER_IISPH, abi.SOLVER_SESPH])
def test_inactive_dynamic_boundary_and_external(gpu_context, oracle_lib, solver):
    """Inactive particles (skipped by every loop, kept as uploaded), a moving
    boundary (non-zero velocity) and external accelerations (IISPH adds and then
    resets them)."""
    O = oracle_lib
    coll = scenes.dam_scene((8, 8, 8), (24, 12, 12), 3, seed=9)
    # keep every cell coordinate positive: with inactive particles present the
    # reference misses cell (-1,-1,-1) (test_inactive_sentinel_collides_with_cell_minus_one)
    coll.position[:] += np.float32(10.0)
    rng = np.random.default_rng(9)
    n = coll.size()
    t = coll.type.copy()
    far = np.nonzero(coll.position[:, 1] > 16.0)[0]
    t[far[:20]] = abi.TYPE_INACTIVE  # inactive fluid away from cell (-1,-1,-1)
    coll.info["type"] = t
    bnd = np.nonzero(t == abi.TYPE_BOUNDARY)[0]
    coll.velocity[bnd[:50]] = np.float32([0.0, 0.0, 0.5])
    fl = np.nonzero(t == abi.TYPE_NORMAL)[0]
    coll.external[fl[:100]] = rng.uniform(-1, 1, (100, 3)).astype(np.float32)
    desc = abi.default_desc(3, solver)
    if solver == abi.SOLVER_SESPH:
        desc.sesph_stiffness_k = 1000.0
    print(f"case inactive/dynamic solver={solver} n={n}", flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    orc = O.OracleScene(desc, coll.copy(), mode=O.MODE_FP32_R64)
    _compare_steps(sc, orc, coll, 4, "mixed", check_iisph=False)
    g = sc.download(coll.copy())
    ina = t == abi.TYPE_INACTIVE
    np.testing.assert_array_equal(g.movement[ina], coll.movement[ina])
    np.testing.assert_array_equal(g.particle_data[ina], coll.particle_data[ina])


def test_simulator_mirror_runs(gpu_context, oracle_lib):
    """The reference-shaped host API (Simulator + IISPHFluidSolver3D) drives the
    device and keeps data.collection in sync after each step."""
    from fluidsolver_amd.solver import ConstantTimestepGenerator, IISPHFluidSolver3D, Simulator

    O = oracle_lib
    coll = scenes.dam_scene((6, 6, 6), (18, 9, 9), 3, seed=12)
    solver = IISPHFluidSolver3D(gpu_context)
    sim = Simulator(solver, ConstantTimestepGenerator(), coll)
    assert solver.create_compatibility_report() == []
    for _ in range(3):
        sim.execute_simulation_step()
    assert solver.stat_last_iteration_count >= 2
    assert sim.timepoint.timestep_number == 3
    orc = O.OracleScene(abi.default_desc(3, abi.SOLVER_IISPH), scenes.dam_scene((6, 6, 6), (18, 9, 9), 3, seed=12),
                        mode=O.MODE_FP32_R64)
    orc.step(3)
    r = orc.download()
    np.testing.assert_array_equal(coll.movement, r.movement)
    nb = solver.neighborhood_search.get_neighbors(0)
    assert 0 in nb


def test_batch_equals_alone(gpu_context):
    """Scenes advanced together in shared launches give bit-identical results to
    each scene stepped alone (uneven sizes, different viscosities, both solvers)."""
    descs, colls = [], []
    for q, (n, nu, solver) in enumerate([(5, 5.0, 0), (9, 0.01, 0), (7, 0.05, 0), (6, 5.0, 1), (4, 1.0, 1)]):
        c = scenes.dam_scene((n, n + 1, n), (3 * n, 2 * n, 2 * n), 3, seed=30 + q)
        d = abi.default_desc(3, solver)
        if solver == abi.SOLVER_IISPH:
            d.iisph_viscosity = nu
        else:
            d.sesph_viscosity = nu
            d.sesph_stiffness_k = 1000.0
        descs.append(d)
        colls.append(c)
    alone = []
    for d, c in zip(descs, colls):
        sc = Scene(gpu_context, d, c.copy())
        stats = [sc.step(1).as_dict() for _ in range(6)]
        alone.append((stats, sc.download(c.copy())))
        sc.close()
    batch = [Scene(gpu_context, d, c.copy()) for d, c in zip(descs, colls)]
    bstats = [[] for _ in batch]
    for _ in range(6):
        for q, st in enumerate(gpu_context.batch_step(batch, 1)):
            bstats[q].append(st.as_dict())
    for q, sc in enumerate(batch):
        assert bstats[q] == alone[q][0], q
        g = sc.download(colls[q].copy())
        np.testing.assert_array_equal(g.movement, alone[q][1].movement)
        np.testing.assert_array_equal(g.particle_data, alone[q][1].particle_data)


def test_scene_without_fluid_alone_and_in_batches(gpu_context, oracle_lib):
    """A 3D IISPH scene with no fluid particle has no phase-1 block, so its
    Jacobi loop state comes from the separate loop-start launch; the scenes with
    fluid get theirs from phase 1.  Stepped alone it equals the oracle; in a
    batch with two dam scenes every scene equals running it alone, in sets that
    skip the first iteration's check (min iterations 2) and sets that keep it."""
    O = oracle_lib
    dam = scenes.dam_scene((6, 7, 6), (18, 12, 12), 3, seed=41)
    walls = ParticleCollection.from_arrays(3, dam.position[dam.type == abi.TYPE_BOUNDARY].copy(),
                                           np.full(int(np.count_nonzero(dam.type == abi.TYPE_BOUNDARY)),
                                                   abi.TYPE_BOUNDARY, np.uint8))
    other = scenes.dam_scene((5, 5, 5), (15, 10, 10), 3, seed=42)
    for min_it in (1, 2):
        d = abi.default_desc(3, abi.SOLVER_IISPH)
        d.iisph_min_iterations = min_it
        print(f"case fluid-free scene, min iterations {min_it}", flush=True)
        sc = Scene(gpu_context, d, walls.copy())
        orc = O.OracleScene(d, walls.copy(), mode=O.MODE_FP32_R64)
        for step in range(3):
            st = sc.step(1)
            ost = orc.step(1)
            assert st.status == abi.FSA_OK
            assert st.iterations == ost.iterations, (min_it, step)
            assert st.average_density_error == ost.average_density_error or (
                np.isnan(st.average_density_error) and np.isnan(ost.average_density_error)), (min_it, step)
        sc.close()
        colls = [dam, walls, other]
        alone = []
        for c in colls:
            a = Scene(gpu_context, d, c.copy())
            alone.append(([a.step(1).as_dict() for _ in range(3)], a.download(c.copy())))
            a.close()
        batch = [Scene(gpu_context, d, c.copy()) for c in colls]
        bstats = [[] for _ in batch]
        for _ in range(3):
            for q, st in enumerate(gpu_context.batch_step(batch, 1)):
                bstats[q].append(st.as_dict())
        for q, (b, c) in enumerate(zip(batch, colls)):
            assert str(bstats[q]) == str(alone[q][0]), (min_it, q)
            g = b.download(c.copy())
            np.testing.assert_array_equal(g.movement, alone[q][1].movement)
            np.testing.assert_array_equal(g.particle_data, alone[q][1].particle_data)
            b.close()


def test_cpp_dropin_program(oracle_lib, tmp_path):
    """A C++ program using include/fluidsolver_amd.hpp (ParticleCollection +
    IISPHFluidSolver3D + Simulator) reproduces the oracle bit for bit."""
    import os
    import subprocess

    from fluidsolver_amd import build as fbuild

    O = oracle_lib
    if not os.path.exists(fbuild.DROPIN_BIN):
        pytest.skip("drop-in example not built")
    out = tmp_path / "movement.bin"
    print("case cpp dropin 6x6x6", flush=True)
    r = subprocess.run([fbuild.DROPIN_BIN, "6", "6", "6", "3", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "iterations" in r.stdout
    mv = np.fromfile(out, dtype=np.float32).reshape(-1, 9)
    coll = scenes.dam_scene((6, 6, 6), (18, 12, 12), 3, seed=0, amplitude=0.0)
    orc = O.OracleScene(abi.default_desc(3, abi.SOLVER_IISPH), coll, mode=O.MODE_FP32_R64)
    orc.step(3)
    np.testing.assert_array_equal(mv, orc.download().movement)


def test_codec_sizes_match_oracle(gpu_context, oracle_lib):
    """NeighborStorage neighbor counts and encoded bytes of the fluid lists,
    counted on the device, equal the oracle's compressed search."""
    O = oracle_lib
    coll = scenes.dam_scene((10, 9, 8), (30, 14, 12), 3, seed=41)
    desc = abi.default_desc(3, abi.SOLVER_IISPH)
    print("case codec sizes", coll.size(), flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    orc = O.OracleScene(desc, coll.copy())
    for step in range(3):
        sc.find_neighbors()
        orc.find_neighbors()
        st = sc.neighbor_stats()
        cnt, byt = orc.codec_sizes()
        fl = coll.type == abi.TYPE_NORMAL
        assert st["total_neighbors"] == int(cnt[fl].sum()), step
        assert st["reference_codec_bytes"] == int(byt[fl].sum()), step
        sc.simulate()
        orc.simulate()


def _codec_probe(m):
    """One fluid particle P whose m neighbors in the next cell are separated,
    in the collection order that the stable sort keeps within a cell, by 260
    boundary particles outside its radius: its list has m - 1 four-byte deltas."""
    from fluidsolver_amd.collection import ParticleCollection

    rows = [np.array([[1.0, 1.0, 1.0]], np.float32)]
    far = scenes.lattice((13, 20, 25)).astype(np.float32) * np.float32([0.038, 0.1, 0.08]) + np.float32([3.5, 0.0, 0.0])
    for q in range(m):
        rows.append(np.array([[2.0 + 0.02 * q, 1.0, 1.0]], np.float32))
        if q + 1 < m:
            rows.append(far[260 * q : 260 * (q + 1)])
    pos = np.concatenate(rows).astype(np.float32)
    types = np.ones(len(pos), np.uint8)
    types[0] = abi.TYPE_NORMAL
    return ParticleCollection.from_arrays(3, pos, types)


def test_codec_capacity_boundary(gpu_context, oracle_lib):
    """The compressed search's fixed NeighborStorage holds 96 payload bytes:
    24 four-byte deltas fit, 25 overflow (FSA_E_CAPACITY); plain lists take both."""
    from fluidsolver_amd.solver import FsaError

    O = oracle_lib
    for m, fits in [(25, True), (26, False)]:
        coll = _codec_probe(m)
        print(f"case codec capacity m={m} n={coll.size()}", flush=True)
        ref = O.Codec()
        ref.set_first_neighbor(0)
        ok = [ref.set_next_neighbor(1)] + [ref.set_next_neighbor(261) for _ in range(m - 1)]
        assert all(ok) == fits
        sc = Scene(gpu_context, abi.default_desc(3, abi.SOLVER_IISPH), coll.copy())
        if fits:
            sc.find_neighbors()
            assert sc.neighbor_stats()["reference_codec_bytes"] == ref.encoded_bytes()
        else:
            with pytest.raises(FsaError) as e:
                sc.find_neighbors()
            assert e.value.code == abi.FSA_E_CAPACITY
        sc.close()
        plain = Scene(gpu_context, abi.default_desc(3, abi.SOLVER_IISPH, neighbor_lists=abi.LIST_PLAIN), coll.copy())
        plain.find_neighbors()
        assert plain.get_neighbors(0) == sorted(plain.get_neighbors(0)) and len(plain.get_neighbors(0)) == m + 1
        plain.close()


def test_tiny_coordinates_take_the_ieee_path(gpu_context, oracle_lib):
    """A coordinate in (0, 2^-36) can make a pair difference smaller than the
    shared-reciprocal division covers; such scenes run the IEEE pair math and
    stay bit-identical (the fast path is checked by every other test)."""
    O = oracle_lib
    coll = scenes.dam_scene((6, 6, 6), (18, 10, 10), 2, seed=71)
    fl = np.nonzero(coll.type == abi.TYPE_NORMAL)[0]
    coll.position[fl[0], 0] = np.float32(3e-20)  # next to particles at x = 0 (jittered lattice)
    desc = abi.default_desc(3, abi.SOLVER_IISPH)
    print("case tiny coordinates", flush=True)
    sc = Scene(gpu_context, desc, coll.copy())
    orc = O.OracleScene(desc, coll.copy(), mode=O.MODE_FP32_R64)
    _compare_steps(sc, orc, coll, 4, "tiny")


def test_set_desc_refuses_a_search_change(gpu_context):
    """The search fixes a scene's buffers (the hashed and quadratic orders keep grid
    state a compressed scene has none of): set_desc refuses to change it on the host,
    and the scene then steps as before."""
    from fluidsolver_amd.solver import FsaError

    c = scenes.dam_scene((5, 5, 5), (15, 10, 10), 3, seed=31)
    d = abi.default_desc(3, abi.SOLVER_IISPH)
    print("case set_desc search change", flush=True)
    sc = Scene(gpu_context, d, c.copy())
    for other in (abi.SEARCH_HASHED_3D, abi.SEARCH_QUADRATIC_3D):
        bad = abi.default_desc(3, abi.SOLVER_IISPH)
        bad.search = other
        with pytest.raises(FsaError) as e:
            sc.set_desc(bad)
        assert e.value.code == abi.FSA_E_INVALID
    ok = abi.default_desc(3, abi.SOLVER_IISPH)
    ok.neighbor_lists = abi.LIST_PLAIN  # a change of settings alone is accepted
    sc.set_desc(ok)
    sc.set_desc(d)
    sc.step(2)
    got = sc.download(c.copy())
    sc.close()
    ref = Scene(gpu_context, d, c.copy())
    ref.step(2)
    np.testing.assert_array_equal(got.movement, ref.download(c.copy()).movement)
    ref.close()
