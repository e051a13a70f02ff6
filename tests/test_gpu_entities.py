"""Simulation entities and sensors on the GPU against the oracle, bit for bit:
BoundaryPreprocessor (each search's neighbor order), VelocityAlterationByTag
(incl. boundary particles, which then move), ParticleRemover3D (members turn
inactive after the step; the scene is re-laid out in the reference's order),
in single and batched steps; the global sensors reduced on the device
(maxima and minima exact, sums in FP64 within the reference's FP32 rounding)."""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.solver import Scene

pytestmark = pytest.mark.gpu

SEARCH = {"compressed": (abi.SEARCH_COMPRESSED_3D, "SEARCH_COMPRESSED3D"),
          "hashed": (abi.SEARCH_HASHED_3D, "SEARCH_HASHED3D"),
          "quadratic": (abi.SEARCH_QUADRATIC_3D, "SEARCH_QUADRATIC3D")}


def shifted(coll):
    """Every cell coordinate positive: with inactive particles present the
    compressed search's sentinel key equals cell (-1,-1,-1)'s, a reference
    defect the device does not reproduce (the parity tests keep that cell empty)."""
    coll.position[:] += np.float32(10.0)
    return coll


def pair(ctx, O, desc, coll, search, entities):
    d = abi.SceneDesc.from_buffer_copy(bytes(desc))
    d.search = SEARCH[search][0]
    sc = Scene(ctx, d, coll.copy())
    orc = O.OracleScene(d, coll.copy(), search=getattr(O, SEARCH[search][1]), mode=O.MODE_FP32_R64)
    if entities:
        sc.set_entities(entities)
        orc.set_entities(entities)
    return sc, orc


def compare_state(sc, orc, coll, label):
    g = sc.download(coll.copy())
    r = orc.download()
    np.testing.assert_array_equal(g.info["type"], r.info["type"], err_msg=label)
    np.testing.assert_array_equal(g.movement, r.movement, err_msg=label)
    np.testing.assert_array_equal(g.particle_data, r.particle_data, err_msg=label)


def run(sc, orc, coll, steps, label):
    for step in range(steps):
        st, ost = sc.step(1), orc.step(1)
        assert (st.desired_time_step, st.actual_time_step, st.iterations, st.average_density_error) == (
            ost.desired_time_step, ost.actual_time_step, ost.iterations, ost.average_density_error), (label, step)
        if step % 4 == 3 or step == steps - 1:
            compare_state(sc, orc, coll, f"{label} step {step}")


@pytest.mark.parametrize("search", ["compressed", "hashed", "quadratic"])
@pytest.mark.parametrize("solver", [abi.SOLVER_IISPH, abi.SOLVER_SESPH])
def test_boundary_preprocessor_bitwise(gpu_context, oracle_lib, search, solver):
    coll = shifted(scenes.dam_scene((5, 6, 5), (15, 10, 10), 2, seed=71))
    d = abi.default_desc(3, solver, timestep=0.003, sesph_stiffness_k=1000.0)
    d.iisph_single_layer_boundary = d.sesph_single_layer_boundary = 1
    print("case boundary preprocessor", search, solver, flush=True)
    sc, orc = pair(gpu_context, oracle_lib, d, coll, search, [abi.boundary_preprocessor()])
    run(sc, orc, coll, 8, f"bp {search} {solver}")
    got = sc.download(coll.copy())
    assert np.any(got.mass[got.type == abi.TYPE_BOUNDARY] != np.float32(1.0))  # corrected
    sc.close()


@pytest.mark.parametrize("search", ["compressed", "hashed"])
def test_velocity_alteration_bitwise(gpu_context, oracle_lib, search):
    coll = shifted(scenes.dam_scene((5, 5, 5), (15, 10, 10), 2, seed=72))
    tags = np.zeros(coll.size(), dtype=np.uint32)
    tags[coll.type == abi.TYPE_NORMAL] = np.arange(int((coll.type == abi.TYPE_NORMAL).sum())) % 4
    b = np.nonzero(coll.type == abi.TYPE_BOUNDARY)[0]
    tags[b[coll.position[b, 1] < 10.5]] = 9  # the floor slides: boundary particles move
    coll.info["tag"] = tags
    coll.info["type"][np.nonzero(tags == 3)[0][::5]] = abi.TYPE_INACTIVE  # inactive rows are altered too
    d = abi.default_desc(3, abi.SOLVER_IISPH, timestep=0.003)
    ents = [abi.velocity_alteration_by_tag(3, [0.4, 0.0, -0.3]), abi.velocity_alteration_by_tag(9, [0.2, 0.0, 0.0])]
    print("case velocity alteration", search, flush=True)
    sc, orc = pair(gpu_context, oracle_lib, d, coll, search, ents)
    run(sc, orc, coll, 10, f"va {search}")
    got = sc.download(coll.copy())
    floor = got.tag == 9
    assert np.all(got.velocity[floor] == np.float32([0.2, 0.0, 0.0]))
    assert not np.array_equal(got.position[floor], coll.position[floor])  # it moved
    sc.close()


@pytest.mark.parametrize("search", ["compressed", "hashed", "quadratic"])
def test_particle_remover_bitwise(gpu_context, oracle_lib, search):
    """A dam break whose front runs out of a kept region: members turn
    inactive step by step and the scene is re-laid out each time."""
    coll = shifted(scenes.dam_scene((6, 8, 6), (18, 12, 12), 2, seed=73))  # fluid in [10, 15] x [10, 17]
    d = abi.default_desc(3, abi.SOLVER_IISPH, timestep=0.01)
    # a drain under the fluid: the floor and the lowest layer go at once, the rest as it falls in
    drain = abi.particle_remover_3d([12.5, 8.8, 15.0], [2.0, 1.8, 10.0], False)
    print("case particle remover", search, flush=True)
    sc, orc = pair(gpu_context, oracle_lib, d, coll, search, [drain])
    removed = []
    for step in range(40):
        st, ost = sc.step(1), orc.step(1)
        assert (st.actual_time_step, st.iterations, st.average_density_error) == (
            ost.actual_time_step, ost.iterations, ost.average_density_error), step
        g = sc.download(coll.copy())
        removed.append(int((g.type == abi.TYPE_INACTIVE).sum()))
        if step % 4 == 3:
            compare_state(sc, orc, coll, f"remover {search} step {step}")
    assert removed[-1] > removed[0] > 0 and len(set(removed)) > 2, removed
    sc.close()


def test_entities_in_batches(gpu_context, oracle_lib):
    O = oracle_lib
    base = shifted(scenes.dam_scene((5, 6, 5), (15, 10, 10), 2, seed=74))
    base.info["tag"] = (np.arange(base.size()) % 5).astype(np.uint32)
    d = abi.default_desc(3, abi.SOLVER_IISPH, timestep=0.005)
    sets = [[abi.boundary_preprocessor()], [abi.velocity_alteration_by_tag(1, [0.0, 0.5, 0.0])],
            [abi.particle_remover_3d([14.0, 15.0, 15.0], [3.9, 9.0, 9.0], True)], []]
    print("case entities batched", flush=True)
    pairs = [pair(gpu_context, O, d, base, "compressed", e) for e in sets]
    scs = [p[0] for p in pairs]
    for step in range(16):
        sts = gpu_context.batch_step(scs, 1)
        for q, (st, (_, orc)) in enumerate(zip(sts, pairs)):
            ost = orc.step(1)
            assert (st.actual_time_step, st.iterations) == (ost.actual_time_step, ost.iterations), (q, step)
    for q, (sc, orc) in enumerate(pairs):
        compare_state(sc, orc, base, f"batch scene {q}")
        sc.close()


def test_statistics_against_oracle(gpu_context, oracle_lib):
    O = oracle_lib
    coll = shifted(scenes.dam_scene((8, 8, 8), (24, 14, 14), 2, seed=75))
    coll.info["type"][np.nonzero(coll.type == abi.TYPE_NORMAL)[0][::17]] = abi.TYPE_INACTIVE
    d = abi.default_desc(3, abi.SOLVER_IISPH, timestep=0.004)
    print("case statistics", flush=True)
    sc, orc = pair(gpu_context, O, d, coll, "compressed", [])
    for _ in range(6):
        sc.step(1)
        orc.step(1)
    zero = 12.5
    g = sc.statistics(zero).as_dict()
    r = orc.statistics(zero).as_dict()
    for k in ("normal_particles", "boundary_particles", "inactive_particles", "density_max", "density_min",
              "pressure_max", "pressure_min", "velocity_max", "velocity_min"):
        assert g[k] == r[k], k
    n = g["normal_particles"]
    # FP64 sums of the per-particle float terms against the reference's FP32 running sums:
    # recursive summation errs by at most (n - 1) * 2^-24 * sum |terms|
    got = sc.download(coll.copy())
    fl = got.type == abi.TYPE_NORMAL
    f32 = np.float32
    vel = got.velocity[fl]
    v = np.sqrt((vel[:, 0] * vel[:, 0] + vel[:, 1] * vel[:, 1]) + vel[:, 2] * vel[:, 2])  # float32, glm order
    m = got.mass[fl]
    # the per-particle float terms in the reference's operation order
    terms = {"density_avg": got.density[fl], "pressure_avg": got.pressure[fl], "velocity_avg": v,
             "kinetic_energy": ((f32(0.5) * m) * v) * v,
             "potential_energy": (f32(d.gravity) * m) * (got.position[fl, 1] - f32(zero))}
    for k, t in terms.items():
        assert t.dtype == np.float32, k
        t = t.astype(np.float64)
        scale = n if k.endswith("_avg") else 1
        exact = t.sum() / scale
        # device: FP64 sums of the float terms (order only)
        assert abs(g[k] - exact) <= 1e-12 * np.abs(t).sum() / scale, k
        # reference: FP32 running sums, (n - 1) * 2^-24 * sum |terms| at most, plus the final division
        bound = (n - 1) * 2.0**-24 * np.abs(t).sum() / scale
        assert abs(r[k] - exact) <= bound + 2 * np.spacing(f32(abs(exact))), k
    sc.close()
