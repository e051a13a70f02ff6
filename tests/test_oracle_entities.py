"""The oracle's simulation entities and sensors against independent statements
(no GPU).

BoundaryPreprocessor (BoundaryPreprocessor.cpp:7-48, 96-117): on a flat
single-layer plane every interior particle sees the 5 x 5 plane sum that
defines gamma_1, so its corrected mass is the cap rest_density * ps^3 (to the
rounding of the two sums' orders); edge particles see less and are capped; in
a thick block the masses follow a float64 brute-force sum.  Fluid masses stay.
VelocityAlterationByTag equals setting the velocities by hand before each
step.  ParticleRemover3D marks exactly the particles numpy finds in the box.
The sensors equal a float32 loop over the collection in the oracle's order."""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.collection import ParticleCollection

f32 = np.float32


def w_cubic(q, h=1.0):
    """CubicSplineKernel3D value (float64) at distance q*h."""
    alpha = 1.0 / (4.0 * np.pi * h**3)
    a, b = 2.0 - q, 1.0 - q
    return alpha * np.where(q < 1.0, a**3 - 4.0 * b**3, np.where(q < 2.0, a**3, 0.0))


def boundary_only(points):
    c = ParticleCollection(3, len(points))
    c.position[:] = np.asarray(points, dtype=np.float32)
    c.particle_data[:, 0] = 1.0
    c.particle_data[:, 2] = 1.0
    c.info["type"] = abi.TYPE_BOUNDARY
    return c


SEARCHES = ["compressed", "hashed", "quadratic"]


def oracle_scene(O, desc, coll, search):
    s = {"compressed": O.SEARCH_COMPRESSED3D, "hashed": O.SEARCH_HASHED3D, "quadratic": O.SEARCH_QUADRATIC3D}[search]
    return O.OracleScene(desc, coll.copy(), search=s, mode=O.MODE_FP32)


def preprocessed_masses(O, coll, search, solver=abi.SOLVER_SESPH):
    d = abi.default_desc(3, solver, sesph_stiffness_k=1000.0, timestep=1e-6)
    orc = oracle_scene(O, d, coll, search)
    orc.set_entities([abi.boundary_preprocessor()])
    orc.find_neighbors()
    orc.simulate()
    return orc.download()


@pytest.mark.parametrize("search", SEARCHES)
def test_boundary_preprocessor_plane(oracle_lib, search):
    yz = np.stack(np.meshgrid(np.arange(15), np.arange(15), indexing="ij"), -1).reshape(-1, 2)
    pts = np.c_[np.zeros(len(yz)), yz].astype(np.float32) + np.float32(0.25)
    out = preprocessed_masses(oracle_lib, boundary_only(pts), search)
    interior = np.all((yz >= 2) & (yz <= 12), axis=1)
    m = out.mass
    cap = f32(1.0)  # rest_density * ps^3
    assert np.all(m <= cap)
    np.testing.assert_allclose(m[interior], cap, rtol=3e-7, atol=0)  # the plane sum in another order
    assert np.all(m[~interior] == cap)  # fewer neighbors: larger correction, capped


@pytest.mark.parametrize("search", SEARCHES)
def test_boundary_preprocessor_block_against_brute_force(oracle_lib, search):
    g = np.stack(np.meshgrid(np.arange(8), np.arange(6), np.arange(3), indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(4)
    pts = (g + rng.uniform(-0.05, 0.05, g.shape)).astype(np.float32)
    out = preprocessed_masses(oracle_lib, boundary_only(pts), search)
    p = pts.astype(np.float64)
    d = np.sqrt(((p[:, None, :] - p[None, :, :]) ** 2).sum(-1))
    vr = np.where(d <= 2.0, w_cubic(d), 0.0).sum(1)
    plane = sum(w_cubic(np.hypot(y, z)) for y in range(-2, 3) for z in range(-2, 3))
    want = np.minimum(1.0, plane / vr)
    np.testing.assert_allclose(out.mass, want, rtol=2e-6)
    assert np.any(out.mass < 0.99)  # thick block: below the cap


def test_boundary_preprocessor_leaves_fluid(oracle_lib):
    coll = scenes.dam_scene((4, 4, 4), (12, 8, 8), 2, seed=3)
    coll.particle_data[coll.type == abi.TYPE_NORMAL, 0] = 0.75
    out = preprocessed_masses(oracle_lib, coll, "compressed", solver=abi.SOLVER_IISPH)
    assert np.all(out.mass[out.type == abi.TYPE_NORMAL] == f32(0.75))
    assert np.any(out.mass[out.type == abi.TYPE_BOUNDARY] != f32(1.0))


@pytest.mark.parametrize("search", SEARCHES)
def test_velocity_alteration_equals_manual_edit(oracle_lib, search):
    O = oracle_lib
    coll = scenes.dam_scene((4, 4, 4), (12, 8, 8), 2, seed=5)
    coll.info["tag"] = (np.arange(coll.size()) % 3).astype(np.uint32)
    d = abi.default_desc(3, abi.SOLVER_IISPH, timestep=0.002)
    vel = np.float32([0.3, -0.2, 0.1])
    a = oracle_scene(O, d, coll, search)
    a.set_entities([abi.velocity_alteration_by_tag(2, vel)])
    b = oracle_scene(O, d, coll, search)
    for _ in range(4):
        a.step(1)
        c = b.download()
        c.velocity[c.tag == 2] = vel
        b.upload(c)
        b.step(1)
    ra, rb = a.download(), b.download()
    np.testing.assert_array_equal(ra.movement, rb.movement)
    np.testing.assert_array_equal(ra.particle_data, rb.particle_data)


@pytest.mark.parametrize("outside", [False, True])
def test_particle_remover_marks_the_box(oracle_lib, outside):
    O = oracle_lib
    coll = scenes.dam_scene((5, 5, 5), (15, 10, 10), 2, seed=6)
    d = abi.default_desc(3, abi.SOLVER_IISPH, timestep=0.004)
    center, dist = np.float32([2.0, 2.0, 2.0]), np.float32([1.5, 1.4, 3.0])
    a = oracle_scene(O, d, coll, "compressed")
    a.set_entities([abi.particle_remover_3d(center, dist, outside)])
    b = oracle_scene(O, d, coll, "compressed")
    a.step(1)
    b.step(1)
    ra, rb = a.download(), b.download()
    np.testing.assert_array_equal(ra.movement, rb.movement)
    lo, hi = center - dist, center + dist
    x = rb.position
    inside = ~((x[:, 0] < lo[0]) | (x[:, 0] > hi[0]) | (x[:, 1] > hi[1]) | (x[:, 1] < lo[1]) |
               (x[:, 2] < lo[2]) | (x[:, 2] > hi[2]))
    member = inside != outside
    assert member.any() and (~member).any()
    want = np.where(member, abi.TYPE_INACTIVE, rb.type)
    np.testing.assert_array_equal(ra.type, want)


def sensors_by_hand(coll, order, gravity, zero):
    t = coll.type[order]
    fl = t == abi.TYPE_NORMAL
    rho, p, m = (coll.density[order][fl], coll.pressure[order][fl], coll.mass[order][fl])
    v3 = coll.velocity[order][fl]
    y = coll.position[order][fl, 1]
    out = {"normal_particles": int(fl.sum()), "boundary_particles": int((t == 1).sum()),
           "inactive_particles": int((t == 2).sum())}
    lo, hi = f32(np.finfo(np.float32).max), f32(np.finfo(np.float32).tiny)
    for name, vals in (("density", rho), ("pressure", p)):
        s, mx, mn = f32(0), hi, lo
        for x in vals:
            s = f32(s + x)
            mx, mn = max(mx, x), min(mn, x)
        out[name + "_max"], out[name + "_min"] = float(mx), float(mn)
        out[name + "_avg"] = float(f32(s / f32(len(vals)))) if len(vals) else 0.0
    s, mx, mn, ke, pe = f32(0), hi, lo, f32(0), f32(0)
    for k in range(len(v3)):
        vv = f32(np.sqrt(f32(f32(f32(v3[k, 0] * v3[k, 0]) + f32(v3[k, 1] * v3[k, 1])) + f32(v3[k, 2] * v3[k, 2]))))
        s = f32(s + vv)
        mx, mn = max(mx, vv), min(mn, vv)
        ke = f32(ke + f32(f32(f32(f32(0.5) * m[k]) * vv) * vv))
        pe = f32(pe + f32(f32(f32(gravity) * m[k]) * f32(y[k] - f32(zero))))
    out["velocity_max"], out["velocity_min"] = float(mx), float(mn)
    out["velocity_avg"] = float(f32(s / f32(len(v3)))) if len(v3) else 0.0
    out["kinetic_energy"], out["potential_energy"] = float(ke), float(pe)
    return out


def test_sensors_against_a_float32_loop(oracle_lib):
    O = oracle_lib
    coll = scenes.dam_scene((5, 5, 5), (15, 10, 10), 2, seed=7)
    coll.info["type"][np.nonzero(coll.type == abi.TYPE_NORMAL)[0][::13]] = abi.TYPE_INACTIVE
    d = abi.default_desc(3, abi.SOLVER_IISPH, timestep=0.004)
    orc = O.OracleScene(d, coll.copy(), search=O.SEARCH_COMPRESSED3D, mode=O.MODE_FP32)
    orc.step(3)
    st = orc.statistics(relative_zero_height=1.5).as_dict()
    got = orc.download()
    order = orc.internal_ids().astype(np.int64)  # the collection order the reference iterates
    want = sensors_by_hand(got, order, d.gravity, 1.5)
    for k, v in want.items():
        assert st[k] == v, k


def test_sensor_maximum_starts_at_flt_min(oracle_lib):
    """MaxMinAvgSensorData starts maximum at numeric_limits<float>::min(): with
    every pressure zero the reported maximum is that tiny positive number."""
    O = oracle_lib
    coll = scenes.dam_scene((3, 3, 3), (9, 6, 6), 2, seed=8)
    orc = O.OracleScene(abi.default_desc(3, abi.SOLVER_IISPH), coll.copy(), search=O.SEARCH_COMPRESSED3D,
                        mode=O.MODE_FP32)
    st = orc.statistics()
    assert st.pressure_max == float(np.finfo(np.float32).tiny) and st.pressure_min == 0.0
