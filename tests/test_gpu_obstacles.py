"""The obstacle configuration (SESPH and IISPH with the compressed search's
NeighborStorage capacity, a fluid block beside seeded sphere and torus
obstacles sampled as boundary particles) on the GPU.

At a size the oracle steps in seconds the device is compared with it bit for
bit.  At the full size (a 200^3 fluid block, about 9.1M particles) the checks are
independent of the oracle: the neighbor sets of sampled fluid particles equal
a float32 brute force over their 27 cells, and the densities of the first step
equal float64 sums over those sets within the 1e-5 relative bound."""
import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.solver import Scene

pytestmark = pytest.mark.gpu

f32 = np.float32


def w_cubic(d, h=1.0):
    """CubicSplineKernel3D value (float64) at distance d."""
    q = d / h
    alpha = 1.0 / (4.0 * np.pi * h**3)
    a, b = 2.0 - q, 1.0 - q
    return alpha * np.where(q < 1.0, a**3 - 4.0 * b**3, np.where(q < 2.0, a**3, 0.0))


def desc_for(solver):
    d = abi.default_desc(3, solver, timestep=0.001, sesph_stiffness_k=1000.0)
    d.neighbor_lists = abi.LIST_COMPRESSED
    return d


@pytest.mark.parametrize("solver", [abi.SOLVER_IISPH, abi.SOLVER_SESPH])
def test_obstacles_small_bitwise(gpu_context, oracle_lib, solver):
    O = oracle_lib
    coll = scenes.config3_obstacles(seed=3, side=16, n_spheres=3, n_tori=2, interior=(96, 40, 40))
    coll.position[:] += f32(10.0)  # every cell coordinate positive
    d = desc_for(solver)
    print("case obstacles small", solver, coll.size(), flush=True)
    sc = Scene(gpu_context, d, coll.copy())
    orc = O.OracleScene(d, coll.copy(), search=O.SEARCH_COMPRESSED3D, mode=O.MODE_FP32_R64)
    for step in range(15):
        st, ost = sc.step(1), orc.step(1)
        assert (st.actual_time_step, st.iterations, st.average_density_error) == (
            ost.actual_time_step, ost.iterations, ost.average_density_error), step
    g, r = sc.download(coll.copy()), orc.download()
    np.testing.assert_array_equal(g.movement, r.movement)
    np.testing.assert_array_equal(g.particle_data, r.particle_data)
    sc.close()


def _cell_index(pos, r):
    """Sorted cell keys of all particles, for 27-cell candidate lookups."""
    c = np.floor(pos / r).astype(np.int64)
    lo = c.min(0) - 1
    c -= lo
    dims = c.max(0) + 2
    key = (c[:, 2] * dims[1] + c[:, 1]) * dims[0] + c[:, 0]
    order = np.argsort(key, kind="stable")
    return c, dims, key[order], order


def _candidates(ci, dims, skeys, order):
    out = []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                k = ((ci[2] + dz) * dims[1] + (ci[1] + dy)) * dims[0] + (ci[0] + dx)
                a, b = np.searchsorted(skeys, k, "left"), np.searchsorted(skeys, k, "right")
                out.append(order[a:b])
    return np.concatenate(out)


@pytest.mark.parametrize("solver", [abi.SOLVER_IISPH, abi.SOLVER_SESPH])
def test_obstacles_full_size_lists_and_densities(gpu_context, solver):
    coll = scenes.config3_obstacles(seed=3)
    n = coll.size()
    nf = int((coll.type == abi.TYPE_NORMAL).sum())
    print("case obstacles full", solver, nf, n, flush=True)
    assert nf == 200**3 and n - nf == 2_991_211
    d = desc_for(solver)
    r = f32(d.particle_size * 2)
    sc = Scene(gpu_context, d, coll.copy())
    sc.find_neighbors()
    pos = coll.position
    c, dims, skeys, order = _cell_index(pos, r)
    rng = np.random.default_rng(11)
    fluid = np.nonzero(coll.type == abi.TYPE_NORMAL)[0]
    # fluid particles near the obstacles and the walls as well as random ones
    near = fluid[pos[fluid, 0] > 190.0]
    sample = np.concatenate([rng.choice(fluid, 300, replace=False), rng.choice(near, 100, replace=False)])
    lists = {}
    for i in sample:
        got = np.asarray(sc.get_neighbors(int(i)), dtype=np.int64)
        cand = _candidates(c[i], dims, skeys, order)
        dd = (pos[cand] - pos[i]).astype(np.float32)
        d2 = (dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]) + dd[:, 2] * dd[:, 2]
        want = cand[d2 <= r * r]
        assert sorted(got.tolist()) == sorted(want.tolist()), i
        lists[int(i)] = want
    # one step: the density of each sampled particle is the sum over its list
    st = sc.simulate()
    assert st.status == abi.FSA_OK
    got = sc.download(coll.copy())
    assert np.all(np.isfinite(got.position)) and np.all(np.isfinite(got.velocity))
    m = coll.mass.astype(np.float64)
    for i, nb in lists.items():
        dist = np.sqrt(((pos[nb].astype(np.float64) - pos[i].astype(np.float64)) ** 2).sum(1))
        want = (m[nb] * w_cubic(dist)).sum()
        assert abs(float(got.density[i]) - want) <= 1e-5 * want, i
    sc.close()
