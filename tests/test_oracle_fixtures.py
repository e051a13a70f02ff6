"""The oracle on the reference's own scenario fixtures (decoded from
scenarios/*.data into tests/golden): counts match the files, the compressed
search equals the brute force, and the scenarios step with their JSON settings."""
import numpy as np
import pytest

from fixtures import FIXTURES_2D, FIXTURES_3D, load

COUNTS = {
    "small_breaking_dam_3d": (1000, 4764),
    "filled_container_single_layer_3d": (2535, 1121),
    "boundary_test_multi_layer_3d": (1, 363),
    "box_2d": (441, 312),
    "boundary_test_2d": (1, 63),
}


@pytest.mark.parametrize("name", FIXTURES_3D + FIXTURES_2D)
def test_fixture_counts_and_defaults(name):
    c, d, s = load(name)
    nf, nb = COUNTS[name]
    assert c.counts() == {"fluid": nf, "boundary": nb, "inactive": 0}
    assert np.all(c.mass == 1.0) and np.all(c.density == 1.0)
    assert d.particle_size == 1.0 and d.rest_density == 1.0


@pytest.mark.parametrize("name", FIXTURES_3D)
def test_fixture_search_sets(oracle_lib, name):
    O = oracle_lib
    c, d, _ = load(name)
    a = O.OracleScene(d, c.copy(), search=O.SEARCH_COMPRESSED3D)
    b = O.OracleScene(d, c.copy(), search=O.SEARCH_QUADRATIC3D)
    a.find_neighbors()
    b.find_neighbors()
    oa, ca = a.export_neighbors()
    ob, cb = b.export_neighbors()
    assert np.array_equal(oa, ob) and np.array_equal(ca, cb)


@pytest.mark.parametrize("name", FIXTURES_3D)
def test_fixture_steps(oracle_lib, name):
    O = oracle_lib
    c, d, _ = load(name)
    s = O.OracleScene(d, c.copy(), mode=O.MODE_FP32)
    for _ in range(5):
        st = s.step(1)
    out = s.download()
    assert np.all(np.isfinite(out.movement))
    fl = c.type == 0
    # boundary particles are static; the fluid moved
    assert np.all(out.position[~fl] == c.position[~fl])
    assert np.any(out.position[fl] != c.position[fl])
    assert st.actual_time_step <= d.timestep
