"""The oracle's timestep generators against a numpy statement of the
reference's formulas (no GPU).

DynamicCflTimestepGenerator::generate_next_timestep (DynamicCflTimestepGenerator.cpp:46-62):
the largest float lengths |v| and |a| over non-inactive particles; min_timestep
unless both exceed FLT_EPSILON, else max(min_timestep, min(max_timestep,
get_non_cfl_validating_timestep(|a|, |v|))) with (:84-109) dt_a = sqrt(double)
of particle_size / |a| times lambda_a, dt_v = particle_size / |v| * lambda_v,
fmax(fmin(dt_a, dt_v), 1e-6).  A mistake in the maximum (inactive rows, a
dropped component), in the clamps or in the fallback fails one of the cases."""
import numpy as np
import pytest

from fluidsolver_amd import abi
from fluidsolver_amd.collection import ParticleCollection

f32 = np.float32


def expected(coll, d):
    act = coll.type != abi.TYPE_INACTIVE
    dim = coll.dimension
    v = coll.velocity[act].astype(f32)
    a = coll.acceleration[act].astype(f32)

    def lengths(x):
        s = x[:, 0] * x[:, 0]
        for k in range(1, dim):
            s = f32(s + x[:, k] * x[:, k])
        return np.sqrt(s.astype(f32)).astype(f32)

    mv = f32(lengths(v).max()) if len(v) else f32(0)
    ma = f32(lengths(a).max()) if len(a) else f32(0)
    ts = f32(d.min_timestep)
    eps = np.finfo(np.float32).eps
    if mv > eps and ma > eps:
        ps = f32(d.particle_size)
        dta = f32(np.sqrt(np.float64(f32(ps / ma))) * np.float64(f32(d.lambda_a)))
        dtv = f32(f32(ps / mv) * f32(d.lambda_v))
        allowed = max(min(dta, dtv), f32(1e-6))
        ts = max(ts, min(f32(d.max_timestep), allowed))
    return float(ts)


def make(dim, n, vscale, ascale, seed, inactive_fast=False):
    rng = np.random.default_rng(seed)
    c = ParticleCollection(dim, n)
    c.position[:] = rng.uniform(0, 10, (n, dim)).astype(f32)
    c.velocity[:] = (rng.standard_normal((n, dim)) * vscale).astype(f32)
    c.acceleration[:] = (rng.standard_normal((n, dim)) * ascale).astype(f32)
    c.particle_data[:, 0] = 1.0
    c.particle_data[:, 2] = 1.0
    c.info["type"] = np.where(rng.random(n) < 0.7, abi.TYPE_NORMAL, abi.TYPE_BOUNDARY).astype(np.uint8)
    if inactive_fast:
        c.info["type"][:3] = abi.TYPE_INACTIVE
        c.velocity[:3] = 1e6
        c.acceleration[:3] = 1e9
    return c


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("case", ["rest", "v_only", "typical", "fast", "slow", "inactive"])
def test_dynamic_cfl_generator(oracle_lib, dim, case):
    O = oracle_lib
    d = abi.default_desc(dim, abi.SOLVER_SESPH, timestep_generator=abi.TIMESTEP_DYNAMIC_CFL,
                         max_timestep=0.01, min_timestep=0.0002)
    vs, as_ = {"rest": (0, 0), "v_only": (2.0, 0), "typical": (30.0, 4000.0), "fast": (800.0, 5e5),
               "slow": (1e-3, 1e-2), "inactive": (30.0, 4000.0)}[case]
    c = make(dim, 400, vs, as_, seed=dim * 10 + len(case), inactive_fast=case == "inactive")
    search = O.SEARCH_HASHED2D if dim == 2 else O.SEARCH_HASHED3D
    orc = O.OracleScene(d, c.copy(), search=search, mode=O.MODE_FP32)
    want = expected(c, d)
    assert orc.next_timestep() == want
    if case in ("rest", "v_only"):
        assert want == float(f32(0.0002))
    if case == "slow":
        assert want == float(f32(0.01))  # clamped to max_timestep
    if case == "fast":
        assert want == float(f32(0.0002))  # clamped to min_timestep
    if case == "typical":
        assert float(f32(0.0002)) < want < float(f32(0.01))


def test_constant_generator(oracle_lib):
    O = oracle_lib
    d = abi.default_desc(3, abi.SOLVER_SESPH, timestep=0.0035)
    c = make(3, 50, 3.0, 40.0, seed=1)
    assert O.OracleScene(d, c, search=O.SEARCH_HASHED3D, mode=O.MODE_FP32).next_timestep() == float(f32(0.0035))
