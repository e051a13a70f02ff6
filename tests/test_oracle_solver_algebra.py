"""Pins the oracle's IISPH solvers (3D and 2D) by the algebra and closed forms
the reference's equations fix, not by restating them.

The reference has no solver tests.  Each check below compares two parts of the
oracle that the reference writes independently, or one part with a value the
mathematics fixes, so that a transcription mistake shows:

* a_ii is the diagonal of the Jacobi operator: with p = e_k, pass A followed by
  pass B's product gives (A e_k)_k, which must equal the a_ii that phase 2 forms
  from its own formula (IISPHFluidSolver3D.hpp:263-336 against :348-440;
  IISPHFluidSolver.hpp:311-371 against :408-466);
* the Jacobi loop's fixed point solves A p = s on every particle with p > 0,
  and keeps p = 0 only where the relaxed update would not raise it
  (IISPHFluidSolver3D.hpp:440-462);
* uniform velocity (every particle, boundary included, moving with one v):
  viscosity is exactly zero, so a_np = gravity and the source term reduces to
  rho0 - rho_i - dt sum_b m_b v*_i . gradW (3D, :228-258) or
  rho0 - rho_i - dt sum_b m_b (v*_i - v*_b) . gradW (2D, IISPHFluidSolver.hpp:283-305),
  computed here from brute-force pairs and a numpy kernel;
* the rest lattice: an interior particle of a perfect lattice has the closed
  form density sum_j W(|x_j|), d_ii = 0 by symmetry, and
  a_ii = -dt^2 m^2/rho0^2 sum_j |gradW_ij|^2;
* 2D: free fall, momentum balance of pass A, the literal nested O(n^2)
  diagonal of IISPHFluidSolver.hpp:311-371 against the oracle's factored one
  (bit for bit in FP32), and the integer abs() of the loop condition
  (IISPHFluidSolver.hpp:400-402).

Every check is also run with a mistake planted in the oracle (set_mutation):
the check must fail on it, which shows it pins that part.
"""
import math

import numpy as np
import pytest

from fluidsolver_amd import abi, scenes
from fluidsolver_amd.collection import ParticleCollection

DT = float(np.float32(0.001))  # the API takes the step as float32


@pytest.fixture
def O(oracle_lib):
    oracle_lib.set_mutation(0)
    yield oracle_lib
    oracle_lib.set_mutation(0)


def alpha(dim):
    return 1.0 / (4.0 * math.pi) if dim == 3 else 5.0 / (14.0 * math.pi)


def kernel_np(d, dim):
    """W and grad_i W (FP64) for d = x_i - x_j, h = 1: the cubic spline's closed form."""
    q = np.linalg.norm(d, axis=-1)
    a = alpha(dim)
    w = np.where(q < 1, (2 - q) ** 3 - 4 * (1 - q) ** 3, np.where(q < 2, (2 - q) ** 3, 0.0)) * a
    fp = np.where(q < 1, -3 * (2 - q) ** 2 + 12 * (1 - q) ** 2, np.where(q < 2, -3 * (2 - q) ** 2, 0.0))
    with np.errstate(invalid="ignore", divide="ignore"):
        unit = np.where(q[..., None] > np.finfo(np.float32).eps, d / q[..., None], 0.0)
    return w, a * fp[..., None] * unit


def scene(dim, seed=5, velocity=None, boundary=True):
    """A jittered fluid block resting over a boundary plate (spacing 1, support 2)."""
    rng = np.random.default_rng(seed)
    if dim == 3:
        f = scenes.lattice((5, 5, 4), origin=(1, 1, 1))
        b = scenes.lattice((8, 8, 2), origin=(-0.5, -0.5, -1.5))
    else:
        f = scenes.lattice((6, 6), origin=(1, 1))
        b = scenes.lattice((10, 2), origin=(-1.5, -1.5))
    f = (f + rng.uniform(-0.05, 0.05, f.shape)).astype(np.float32)
    pos = np.concatenate([f, b]) if boundary else f
    types = np.concatenate([np.zeros(len(f), np.uint8), np.ones(len(pos) - len(f), np.uint8)])
    c = ParticleCollection.from_arrays(dim, pos.astype(np.float32), types)
    if velocity is None:
        c.velocity[: len(f)] = rng.uniform(-0.3, 0.3, (len(f), dim)).astype(np.float32)
    else:
        c.velocity[:] = np.asarray(velocity, np.float32)
    c.density[: len(f)] = (1.0 + rng.uniform(-0.01, 0.01, len(f))).astype(np.float32)
    return c


def prepared(O, c, mode=None, solve=False, **kw):
    desc = abi.default_desc(c.dimension, abi.SOLVER_IISPH, **kw)
    s = O.OracleScene(desc, c, mode=O.MODE_FP64 if mode is None else mode)
    s.find_neighbors()
    it = s.iisph_prepare(DT, solve=solve)
    return s, it


# ---- (a) a_ii is the diagonal of the Jacobi operator --------------------------------------
def diagonal_mismatch(O, dim):
    c = scene(dim)
    s, _ = prepared(O, c)
    diag = s.iisph_fields()["diag"]
    fluid = np.flatnonzero(c.type == abi.TYPE_NORMAL)
    col = np.zeros(len(fluid))
    for q, k in enumerate(fluid):
        e = np.zeros(c.size())
        e[k] = 1.0
        ap, _ = s.iisph_apply(DT, e)
        col[q] = ap[k]
    scale = np.abs(diag[fluid]).max()
    assert scale > 0
    return np.abs(col - diag[fluid]).max() / scale


@pytest.mark.parametrize("dim", [3, 2])
def test_diagonal_is_the_operator_diagonal(O, dim):
    assert diagonal_mismatch(O, dim) < 1e-12


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("mutation", [1, 2, 3, 6])
def test_diagonal_check_detects_planted_mistakes(O, dim, mutation):
    O.set_mutation(mutation)
    assert diagonal_mismatch(O, dim) > 1e-4


# ---- (b) the Jacobi fixed point solves A p = s ---------------------------------------------
def fixed_point_residual(O, dim):
    c = scene(dim, seed=7)
    c.velocity[c.type == abi.TYPE_NORMAL] -= np.float32(2.0) * (np.arange(dim) == dim - 1).astype(np.float32)
    # a float abs() in the 2D loop condition (flags = 0): the integer one stops at the minimum
    s, it = prepared(O, c, solve=True, iisph_max_density_error_allowed=1e-13, iisph_max_iterations=20000, flags=0)
    f = s.iisph_fields()
    p = f["pressure"]
    ap, _ = s.iisph_apply(DT, p)
    fl = c.type == abi.TYPE_NORMAL
    r = (f["source"] - ap)[fl]
    pos = p[fl] > np.finfo(np.float32).eps
    scale = np.abs(f["source"][fl]).max()
    # p > 0: A p = s; p = 0: the update p + omega (s - A p) / a_ii would not raise it
    res_pos = np.abs(r[pos]).max() / scale if pos.any() else 0.0
    zero_ok = np.all((r[~pos] / f["diag"][fl][~pos]) <= 1e-9 * scale / np.abs(f["diag"][fl]).max())
    return it, res_pos, zero_ok, int(pos.sum())


@pytest.mark.parametrize("dim", [3, 2])
def test_jacobi_fixed_point_solves_the_system(O, dim):
    it, res, zero_ok, npos = fixed_point_residual(O, dim)
    assert npos >= 5, "the block should hold a pressure field"
    assert 2 < it < 20000
    assert res < 1e-8
    assert zero_ok


@pytest.mark.parametrize("dim", [3, 2])
def test_fixed_point_check_detects_a_flipped_update(O, dim):
    O.set_mutation(5)
    it, res, zero_ok, _ = fixed_point_residual(O, dim)
    assert res > 1e-4 or not zero_ok


# ---- (c) uniform velocity: no viscosity, the source term's boundary part --------------------
def uniform_velocity_errors(O, dim):
    v = np.array([0.3, -0.2, 0.1][:dim], np.float64)
    c = scene(dim, seed=9, velocity=v)
    s, _ = prepared(O, c)
    f = s.iisph_fields()
    fl = c.type == abi.TYPE_NORMAL
    bd = c.type == abi.TYPE_BOUNDARY
    x = c.position.astype(np.float64)
    m = c.mass.astype(np.float64)
    g = np.zeros(dim)
    g[1] = -float(np.float32(9.81))
    vstar_f = v.astype(np.float32).astype(np.float64) + DT * g
    # a_np = gravity exactly (3D records it), v* = v + dt g on fluid, v on boundary
    if dim == 3:
        assert np.array_equal(f["anp"][fl], np.broadcast_to(g, (fl.sum(), dim)))
    assert np.array_equal(f["vstar"][fl], np.broadcast_to(vstar_f, (fl.sum(), dim)))
    d = x[fl][:, None, :] - x[None, :, :]
    inside = np.sum(d * d, -1) <= 4.0
    w, gw = kernel_np(d, dim)
    rho = np.sum(np.where(inside, m[None, :] * w, 0.0), 1)
    # boundary v* = v (no forces act on it): 3D weighs v*_i alone, 2D v*_i - v*_b = dt g
    v32 = v.astype(np.float32).astype(np.float64)
    rel = (vstar_f - v32) if dim == 2 else vstar_f
    bterm = np.sum(np.where(inside & bd[None, :], m[None, :] * (gw @ rel), 0.0), 1)
    src = float(np.float32(1.0)) - rho - DT * bterm
    rho_err = np.abs(f["density"][fl] - rho).max()
    src_err = np.abs(f["source"][fl] - src).max() / np.abs(DT * bterm).max()
    return rho_err, src_err


@pytest.mark.parametrize("dim", [3, 2])
def test_uniform_velocity_source_term(O, dim):
    rho_err, src_err = uniform_velocity_errors(O, dim)
    assert rho_err < 1e-12
    assert src_err < 1e-6


@pytest.mark.parametrize("dim", [3, 2])
def test_uniform_velocity_check_detects_the_other_boundary_form(O, dim):
    O.set_mutation(4)
    _, src_err = uniform_velocity_errors(O, dim)
    assert src_err > 1e-2


# ---- (d) rest lattice: closed forms of rho, s and a_ii ----------------------------------------
def lattice_closed_forms(dim):
    """Interior of the unit lattice, support 2: neighbors at |x| in {0, 1, sqrt2, sqrt3, 2}."""
    shells = {3: [(0.0, 1), (1.0, 6), (math.sqrt(2), 12), (math.sqrt(3), 8), (2.0, 6)],
              2: [(0.0, 1), (1.0, 4), (math.sqrt(2), 4), (2.0, 4)]}[dim]
    a = alpha(dim)
    rho = 0.0
    grad2 = 0.0
    for r, cnt in shells:
        w = (2 - r) ** 3 - 4 * (1 - r) ** 3 if r < 1 else (2 - r) ** 3
        fp = -3 * (2 - r) ** 2 + 12 * (1 - r) ** 2 if r < 1 else -3 * (2 - r) ** 2
        rho += cnt * a * w
        if r > 0:
            grad2 += cnt * (a * fp) ** 2
    return rho, -DT * DT * grad2


def rest_lattice_errors(O, dim):
    n = 7 if dim == 3 else 9
    pos = scenes.lattice((n,) * dim)
    c = ParticleCollection.from_arrays(dim, pos, np.zeros(len(pos), np.uint8))
    s, _ = prepared(O, c, gravity=0.0)
    f = s.iisph_fields()
    interior = np.all((pos >= 2) & (pos <= n - 3), 1)
    rho, aii = lattice_closed_forms(dim)
    return (np.abs(f["density"][interior] - rho).max(), np.abs(f["source"][interior] - (1.0 - rho)).max(),
            np.abs(f["diag"][interior] - aii).max() / abs(aii), rho)


@pytest.mark.parametrize("dim", [3, 2])
def test_rest_lattice_closed_forms(O, dim):
    e_rho, e_src, e_aii, rho = rest_lattice_errors(O, dim)
    assert rho == pytest.approx({3: 0.999972, 2: 1.000862}[dim], abs=2e-6)
    assert e_rho < 1e-14 and e_src < 1e-14
    assert e_aii < 1e-12


@pytest.mark.parametrize("dim", [3, 2])
def test_rest_lattice_check_detects_a_swapped_gradient(O, dim):
    O.set_mutation(2)
    assert rest_lattice_errors(O, dim)[2] > 1.0


# ---- 2D: free fall, momentum balance, literal diagonal, integer abs ------------------------------
@pytest.mark.parametrize("solver", [abi.SOLVER_IISPH, abi.SOLVER_SESPH])
def test_2d_free_fall_of_isolated_particle(O, solver):
    """rho = m W(0) = 4 alpha_2D = 10/(7 pi), no pressure, v = -g dt, x += dt v (Euler-Cromer), in float32."""
    c = ParticleCollection.from_arrays(2, np.array([[10.0, 20.0]], np.float32), np.zeros(1, np.uint8))
    s = O.OracleScene(abi.default_desc(2, solver), c)
    st = s.step(1)
    out = s.download()
    dt, g = np.float32(0.001), np.float32(9.81)
    vy = np.float32(np.float32(0.0) + dt * np.float32(np.float32(0.0) + -g))
    assert out.density[0] == pytest.approx(10.0 / (7.0 * math.pi), rel=1e-6)
    assert out.pressure[0] == 0.0
    assert out.velocity[0, 1] == vy and out.velocity[0, 0] == 0.0
    assert out.position[0, 1] == np.float32(np.float32(20.0) + dt * vy)
    assert out.position[0, 0] == np.float32(10.0)
    if solver == abi.SOLVER_IISPH:
        assert st.iterations == 2


def momentum_imbalance(O, dim):
    """Fluid-only blob, random pressures: pass A's pair forces m_i m_j (p_i + p_j) gradW_ij
    are antisymmetric, so sum_i m_i a_i = 0."""
    c = scene(dim, seed=3, boundary=False)
    s, _ = prepared(O, c)
    p = np.random.default_rng(2).uniform(0.0, 5.0, c.size())
    _, acc = s.iisph_apply(DT, p)
    ma = c.mass.astype(np.float64)[:, None] * acc
    return np.abs(ma.sum(0)).max() / np.abs(ma).sum()


@pytest.mark.parametrize("dim", [3, 2])
def test_pass_a_conserves_momentum(O, dim):
    assert momentum_imbalance(O, dim) < 1e-13


@pytest.mark.parametrize("dim", [3, 2])
def test_momentum_check_detects_a_wrong_pressure_index(O, dim):
    O.set_mutation(6)
    assert momentum_imbalance(O, dim) > 1e-3


def test_2d_literal_nested_diagonal_equals_factored(O):
    """IISPHFluidSolver.hpp:311-371 forms the inner sum inside the outer neighbor loop;
    the oracle forms it once.  The inner sum does not depend on the outer neighbor, so
    the literal O(n^2) loop in float32 must give the same bits."""
    c = scene(2, seed=12)
    desc = abi.default_desc(2, abi.SOLVER_IISPH)
    s = O.OracleScene(desc, c, mode=O.MODE_FP32)
    s.find_neighbors()
    s.iisph_prepare(DT)
    diag = s.iisph_fields()["diag"]
    # the lists as stored (list order), mapped to caller rows
    roff, rcols = s.export_raw()
    ids = s.internal_ids().astype(np.int64)
    lists = {int(ids[r]): [int(ids[q]) for q in rcols[roff[r]:roff[r + 1]]] for r in range(c.size())}
    f32 = np.float32
    pos, m, typ = c.position, c.mass, c.type
    rho0, gam, dt = f32(desc.rest_density), f32(desc.iisph_gamma), f32(DT)

    def grad(nb, p):  # GetKernelDerivativeReversedValue(nb, p) = -GetKernelDerivativeValue(p - nb)
        return -O.kernel_derivative(2, (pos[p] - pos[nb]).astype(np.float32))

    def dot(a, b):
        return f32(a[0] * b[0]) + f32(a[1] * b[1])

    checked = 0
    for i in np.flatnonzero(typ == abi.TYPE_NORMAL):
        nbrs = [j for j in lists[int(i)] if typ[j] != abi.TYPE_INACTIVE]
        total = f32(0.0)
        for j in nbrs:
            inner = np.zeros(2, np.float32)
            for k in nbrs:
                if typ[k] == abi.TYPE_NORMAL:
                    coef = f32(f32(m[k] / rho0) / rho0)
                else:
                    coef = f32(f32(f32(f32(f32(2.0) * gam) * m[k]) / rho0) / rho0)
                inner = (inner + (coef * grad(k, i)).astype(np.float32)).astype(np.float32)
            total = f32(total + f32(m[j] * dot(-inner, grad(j, i))))
        for j in nbrs:
            if typ[j] != abi.TYPE_NORMAL:
                continue
            a = (f32(f32(m[i] / rho0) / rho0) * grad_pair(O, pos, i, j)).astype(np.float32)
            total = f32(total + f32(m[j] * dot(a, grad(j, i))))
        assert f32(f32(dt * dt) * total) == f32(diag[i]), i
        checked += 1
    assert checked == 36


def grad_pair(O, pos, i, j):
    """GetKernelDerivativeReversedValue(x_i, x_j): the gradient with the roles swapped."""
    return -O.kernel_derivative(2, (pos[j] - pos[i]).astype(np.float32))


def test_x86_int_abs():
    """std::abs((int)e) on x86-64: truncation toward zero, INT_MIN for NaN and out-of-range values."""
    from oracle import oracle as O

    assert O.x86_int_abs(0.7) == 0.0
    assert O.x86_int_abs(-2.5) == 2.0
    assert O.x86_int_abs(2147483520.0) == 2147483520.0
    assert O.x86_int_abs(-2147483648.0) == -2147483648.0
    assert O.x86_int_abs(3.0e9) == -2147483648.0
    assert O.x86_int_abs(-3.0e9) == -2147483648.0
    assert O.x86_int_abs(float("nan")) == -2147483648.0
    assert O.x86_int_abs(float("inf")) == -2147483648.0


def test_2d_int_abs_loop_runs_min_iterations(O):
    """With the integer abs() of IISPHFluidSolver.hpp:400-402 an average error below 1 truncates
    to 0, so the loop runs exactly MinNumberOfIterations; with a float abs() and a tight
    tolerance it runs on."""
    c = scene(2, seed=4)
    c.velocity[c.type == abi.TYPE_NORMAL, 1] -= np.float32(3.0)
    for min_it in (2, 5):
        emu = abi.default_desc(2, abi.SOLVER_IISPH, iisph_min_iterations=min_it, iisph_max_density_error_allowed=1e-7)
        s = O.OracleScene(emu, c.copy(), mode=O.MODE_FP32)
        st = s.step(1)
        assert st.iterations == min_it
        assert 0.0 < st.average_density_error < 1.0
        plain = abi.default_desc(2, abi.SOLVER_IISPH, iisph_min_iterations=min_it, iisph_max_density_error_allowed=1e-7)
        plain.flags = 0
        s2 = O.OracleScene(plain, c.copy(), mode=O.MODE_FP32)
        assert s2.step(1).iterations > min_it


# ---- SESPH (3D and 2D): rest lattice and momentum balance ------------------------------------------
def sesph_rest_lattice_errors(O, dim):
    """Perfect lattice at rest, no gravity: interior density has the closed form, pressure is
    max(0, k (rho / rho0 - 1)) (SESPHFluidSolver3D.hpp:196-200), and a particle whose whole
    neighborhood is interior feels no net pressure force (symmetric neighbors, equal p / rho^2)."""
    n = 11 if dim == 3 else 13
    pos = scenes.lattice((n,) * dim)
    c = ParticleCollection.from_arrays(dim, pos, np.zeros(len(pos), np.uint8))
    desc = abi.default_desc(dim, abi.SOLVER_SESPH, gravity=0.0, sesph_stiffness_k=1000.0)
    s = O.OracleScene(desc, c, mode=O.MODE_FP64)
    s.step(1, dt=DT)
    out = s.download()
    rho, _ = lattice_closed_forms(dim)
    k = float(np.float32(1000.0))
    interior = np.all((pos >= 2) & (pos <= n - 3), 1)
    deep = np.all((pos >= 4) & (pos <= n - 5), 1)
    e_rho = np.abs(out.density[interior].astype(np.float64) - rho).max() / rho
    e_p = np.abs(out.pressure[interior].astype(np.float64) - max(0.0, k * (rho - 1.0))).max()
    a_deep = np.abs(out.acceleration[deep].astype(np.float64)).max()
    return e_rho, e_p, a_deep


@pytest.mark.parametrize("dim", [3, 2])
def test_sesph_rest_lattice(O, dim):
    e_rho, e_p, a_deep = sesph_rest_lattice_errors(O, dim)
    assert e_rho < 1e-7  # the download is float32
    assert e_p < 1e-4
    assert a_deep < 1e-6


def test_sesph_rest_lattice_check_detects_a_missing_clamp(O):
    """3D only: the 3D lattice density is below rho0 (the 2D one above it)."""
    O.set_mutation(8)
    assert sesph_rest_lattice_errors(O, 3)[1] > 1e-3


def sesph_momentum_imbalance(O, dim):
    """Fluid-only compressed blob at rest, no gravity: the pair forces
    m_i m_j (p_i / rho_i^2 + p_j / rho_j^2) gradW_ij are antisymmetric (SESPHFluidSolver3D.hpp:245-290)."""
    rng = np.random.default_rng(21)
    pos = scenes.lattice((8,) * dim) * np.float32(0.9)
    pos = (pos + rng.uniform(-0.05, 0.05, pos.shape)).astype(np.float32)
    c = ParticleCollection.from_arrays(dim, pos, np.zeros(len(pos), np.uint8))
    desc = abi.default_desc(dim, abi.SOLVER_SESPH, gravity=0.0, sesph_stiffness_k=1000.0)
    s = O.OracleScene(desc, c, mode=O.MODE_FP64)
    s.step(1, dt=DT)
    out = s.download()
    ma = out.mass.astype(np.float64)[:, None] * out.acceleration.astype(np.float64)
    assert np.abs(ma).max() > 1.0
    return np.abs(ma.sum(0)).max() / np.abs(ma).sum()


@pytest.mark.parametrize("dim", [3, 2])
def test_sesph_pressure_forces_conserve_momentum(O, dim):
    assert sesph_momentum_imbalance(O, dim) < 1e-6  # float32 download of the accelerations


@pytest.mark.parametrize("dim", [3, 2])
def test_sesph_momentum_check_detects_a_wrong_pressure_term(O, dim):
    O.set_mutation(7)
    assert sesph_momentum_imbalance(O, dim) > 1e-3
