"""Loads the committed reference fixtures (tests/golden/*.npz) as
ParticleCollections plus the scene description of the scenario's JSON."""
import json
import os

import numpy as np

from fluidsolver_amd import abi
from fluidsolver_amd.collection import ParticleCollection

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES_3D = ["small_breaking_dam_3d", "filled_container_single_layer_3d", "boundary_test_multi_layer_3d"]
FIXTURES_2D = ["box_2d", "boundary_test_2d"]


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    dim = int(z["dimension"])
    n = z["type"].shape[0]
    c = ParticleCollection(dim, n)
    c.movement[:] = z["movement"]
    c.particle_data[:] = z["particle_data"]
    c.info["type"] = z["type"]
    c.info["tag"] = z["tag"]
    c.external[:] = z["external"]
    s = json.loads(bytes(z["settings_json"]).decode())
    solver = abi.SOLVER_IISPH if s["solver"].startswith("iisph") else abi.SOLVER_SESPH
    d = abi.default_desc(dim, solver)
    d.gravity = s["gravity"]
    d.particle_size = s["particle_size"]
    d.rest_density = s["rest_density"]
    d.timestep = s["timestep"]
    if solver == abi.SOLVER_SESPH:
        d.sesph_stiffness_k = s.get("stiffness", d.sesph_stiffness_k)
        d.sesph_viscosity = s.get("viscosity", d.sesph_viscosity)
        d.sesph_single_layer_boundary = 1 if s["single_layer"] else 0
    else:
        d.iisph_viscosity = s.get("viscosity", d.iisph_viscosity)
        d.iisph_gamma = s.get("gamma", d.iisph_gamma)
        d.iisph_omega = s.get("omega", d.iisph_omega)
        d.iisph_max_density_error_allowed = s.get("max_density_error", d.iisph_max_density_error_allowed)
        d.iisph_min_iterations = s.get("min_iterations", d.iisph_min_iterations)
        d.iisph_max_iterations = s.get("max_iterations", d.iisph_max_iterations)
        d.iisph_single_layer_boundary = 1 if s["single_layer"] else 0
        d.iisph_single_layer_gamma_1 = s.get("single_layer_gamma_1", d.iisph_single_layer_gamma_1)
    return c, d, s
