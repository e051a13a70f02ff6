"""Host-side ParticleCollection with the reference's component records.

Mirrors LibFluid::ParticleCollection (src/libFluid/fluidSolver/ParticleCollection.hpp:11-177)
for the components the hot path touches (Particle.hpp:15-50): one array per
component, row i = particle i.  The arrays are handed to the device library as
strided views without repacking.
"""
import numpy as np

from .abi import INFO_DTYPE, TYPE_BOUNDARY, TYPE_INACTIVE, TYPE_NORMAL


class ParticleCollection:
    """Component store: movement (MovementData/MovementData3D), particle_data
    (ParticleData: mass, pressure, density), info (ParticleInfo: tag, type),
    external (ExternalForces/ExternalForces3D)."""

    def __init__(self, dimension, n=0):
        if dimension not in (2, 3):
            raise ValueError("dimension must be 2 or 3")
        self.dimension = dimension
        self.movement = np.zeros((n, 3 * dimension), dtype=np.float32)
        self.particle_data = np.zeros((n, 3), dtype=np.float32)
        self.info = np.zeros(n, dtype=INFO_DTYPE)
        self.external = np.zeros((n, dimension), dtype=np.float32)

    # ---- reference-shaped accessors
    def size(self):
        return self.movement.shape[0]

    @property
    def position(self):
        return self.movement[:, 0 : self.dimension]

    @property
    def velocity(self):
        return self.movement[:, self.dimension : 2 * self.dimension]

    @property
    def acceleration(self):
        return self.movement[:, 2 * self.dimension : 3 * self.dimension]

    @property
    def mass(self):
        return self.particle_data[:, 0]

    @property
    def pressure(self):
        return self.particle_data[:, 1]

    @property
    def density(self):
        return self.particle_data[:, 2]

    @property
    def type(self):
        return self.info["type"]

    @property
    def tag(self):
        return self.info["tag"]

    def counts(self):
        t = self.type
        return {
            "fluid": int(np.count_nonzero(t == TYPE_NORMAL)),
            "boundary": int(np.count_nonzero(t == TYPE_BOUNDARY)),
            "inactive": int(np.count_nonzero(t == TYPE_INACTIVE)),
        }

    def copy(self):
        c = ParticleCollection(self.dimension, 0)
        c.movement = self.movement.copy()
        c.particle_data = self.particle_data.copy()
        c.info = self.info.copy()
        c.external = None if self.external is None else self.external.copy()
        return c

    @staticmethod
    def from_arrays(dimension, position, particle_type, mass=None, density=None, velocity=None, tag=None):
        n = position.shape[0]
        c = ParticleCollection(dimension, n)
        c.position[:] = position
        if velocity is not None:
            c.velocity[:] = velocity
        c.mass[:] = 1.0 if mass is None else mass
        c.density[:] = 1.0 if density is None else density
        c.info["type"] = particle_type
        if tag is not None:
            c.info["tag"] = tag
        return c

    @staticmethod
    def concatenate(parts):
        dim = parts[0].dimension
        c = ParticleCollection(dim, 0)
        c.movement = np.concatenate([p.movement for p in parts])
        c.particle_data = np.concatenate([p.particle_data for p in parts])
        c.info = np.concatenate([p.info for p in parts])
        c.external = np.concatenate([p.external for p in parts])
        return c
