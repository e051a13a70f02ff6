"""Scenario files: the JSON version-2 scenario and the LZ4-framed particle
.data files, read and written by the device library's host code.

Mirrors LibFluid::Serialization:
  MainSerializer::deserialize / serialize_bundle   src/libFluid/serialization/MainSerializer.cpp:21-223
  ParticleSerializer::deserialize / serialize      src/libFluid/serialization/ParticleSerializer.cpp:41-141
  SolverSerializer / TimestepGeneratorSerializer   src/libFluid/serialization/serializers/*.cpp
A scenario names a particle file relative to its JSON file; load_scenario()
resolves it the way MainSerializer::get_full_particle_data_path does and reads
it into a ParticleCollection.  Entities, sensors and visualizers are not run
by this library; their presence is reported on the Scenario.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from ._native import lib
from .collection import ParticleCollection
from .solver import check


def _path(p):
    return os.fsencode(os.fspath(p))


def read_particle_header(path):
    """u32 version, component ids, u64 count of a particle file."""
    h = abi.DataHeader()
    check(lib().fsa_data_read_header(_path(path), C.byref(h)))
    return h


def read_particles(path):
    """A ParticleCollection holding a particle file (ParticleSerializer::deserialize)."""
    h = read_particle_header(path)
    dim = h.dimension or (2 if h.components & (1 << (abi.COMPONENT_EXTERNAL_2D - 1)) else 3)
    coll = ParticleCollection(dim, int(h.count))
    v = abi.make_view(coll)
    check(lib().fsa_data_read(_path(path), C.byref(v)))
    return coll


def write_particles(path, coll, external=True):
    """Writes coll as a particle file (ParticleSerializer::serialize): movement,
    info, particle data and, unless external is False, external forces."""
    v = abi.make_view(coll)
    if not external:
        v.external = None
    check(lib().fsa_data_write(_path(path), coll.dimension, C.byref(v)))


def lz4_frame_decode(data):
    """Decoded content of an LZ4 frame (bytes)."""
    src = bytes(data)
    need = C.c_uint64(0)
    rc = lib().fsa_lz4_frame_decode(src, len(src), None, 0, C.byref(need))
    if rc not in (abi.FSA_OK, abi.FSA_E_CAPACITY):
        check(rc)
    buf = C.create_string_buffer(max(int(need.value), 1))
    check(lib().fsa_lz4_frame_decode(src, len(src), buf, need.value, C.byref(need)))
    return buf.raw[: need.value]


def lz4_frame_encode(data):
    """An LZ4 frame (independent 64 KiB blocks, content checksum) holding data."""
    src = bytes(data)
    cap = len(src) + len(src) // 255 + 64 + 8 * (len(src) // 65536 + 1)
    buf = C.create_string_buffer(cap)
    out = C.c_uint64(0)
    check(lib().fsa_lz4_frame_encode(src, len(src), buf, cap, C.byref(out)))
    return buf.raw[: out.value]


def xxh32(data, seed=0):
    src = bytes(data)
    return int(lib().fsa_xxh32(src, len(src), seed))


class Scenario:
    """A scenario file's content: the scene description (solver, search and
    settings), the timestep generator, the particle file and, once loaded,
    the particles."""

    def __init__(self, raw, collection=None):
        self.raw = raw
        self.collection = collection

    @property
    def desc(self):
        return self.raw.desc

    @property
    def timestep_generator(self):
        return self.raw.desc.timestep_generator

    @property
    def particles(self):
        return self.raw.particles.decode()

    @property
    def particles_path(self):
        return self.raw.particles_path.decode()

    @property
    def dynamic_cfl(self):
        d = self.raw.desc
        if d.timestep_generator != abi.TIMESTEP_DYNAMIC_CFL:
            return None
        return {"max_timestep": d.max_timestep, "min_timestep": d.min_timestep, "cfl": d.cfl_number}

    @property
    def entities(self):
        """Entries in the file's entity list."""
        return self.raw.entities

    @property
    def entity_list(self):
        """The entities this library runs (abi.Entity), in file order."""
        return [self.raw.entity_list[k] for k in range(self.raw.entity_count)]

    def set_entity_list(self, entities):
        if len(entities) > len(self.raw.entity_list):
            raise ValueError("at most 32 entities")
        for k, e in enumerate(entities):
            self.raw.entity_list[k] = e
        self.raw.entity_count = len(entities)

    @property
    def sensors(self):
        return self.raw.sensors


def create_scene(context, scenario):
    """A device scene for a loaded scenario, with its entities."""
    from .solver import Scene

    if scenario.collection is None:
        raise ValueError("the scenario's particles are not loaded")
    sc = Scene(context, scenario.desc, scenario.collection.copy())
    if scenario.raw.entity_count:
        sc.set_entities(scenario.entity_list)
    return sc


def default_scenario(dimension=3, solver=abi.SOLVER_IISPH):
    raw = abi.Scenario()
    lib().fsa_scenario_defaults(C.byref(raw), dimension, solver)
    return Scenario(raw)


def load_scenario(json_path, load_particles=True):
    """Parses a scenario file and reads its particle file (MainSerializer::deserialize)."""
    raw = abi.Scenario()
    check(lib().fsa_scenario_read(_path(json_path), C.byref(raw)))
    sc = Scenario(raw)
    if load_particles:
        p = sc.particles_path
        if not os.path.exists(p):
            raise FileNotFoundError(f"Could not find particle file! ({p})")
        sc.collection = read_particles(p)
        if sc.collection.dimension != raw.desc.dimension:
            raise ValueError("the particle file's dimension differs from the solver's")
    return sc


def save_scenario(json_path, scenario, collection=None, particles=None):
    """Writes a scenario file and, when a collection is given, its particle file
    beside it (MainSerializer::serialize_bundle with save_particle_data)."""
    raw = abi.Scenario()
    C.memmove(C.byref(raw), C.byref(scenario.raw), C.sizeof(abi.Scenario))
    if particles is not None:
        raw.particles = particles.encode()
    name = raw.particles.decode()
    if not name:
        name = os.path.splitext(os.path.basename(os.fspath(json_path)))[0] + ".data"
        raw.particles = name.encode()
    check(lib().fsa_scenario_write(_path(json_path), C.byref(raw)))
    coll = collection if collection is not None else scenario.collection
    if coll is not None:
        target = name if os.path.isabs(name) else os.path.join(os.path.dirname(os.fspath(json_path)), name)
        write_particles(target, coll)
    return Scenario(raw, coll)


def records_equal(a, b):
    """Component-wise bitwise equality of two collections."""
    return (a.dimension == b.dimension and np.array_equal(a.movement.view(np.uint32), b.movement.view(np.uint32))
            and np.array_equal(a.particle_data.view(np.uint32), b.particle_data.view(np.uint32))
            and np.array_equal(a.info, b.info)
            and np.array_equal(a.external.view(np.uint32), b.external.view(np.uint32)))
