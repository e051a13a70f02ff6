"""ctypes mirror of include/fluidsolver_amd.h (structs, enums, status codes).

Pure data definitions: importing this module loads no native library.
"""
import ctypes as C

import numpy as np

FSA_OK = 0
FSA_E_INVALID = 1
FSA_E_NOMEM = 2
FSA_E_DEVICE = 3
FSA_E_CAPACITY = 4
FSA_E_STATE = 5
FSA_E_BUDGET = 6
FSA_E_NONFINITE = 7
FSA_E_IO = 8

STATUS_NAMES = {
    FSA_OK: "FSA_OK",
    FSA_E_INVALID: "FSA_E_INVALID",
    FSA_E_NOMEM: "FSA_E_NOMEM",
    FSA_E_DEVICE: "FSA_E_DEVICE",
    FSA_E_CAPACITY: "FSA_E_CAPACITY",
    FSA_E_STATE: "FSA_E_STATE",
    FSA_E_BUDGET: "FSA_E_BUDGET",
    FSA_E_NONFINITE: "FSA_E_NONFINITE",
    FSA_E_IO: "FSA_E_IO",
}

# neighbor-walk variants of the 3D IISPH kernels (autotuner candidates)
WALK_U2 = 0x02
WALK_U4 = 0x04
WALK_U8 = 0x08
WALK_U4_W8 = 0x84
WALK_U8_W6 = 0x68
WALK_VARIANTS = (WALK_U2, WALK_U4, WALK_U8, WALK_U4_W8, WALK_U8_W6)

SOLVER_IISPH = 0
SOLVER_SESPH = 1
LIST_COMPRESSED = 0
LIST_PLAIN = 1
LIST_DELTA = 2
SEARCH_DEFAULT = 0
SEARCH_COMPRESSED_3D = 1
SEARCH_HASHED_3D = 2
SEARCH_QUADRATIC_3D = 3
SEARCH_HASHED_2D = 4
SEARCH_QUADRATIC_2D = 5
TIMESTEP_CONSTANT = 0
TIMESTEP_DYNAMIC_CFL = 1
TYPE_NORMAL = 0
TYPE_BOUNDARY = 1
TYPE_INACTIVE = 2
FLAG_EMULATE_INT_ABS = 1
MAX_NEIGHBORS = 97

JOB_QUEUED = 0
JOB_RUNNING = 1
JOB_DONE = 2
JOB_FAILED = 3
JOB_CANCELLED = 4
JOB_STATE_NAMES = {JOB_QUEUED: "queued", JOB_RUNNING: "running", JOB_DONE: "done", JOB_FAILED: "failed",
                   JOB_CANCELLED: "cancelled"}

TIMER_NAMES = ["search", "density", "diagonal", "pressure_a", "pressure_b", "integrate", "pressure_first", "lists",
               "pressure_tail"]


class SceneDesc(C.Structure):
    """fsa_scene_desc: SimulationParameters + IISPH/SESPH settings + dt generator."""

    _fields_ = [
        ("dimension", C.c_uint32),
        ("solver", C.c_uint32),
        ("neighbor_lists", C.c_uint32),
        ("flags", C.c_uint32),
        ("rest_density", C.c_float),
        ("gravity", C.c_float),
        ("particle_size", C.c_float),
        ("iisph_max_density_error_allowed", C.c_float),
        ("iisph_min_iterations", C.c_uint32),
        ("iisph_max_iterations", C.c_uint32),
        ("iisph_omega", C.c_float),
        ("iisph_gamma", C.c_float),
        ("iisph_viscosity", C.c_float),
        ("iisph_single_layer_boundary", C.c_uint32),
        ("iisph_single_layer_gamma_1", C.c_float),
        ("sesph_stiffness_k", C.c_float),
        ("sesph_viscosity", C.c_float),
        ("sesph_single_layer_boundary", C.c_uint32),
        ("sesph_single_layer_gamma_1", C.c_float),
        ("sesph_single_layer_gamma_2", C.c_float),
        ("timestep", C.c_float),
        ("lambda_v", C.c_float),
        ("lambda_a", C.c_float),
        ("search", C.c_uint32),
        ("timestep_generator", C.c_uint32),
        ("max_timestep", C.c_float),
        ("min_timestep", C.c_float),
        ("cfl_number", C.c_float),
    ]


def default_desc(dimension=3, solver=SOLVER_IISPH, **overrides):
    """Reference defaults (IISPHSettings3D.hpp:7-17, SESPHSettings3D.hpp:7-12,
    ConstantTimestepGenerator.hpp:9-12, DynamicCflTimestepGenerator.hpp:10-17) with Simulator's overrides
    particle_size = 1, rest_density = 1 (Simulator.hpp:33-35)."""
    d = SceneDesc()
    d.dimension = dimension
    d.solver = solver
    d.neighbor_lists = LIST_COMPRESSED
    d.flags = FLAG_EMULATE_INT_ABS if dimension == 2 else 0
    d.rest_density = 1.0
    d.gravity = 9.81
    d.particle_size = 1.0
    d.iisph_max_density_error_allowed = 0.001
    d.iisph_min_iterations = 2
    d.iisph_max_iterations = 100
    d.iisph_omega = 0.5
    d.iisph_gamma = 0.7
    d.iisph_viscosity = 5.0
    d.iisph_single_layer_boundary = 0
    d.iisph_single_layer_gamma_1 = 1.1
    d.sesph_stiffness_k = 100000.0
    d.sesph_viscosity = 5.0
    d.sesph_single_layer_boundary = 0
    d.sesph_single_layer_gamma_1 = 1.1
    d.sesph_single_layer_gamma_2 = 1.1
    d.timestep = 0.001
    d.lambda_v = 0.4
    d.lambda_a = 0.25
    d.search = SEARCH_HASHED_2D if dimension == 2 else SEARCH_COMPRESSED_3D
    d.timestep_generator = TIMESTEP_CONSTANT
    d.max_timestep = 0.02  # DynamicCflTimestepGenerator.hpp:10-17
    d.min_timestep = 0.0001
    d.cfl_number = 0.8
    for k, v in overrides.items():
        setattr(d, k, v)
    return d


class ParticleView(C.Structure):
    _fields_ = [
        ("count", C.c_uint64),
        ("movement", C.c_void_p),
        ("movement_stride", C.c_uint64),
        ("particle_data", C.c_void_p),
        ("particle_data_stride", C.c_uint64),
        ("info", C.c_void_p),
        ("info_stride", C.c_uint64),
        ("external", C.c_void_p),
        ("external_stride", C.c_uint64),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("desired_time_step", C.c_float),
        ("actual_time_step", C.c_float),
        ("iterations", C.c_uint32),
        ("status", C.c_uint32),
        ("average_density_error", C.c_double),
        ("max_velocity", C.c_float),
        ("max_acceleration", C.c_float),
    ]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("max_running", C.c_uint32),
        ("serial_ticks", C.c_uint32),
        ("reserved", C.c_uint32 * 4),
    ]


class JobInfo(C.Structure):
    _fields_ = [
        ("job_id", C.c_uint64),
        ("state", C.c_uint32),
        ("status", C.c_uint32),
        ("steps_done", C.c_uint32),
        ("steps_total", C.c_uint32),
        ("fluid_particles", C.c_uint64),
        ("submit_s", C.c_double),
        ("admit_s", C.c_double),
        ("finish_s", C.c_double),
        ("last", StepStats),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("ticks", C.c_uint64),
        ("scene_steps", C.c_uint64),
        ("particle_updates", C.c_uint64),
        ("jobs_submitted", C.c_uint64),
        ("jobs_done", C.c_uint64),
        ("jobs_failed", C.c_uint64),
        ("jobs_queued", C.c_uint64),
        ("jobs_running", C.c_uint64),
        ("hbm_in_use", C.c_uint64),
        ("busy_s", C.c_double),
        ("p50_step_ms", C.c_double),
        ("p99_step_ms", C.c_double),
        ("mean_step_ms", C.c_double),
        ("max_step_ms", C.c_double),
        ("admit_s", C.c_double),
        ("step_s", C.c_double),
        ("retire_s", C.c_double),
    ]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


# entities and sensors
ENTITY_BOUNDARY_PREPROCESSOR = 1
ENTITY_VELOCITY_ALTERATION_BY_TAG = 2
ENTITY_PARTICLE_REMOVER_3D = 3


class Entity(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("tag", C.c_uint32),
        ("remove_if_outside", C.c_uint32),
        ("reserved", C.c_uint32),
        ("velocity", C.c_float * 3),
        ("center", C.c_float * 3),
        ("distance_from_center", C.c_float * 3),
    ]


def boundary_preprocessor():
    e = Entity()
    e.type = ENTITY_BOUNDARY_PREPROCESSOR
    return e


def velocity_alteration_by_tag(tag, velocity):
    e = Entity()
    e.type = ENTITY_VELOCITY_ALTERATION_BY_TAG
    e.tag = tag
    e.velocity[:] = [float(x) for x in velocity]
    return e


def particle_remover_3d(center, distance_from_center, remove_if_outside):
    e = Entity()
    e.type = ENTITY_PARTICLE_REMOVER_3D
    e.center[:] = [float(x) for x in center]
    e.distance_from_center[:] = [float(x) for x in distance_from_center]
    e.remove_if_outside = 1 if remove_if_outside else 0
    return e


def entity_array(entities):
    arr = (Entity * max(1, len(entities)))()
    for k, e in enumerate(entities):
        arr[k] = e
    return arr


class SceneStatistics(C.Structure):
    _fields_ = [
        ("normal_particles", C.c_uint64),
        ("boundary_particles", C.c_uint64),
        ("inactive_particles", C.c_uint64),
        ("density_max", C.c_float),
        ("density_min", C.c_float),
        ("density_avg", C.c_double),
        ("pressure_max", C.c_float),
        ("pressure_min", C.c_float),
        ("pressure_avg", C.c_double),
        ("velocity_max", C.c_float),
        ("velocity_min", C.c_float),
        ("velocity_avg", C.c_double),
        ("kinetic_energy", C.c_double),
        ("potential_energy", C.c_double),
    ]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


# position queries, SensorPlane and the raytracer's volume probe (3D)
SENSOR_PLANE_DENSITY = 0
SENSOR_PLANE_VELOCITY = 1
SENSOR_PLANE_PRESSURE = 2


class SensorPlane(C.Structure):
    """SensorPlane::SensorPlaneSettings (SensorPlane.hpp:17-37)."""
    _fields_ = [
        ("origin", C.c_float * 3),
        ("span_x", C.c_float * 3),
        ("span_y", C.c_float * 3),
        ("width", C.c_float),
        ("height", C.c_float),
        ("number_of_samples_x", C.c_uint64),
        ("number_of_samples_y", C.c_uint64),
        ("max_image_value", C.c_float),
        ("min_image_value", C.c_float),
        ("sensor_type", C.c_uint32),
        ("reserved", C.c_uint32),
        ("sub_sample_grid_size", C.c_uint64),
        ("calculate_plane_every_nth_step", C.c_uint64),
    ]


def sensor_plane(origin=(0.0, 0.0, 0.0), span_x=(1.0, 0.0, 0.0), span_y=(0.0, 1.0, 0.0), width=30.0, height=10.0,
                 samples_x=300, samples_y=100, sensor_type=SENSOR_PLANE_DENSITY, sub_sample_grid_size=2,
                 max_image_value=1.02, min_image_value=0.98, every_nth_step=1):
    """A SensorPlane settings record; the defaults are the reference's."""
    p = SensorPlane()
    p.origin[:] = [float(v) for v in origin]
    p.span_x[:] = [float(v) for v in span_x]
    p.span_y[:] = [float(v) for v in span_y]
    p.width, p.height = float(width), float(height)
    p.number_of_samples_x, p.number_of_samples_y = int(samples_x), int(samples_y)
    p.max_image_value, p.min_image_value = float(max_image_value), float(min_image_value)
    p.sensor_type = int(sensor_type)
    p.sub_sample_grid_size = int(sub_sample_grid_size)
    p.calculate_plane_every_nth_step = int(every_nth_step)
    return p


class VolumeSample(C.Structure):
    """ParticleIntersectionAccelerator::VolumeEvaluationResult without the position."""
    _fields_ = [
        ("density", C.c_float),
        ("boundary_only_density", C.c_float),
        ("fluid_only_density", C.c_float),
        ("fluid_only_normal", C.c_float * 3),
        ("boundary_only_normal", C.c_float * 3),
    ]


def query_positions(positions):
    """Query positions as a contiguous (n, 3) float32 array."""
    q = np.ascontiguousarray(np.asarray(positions, dtype=np.float32).reshape(-1, 3))
    return q


# scenario files
COMPONENT_MOVEMENT_2D = 1
COMPONENT_MOVEMENT_3D = 2
COMPONENT_INFO = 3
COMPONENT_PARTICLE_DATA = 4
COMPONENT_EXTERNAL_2D = 5
COMPONENT_EXTERNAL_3D = 6


class DataHeader(C.Structure):
    _fields_ = [
        ("version", C.c_uint32),
        ("dimension", C.c_uint32),
        ("components", C.c_uint32),
        ("reserved", C.c_uint32),
        ("count", C.c_uint64),
    ]

    def component_ids(self):
        return [k + 1 for k in range(6) if self.components & (1 << k)]


class Scenario(C.Structure):
    _fields_ = [
        ("desc", SceneDesc),
        ("entities", C.c_uint32),
        ("sensors", C.c_uint32),
        ("entity_count", C.c_uint32),
        ("reserved", C.c_uint32),
        ("entity_list", Entity * 32),
        ("particles", C.c_char * 1024),
        ("particles_path", C.c_char * 4096),
    ]


# numpy records with the reference component layouts (Particle.hpp:15-50)
INFO_DTYPE = np.dtype({"names": ["tag", "type"], "formats": ["<u4", "u1"], "offsets": [0, 4], "itemsize": 8})


def iisph_record_floats(dimension):
    """Floats per IISPHParticleData3D (9) / IISPHParticleData (4) record."""
    return 9 if dimension == 3 else 4


def make_view(coll):
    """ParticleView over a ParticleCollection's component arrays (no copy)."""
    v = ParticleView()
    v.count = coll.size()
    v.movement = coll.movement.ctypes.data
    v.movement_stride = coll.movement.strides[0]
    v.particle_data = coll.particle_data.ctypes.data
    v.particle_data_stride = coll.particle_data.strides[0]
    v.info = coll.info.ctypes.data
    v.info_stride = coll.info.strides[0]
    if coll.external is not None:
        v.external = coll.external.ctypes.data
        v.external_stride = coll.external.strides[0]
    else:
        v.external = None
        v.external_stride = 0
    return v
