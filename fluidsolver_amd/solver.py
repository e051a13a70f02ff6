"""Host-side mirror of libFluid's solver, neighborhood-search and simulator
interfaces, backed by the gfx950 library.

Names, fields and call order follow the reference so a caller written against
libFluid reads the same:
  IFluidSolverBase            src/libFluid/fluidSolver/IFluidSolver.hpp:16-34
  IISPHFluidSolver3D          src/libFluid/fluidSolver/solver/IISPHFluidSolver3D.hpp:22-609
  SESPHFluidSolver3D          src/libFluid/fluidSolver/solver/SESPHFluidSolver3D.hpp:12-337
  CompressedNeighborhoodSearch src/libFluid/fluidSolver/neighborhoodSearch/CompressedNeighbors.hpp:20-180
  ConstantTimestepGenerator   src/libFluid/time/ConstantTimestepGenerator.cpp
  DynamicCflTimestepGenerator src/libFluid/time/DynamicCflTimestepGenerator.cpp:6-109
  Timepoint                   src/libFluid/time/Timepoint.hpp:7-14
  Simulator                   src/libFluid/Simulator.cpp:7-57
Errors: the C ABI returns status codes; here they raise FsaError carrying the
code and the library's message.  Pre-flight problems are reported through
create_compatibility_report() as a list of issues, like CompatibilityReport.
"""
import ctypes as C
import weakref

import numpy as np

from . import abi
from ._native import lib


class FsaError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{abi.STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


def check(rc):
    if rc != abi.FSA_OK:
        raise FsaError(rc, lib().fsa_last_error().decode())


class Context:
    """Device, stream and HBM budget shared by scenes (one per GPU)."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        h = C.c_void_p()
        check(lib().fsa_context_create(int(device), int(hbm_budget_bytes), C.byref(h)))
        self._h = h
        self._scenes = weakref.WeakSet()
        self._servers = weakref.WeakSet()

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            for srv in list(self._servers):
                srv.close()
            for sc in list(self._scenes):
                sc.close()
            lib().fsa_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().fsa_context_synchronize(self._h))

    def memory_in_use(self):
        v = C.c_uint64(0)
        check(lib().fsa_context_memory_in_use(self._h, C.byref(v)))
        return v.value

    def set_timing(self, enabled):
        check(lib().fsa_context_set_timing(self._h, 1 if enabled else 0))

    def timing(self):
        n = len(abi.TIMER_NAMES)
        ms = (C.c_double * n)()
        cnt = (C.c_uint64 * n)()
        check(lib().fsa_context_get_timing(self._h, ms, cnt, n))
        return {abi.TIMER_NAMES[k]: (ms[k], cnt[k]) for k in range(n)}

    def set_graph(self, enabled):
        """Replay the head of batched steps from captured hipGraphs."""
        check(lib().fsa_batch_set_graph(self._h, 1 if enabled else 0))

    def set_autotune(self, enabled):
        check(lib().fsa_context_set_autotune(self._h, 1 if enabled else 0))

    def set_walk_variant(self, variant):
        """Force one abi.WALK_* neighbor-walk variant on every 3D IISPH kernel
        family (0: the autotuner chooses)."""
        check(lib().fsa_context_set_walk_variant(self._h, int(variant)))

    def tuning(self):
        """Text report of the autotuned walk variants and their timings."""
        n = C.c_uint64(0)
        check(lib().fsa_context_get_tuning(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib().fsa_context_get_tuning(self._h, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def batch_step(self, scenes, steps=1):
        arr = (C.c_void_p * len(scenes))(*[s.handle.value for s in scenes])
        st = (abi.StepStats * len(scenes))()
        check(lib().fsa_batch_step(self._h, arr, len(scenes), int(steps), st))
        return [st[k] for k in range(len(scenes))]


def memory_estimate(desc, n_fluid, n_boundary):
    v = C.c_uint64(0)
    check(lib().fsa_scene_memory_estimate(C.byref(desc), int(n_fluid), int(n_boundary), C.byref(v)))
    return v.value


class Scene:
    """One device-resident scene: particle state, neighbor lists, solver settings."""

    def __init__(self, context, desc, collection):
        self.context = context
        self.desc = desc
        self.collection = collection
        self.n = collection.size()
        v = abi.make_view(collection)
        h = C.c_void_p()
        check(lib().fsa_scene_create(context.handle, C.byref(desc), C.byref(v), C.byref(h)))
        self._h = h
        context._scenes.add(self)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().fsa_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, collection=None):
        c = self.collection if collection is None else collection
        check(lib().fsa_scene_upload(self._h, C.byref(abi.make_view(c))))

    def download(self, collection=None):
        c = self.collection if collection is None else collection
        check(lib().fsa_scene_download(self._h, C.byref(abi.make_view(c))))
        return c

    def download_iisph(self):
        rec = np.zeros((self.n, abi.iisph_record_floats(self.desc.dimension)), dtype=np.float32)
        check(lib().fsa_scene_download_iisph(self._h, rec.ctypes.data, rec.strides[0]))
        return rec

    def set_desc(self, desc):
        check(lib().fsa_scene_set_desc(self._h, C.byref(desc)))
        self.desc = desc

    def find_neighbors(self):
        check(lib().fsa_scene_find_neighbors(self._h))

    def get_neighbors(self, index):
        cnt = C.c_uint64(0)
        check(lib().fsa_scene_get_neighbors(self._h, int(index), None, 0, C.byref(cnt)))
        out = (C.c_uint64 * max(1, cnt.value))()
        check(lib().fsa_scene_get_neighbors(self._h, int(index), out, cnt.value, C.byref(cnt)))
        return [out[k] for k in range(cnt.value)]

    def export_neighbors(self):
        nnz = C.c_uint64(0)
        off = np.zeros(self.n + 1, dtype=np.uint64)
        check(lib().fsa_scene_export_neighbors(self._h, off.ctypes.data, None, 0, C.byref(nnz)))
        cols = np.zeros(max(1, nnz.value), dtype=np.uint64)
        check(lib().fsa_scene_export_neighbors(self._h, off.ctypes.data, cols.ctypes.data, nnz.value, C.byref(nnz)))
        return off, cols[: nnz.value]

    def neighbor_stats(self):
        t, d, r = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        check(lib().fsa_scene_neighbor_stats(self._h, C.byref(t), C.byref(d), C.byref(r)))
        return {"total_neighbors": t.value, "device_list_bytes": d.value, "reference_codec_bytes": r.value}

    def set_entities(self, entities):
        """Simulation entities run around each step (abi.Entity records: boundary
        preprocessor, velocity alteration by tag, particle remover 3D)."""
        arr = abi.entity_array(list(entities))
        check(lib().fsa_scene_set_entities(self._h, arr, len(entities)))

    def statistics(self, relative_zero_height=0.0, no_alloc=False):
        """The global sensors (density, pressure, |v|, energies, counts) reduced on the device;
        no_alloc: through the entry point that allocates no device memory."""
        out = abi.SceneStatistics()
        fn = lib().fsa_scene_get_statistics_noalloc if no_alloc else lib().fsa_scene_get_statistics
        check(fn(self._h, float(relative_zero_height), C.byref(out)))
        return out

    def query_neighbors(self, positions):
        """NeighborhoodInterface::get_neighbors(position) for each row of
        positions (n x 3), on the device: CSR (offsets, rows) of the caller
        rows near each position, in the search's iteration order."""
        q = abi.query_positions(positions)
        n = q.shape[0]
        off = np.zeros(n + 1, dtype=np.uint64)
        nnz = C.c_uint64(0)
        check(lib().fsa_scene_query_neighbors(self._h, q.ctypes.data, n, off.ctypes.data, None, 0, C.byref(nnz)))
        rows = np.zeros(max(nnz.value, 1), dtype=np.uint64)
        check(lib().fsa_scene_query_neighbors(self._h, q.ctypes.data, n, off.ctypes.data, rows.ctypes.data,
                                              rows.size, C.byref(nnz)))
        return off, rows[:nnz.value]

    def sensor_plane(self, plane):
        """SensorPlane values (abi.SensorPlane settings) over the current state:
        (samples_x * samples_y, 3) float32 in the reference's sample order."""
        out = np.zeros((plane.number_of_samples_x * plane.number_of_samples_y, 3), dtype=np.float32)
        check(lib().fsa_scene_sensor_plane(self._h, C.byref(plane), out.ctypes.data))
        return out

    def evaluate_volume(self, positions):
        """The raytracer's volume probe at each position (abi.VolumeSample array)."""
        q = abi.query_positions(positions)
        out = (abi.VolumeSample * max(q.shape[0], 1))()
        check(lib().fsa_scene_evaluate_volume(self._h, q.ctypes.data, q.shape[0], out))
        return out[:q.shape[0]]

    def next_timestep(self):
        """The desired step of the scene's timestep generator (constant or dynamic CFL)."""
        dt = C.c_float(0.0)
        check(lib().fsa_scene_next_timestep(self._h, C.byref(dt)))
        return dt.value

    def simulate(self, dt=None):
        st = abi.StepStats()
        check(lib().fsa_scene_simulate(self._h, float(self.next_timestep() if dt is None else dt), C.byref(st)))
        return st

    def step(self, steps=1):
        st = abi.StepStats()
        check(lib().fsa_scene_step(self._h, int(steps), C.byref(st)))
        return st


# ---- reference-shaped configuration objects ----------------------------------------
class SimulationParameters:
    """IFluidSolverBase::SimulationParameters with Simulator's values (Simulator.hpp:33-35)."""

    def __init__(self, rest_density=1.0, gravity=9.81, particle_size=1.0):
        self.rest_density = rest_density
        self.gravity = gravity
        self.particle_size = particle_size


class IISPHSettings3D:
    """settings/IISPHSettings3D.hpp:6-20."""

    def __init__(self):
        self.max_density_error_allowed = 0.001
        self.min_number_of_iterations = 2
        self.max_number_of_iterations = 100
        self.omega = 0.5
        self.gamma = 0.7
        self.viscosity = 5.0
        self.single_layer_boundary = False
        self.single_layer_boundary_gamma_1 = 1.1


class SESPHSettings3D:
    """settings/SESPHSettings3D.hpp:6-12."""

    def __init__(self):
        self.StiffnessK = 100000.0
        self.Viscosity = 5.0
        self.single_layer_boundary = False
        self.single_layer_boundary_gamma_1 = 1.1
        self.single_layer_boundary_gamma_2 = 1.1


class ConstantTimestepGenerator:
    """time/ConstantTimestepGenerator.cpp; the CFL correction itself runs on device."""

    class Settings:
        def __init__(self):
            self.timestep = 0.001
            self.lambda_v = 0.4
            self.lambda_a = 0.25

    def __init__(self):
        self.settings = ConstantTimestepGenerator.Settings()
        self._generated = 0.0

    def generate_next_timestep(self):
        self._generated = self.settings.timestep

    def get_generated_timestep(self):
        return self._generated


class DynamicCflTimestepGenerator:
    """time/DynamicCflTimestepGenerator.cpp: the next desired step follows the
    largest |v| and |a| of the active particles (reduced on the device during
    the integration), clamped to [min_timestep, max_timestep]; max_timestep is
    also the fallback of the solver's in-step CFL correction."""

    class Settings:
        def __init__(self):  # DynamicCflTimestepGenerator.hpp:10-17
            self.max_timestep = 0.02
            self.min_timestep = 0.0001
            self.cfl_number = 0.8
            self.lambda_v = 0.4
            self.lambda_a = 0.25

    def __init__(self):
        self.settings = DynamicCflTimestepGenerator.Settings()
        self._generated = 0.0
        self._solver = None  # set when a solver takes this generator

    def generate_next_timestep(self):
        if self._solver is None:
            raise RuntimeError("DynamicCflTimestepGenerator has no particle collection (no solver)")
        self._generated = self._solver._scene_or_create().next_timestep()

    def get_generated_timestep(self):
        return self._generated


class Timepoint:
    """time/Timepoint.hpp:7-14."""

    def __init__(self):
        self.timestep_number = 0
        self.simulation_time = 0.0
        self.desired_time_step = 0.0
        self.actual_time_step = 0.0


class _SolverData:
    def __init__(self):
        self.timestep_generator = None
        self.collection = None


class HipNeighborhoodSearch:
    """CompressedNeighborhoodSearch-shaped view of a solver's device lists:
    find_neighbors() / get_neighbors(i) in the caller's index space."""

    def __init__(self, solver):
        self._solver = solver
        self.collection = None
        self.search_radius = 0.0

    def initialize(self):
        self.collection = self._solver.data.collection
        self.search_radius = self._solver.parameters.particle_size * 2.0

    def find_neighbors(self):
        self._solver._scene_or_create().find_neighbors()

    def get_neighbors(self, index_or_position):
        """NeighborhoodInterface::get_neighbors (NeighborhoodInterface.hpp:86-89):
        by particle index, or by a 3D position (link.get_by_position_3d),
        answered on the device in the search's order."""
        scene = self._solver._scene_or_create()
        if np.ndim(index_or_position) == 0:
            return scene.get_neighbors(int(index_or_position))
        off, rows = scene.query_neighbors(np.asarray(index_or_position, dtype=np.float32).reshape(1, 3))
        return rows


class _HipFluidSolverBase:
    """IFluidSolverBase on the device: execute_neighborhood_search() then
    execute_simulation_step(timepoint).  The host collection is refreshed from
    the device after every step when `synchronize_collection` is true (the
    reference's collection is the source of truth after each step)."""

    solver_kind = None

    def __init__(self, context=None):
        self.parameters = SimulationParameters()
        self.data = _SolverData()
        self.context = context
        self.neighborhood_search = HipNeighborhoodSearch(self)
        self.synchronize_collection = True
        self._scene = None
        self._desc_key = None
        self.stat_last_iteration_count = 0
        self.stat_last_average_predicted_density_error = 0.0

    # -- desc from the reference-shaped settings
    def _desc(self):
        d = abi.default_desc(3, self.solver_kind)
        d.rest_density = self.parameters.rest_density
        d.gravity = self.parameters.gravity
        d.particle_size = self.parameters.particle_size
        tg = self.data.timestep_generator
        if isinstance(tg, DynamicCflTimestepGenerator):
            tg._solver = self
            d.timestep_generator = abi.TIMESTEP_DYNAMIC_CFL
            d.max_timestep = tg.settings.max_timestep
            d.min_timestep = tg.settings.min_timestep
            d.cfl_number = tg.settings.cfl_number
            d.lambda_v = tg.settings.lambda_v
            d.lambda_a = tg.settings.lambda_a
        elif tg is not None:
            d.timestep = tg.settings.timestep
            d.lambda_v = tg.settings.lambda_v
            d.lambda_a = tg.settings.lambda_a
        self._fill_settings(d)
        return d

    def _fill_settings(self, d):
        raise NotImplementedError

    def create_compatibility_report(self):
        issues = []
        if self.data.collection is None:
            issues.append("Particle collection is null.")
        if self.data.timestep_generator is None:
            issues.append("Timestep generator is null")
        if self.parameters.particle_size <= 0.0:
            issues.append("Search radius is smaller or equal to zero.")
        return issues

    def initialize(self):
        if self.context is None:
            self.context = Context()
        self.neighborhood_search.initialize()
        d = self._desc()
        key = bytes(d)
        if self._scene is None:
            self._scene = Scene(self.context, d, self.data.collection)
        elif key != self._desc_key:
            self._scene.set_desc(d)
        self._desc_key = key

    def _scene_or_create(self):
        if self._scene is None:
            self.initialize()
        return self._scene

    def execute_neighborhood_search(self):
        self.initialize()
        self._scene.find_neighbors()

    def execute_simulation_step(self, timepoint):
        self.initialize()
        st = self._scene.simulate(timepoint.desired_time_step)
        timepoint.actual_time_step = st.actual_time_step
        self.stat_last_iteration_count = st.iterations
        self.stat_last_average_predicted_density_error = st.average_density_error
        if self.synchronize_collection:
            self._scene.download(self.data.collection)
        return st

    def create_neighborhood_interface(self):
        return self.neighborhood_search

    def upload_collection(self):
        """Push host-side edits of data.collection to the device."""
        self._scene_or_create().upload(self.data.collection)


class IISPHFluidSolver3D(_HipFluidSolverBase):
    solver_kind = abi.SOLVER_IISPH

    def __init__(self, context=None):
        super().__init__(context)
        self.settings = IISPHSettings3D()

    def _fill_settings(self, d):
        s = self.settings
        d.iisph_max_density_error_allowed = s.max_density_error_allowed
        d.iisph_min_iterations = s.min_number_of_iterations
        d.iisph_max_iterations = s.max_number_of_iterations
        d.iisph_omega = s.omega
        d.iisph_gamma = s.gamma
        d.iisph_viscosity = s.viscosity
        d.iisph_single_layer_boundary = 1 if s.single_layer_boundary else 0
        d.iisph_single_layer_gamma_1 = s.single_layer_boundary_gamma_1

    def create_compatibility_report(self):
        issues = super().create_compatibility_report()
        if self.settings.max_number_of_iterations < self.settings.min_number_of_iterations:
            issues.append("Max iterations are less than min number of iterations.")
        if self.settings.max_density_error_allowed <= 0.0:
            issues.append("MaxDensityErrorAllowed is smaller or equal to zero.")
        return issues


class SESPHFluidSolver3D(_HipFluidSolverBase):
    solver_kind = abi.SOLVER_SESPH

    def __init__(self, context=None):
        super().__init__(context)
        self.settings = SESPHSettings3D()

    def _fill_settings(self, d):
        s = self.settings
        d.sesph_stiffness_k = s.StiffnessK
        d.sesph_viscosity = s.Viscosity
        d.sesph_single_layer_boundary = 1 if s.single_layer_boundary else 0
        d.sesph_single_layer_gamma_1 = s.single_layer_boundary_gamma_1
        d.sesph_single_layer_gamma_2 = s.single_layer_boundary_gamma_2


class Simulator:
    """Simulator::execute_simulation_step without entities and sensors
    (Simulator.cpp:7-57): dt -> neighborhood search -> solver -> time."""

    def __init__(self, fluid_solver, timestep_generator, collection):
        self.timepoint = Timepoint()
        self.fluid_solver = fluid_solver
        self.timestep_generator = timestep_generator
        fluid_solver.data.collection = collection
        fluid_solver.data.timestep_generator = timestep_generator
        if isinstance(timestep_generator, DynamicCflTimestepGenerator):
            timestep_generator._solver = fluid_solver

    def execute_simulation_step(self):
        self.timestep_generator.generate_next_timestep()
        dt = self.timestep_generator.get_generated_timestep()
        self.timepoint.desired_time_step = dt
        self.timepoint.actual_time_step = dt
        self.fluid_solver.execute_neighborhood_search()
        self.fluid_solver.execute_simulation_step(self.timepoint)
        self.timepoint.simulation_time += self.timepoint.actual_time_step
        self.timepoint.timestep_number += 1
