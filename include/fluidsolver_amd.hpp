// fluidsolver_amd.hpp — header-only C++ host layer over the C ABI.
//
// Mirrors the reference's solver-side interfaces so C++ code written against
// libFluid reads the same, with the device library underneath:
//   fsa::IISPHFluidSolver3D / fsa::SESPHFluidSolver3D   IFluidSolverBase (IFluidSolver.hpp:16-34)
//   parameters / settings / stat_last_*                 IISPHFluidSolver3D.hpp:26-33; settings/*.hpp
//   execute_neighborhood_search()                       IISPHFluidSolver3D.hpp:105-118
//   execute_simulation_step(Timepoint&)                 IISPHFluidSolver3D.hpp:120-562
//   get_neighbors(i)                                    CompressedNeighbors.cpp:242-248
//   Timepoint                                           time/Timepoint.hpp:7-14
// Particle data is passed as strided views over component arrays laid out
// like the reference's MovementData3D / ParticleData / ParticleInfo /
// ExternalForces3D (Particle.hpp:22-50); see fsa::ComponentViews.
// Errors raise fsa::Error carrying the C status code.
#ifndef FLUIDSOLVER_AMD_HPP
#define FLUIDSOLVER_AMD_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fluidsolver_amd.h"

namespace fsa {

struct Error : std::runtime_error {
    int32_t code;
    Error(int32_t c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int32_t rc) {
    if (rc != FSA_OK) throw Error(rc, fsa_last_error());
}

// Component records with the reference's layouts (Particle.hpp:22-50).
struct Vec3 {
    float x, y, z;
};
struct MovementData3D {
    Vec3 position, velocity, acceleration;
};
struct ParticleData {
    float mass, pressure, density;
};
struct ParticleInfo {
    uint32_t tag;
    uint8_t type;
};
struct ExternalForces3D {
    Vec3 non_pressure_acceleration;
};
static_assert(sizeof(MovementData3D) == 36, "MovementData3D layout");
static_assert(sizeof(ParticleData) == 12, "ParticleData layout");
static_assert(sizeof(ParticleInfo) == 8, "ParticleInfo layout");
static_assert(sizeof(ExternalForces3D) == 12, "ExternalForces3D layout");

// A minimal ParticleCollection: one vector per component, row i = particle i.
struct ParticleCollection {
    std::vector<MovementData3D> movement;
    std::vector<ParticleData> particle_data;
    std::vector<ParticleInfo> info;
    std::vector<ExternalForces3D> external;

    size_t size() const { return movement.size(); }
    size_t add() {
        movement.push_back({});
        particle_data.push_back({1.0f, 0.0f, 1.0f});
        info.push_back({0u, 0u});
        external.push_back({});
        return movement.size() - 1;
    }
    fsa_particle_view view() {
        fsa_particle_view v{};
        v.count = size();
        v.movement = movement.data();
        v.movement_stride = sizeof(MovementData3D);
        v.particle_data = particle_data.data();
        v.particle_data_stride = sizeof(ParticleData);
        v.info = info.data();
        v.info_stride = sizeof(ParticleInfo);
        v.external = external.data();
        v.external_stride = sizeof(ExternalForces3D);
        return v;
    }
};

class Context {
  public:
    explicit Context(int32_t device = 0, uint64_t hbm_budget_bytes = 0) {
        check(fsa_context_create(device, hbm_budget_bytes, &h_));
    }
    ~Context() {
        if (h_) fsa_context_destroy(h_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    fsa_context handle() const { return h_; }
    void synchronize() { check(fsa_context_synchronize(h_)); }

  private:
    fsa_context h_ = nullptr;
};

struct Timepoint {
    size_t timestep_number = 0;
    float simulation_time = 0.0f;
    float desired_time_step = 0.0f;
    float actual_time_step = 0.0f;
};

struct SimulationParameters {
    float rest_density = 1.0f;
    float gravity = 9.81f;
    float particle_size = 1.0f;
};

struct IISPHSettings3D {
    float max_density_error_allowed = 0.001f;
    size_t min_number_of_iterations = 2;
    size_t max_number_of_iterations = 100;
    float omega = 0.5f;
    float gamma = 0.7f;
    float viscosity = 5.0f;
    bool single_layer_boundary = false;
    float single_layer_boundary_gamma_1 = 1.1f;
};

struct SESPHSettings3D {
    float StiffnessK = 100000.0f;
    float Viscosity = 5.0f;
    bool single_layer_boundary = false;
    float single_layer_boundary_gamma_1 = 1.1f;
    float single_layer_boundary_gamma_2 = 1.1f;
};

struct ConstantTimestepGeneratorSettings {
    float timestep = 0.001f;
    float lambda_v = 0.4f;
    float lambda_a = 0.25f;
};

// Shared part of the device-backed solvers: IFluidSolverBase's call shape.
class DeviceSolverBase {
  public:
    SimulationParameters parameters;
    ConstantTimestepGeneratorSettings timestep_generator;
    ParticleCollection* collection = nullptr;  // data.collection
    bool synchronize_collection = true;        // download after every step

    explicit DeviceSolverBase(Context& ctx) : ctx_(ctx) {}
    virtual ~DeviceSolverBase() {
        if (scene_) fsa_scene_destroy(scene_);
    }
    DeviceSolverBase(const DeviceSolverBase&) = delete;
    DeviceSolverBase& operator=(const DeviceSolverBase&) = delete;

    void initialize() {
        fsa_scene_desc d = desc();
        if (!scene_) {
            fsa_particle_view v = collection->view();
            check(fsa_scene_create(ctx_.handle(), &d, &v, &scene_));
        } else {
            check(fsa_scene_set_desc(scene_, &d));
        }
    }
    void execute_neighborhood_search() {
        if (!scene_) initialize();
        check(fsa_scene_find_neighbors(scene_));
    }
    fsa_step_stats execute_simulation_step(Timepoint& tp) {
        if (!scene_) initialize();
        fsa_step_stats st{};
        check(fsa_scene_simulate(scene_, tp.desired_time_step, &st));
        tp.actual_time_step = st.actual_time_step;
        if (synchronize_collection) download();
        return st;
    }
    void download() {
        fsa_particle_view v = collection->view();
        check(fsa_scene_download(scene_, &v));
    }
    void upload() {
        fsa_particle_view v = collection->view();
        check(fsa_scene_upload(scene_, &v));
    }
    std::vector<uint64_t> get_neighbors(size_t index) const {
        uint64_t n = 0;
        check(fsa_scene_get_neighbors(scene_, index, nullptr, 0, &n));
        std::vector<uint64_t> out(n);
        check(fsa_scene_get_neighbors(scene_, index, out.data(), n, &n));
        return out;
    }

  protected:
    virtual fsa_scene_desc desc() const = 0;
    fsa_scene_desc base_desc(uint32_t solver) const {
        fsa_scene_desc d;
        fsa_scene_desc_defaults(&d, 3, solver);
        d.rest_density = parameters.rest_density;
        d.gravity = parameters.gravity;
        d.particle_size = parameters.particle_size;
        d.timestep = timestep_generator.timestep;
        d.lambda_v = timestep_generator.lambda_v;
        d.lambda_a = timestep_generator.lambda_a;
        return d;
    }
    Context& ctx_;
    fsa_scene scene_ = nullptr;
};

class IISPHFluidSolver3D : public DeviceSolverBase {
  public:
    IISPHSettings3D settings;
    size_t stat_last_iteration_count = 0;
    float stat_last_average_predicted_density_error = 0.0f;

    explicit IISPHFluidSolver3D(Context& ctx) : DeviceSolverBase(ctx) {}

    fsa_step_stats execute_simulation_step(Timepoint& tp) {
        fsa_step_stats st = DeviceSolverBase::execute_simulation_step(tp);
        stat_last_iteration_count = st.iterations;
        stat_last_average_predicted_density_error = (float)st.average_density_error;
        return st;
    }

  protected:
    fsa_scene_desc desc() const override {
        fsa_scene_desc d = base_desc(FSA_SOLVER_IISPH);
        d.iisph_max_density_error_allowed = settings.max_density_error_allowed;
        d.iisph_min_iterations = (uint32_t)settings.min_number_of_iterations;
        d.iisph_max_iterations = (uint32_t)settings.max_number_of_iterations;
        d.iisph_omega = settings.omega;
        d.iisph_gamma = settings.gamma;
        d.iisph_viscosity = settings.viscosity;
        d.iisph_single_layer_boundary = settings.single_layer_boundary ? 1u : 0u;
        d.iisph_single_layer_gamma_1 = settings.single_layer_boundary_gamma_1;
        return d;
    }
};

class SESPHFluidSolver3D : public DeviceSolverBase {
  public:
    SESPHSettings3D settings;

    explicit SESPHFluidSolver3D(Context& ctx) : DeviceSolverBase(ctx) {}

  protected:
    fsa_scene_desc desc() const override {
        fsa_scene_desc d = base_desc(FSA_SOLVER_SESPH);
        d.sesph_stiffness_k = settings.StiffnessK;
        d.sesph_viscosity = settings.Viscosity;
        d.sesph_single_layer_boundary = settings.single_layer_boundary ? 1u : 0u;
        d.sesph_single_layer_gamma_1 = settings.single_layer_boundary_gamma_1;
        d.sesph_single_layer_gamma_2 = settings.single_layer_boundary_gamma_2;
        return d;
    }
};

// Serving loop over the C ABI (fsa_server_*): jobs are copied at submit, admitted
// against the HBM budget and slot limit, stepped together, and collected in the
// caller's row order.
class Server {
  public:
    explicit Server(Context& ctx, uint64_t hbm_budget_bytes = 0, uint32_t max_running = 0) {
        fsa_server_config cfg{};
        cfg.hbm_budget_bytes = hbm_budget_bytes;
        cfg.max_running = max_running;
        check(fsa_server_create(ctx.handle(), &cfg, &h_));
    }
    ~Server() {
        if (h_) fsa_server_destroy(h_);
    }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    uint64_t submit(const fsa_scene_desc& desc, ParticleCollection& c, uint32_t steps) {
        uint64_t id = 0;
        fsa_particle_view v = c.view();
        check(fsa_server_submit(h_, &desc, &v, steps, &id));
        return id;
    }
    uint32_t tick() {
        uint32_t n = 0;
        check(fsa_server_tick(h_, &n));
        return n;
    }
    fsa_server_stats stats() const {
        fsa_server_stats st{};
        check(fsa_server_get_stats(h_, &st));
        return st;
    }
    bool idle() const {
        const fsa_server_stats st = stats();
        return st.jobs_queued == 0 && st.jobs_running == 0;
    }
    fsa_job_info job(uint64_t id) const {
        fsa_job_info info{};
        check(fsa_server_job(h_, id, &info));
        return info;
    }
    void collect(uint64_t id, ParticleCollection& c) {
        fsa_particle_view v = c.view();
        check(fsa_server_collect(h_, id, &v));
    }
    void cancel(uint64_t id) { check(fsa_server_cancel(h_, id)); }

  private:
    fsa_server h_ = nullptr;
};

// Simulator::execute_simulation_step without entities and sensors (Simulator.cpp:7-57).
template <class Solver>
class Simulator {
  public:
    Timepoint timepoint;
    Solver& fluid_solver;
    explicit Simulator(Solver& s) : fluid_solver(s) {}
    void execute_simulation_step() {
        timepoint.desired_time_step = fluid_solver.timestep_generator.timestep;
        timepoint.actual_time_step = timepoint.desired_time_step;
        fluid_solver.execute_neighborhood_search();
        fluid_solver.execute_simulation_step(timepoint);
        timepoint.simulation_time += timepoint.actual_time_step;
        timepoint.timestep_number++;
    }
};

}  // namespace fsa

#endif  // FLUIDSOLVER_AMD_HPP
