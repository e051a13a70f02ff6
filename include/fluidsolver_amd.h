/*
 * fluidsolver_amd.h — C ABI of the MI355X (gfx950) SPH hot path.
 *
 * The library replaces libFluid's per-timestep hot path: the neighborhood
 * search plus the IISPH and SESPH particle updates.  Every entry point below
 * takes plain pointers and sizes, returns an int32 status (FSA_OK on success)
 * and never throws or aborts across the ABI.  The text of the last error of
 * the calling thread is returned by fsa_last_error().
 *
 * Reference interfaces replaced (paths relative to the FluidSolver tree):
 *   IFluidSolverBase                 src/libFluid/fluidSolver/IFluidSolver.hpp:16-34
 *   NeighborhoodSearch concept       src/libFluid/fluidSolver/IFluidSolver.hpp:36-73
 *   CompressedNeighborhoodSearch     src/libFluid/fluidSolver/neighborhoodSearch/CompressedNeighbors.hpp:20-180
 *   IISPHFluidSolver3D               src/libFluid/fluidSolver/solver/IISPHFluidSolver3D.hpp:22-609
 *   SESPHFluidSolver3D               src/libFluid/fluidSolver/solver/SESPHFluidSolver3D.hpp:12-337
 *   IISPHFluidSolver (2D)            src/libFluid/fluidSolver/solver/IISPHFluidSolver.hpp:21-548
 *   SESPHFluidSolver (2D)            src/libFluid/fluidSolver/solver/SESPHFluidSolver.hpp:49-299
 *   HashedNeighborhoodSearch (2D)    src/libFluid/fluidSolver/neighborhoodSearch/HashedNeighborhoodSearch.cpp:9-161
 *   ConstantTimestepGenerator        src/libFluid/time/ConstantTimestepGenerator.cpp:17-41
 *
 * Particle data crosses the boundary as strided views over the reference's
 * component records (src/libFluid/fluidSolver/Particle.hpp:15-50), so a
 * ParticleCollection's component vectors are handed over without repacking:
 *   MovementData3D {vec3 position, velocity, acceleration}   36 B
 *   MovementData   {vec2 position, velocity, acceleration}   24 B
 *   ParticleData   {float mass, pressure, density}           12 B
 *   ParticleInfo   {uint32 tag; uint8 type}                   8 B (type at offset 4)
 *   ExternalForces3D / ExternalForces {vec3 / vec2}       12 / 8 B
 *
 * Index space: particle indices passed to or returned by the library are the
 * caller's indices (the row order of the uploaded views).  The device keeps
 * its own cell-sorted order internally and maps back on every download.
 *
 * Threading: a context may be used from several host threads; a scene must
 * be driven by one host thread at a time.  All device work of a context is
 * ordered on the context's stream.
 */
#ifndef FLUIDSOLVER_AMD_H
#define FLUIDSOLVER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSA_API __attribute__((visibility("default")))

/* ---- status codes ------------------------------------------------------ */
#define FSA_OK 0
#define FSA_E_INVALID 1      /* bad argument or inconsistent description          */
#define FSA_E_NOMEM 2        /* device or host allocation failed                  */
#define FSA_E_DEVICE 3       /* HIP runtime error                                 */
#define FSA_E_CAPACITY 4     /* a particle has more neighbors than the list holds */
#define FSA_E_STATE 5        /* call order violated (e.g. simulate before search) */
#define FSA_E_BUDGET 6       /* scene does not fit the context's HBM budget       */
#define FSA_E_NONFINITE 7    /* NaN/Inf detected in the particle state            */
#define FSA_E_IO 8           /* a file could not be read or written, or is malformed */

/* ---- enumerations ------------------------------------------------------ */
#define FSA_SOLVER_IISPH 0   /* IISPHFluidSolver3D / IISPHFluidSolver            */
#define FSA_SOLVER_SESPH 1   /* SESPHFluidSolver3D / SESPHFluidSolver            */

/* fsa_scene_desc.neighbor_lists (3D).  The search writes 32-bit neighbor
 * lists in every case; COMPRESSED also sizes every list in the compressed
 * search's NeighborStorage delta code each search (CompressedNeighbors.cpp:260-330)
 * and fails the step with FSA_E_CAPACITY where that fixed-size storage would
 * overflow (more than 96 deltas or 96 payload bytes), as the reference does.
 * DELTA (compressed-3d search only; other searches treat it as COMPRESSED) does
 * the same and also re-encodes each list as 16-bit deltas to the previous
 * neighbor (the compressed search's NeighborStorage idea, CompressedNeighbors.hpp:26-60,
 * in a layout the GPU reads two entries at a time), which the solver walks
 * read instead of the 32-bit entries; the lists and results are the same.
 * A scene cannot switch to or from DELTA after creation. */
#define FSA_LIST_COMPRESSED 0 /* NeighborStorage capacity and sizes (default)    */
#define FSA_LIST_PLAIN 1      /* 97-neighbor capacity only                       */
#define FSA_LIST_DELTA 2      /* COMPRESSED + walks over 16-bit delta-coded lists */

/* fsa_scene_desc.search: the reference search whose neighbor order the device
 * reproduces (registration strings of SolverSerializer.cpp:137-219). */
#define FSA_SEARCH_DEFAULT 0       /* compressed-3d in 3D, hashed in 2D               */
#define FSA_SEARCH_COMPRESSED_3D 1 /* "compressed-3d": CompressedNeighborhoodSearch  */
#define FSA_SEARCH_HASHED_3D 2     /* "hashed-3d": HashedNeighborhoodSearch3D         */
#define FSA_SEARCH_QUADRATIC_3D 3  /* "quadratic-dynamic-allocated-3d" (brute force)  */
#define FSA_SEARCH_HASHED_2D 4     /* "hashed": HashedNeighborhoodSearch (2D)         */
#define FSA_SEARCH_QUADRATIC_2D 5  /* "quadratic-dynamic-allocated" (2D)             */

/* fsa_scene_desc.timestep_generator: how fsa_scene_step / fsa_batch_step /
 * served jobs choose each step's desired time step (Simulator.cpp:17-24). */
#define FSA_TIMESTEP_CONSTANT 0     /* ConstantTimestepGenerator: desc.timestep               */
#define FSA_TIMESTEP_DYNAMIC_CFL 1  /* DynamicCflTimestepGenerator (DynamicCflTimestepGenerator.cpp:6-62):
                                       from the largest |v| and |a| of the active particles,
                                       clamped to [min_timestep, max_timestep]             */

#define FSA_TYPE_NORMAL 0    /* ParticleTypeNormal   (Particle.hpp:9)            */
#define FSA_TYPE_BOUNDARY 1  /* ParticleTypeBoundary (Particle.hpp:10)           */
#define FSA_TYPE_INACTIVE 2  /* ParticleTypeInactive (Particle.hpp:11)           */

/* fsa_scene_desc.flags */
#define FSA_FLAG_EMULATE_INT_ABS 1u /* 2D IISPH: reproduce the integer abs() of the
                                       convergence test (IISPHFluidSolver.hpp:400-402) */

/* Maximum neighbors per particle: first neighbor + 96 deltas, the capacity of
 * CompressedNeighborhoodSearch::NeighborStorage (CompressedNeighbors.hpp:29-34). */
#define FSA_MAX_NEIGHBORS 97

/* ---- scene description -------------------------------------------------- */
/* One struct carries what the reference spreads over
 * IFluidSolverBase::SimulationParameters (IFluidSolver.hpp:18-23),
 * IISPHSettings3D (settings/IISPHSettings3D.hpp:6-20), IISPHSettings
 * (settings/IISPHSettings.hpp:6-15), SESPHSettings3D (settings/SESPHSettings3D.hpp:6-12)
 * and ConstantTimestepGenerator::ConstantTimestepGeneratorSettings
 * (time/ConstantTimestepGenerator.hpp:9-12). fsa_scene_desc_defaults() fills
 * the reference defaults with Simulator's overrides (Simulator.hpp:33-35). */
typedef struct fsa_scene_desc {
    uint32_t dimension;      /* 2 or 3 */
    uint32_t solver;         /* FSA_SOLVER_* */
    uint32_t neighbor_lists; /* FSA_LIST_* */
    uint32_t flags;          /* FSA_FLAG_* */

    float rest_density;
    float gravity; /* applied as (0, -gravity, 0) */
    float particle_size; /* lattice spacing; kernel support = 2 * particle_size */

    float iisph_max_density_error_allowed;
    uint32_t iisph_min_iterations;
    uint32_t iisph_max_iterations;
    float iisph_omega;
    float iisph_gamma;
    float iisph_viscosity;
    uint32_t iisph_single_layer_boundary;
    float iisph_single_layer_gamma_1;

    float sesph_stiffness_k;
    float sesph_viscosity;
    uint32_t sesph_single_layer_boundary;
    float sesph_single_layer_gamma_1;
    float sesph_single_layer_gamma_2;

    float timestep; /* ConstantTimestepGenerator: generated dt */
    float lambda_v;
    float lambda_a;

    uint32_t search;         /* FSA_SEARCH_* */
    uint32_t timestep_generator; /* FSA_TIMESTEP_*                                     */
    float max_timestep;      /* DynamicCflTimestepGenerator settings (DynamicCflTimestepGenerator.hpp:10-17); */
    float min_timestep;      /* max_timestep is also the fallback of the in-step CFL correction */
    float cfl_number;        /* carried for the scenario file; the reference does not use it  */
} fsa_scene_desc;

/* Strided views over the caller's component arrays. Row i of every view is
 * particle i. external may be NULL (treated as zero on upload, skipped on
 * download). For downloads, any NULL pointer is skipped. */
typedef struct fsa_particle_view {
    uint64_t count;
    void* movement;
    uint64_t movement_stride;
    void* particle_data;
    uint64_t particle_data_stride;
    void* info;
    uint64_t info_stride;
    void* external;
    uint64_t external_stride;
} fsa_particle_view;

/* Result of one fsa_scene_simulate(): the reference's Timepoint update
 * (time/Timepoint.hpp:7-14) plus IISPHFluidSolver3D::stat_last_* fields
 * (IISPHFluidSolver3D.hpp:32-33). */
typedef struct fsa_step_stats {
    float desired_time_step;
    float actual_time_step;
    uint32_t iterations;           /* IISPH Jacobi iterations; 0 for SESPH */
    uint32_t status;               /* FSA_OK or the first error raised on device */
    double average_density_error;  /* FP64 average predicted density error */
    float max_velocity;            /* sqrt of the CFL velocity maximum */
    float max_acceleration;        /* sqrt of the CFL acceleration maximum */
} fsa_step_stats;

typedef struct fsa_context_s* fsa_context;
typedef struct fsa_scene_s* fsa_scene;

/* ---- library ----------------------------------------------------------- */
FSA_API int32_t fsa_version(void);
FSA_API const char* fsa_last_error(void);
FSA_API void fsa_scene_desc_defaults(fsa_scene_desc* desc, uint32_t dimension, uint32_t solver);

/* ---- context: device, stream, HBM budget --------------------------------- */
/* hbm_budget_bytes = 0 means "no budget beyond the device". */
FSA_API int32_t fsa_context_create(int32_t device, uint64_t hbm_budget_bytes, fsa_context* out);
FSA_API int32_t fsa_context_destroy(fsa_context ctx);
FSA_API int32_t fsa_context_synchronize(fsa_context ctx);
/* Bytes of device memory currently held by the context's scenes. */
FSA_API int32_t fsa_context_memory_in_use(fsa_context ctx, uint64_t* bytes);
/* Device bytes a scene of this shape would hold (for admission control). */
FSA_API int32_t fsa_scene_memory_estimate(const fsa_scene_desc* desc, uint64_t n_fluid, uint64_t n_boundary,
                                          uint64_t* bytes);

/* ---- scene: IFluidSolverBase + ParticleCollection bridge ---------------- */
/* Creates a scene on the context and uploads the particles of `view`
 * (Simulator::initialize wiring, Simulator.cpp:59-139). Fails with
 * FSA_E_BUDGET when the scene would exceed the context's HBM budget. */
FSA_API int32_t fsa_scene_create(fsa_context ctx, const fsa_scene_desc* desc, const fsa_particle_view* view,
                                 fsa_scene* out);
FSA_API int32_t fsa_scene_destroy(fsa_scene scene);
/* Replaces the particle state (same count and same particle types). */
FSA_API int32_t fsa_scene_upload(fsa_scene scene, const fsa_particle_view* view);
/* Writes the current state back into the caller's views (caller order). */
FSA_API int32_t fsa_scene_download(fsa_scene scene, const fsa_particle_view* view);
/* Writes IISPHParticleData3D / IISPHParticleData records
 * (IISPHFluidSolver3D.hpp:13-19; IISPHFluidSolver.hpp:14-18) of every
 * non-inactive particle into records + i * stride. */
FSA_API int32_t fsa_scene_download_iisph(fsa_scene scene, void* records, uint64_t stride);
/* Updates parameters and settings (DataChangeStruct semantics); the particle
 * counts, the dimension and the neighborhood search must not change
 * (FSA_E_INVALID otherwise: the search is a template argument of the
 * reference solver, IISPHFluidSolver3D.hpp:22-24, and fixes the scene's
 * buffers). */
FSA_API int32_t fsa_scene_set_desc(fsa_scene scene, const fsa_scene_desc* desc);

/* execute_neighborhood_search() — CompressedNeighborhoodSearch::find_neighbors
 * (CompressedNeighbors.cpp:168-240). Lists are built for fluid particles. */
FSA_API int32_t fsa_scene_find_neighbors(fsa_scene scene);
/* get_neighbors(i) (CompressedNeighbors.cpp:242-248): ascending caller
 * indices of the neighbors of fluid particle `index`, self included. */
FSA_API int32_t fsa_scene_get_neighbors(fsa_scene scene, uint64_t index, uint64_t* out, uint64_t capacity,
                                        uint64_t* count);
/* All fluid lists at once in CSR form over caller indices: row_offsets has
 * view.count + 1 entries; rows of non-fluid particles are empty. */
FSA_API int32_t fsa_scene_export_neighbors(fsa_scene scene, uint64_t* row_offsets, uint64_t* columns,
                                           uint64_t capacity, uint64_t* nnz);
/* Encoded neighbor-list bytes of the last search (reference codec size model
 * and the device's own), for roofline accounting. */
FSA_API int32_t fsa_scene_neighbor_stats(fsa_scene scene, uint64_t* total_neighbors, uint64_t* device_list_bytes,
                                         uint64_t* reference_codec_bytes);

/* execute_simulation_step(Timepoint&) (IISPHFluidSolver3D.hpp:120-562;
 * SESPHFluidSolver3D.hpp:49-130; 2D: IISPHFluidSolver.hpp:139-180,
 * SESPHFluidSolver.hpp:49-105). Requires a prior fsa_scene_find_neighbors. */
FSA_API int32_t fsa_scene_simulate(fsa_scene scene, float desired_time_step, fsa_step_stats* stats);
/* `steps` x (find_neighbors + simulate), as Simulator::execute_simulation_step
 * does (Simulator.cpp:7-57) without entities/sensors; desired dt from the
 * scene's timestep generator (desc.timestep_generator). stats (may be NULL)
 * receives the last step. */
FSA_API int32_t fsa_scene_step(fsa_scene scene, uint32_t steps, fsa_step_stats* stats);
/* The desired time step the scene's generator gives for its current state
 * (TimestepGenerator::generate_next_timestep + get_generated_timestep,
 * ConstantTimestepGenerator.cpp:4-7, DynamicCflTimestepGenerator.cpp:46-62). */
FSA_API int32_t fsa_scene_next_timestep(fsa_scene scene, float* dt);

/* ---- batching: many scenes advanced together ------------------------------ */
/* Advances `n` scenes of one context by `steps` steps each. All scenes are
 * advanced in the same kernel launches over a segmented layout; every scene
 * keeps its own Jacobi convergence, CFL dt and statistics, and its results
 * are identical to stepping it alone. stats (may be NULL) holds n entries. */
FSA_API int32_t fsa_batch_step(fsa_context ctx, const fsa_scene* scenes, uint32_t n, uint32_t steps,
                               fsa_step_stats* stats);
/* With graphs on (0 = off, the default), the head of every batched step and of
 * fsa_scene_step (search, solver phases and the predicted number of Jacobi
 * iterations) is captured once per scene list and launch configuration into
 * a HIP graph and replayed; further iterations and the integration are
 * launched directly.  Results are identical either way. */
FSA_API int32_t fsa_batch_set_graph(fsa_context ctx, uint32_t enabled);
/* Autotuning (no counterpart in the reference): on the first batched step of
 * a size class (kernel spec, total fluid rounded to a power of two) the
 * neighbor-walk variant of each 3D IISPH kernel family is chosen by timing the
 * candidates on the live neighbor lists, writing only to scratch copies.  The
 * candidates are FSA_WALK_* below: the unroll of the neighbor loop, with or
 * without an occupancy hint (at least that many waves per SIMD).  Every
 * variant gives the same results bit for bit.  Enabled by default; the choices
 * and their timings are reported as text. */
#define FSA_WALK_U2 0x02u
#define FSA_WALK_U4 0x04u
#define FSA_WALK_U8 0x08u
#define FSA_WALK_U4_W8 0x84u
#define FSA_WALK_U8_W6 0x68u
FSA_API int32_t fsa_context_set_autotune(fsa_context ctx, uint32_t enabled);
FSA_API int32_t fsa_context_get_tuning(fsa_context ctx, char* buf, uint64_t cap, uint64_t* len);
/* Forces one FSA_WALK_* variant on every kernel family and skips the
 * autotuner (0 hands the choice back to it).  Other values: FSA_E_INVALID.
 * No counterpart in the reference: for callers that want a fixed launch
 * configuration (reproducible timings, a graph captured once). */
FSA_API int32_t fsa_context_set_walk_variant(fsa_context ctx, uint32_t variant);
/* Threads per block of the per-particle kernels: 0 (the default) or 256, the
 * one shape this build's block maps are laid out for. */
FSA_API int32_t fsa_context_set_block_size(fsa_context ctx, uint32_t threads);

/* ---- serving: continuous batching of scene jobs ---------------------------- */
/* A server owns a FIFO of submitted scene jobs.  Each fsa_server_tick admits
 * queued jobs while the HBM budget and the slot limit allow (admission is by
 * fsa_scene_memory_estimate), advances every running scene by one step in
 * shared batched launches, and retires finished jobs: their final state moves
 * to host memory and their device memory is released, so jobs that arrive
 * while others run are admitted as memory frees up.  Results of a job are
 * identical to running it alone. */
#define FSA_JOB_QUEUED 0
#define FSA_JOB_RUNNING 1
#define FSA_JOB_DONE 2
#define FSA_JOB_FAILED 3
#define FSA_JOB_CANCELLED 4

typedef struct fsa_server_config {
    uint64_t hbm_budget_bytes;   /* admission budget; 0 = the context's budget */
    uint32_t max_running;        /* slot limit; 0 = no limit */
    /* 0: while a tick steps the running scenes (on a worker thread, context
     * stream), the tick uploads admitted jobs and downloads the jobs whose steps
     * ended in the previous tick on a second stream; admitted jobs take their
     * first step in the next tick.  1: admit, step and retire one after another. */
    uint32_t serial_ticks;
    uint32_t reserved[4];
} fsa_server_config;

typedef struct fsa_job_info {
    uint64_t job_id;
    uint32_t state;              /* FSA_JOB_* */
    uint32_t status;             /* FSA_OK or the error that failed the job */
    uint32_t steps_done;
    uint32_t steps_total;
    uint64_t fluid_particles;
    double submit_s, admit_s, finish_s;  /* server clock, seconds */
    fsa_step_stats last;
} fsa_job_info;

typedef struct fsa_server_stats {
    uint64_t ticks;
    uint64_t scene_steps;        /* sum over ticks of running scenes */
    uint64_t particle_updates;   /* fluid particles x steps */
    uint64_t jobs_submitted, jobs_done, jobs_failed, jobs_queued, jobs_running;
    uint64_t hbm_in_use;
    double busy_s;               /* time spent inside ticks */
    double p50_step_ms, p99_step_ms, mean_step_ms, max_step_ms;  /* per scene-step latency */
    double admit_s, step_s, retire_s;  /* time inside ticks spent admitting, stepping and retiring jobs */
} fsa_server_stats;

typedef struct fsa_server_s* fsa_server;

/* A server is destroyed with its context if not destroyed before. */
FSA_API int32_t fsa_server_create(fsa_context ctx, const fsa_server_config* cfg, fsa_server* out);
FSA_API int32_t fsa_server_destroy(fsa_server srv);
/* Copies the particles of the job; the caller's view may be released after the call. */
FSA_API int32_t fsa_server_submit(fsa_server srv, const fsa_scene_desc* desc, const fsa_particle_view* view,
                                  uint32_t steps, uint64_t* job_id);
/* A queued job, or a running one with steps left, is CANCELLED; a running job
 * whose steps have all run is retired DONE. */
FSA_API int32_t fsa_server_cancel(fsa_server srv, uint64_t job_id);
/* Admit, step every running scene once, retire; *running receives the number
 * of scenes stepped (0 when nothing could run).  With overlapped ticks a job
 * whose last step ran in this tick is retired (DONE) in the next one. */
FSA_API int32_t fsa_server_tick(fsa_server srv, uint32_t* running);
FSA_API int32_t fsa_server_job(fsa_server srv, uint64_t job_id, fsa_job_info* info);
/* Forgets a job in a terminal state.  For a DONE job, first writes its final
 * state into the caller's view (caller order; view may be NULL); a FAILED or
 * CANCELLED job has no final state and view is not written.  FSA_E_STATE for a
 * queued or running job. */
FSA_API int32_t fsa_server_collect(fsa_server srv, uint64_t job_id, const fsa_particle_view* view);
FSA_API int32_t fsa_server_get_stats(fsa_server srv, fsa_server_stats* stats);

/* ---- timing -------------------------------------------------------------- */
/* Average device time (ms) of each kernel class since the last reset, measured
 * with HIP events on the context stream when timing is enabled. */
#define FSA_TIMER_SEARCH 0
#define FSA_TIMER_DENSITY 1   /* IISPH phase 1 / SESPH density+pressure */
#define FSA_TIMER_DIAGONAL 2  /* IISPH phase 2 / SESPH accelerations     */
#define FSA_TIMER_PRESSURE_A 3
#define FSA_TIMER_PRESSURE_B 4
#define FSA_TIMER_INTEGRATE 5
#define FSA_TIMER_PRESSURE_FIRST 6  /* the first Jacobi iteration (3D IISPH: no neighbor walk) */
#define FSA_TIMER_LISTS 7  /* 3D compressed search: the list kernel alone (inside FSA_TIMER_SEARCH) */
#define FSA_TIMER_PRESSURE_TAIL 8  /* 3D IISPH: the gated Jacobi pass beyond the predicted ones (passes A and B) */
#define FSA_TIMER_COUNT 9
FSA_API int32_t fsa_context_set_timing(fsa_context ctx, uint32_t enabled);
FSA_API int32_t fsa_context_get_timing(fsa_context ctx, double* total_ms, uint64_t* launches, uint32_t n);

/* ---- entities and sensors around the solver ------------------------------ */
/* Simulation entities run inside fsa_scene_step / fsa_batch_step / served
 * jobs in the reference's order (Simulator.cpp:26-45): after the neighborhood
 * search, the entities whose execution point is before the solver, in list
 * order; after the solver, the others.  fsa_scene_simulate runs them too. */
#define FSA_ENTITY_BOUNDARY_PREPROCESSOR 1    /* BoundaryPreprocessor (BoundaryPreprocessor.cpp:7-48, 96-117):
                                                 before the solver, each boundary particle's mass becomes
                                                 min(rest_density * particle_size^3,
                                                     rest_density * gamma_1 / sum of W over its boundary
                                                     neighbors), gamma_1 from a perfectly sampled plane; 3D */
#define FSA_ENTITY_VELOCITY_ALTERATION_BY_TAG 2 /* VelocityAlterationByTag (VelocityAlterationByTag.cpp:7-22):
                                                 before the solver, particles with `tag` get `velocity`; 3D */
#define FSA_ENTITY_PARTICLE_REMOVER_3D 3      /* ParticleRemover3D (ParticleRemover3D.cpp:8-16, VolumeGroup.cpp:43-58):
                                                 after the solver, particles inside (or, with
                                                 remove_if_outside, outside) the box become inactive */

typedef struct fsa_entity {
    uint32_t type;                 /* FSA_ENTITY_*                                          */
    uint32_t tag;                  /* velocity alteration: the tag it applies to            */
    uint32_t remove_if_outside;    /* particle remover: 1 = remove outside the volume        */
    uint32_t reserved;
    float velocity[3];             /* velocity alteration                                   */
    float center[3];               /* particle remover: Volume {center, distance_from_center} */
    float distance_from_center[3];
} fsa_entity;

/* Replaces the scene's entity list (count 0 clears it). */
FSA_API int32_t fsa_scene_set_entities(fsa_scene scene, const fsa_entity* entities, uint32_t count);

/* The global sensors over the scene's current state, reduced on the device
 * (ParticleStatistics.cpp:17-207): density, pressure and |v| over normal
 * particles, kinetic and potential energy, particle counts.  Maxima start
 * at FLT_MIN and minima at FLT_MAX, as the reference's MaxMinAvgSensorData
 * does (ParticleStatistics.hpp:10-14).  Sums are formed in FP64 (the
 * reference adds in FP32 in collection order); maxima and minima are exact. */
typedef struct fsa_scene_statistics {
    uint64_t normal_particles;
    uint64_t boundary_particles;
    uint64_t inactive_particles;
    float density_max, density_min;
    double density_avg;
    float pressure_max, pressure_min;
    double pressure_avg;
    float velocity_max, velocity_min;  /* |v| (glm::length) */
    double velocity_avg;
    double kinetic_energy;             /* sum of 0.5f * m * |v| * |v|                 */
    double potential_energy;           /* sum of gravity * m * (y - relative_zero_height) */
} fsa_scene_statistics;

FSA_API int32_t fsa_scene_get_statistics(fsa_scene scene, float relative_zero_height, fsa_scene_statistics* out);
/* fsa_scene_get_statistics without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code in a tight
 * serving loop, harnesses that forbid hipMalloc).  The per-block partials and
 * the result live in a fixed device array the library holds on each GPU until
 * the process ends (65,537 x 80 bytes, 5.2 MB); scenes of more than 16,777,216
 * resident particles are refused with FSA_E_INVALID.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (several host threads, or
 * contexts on different streams) race and must be serialized by the caller.
 * Results equal fsa_scene_get_statistics bit for bit. */
FSA_API int32_t fsa_scene_get_statistics_noalloc(fsa_scene scene, float relative_zero_height,
                                                 fsa_scene_statistics* out);

/* ---- position queries (3D) ------------------------------------------------ */
/* NeighborhoodInterface::get_neighbors(position) (NeighborhoodInterface.cpp; the
 * position iterators of CompressedNeighbors.cpp:529-595,
 * HashedNeighborhoodSearch3D.cpp:272-349, QuadraticNeighborhoodSearch3D.cpp:199-219)
 * for many positions at once, on the device.  The particles near each
 * position in the scene search's iteration order, as caller rows:
 *   compressed  (p - x).(p - x) <= search_radius^2, 27 cells by ascending key;
 *   hashed      length(x - p) <= search_radius, cells dx-outer, members ascending;
 *   quadratic   length(x - p) <= search_radius, ascending particle index.
 * Queries see the cell structure of the last search and the current positions:
 * the reference's state between a step and the next search.  Inactive particles
 * are never reported (the reference's quadratic iterator lists them; its
 * callers either skip them or, SensorPlane, would weigh them at their old
 * positions).  FSA_E_STATE when the layout changed since the last search (an
 * upload, or particles a remover took out): call fsa_scene_find_neighbors.
 * positions: count x 3 floats, finite.  offsets (count + 1 entries, may be
 * NULL) receives the CSR offsets, rows (may be NULL: count only) the rows;
 * *nnz = total.  FSA_E_INVALID when capacity < *nnz (offsets and *nnz are
 * still written). */
FSA_API int32_t fsa_scene_query_neighbors(fsa_scene scene, const float* positions, uint64_t count, uint64_t* offsets,
                                          uint64_t* rows, uint64_t capacity, uint64_t* nnz);

/* SensorPlane (SensorPlane.cpp:75-137; settings SensorPlane.hpp:17-37): a grid
 * of samples on a plane, each the mean over (sub_sample_grid_size + 1)^2
 * sub-samples of sum over the neighbors (as fsa_scene_query_neighbors) of
 * (mass, 0, 0), velocity or (0, 0, pressure) times W(|sample - p|), in float in
 * the search's order.  values: number_of_samples_x * number_of_samples_y x 3
 * floats, sample i at row y = number_of_samples_y - 1 - i / number_of_samples_x,
 * column x = i % number_of_samples_x (the reference's output order). */
#define FSA_SENSOR_PLANE_DENSITY 0
#define FSA_SENSOR_PLANE_VELOCITY 1
#define FSA_SENSOR_PLANE_PRESSURE 2
typedef struct fsa_sensor_plane {
    float origin[3];
    float span_x[3];              /* normalized by the call                        */
    float span_y[3];
    float width, height;
    uint64_t number_of_samples_x, number_of_samples_y;
    float max_image_value, min_image_value;  /* image range (not used by the values) */
    uint32_t sensor_type;         /* FSA_SENSOR_PLANE_*                             */
    uint32_t reserved;
    uint64_t sub_sample_grid_size;
    uint64_t calculate_plane_every_nth_step; /* not used by the call                */
} fsa_sensor_plane;
FSA_API void fsa_sensor_plane_defaults(fsa_sensor_plane* plane);
FSA_API int32_t fsa_scene_sensor_plane(fsa_scene scene, const fsa_sensor_plane* plane, float* values);

/* ParticleIntersectionAccelerator::evaluate_volume_at_position
 * (ParticleIntersectionAccelerator.cpp:97-131), the raytracer's volume probe:
 * over the neighbors of each position, in the search's order, the density
 * sum of m W(|x - p|), split into fluid-only and boundary-only parts, and the
 * normals sum of m * GetKernelDerivativeValue(p, x) by type.  The raytracer
 * builds its own compressed search over the current positions: call
 * fsa_scene_find_neighbors on a compressed-search scene first for the same
 * neighbor order. */
typedef struct fsa_volume_sample {
    float density;
    float boundary_only_density;
    float fluid_only_density;
    float fluid_only_normal[3];
    float boundary_only_normal[3];
} fsa_volume_sample;
FSA_API int32_t fsa_scene_evaluate_volume(fsa_scene scene, const float* positions, uint64_t count,
                                          fsa_volume_sample* out);

/* ---- scenario files ----------------------------------------------------- */
/* Particle data files (.data): an LZ4 frame holding u32 version = 1, u32 number
 * of components, u32 component ids, u64 particle count, then one little-endian
 * record per particle with the components in id order
 * (ParticleSerializer::serialize / deserialize, ParticleSerializer.cpp:41-141;
 * Lz4CompressedStream.cpp:7-187; EndianSafeBinaryStream.hpp:29-30).
 * Component ids and their record fields: */
#define FSA_COMPONENT_MOVEMENT_2D 1  /* vec2 position, velocity, acceleration        */
#define FSA_COMPONENT_MOVEMENT_3D 2  /* vec3 position, velocity, acceleration        */
#define FSA_COMPONENT_INFO 3         /* u8 type, u32 tag                             */
#define FSA_COMPONENT_PARTICLE_DATA 4 /* float density, mass, pressure (file order)  */
#define FSA_COMPONENT_EXTERNAL_2D 5  /* vec2 non-pressure acceleration               */
#define FSA_COMPONENT_EXTERNAL_3D 6  /* vec3 non-pressure acceleration               */

typedef struct fsa_data_header {
    uint32_t version;     /* 1 */
    uint32_t dimension;   /* 2 or 3, from the movement component (0 if none)      */
    uint32_t components;  /* bit (id - 1) set for each component id in the file   */
    uint32_t reserved;
    uint64_t count;       /* particles                                            */
} fsa_data_header;

/* Reads the header of a particle file. */
FSA_API int32_t fsa_data_read_header(const char* path, fsa_data_header* out);
/* Reads a particle file into the caller's view (view->count must equal the
 * file's count; movement rows are 3 * dimension floats).  Components missing
 * from the file are zeroed in the view; NULL view pointers skip a component. */
FSA_API int32_t fsa_data_read(const char* path, const fsa_particle_view* view);
/* Writes the view as a particle file of the given dimension, with the
 * movement, info and particle-data components and, when view->external is
 * not NULL, the external-force component (ParticleSerializer::serialize). */
FSA_API int32_t fsa_data_write(const char* path, uint32_t dimension, const fsa_particle_view* view);

/* Scenario files: JSON version 2 with the "scenario", "solver" and
 * "timestep-generator" nodes (MainSerializer.cpp:114-171; ScenarioSerializer.cpp:43-79;
 * SolverSerializer.cpp:137-259; TimestepGeneratorSerializer.cpp:47-72).
 * Sensors and the visualizer are not simulated by this library, nor the 2D
 * particle remover and spawner entities; their presence is reported. */
typedef struct fsa_scenario {
    fsa_scene_desc desc;          /* dimension, solver, search, settings and timestep generator */
    uint32_t entities;            /* entries in scenario.entities                            */
    uint32_t sensors;             /* entries in scenario.sensors (not simulated here)        */
    uint32_t entity_count;        /* the entities this library runs, in file order
                                     (particle-remover-3d, boundary-preprocessor,
                                     velocity-alteration-by-tag; EntitySerializer.cpp:66-173) */
    uint32_t reserved;
    fsa_entity entity_list[32];
    char particles[1024];         /* scenario.particles as written in the file               */
    char particles_path[4096];    /* the file the reference opens for it
                                     (MainSerializer::get_full_particle_data_path, :216-223)  */
} fsa_scenario;

/* Parses a scenario file.  Missing required keys, unknown solver / search /
 * kernel combinations and versions other than 2 are errors. */
FSA_API int32_t fsa_scenario_read(const char* json_path, fsa_scenario* out);
/* Writes a scenario file (version 2, visualizer "no-visualizer") naming
 * in->particles as its particle file, with in->entity_list as its entities. */
FSA_API int32_t fsa_scenario_write(const char* json_path, const fsa_scenario* in);
/* Scene defaults for the dimension / solver and no particle file. */
FSA_API void fsa_scenario_defaults(fsa_scenario* out, uint32_t dimension, uint32_t solver);

/* LZ4 frame codec and xxHash32 used by the particle files (the lz4 library's
 * frame format, which Lz4CompressedStream drives through LZ4F_compressUpdate /
 * LZ4F_decompress).  Decoding accepts linked and independent blocks, block and
 * content checksums, a content size and skippable frames; encoding writes
 * independent 64 KiB blocks with a content checksum.  *out_size receives the
 * bytes produced; FSA_E_CAPACITY when they exceed capacity (*out_size is then
 * the size needed). */
FSA_API int32_t fsa_lz4_frame_decode(const void* src, uint64_t src_size, void* dst, uint64_t capacity,
                                     uint64_t* out_size);
FSA_API int32_t fsa_lz4_frame_encode(const void* src, uint64_t src_size, void* dst, uint64_t capacity,
                                     uint64_t* out_size);
FSA_API uint32_t fsa_xxh32(const void* data, uint64_t size, uint32_t seed);

#ifdef __cplusplus
}
#endif
#endif /* FLUIDSOLVER_AMD_H */
