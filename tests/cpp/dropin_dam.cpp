// C++ caller of the device library through include/fluidsolver_amd.hpp, shaped
// like a libFluid program: ParticleCollection + IISPHFluidSolver3D + Simulator.
// Builds an nx x ny x nz fluid block on the unit lattice with a 3-layer closed
// box (interior 3nx x 2ny x 2nz), runs `steps` steps and writes the final
// MovementData3D records (caller order) to `out` as raw little-endian floats.
// usage: dropin_dam nx ny nz steps out
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fluidsolver_amd.hpp"

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s nx ny nz steps out\n", argv[0]);
        return 2;
    }
    const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]);
    const int steps = std::atoi(argv[4]);
    const int L[3] = {3 * nx, 2 * ny, 2 * nz};
    fsa::ParticleCollection coll;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                size_t i = coll.add();
                coll.movement[i].position = {(float)x, (float)y, (float)z};
                coll.info[i].type = FSA_TYPE_NORMAL;
            }
    for (int z = -3; z < L[2] + 3; z++)
        for (int y = -3; y < L[1] + 3; y++)
            for (int x = -3; x < L[0] + 3; x++) {
                bool inside = x >= 0 && y >= 0 && z >= 0 && x < L[0] && y < L[1] && z < L[2];
                if (inside) continue;
                size_t i = coll.add();
                coll.movement[i].position = {(float)x, (float)y, (float)z};
                coll.info[i].type = FSA_TYPE_BOUNDARY;
            }
    try {
        fsa::Context ctx(0);
        fsa::IISPHFluidSolver3D solver(ctx);
        solver.collection = &coll;
        fsa::Simulator<fsa::IISPHFluidSolver3D> sim(solver);
        for (int s = 0; s < steps; s++) {
            sim.execute_simulation_step();
            std::printf("step %zu dt %.6g iterations %zu avg_err %.6g\n", sim.timepoint.timestep_number,
                        sim.timepoint.actual_time_step, solver.stat_last_iteration_count,
                        solver.stat_last_average_predicted_density_error);
        }
        std::vector<uint64_t> nb = solver.get_neighbors(0);
        std::printf("particle 0 has %zu neighbors\n", nb.size());
        FILE* f = std::fopen(argv[5], "wb");
        if (!f) return 3;
        std::fwrite(coll.movement.data(), sizeof(fsa::MovementData3D), coll.size(), f);
        std::fclose(f);
    } catch (const fsa::Error& e) {
        std::fprintf(stderr, "fsa error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
