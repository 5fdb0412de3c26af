// model_driver.cpp — a Model::Step-style host driver (simulation/model.cpp:7-78)
// written against the drop-in GPU_Implementation5 shim, unchanged in shape:
// per frame clear_force_accumulator, then reset_grid / p2g / update_nodes /
// g2p / record_timings per substep, then transfer_from_device, the frame
// arrays (SnapshotManager::PrepareFrameArrays, on the device) and the timing
// table.  Scene: a rectangular ice sheet, 4 points per cell, on a grid with a
// two-node wall ring, pushed by a uniform wind stress.
//
//   model_driver [gx gy frames steps_per_frame wind_Pa]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "gpu_implementation5.hpp"

using namespace icempm;

int main(int argc, char **argv) {
    const int gx = argc > 1 ? std::atoi(argv[1]) : 256;
    const int gy = argc > 2 ? std::atoi(argv[2]) : 128;
    const int frames = argc > 3 ? std::atoi(argv[3]) : 2;
    const int steps_per_frame = argc > 4 ? std::atoi(argv[4]) : 100;
    const double wind = argc > 5 ? std::atof(argv[5]) : 0.1;
    try {
        SimParams prms;
        prms.GridXTotal = gx;
        prms.GridYTotal = gy;
        prms.cellsize = 1.0;
        prms.InitialTimeStep = 1e-4;
        prms.AnimationFramePeriod = steps_per_frame * 1e-4 * (1 + 1e-9);
        prms.YoungsModulus = 1e9;
        prms.IceTensileStrength = prms.IceTensileStrength2 = 3e6;
        prms.IceShearStrength = 2.5e6;
        prms.DP_phi = 30;
        prms.DP_threshold_p = -1e5 / std::tan(30 * M_PI / 180);
        prms.RidgeFormationCoeff = 0.5;
        prms.waterDragEffectiveLinear = 0;
        prms.waterDragEffectiveQuadratic = 1030 * 0.005 / 916.0;
        prms.windStressToAccel = 1 / 916.0;

        // points: 2x2 jittered per cell over the left 60 % (snapshotmanager.cpp:150-270 conventions)
        std::mt19937 rng(12345);
        std::uniform_real_distribution<float> U(0.f, 1.f);
        std::vector<float> X, Y;
        const int n_gen = 4 * (gx - 1) * (gy - 1);
        for (int i = 0; i < gx - 1; i++)
            for (int j = 0; j < gy - 1; j++)
                for (int k = 0; k < 4; k++) {
                    float x = i + ((k / 2) + U(rng)) * 0.5f, y = j + ((k % 2) + U(rng)) * 0.5f;
                    int xi = (int)(x + 0.5f), yi = (int)(y + 0.5f);
                    if (x >= 0.6f * (gx - 1) || xi <= 1 || yi <= 1 || xi >= gx - 2 || yi >= gy - 2) continue;
                    X.push_back(x);
                    Y.push_back(y);
                }
        prms.ParticleVolume = prms.cellsize * prms.cellsize * gx * gy / n_gen;
        prms.ComputeHelperVariables();

        GPU_Implementation5 gpu(&prms);
        gpu.allocate_host_arrays_grid();
        for (int i = 0; i < gx; i++)
            for (int j = 0; j < gy; j++)
                if (i < 2 || j < 2 || i >= gx - 2 || j >= gy - 2) gpu.grid_status_buffer[j + i * gy] = 1;
        gpu.allocate_host_arrays_points((int)X.size());
        gpu.hssoa.size = (unsigned)X.size();
        for (unsigned k = 0; k < gpu.hssoa.size; k++) {
            int xi = (int)(X[k] + 0.5f), yi = (int)(Y[k] + 0.5f);
            gpu.hssoa.setValueInt(SimParams::idx_utility_data, k, 0);
            gpu.hssoa.setValueInt(SimParams::integer_cell_idx, k, ((uint32_t)yi << 16) | (uint32_t)xi);
            gpu.hssoa.setValueInt(SimParams::integer_point_idx, k, k);
            gpu.hssoa.getPointerToLine(SimParams::idx_Jp_inv)[k] = 1.f;
            gpu.hssoa.getPointerToLine(SimParams::posx)[k] = X[k] - xi;
            gpu.hssoa.getPointerToLine(SimParams::posx + 1)[k] = Y[k] - yi;
            gpu.hssoa.getPointerToLine(SimParams::Fe00)[k] = 1.f;
            gpu.hssoa.getPointerToLine(SimParams::Fe00 + 3)[k] = 1.f;
            gpu.hssoa.getPointerToLine(SimParams::idx_thickness)[k] = 1.f;
        }
        gpu.wind_stress_field.assign(2 * (size_t)gx * gy, 0.f);
        for (size_t i = 0; i < (size_t)gx * gy; i++) gpu.wind_stress_field[i] = (float)wind;

        // SnapshotManager::SplitIntoPartitionsAndTransferToDevice + Model::Prepare
        gpu.initialize(0);
        gpu.split_hssoa_into_partitions();
        gpu.allocate_device_arrays();
        gpu.transfer_to_device();
        gpu.update_constants();
        gpu.transfer_wind_and_current_data_to_device();

        int SimulationStep = 0;
        double simulation_time = 0;
        FrameArrays frame;
        for (int f = 0; f < frames; f++) {
            gpu.reset_timings();
            gpu.clear_force_accumulator();
            int count = 0;
            do {
                const int step = SimulationStep + count;
                simulation_time = prms.InitialTimeStep * step;
                gpu.reset_grid();
                gpu.p2g();
                gpu.update_nodes((float)simulation_time, 0, 0);
                gpu.g2p(step % prms.UpdateEveryNthStep == 0, false, 0);
                count++;
            } while ((SimulationStep + count) % prms.UpdateEveryNthStep != 0);
            gpu.record_timings(false);
            gpu.transfer_from_device(count);
            SimulationStep += count;
            GPU_Partition &p = gpu.partitions.front();
            p.normalize_timings(count);
            double vx = 0;
            const float *v = gpu.hssoa.getPointerToLine(SimParams::velx);
            for (unsigned k = 0; k < gpu.hssoa.size; k++) vx += v[k];
            std::printf("frame %d step %d t %.4f pts %zu dis %d err 0x%x | p2g %.1f upd %.1f g2p %.1f tot %.1f us | "
                        "mean vx %.3e\n",
                        f, SimulationStep, simulation_time, p.nPts_partition, p.disabled_points_count, gpu.error_code,
                        p.timing_10_P2GAndHalo, p.timing_30_updateGrid, p.timing_40_G2P, p.timing_stepTotal,
                        vx / std::max(1u, gpu.hssoa.size));
            if (gpu.error_code) return 2;
            frame.prepare(gpu, count);
            size_t ice_nodes = 0;
            double mass = 0;
            for (size_t i = 0; i < frame.mass_mask.size(); i++) {
                ice_nodes += frame.mass_mask[i];
                mass += frame.vis_mass[i];
            }
            std::printf("frame %d arrays: ice nodes %zu, mass on water nodes %.6e, wall force region 1 (%.4e, %.4e)\n",
                        f, ice_nodes, mass, frame.forces_per_region[1][0], frame.forces_per_region[1][1]);
        }
        gpu.synchronize();
        std::printf("model_driver done\n");
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "model_driver: %s\n", e.what());
        return 1;
    }
}
