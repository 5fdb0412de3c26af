// model_driver.cpp — a Model::Step-style host driver (simulation/model.cpp:7-78)
// written against the drop-in GPU_Implementation5 shim, unchanged in shape:
// per frame clear_force_accumulator, then reset_grid / p2g / update_nodes /
// g2p / record_timings per substep, then transfer_from_device, the frame
// arrays (SnapshotManager::PrepareFrameArrays, on the device) and the timing
// table.  Scene: a rectangular ice sheet, 4 points per cell, on a grid with a
// two-node wall ring, pushed by a uniform wind stress, or read from a scene
// file; the final points can be written out.
//
//   model_driver [gx gy frames steps_per_frame wind_Pa [scene_in|- [points_out]]]
//
// Scene file (little-endian): int32 gx, gy, nlines, n; float64
// ParticleVolume; uint8 status[gx*gy]; float32 wind stress [2][gx*gy];
// float32 points [nlines][n] (the hssoa lines).  Points file: int32 nlines, n;
// float32 [nlines][n].
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

#include "gpu_implementation5.hpp"

using namespace icempm;

int main(int argc, char **argv) {
    const int gx = argc > 1 ? std::atoi(argv[1]) : 256;
    const int gy = argc > 2 ? std::atoi(argv[2]) : 128;
    const int frames = argc > 3 ? std::atoi(argv[3]) : 2;
    const int steps_per_frame = argc > 4 ? std::atoi(argv[4]) : 100;
    const double wind = argc > 5 ? std::atof(argv[5]) : 0.1;
    const std::string scene_in = argc > 6 ? argv[6] : "-";
    const std::string points_out = argc > 7 ? argv[7] : "";
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

        GPU_Implementation5 gpu(&prms);
        if (scene_in != "-") {
            FILE *f = std::fopen(scene_in.c_str(), "rb");
            if (!f) throw std::runtime_error("cannot open " + scene_in);
            int32_t hdr[4];
            double vol;
            bool ok = std::fread(hdr, sizeof(hdr), 1, f) == 1 && std::fread(&vol, sizeof(vol), 1, f) == 1;
            if (!ok || hdr[0] != gx || hdr[1] != gy || hdr[2] < ICE_N_LINES || hdr[2] > ICE_N_LINES_MAX || hdr[3] < 0)
                throw std::runtime_error("scene file header does not match the grid");
            const size_t N = (size_t)gx * gy, n = (size_t)hdr[3];
            prms.ParticleVolume = vol;
            if (hdr[2] == ICE_N_LINES_MAX) prms.flags |= ICE_FLAG_COHESION_LINE;
            prms.ComputeHelperVariables();
            gpu.allocate_host_arrays_grid();
            gpu.wind_stress_field.assign(2 * N, 0.f);
            ok = std::fread(gpu.grid_status_buffer.data(), 1, N, f) == N &&
                 std::fread(gpu.wind_stress_field.data(), sizeof(float), 2 * N, f) == 2 * N;
            gpu.allocate_host_arrays_points((int)n, hdr[2]);
            gpu.hssoa.size = (unsigned)n;
            for (int l = 0; ok && l < hdr[2]; l++)
                ok = std::fread(gpu.hssoa.getPointerToLine(l), sizeof(float), n, f) == n;
            std::fclose(f);
            if (!ok) throw std::runtime_error("scene file truncated");
        } else {
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
        }

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
            gpu.synchronize();
            const auto t0 = std::chrono::steady_clock::now();
            do {
                const int step = SimulationStep + count;
                simulation_time = prms.InitialTimeStep * step;
                gpu.reset_grid();
                gpu.p2g();
                gpu.update_nodes((float)simulation_time, 0, 0);
                gpu.g2p(step % prms.UpdateEveryNthStep == 0, false, 0);
                count++;
            } while ((SimulationStep + count) % prms.UpdateEveryNthStep != 0);
            gpu.synchronize();
            const double wall_ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            std::printf("frame %d wall %.4f ms per substep\n", f, wall_ms / count);
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
        if (!points_out.empty()) {
            FILE *f = std::fopen(points_out.c_str(), "wb");
            if (!f) throw std::runtime_error("cannot write " + points_out);
            const int32_t hdr[2] = {gpu.hssoa.nlines, (int32_t)gpu.hssoa.size};
            bool ok = std::fwrite(hdr, sizeof(hdr), 1, f) == 1;
            for (int l = 0; ok && l < gpu.hssoa.nlines; l++)
                ok = std::fwrite(gpu.hssoa.getPointerToLine(l), sizeof(float), gpu.hssoa.size, f) == gpu.hssoa.size;
            if (std::fclose(f) != 0 || !ok) throw std::runtime_error("writing " + points_out + " failed");
        }
        std::printf("model_driver done\n");
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "model_driver: %s\n", e.what());
        return 1;
    }
}
