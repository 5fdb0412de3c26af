// gpu_implementation5.hpp — C++ drop-in for the reference's GPU_Partition and
// GPU_Implementation5 (simulation/gpu_partition.h:63-134,
// simulation/gpu_implementation5.h:23-68) over the C-ABI of icempm.h.
//
// Header-only.  Same class, member and method names and the same argument
// meaning; failures throw std::runtime_error as the reference does.  One
// partition on one GPU; the host SoA is a float [nPtsArrays][capacity] buffer
// like HostSideSOA (hssoa/host_side_soa.h:51-74) with t_PointReal = float.
#pragma once
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/icempm.h"

namespace icempm {

// SimParams (parameters_sim.h:36-127): the substep's fields, FP64.
struct SimParams : IceParams {
    static constexpr int nPtsArrays = ICE_N_LINES;
    static constexpr int idx_utility_data = ICE_IDX_UTILITY, integer_cell_idx = ICE_IDX_CELL,
                         integer_point_idx = ICE_IDX_POINT, idx_P = ICE_IDX_P, idx_Q = ICE_IDX_Q,
                         idx_Jp_inv = ICE_IDX_JP_INV, posx = ICE_IDX_POSX, velx = ICE_IDX_VELX,
                         Fe00 = ICE_IDX_FE00, Bp00 = ICE_IDX_BP00, idx_thickness = ICE_IDX_THICKNESS;
    static constexpr float disabled_pts_proportion_threshold = 0.05f;   // parameters_sim.h:39
    SimParams() { Reset(); }
    void Reset() { ice_params_reset(this); }
    void ComputeHelperVariables() { ice_params_derive(this); }
    void ComputeLame() { ice_params_derive(this); }
    int AnimationFrameNumber(int SimulationStep) const { return SimulationStep / UpdateEveryNthStep; }
};

class HostSideSOA {
public:
    std::vector<float> host_buffer;
    unsigned capacity = 0;
    unsigned size = 0;
    int nlines = ICE_N_LINES;

    void Allocate(int pts_capacity, int n_lines = ICE_N_LINES) {
        capacity = (unsigned)pts_capacity;
        nlines = n_lines;
        size = 0;
        host_buffer.assign((size_t)capacity * nlines, 0.f);
    }
    float *getPointerToLine(int idxLine) { return host_buffer.data() + (size_t)capacity * idxLine; }
    uint32_t getValueInt(int line, unsigned i) const {
        uint32_t u;
        std::memcpy(&u, &host_buffer[(size_t)capacity * line + i], 4);
        return u;
    }
    void setValueInt(int line, unsigned i, uint32_t v) { std::memcpy(&host_buffer[(size_t)capacity * line + i], &v, 4); }
};

inline void check(int rc, IceHandle *h, const char *what) {
    if (rc != ICE_OK) throw std::runtime_error(std::string(what) + ": " + ice_last_error(h));
}

class GPU_Implementation5;

struct PartitionFields {
    int PartitionID = 0;
    int Device = 0;
    size_t nPts_partition = 0;
    size_t GridX_partition = 0;
    size_t GridX_offset = 0;
    uint32_t error_code = 0;
    int disabled_points_count = 0;
    size_t nPtsPitch = 0, nGridPitch = 0;
    float timing_10_P2GAndHalo = 0, timing_20_acceptHalo = 0, timing_30_updateGrid = 0, timing_40_G2P = 0,
          timing_60_ptsSent = 0, timing_70_ptsAccepted = 0, timing_stepTotal = 0;
};

struct GPU_Partition : PartitionFields {
    static SimParams *prms;
    IceHandle *handle = nullptr;   // owns the device buffers, stream and events

    GPU_Partition() = default;
    GPU_Partition(const GPU_Partition &) = delete;
    GPU_Partition &operator=(const GPU_Partition &) = delete;
    GPU_Partition(GPU_Partition &&o) noexcept : PartitionFields(o), handle(o.handle) { o.handle = nullptr; }
    GPU_Partition &operator=(GPU_Partition &&o) noexcept {
        if (this != &o) {
            if (handle) ice_destroy(handle);
            PartitionFields::operator=(o);
            handle = o.handle;
            o.handle = nullptr;
        }
        return *this;
    }
    ~GPU_Partition() {
        if (handle) ice_destroy(handle);
    }

    void initialize(int device, int partition) {
        if (handle) throw std::runtime_error("GPU_Partition double initialization");
        PartitionID = partition;
        Device = device;
        check(ice_create(prms, device, &handle), nullptr, "GPU_Partition initialization failure");
        check(ice_enable_timings(handle, 1), handle, "timings");
    }
    void allocate(int n_points_capacity, int /*grid_x_capacity*/) {
        check(ice_allocate(handle, (uint32_t)n_points_capacity), handle, "GPU_Partition allocate");
        IceInfo info;
        check(ice_get_info(handle, &info), handle, "GPU_Partition allocate");
        nPtsPitch = info.point_pitch;
        nGridPitch = info.grid_pitch;
    }
    void transfer_points_from_soa_to_device(HostSideSOA &hssoa, int point_idx_offset) {
        // the host buffer is line-major with pitch = capacity; offset inside each line
        std::vector<float> view((size_t)hssoa.nlines * nPts_partition);
        for (int l = 0; l < hssoa.nlines; l++)
            std::memcpy(&view[(size_t)l * nPts_partition], hssoa.getPointerToLine(l) + point_idx_offset,
                        nPts_partition * sizeof(float));
        check(ice_upload_points(handle, view.data(), nPts_partition, (uint32_t)nPts_partition, hssoa.nlines), handle,
              "transfer_points_from_soa_to_device");
        disabled_points_count = 0;
    }
    void transfer_grid_data_to_device(GPU_Implementation5 *gpu);
    void update_constants() {
        check(ice_update_constants(handle, prms), handle, "cuda_update_constants");
        error_code = 0;
    }
    void update_current_field(const float *current, const float *wind_stress) {
        check(ice_upload_forcing(handle, current, wind_stress), handle, "update_current_field");
    }
    void transfer_from_device(HostSideSOA &hssoa, int point_idx_offset, std::vector<float> &boundary_forces) {
        IceInfo info;
        check(ice_get_info(handle, &info), handle, "transfer_from_device()");
        nPts_partition = info.n_points;
        if (point_idx_offset + nPts_partition > hssoa.capacity)
            throw std::runtime_error("transfer_from_device() HSSOA capacity");
        std::vector<float> buf((size_t)hssoa.nlines * std::max<size_t>(nPts_partition, 1));
        int32_t dis = 0;
        check(ice_download(handle, buf.data(), std::max<size_t>(nPts_partition, 1), hssoa.nlines,
                           boundary_forces.data(), &error_code, &dis),
              handle, "transfer_from_device()");
        for (int l = 0; l < hssoa.nlines; l++)
            std::memcpy(hssoa.getPointerToLine(l) + point_idx_offset, &buf[(size_t)l * std::max<size_t>(nPts_partition, 1)],
                        nPts_partition * sizeof(float));
        disabled_points_count = dis;
    }
    void reset_grid() { check(ice_reset_grid(handle), handle, "cuda_reset_grid error"); }
    void clear_force_accumulator() { check(ice_clear_force_accumulator(handle), handle, "clear_force_accumulator()"); }
    void p2g() { check(ice_p2g(handle), handle, "p2g kernel"); }
    void update_nodes(float simulation_time, float windSpeed = 0, float windAngle = 0) {
        check(ice_update_nodes(handle, simulation_time, windSpeed, windAngle), handle, "update_nodes");
    }
    void g2p(bool recordPQ, bool enablePointTransfer, int applyGlensLaw) {
        check(ice_g2p(handle, recordPQ, enablePointTransfer, applyGlensLaw), handle, "g2p kernel");
    }
    void reset_timings() {
        check(ice_reset_timings(handle), handle, "reset_timings");
        timing_10_P2GAndHalo = timing_20_acceptHalo = timing_30_updateGrid = timing_40_G2P = 0;
        timing_60_ptsSent = timing_70_ptsAccepted = timing_stepTotal = 0;
    }
    void record_timings(bool /*enablePointTransfer*/) {
        double t[8];
        check(ice_get_timings(handle, t), handle, "record_timings");
        timing_10_P2GAndHalo = (float)t[1];
        timing_30_updateGrid = (float)t[2];
        timing_40_G2P = (float)t[3];
        timing_stepTotal = (float)t[4];
    }
    void normalize_timings(int cycles) {
        const float c = 1000.f / (float)cycles;
        timing_10_P2GAndHalo *= c; timing_20_acceptHalo *= c; timing_30_updateGrid *= c; timing_40_G2P *= c;
        timing_60_ptsSent *= c; timing_70_ptsAccepted *= c; timing_stepTotal *= c;
    }
};

inline SimParams *GPU_Partition::prms = nullptr;

class GPU_Implementation5 {
public:
    SimParams *prms = nullptr;
    std::vector<GPU_Partition> partitions;
    HostSideSOA hssoa;
    uint32_t error_code = 0;
    std::vector<uint8_t> grid_status_buffer;
    std::vector<float> grid_boundary_normals, grid_boundary_forces;
    std::vector<float> current_field, wind_stress_field;   // [2][gx*gy] each, empty = none
    std::function<void()> transfer_completion_callback;

    explicit GPU_Implementation5(SimParams *p) : prms(p) { GPU_Partition::prms = p; }

    void allocate_host_arrays_grid() {
        const size_t N = (size_t)prms->GridXTotal * prms->GridYTotal;
        grid_status_buffer.assign(N, ICE_GRID_STATUS_WATER);
        grid_boundary_normals.assign(2 * N, 0.f);
        grid_boundary_forces.assign(2 * N, 0.f);
    }
    void allocate_host_arrays_points(int nPtsInitial, int n_lines = ICE_N_LINES) { hssoa.Allocate(nPtsInitial, n_lines); }
    void initialize(int device = 0) {
        partitions.clear();
        partitions.emplace_back();
        partitions.front().initialize(device, 0);
    }
    void split_hssoa_into_partitions() {
        GPU_Partition &p = partitions.front();
        p.nPts_partition = hssoa.size;
        p.GridX_partition = prms->GridXTotal;
        p.disabled_points_count = 0;
    }
    void allocate_device_arrays() { partitions.front().allocate((int)hssoa.capacity, (int)partitions.front().GridX_partition); }
    void transfer_to_device() {
        GPU_Partition &p = partitions.front();
        p.transfer_points_from_soa_to_device(hssoa, 0);
        p.transfer_grid_data_to_device(this);
    }
    void transfer_wind_and_current_data_to_device() {
        for (auto &p : partitions)
            p.update_current_field(current_field.empty() ? nullptr : current_field.data(),
                                   wind_stress_field.empty() ? nullptr : wind_stress_field.data());
    }
    void transfer_from_device(int elapsed_cycles) {
        unsigned offset = 0;
        for (auto &p : partitions) {
            p.transfer_from_device(hssoa, (int)offset, grid_boundary_forces);
            offset += (unsigned)p.nPts_partition;
            if (p.error_code) error_code = p.error_code;
        }
        hssoa.size = offset;
        // boundary-force normalisation (gpu_implementation5.cpp:219-221)
        const double k = 1.0 / (elapsed_cycles * prms->InitialTimeStep);
        for (float &f : grid_boundary_forces) f = (float)(f * k);
        if (transfer_completion_callback) transfer_completion_callback();
    }
    void synchronize() { for (auto &p : partitions) check(ice_synchronize(p.handle), p.handle, "synchronize"); }
    void update_constants() {
        error_code = 0;
        for (auto &p : partitions) p.update_constants();
    }
    void reset_grid() { for (auto &p : partitions) p.reset_grid(); }
    void reset_timings() { for (auto &p : partitions) p.reset_timings(); }
    void clear_force_accumulator() { for (auto &p : partitions) p.clear_force_accumulator(); }
    void p2g() { partitions.front().p2g(); }
    void update_nodes(float simulation_time, float windSpeed, float windAngle) {
        for (auto &p : partitions) p.update_nodes(simulation_time, windSpeed, windAngle);
    }
    void g2p(bool recordPQ, bool enablePointTransfer, int applyGlensLaw) {
        for (auto &p : partitions) p.g2p(recordPQ, enablePointTransfer, applyGlensLaw);
    }
    void record_timings(bool enablePointTransfer) { for (auto &p : partitions) p.record_timings(enablePointTransfer); }
};

// The per-frame arrays of SnapshotManager (data_manager/snapshotmanager.h:53-64),
// computed on the device in place of SnapshotManager::PrepareFrameArrays
// (snapshotmanager.cpp:623-765); call after transfer_from_device with the same
// elapsed_cycles.
struct FrameArrays {
    std::vector<uint8_t> count, rgb, mass_mask;
    std::vector<float> vis_point_density, vis_mass, vis_vx, vis_vy, vis_r, vis_g, vis_b, vis_Jpinv, vis_P, vis_Q;
    float forces_per_region[ICE_MAX_REGIONS][2] = {};

    void prepare(GPU_Implementation5 &gpu, int elapsed_cycles) {
        GPU_Partition &p = gpu.partitions.front();
        const size_t N = (size_t)gpu.prms->GridXTotal * gpu.prms->GridYTotal;
        count.assign(N, 0);
        rgb.assign(3 * N, 0);
        mass_mask.assign(N, 0);
        std::vector<float> fields((size_t)ICE_FRAME_FIELDS * N);
        IceFrameArrays out{count.data(), fields.data(), rgb.data(), mass_mask.data(), &forces_per_region[0][0]};
        check(ice_prepare_frame_arrays(p.handle, elapsed_cycles, &out), p.handle, "PrepareFrameArrays");
        std::vector<float> *dst[ICE_FRAME_FIELDS] = {&vis_point_density, &vis_mass, &vis_vx, &vis_vy, &vis_r,
                                                      &vis_g, &vis_b, &vis_Jpinv, &vis_P, &vis_Q};
        for (int f = 0; f < ICE_FRAME_FIELDS; f++) dst[f]->assign(fields.begin() + f * N, fields.begin() + (f + 1) * N);
    }
};

inline void GPU_Partition::transfer_grid_data_to_device(GPU_Implementation5 *gpu) {
    check(ice_upload_grid(handle, gpu->grid_status_buffer.data(),
                          gpu->grid_boundary_normals.empty() ? nullptr : gpu->grid_boundary_normals.data()),
          handle, "transfer_grid_data_to_device");
}

}  // namespace icempm
