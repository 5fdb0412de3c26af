// kernels_frame.hpp — SnapshotManager::PrepareFrameArrays
// (simulation/data_manager/snapshotmanager.cpp:623-765) on the device: the
// per-frame splat of the points onto the grid for visualisation (point
// density, mass, mass-weighted velocity, colour, Jp_inv, P, Q), its
// normalisation and masks, the colour image blended with the water colour,
// the points per cell, and the boundary forces summed per region.
//
// The reference splats point by point on the host.  Here the points are
// sorted by cell first, and one lane per node gathers the points of the (up
// to) 3 x 3 cells whose stencils hold the node, in a fixed order: no atomics,
// the same result on every run.
#pragma once
#include "tile_common.hpp"

namespace ice {

constexpr int kMaxRegions = 100;                  // SimParams::MAX_REGIONS, parameters_sim.h:46
constexpr int kFrameFields = 10;                  // density, mass, vx, vy, r, g, b, Jp_inv - 1, P, Q
constexpr float kWater[3] = {21.f, 31.f, 47.f};   // ColorMap::rgb_water {0x15, 0x1f, 0x2f}, gui/colormap.h:36

// Output arrays of one frame (device memory, gx * gy nodes each).
struct FrameOut {
    uint8_t *count;   // points per cell, modulo 256 (the reference's counter is uint8)
    float *fields;    // [kFrameFields][N]
    uint8_t *rgb;     // [N][3]
    uint8_t *mask;    // density > 0.1
};

__global__ __launch_bounds__(256) void k_frame_fields(DomainDesc dd, Lines L, const uint32_t *__restrict__ cs,
                                                      const uint8_t *__restrict__ status,
                                                      const uint32_t *__restrict__ colors, uint32_t ncolors,
                                                      float particle_mass, FrameOut o) {
    const int gx = dd.gx, gy = dd.gy;
    const size_t N = (size_t)gx * gy;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < N; k += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(k / gy), y = (int)(k - (size_t)x * gy);
        float v[kFrameFields] = {};
        uint32_t cnt = 0;
        // the cell (x + 1 - a, y + 1 - b) holds the node at stencil offset (a - 1, b - 1)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int cx = x + 1 - a;
            if (cx < 0 || cx >= gx) continue;
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const int cy = y + 1 - b;
                if (cy < 0 || cy >= gy) continue;
                const size_t c = (size_t)cx * gy + cy;
                const uint32_t s0 = cs[c], s1 = cs[c + 1];
                for (uint32_t i = s0; i < s1; i++) {
                    if (L.u(ICE_IDX_UTILITY, i) & kDisabled) continue;
                    if (a == 1 && b == 1) cnt++;
                    float wx[3], wy[3];
                    weights(L(ICE_IDX_POSX, i), wx);
                    weights(L(ICE_IDX_POSX + 1, i), wy);
                    const float w = wx[a] * wy[b];
                    const float m = w * particle_mass * L(ICE_IDX_THICKNESS, i);
                    const uint32_t pt = L.u(ICE_IDX_POINT, i);
                    const uint32_t col = pt < ncolors ? colors[pt] : 0u;
                    v[0] += w;
                    v[1] += m;
                    v[2] += m * L(ICE_IDX_VELX, i);
                    v[3] += m * L(ICE_IDX_VELX + 1, i);
                    v[4] += w * (float)((col >> 16) & 0xffu);
                    v[5] += w * (float)((col >> 8) & 0xffu);
                    v[6] += w * (float)(col & 0xffu);
                    v[7] += w * L(ICE_IDX_JP_INV, i);
                    v[8] += w * L(ICE_IDX_P, i);
                    v[9] += w * L(ICE_IDX_Q, i);
                }
            }
        }
        // normalisation and masks (snapshotmanager.cpp:707-745)
        const float density = v[0], mass = v[1];
        if (density > 0.f) {
            const float coeff = 1.f / density;
            v[2] /= mass;
            v[3] /= mass;
#pragma unroll
            for (int f = 4; f < kFrameFields; f++) v[f] *= coeff;
            v[7] -= 1.f;
        }
        const bool mk = density > 0.1f;
        if (!mk || status[k] != ICE_GRID_STATUS_WATER) v[0] = v[1] = v[2] = v[3] = v[7] = v[8] = v[9] = 0.f;
        const float alpha = fminf(fmaxf(density, 0.f), 1.f);
        if (o.rgb)
#pragma unroll
            for (int ch = 0; ch < 3; ch++) {
                const float c8 = (float)(uint32_t)fminf(fmaxf(v[4 + ch], 0.f), 255.f);
                o.rgb[k * 3 + ch] = (uint8_t)(uint32_t)(alpha * c8 + (1.f - alpha) * kWater[ch]);
            }
        if (o.fields)
#pragma unroll
            for (int f = 0; f < kFrameFields; f++) o.fields[(size_t)f * N + k] = v[f];
        if (o.mask) o.mask[k] = mk ? 1 : 0;
        if (o.count) o.count[k] = (uint8_t)cnt;
    }
}

// Boundary forces summed per region (snapshotmanager.cpp:748-756): region
// index status >= 100 ? status - 99 : status; statuses past the reference's
// 100 regions are skipped.  Per-block FP64 partial sums (only nodes with a
// nonzero force add), part [gridDim.x][kMaxRegions][2].
__global__ __launch_bounds__(256) void k_region_forces(DomainDesc dd, const float *__restrict__ fx,
                                                       const float *__restrict__ fy,
                                                       const uint8_t *__restrict__ status, double scale,
                                                       double *__restrict__ part) {
    __shared__ double s[2 * kMaxRegions];
    for (int j = threadIdx.x; j < 2 * kMaxRegions; j += blockDim.x) s[j] = 0.0;
    __syncthreads();
    const size_t N = (size_t)dd.gx * dd.gy;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < N; k += (size_t)gridDim.x * blockDim.x) {
        const float a = fx[k], b = fy[k];
        if (a == 0.f && b == 0.f) continue;
        const int region = status[k];
        const int idx = region >= 100 ? region - 99 : region;
        if (idx >= kMaxRegions) continue;
        atomicAdd(&s[2 * idx], (double)a * scale);
        atomicAdd(&s[2 * idx + 1], (double)b * scale);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < 2 * kMaxRegions; j += blockDim.x) part[(size_t)blockIdx.x * 2 * kMaxRegions + j] = s[j];
}

// The partial sums of nblocks blocks, in block order, as float.
__global__ __launch_bounds__(256) void k_region_sum(const double *__restrict__ part, int nblocks,
                                                    float *__restrict__ out) {
    const int j = threadIdx.x;
    if (j >= 2 * kMaxRegions) return;
    double t = 0.0;
    for (int b = 0; b < nblocks; b++) t += part[(size_t)b * 2 * kMaxRegions + j];
    out[j] = (float)t;
}

}  // namespace ice
