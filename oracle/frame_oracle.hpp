// frame_oracle.hpp — TEST INFRASTRUCTURE ONLY.  Plain C++ restatement of
// SnapshotManager::PrepareFrameArrays (simulation/data_manager/
// snapshotmanager.cpp:623-765): the per-frame splat of the points onto the
// grid for visualisation, its normalisation and masks, the blended colour
// image, and the boundary forces summed per region.  Sums are FP64 here (the
// reference sums into float vectors in point order); outputs are float.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc_frame {

constexpr int MAX_REGIONS = 100;                          // SimParams::MAX_REGIONS, parameters_sim.h:46
constexpr uint8_t WATER[3] = {0x15, 0x1f, 0x2f};         // ColorMap::rgb_water, gui/colormap.h:36
constexpr uint32_t DISABLED = 0x20000u;                   // utility bit 17
constexpr int NF = 10;   // density, mass, vx, vy, r, g, b, Jp_inv - 1, P, Q

// soa: float lines [>= 19][pitch] as the device stores them (utility, cell and
// point index bit patterns); forces [2][gx*gy] before scaling by fscale.
inline void prepare(const float *soa, size_t pitch, int n, int gx, int gy, double particle_mass,
                    const uint8_t *status, const uint32_t *colors, int ncolors, const float *forces, double fscale,
                    uint8_t *count, float *fields, uint8_t *rgb, uint8_t *mask, float *fpr) {
    const size_t N = (size_t)gx * gy;
    auto line = [&](int l, int i) { return soa[(size_t)l * pitch + i]; };
    auto uline = [&](int l, int i) {
        uint32_t u;
        const float f = line(l, i);
        std::memcpy(&u, &f, 4);
        return u;
    };
    std::vector<double> acc((size_t)NF * N, 0.0);
    std::vector<uint8_t> cnt(N, 0);
    for (int i = 0; i < n; i++) {                         // snapshotmanager.cpp:652-704
        if (uline(0, i) & DISABLED) continue;
        const uint32_t cw = uline(1, i);
        const int cx = (int)(cw & 0xffff), cy = (int)(cw >> 16);
        const double px = line(6, i), py = line(7, i);
        const double vx = line(8, i), vy = line(9, i);
        const double jp = line(5, i), P = line(3, i), Q = line(4, i);
        const double pm = particle_mass * line(18, i);
        const uint32_t pt = uline(2, i);
        const uint32_t c = (pt < (uint32_t)ncolors) ? colors[pt] : 0u;
        const double cr = (c >> 16) & 0xff, cg = (c >> 8) & 0xff, cb = c & 0xff;
        double wx[3], wy[3];                              // CalculateWeightCoeffs, :611-620
        wx[0] = 0.5 * (0.5 - px) * (0.5 - px);
        wx[1] = 0.75 - px * px;
        wx[2] = 0.5 * (px + 0.5) * (px + 0.5);
        wy[0] = 0.5 * (0.5 - py) * (0.5 - py);
        wy[1] = 0.75 - py * py;
        wy[2] = 0.5 * (py + 0.5) * (py + 0.5);
        for (int a = -1; a <= 1; a++)
            for (int b = -1; b <= 1; b++) {
                const size_t k = (size_t)(b + cy) + (size_t)(a + cx) * gy;
                const double w = wx[a + 1] * wy[b + 1];
                const double m = w * pm;
                acc[0 * N + k] += w;
                acc[1 * N + k] += m;
                acc[2 * N + k] += m * vx;
                acc[3 * N + k] += m * vy;
                acc[4 * N + k] += w * cr;
                acc[5 * N + k] += w * cg;
                acc[6 * N + k] += w * cb;
                acc[7 * N + k] += w * jp;
                acc[8 * N + k] += w * P;
                acc[9 * N + k] += w * Q;
            }
        cnt[(size_t)cx * gy + cy]++;                      // uint8, wraps as the reference's counter
    }
    for (size_t k = 0; k < N; k++) {                      // :707-745
        double v[NF];
        for (int f = 0; f < NF; f++) v[f] = acc[f * N + k];
        const double density = v[0], mass = v[1];
        if (density > 0) {
            const double coeff = 1.0 / density;
            v[2] /= mass;
            v[3] /= mass;
            for (int f = 4; f < NF; f++) v[f] *= coeff;
            v[7] -= 1.0;
        }
        const bool mk = density > 1e-1;
        if (!mk || status[k] != 100) {
            v[0] = v[1] = v[2] = v[3] = v[7] = v[8] = v[9] = 0.0;
        }
        const double alpha = std::clamp(density, 0.0, 1.0);
        for (int ch = 0; ch < 3; ch++) {
            const uint8_t c8 = (uint8_t)std::clamp(v[4 + ch], 0.0, 255.0);
            if (rgb) rgb[k * 3 + ch] = (uint8_t)(alpha * c8 + (1.0 - alpha) * WATER[ch]);
        }
        if (fields)
            for (int f = 0; f < NF; f++) fields[f * N + k] = (float)v[f];
        if (mask) mask[k] = mk ? 1 : 0;
        if (count) count[k] = cnt[k];
    }
    if (fpr) {                                            // :748-756
        std::vector<double> s(2 * MAX_REGIONS, 0.0);
        for (size_t k = 0; k < N; k++) {
            const int region = status[k];
            const int idx = region >= 100 ? region - 99 : region;
            if (idx >= MAX_REGIONS) continue;             // out of the reference's array
            s[2 * idx] += forces[k] * fscale;
            s[2 * idx + 1] += forces[k + N] * fscale;
        }
        for (int j = 0; j < 2 * MAX_REGIONS; j++) fpr[j] = (float)s[j];
    }
}

}  // namespace orc_frame
