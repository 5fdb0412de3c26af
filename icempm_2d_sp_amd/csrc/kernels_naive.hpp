// kernels_naive.hpp — the reference's launch structure on unsorted points:
// P2G with 27 global FP32 atomics per point, a node-update pass, and a G2P
// pass (simulation/kernels.cuh:39-315).  Counts are kernel arguments (not
// constant symbols) and error bits are OR-ed atomically.  Used for unsorted
// inputs, as the deterministic-order baseline and in tests.
#pragma once
#include "point_ops.hpp"

namespace ice {

__global__ __launch_bounds__(256) void k_fill_zero(float *__restrict__ a, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = 0.f;
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *__restrict__ a, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = v;
}

__global__ void k_step_advance(Counters *cnt, int by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) cnt->step += by;
}

__global__ void k_step_set(Counters *cnt, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) cnt->step = v;
}

// The wind-stress lines of a domain at the substep's time from its two
// frames: w0 + s (w1 - w0), s = clamp((t - t0) / (t1 - t0), 0, 1), the time
// interpolation the reference sketches and leaves unfinished (kernels.cuh:580-622,
// (1 - tb) w0 + tb w1; windandcurrentinterpolator.cpp:66-70).
__global__ __launch_bounds__(256) void k_wind_frames(StepArg S, DevParams P, float *__restrict__ wx,
                                                     float *__restrict__ wy, const float *__restrict__ w0x,
                                                     const float *__restrict__ w0y, const float *__restrict__ dwx,
                                                     const float *__restrict__ dwy, size_t n, double t0,
                                                     double inv_span) {
    double t;
    bool rec;
    resolve_step(S, P, t, rec);
    const float s = (float)fmin(fmax((t - t0) * inv_span, 0.0), 1.0);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        wx[i] = w0x[i] + s * dwx[i];
        wy[i] = w0y[i] + s * dwy[i];
    }
}

// Fe -> Fe - I on the first n points of an uploaded domain (the device keeps
// e = Fe - I in the Fe lines; download adds I back on the host).
__global__ __launch_bounds__(256) void k_fe_minus_identity(float *__restrict__ pts, size_t pitch, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        pts[(size_t)ICE_IDX_FE00 * pitch + i] -= 1.f;
        pts[(size_t)(ICE_IDX_FE00 + 3) * pitch + i] -= 1.f;
    }
}

// partition_kernel_p2g (kernels.cuh:39-104)
__global__ __launch_bounds__(256) void k_p2g_naive(DevParams P, Lines L, uint32_t n, float *__restrict__ gm,
                                                   float *__restrict__ gpx, float *__restrict__ gpy,
                                                   Counters *cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t err = 0;
    if (i < n) {
        const uint32_t u = L.u(ICE_IDX_UTILITY, i);
        if (!(u & kDisabled)) {
            P2GPoint q;
            load_p2g_point(L, i, L.u(ICE_IDX_CELL, i), P, q);
            if (!stencil_inside(q.cx, q.cy, P)) {
                err |= ICE_ERR_POINT_LEFT_AREA;
            } else {
                float wx[3], wy[3];
                weights(q.x, wx);
                weights(q.y, wy);
                for (int a = -1; a <= 1; a++)
                    for (int b = -1; b <= 1; b++) {
                        float dm, dpx, dpy;
                        p2g_contrib(q, a, b, wx[a + 1] * wy[b + 1], dm, dpx, dpy);
                        const size_t node = (size_t)(q.cy + b) + (size_t)(q.cx + a) * P.gy;
                        unsafeAtomicAdd(&gm[node], dm);
                        unsafeAtomicAdd(&gpx[node], dpx);
                        unsafeAtomicAdd(&gpy[node], dpy);
                        if (__builtin_isnan(dpx) || __builtin_isnan(dpy)) err |= ICE_ERR_GRID_P2G_NAN_VEL;
                        if (__builtin_isnan(dm)) err |= ICE_ERR_GRID_P2G_NAN_MASS;
                    }
            }
        }
    }
    flush_err(cnt, err);
}

// partition_kernel_update_nodes (kernels.cuh:109-177): velocity written over
// (px, py), wall momentum added to (fx, fy).
__global__ __launch_bounds__(256) void k_update_naive(DevParams P, StepArg S, float *__restrict__ gm,
                                                      float *__restrict__ gpx, float *__restrict__ gpy,
                                                      const uint8_t *__restrict__ status,
                                                      const float *__restrict__ cur_x, const float *__restrict__ cur_y,
                                                      const float *__restrict__ wind_x,
                                                      const float *__restrict__ wind_y, float *__restrict__ gfx,
                                                      float *__restrict__ gfy, Counters *cnt) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t err = 0;
    const size_t N = (size_t)P.gx * P.gy;
    if (idx < N) {
        const float m = gm[idx];
        if (m != 0.f) {
            double t;
            bool rec;
            resolve_step(S, P, t, rec);
            const float ramp = current_ramp(t);
            const float mx = gpx[idx], my = gpy[idx];
            float cx = 0.f, cy = 0.f;
            if (P.has_current) { cx = cur_x[idx]; cy = cur_y[idx]; }
            float vx, vy;
            bool wall;
            node_velocity(m, mx, my, status[idx], cx, cy, wind_x[idx], wind_y[idx], ramp, P, vx, vy, wall);
            if (wall) { gfx[idx] += mx; gfy[idx] += my; }
            gpx[idx] = vx;
            gpy[idx] = vy;
            if (__builtin_isnan(vx) || __builtin_isnan(vy)) err |= ICE_ERR_GRID_NAN;
        }
    }
    flush_err(cnt, err);
}

// partition_kernel_g2p (kernels.cuh:181-315), node velocities read from (px, py).
__global__ __launch_bounds__(256) void k_g2p_naive(DevParams P, StepArg S, Lines L, uint32_t n,
                                                   const float *__restrict__ gvx, const float *__restrict__ gvy,
                                                   Counters *cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t err = 0;
    if (i < n) {
        const uint32_t u = L.u(ICE_IDX_UTILITY, i);
        G2PState s;
        load_g2p_state(L, i, u, s);
        if (!(u & kDisabled) && !stencil_inside(s.cx, s.cy, P)) err |= ICE_ERR_POINT_LEFT_AREA;
        else if (!(u & kDisabled)) {
            double t;
            bool rec;
            resolve_step(S, P, t, rec);
            const float th = L(ICE_IDX_THICKNESS, i);
            float wx[3], wy[3];
            weights(s.x, wx);
            weights(s.y, wy);
            float2 nv[3][3];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    const size_t node = (size_t)(s.cy + b - 1) + (size_t)(s.cx + a - 1) * P.gy;
                    nv[a][b] = make_float2(gvx[node], gvy[node]);
                }
            float pvx, pvy;
            M2 C;
            g2p_gather(nv, s.x, s.y, wx, wy, P.four_h_inv, pvx, pvy, C);
            FInv vf;
            g2p_finish(P, s, pvx, pvy, C, th, dp_threshold(P, L, i), err, vf);
            if (s.newly_disabled) atomicAdd(&cnt->disabled, 1);
            g2p_store(L, i, s, pvx, pvy, C, rec);
        }
    }
    flush_err(cnt, err);
}

}  // namespace ice
