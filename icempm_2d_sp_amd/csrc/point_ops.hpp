// point_ops.hpp — per-point P2G and G2P work shared by every kernel variant.
#pragma once
#include "device_math.hpp"
#include "../../include/icempm.h"

namespace ice {

// Simulation time and P/Q recording of one substep.  Either given directly
// (phase-by-phase calls) or derived from a device step counter, so a captured
// HIP graph replays correctly at any step.
struct StepArg {
    double t;
    int recordPQ;
    const int *step_ptr;   // when set: step = *step_ptr + step_off
    int step_off;
    // Pipelined path only: 0 lets the substep skip storing what no later
    // substep reads (velocity and Bp, which the next P2G takes from registers);
    // the last substep before control returns to the caller stores everything.
    int complete = 1;
};

__device__ __forceinline__ void resolve_step(const StepArg &s, const DevParams &P, double &t, bool &rec) {
    if (s.step_ptr) {
        const int step = *s.step_ptr + s.step_off;
        t = P.dt_d * (double)step;
        rec = P.update_every > 0 && (step % P.update_every) == 0;
    } else {
        t = s.t;
        rec = s.recordPQ != 0;
    }
}

// Device-side counters of one partition.
struct Counters {
    uint32_t err;        // gpu_error_indicator (kernels.cuh:16)
    int32_t disabled;    // gpu_disabled_points_count (kernels.cuh:29)
    uint32_t n_esc;      // points on the off-tile list
    int32_t step;        // device step counter for graph replays
    uint32_t pad[2];     // learned projection: accepted, fallbacks
    uint32_t n_drift;    // points off their sorted cell, counted on a run chunk's last substep
    uint32_t pad3;
};

// The SoA lines of a point range.  Slot i of a line is addressed as the
// line's (wave-uniform) base plus a 32-bit byte offset, so every access of a
// lane shares one offset register (the library keeps pitch below 2^30).
struct Lines {
    float *base;
    size_t pitch;
    __device__ __forceinline__ float &operator()(int line, uint32_t i) const {
        return *reinterpret_cast<float *>(reinterpret_cast<char *>(base + (size_t)line * pitch) + (uint32_t)(i << 2));
    }
    __device__ __forceinline__ uint32_t u(int line, uint32_t i) const { return __float_as_uint((*this)(line, i)); }
    __device__ __forceinline__ void set_u(int line, uint32_t i, uint32_t v) const { (*this)(line, i) = __uint_as_float(v); }
};

// P2G inputs of one point: mass and the APIC affine momentum terms
// (kernels.cuh:54-76).  Contribution to node offset (a, b):
//   m   += w * mass
//   p   += w * (mass v + S * ((a - x) h, (b - y) h))
struct P2GPoint {
    float x, y;            // cell-local position
    int cx, cy;            // cell (nearest node)
    float mass, mvx, mvy;  // mass, mass * velocity
    float s00, s01, s10, s11;   // subterm2 * h
};

__device__ __forceinline__ void load_p2g_point(const Lines &L, uint32_t i, uint32_t cell, const DevParams &P,
                                               P2GPoint &q) {
    q.cx = (int)(cell & 0xffff);
    q.cy = (int)(cell >> 16);
    q.x = L(ICE_IDX_POSX, i);
    q.y = L(ICE_IDX_POSX + 1, i);
    const float vx = L(ICE_IDX_VELX, i), vy = L(ICE_IDX_VELX + 1, i);
    M2 F{L(ICE_IDX_FE00, i), L(ICE_IDX_FE00 + 1, i), L(ICE_IDX_FE00 + 2, i), L(ICE_IDX_FE00 + 3, i)};
    const float b00 = L(ICE_IDX_BP00, i), b01 = L(ICE_IDX_BP00 + 1, i);
    const float b10 = L(ICE_IDX_BP00 + 2, i), b11 = L(ICE_IDX_BP00 + 3, i);
    const float Jp = L(ICE_IDX_JP_INV, i);
    const float th = L(ICE_IDX_THICKNESS, i);
    q.mass = P.particle_mass * th;
    q.mvx = q.mass * vx;
    q.mvy = q.mass * vy;
    const FInv v = invariants(F);
    float t00, t01, t11;
    kirchhoff(v, P, t00, t01, t11);
    const float k = P.dt_vol_Dpinv * Jp * th;
    q.s00 = (q.mass * b00 - k * t00) * P.h;
    q.s01 = (q.mass * b01 - k * t01) * P.h;
    q.s10 = (q.mass * b10 - k * t01) * P.h;
    q.s11 = (q.mass * b11 - k * t11) * P.h;
}

// Contribution of a point to node offset (a, b) with weights wx[a+1]*wy[b+1].
__device__ __forceinline__ void p2g_contrib(const P2GPoint &q, int a, int b, float w, float &dm, float &dpx,
                                            float &dpy) {
    const float dx = (float)a - q.x, dy = (float)b - q.y;
    dm = w * q.mass;
    dpx = w * (q.mvx + q.s00 * dx + q.s01 * dy);
    dpy = w * (q.mvy + q.s10 * dx + q.s11 * dy);
}

// Everything G2P does after the stencil gather (kernels.cuh:239-314).
struct G2PState {
    uint32_t u;
    int cx, cy;
    float x, y;
    M2 Fe;
    float Jp;
    float p_tr, q_tr;
    bool cell_updated;
    bool newly_disabled;
};

// G2P after the stencil gather up to the return map (kernels.cuh:239-283):
// advect, cell crossing and disable band, F update, PQ, the crushed latch.  vf
// receives the invariants of the trial Fe.
__device__ __forceinline__ void g2p_predict(const DevParams &P, G2PState &s, float pvx, float pvy, const M2 &C,
                                            uint32_t &err, FInv &vf) {
    s.x += pvx * P.dt_h_inv;
    s.y += pvy * P.dt_h_inv;
    if (__builtin_isnan(pvx) || __builtin_isnan(pvy)) err |= ICE_ERR_POINT_VEL_NAN;
    if (__builtin_isnan(s.x) || __builtin_isnan(s.y)) err |= ICE_ERR_POINT_POS_NAN;
    if (isnan4(C)) err |= ICE_ERR_POINT_BP_NAN;
    int dx = 0, dy = 0;
    if (s.x > 0.5f) { s.x -= 1.f; dx = 1; }
    else if (s.x < -0.5f) { s.x += 1.f; dx = -1; }
    if (s.y > 0.5f) { s.y -= 1.f; dy = 1; }
    else if (s.y < -0.5f) { s.y += 1.f; dy = -1; }
    s.cell_updated = (dx | dy) != 0;
    if (s.cell_updated) {
        s.cx += dx;
        s.cy += dy;
        drift_cell(s.u, dx, dy);   // the sorted path's bookkeeping; no effect on the state
    }
    s.newly_disabled = false;
    if (s.cell_updated && (s.cx <= 1 || s.cx >= P.gx - 2 || s.cy <= 1 || s.cy >= P.gy - 2)) {
        s.u |= kDisabled;
        s.newly_disabled = true;
    }
    if (s.x > 0.5f || s.x < -0.5f || s.y > 0.5f || s.y < -0.5f) err |= ICE_ERR_POINT_JUMP_CELLS;
    // Fe = (I + dt C) Fe
    const M2 F = s.Fe;
    const float c00 = P.dt * C.a00, c01 = P.dt * C.a01, c10 = P.dt * C.a10, c11 = P.dt * C.a11;
    s.Fe.a00 = F.a00 + (c00 * F.a00 + c01 * F.a10);
    s.Fe.a01 = F.a01 + (c00 * F.a01 + c01 * F.a11);
    s.Fe.a10 = F.a10 + (c10 * F.a00 + c11 * F.a10);
    s.Fe.a11 = F.a11 + (c10 * F.a01 + c11 * F.a11);
    if (isnan4(s.Fe)) err |= ICE_ERR_POINT_FE_NAN;
    vf = invariants(s.Fe);
    compute_pq(vf, P, s.p_tr, s.q_tr);
    if (!(s.u & kCrushed) && outside_failure_surface(s.p_tr, s.q_tr, P)) s.u |= kCrushed;
}

// The return map of a crushed point (kernels.cuh:284-285); vf is updated to
// the invariants of the final Fe.
__device__ __forceinline__ void g2p_return_map(const DevParams &P, G2PState &s, float thickness, float dpt,
                                               uint32_t &err, FInv &vf) {
    if (return_map(vf, s.p_tr, s.q_tr, thickness, dpt, P, s.Fe, s.Jp) != 2) vf = invariants(s.Fe);
    if (isnan4(s.Fe)) err |= ICE_ERR_POINT_FE_NAN;
}

// Everything G2P does after the stencil gather (kernels.cuh:239-314); vf
// receives the invariants of the final Fe (for the next P2G's stress).
__device__ __forceinline__ void g2p_finish(const DevParams &P, G2PState &s, float pvx, float pvy, const M2 &C,
                                           float thickness, float dpt, uint32_t &err, FInv &vf) {
    g2p_predict(P, s, pvx, pvy, C, err, vf);
    if (s.u & kCrushed) g2p_return_map(P, s, thickness, dpt, err, vf);
}

__device__ __forceinline__ float dp_threshold(const DevParams &P, const Lines &L, uint32_t i) {
    if (P.has_cohesion) return -L(ICE_IDX_COHESION, i) * P.inv_tan_phi;
    return P.DP_threshold_p;
}

// Write back what G2P changed (kernels.cuh:286-314).
__device__ __forceinline__ void g2p_store(const Lines &L, uint32_t i, const G2PState &s, float pvx, float pvy,
                                          const M2 &C, bool rec) {
    L(ICE_IDX_POSX, i) = s.x;
    L(ICE_IDX_POSX + 1, i) = s.y;
    L(ICE_IDX_VELX, i) = pvx;
    L(ICE_IDX_VELX + 1, i) = pvy;
    L(ICE_IDX_FE00, i) = s.Fe.a00;
    L(ICE_IDX_FE00 + 1, i) = s.Fe.a01;
    L(ICE_IDX_FE00 + 2, i) = s.Fe.a10;
    L(ICE_IDX_FE00 + 3, i) = s.Fe.a11;
    L(ICE_IDX_BP00, i) = C.a00;
    L(ICE_IDX_BP00 + 1, i) = C.a01;
    L(ICE_IDX_BP00 + 2, i) = C.a10;
    L(ICE_IDX_BP00 + 3, i) = C.a11;
    L(ICE_IDX_JP_INV, i) = s.Jp;
    L.set_u(ICE_IDX_UTILITY, i, s.u);
    if (s.cell_updated) L.set_u(ICE_IDX_CELL, i, ((uint32_t)s.cy << 16) | (uint32_t)s.cx);
    if (rec) {
        L(ICE_IDX_P, i) = s.p_tr;
        L(ICE_IDX_Q, i) = s.q_tr;
    }
}

// The pipelined path's store: position and Fe every substep; Jp_inv, the
// utility word and the cell word when they changed; velocity and Bp (read by
// no later substep of the pipeline) only when `complete`; P and Q when rec.
__device__ __forceinline__ void g2p_store_pipelined(const Lines &L, uint32_t i, const G2PState &s, uint32_t u0,
                                                   float jp0, float pvx, float pvy, const M2 &C, bool rec,
                                                   bool complete) {
    L(ICE_IDX_POSX, i) = s.x;
    L(ICE_IDX_POSX + 1, i) = s.y;
    L(ICE_IDX_FE00, i) = s.Fe.a00;
    L(ICE_IDX_FE00 + 1, i) = s.Fe.a01;
    L(ICE_IDX_FE00 + 2, i) = s.Fe.a10;
    L(ICE_IDX_FE00 + 3, i) = s.Fe.a11;
    if (__float_as_uint(s.Jp) != __float_as_uint(jp0)) L(ICE_IDX_JP_INV, i) = s.Jp;
    if (s.u != u0) L.set_u(ICE_IDX_UTILITY, i, s.u);
    if (s.cell_updated) L.set_u(ICE_IDX_CELL, i, ((uint32_t)s.cy << 16) | (uint32_t)s.cx);
    if (complete) {
        L(ICE_IDX_VELX, i) = pvx;
        L(ICE_IDX_VELX + 1, i) = pvy;
        L(ICE_IDX_BP00, i) = C.a00;
        L(ICE_IDX_BP00 + 1, i) = C.a01;
        L(ICE_IDX_BP00 + 2, i) = C.a10;
        L(ICE_IDX_BP00 + 3, i) = C.a11;
    }
    if (rec) {
        L(ICE_IDX_P, i) = s.p_tr;
        L(ICE_IDX_Q, i) = s.q_tr;
    }
}

__device__ __forceinline__ void load_g2p_state(const Lines &L, uint32_t i, uint32_t u, G2PState &s) {
    s.u = u;
    const uint32_t cell = L.u(ICE_IDX_CELL, i);
    s.cx = (int)(cell & 0xffff);
    s.cy = (int)(cell >> 16);
    s.x = L(ICE_IDX_POSX, i);
    s.y = L(ICE_IDX_POSX + 1, i);
    s.Fe = M2{L(ICE_IDX_FE00, i), L(ICE_IDX_FE00 + 1, i), L(ICE_IDX_FE00 + 2, i), L(ICE_IDX_FE00 + 3, i)};
    s.Jp = L(ICE_IDX_JP_INV, i);
}

// The 3x3 stencil of cell (cx, cy) lies inside the grid.  The reference
// raises error_code_point_left_area for the complement (kernels.cuh:102-103)
// but still scatters; here such points are flagged and skipped, so no access
// leaves the grid.
__device__ __forceinline__ bool stencil_inside(int cx, int cy, const DevParams &P) {
    return cx >= 1 && cy >= 1 && cx <= P.gx - 2 && cy <= P.gy - 2;
}

// OR an error word into the device counter once per wave.
__device__ __forceinline__ void flush_err(Counters *cnt, uint32_t err) {
    // wave-wide OR through a reduction over lanes
    for (int off = 32; off > 0; off >>= 1) err |= __shfl_xor(err, off, 64);
    if ((threadIdx.x & 63) == 0 && err) atomicOr(&cnt->err, err);
}

}  // namespace ice
