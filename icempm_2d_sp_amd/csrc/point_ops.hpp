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
    __device__ __forceinline__ float get(int line, uint32_t i) const { return (*this)(line, i); }
    __device__ __forceinline__ void set(int line, uint32_t i, float v) const { (*this)(line, i) = v; }
};

// The same lines through one buffer resource over a domain's point range:
// slot i of line l is byte i * 4 + l * pitch * 4, the line's part a scalar
// offset (SALU), so the lines need no 64-bit base each.  Valid while
// 20 * pitch * 4 < 2^32 (the host picks this form only then).
struct BufLines {
    __amdgpu_buffer_rsrc_t r;
    uint32_t pitch4;
    __device__ __forceinline__ static BufLines make(float *base, size_t pitch) {
        BufLines b;
        // word 3 of the gfx9 buffer descriptor for raw 32-bit accesses
        b.r = __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)0xffffffff, 0x00020000);
        b.pitch4 = (uint32_t)(pitch * 4);
        return b;
    }
    __device__ __forceinline__ uint32_t u(int line, uint32_t i) const {
        return __builtin_amdgcn_raw_buffer_load_b32(r, (int)(i << 2), (int)((uint32_t)line * pitch4), 0);
    }
    __device__ __forceinline__ void set_u(int line, uint32_t i, uint32_t v) const {
        __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)(i << 2), (int)((uint32_t)line * pitch4), 0);
    }
    __device__ __forceinline__ float get(int line, uint32_t i) const { return __uint_as_float(u(line, i)); }
    __device__ __forceinline__ void set(int line, uint32_t i, float v) const { set_u(line, i, __float_as_uint(v)); }
};

// Whether a point buffer of `pitch` slots per line can use BufLines.
__host__ __device__ constexpr bool buf_lines_ok(size_t pitch) {
    return (uint64_t)pitch * 4u * ICE_N_LINES_MAX <= 0xffffffffull;
}

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

template <class LT>
__device__ __forceinline__ void load_p2g_point(const LT &L, uint32_t i, uint32_t cell, const DevParams &P,
                                               P2GPoint &q) {
    q.cx = (int)(cell & 0xffff);
    q.cy = (int)(cell >> 16);
    q.x = L.get(ICE_IDX_POSX, i);
    q.y = L.get(ICE_IDX_POSX + 1, i);
    const float vx = L.get(ICE_IDX_VELX, i), vy = L.get(ICE_IDX_VELX + 1, i);
    M2 F{L.get(ICE_IDX_FE00, i), L.get(ICE_IDX_FE00 + 1, i), L.get(ICE_IDX_FE00 + 2, i), L.get(ICE_IDX_FE00 + 3, i)};
    const float b00 = L.get(ICE_IDX_BP00, i), b01 = L.get(ICE_IDX_BP00 + 1, i);
    const float b10 = L.get(ICE_IDX_BP00 + 2, i), b11 = L.get(ICE_IDX_BP00 + 3, i);
    const float Jp = L.get(ICE_IDX_JP_INV, i);
    const float th = L.get(ICE_IDX_THICKNESS, i);
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

// A point's whole 3 x 3 P2G stencil added to register sums, node offset
// (a - 1, b - 1) in [a][b].  The affine momentum w (m v + S ((a-1) - x,
// (b-1) - y)) is evaluated as w (A + S (a-1, b-1)) with A = m v - S (x, y),
// and the mass term as (wx[a] m) wy[b], so each node costs three FMAs.
__device__ __forceinline__ void p2g_accumulate(const P2GPoint &t, const float wx[3], const float wy[3],
                                               float am[3][3], float ax[3][3], float ay[3][3]) {
    const float Ax = t.mvx - (t.s00 * t.x + t.s01 * t.y);
    const float Ay = t.mvy - (t.s10 * t.x + t.s11 * t.y);
    const float cx[3] = {Ax - t.s00, Ax, Ax + t.s00};
    const float cy[3] = {Ay - t.s10, Ay, Ay + t.s10};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float wm = wx[a] * t.mass;
        const float vx[3] = {cx[a] - t.s01, cx[a], cx[a] + t.s01};
        const float vy[3] = {cy[a] - t.s11, cy[a], cy[a] + t.s11};
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const float w = wx[a] * wy[b];
            am[a][b] += wm * wy[b];
            ax[a][b] += w * vx[b];
            ay[a][b] += w * vy[b];
        }
    }
}

// The G2P gather of one point (kernels.cuh:224-236) from the velocities of its
// 3 x 3 nodes, v[a][b] at offset (a - 1, b - 1), factored by the separable
// weights w = wx[a] wy[b]:
//   v_p = sum_a wx[a] sum_b wy[b] v[a][b]
//   C   = (4/h) sum_ab w v ((a-1) - x, (b-1) - y)^T
//       = (4/h) (sum_ab w v (a-1) - x v_p, sum_ab w v (b-1) - y v_p)
// (sum_a wx[a] (a-1) = x for the quadratic B-spline weights).
__device__ __forceinline__ void g2p_gather(const float2 v[3][3], float x, float y, const float wx[3],
                                           const float wy[3], float four_h_inv, float &pvx, float &pvy, M2 &C) {
    float sx[3], sy[3], tx[3], ty[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        sx[a] = wy[0] * v[a][0].x + wy[1] * v[a][1].x + wy[2] * v[a][2].x;
        sy[a] = wy[0] * v[a][0].y + wy[1] * v[a][1].y + wy[2] * v[a][2].y;
        tx[a] = wy[2] * v[a][2].x - wy[0] * v[a][0].x;
        ty[a] = wy[2] * v[a][2].y - wy[0] * v[a][0].y;
    }
    pvx = wx[0] * sx[0] + wx[1] * sx[1] + wx[2] * sx[2];
    pvy = wx[0] * sy[0] + wx[1] * sy[1] + wx[2] * sy[2];
    const float ax = wx[2] * sx[2] - wx[0] * sx[0], ay = wx[2] * sy[2] - wx[0] * sy[0];
    const float bx = wx[0] * tx[0] + wx[1] * tx[1] + wx[2] * tx[2];
    const float by = wx[0] * ty[0] + wx[1] * ty[1] + wx[2] * ty[2];
    C.a00 = four_h_inv * (ax - x * pvx);
    C.a01 = four_h_inv * (bx - y * pvx);
    C.a10 = four_h_inv * (ay - x * pvy);
    C.a11 = four_h_inv * (by - y * pvy);
}

// Everything G2P does after the stencil gather (kernels.cuh:239-314).
struct G2PState {
    uint32_t u;
    int cx, cy;
    float x, y;
    M2 Fe;          // Fe - I, as the device stores it
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
    // the disable band, c <= 1 or c >= g - 2, as one unsigned compare per axis
    s.newly_disabled = s.cell_updated && ((uint32_t)(s.cx - 2) > (uint32_t)(P.gx - 5) ||
                                          (uint32_t)(s.cy - 2) > (uint32_t)(P.gy - 5));
    if (s.newly_disabled) s.u |= kDisabled;
    if (fmaxf(fabsf(s.x), fabsf(s.y)) > 0.5f) err |= ICE_ERR_POINT_JUMP_CELLS;
    // Fe = (I + dt C) Fe on e = Fe - I: e' = e + dt C + dt C e
    const M2 e = s.Fe;
    const float c00 = P.dt * C.a00, c01 = P.dt * C.a01, c10 = P.dt * C.a10, c11 = P.dt * C.a11;
    s.Fe.a00 = e.a00 + (c00 + (c00 * e.a00 + c01 * e.a10));
    s.Fe.a01 = e.a01 + (c01 + (c00 * e.a01 + c01 * e.a11));
    s.Fe.a10 = e.a10 + (c10 + (c10 * e.a00 + c11 * e.a10));
    s.Fe.a11 = e.a11 + (c11 + (c10 * e.a01 + c11 * e.a11));
    // the NaN flags of velocity, position, Bp and Fe (kernels.cuh:242-244,
    // 273): one test of their sum on the common path (a NaN anywhere makes it
    // NaN, and the cell shift above leaves a NaN position NaN; an inf - inf
    // only sends the point to the exact tests), the flags themselves only
    // where it fails
    if (__builtin_isnan(((pvx + pvy) + (s.x + s.y)) + ((C.a00 + C.a01) + (C.a10 + C.a11)) +
                        ((s.Fe.a00 + s.Fe.a01) + (s.Fe.a10 + s.Fe.a11)))) {
        if (__builtin_isnan(pvx) || __builtin_isnan(pvy)) err |= ICE_ERR_POINT_VEL_NAN;
        if (__builtin_isnan(s.x) || __builtin_isnan(s.y)) err |= ICE_ERR_POINT_POS_NAN;
        if (isnan4(C)) err |= ICE_ERR_POINT_BP_NAN;
        if (isnan4(s.Fe)) err |= ICE_ERR_POINT_FE_NAN;
    }
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

template <class LT>
__device__ __forceinline__ float dp_threshold(const DevParams &P, const LT &L, uint32_t i) {
    if (P.has_cohesion) return -L.get(ICE_IDX_COHESION, i) * P.inv_tan_phi;
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
template <class LT>
__device__ __forceinline__ void g2p_store_pipelined(const LT &L, uint32_t i, const G2PState &s, uint32_t u0,
                                                   float jp0, float pvx, float pvy, const M2 &C, bool rec,
                                                   bool complete) {
    L.set(ICE_IDX_POSX, i, s.x);
    L.set(ICE_IDX_POSX + 1, i, s.y);
    L.set(ICE_IDX_FE00, i, s.Fe.a00);
    L.set(ICE_IDX_FE00 + 1, i, s.Fe.a01);
    L.set(ICE_IDX_FE00 + 2, i, s.Fe.a10);
    L.set(ICE_IDX_FE00 + 3, i, s.Fe.a11);
    if (__float_as_uint(s.Jp) != __float_as_uint(jp0)) L.set(ICE_IDX_JP_INV, i, s.Jp);
    if (s.u != u0) L.set_u(ICE_IDX_UTILITY, i, s.u);
    if (s.cell_updated) L.set_u(ICE_IDX_CELL, i, ((uint32_t)s.cy << 16) | (uint32_t)s.cx);
    if (complete) {
        L.set(ICE_IDX_VELX, i, pvx);
        L.set(ICE_IDX_VELX + 1, i, pvy);
        L.set(ICE_IDX_BP00, i, C.a00);
        L.set(ICE_IDX_BP00 + 1, i, C.a01);
        L.set(ICE_IDX_BP00 + 2, i, C.a10);
        L.set(ICE_IDX_BP00 + 3, i, C.a11);
    }
    if (rec) {
        L.set(ICE_IDX_P, i, s.p_tr);
        L.set(ICE_IDX_Q, i, s.q_tr);
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
    // 1 <= c <= g - 2 as one unsigned compare per axis (g >= 5 for every grid)
    return (uint32_t)(cx - 1) <= (uint32_t)(P.gx - 3) && (uint32_t)(cy - 1) <= (uint32_t)(P.gy - 3);
}

// OR an error word into the device counter once per wave.
__device__ __forceinline__ void flush_err(Counters *cnt, uint32_t err) {
    // wave-wide OR through a reduction over lanes
    for (int off = 32; off > 0; off >>= 1) err |= __shfl_xor(err, off, 64);
    if ((threadIdx.x & 63) == 0 && err) atomicOr(&cnt->err, err);
}

}  // namespace ice
