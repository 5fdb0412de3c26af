// kernels_fused.hpp — the pipelined sorted path: one pass over each tile's
// points per substep runs the grid update and G2P of substep n and the P2G of
// substep n + 1 (partition_kernel_update_nodes / _g2p / _p2g,
// simulation/kernels.cuh:109-177, 181-315, 39-104), so a point's state is read
// once and written once per substep.
//
// Node sums live in frames: tile T (TW x TH nodes) owns a frame of the
// (TW + 4) x (TH + 4) nodes T +- 2, into which the points sorted into T scatter
// (they sit at most one cell from T, so their stencils stay inside it).  A
// node's sum is the sum of every frame that covers it (at most 2 x 2), taken
// in ascending tile order, so any reader sees the same value.  Frames come in
// two buffers: substep n reads buffer n % 2 and writes buffer (n + 1) % 2.
//
// Inside a tile, a workgroup walks T's cells in rounds of 256 (one lane per
// cell).  Lanes take the round's points one per lane in sorted order
// (coalesced SoA loads and stores), run G2P, and stage the point's P2G terms
// in LDS; then each cell lane sums the terms of its own points into a
// register stencil, writes it back to LDS, and each node the round touches
// gathers the stencils of its (up to) nine cells in a fixed order: plain LDS
// adds, four barriers per batch of staged points.  A point that changed cell
// since the sort (one cell at most) is still summed by the lane of its sorted
// cell: the part of its stencil inside that lane's 3 x 3 goes into the
// registers with shifted weights, the rest (three nodes for a move along an
// axis, five for a diagonal one) into the LDS frame with atomics.  Points that
// leave the 3 x 3 around their sorted cell (the drift kept in the utility
// word) leave the tile pass (kEscaped + kFresh) and the off-tile kernel does
// their P2G, with global atomics into the frame of the tile that now holds
// them.
#pragma once
#include "tile_common.hpp"
#include "mlp.hpp"

namespace ice {

template <int TW, int TH>
struct Frame {
    static constexpr int W = TW + 4, H = TH + 4, N = W * H;   // frame nodes (global layout: lx * H + ly)
    static constexpr int LP = H + 1;                          // LDS column pitch
    static constexpr int NN = W * LP;                         // LDS floats per component
};

// Offset (floats) of frame (tx, ty) of a domain.  IdxT = uint32_t when every
// frame offset fits 32 bits (the host checks the frame buffer size), which
// keeps the index arithmetic off the 64-bit path.
template <int TW, int TH, class IdxT = size_t>
__device__ __forceinline__ IdxT frame_offset(const DomainDesc &dd, int tx, int ty) {
    return (IdxT)(dd.frame_base + (uint64_t)tx * dd.nty + ty) * (IdxT)(3 * Frame<TW, TH>::N);
}

// F[o]; with 32-bit offsets as a 32-bit byte offset from F (frames below 2^30
// floats), which the loads take as a register offset to a scalar base.
__device__ __forceinline__ float frame_ld(const float *__restrict__ F, size_t o) { return F[o]; }
__device__ __forceinline__ float frame_ld(const float *__restrict__ F, uint32_t o) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(F) + (o << 2));
}

// Sum of node (x, y) of a domain over every frame that covers it, ascending tx
// then ty.  (x, y) must lie inside the domain's grid.  Branch-free: the four
// candidate frames are always read (a covered-or-not flag selects the value,
// uncovered candidates read the frame of the tile holding the node), so the
// loads of several nodes can be in flight together.
template <int TW, int TH, class IdxT = size_t>
__device__ __forceinline__ void frame_node_sum(const float *__restrict__ F, const DomainDesc &dd, int x, int y,
                                               float &m, float &px, float &py) {
    using Fr = Frame<TW, TH>;
    m = px = py = 0.f;
    const int ta = (x + 2) / TW, ua = (y + 2) / TH;
    const IdxT home = frame_offset<TW, TH, IdxT>(dd, x / TW, y / TH) + (x % TW + 2) * Fr::H + (y % TH + 2);
#pragma unroll
    for (int dx = 1; dx >= 0; dx--) {
        const int tx = ta - dx;
        const int lx = x - tx * TW + 2;
        const bool okx = tx >= 0 && tx < dd.ntx && lx < Fr::W;
#pragma unroll
        for (int dy = 1; dy >= 0; dy--) {
            const int ty = ua - dy;
            const int ly = y - ty * TH + 2;
            const bool ok = okx && ty >= 0 && ty < dd.nty && ly < Fr::H;
            const IdxT o = ok ? frame_offset<TW, TH, IdxT>(dd, tx, ty) + lx * Fr::H + ly : home;
            const float a = frame_ld(F, o), b = frame_ld(F, (IdxT)(o + Fr::N)), c = frame_ld(F, (IdxT)(o + 2 * Fr::N));
            m += ok ? a : 0.f;
            px += ok ? b : 0.f;
            py += ok ? c : 0.f;
        }
    }
}

// Node velocity of node (x, y) from frame sums (kernels.cuh:109-177 plus wind).
template <int TW, int TH>
__device__ __forceinline__ void frame_node_velocity(const float *__restrict__ F, const DomainDesc &dd,
                                                    const GridLines &G, const DevParams &P, float ramp, int x, int y,
                                                    float &vx, float &vy, float &mx, float &my, bool &wall,
                                                    float &m) {
    frame_node_sum<TW, TH>(F, dd, x, y, m, mx, my);
    const size_t node = (size_t)x * P.gy + y;
    float cx = 0.f, cy = 0.f;
    if (P.has_current) { cx = G.cur_x[node]; cy = G.cur_y[node]; }
    node_velocity(m, mx, my, G.status[node], cx, cy, G.wind_x[node], G.wind_y[node], ramp, P, vx, vy, wall);
}

// P2G terms of a point from its state in registers (kernels.cuh:54-76); v
// holds the invariants of the point's Fe.
__device__ __forceinline__ void p2g_terms(const DevParams &P, float x, float y, int cx, int cy, float vx, float vy,
                                          const FInv &v, const M2 &Bp, float Jp, float th, P2GPoint &q) {
    q.x = x;
    q.y = y;
    q.cx = cx;
    q.cy = cy;
    q.mass = P.particle_mass * th;
    q.mvx = q.mass * vx;
    q.mvy = q.mass * vy;
    float t00, t01, t11;
    kirchhoff(v, P, t00, t01, t11);
    const float k = P.dt_vol_Dpinv * Jp * th;
    q.s00 = (q.mass * Bp.a00 - k * t00) * P.h;
    q.s01 = (q.mass * Bp.a01 - k * t01) * P.h;
    q.s10 = (q.mass * Bp.a10 - k * t01) * P.h;
    q.s11 = (q.mass * Bp.a11 - k * t11) * P.h;
}

// A staged point whose cell (cx + sx, cy + sy) differs from the sorted cell
// (cx, cy) of the lane that sums it: the nodes of its stencil outside the
// lane's 3 x 3 go into the LDS frame (lane's cell at frame column lx, row ly)
// with atomics; then t, wx, wy are re-expressed from the lane's cell (weights
// of the lane's 3 x 3 nodes, zero where the stencil does not reach them).
// Returns false, flagging it, for a point further than one cell away.
template <int TW, int TH>
__device__ __forceinline__ bool moved_stencil(P2GPoint &t, int sx, int sy, int lx, int ly, float wx[3], float wy[3],
                                              float *acc, uint32_t &err) {
    using Fr = Frame<TW, TH>;
    if (sx < -1 || sx > 1 || sy < -1 || sy > 1) {
        err |= ICE_ERR_POINT_LEFT_AREA;   // cannot happen while the drift invariant holds
        return false;
    }
    // the stencil's column at x offset 2 sx and its row at y offset 2 sy
    // (new-cell stencil indices ra, rb) lie outside the lane's 3 x 3
    const int ra = sx > 0 ? 2 : 0, rb = sy > 0 ? 2 : 0;
    for (int k = 0; k < 6; k++) {
        const bool col = k < 3;
        const int a = col ? ra : k - 3, b = col ? k : rb;
        if (col ? sx == 0 : (sy == 0 || (sx != 0 && a == ra))) continue;   // inside, or the corner again
        const float w = (col ? (ra ? wx[2] : wx[0]) : (a == 0 ? wx[0] : a == 1 ? wx[1] : wx[2])) *
                        (col ? (b == 0 ? wy[0] : b == 1 ? wy[1] : wy[2]) : (rb ? wy[2] : wy[0]));
        float dm, dpx, dpy;
        p2g_contrib(t, a - 1, b - 1, w, dm, dpx, dpy);
        if (__builtin_isnan(dpx) || __builtin_isnan(dpy)) err |= ICE_ERR_GRID_P2G_NAN_VEL;
        if (__builtin_isnan(dm)) err |= ICE_ERR_GRID_P2G_NAN_MASS;
        const int o = (lx + a - 1 + sx) * Fr::LP + (ly + b - 1 + sy);
        __hip_atomic_fetch_add(&acc[o], dm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&acc[Fr::NN + o], dpx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&acc[2 * Fr::NN + o], dpy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    float sw[3], sh[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        sw[a] = sx == 0 ? wx[a] : (sx > 0 ? (a > 0 ? wx[a - 1] : 0.f) : (a < 2 ? wx[a + 1] : 0.f));
        sh[a] = sy == 0 ? wy[a] : (sy > 0 ? (a > 0 ? wy[a - 1] : 0.f) : (a < 2 ? wy[a + 1] : 0.f));
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        wx[a] = sw[a];
        wy[a] = sh[a];
    }
    t.x += (float)sx;
    t.y += (float)sy;
    return true;
}

// A point's whole stencil into the global frame of the tile holding its cell.
template <int TW, int TH>
__device__ __forceinline__ void frame_scatter_global(const P2GPoint &q, const DomainDesc &dd, float *F,
                                                     uint32_t &err) {
    using Fr = Frame<TW, TH>;
    const int tx = q.cx / TW, ty = q.cy / TH;
    float *f = F + frame_offset<TW, TH>(dd, tx, ty);
    const int lx0 = q.cx - tx * TW + 1, ly0 = q.cy - ty * TH + 1;   // node cx - 1 in frame coordinates
    float wx[3], wy[3];
    weights(q.x, wx);
    weights(q.y, wy);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            float dm, dpx, dpy;
            p2g_contrib(q, a - 1, b - 1, wx[a] * wy[b], dm, dpx, dpy);
            if (__builtin_isnan(dpx) || __builtin_isnan(dpy)) err |= ICE_ERR_GRID_P2G_NAN_VEL;
            if (__builtin_isnan(dm)) err |= ICE_ERR_GRID_P2G_NAN_MASS;
            const int o = (lx0 + a) * Fr::H + (ly0 + b);
            unsafeAtomicAdd(&f[o], dm);
            unsafeAtomicAdd(&f[Fr::N + o], dpx);
            unsafeAtomicAdd(&f[2 * Fr::N + o], dpy);
        }
}

// Points per lane whose loads are in flight together (one batch of
// kPasses * 256 points is loaded, then computed), and staged per sub-round.
#ifndef ICE_PIPE_PASSES
#define ICE_PIPE_PASSES 4
#endif
// 1: a batch's point loads are issued while the previous batch is summed.
#ifndef ICE_PIPE_PREFETCH
#define ICE_PIPE_PREFETCH 1
#endif
constexpr int kPasses = ICE_PIPE_PASSES;
constexpr int kStage = kPasses * kTileThreads;
constexpr int kStageFields = 9;   // x, y, mass, mvx, mvy, s00, s01, s10, s11 (+ the cell, 16 bits)
// After a batch is summed, the staging area holds the 27 stencil sums of each
// of the round's 256 cells.
static_assert(kStageFields * kStage >= 27 * kTileThreads, "staging area too small for the cell stencils");

// What G2P reads of a point, loaded ahead of the computation.
// A point's G2P inputs as loaded.  coh is the cohesion line, or (no
// cohesion line) the thickness line: one unconditional load from a uniformly
// chosen line, turned into the Drucker-Prager threshold only where it is
// used, so that no computation right after the loads makes the wave wait for
// them (the passes' loads stay in flight together).
struct G2PRaw {
    uint32_t u, cell;
    float x, y, f00, f01, f10, f11, jp, th, coh;
};

// dp_threshold from the raw line value
__device__ __forceinline__ float raw_dpt(const DevParams &P, float coh) {
    return P.has_cohesion ? -coh * P.inv_tan_phi : P.DP_threshold_p;
}

template <class LT>
__device__ __forceinline__ void load_g2p_raw(const LT &L, const DevParams &P, uint32_t i, G2PRaw &r) {
    r.u = L.u(ICE_IDX_UTILITY, i);
    r.cell = L.u(ICE_IDX_CELL, i);
    r.x = L.get(ICE_IDX_POSX, i);
    r.y = L.get(ICE_IDX_POSX + 1, i);
    r.f00 = L.get(ICE_IDX_FE00, i);
    r.f01 = L.get(ICE_IDX_FE00 + 1, i);
    r.f10 = L.get(ICE_IDX_FE00 + 2, i);
    r.f11 = L.get(ICE_IDX_FE00 + 3, i);
    r.jp = L.get(ICE_IDX_JP_INV, i);
    r.th = L.get(ICE_IDX_THICKNESS, i);
    r.coh = L.get(P.has_cohesion ? ICE_IDX_COHESION : ICE_IDX_THICKNESS, i);
}

// A point's P2G terms into staging slot jj (mass 0 marks an empty slot), and
// its cell in frame coordinates (column | row << 8; 0xffff if outside).
// Staging position of the batch's k-th point.  The cell lanes read the points
// of 64 neighbouring cells, about four apart, so that lanes 8 apart would hit
// one LDS bank; an XOR of bits 5-6 into bits 0-1 (a permutation inside each
// block of 32) spreads them over distinct banks, and the staging writes of
// consecutive lanes stay conflict-free.
// (Not in the surrogate variant, which is at its register budget.)
template <bool SWZ>
__device__ __forceinline__ uint32_t stage_slot(uint32_t k) { return SWZ ? k ^ ((k >> 5) & 3u) : k; }

template <bool SWZ>
__device__ __forceinline__ void stage_terms(float (*s_st)[kStage], uint16_t *s_cw, uint32_t jj0, bool stage,
                                            const P2GPoint &q, int fx0, int fy0) {
    const uint32_t jj = stage_slot<SWZ>(jj0);
    if (stage) {
        const uint32_t lx = (uint32_t)(q.cx - fx0), ly = (uint32_t)(q.cy - fy0);
        s_cw[jj] = (lx < 256u && ly < 256u) ? (uint16_t)(lx | (ly << 8)) : (uint16_t)0xffffu;
        s_st[0][jj] = q.x;
        s_st[1][jj] = q.y;
        s_st[2][jj] = q.mass;
        s_st[3][jj] = q.mvx;
        s_st[4][jj] = q.mvy;
        s_st[5][jj] = q.s00;
        s_st[6][jj] = q.s01;
        s_st[7][jj] = q.s10;
        s_st[8][jj] = q.s11;
    } else {
        s_st[2][jj] = 0.f;
    }
}

// FUSED = false: P2G of the current point state into frames Fn (fills the
// pipeline).  FUSED = true: update + G2P of substep S from frames Fc, then P2G
// of the new state into Fn.
// The surrogate's deployment arguments (used when SURR).
struct SurrArg {
    MlpDev M;
    MlpNorm N;
    float tol;
    float lj0, lj1, d0, d1, jp0, jp1;   // feature ranges it was trained on
    int infer;                          // 0: recording only, the exact map projects every point
    // recording of the features of ridge-capped plastic trial states (rec_x
    // null: off): a state is kept when a hash of (slot, step) falls below
    // rec_keep (of 2^32), appended at rec_count[1] while it is below rec_cap;
    // rec_count[0] counts every such state seen (64-bit)
    float *rec_x;
    unsigned long long *rec_count;
    uint32_t rec_cap, rec_keep;
};

// BUF: 32-bit addressing of the points (BufLines) and frames, chosen by the
// host when the point buffer and the frames are small enough (buf_lines_ok,
// frames below 2^30 floats).
// Workgroups per CU the register budget is sized for.  Tiles up to 512 cells
// leave room in LDS for 3 (<= 168 VGPRs); the surrogate variant is held to
// that budget (its network would otherwise take it past 168 and to 2 waves
// per SIMD, ~30 % slower), the plain one is not (it stays below 168 by itself
// and schedules better without the bound).
template <int TW, int TH, bool SURR>
constexpr int tile_min_blocks() { return SURR && TW * TH <= 512 ? 3 : 2; }

template <int TW, int TH, bool FUSED, bool SURR = false, bool PERSIST = false, bool BUF = false>
__global__ __launch_bounds__(kTileThreads, (tile_min_blocks<TW, TH, SURR>())) void k_tile_pipeline(DevParams P0, StepArg S, SurrArg SA,
                                                                const DomainDesc *__restrict__ D,
                                                                const TileRef *__restrict__ T, int ntiles, Lines L0,
                                                                const uint32_t *__restrict__ cs0, GridLines G0,
                                                                Counters *cnt0, uint32_t *off0,
                                                                const float *__restrict__ Fc, float *__restrict__ Fn,
                                                                uint32_t *queue) {
    using Fr = Frame<TW, TH>;
    using LT = typename std::conditional<BUF, BufLines, Lines>::type;
    using IdxT = typename std::conditional<BUF, uint32_t, size_t>::type;
    static_assert(kTileThreads % TH == 0 && TH <= kTileThreads, "a round holds whole tile columns");
    constexpr int RC = kTileThreads / TH;              // tile columns per round
    constexpr int NROUND = (TW + RC - 1) / RC;
    __shared__ float s_acc[3 * Fr::NN];
    __shared__ float2 s_v[FUSED ? Fr::N : 1];
    __shared__ float s_st[kStageFields][kStage];
    __shared__ uint16_t s_cw[kStage];
    __shared__ uint32_t s_rb[RC];
    __shared__ uint32_t s_pre[RC + 1];
    __shared__ int s_next;
    __shared__ uint32_t s_tally[4];   // the tile's drift, accepted, fallback and recorded-state counts
    // recording: each wave's kept states and the workgroup's first record
    // slot, double-buffered over the passes of a batch
    __shared__ uint32_t s_rec[SURR ? 2 : 1][kTileThreads / 64];
    __shared__ unsigned long long s_rec_base[SURR ? 2 : 1];

    // Workgroup b takes tile-table entry xcd_remap(b).  PERSIST: the launch is
    // persistent, workgroup b (on XCD b % 8) takes the next entry of that
    // XCD's contiguous run of the table from queue[b % 8] (zeroed before the
    // launch) until the run is exhausted; the grid is a multiple of 8.
    for (int iter = 0;; iter++) {
        int entry;
        if constexpr (PERSIST) {
            const int xcd = (int)(blockIdx.x % 8u), per = ntiles / 8, rem = ntiles % 8;
            __syncthreads();   // the previous tile is done with every shared array
            if (threadIdx.x == 0) s_next = (int)atomicAdd(&queue[xcd], 1u);
            __syncthreads();
            const int j = s_next;   // uniform
            if (j >= per + (xcd < rem ? 1 : 0)) break;
            entry = xcd * per + min(xcd, rem) + j;
        } else {
            if (iter > 0) break;
            entry = xcd_remap(blockIdx.x, ntiles);
        }
        const TileRef tr = T[entry];
        const DomainDesc dd = D[tr.d];
        DevParams P = P0;
        P.gx = dd.gx;
        P.gy = dd.gy;
        P.has_current = dd.has_current;
        LT L;
        if constexpr (BUF) L = BufLines::make(L0.base + dd.pt_off, L0.pitch);
        else L = Lines{L0.base + dd.pt_off, L0.pitch};
        const uint32_t *__restrict__ cs = cs0 + dd.cs_base;
        const size_t nb = dd.node_base;
        const GridLines G{G0.m + nb, G0.px + nb, G0.py + nb, G0.cur_x + nb, G0.cur_y + nb, G0.fx + nb, G0.fy + nb,
                          G0.wind_x + nb, G0.wind_y + nb, G0.status + nb};
        Counters *cnt = cnt0 + tr.d;
        uint32_t *__restrict__ offtile = off0 + dd.off_base;
        const int x0 = tr.tx * TW, y0 = tr.ty * TH;
        const int fx0 = x0 - 2, fy0 = y0 - 2;                // frame origin

        uint32_t err = 0;
        double time = 0.0;
        bool rec = false;
        float ramp = 1.f;

        if constexpr (FUSED) {
            resolve_step(S, P, time, rec);
            ramp = current_ramp(time);
            // node velocities of T +- 2: every node's loads first (coordinates
            // clamped into the grid, so all loads are unconditional), then the update
            constexpr int NIT = (Fr::N + kTileThreads - 1) / kTileThreads;
            float nm[NIT], nmx[NIT], nmy[NIT], nwx[NIT], nwy[NIT], ncx[NIT], ncy[NIT];
            uint8_t nst[NIT];
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                const int k = min(it * kTileThreads + (int)threadIdx.x, Fr::N - 1);
                const int lx = k / Fr::H, ly = k - (k / Fr::H) * Fr::H;
                const int x = min(max(fx0 + lx, 0), P.gx - 1), y = min(max(fy0 + ly, 0), P.gy - 1);
                frame_node_sum<TW, TH, IdxT>(Fc, dd, x, y, nm[it], nmx[it], nmy[it]);
                const size_t node = (size_t)x * P.gy + y;
                nst[it] = G.status[node];
                nwx[it] = G.wind_x[node];
                nwy[it] = G.wind_y[node];
                ncx[it] = P.has_current ? G.cur_x[node] : 0.f;
                ncy[it] = P.has_current ? G.cur_y[node] : 0.f;
            }
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                const int k = it * kTileThreads + (int)threadIdx.x;
                if (k < Fr::N) {
                    const int lx = k / Fr::H, ly = k - (k / Fr::H) * Fr::H;
                    const int x = fx0 + lx, y = fy0 + ly;
                    float2 v = make_float2(0.f, 0.f);
                    if (x >= 0 && x < P.gx && y >= 0 && y < P.gy) {
                        bool wall;
                        node_velocity(nm[it], nmx[it], nmy[it], nst[it], ncx[it], ncy[it], nwx[it], nwy[it], ramp, P, v.x,
                                      v.y, wall);
                        if (lx >= 2 && lx < TW + 2 && ly >= 2 && ly < TH + 2) {   // T's own node
                            const size_t node = (size_t)x * P.gy + y;
                            if (wall) { G.fx[node] += nmx[it]; G.fy[node] += nmy[it]; }
                            if (nm[it] != 0.f && (__builtin_isnan(v.x) || __builtin_isnan(v.y))) err |= ICE_ERR_GRID_NAN;
                        }
                    }
                    s_v[k] = v;
                }
            }
        }
        for (int k = threadIdx.x; k < 3 * Fr::NN; k += kTileThreads) s_acc[k] = 0.f;
        if (threadIdx.x < 4) s_tally[threadIdx.x] = 0u;
        // wave-uniform counts, added to the domain's counters once per tile:
        // one global atomic per wave on a domain-wide address serialises at
        // the L2 when most waves count (fast flow, the learned projection)
        uint32_t w_drift = 0, w_acc = 0, w_fb = 0, w_seen = 0;

        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int rd = 0; rd < NROUND; rd++) {
            const int col0 = rd * RC;
            const int ncol = min(RC, TW - col0);
            column_ranges(cs, P, x0 + col0, ncol, y0, y0 + TH, s_rb, s_pre);   // contains __syncthreads
            const uint32_t n_round = s_pre[ncol];
            // this lane's cell and the flattened range of its points
            const int cc = (int)threadIdx.x / TH, rr = (int)threadIdx.x - cc * TH;
            const int cx = x0 + col0 + cc, cy = y0 + rr;
            uint32_t f0 = 0, flen = 0;
            if (cc < ncol && cx < P.gx && cy < P.gy) {
                const size_t cell = (size_t)cx * P.gy + cy;
                const uint32_t b = cs[cell];
                f0 = s_pre[cc] + (b - s_rb[cc]);
                flen = cs[cell + 1] - b;
            }
            // REG: the round's column starts in registers (rounds of at most 8
            // columns, i.e. tiles 32 cells high): a point's column is the count
            // of starts at or below its flattened index, with no chain of
            // dependent LDS reads (the binary and linear searches over s_pre
            // cost about five LDS round trips per pass; 0.738 -> 0.709 ms per
            // launch on the ice field).  Not in the surrogate variant, which is
            // at its register budget.
            constexpr bool REG = !SURR && RC <= 8;
            uint32_t pre_r[REG ? RC : 1];
            if constexpr (REG) {
#pragma unroll
                for (int q = 1; q < RC; q++) pre_r[q] = q < ncol ? s_pre[q] : 0xffffffffu;
            }
            // this lane's point of each pass of the batch at `base` (slot, or ~0u for none)
            auto locate = [&](uint32_t base, uint32_t (&sl)[kPasses]) {
#pragma unroll
                for (int p = 0; p < kPasses; p++) {
                    sl[p] = ~0u;
                    const uint32_t jw = base + (uint32_t)(p * kTileThreads + wave * 64);
                    if (jw < n_round) {   // wave-uniform
                        const uint32_t j = jw + lane;
                        if constexpr (REG) {
                            if (j < n_round) {
                                int c = 0;
#pragma unroll
                                for (int q = 1; q < RC; q++) c += pre_r[q] <= j ? 1 : 0;
                                sl[p] = s_rb[c] + (j - s_pre[c]);
                            }
                        } else {
                            const int cb = flat_col_n(jw, s_pre, ncol);
                            if (j < n_round) {
                                const int c = col_from_n(j, cb, s_pre, ncol);
                                sl[p] = s_rb[c] + (j - s_pre[c]);
                            }
                        }
                    }
                }
            };
            // PF: the point loads of the next batch are issued right after the
            // staging barrier, so that their latency passes under the batch's
            // three LDS-only P2G phases (barriers wait on LDS counts only)
            constexpr bool PF = FUSED && !SURR && ICE_PIPE_PREFETCH;
            uint32_t slot_n[kPasses];
            G2PRaw raw_n[PF ? kPasses : 1];
            if constexpr (PF) {
                locate(0u, slot_n);
#pragma unroll
                for (int p = 0; p < kPasses; p++)
                    if (slot_n[p] != ~0u) load_g2p_raw(L, P, slot_n[p], raw_n[p]);
            }
            for (uint32_t sub = 0; sub < n_round; sub += kStage) {   // uniform bound
                uint32_t slot[kPasses];
                if constexpr (PF) {
#pragma unroll
                    for (int p = 0; p < kPasses; p++) slot[p] = slot_n[p];
                } else {
                    locate(sub, slot);
                }
                if constexpr (FUSED) {
                    // every pass's loads first, then the computation (the surrogate
                    // variant, short of registers, loads each pass when it computes it)
                    G2PRaw raw[kPasses];
                    if constexpr (PF) {
#pragma unroll
                        for (int p = 0; p < kPasses; p++) raw[p] = raw_n[p];
                    } else if constexpr (!SURR) {
#pragma unroll
                        for (int p = 0; p < kPasses; p++)
                            if (slot[p] != ~0u) load_g2p_raw(L, P, slot[p], raw[p]);
                    }
#pragma unroll
                    for (int p = 0; p < kPasses; p++) {
                        const uint32_t jj = (uint32_t)(p * kTileThreads) + threadIdx.x;   // staging slot
                        bool stage = false;
                        P2GPoint q;
                        const uint32_t i = slot[p];
                        G2PRaw rw;
                        if constexpr (SURR) {
                            if (i != ~0u) load_g2p_raw(L, P, i, rw);
                        } else {
                            rw = raw[p];
                        }
                        bool act = i != ~0u && !(rw.u & (kDisabled | kEscaped));
                        G2PState s;
                        s.u = rw.u;
                        s.cx = (int)(rw.cell & 0xffff);
                        s.cy = (int)(rw.cell >> 16);
                        s.x = rw.x;
                        s.y = rw.y;
                        s.Fe = M2{rw.f00, rw.f01, rw.f10, rw.f11};
                        s.Jp = rw.jp;
                        const float th = rw.th;
                        const float dpt = raw_dpt(P, rw.coh);
                        float pvx = 0.f, pvy = 0.f;
                        M2 C{0.f, 0.f, 0.f, 0.f};
                        FInv vf;
                        // outside the frame cannot happen while the drift invariant
                        // holds; a stencil outside the grid (an uploaded point at the
                        // edge) is skipped, as by the other paths.  (One flat test and
                        // one branch: a nested branch here costs register copies.)
                        const int rx = s.cx - fx0, ry = s.cy - fy0;
                        const bool outside = (uint32_t)(rx - 1) > (uint32_t)(Fr::W - 3) ||
                                             (uint32_t)(ry - 1) > (uint32_t)(Fr::H - 3) || !stencil_inside(s.cx, s.cy, P);
                        if (act && outside) err |= ICE_ERR_POINT_LEFT_AREA;
                        act = act && !outside;
                        if (act) {
                            float wx[3], wy[3];
                            weights(s.x, wx);
                            weights(s.y, wy);
                            float2 nv[3][3];
#pragma unroll
                            for (int a = 0; a < 3; a++)
#pragma unroll
                                for (int b = 0; b < 3; b++) nv[a][b] = s_v[(rx + a - 1) * Fr::H + (ry + b - 1)];
                            g2p_gather(nv, s.x, s.y, wx, wy, P.four_h_inv, pvx, pvy, C);
                            g2p_predict(P, s, pvx, pvy, C, err, vf);
                        }
                        bool exact_rm = act && (s.u & kCrushed);
                        if constexpr (SURR) {
                            // the plastic branch through the learned projection, one
                            // wave of points per MFMA batch (all lanes take part)
                            // the plastic branch with the ridge cap active (what it was trained on)
                            bool want = exact_rm && th == 1.f && s.p_tr >= dpt &&
                                        s.p_tr >= P.ridge * th * s.Jp && s.q_tr > q_yield_at(s.p_tr, dpt, P);
                            float xr[kMlpIn] = {0.f, 0.f, 1.f, 0.f};
                            if (want) mlp_features(vf, s.Jp, -dpt * P.tan_phi, xr);
                            if (SA.rec_x) {   // uniform: the recording of trial states
                                const uint32_t hsh = (uint32_t)mix64(((uint64_t)i << 24) ^
                                                                     (uint64_t)__double_as_longlong(time));
                                const bool keep = want && hsh < SA.rec_keep;
                                const uint64_t bw = __ballot(want), bk = __ballot(keep);
                                w_seen += (uint32_t)__popcll(bw);
                                // the workgroup's kept states take one block of record
                                // slots (one atomic per workgroup and pass, not per wave)
                                const int rb = p & 1;
                                if (lane == 0) s_rec[rb][wave] = (uint32_t)__popcll(bk);
                                __syncthreads();
                                if (threadIdx.x == 0) {
                                    uint32_t tot = 0;
#pragma unroll
                                    for (int w = 0; w < kTileThreads / 64; w++) tot += s_rec[rb][w];
                                    s_rec_base[rb] = tot ? atomicAdd(&SA.rec_count[1], (unsigned long long)tot) : 0ull;
                                }
                                __syncthreads();
                                unsigned long long base = s_rec_base[rb];
                                for (int w = 0; w < wave; w++) base += s_rec[rb][w];
                                const uint64_t slot_rec = base + __popcll(bk & ((1ull << lane) - 1ull));
                                if (keep && slot_rec < SA.rec_cap)
                                    for (int k = 0; k < kMlpIn; k++) SA.rec_x[slot_rec * kMlpIn + k] = xr[k];
                            }
                            if (want)
                                want = SA.infer && xr[0] >= SA.lj0 && xr[0] <= SA.lj1 && xr[1] >= SA.d0 &&
                                       xr[1] <= SA.d1 && xr[2] >= SA.jp0 && xr[2] <= SA.jp1;
                            if (__any(want)) {
                                float xs[kMlpIn];
#pragma unroll
                                for (int k = 0; k < kMlpIn; k++) xs[k] = want ? (xr[k] - SA.N.xm[k]) / SA.N.xs[k] : 0.f;
                                float y[kMlpOut];
                                mlp_infer(SA.M, xs, y);
                                bool accepted = false;
                                if (want) {
                                    float m0, m1, jpf, pq, qq, viol;
                                    accepted = mlp_decode(SA.N, y, xr[0], s.Jp, dpt, th, P, SA.tol, m0, m1, jpf, pq, qq, viol);
                                    if (accepted) {
                                        s.Fe = rebuild(vf, m0, m1);
                                        s.Jp *= jpf;
                                        vf = invariants(s.Fe);
                                        if (isnan4(s.Fe)) err |= ICE_ERR_POINT_FE_NAN;
                                        exact_rm = false;
                                    }
                                }
                                const uint64_t ba = __ballot(accepted), bw = __ballot(want);
                                w_acc += (uint32_t)__popcll(ba);
                                w_fb += (uint32_t)__popcll(bw & ~ba);
                            }
                        }
                        if (act) {
                            if (exact_rm) g2p_return_map(P, s, th, dpt, err, vf);
                            if (s.newly_disabled) atomicAdd(&cnt->disabled, 1);
                            bool escaped = false;
                            if (s.cell_updated && (s.u & (kDisabled | kFar)) == kFar) {
                                s.u |= kEscaped | kFresh;
                                const uint32_t e = atomicAdd(&cnt->n_esc, 1u);
                                offtile[e] = i;
                                escaped = true;
                            }
                            // escaped and disabled points leave the pipeline: store all
                            g2p_store_pipelined(L, i, s, rw.u, rw.jp, pvx, pvy, C, rec,
                                                S.complete || escaped || (s.u & kDisabled));
                            if (!(s.u & kDisabled) && !escaped) {
                                p2g_terms(P, s.x, s.y, s.cx, s.cy, pvx, pvy, vf, C, s.Jp, th, q);
                                stage = true;
                            }
                        }
                        if (S.complete)   // points off their sorted cell at a chunk's end (re-sort policy)
                            w_drift += (uint32_t)__popcll(__ballot(stage && (s.u & kMoved)));
                        stage_terms<!SURR>(s_st, s_cw, jj, stage, q, fx0, fy0);
                    }
                } else {
#pragma unroll
                    for (int p = 0; p < kPasses; p++) {
                        const uint32_t jj = (uint32_t)(p * kTileThreads) + threadIdx.x;
                        bool stage = false;
                        P2GPoint q;
                        const uint32_t i = slot[p];
                        if (i != ~0u) {
                            const uint32_t u = L.u(ICE_IDX_UTILITY, i);
                            if (!(u & (kDisabled | kEscaped))) {
                                load_p2g_point(L, i, L.u(ICE_IDX_CELL, i), P, q);
                                if (!stencil_inside(q.cx, q.cy, P)) err |= ICE_ERR_POINT_LEFT_AREA;
                                else stage = true;
                            }
                        }
                        stage_terms<!SURR>(s_st, s_cw, jj, stage, q, fx0, fy0);
                    }
                }
                __syncthreads();
                if constexpr (PF) {
                    if (sub + (uint32_t)kStage < n_round) {   // uniform
                        locate(sub + (uint32_t)kStage, slot_n);
#pragma unroll
                        for (int p = 0; p < kPasses; p++)
                            if (slot_n[p] != ~0u) load_g2p_raw(L, P, slot_n[p], raw_n[p]);
                    }
                }
                // each cell lane sums the staged points sorted into its cell
                float am[3][3], ax[3][3], ay[3][3];
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) am[a][b] = ax[a][b] = ay[a][b] = 0.f;
                const uint32_t lo = max(f0, sub), hi = min(f0 + flen, sub + (uint32_t)kStage);
                for (uint32_t jf = lo; jf < hi; jf++) {
                    const uint32_t k = stage_slot<!SURR>(jf - sub);
                    P2GPoint t;
                    t.mass = s_st[2][k];
                    if (t.mass == 0.f) continue;
                    t.x = s_st[0][k];
                    t.y = s_st[1][k];
                    t.mvx = s_st[3][k];
                    t.mvy = s_st[4][k];
                    t.s00 = s_st[5][k];
                    t.s01 = s_st[6][k];
                    t.s10 = s_st[7][k];
                    t.s11 = s_st[8][k];
                    float wx[3], wy[3];
                    weights(t.x, wx);
                    weights(t.y, wy);
                    const uint32_t pc = s_cw[k];
                    const int sx = (int)(pc & 0xffu) - (cx - fx0), sy = (int)(pc >> 8) - (cy - fy0);
                    if ((sx | sy) != 0 && !moved_stencil<TW, TH>(t, sx, sy, cx - fx0, cy - fy0, wx, wy, s_acc, err))
                        continue;
                    p2g_accumulate(t, wx, wy, am, ax, ay);
                }
                __syncthreads();   // staged terms consumed: the area now takes the cell stencils
                float *s_cs = &s_st[0][0];   // [27][kTileThreads]: component-major, offset (a, b), cell lane
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        const int q = a * 3 + b;
                        s_cs[q * kTileThreads + threadIdx.x] = am[a][b];
                        s_cs[(9 + q) * kTileThreads + threadIdx.x] = ax[a][b];
                        s_cs[(18 + q) * kTileThreads + threadIdx.x] = ay[a][b];
                    }
                __syncthreads();
                // each node the round touches gathers its (up to) nine cell stencils
                // in a fixed order: frame columns col0 + 1 .. col0 + ncol + 2, rows 1 .. TH + 2
                const int nyn = TH + 2;
                for (int k = threadIdx.x; k < (ncol + 2) * nyn; k += kTileThreads) {
                    const int ix = k / nyn, iy = k - ix * nyn;
                    float sm = 0.f, sx = 0.f, sy = 0.f;
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        const int ccn = ix - a;
                        if (ccn < 0 || ccn >= ncol) continue;
#pragma unroll
                        for (int b = 0; b < 3; b++) {
                            const int rrn = iy - b;
                            if (rrn < 0 || rrn >= TH) continue;
                            const int c = ccn * TH + rrn, q = a * 3 + b;
                            sm += s_cs[q * kTileThreads + c];
                            sx += s_cs[(9 + q) * kTileThreads + c];
                            sy += s_cs[(18 + q) * kTileThreads + c];
                        }
                    }
                    const int o = (col0 + 1 + ix) * Fr::LP + (1 + iy);
                    s_acc[o] += sm;
                    s_acc[Fr::NN + o] += sx;
                    s_acc[2 * Fr::NN + o] += sy;
                }
                __syncthreads();   // stencils read: the area takes the next staged terms
            }
            __syncthreads();   // the round's ranges are read; the next round may overwrite them
        }
        if (S.complete || SURR) {   // uniform
            if (lane == 0) {
                if (w_drift) atomicAdd(&s_tally[0], w_drift);
                if (w_acc) atomicAdd(&s_tally[1], w_acc);
                if (w_fb) atomicAdd(&s_tally[2], w_fb);
                if (w_seen) atomicAdd(&s_tally[3], w_seen);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                if (s_tally[0]) atomicAdd(&cnt->n_drift, s_tally[0]);
                if (s_tally[1]) atomicAdd(&cnt->pad[0], s_tally[1]);
                if (s_tally[2]) atomicAdd(&cnt->pad[1], s_tally[2]);
                if (s_tally[3]) atomicAdd(&SA.rec_count[0], (unsigned long long)s_tally[3]);
            }
        }
        // the frame, every node once (coalesced).  A NaN P2G contribution
        // (kernels.cuh:97-98) makes its node's sum NaN, so the sums are checked
        // here, once per node, rather than per cell stencil and batch.
        float *__restrict__ out = Fn + frame_offset<TW, TH>(dd, tr.tx, tr.ty);
        for (int k = threadIdx.x; k < 3 * Fr::N; k += kTileThreads) {
            const int comp = k / Fr::N, nd = k - comp * Fr::N;
            const int lx = nd / Fr::H, ly = nd - lx * Fr::H;
            const float v = s_acc[comp * Fr::NN + lx * Fr::LP + ly];
            if (__builtin_isnan(v)) err |= comp == 0 ? ICE_ERR_GRID_P2G_NAN_MASS : ICE_ERR_GRID_P2G_NAN_VEL;
            out[k] = v;
        }
        flush_err(cnt, err);
    }
}

// One off-tile point i of a domain (k_offtile_pipeline).
template <int TW, int TH>
__device__ __forceinline__ void offtile_point(const DevParams &P, const DomainDesc &dd, const Lines &L,
                                              const GridLines &G, Counters *cnt, uint32_t i, const float *Fc,
                                              float *Fn, int do_g2p, bool rec, float ramp, uint32_t &err) {
    uint32_t u = L.u(ICE_IDX_UTILITY, i);
    if (u & kDisabled) return;
    if (u & kFresh) {
        u &= ~kFresh;
        L.set_u(ICE_IDX_UTILITY, i, u);
    } else if (do_g2p) {
        G2PState s;
        load_g2p_state(L, i, u, s);
        if (!stencil_inside(s.cx, s.cy, P)) {
            err |= ICE_ERR_POINT_LEFT_AREA;
            return;
        }
        const float th = L(ICE_IDX_THICKNESS, i);
        float wx[3], wy[3];
        weights(s.x, wx);
        weights(s.y, wy);
        float2 nv[3][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                float mx, my, m;
                bool wall;
                frame_node_velocity<TW, TH>(Fc, dd, G, P, ramp, s.cx + a - 1, s.cy + b - 1, nv[a][b].x, nv[a][b].y, mx,
                                            my, wall, m);
            }
        float pvx, pvy;
        M2 C;
        g2p_gather(nv, s.x, s.y, wx, wy, P.four_h_inv, pvx, pvy, C);
        FInv vf;
        g2p_finish(P, s, pvx, pvy, C, th, dp_threshold(P, L, i), err, vf);
        if (s.newly_disabled) atomicAdd(&cnt->disabled, 1);
        g2p_store(L, i, s, pvx, pvy, C, rec);
        if (s.u & kDisabled) return;
    }
    P2GPoint q;
    load_p2g_point(L, i, L.u(ICE_IDX_CELL, i), P, q);
    if (!stencil_inside(q.cx, q.cy, P)) {
        err |= ICE_ERR_POINT_LEFT_AREA;
        return;
    }
    frame_scatter_global<TW, TH>(q, dd, Fn, err);
}

// The off-tile points after a pipeline pass.  do_g2p = 1 (fused substep):
// points that escaped in an earlier substep run update + G2P from frames Fc;
// every off-tile point then adds its P2G into Fn with global atomics.  Points
// marked kFresh (escaped in this substep's tile pass) only need the P2G.
// The domains' lists form one item space (pre: exclusive prefix of their
// lengths, from k_offtile_prefix; null for a single domain), so a long list in
// one domain is spread over the whole launch and a batch of empty lists costs
// one small grid.
template <int TW, int TH>
__global__ __launch_bounds__(256) void k_offtile_pipeline(DevParams P0, StepArg S, const DomainDesc *__restrict__ D,
                                                          int ndom, const uint32_t *__restrict__ pre, Lines L0,
                                                          GridLines G0, Counters *cnt0,
                                                          const uint32_t *__restrict__ off0, const float *Fc,
                                                          float *Fn, int do_g2p) {
    const uint32_t total = offtile_total(pre, ndom, cnt0);
    double time = 0.0;
    bool rec = false;
    if (do_g2p) resolve_step(S, P0, time, rec);
    const float ramp = current_ramp(time);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < total; j += gridDim.x * blockDim.x) {
        const int d = offtile_domain(pre, ndom, j);
        const DomainDesc dd = D[d];
        Counters *cnt = cnt0 + d;
        DevParams P = P0;
        P.gx = dd.gx;
        P.gy = dd.gy;
        P.has_current = dd.has_current;
        const Lines L{L0.base + dd.pt_off, L0.pitch};
        const size_t nb = dd.node_base;
        const GridLines G{G0.m + nb, G0.px + nb, G0.py + nb, G0.cur_x + nb, G0.cur_y + nb, G0.fx + nb, G0.fy + nb,
                          G0.wind_x + nb, G0.wind_y + nb, G0.status + nb};
        uint32_t err = 0;
        const uint32_t i = off0[dd.off_base + (j - (pre ? pre[d] : 0u))];
        offtile_point<TW, TH>(P, dd, L, G, cnt, i, Fc, Fn, do_g2p, rec, ramp, err);
        if (err) atomicOr(&cnt->err, err);   // rare: per item, to the item's domain
    }
}

// Clears every domain's drift count before a run chunk.
__global__ __launch_bounds__(256) void k_clear_drift(Counters *cnt, int ndom) {
    for (int d = blockIdx.x * blockDim.x + threadIdx.x; d < ndom; d += gridDim.x * blockDim.x) cnt[d].n_drift = 0;
}

// Exclusive prefix of the domains' off-tile list lengths, pre[0..ndom], one
// workgroup (the item space of k_offtile_pipeline for a batch).
__global__ __launch_bounds__(256) void k_offtile_prefix(const Counters *__restrict__ cnt, int ndom, uint32_t *pre) {
    __shared__ uint32_t s[256];
    uint32_t carry = 0;
    for (int base = 0; base < ndom; base += 256) {
        const int d = base + (int)threadIdx.x;
        const uint32_t v = d < ndom ? cnt[d].n_esc : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const uint32_t t = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (d < ndom) pre[d] = carry + s[threadIdx.x] - v;
        carry += s[255];
        __syncthreads();
    }
    if (threadIdx.x == 0) pre[ndom] = carry;
}

// Node sums of domain d from frames F into the grid lines (m, px, py).
template <int TW, int TH>
__global__ __launch_bounds__(256) void k_frames_to_grid(const DomainDesc *__restrict__ D, int d, const float *F,
                                                        GridLines G0) {
    const DomainDesc dd = D[d];
    const size_t N = (size_t)dd.gx * dd.gy, nb = dd.node_base;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < N; k += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(k / dd.gy), y = (int)(k - (size_t)x * dd.gy);
        float m, px, py;
        frame_node_sum<TW, TH>(F, dd, x, y, m, px, py);
        G0.m[nb + k] = m;
        G0.px[nb + k] = px;
        G0.py[nb + k] = py;
    }
}

}  // namespace ice
