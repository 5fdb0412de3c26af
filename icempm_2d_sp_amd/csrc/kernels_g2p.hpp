// kernels_g2p.hpp — the sorted path's fused node update + G2P, one
// workgroup per tile of TW x TH nodes, plus the off-tile point kernel.
//
// The tile computes node velocities (partition_kernel_update_nodes,
// simulation/kernels.cuh:109-177, plus the wind-stress term) for nodes T +- 2
// into LDS, adds wall momentum of T's own nodes to (fx, fy), then runs G2P
// (kernels.cuh:181-315) for the points sorted into T with coalesced SoA loads
// and stores.  A point whose cell changes keeps its drift from the cell it
// was sorted into (kMoved while nonzero); one that leaves the 3 x 3 around
// that cell sets kEscaped and is appended to the domain's off-tile list.
#pragma once
#include "tile_common.hpp"

namespace ice {

template <int TW, int TH>
__global__ __launch_bounds__(kTileThreads) void k_g2p_tiles(DevParams P0, StepArg S, const DomainDesc *__restrict__ D,
                                                            const TileRef *__restrict__ T, int ntiles, Lines L0,
                                                            const uint32_t *__restrict__ cs0, GridLines G0,
                                                            Counters *cnt0, uint32_t *off0) {
    const TileCtx tc = tile_ctx<TW, TH>(P0, D, T, ntiles, L0, cs0, G0, cnt0, off0);
    const DevParams &P = tc.P;
    const Lines &L = tc.L;
    const uint32_t *__restrict__ cs = tc.cs;
    const GridLines &G = tc.G;
    Counters *cnt = tc.cnt;
    uint32_t *__restrict__ offtile = tc.offtile;
    const int x0 = tc.x0, y0 = tc.y0;
    constexpr int RW = TW + 4, RH = TH + 4;
    __shared__ float2 s_v[RW * RH];
    __shared__ uint32_t s_rb[TW];
    __shared__ uint32_t s_pre[TW + 1];

    double time;
    bool rec;
    resolve_step(S, P, time, rec);
    const float ramp = current_ramp(time);
    uint32_t err = 0;

    for (int k = threadIdx.x; k < RW * RH; k += kTileThreads) {
        const int lx = k / RH, ly = k - (k / RH) * RH;
        const int x = x0 - 2 + lx, y = y0 - 2 + ly;
        float2 v = make_float2(0.f, 0.f);
        if (x >= 0 && x < P.gx && y >= 0 && y < P.gy) {
            const size_t node = (size_t)x * P.gy + y;
            const float m = G.m[node], mx = G.px[node], my = G.py[node];
            float cx = 0.f, cy = 0.f;
            if (P.has_current) { cx = G.cur_x[node]; cy = G.cur_y[node]; }
            bool wall;
            node_velocity(m, mx, my, G.status[node], cx, cy, G.wind_x[node], G.wind_y[node], ramp, P, v.x, v.y, wall);
            const bool own = lx >= 2 && lx < TW + 2 && ly >= 2 && ly < TH + 2;
            if (own) {
                if (wall) { G.fx[node] += mx; G.fy[node] += my; }
                if (m != 0.f && (__builtin_isnan(v.x) || __builtin_isnan(v.y))) err |= ICE_ERR_GRID_NAN;
            }
        }
        s_v[k] = v;
    }
    column_ranges(cs, P, x0, TW, y0, y0 + TH, s_rb, s_pre);   // contains __syncthreads

    const uint32_t total = s_pre[TW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // each wave takes every fourth chunk of 64 consecutive items
    for (uint32_t kb = (uint32_t)wave * 64; kb < total; kb += kTileThreads) {
        const int cb = flat_col<TW>(kb, s_pre);
        const uint32_t k = kb + lane;
        if (k >= total) continue;
        const int c = col_from<TW>(k, cb, s_pre);
        const uint32_t i = s_rb[c] + (k - s_pre[c]);
        // issue every load before testing the flags
        const uint32_t u = L.u(ICE_IDX_UTILITY, i);
        G2PState s;
        load_g2p_state(L, i, u, s);
        const float th = L(ICE_IDX_THICKNESS, i);
        const float dpt = dp_threshold(P, L, i);
        if (u & (kDisabled | kEscaped)) continue;
        const int rx = s.cx - x0 + 2, ry = s.cy - y0 + 2;
        // outside the frame cannot happen while the drift invariant holds; a
        // stencil outside the grid (an uploaded point at the edge) is skipped
        if (rx < 1 || rx > RW - 2 || ry < 1 || ry > RH - 2 || !stencil_inside(s.cx, s.cy, P)) {
            err |= ICE_ERR_POINT_LEFT_AREA;
            continue;
        }
        float wx[3], wy[3];
        weights(s.x, wx);
        weights(s.y, wy);
        float2 nv[3][3];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) nv[a][b] = s_v[(rx + a - 1) * RH + (ry + b - 1)];
        float pvx, pvy;
        M2 C;
        g2p_gather(nv, s.x, s.y, wx, wy, P.four_h_inv, pvx, pvy, C);
        FInv vf;
        g2p_finish(P, s, pvx, pvy, C, th, dpt, err, vf);
        if (s.newly_disabled) atomicAdd(&cnt->disabled, 1);
        if (s.cell_updated && (s.u & (kDisabled | kFar)) == kFar) {
            s.u |= kEscaped;
            const uint32_t j = atomicAdd(&cnt->n_esc, 1u);
            offtile[j] = i;
        }
        g2p_store(L, i, s, pvx, pvy, C, rec);
    }
    flush_err(cnt, err);
}

// G2P of the off-tile points, node velocities recomputed from the P2G output
// (the same function the tiles evaluate).  Runs before k_g2p_tiles so points
// that the tile G2P appends to the list are not processed twice.  All
// domains' lists as one item space (offtile_domain).
__global__ __launch_bounds__(256) void k_g2p_offtile(DevParams P0, StepArg S, const DomainDesc *__restrict__ D,
                                                     int ndom, const uint32_t *__restrict__ pre, Lines L0,
                                                     GridLines G0, Counters *cnt0,
                                                     const uint32_t *__restrict__ off0) {
    const uint32_t total = offtile_total(pre, ndom, cnt0);
    double time;
    bool rec;
    resolve_step(S, P0, time, rec);
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
        const float *gm = G0.m + nb, *gpx = G0.px + nb, *gpy = G0.py + nb;
        const float *ccx = G0.cur_x + nb, *ccy = G0.cur_y + nb, *wwx = G0.wind_x + nb, *wwy = G0.wind_y + nb;
        const uint8_t *st = G0.status + nb;
        const uint32_t i = off0[dd.off_base + (j - (pre ? pre[d] : 0u))];
        const uint32_t u = L.u(ICE_IDX_UTILITY, i);
        if (u & kDisabled) continue;
        uint32_t err = 0;
        G2PState s;
        load_g2p_state(L, i, u, s);
        if (!stencil_inside(s.cx, s.cy, P)) {
            atomicOr(&cnt->err, ICE_ERR_POINT_LEFT_AREA);
            continue;
        }
        const float th = L(ICE_IDX_THICKNESS, i);
        float wx[3], wy[3];
        weights(s.x, wx);
        weights(s.y, wy);
        float2 nv[3][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                const size_t node = (size_t)(s.cy + b - 1) + (size_t)(s.cx + a - 1) * P.gy;
                float cx = 0.f, cy = 0.f;
                if (P.has_current) { cx = ccx[node]; cy = ccy[node]; }
                bool wall;
                node_velocity(gm[node], gpx[node], gpy[node], st[node], cx, cy, wwx[node], wwy[node], ramp, P,
                              nv[a][b].x, nv[a][b].y, wall);
            }
        float pvx, pvy;
        M2 C;
        g2p_gather(nv, s.x, s.y, wx, wy, P.four_h_inv, pvx, pvy, C);
        FInv vf;
        g2p_finish(P, s, pvx, pvy, C, th, dp_threshold(P, L, i), err, vf);
        if (s.newly_disabled) atomicAdd(&cnt->disabled, 1);
        g2p_store(L, i, s, pvx, pvy, C, rec);
        if (err) atomicOr(&cnt->err, err);   // rare: per item, to the item's domain
    }
}

// Node velocities of domain d for time t into (vx, vy), for inspection.
__global__ __launch_bounds__(256) void k_node_velocity_out(DevParams P0, double t, const DomainDesc *__restrict__ D,
                                                           int d, GridLines G0, float *__restrict__ vx,
                                                           float *__restrict__ vy) {
    const DomainDesc dd = D[d];
    DevParams P = P0;
    P.has_current = dd.has_current;
    const size_t N = (size_t)dd.gx * dd.gy, nb = dd.node_base;
    const float ramp = current_ramp(t);
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < N; k += (size_t)gridDim.x * blockDim.x) {
        const size_t g = nb + k;
        float cx = 0.f, cy = 0.f;
        if (P.has_current) { cx = G0.cur_x[g]; cy = G0.cur_y[g]; }
        bool wall;
        float a, b;
        node_velocity(G0.m[g], G0.px[g], G0.py[g], G0.status[g], cx, cy, G0.wind_x[g], G0.wind_y[g], ramp, P, a, b,
                      wall);
        vx[k] = a;
        vy[k] = b;
    }
}

}  // namespace ice
