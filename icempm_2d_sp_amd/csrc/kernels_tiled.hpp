// kernels_tiled.hpp — the sorted-points substep: pull P2G staged in LDS and a
// fused node-update + G2P, one workgroup per TW x TH tile of grid nodes.
//
// Invariant (kept by the device sort and by the kMoved/kEscaped bits): points
// are in cell order as of the last sort, cs[k] is the first sorted slot of
// cell k (k = x*GridY + y), and every point that is not on the off-tile list
// sits at most one cell (Chebyshev) from the cell it was sorted into.  A
// point that changes cell a second time since the sort is flagged kEscaped,
// appended to the off-tile list and handled by k_p2g_offtile/k_g2p_offtile
// until the next sort.
//
// P2G "pulls": the workgroup of tile T reads the points sorted into cells
// T +- 2 (their contiguous per-column ranges of cs), accumulates into LDS only
// the contributions to T's own nodes, and stores those nodes once.  Every
// node has exactly one writer, so no global atomics and no grid clear are
// needed.  With DET the LDS sums are int64 fixed point, which makes the result
// independent of accumulation order.
//
// G2P computes node velocities (the reference's update kernel) for T +- 2
// into LDS, adds wall momentum of T's own nodes to (fx, fy), then runs G2P for
// the points sorted into T.
#pragma once
#include "point_ops.hpp"

namespace ice {

constexpr int kTileThreads = 256;

// Workgroup b runs on XCD b % 8; give each XCD one contiguous run of tiles so
// neighbouring tiles, which share halo points and nodes, meet in one L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    const int nx = 8;
    const int xcd = bid % nx, j = bid / nx;
    const int per = nb / nx, rem = nb % nx;
    return xcd * per + min(xcd, rem) + j;
}

// Tile of workgroup `bid`: XCD-contiguous runs (xcd_remap) of a band order
// (bands of `band` tile columns, rows outer, columns inner), so the tiles in
// flight on one XCD form a compact block whose x and y neighbours share halo
// points and nodes in that XCD's L2.
__device__ __forceinline__ void tile_coords(int bid, int ntx, int nty, int band, int &tx, int &ty) {
    const int t = xcd_remap(bid, ntx * nty);
    const int b = t / (band * nty);
    const int wb = min(band, ntx - b * band);
    const int r = t - b * band * nty;
    ty = r / wb;
    tx = b * band + (r - ty * wb);
}

// Column of flattened item k (last column whose prefix is <= k), searched
// from a wave-uniform start column cb.
template <int NCOL>
__device__ __forceinline__ int col_from(uint32_t k, int cb, const uint32_t *s_pre) {
    int c = cb;
    while (c < NCOL - 1 && s_pre[c + 1] <= k) c++;
    return c;
}

struct FixScale {
    float m, p;            // float -> fixed point
    double inv_m, inv_p;   // fixed point -> value
};

// Column ranges of sorted points for columns [xa, xa + ncol), rows [ya, yb)
// into s_rb (first slot) and s_pre (exclusive prefix of lengths, ncol + 1).
__device__ __forceinline__ void column_ranges(const uint32_t *__restrict__ cs, const DevParams &P, int xa, int ncol,
                                              int ya, int yb, uint32_t *s_rb, uint32_t *s_pre) {
    ya = max(ya, 0);
    yb = min(yb, P.gy);
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
        const int x = xa + c;
        uint32_t b = 0, len = 0;
        if (x >= 0 && x < P.gx && ya < yb) {
            const size_t base = (size_t)x * P.gy;
            b = cs[base + ya];
            len = cs[base + yb] - b;
        }
        s_rb[c] = b;
        s_pre[c] = len;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int c = 0; c < ncol; c++) {
            const uint32_t l = s_pre[c];
            s_pre[c] = acc;
            acc += l;
        }
        s_pre[ncol] = acc;
    }
    __syncthreads();
}

// Column of flattened item k: the last column whose prefix is <= k.
template <int NCOL>
__device__ __forceinline__ int flat_col(uint32_t k, const uint32_t *s_pre) {
    int lo = 0, hi = NCOL - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pre[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Slot of flattened item k: the last column whose prefix is <= k.
template <int NCOL>
__device__ __forceinline__ uint32_t flat_slot(uint32_t k, const uint32_t *s_rb, const uint32_t *s_pre) {
    int lo = 0, hi = NCOL - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pre[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return s_rb[lo] + (k - s_pre[lo]);
}

template <int TW, int TH, bool DET>
__global__ __launch_bounds__(kTileThreads) void k_p2g_tiled(DevParams P, Lines L, const uint32_t *__restrict__ cs,
                                                            float *__restrict__ gm, float *__restrict__ gpx,
                                                            float *__restrict__ gpy, Counters *cnt, int ntx, int nty,
                                                            FixScale fs) {
    constexpr int NCOL = TW + 4;
    constexpr int NN = TW * TH;
    using Acc = typename std::conditional<DET, unsigned long long, float>::type;
    __shared__ Acc s_acc[3 * NN];
    __shared__ uint32_t s_rb[NCOL];
    __shared__ uint32_t s_pre[NCOL + 1];

    const int t = xcd_remap(blockIdx.x, ntx * nty);
    const int tx = t / nty, ty = t - tx * nty;
    const int x0 = tx * TW, y0 = ty * TH;
    for (int k = threadIdx.x; k < 3 * NN; k += kTileThreads) s_acc[k] = 0;
    column_ranges(cs, P, x0 - 2, NCOL, y0 - 2, y0 + TH + 2, s_rb, s_pre);

    const uint32_t total = s_pre[NCOL];
    uint32_t err = 0;
    for (uint32_t k = threadIdx.x; k < total; k += kTileThreads) {
        const uint32_t i = flat_slot<NCOL>(k, s_rb, s_pre);
        const uint32_t u = L.u(ICE_IDX_UTILITY, i);
        if (u & (kDisabled | kEscaped)) continue;
        const uint32_t cell = L.u(ICE_IDX_CELL, i);
        const int cx = (int)(cell & 0xffff), cy = (int)(cell >> 16);
        if (cx + 1 < x0 || cx - 1 >= x0 + TW || cy + 1 < y0 || cy - 1 >= y0 + TH) continue;
        P2GPoint q;
        load_p2g_point(L, i, cell, P, q);
        float wx[3], wy[3];
        weights(q.x, wx);
        weights(q.y, wy);
#pragma unroll
        for (int a = -1; a <= 1; a++) {
            const int lx = cx + a - x0;
#pragma unroll
            for (int b = -1; b <= 1; b++) {
                const int ly = cy + b - y0;
                float dm, dpx, dpy;
                p2g_contrib(q, a, b, wx[a + 1] * wy[b + 1], dm, dpx, dpy);
                if (__builtin_isnan(dpx) || __builtin_isnan(dpy)) err |= ICE_ERR_GRID_P2G_NAN_VEL;
                if (__builtin_isnan(dm)) err |= ICE_ERR_GRID_P2G_NAN_MASS;
                if ((unsigned)lx < (unsigned)TW && (unsigned)ly < (unsigned)TH) {
                    const int o = lx * TH + ly;
                    if constexpr (DET) {
                        atomicAdd(&s_acc[o], (unsigned long long)(long long)(dm * fs.m));
                        atomicAdd(&s_acc[NN + o], (unsigned long long)(long long)(dpx * fs.p));
                        atomicAdd(&s_acc[2 * NN + o], (unsigned long long)(long long)(dpy * fs.p));
                    } else {
                        __hip_atomic_fetch_add(&s_acc[o], dm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        __hip_atomic_fetch_add(&s_acc[NN + o], dpx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        __hip_atomic_fetch_add(&s_acc[2 * NN + o], dpy, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                }
            }
        }
        if (cx < 1 || cy < 1 || cx > P.gx - 2 || cy > P.gy - 2) err |= ICE_ERR_POINT_LEFT_AREA;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NN; k += kTileThreads) {
        const int lx = k / TH, ly = k - (k / TH) * TH;
        const int x = x0 + lx, y = y0 + ly;
        if (x < P.gx && y < P.gy) {
            const size_t node = (size_t)x * P.gy + y;
            if constexpr (DET) {
                gm[node] = (float)((double)(long long)s_acc[k] * fs.inv_m);
                gpx[node] = (float)((double)(long long)s_acc[NN + k] * fs.inv_p);
                gpy[node] = (float)((double)(long long)s_acc[2 * NN + k] * fs.inv_p);
            } else {
                gm[node] = s_acc[k];
                gpx[node] = s_acc[NN + k];
                gpy[node] = s_acc[2 * NN + k];
            }
        }
    }
    flush_err(cnt, err);
}

struct GridIn {
    const float *m, *px, *py, *cur_x, *cur_y, *wind_x, *wind_y;
    const uint8_t *status;
};

template <int TW, int TH>
__global__ __launch_bounds__(kTileThreads) void k_g2p_tiled(DevParams P, StepArg S, Lines L,
                                                            const uint32_t *__restrict__ cs, GridIn G,
                                                            float *__restrict__ gfx, float *__restrict__ gfy,
                                                            Counters *cnt, uint32_t *__restrict__ offtile, int ntx,
                                                            int nty, int band) {
    constexpr int RW = TW + 4, RH = TH + 4;
    __shared__ float2 s_v[RW * RH];
    __shared__ uint32_t s_rb[TW];
    __shared__ uint32_t s_pre[TW + 1];

    int tx, ty;
    tile_coords(blockIdx.x, ntx, nty, band, tx, ty);
    const int x0 = tx * TW, y0 = ty * TH;
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
                if (wall) { gfx[node] += mx; gfy[node] += my; }
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
        if (rx < 1 || rx > RW - 2 || ry < 1 || ry > RH - 2) {
            err |= ICE_ERR_POINT_LEFT_AREA;   // cannot happen while the drift invariant holds
            continue;
        }
        float wx[3], wy[3];
        weights(s.x, wx);
        weights(s.y, wy);
        float pvx = 0.f, pvy = 0.f;
        M2 C{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int a = -1; a <= 1; a++)
#pragma unroll
            for (int b = -1; b <= 1; b++) {
                const float w = wx[a + 1] * wy[b + 1];
                const float2 nv = s_v[(rx + a) * RH + (ry + b)];
                const float dx = (float)a - s.x, dy = (float)b - s.y;
                pvx += w * nv.x;
                pvy += w * nv.y;
                const float f = P.four_h_inv * w;
                C.a00 += f * (nv.x * dx);
                C.a01 += f * (nv.x * dy);
                C.a10 += f * (nv.y * dx);
                C.a11 += f * (nv.y * dy);
            }
        g2p_finish(P, s, pvx, pvy, C, th, dpt, err);
        if (s.newly_disabled) atomicAdd(&cnt->disabled, 1);
        if (s.cell_updated && !(s.u & kDisabled)) {
            if (s.u & kMoved) {
                s.u |= kEscaped;
                const uint32_t j = atomicAdd(&cnt->n_esc, 1u);
                offtile[j] = i;
            } else {
                s.u |= kMoved;
            }
        }
        g2p_store(L, i, s, pvx, pvy, C, rec);
    }
    flush_err(cnt, err);
}

// P2G of the off-tile points: global atomics on top of the tile stores.
__global__ __launch_bounds__(256) void k_p2g_offtile(DevParams P, Lines L, const uint32_t *__restrict__ offtile,
                                                     float *__restrict__ gm, float *__restrict__ gpx,
                                                     float *__restrict__ gpy, Counters *cnt) {
    const uint32_t n = cnt->n_esc;
    uint32_t err = 0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t i = offtile[j];
        const uint32_t u = L.u(ICE_IDX_UTILITY, i);
        if (u & kDisabled) continue;
        P2GPoint q;
        load_p2g_point(L, i, L.u(ICE_IDX_CELL, i), P, q);
        if (!stencil_inside(q.cx, q.cy, P)) {
            err |= ICE_ERR_POINT_LEFT_AREA;
            continue;
        }
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
    flush_err(cnt, err);
}

// G2P of the off-tile points, node velocities recomputed from the P2G output
// (the same function the tiles evaluate).  Runs before k_g2p_tiled so points
// that tile G2P appends to the list are not processed twice.
__global__ __launch_bounds__(256) void k_g2p_offtile(DevParams P, StepArg S, Lines L,
                                                     const uint32_t *__restrict__ offtile, GridIn G, Counters *cnt) {
    const uint32_t n = cnt->n_esc;
    double time;
    bool rec;
    resolve_step(S, P, time, rec);
    const float ramp = current_ramp(time);
    uint32_t err = 0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t i = offtile[j];
        const uint32_t u = L.u(ICE_IDX_UTILITY, i);
        if (u & kDisabled) continue;
        G2PState s;
        load_g2p_state(L, i, u, s);
        if (!stencil_inside(s.cx, s.cy, P)) {
            err |= ICE_ERR_POINT_LEFT_AREA;
            continue;
        }
        const float th = L(ICE_IDX_THICKNESS, i);
        float wx[3], wy[3];
        weights(s.x, wx);
        weights(s.y, wy);
        float pvx = 0.f, pvy = 0.f;
        M2 C{0.f, 0.f, 0.f, 0.f};
        for (int a = -1; a <= 1; a++)
            for (int b = -1; b <= 1; b++) {
                const float w = wx[a + 1] * wy[b + 1];
                const size_t node = (size_t)(s.cy + b) + (size_t)(s.cx + a) * P.gy;
                float cx = 0.f, cy = 0.f;
                if (P.has_current) { cx = G.cur_x[node]; cy = G.cur_y[node]; }
                float nvx, nvy;
                bool wall;
                node_velocity(G.m[node], G.px[node], G.py[node], G.status[node], cx, cy, G.wind_x[node],
                              G.wind_y[node], ramp, P, nvx, nvy, wall);
                const float dx = (float)a - s.x, dy = (float)b - s.y;
                pvx += w * nvx;
                pvy += w * nvy;
                const float f = P.four_h_inv * w;
                C.a00 += f * (nvx * dx);
                C.a01 += f * (nvx * dy);
                C.a10 += f * (nvy * dx);
                C.a11 += f * (nvy * dy);
            }
        g2p_finish(P, s, pvx, pvy, C, th, dp_threshold(P, L, i), err);
        if (s.newly_disabled) atomicAdd(&cnt->disabled, 1);
        g2p_store(L, i, s, pvx, pvy, C, rec);
    }
    flush_err(cnt, err);
}

// Node velocities of the last substep into (vx, vy), for inspection.
__global__ __launch_bounds__(256) void k_node_velocity_out(DevParams P, double t, GridIn G, float *__restrict__ vx,
                                                           float *__restrict__ vy) {
    const size_t N = (size_t)P.gx * P.gy;
    const float ramp = current_ramp(t);
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < N; k += (size_t)gridDim.x * blockDim.x) {
        float cx = 0.f, cy = 0.f;
        if (P.has_current) { cx = G.cur_x[k]; cy = G.cur_y[k]; }
        bool wall;
        float a, b;
        node_velocity(G.m[k], G.px[k], G.py[k], G.status[k], cx, cy, G.wind_x[k], G.wind_y[k], ramp, P, a, b, wall);
        vx[k] = a;
        vy[k] = b;
    }
}

}  // namespace ice
