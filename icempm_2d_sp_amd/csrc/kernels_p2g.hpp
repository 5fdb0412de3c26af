// kernels_p2g.hpp — the deterministic mode's P2G: one workgroup per tile of
// TW x TH nodes, pulling the points sorted into cells T +- 2, one lane per
// cell (simulation/kernels.cuh:39-104).
//
// Every contribution is converted to int64 fixed point on its own and the
// sums are integer (independent of order and of how points are grouped into
// cells): a lane walks the points sorted into its cell and sums their 3x3
// stencil contributions in registers (unmoved points need no cell-word load),
// then adds the stencil to the tile's nodes with LDS integer atomics; points
// that changed cell since the sort add their own stencil, and the tile kernel
// also adds the domain's off-tile points, so results are bit-identical for
// any schedule and any sort history.  Every node of the tile is then stored
// once: no global atomics and no grid clear.  (The float P2G of the other
// sorted paths is the tile pass's: points loaded one per lane and their terms
// staged in LDS for the cell lanes, kernels_fused.hpp.)
#pragma once
#include "tile_common.hpp"

namespace ice {

// One contribution in fixed point.  A NaN contribution (flagged by the
// caller) is dropped: integer sums cannot carry it.
__device__ __forceinline__ long long to_fix(float v, float scale) {
    return __builtin_isnan(v) ? 0ll : (long long)(v * scale);
}

template <int NN>
__device__ __forceinline__ void lds_add3(unsigned long long *acc, int o, float dm, float dpx, float dpy,
                                         const FixScale &fs) {
    atomicAdd(&acc[o], (unsigned long long)to_fix(dm, fs.m));
    atomicAdd(&acc[NN + o], (unsigned long long)to_fix(dpx, fs.p));
    atomicAdd(&acc[2 * NN + o], (unsigned long long)to_fix(dpy, fs.p));
}

// Add one point's whole stencil (cell (cx, cy)) to the tile's nodes.
template <int TW, int TH, int LP>
__device__ __forceinline__ void scatter_point(const P2GPoint &q, int x0, int y0, unsigned long long *acc,
                                              const FixScale &fs, uint32_t &err) {
    constexpr int NN = TW * LP;
    float wx[3], wy[3];
    weights(q.x, wx);
    weights(q.y, wy);
#pragma unroll
    for (int a = -1; a <= 1; a++)
#pragma unroll
        for (int b = -1; b <= 1; b++) {
            float dm, dpx, dpy;
            p2g_contrib(q, a, b, wx[a + 1] * wy[b + 1], dm, dpx, dpy);
            if (__builtin_isnan(dpx) || __builtin_isnan(dpy)) err |= ICE_ERR_GRID_P2G_NAN_VEL;
            if (__builtin_isnan(dm)) err |= ICE_ERR_GRID_P2G_NAN_MASS;
            const int lx = q.cx + a - x0, ly = q.cy + b - y0;
            if ((unsigned)lx < (unsigned)TW && (unsigned)ly < (unsigned)TH)
                lds_add3<NN>(acc, lx * LP + ly, dm, dpx, dpy, fs);
        }
}

// Sum the unmoved points of cell (x, y) into the register stencil, each
// contribution in fixed point; returns whether any point contributed and sets
// *moved when the cell also holds points that changed cell since the sort.
__device__ __forceinline__ bool cell_stencil(const Lines &L, const DevParams &P, uint32_t b, uint32_t e, int x, int y,
                                             bool ring2, long long am[3][3], long long ax[3][3], long long ay[3][3],
                                             bool &moved, uint32_t &err, const FixScale &fs) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) am[i][j] = ax[i][j] = ay[i][j] = 0;
    bool any = false;
    moved = false;
    const bool inside = x >= 1 && y >= 1 && x <= P.gx - 2 && y <= P.gy - 2;
    for (uint32_t i = b; i < e; i++) {
        const uint32_t u = L.u(ICE_IDX_UTILITY, i);
        if (u & (kDisabled | kEscaped)) continue;
        if (u & kMoved) { moved = true; continue; }
        if (ring2) continue;   // an unmoved point two cells out cannot reach the tile
        if (!inside) { err |= ICE_ERR_POINT_LEFT_AREA; continue; }
        P2GPoint q;
        load_p2g_point(L, i, ((uint32_t)y << 16) | (uint32_t)x, P, q);
        float wx[3], wy[3];
        weights(q.x, wx);
        weights(q.y, wy);
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int bb = 0; bb < 3; bb++) {
                float dm, dpx, dpy;
                p2g_contrib(q, a - 1, bb - 1, wx[a] * wy[bb], dm, dpx, dpy);
                // the conversion scatter_point's contributions take
                if (__builtin_isnan(dpx) || __builtin_isnan(dpy)) err |= ICE_ERR_GRID_P2G_NAN_VEL;
                if (__builtin_isnan(dm)) err |= ICE_ERR_GRID_P2G_NAN_MASS;
                am[a][bb] += to_fix(dm, fs.m);
                ax[a][bb] += to_fix(dpx, fs.p);
                ay[a][bb] += to_fix(dpy, fs.p);
            }
        any = true;
    }
    return any;
}

// Moved points of cell [b, e): their own stencil, added with LDS atomics.
template <int TW, int TH, int LP>
__device__ __forceinline__ void scatter_moved(const Lines &L, const DevParams &P, uint32_t b, uint32_t e, int x0, int y0,
                                              unsigned long long *acc, const FixScale &fs, uint32_t &err) {
    for (uint32_t i = b; i < e; i++) {
        const uint32_t u = L.u(ICE_IDX_UTILITY, i);
        if ((u & (kDisabled | kEscaped)) || !(u & kMoved)) continue;
        const uint32_t cw = L.u(ICE_IDX_CELL, i);
        P2GPoint q;
        load_p2g_point(L, i, cw, P, q);
        if (!stencil_inside(q.cx, q.cy, P)) { err |= ICE_ERR_POINT_LEFT_AREA; continue; }
        scatter_point<TW, TH, LP>(q, x0, y0, acc, fs, err);
    }
}

template <int TW, int TH>
__global__ __launch_bounds__(kTileThreads) void k_p2g_tiles_det(DevParams P0, const DomainDesc *__restrict__ D,
                                                            const TileRef *__restrict__ T, int ntiles, Lines L0,
                                                            const uint32_t *__restrict__ cs0, GridLines G0,
                                                            Counters *cnt0, uint32_t *off0, FixScale fs) {
    const TileCtx tc = tile_ctx<TW, TH>(P0, D, T, ntiles, L0, cs0, G0, cnt0, off0);
    const DevParams &P = tc.P;
    const Lines &L = tc.L;
    const uint32_t *__restrict__ cs = tc.cs;
    Counters *cnt = tc.cnt;
    const uint32_t *__restrict__ offtile = tc.offtile;
    const int x0 = tc.x0, y0 = tc.y0;
    constexpr int RW = TW + 4, RH = TH + 4;   // cells pulled
    constexpr int NCELL = RW * RH;
    constexpr int LP = TH + 1;                // padded LDS column pitch
    constexpr int NN = TW * LP;
    __shared__ unsigned long long s_acc[3 * NN];

    for (int k = threadIdx.x; k < 3 * NN; k += kTileThreads) s_acc[k] = 0;
    __syncthreads();

    uint32_t err = 0;
    // Rounds of kTileThreads cells; the loop bound is uniform so every thread
    // reaches every barrier.
    for (int base = 0; base < NCELL; base += kTileThreads) {
        const int k = base + (int)threadIdx.x;
        const int c = k / RH, r = k - (k / RH) * RH;
        const int x = x0 - 2 + c, y = y0 - 2 + r;
        const bool valid = k < NCELL && x >= 0 && x < P.gx && y >= 0 && y < P.gy;
        uint32_t b = 0, e = 0;
        if (valid) {
            const size_t cell = (size_t)x * P.gy + y;
            b = cs[cell];
            e = cs[cell + 1];
        }
        const bool ring2 = (c == 0 || c == RW - 1 || r == 0 || r == RH - 1);
        long long am[3][3], ax[3][3], ay[3][3];
        bool moved = false;
        const bool any = valid && cell_stencil(L, P, b, e, x, y, ring2, am, ax, ay, moved, err, fs);
        if (any) {
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int bb = 0; bb < 3; bb++) {
                    const int lx = x + a - 1 - x0, ly = y + bb - 1 - y0;
                    if ((unsigned)lx < (unsigned)TW && (unsigned)ly < (unsigned)TH) {
                        const int o = lx * LP + ly;
                        atomicAdd(&s_acc[o], (unsigned long long)am[a][bb]);
                        atomicAdd(&s_acc[NN + o], (unsigned long long)ax[a][bb]);
                        atomicAdd(&s_acc[2 * NN + o], (unsigned long long)ay[a][bb]);
                    }
                }
        }
        if (moved) scatter_moved<TW, TH, LP>(L, P, b, e, x0, y0, s_acc, fs, err);
    }
    {
        // off-tile points, order-independent in fixed point
        const uint32_t n_off = cnt->n_esc;
        for (uint32_t j = threadIdx.x; j < n_off; j += kTileThreads) {
            const uint32_t i = offtile[j];
            const uint32_t u = L.u(ICE_IDX_UTILITY, i);
            if (u & kDisabled) continue;
            const uint32_t cw = L.u(ICE_IDX_CELL, i);
            const int cx = (int)(cw & 0xffff), cy = (int)(cw >> 16);
            if (cx + 1 < x0 || cx - 1 >= x0 + TW || cy + 1 < y0 || cy - 1 >= y0 + TH) continue;
            P2GPoint q;
            load_p2g_point(L, i, cw, P, q);
            if (!stencil_inside(q.cx, q.cy, P)) { err |= ICE_ERR_POINT_LEFT_AREA; continue; }
            scatter_point<TW, TH, LP>(q, x0, y0, s_acc, fs, err);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TW * TH; k += kTileThreads) {
        const int lx = k / TH, ly = k - (k / TH) * TH;
        const int x = x0 + lx, y = y0 + ly;
        const int o = lx * LP + ly;
        if (x < P.gx && y < P.gy) {
            const size_t node = (size_t)x * P.gy + y;
            tc.G.m[node] = (float)((double)(long long)s_acc[o] * fs.inv_m);
            tc.G.px[node] = (float)((double)(long long)s_acc[NN + o] * fs.inv_p);
            tc.G.py[node] = (float)((double)(long long)s_acc[2 * NN + o] * fs.inv_p);
        }
    }
    flush_err(cnt, err);
}

}  // namespace ice
