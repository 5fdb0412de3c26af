// kernels_p2g_cells.hpp — pull P2G with one lane per cell.
//
// Same tile contract as k_p2g_tiled (tile T of TW x TH nodes, points sorted
// into cells T +- 2 are pulled, every node of T is written once by its tile),
// but lanes map to cells, not points: a lane walks the points sorted into its
// cell, sums their 3x3 stencil contributions in registers, and issues 27 LDS
// atomics per cell.  Consecutive lanes own consecutive cells, so one atomic
// instruction never has two lanes on the same node, and there are ~4x fewer
// atomics than with one lane per point (4 points per cell).
//
// Points that changed cell since the sort (kMoved) may contribute around a
// cell next to their lane's; they add their own stencil directly.  In the
// deterministic mode the off-tile points are also added here (int64 sums do
// not depend on order), instead of by k_p2g_offtile's float atomics.
#pragma once
#include "kernels_tiled.hpp"

namespace ice {

template <bool DET, int NN>
__device__ __forceinline__ void lds_add3(typename std::conditional<DET, unsigned long long, float>::type *acc, int o,
                                         float dm, float dpx, float dpy, const FixScale &fs) {
    if constexpr (DET) {
        atomicAdd(&acc[o], (unsigned long long)(long long)(dm * fs.m));
        atomicAdd(&acc[NN + o], (unsigned long long)(long long)(dpx * fs.p));
        atomicAdd(&acc[2 * NN + o], (unsigned long long)(long long)(dpy * fs.p));
    } else {
        __hip_atomic_fetch_add(&acc[o], dm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&acc[NN + o], dpx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&acc[2 * NN + o], dpy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// Add one point's whole stencil (cell (cx, cy)) to the tile's nodes.
template <int TW, int TH, int LP, bool DET>
__device__ __forceinline__ void scatter_point(const P2GPoint &q, int x0, int y0,
                                              typename std::conditional<DET, unsigned long long, float>::type *acc,
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
                lds_add3<DET, NN>(acc, lx * LP + ly, dm, dpx, dpy, fs);
        }
}

template <int TW, int TH, bool DET>
__global__ __launch_bounds__(kTileThreads) void k_p2g_cells(DevParams P, Lines L, const uint32_t *__restrict__ cs,
                                                            float *__restrict__ gm, float *__restrict__ gpx,
                                                            float *__restrict__ gpy, Counters *cnt, int ntx, int nty,
                                                            FixScale fs, const uint32_t *__restrict__ offtile) {
    constexpr int RW = TW + 4, RH = TH + 4;   // cells pulled
    constexpr int LP = TH + 1;                // padded LDS column pitch (no 2-way bank conflicts)
    constexpr int NN = TW * LP;
    using Acc = typename std::conditional<DET, unsigned long long, float>::type;
    __shared__ Acc s_acc[3 * NN];

    const int t = xcd_remap(blockIdx.x, ntx * nty);
    const int tx = t / nty, ty = t - tx * nty;
    const int x0 = tx * TW, y0 = ty * TH;
    for (int k = threadIdx.x; k < 3 * NN; k += kTileThreads) s_acc[k] = 0;
    __syncthreads();

    uint32_t err = 0;
    for (int k = threadIdx.x; k < RW * RH; k += kTileThreads) {
        const int c = k / RH, r = k - (k / RH) * RH;
        const int x = x0 - 2 + c, y = y0 - 2 + r;
        if (x < 0 || x >= P.gx || y < 0 || y >= P.gy) continue;
        const bool ring2 = (c == 0 || c == RW - 1 || r == 0 || r == RH - 1);
        const size_t cell = (size_t)x * P.gy + y;
        const uint32_t b = cs[cell], e = cs[cell + 1];
        float am[3][3], ax[3][3], ay[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) am[i][j] = ax[i][j] = ay[i][j] = 0.f;
        bool any = false;
        const bool inside = x >= 1 && y >= 1 && x <= P.gx - 2 && y <= P.gy - 2;
        for (uint32_t i = b; i < e; i++) {
            const uint32_t u = L.u(ICE_IDX_UTILITY, i);
            if (u & (kDisabled | kEscaped)) continue;
            if (u & kMoved) {
                const uint32_t cw = L.u(ICE_IDX_CELL, i);
                P2GPoint q;
                load_p2g_point(L, i, cw, P, q);
                if (!stencil_inside(q.cx, q.cy, P)) { err |= ICE_ERR_POINT_LEFT_AREA; continue; }
                scatter_point<TW, TH, LP, DET>(q, x0, y0, s_acc, fs, err);
                continue;
            }
            if (ring2) continue;   // an unmoved point two cells out cannot reach T
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
                    am[a][bb] += dm;
                    ax[a][bb] += dpx;
                    ay[a][bb] += dpy;
                }
            any = true;
        }
        if (any) {
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int bb = 0; bb < 3; bb++) {
                    if (__builtin_isnan(ax[a][bb]) || __builtin_isnan(ay[a][bb])) err |= ICE_ERR_GRID_P2G_NAN_VEL;
                    if (__builtin_isnan(am[a][bb])) err |= ICE_ERR_GRID_P2G_NAN_MASS;
                    const int lx = x + a - 1 - x0, ly = y + bb - 1 - y0;
                    if ((unsigned)lx < (unsigned)TW && (unsigned)ly < (unsigned)TH)
                        lds_add3<DET, NN>(s_acc, lx * LP + ly, am[a][bb], ax[a][bb], ay[a][bb], fs);
                }
        }
    }
    if constexpr (DET) {
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
            scatter_point<TW, TH, LP, DET>(q, x0, y0, s_acc, fs, err);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < TW * TH; k += kTileThreads) {
        const int lx = k / TH, ly = k - (k / TH) * TH;
        const int x = x0 + lx, y = y0 + ly;
        const int o = lx * LP + ly;
        if (x < P.gx && y < P.gy) {
            const size_t node = (size_t)x * P.gy + y;
            if constexpr (DET) {
                gm[node] = (float)((double)(long long)s_acc[o] * fs.inv_m);
                gpx[node] = (float)((double)(long long)s_acc[NN + o] * fs.inv_p);
                gpy[node] = (float)((double)(long long)s_acc[2 * NN + o] * fs.inv_p);
            } else {
                gm[node] = s_acc[o];
                gpx[node] = s_acc[NN + o];
                gpy[node] = s_acc[2 * NN + o];
            }
        }
    }
    flush_err(cnt, err);
}

}  // namespace ice
