// tile_common.hpp — what the tile kernels share: domain and tile tables,
// the XCD-aware workgroup -> tile order, per-column point ranges.
//
// A handle holds one or more independent domains.  All domains' points live
// in one SoA (each domain a contiguous slot range), all their grids in one set
// of grid lines (each domain a contiguous node range), all their cell-start
// arrays in one buffer.  One launch covers every tile of every domain: the
// host builds a tile table (largest domains first, band order inside a
// domain), and workgroup b takes entry xcd_remap(b), so each XCD walks one
// contiguous run of the table.
//
// Sorted-path invariant (kept by the device sort and the kMoved/kEscaped
// bits): each domain's points are in cell order as of its last sort, cs[k] is
// the first sorted slot of cell k (k = x*GridY + y, slots relative to the
// domain), and every point not on the domain's off-tile list sits at most one
// cell (Chebyshev) from the cell it was sorted into.
#pragma once
#include "point_ops.hpp"

namespace ice {

constexpr int kTileThreads = 256;

struct DomainDesc {
    int gx, gy;            // nodes
    int has_current;
    uint32_t off_base;     // first slot of the domain's off-tile list
    uint64_t pt_off;       // first point slot of the domain from the point buffer's base (its half included)
    uint64_t node_base;    // first node of the domain in the grid lines
    uint64_t cs_base;      // first entry of the domain's cell-start array
    int ntx, nty;          // tiles per axis
    uint64_t frame_base;   // first frame of the domain (frames in tile-grid order, ty fastest)
};

// The end of a domain's sort, in stream order: its point slots moved to the
// other half of the point buffer (the table entry, when one is given), its
// off-tile list emptied, and its disabled count cleared when the sort
// dropped the disabled points.
__global__ void k_sort_done(DomainDesc *d, uint64_t pt_off, Counters *c, int clear_disabled) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (d) d->pt_off = pt_off;
        c->n_esc = 0;
        if (clear_disabled) c->disabled = 0;
    }
}

// The off-tile points of every domain as one item space: pre holds the
// exclusive prefix of the domains' list lengths (pre[ndom] = total), or is
// null for a single domain.  The length of the space, and the domain of item
// j: the last d with pre[d] <= j (domains with empty lists are skipped).
__device__ __forceinline__ uint32_t offtile_total(const uint32_t *pre, int ndom, const Counters *cnt0) {
    return pre ? pre[ndom] : cnt0[0].n_esc;
}
__device__ __forceinline__ int offtile_domain(const uint32_t *pre, int ndom, uint32_t j) {
    int d = 0;
    if (pre) {
        int hi = ndom - 1;
        while (d < hi) {
            const int mid = (d + hi + 1) >> 1;
            if (pre[mid] <= j) d = mid;
            else hi = mid - 1;
        }
    }
    return d;
}

struct TileRef {
    uint16_t d, tx, ty, pad;
};

struct FixScale {
    float m, p;            // float -> fixed point
    double inv_m, inv_p;   // fixed point -> value
};

struct GridLines {
    float *m, *px, *py, *cur_x, *cur_y, *fx, *fy, *wind_x, *wind_y;
    uint8_t *status;
};

// Workgroup b runs on XCD b % 8; give each XCD one contiguous run of the tile
// table so neighbouring tiles, which share halo points and nodes, meet in one L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    const int nx = 8;
    const int xcd = bid % nx, j = bid / nx;
    const int per = nb / nx, rem = nb % nx;
    return xcd * per + min(xcd, rem) + j;
}

// The tile a workgroup works on, with every pointer moved to its domain.
struct TileCtx {
    int d, x0, y0;
    DevParams P;
    Lines L;
    const uint32_t *cs;
    GridLines G;
    Counters *cnt;
    uint32_t *offtile;
};

template <int TW, int TH>
__device__ __forceinline__ TileCtx tile_ctx(const DevParams &P0, const DomainDesc *__restrict__ D,
                                            const TileRef *__restrict__ T, int ntiles, const Lines &L0,
                                            const uint32_t *cs0, const GridLines &G0, Counters *cnt0,
                                            uint32_t *off0) {
    const TileRef tr = T[xcd_remap(blockIdx.x, ntiles)];
    const DomainDesc dd = D[tr.d];
    TileCtx c;
    c.d = tr.d;
    c.x0 = tr.tx * TW;
    c.y0 = tr.ty * TH;
    c.P = P0;
    c.P.gx = dd.gx;
    c.P.gy = dd.gy;
    c.P.has_current = dd.has_current;
    c.L = Lines{L0.base + dd.pt_off, L0.pitch};
    c.cs = cs0 + dd.cs_base;
    const size_t nb = dd.node_base;
    c.G = GridLines{G0.m + nb, G0.px + nb, G0.py + nb, G0.cur_x + nb, G0.cur_y + nb, G0.fx + nb, G0.fy + nb,
                    G0.wind_x + nb, G0.wind_y + nb, G0.status + nb};
    c.cnt = cnt0 + tr.d;
    c.offtile = off0 + dd.off_base;
    return c;
}

// Column ranges of sorted points for columns [xa, xa + ncol), rows [ya, yb)
// into s_rb (first slot) and s_pre (exclusive prefix of lengths, ncol + 1).
// Contains barriers: call from every thread.
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

// The same, walking forward from a wave-uniform start column cb.
template <int NCOL>
__device__ __forceinline__ int col_from(uint32_t k, int cb, const uint32_t *s_pre) {
    int c = cb;
    while (c < NCOL - 1 && s_pre[c + 1] <= k) c++;
    return c;
}

// Runtime-count versions of flat_col / col_from (ncol >= 1 columns).
__device__ __forceinline__ int flat_col_n(uint32_t k, const uint32_t *s_pre, int ncol) {
    int lo = 0, hi = ncol - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pre[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int col_from_n(uint32_t k, int cb, const uint32_t *s_pre, int ncol) {
    int c = cb;
    while (c < ncol - 1 && s_pre[c + 1] <= k) c++;
    return c;
}

}  // namespace ice
