// kernels_sort.hpp — device cell-binned sort with the order of std::stable_sort.
//
// HostSideSOA::RemoveDisabledAndSort (simulation/hssoa/host_side_soa.cpp:36-47)
// removes disabled points and sorts the rest by ProxyPoint::getCellIndex =
// x*GridY + y (proxypoint2d.cpp:117-123).  Here, per sort:
//   1. key per point (disabled points go to bucket NC = GridX*GridY) and a slot
//      inside its cell from the cell's counter (one atomic per run of equal
//      keys in a wave),
//   2. exclusive scans: cell counts -> cs (cell start), disabled flags -> rank,
//   3. scatter slot indices, then sort each cell's short index run ascending,
//      which is exactly the stable order (ties keep input order),
//   4. gather every SoA line into the domain's range of the other half of the
//      point buffer, in the new order.
// Disabled points either follow the sorted ones (in input order) or are
// dropped, as the caller asks.  The re-sorts of many domains between run
// chunks go through the same steps as one set of launches (the *_many
// kernels): each launch flattens the domains' blocks, and a block finds its
// domain ("job") in a prefix of the jobs' block counts.
#pragma once
#include "point_ops.hpp"
#include "tile_common.hpp"

namespace ice {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *s_wsum, uint32_t &total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wsum[wid] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int w = 0; w < kScanThreads / 64; w++) {
        const uint32_t ws = s_wsum[w];
        if (w < wid) pre += ws;
        tot += ws;
    }
    __syncthreads();
    total = tot;
    return pre + x - v;
}

// One exclusive scan of a three-kernel pass: in[0..n) -> out[0..n], out[n] =
// total; bsum holds nb + 1 block sums.  A pass runs one or two scans (grid
// row blockIdx.y picks the job; workgroups past a job's blocks leave at once).
struct ScanJob {
    const uint32_t *in;
    size_t n;
    uint32_t *bsum;
    uint32_t nb;
    uint32_t *out;
};
struct ScanJobs {
    ScanJob j[2];
};

// The last job j with blk[j] <= b: the job of flattened block b (jobs with
// no blocks share their start with the next job and are never chosen).
__device__ __forceinline__ int job_of(const uint32_t *__restrict__ blk, int njobs, uint32_t b) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void scan_reduce_block(const ScanJob &jb, uint32_t b, uint32_t *s_w) {
    const uint32_t *__restrict__ in = jb.in;
    const size_t n = jb.n;
    const size_t base = (size_t)b * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) s += in[base + k];
    uint32_t tot;
    block_exclusive_scan(s, s_w, tot);
    if (threadIdx.x == 0) jb.bsum[b] = tot;
}

__device__ __forceinline__ void scan_blocks_job(const ScanJob &jb, uint32_t *s_w) {
    uint32_t *__restrict__ bsum = jb.bsum;
    const uint32_t nb = jb.nb;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += kScanTile) {
        uint32_t v[kScanItems];
        uint32_t s = 0;
        for (int k = 0; k < kScanItems; k++) {
            const uint32_t idx = base + threadIdx.x * kScanItems + k;
            v[k] = idx < nb ? bsum[idx] : 0u;
            s += v[k];
        }
        uint32_t tot;
        uint32_t pre = block_exclusive_scan(s, s_w, tot) + carry;
        for (int k = 0; k < kScanItems; k++) {
            const uint32_t idx = base + threadIdx.x * kScanItems + k;
            if (idx < nb) bsum[idx] = pre;
            pre += v[k];
        }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

__device__ __forceinline__ void scan_apply_block(const ScanJob &jb, uint32_t b, uint32_t *s_w) {
    const uint32_t *__restrict__ in = jb.in;
    const size_t n = jb.n;
    const uint32_t *__restrict__ bsum = jb.bsum;
    uint32_t *__restrict__ out = jb.out;
    const size_t base = (size_t)b * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; k++) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t pre = block_exclusive_scan(s, s_w, tot) + bsum[b];
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (b == 0 && threadIdx.x == 0) out[n] = bsum[jb.nb];
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(ScanJobs J) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    const ScanJob &jb = J.j[blockIdx.y];
    if (blockIdx.x >= jb.nb) return;   // whole workgroup
    scan_reduce_block(jb, blockIdx.x, s_w);
}

// One workgroup per job: exclusive scan of bsum[0..nb) in place, bsum[nb] = total.
__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(ScanJobs J) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    scan_blocks_job(J.j[blockIdx.y], s_w);
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(ScanJobs J) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    const ScanJob &jb = J.j[blockIdx.y];
    if (blockIdx.x >= jb.nb) return;   // whole workgroup
    scan_apply_block(jb, blockIdx.x, s_w);
}

// The same three passes over many jobs: blk[j] = first flattened block of job j.
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce_many(const ScanJob *__restrict__ jobs,
                                                                   const uint32_t *__restrict__ blk, int njobs) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    const int j = job_of(blk, njobs, blockIdx.x);
    scan_reduce_block(jobs[j], blockIdx.x - blk[j], s_w);
}

__global__ __launch_bounds__(kScanThreads) void k_scan_blocks_many(const ScanJob *__restrict__ jobs) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    scan_blocks_job(jobs[blockIdx.x], s_w);
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply_many(const ScanJob *__restrict__ jobs,
                                                                  const uint32_t *__restrict__ blk, int njobs) {
    __shared__ uint32_t s_w[kScanThreads / 64];
    const int j = job_of(blk, njobs, blockIdx.x);
    scan_apply_block(jobs[j], blockIdx.x - blk[j], s_w);
}

// Points whose cell word lies outside the grid cannot be binned; they join
// the disabled tail (and raise the left-area error bit).  Points arrive
// nearly sorted, so neighbouring lanes mostly share a cell: each run of equal
// keys in a wave takes its slots with one atomic (the slot order inside a
// cell does not matter, k_sort_cells orders each cell by input index).
__device__ __forceinline__ void sort_keys_block(const Lines &L, uint32_t n, int gx, int gy, uint32_t NC,
                                                uint32_t *__restrict__ count, uint32_t *__restrict__ key,
                                                uint32_t *__restrict__ slot, uint32_t *__restrict__ dflag,
                                                Counters *cnt, uint32_t b) {
    const uint32_t i = b * blockDim.x + threadIdx.x;
    const int lane = (int)(threadIdx.x & 63);
    bool binned = false;
    uint32_t k = 0xffffffffu;   // never a cell key: lanes not binned start runs of their own
    if (i < n) {
        const uint32_t u = L.u(ICE_IDX_UTILITY, i);
        const uint32_t c = L.u(ICE_IDX_CELL, i);
        const bool outside = (int)(c & 0xffffu) >= gx || (int)(c >> 16) >= gy;
        if (outside && !(u & kDisabled)) atomicOr(&cnt->err, ICE_ERR_POINT_LEFT_AREA);
        binned = !(u & kDisabled) && !outside;
        if (binned) k = (c & 0xffffu) * (uint32_t)gy + (c >> 16);
    }
    const uint32_t prev = __shfl_up(k, 1, 64);
    const bool head = lane == 0 || prev != k;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);   // lanes 0..lane
    const int first = 63 - __clzll(heads & upto);                      // this run's head lane
    const uint64_t later = heads & ~upto;
    const int end = later ? __ffsll((unsigned long long)later) - 1 : 64;   // next run's head lane
    uint32_t base = 0;
    if (head && binned) base = atomicAdd(&count[k], (uint32_t)(end - lane));
    base = __shfl(base, first, 64);
    if (i < n) {
        key[i] = binned ? k : NC;
        slot[i] = binned ? base + (uint32_t)(lane - first) : 0u;
        dflag[i] = binned ? 0u : 1u;
    }
}

__global__ __launch_bounds__(256) void k_sort_keys(Lines L, uint32_t n, int gx, int gy, uint32_t NC,
                                                   uint32_t *__restrict__ count, uint32_t *__restrict__ key,
                                                   uint32_t *__restrict__ slot, uint32_t *__restrict__ dflag,
                                                   Counters *cnt) {
    sort_keys_block(L, n, gx, gy, NC, count, key, slot, dflag, cnt, blockIdx.x);
}

__device__ __forceinline__ void sort_scatter_item(uint32_t i, uint32_t n, uint32_t NC,
                                                  const uint32_t *__restrict__ key,
                                                  const uint32_t *__restrict__ slot, const uint32_t *__restrict__ cs,
                                                  const uint32_t *__restrict__ drank, uint32_t *__restrict__ order) {
    if (i >= n) return;
    const uint32_t k = key[i];
    const uint32_t pos = (k == NC) ? cs[NC] + drank[i] : cs[k] + slot[i];
    order[pos] = i;
}

__global__ __launch_bounds__(256) void k_sort_scatter(uint32_t n, uint32_t NC, const uint32_t *__restrict__ key,
                                                      const uint32_t *__restrict__ slot,
                                                      const uint32_t *__restrict__ cs,
                                                      const uint32_t *__restrict__ drank,
                                                      uint32_t *__restrict__ order) {
    sort_scatter_item(blockIdx.x * blockDim.x + threadIdx.x, n, NC, key, slot, cs, drank, order);
}

// Ascending input index inside each cell = the stable order.  Also clears
// the cell counters the key pass filled (the scan has consumed them), so the
// next sort starts from zeros without a fill.
__device__ __forceinline__ void sort_cells_item(size_t k, uint32_t NC, const uint32_t *__restrict__ cs,
                                                uint32_t *__restrict__ order, uint32_t *__restrict__ count) {
    if (k >= NC) return;
    count[k] = 0;
    const uint32_t b = cs[k], e = cs[k + 1];
    for (uint32_t p = b + 1; p < e; p++) {
        const uint32_t v = order[p];
        uint32_t q = p;
        while (q > b && order[q - 1] > v) {
            order[q] = order[q - 1];
            q--;
        }
        order[q] = v;
    }
}

__global__ __launch_bounds__(256) void k_sort_cells(uint32_t NC, const uint32_t *__restrict__ cs,
                                                    uint32_t *__restrict__ order, uint32_t *__restrict__ count) {
    sort_cells_item((size_t)blockIdx.x * blockDim.x + threadIdx.x, NC, cs, order, count);
}

__device__ __forceinline__ void sort_gather_item(uint32_t p, const Lines &src, const Lines &dst, int nlines,
                                                 uint32_t n_out, const uint32_t *__restrict__ order,
                                                 uint32_t *__restrict__ perm) {
    if (p >= n_out) return;
    const uint32_t s = order[p];
    dst.set_u(ICE_IDX_UTILITY, p, src.u(ICE_IDX_UTILITY, s) & ~kInternalBits);
    for (int l = 1; l < nlines; l++) dst(l, p) = src(l, s);
    if (perm) perm[p] = s;
}

__global__ __launch_bounds__(256) void k_sort_gather(Lines src, Lines dst, int nlines, uint32_t n_out,
                                                     const uint32_t *__restrict__ order, uint32_t *__restrict__ perm) {
    sort_gather_item(blockIdx.x * blockDim.x + threadIdx.x, src, dst, nlines, n_out, order, perm);
}

// One domain's sort in a batch of sorts (disabled points kept, behind the
// sorted ones).  Scratch arrays point at the domain's own ranges.
struct SortJob {
    Lines src, dst;            // the domain's lines in its current half and in the other
    uint32_t n, NC;
    int gx, gy;
    uint32_t *count, *cs, *key, *slot, *dflag, *drank, *order;
    Counters *cnt;
    DomainDesc *entry;         // the domain's table entry to repoint, or null
    uint64_t pt_off;           // the domain's point offset after the sort
};

// pblk: first block of each job over its points; cblk: over its cells.
__global__ __launch_bounds__(256) void k_sort_keys_many(const SortJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ pblk, int njobs) {
    const int j = job_of(pblk, njobs, blockIdx.x);
    const SortJob &J = jobs[j];
    sort_keys_block(J.src, J.n, J.gx, J.gy, J.NC, J.count, J.key, J.slot, J.dflag, J.cnt, blockIdx.x - pblk[j]);
}

__global__ __launch_bounds__(256) void k_sort_scatter_many(const SortJob *__restrict__ jobs,
                                                           const uint32_t *__restrict__ pblk, int njobs) {
    const int j = job_of(pblk, njobs, blockIdx.x);
    const SortJob &J = jobs[j];
    sort_scatter_item((blockIdx.x - pblk[j]) * blockDim.x + threadIdx.x, J.n, J.NC, J.key, J.slot, J.cs, J.drank,
                      J.order);
}

__global__ __launch_bounds__(256) void k_sort_cells_many(const SortJob *__restrict__ jobs,
                                                         const uint32_t *__restrict__ cblk, int njobs) {
    const int j = job_of(cblk, njobs, blockIdx.x);
    const SortJob &J = jobs[j];
    sort_cells_item((size_t)(blockIdx.x - cblk[j]) * blockDim.x + threadIdx.x, J.NC, J.cs, J.order, J.count);
}

__global__ __launch_bounds__(256) void k_sort_gather_many(const SortJob *__restrict__ jobs,
                                                          const uint32_t *__restrict__ pblk, int njobs, int nlines) {
    const int j = job_of(pblk, njobs, blockIdx.x);
    const SortJob &J = jobs[j];
    sort_gather_item((blockIdx.x - pblk[j]) * blockDim.x + threadIdx.x, J.src, J.dst, nlines, J.n, J.order,
                     nullptr);
}

// The end of every job's sort (k_sort_done for each): one thread per job.
__global__ __launch_bounds__(64) void k_sort_done_many(const SortJob *__restrict__ jobs, int njobs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    const SortJob &J = jobs[j];
    if (J.entry) J.entry->pt_off = J.pt_off;
    J.cnt->n_esc = 0;
}

}  // namespace ice
