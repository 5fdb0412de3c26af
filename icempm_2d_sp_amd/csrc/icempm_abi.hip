// icempm_abi.hip — the C-ABI of include/icempm.h on one MI355X (gfx950).
//
// One IceHandle stands for the reference's GPU_Partition (gpu_partition.h:63-134)
// and owns the device SoA (two buffers, for the sort), the grid lines, the
// status bytes, the stream and the timing events.  A handle holds one domain
// (ice_create) or several independent domains (ice_create_batch) that share
// every buffer and every launch: one launch per phase covers all tiles of all
// domains.  Nothing here throws across the ABI; failures return ICE_E_* and
// leave a message in ice_last_error.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/icempm.h"
#include "device_math.hpp"
#include "point_ops.hpp"
#include "kernels_naive.hpp"
#include "kernels_sort.hpp"
#include "tile_common.hpp"
#include "kernels_p2g.hpp"
#include "kernels_g2p.hpp"
#include "kernels_fused.hpp"
#include "kernels_frame.hpp"

using namespace ice;

namespace {

std::string g_create_error;

// m px py nx ny cur_x cur_y fx fy wind_x wind_y, and the time-varying wind's
// frames: w0 and w1 - w0 (ice_upload_wind_frames)
constexpr int kGridLines = 15;
enum { G_M = 0, G_PX, G_PY, G_NX, G_NY, G_CX, G_CY, G_FX, G_FY, G_WX, G_WY, G_W0X, G_W0Y, G_DWX, G_DWY };
constexpr int kChunk = 16;       // substeps per captured graph

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct TimingSet {
    hipEvent_t e[6];
};

struct Domain {
    int gx = 0, gy = 0;
    size_t N = 0;
    uint32_t capacity = 0, n = 0;
    uint32_t pt_base = 0;          // also the base of the domain's off-tile list
    int half = 0;                  // which half of the point buffer holds the points
    size_t node_base = 0, cs_base = 0;
    bool has_current = false;
    bool sorted = false;
    int steps_since_sort = 0, n_sorts = 0;
    // time-varying wind (ice_upload_wind_frames): wind = w0 + s (w1 - w0),
    // s = clamp((t - wind_t0) * wind_inv_span, 0, 1)
    bool wind_frames = false;
    double wind_t0 = 0.0, wind_inv_span = 0.0;
};

}  // namespace

struct IceHandle {
    IceParams prm{};
    DevParams dp{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool batch = false;
    std::vector<Domain> dom;

    // points (all domains): one allocation of two halves of [ICE_N_LINES_MAX][pitch];
    // a domain's sort gathers its points into the other half and flips Domain::half
    float *pts = nullptr;
    size_t half_floats = 0;
    size_t pitch = 0;
    uint32_t total_capacity = 0;
    int nlines = ICE_N_LINES;
    // grid (all domains)
    float *grid = nullptr;
    size_t gpitch = 0, total_nodes = 0;
    uint8_t *status = nullptr;
    Counters *cnt = nullptr;       // [ndom + 1]; cnt[ndom].step is the device step counter
    Counters *host_cnt = nullptr;  // pinned copy
    uint32_t *offtile = nullptr;
    uint32_t *off_pre = nullptr;   // [ndom + 1]: prefix of the off-tile list lengths (batches)
    uint32_t *cs = nullptr;        // cell starts of every domain
    size_t cs_total = 0;
    DomainDesc *d_desc = nullptr;
    TileRef *d_tiles = nullptr;
    size_t tiles_cap = 0;
    int ntiles = 0;
    bool tables_dirty = true;
    // sort scratch, sized for the largest domain
    bool s_count_clean = false;   // every cell counter zero (k_sort_cells clears what a sort counted)
    uint32_t *s_count = nullptr, *s_key = nullptr, *s_slot = nullptr, *s_dflag = nullptr, *s_drank = nullptr,
             *s_order = nullptr, *s_bsum = nullptr;
    size_t s_bsum_len = 0;
    size_t s_len = 0;   // entries of s_key, s_slot, s_dflag, s_order (s_drank: one more)
    // batched sorts: the job table on the device, its pinned host staging and
    // the event after its last upload (the staging is rewritten only after it)
    void *sort_tab = nullptr, *sort_tab_host = nullptr;
    size_t sort_tab_bytes = 0;
    hipEvent_t sort_tab_ev = nullptr;
    bool sort_tab_pending = false;
    // tiles
    // band 1: tiles walk down each column (y fastest), the order of the point
    // and cell layout (x * GridY + y), so neighbouring workgroups read
    // neighbouring memory (1-1.3 % faster than bands of 4 columns)
    int tile_w = 16, tile_h = 32, band = 1;
    FixScale fix{};
    float max_thickness = 1.f;
    // phase state
    double pending_t = 0.0;
    double last_t = 0.0;
    // timings
    bool timing = false;
    std::vector<TimingSet> ev_pool;
    size_t ev_used = 0;
    double acc[8] = {0};
    // frames of the pipelined path (two buffers, substep n reads fr[fr_cur])
    float *fr[2] = {nullptr, nullptr};
    size_t fr_cap = 0;             // floats per buffer
    size_t fr_used = 0;            // floats per buffer of the current tile shape
    int fr_cur = 0;
    bool frames_valid = false;     // fr[fr_cur] holds the P2G of the current point state
    bool last_fused = false;       // the last substep ran pipelined: its grid is fr[fr_cur ^ 1]
    bool p2g_called = false;       // ice_p2g on the pipelined path: the grid is fr[fr_cur]
    bool fused_opt = true;
    bool persistent = false;          // option "persistent": tile passes as persistent launches
    int resort_drift = 30;            // option "resort_drift": re-sort when this % of points left their cell
    uint32_t *tile_queue = nullptr;   // their per-XCD tile counters
    int n_cu = 0;
    // graph
    hipGraphExec_t graph_exec = nullptr;
    int graph_chunk = 0;
    int graph_fr0 = 0;
    // learned plastic projection
    float *mlp_p = nullptr, *mlp_g = nullptr, *mlp_m = nullptr, *mlp_v = nullptr, *mlp_loss = nullptr;
    bf16x4 *mlp_pk = nullptr;
    bf16x8 *mlp_ipk = nullptr;
    uint32_t *colors = nullptr;   // point colours by original point index
    uint32_t ncolors = 0;
    uint32_t *mlp_fail = nullptr;
    float *mlp_x = nullptr, *mlp_t = nullptr;   // generated pairs
    uint32_t mlp_n = 0;
    size_t mlp_loss_cap = 0;
    MlpNorm mlp_norm{};
    IceMlpSampling mlp_samp{};
    bool mlp_have_norm = false, mlp_have_params = false;
    uint32_t mlp_adam_t = 0;
    bool mlp_deployed = false;
    float mlp_tol = 1e-4f;
    // recording of ridge-capped plastic trial states (ice_mlp_record)
    float *mlp_rec = nullptr;
    unsigned long long *mlp_rec_cnt = nullptr;   // [0] seen, [1] kept
    uint32_t mlp_rec_cap = 0, mlp_rec_keep = 0;

    int ndom() const { return (int)dom.size(); }
    float *gline(int l) const { return grid + (size_t)l * gpitch; }
    Lines lines() const { return Lines{pts, pitch}; }
    uint64_t pt_off(int d, int half) const { return (uint64_t)half * half_floats + dom[d].pt_base; }
    float *dpts(int d) const { return pts + pt_off(d, dom[d].half); }
    Lines dlines(int d) const { return Lines{dpts(d), pitch}; }
    float *dgline(int l, int d) const { return gline(l) + dom[d].node_base; }
    Counters *dcnt(int d) const { return cnt + d; }
    Counters *step_cnt() const { return cnt + dom.size(); }
    bool naive() const { return (prm.flags & ICE_FLAG_NAIVE) != 0; }
    bool det() const { return (prm.flags & ICE_FLAG_DETERMINISTIC) != 0; }
    bool pipelined() const { return fused_opt && !naive() && !det(); }
    uint32_t max_capacity() const {
        uint32_t m = 1;
        for (auto &d : dom) m = std::max(m, d.capacity);
        return m;
    }
    size_t max_nodes() const {
        size_t m = 1;
        for (auto &d : dom) m = std::max(m, d.N);
        return m;
    }
};

#define HIPCHK(h, expr)                                                                 \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);               \
            return ICE_E_HIP;                                                           \
        }                                                                               \
    } while (0)

#define LAUNCHCHK(h, what)                                                              \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess) {                                                         \
            (h)->err = std::string(what) + ": " + hipGetErrorString(e_);                \
            return ICE_E_HIP;                                                           \
        }                                                                               \
    } while (0)

namespace {

void derive_params(IceParams &p) {
    // SimParams::ComputeLame / ComputeHelperVariables (parameters_sim.cpp:122-140)
    p.lambda = p.YoungsModulus * p.PoissonsRatio / ((1 + p.PoissonsRatio) * (1 - 2 * p.PoissonsRatio));
    p.mu = p.YoungsModulus / (2 * (1 + p.PoissonsRatio));
    p.kappa = p.mu * 2. / 3. + p.lambda;
    p.ParticleMass = p.ParticleVolume * p.IceDensity;
    p.UpdateEveryNthStep = (int)(p.AnimationFramePeriod / p.InitialTimeStep);
    p.cellsize_inv = 1. / p.cellsize;
    p.Dp_inv = 4. / (p.cellsize * p.cellsize);
    p.dt_vol_Dpinv = p.InitialTimeStep * p.ParticleVolume * p.Dp_inv;
    p.vmax = 0.25 * p.cellsize / p.InitialTimeStep;
}

DevParams make_dev_params(const IceParams &p) {
    DevParams d{};
    d.gx = p.GridXTotal;
    d.gy = p.GridYTotal;
    d.h = (float)p.cellsize;
    d.h_inv = (float)(1.0 / p.cellsize);
    d.dt = (float)p.InitialTimeStep;
    d.dt_h_inv = (float)(p.InitialTimeStep / p.cellsize);
    d.four_h_inv = (float)(4.0 / p.cellsize);
    d.mu = (float)p.mu;
    d.kappa = (float)p.kappa;
    d.half_kappa = (float)(0.5 * p.kappa);
    d.inv_2mu = (float)(0.5 / p.mu);
    d.inv_half_kappa = (float)(2.0 / p.kappa);
    d.particle_mass = (float)p.ParticleMass;
    d.dt_vol_Dpinv = (float)p.dt_vol_Dpinv;
    d.vmax2_01 = (float)(0.1 * p.vmax * p.vmax);
    d.kL = (float)(p.waterDragEffectiveLinear * p.InitialTimeStep);
    d.kQ = (float)(p.waterDragEffectiveQuadratic * p.InitialTimeStep);
    d.wind_dt = (float)(p.InitialTimeStep * p.windStressToAccel);
    const double pmax = p.IceCompressiveStrength, pmin = -p.IceTensileStrength;
    const double qmax = p.IceShearStrength, pmin2 = -p.IceTensileStrength2;
    const double beta = p.IceTensileStrength / p.IceCompressiveStrength;
    d.pmax = (float)pmax;
    d.pmin = (float)pmin;
    d.qmax = (float)qmax;
    d.pmin2 = (float)pmin2;
    d.beta = (float)beta;
    d.mSq = (float)((4. * qmax * qmax * (1. + 2. * beta)) / ((pmax - pmin) * (pmax - pmin)));
    const double q0 = 2 * std::sqrt(-pmax * pmin) * qmax / (pmax - pmin);
    d.fs_k = (float)(-q0 / pmin2);
    d.ellipse_c = (float)(2 * qmax / (pmax - pmin));
    const double tan_phi = std::tan(p.DP_phi * 3.14159265358979323846 / 180);
    d.tan_phi = (float)tan_phi;
    d.inv_tan_phi = (float)(1.0 / tan_phi);
    d.DP_threshold_p = (float)p.DP_threshold_p;
    d.ridge = (float)(p.RidgeFormationCoeff * 9.8 * p.IceDensity);
    d.has_cohesion = (p.flags & ICE_FLAG_COHESION_LINE) ? 1 : 0;
    d.has_current = 0;   // per domain (DomainDesc)
    d.update_every = p.UpdateEveryNthStep;
    d.dt_d = p.InitialTimeStep;
    return d;
}

// Fixed-point scales for the deterministic P2G: each node sum must stay below
// 2^62 for up to 4096 point-equivalents of the heaviest point at |v| = vmax,
// plus a stress term at the compressive strength.
FixScale make_fix_scale(const IceParams &p, float max_thickness) {
    const double th = std::max(1.0f, max_thickness);
    const double m_bound = 4096.0 * p.ParticleMass * th;
    const double stress = p.dt_vol_Dpinv * 4.0 * std::max(p.IceCompressiveStrength, p.kappa) * p.cellsize;
    const double p_bound = m_bound * p.vmax * 2.0 + 4096.0 * stress * th;
    FixScale f;
    const double lim = std::ldexp(1.0, 62);
    const int em = (int)std::floor(std::log2(lim / m_bound));
    const int ep = (int)std::floor(std::log2(lim / p_bound));
    f.m = (float)std::ldexp(1.0, em);
    f.p = (float)std::ldexp(1.0, ep);
    f.inv_m = std::ldexp(1.0, -em);
    f.inv_p = std::ldexp(1.0, -ep);
    return f;
}

unsigned blocks_for(size_t n, unsigned tpb) { return (unsigned)std::max<size_t>(1, (n + tpb - 1) / tpb); }

int launch_fill(IceHandle *h, float *a, size_t n) {
    if (!n) return ICE_OK;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<size_t>(blocks_for(n, 256), 65536)), dim3(256), 0,
                       h->stream, a, n);
    LAUNCHCHK(h, "k_fill_zero");
    return ICE_OK;
}

int launch_fill_u32(IceHandle *h, uint32_t *a, size_t n, uint32_t v) {
    if (!n) return ICE_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<size_t>(blocks_for(n, 256), 65536)), dim3(256), 0,
                       h->stream, a, n, v);
    LAUNCHCHK(h, "k_fill_u32");
    return ICE_OK;
}

// ---- tile dispatch ----------------------------------------------------------

using P2GFn = void (*)(DevParams, const DomainDesc *, const TileRef *, int, Lines, const uint32_t *, GridLines,
                       Counters *, uint32_t *, FixScale);
using G2PFn = void (*)(DevParams, StepArg, const DomainDesc *, const TileRef *, int, Lines, const uint32_t *,
                       GridLines, Counters *, uint32_t *);
using PipeFn = void (*)(DevParams, StepArg, SurrArg, const DomainDesc *, const TileRef *, int, Lines,
                        const uint32_t *, GridLines, Counters *, uint32_t *, const float *, float *, uint32_t *);
using OffFn = void (*)(DevParams, StepArg, const DomainDesc *, int, const uint32_t *, Lines, GridLines, Counters *,
                       const uint32_t *, const float *, float *, int);
using ToGridFn = void (*)(const DomainDesc *, int, const float *, GridLines);

struct TileKernels {
    int w, h;
    P2GFn p2g_det;   // deterministic mode's P2G (fixed point)
    G2PFn g2p;
    PipeFn fill, fused, fused_surr, fill_persistent, fused_persistent;
    PipeFn fused_buf, fused_surr_buf;   // 32-bit addressing (BufLines, 32-bit frame offsets)
    OffFn offtile;
    ToGridFn to_grid;
};

#define TK(W, H)                                                                                          \
    {W, H, k_p2g_tiles_det<W, H>, k_g2p_tiles<W, H>, k_tile_pipeline<W, H, false>,                          \
     k_tile_pipeline<W, H, true>, k_tile_pipeline<W, H, true, true>, k_tile_pipeline<W, H, false, false, true>, \
     k_tile_pipeline<W, H, true, false, true>, k_tile_pipeline<W, H, true, false, false, true>,             \
     k_tile_pipeline<W, H, true, true, false, true>, k_offtile_pipeline<W, H>, k_frames_to_grid<W, H>}
const TileKernels kTiles[] = {TK(8, 8),   TK(8, 16),  TK(16, 8),  TK(16, 16), TK(16, 32),
                              TK(32, 16), TK(32, 32), TK(32, 64), TK(64, 32), TK(64, 64)};
#undef TK

const TileKernels *find_tile(int w, int h) {
    for (const auto &t : kTiles)
        if (t.w == w && t.h == h) return &t;
    return nullptr;
}

GridLines grid_lines(const IceHandle *h) {
    GridLines G;
    G.m = h->gline(G_M);
    G.px = h->gline(G_PX);
    G.py = h->gline(G_PY);
    G.cur_x = h->gline(G_CX);
    G.cur_y = h->gline(G_CY);
    G.fx = h->gline(G_FX);
    G.fy = h->gline(G_FY);
    G.wind_x = h->gline(G_WX);
    G.wind_y = h->gline(G_WY);
    G.status = h->status;
    return G;
}

// Destroy the captured graph.  A launch of it may still be in flight (ice_run
// returns without waiting on paths that need no counters), so the stream is
// drained first: an executable graph is never destroyed under its own launch.
void drop_graph(IceHandle *h) {
    if (h->graph_exec) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipGraphExecDestroy(h->graph_exec);
    }
    h->graph_exec = nullptr;
    h->graph_chunk = 0;
}

// Domain and tile tables: domains by decreasing node count, inside a domain
// bands of `band` tile columns, rows outer, columns inner.
int upload_tables(IceHandle *h) {
    if (!h->tables_dirty) return ICE_OK;
    const int nd = h->ndom();
    std::vector<DomainDesc> desc(nd);
    uint64_t nframes = 0;
    for (int d = 0; d < nd; d++) {
        const Domain &D = h->dom[d];
        const int ntx = (D.gx + h->tile_w - 1) / h->tile_w, nty = (D.gy + h->tile_h - 1) / h->tile_h;
        desc[d] = DomainDesc{D.gx, D.gy, D.has_current ? 1 : 0, D.pt_base, h->pt_off(d, D.half), D.node_base,
                             D.cs_base, ntx, nty, nframes};
        nframes += (uint64_t)ntx * nty;
    }
    // frames: 3 components of (tile_w + 4) x (tile_h + 4) nodes per tile
    const size_t fr_need = (size_t)nframes * 3 * (h->tile_w + 4) * (h->tile_h + 4);
    if (fr_need > h->fr_cap) {
        HIPCHK(h, hipStreamSynchronize(h->stream));
        for (auto &f : h->fr) {
            if (f) (void)hipFree(f);
            f = nullptr;
        }
        h->fr_cap = 0;
        HIPCHK(h, hipMalloc(&h->fr[0], fr_need * sizeof(float)));
        HIPCHK(h, hipMalloc(&h->fr[1], fr_need * sizeof(float)));
        h->fr_cap = fr_need;
    }
    h->fr_used = fr_need;
    h->frames_valid = false;   // the frame layout follows the tile shape
    h->last_fused = false;
    h->p2g_called = false;
    std::vector<int> order(nd);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h->dom[a].N > h->dom[b].N; });
    std::vector<TileRef> tiles;
    for (int d : order) {
        const int ntx = (h->dom[d].gx + h->tile_w - 1) / h->tile_w;
        const int nty = (h->dom[d].gy + h->tile_h - 1) / h->tile_h;
        for (int b0 = 0; b0 < ntx; b0 += h->band) {
            const int wb = std::min(h->band, ntx - b0);
            for (int ty = 0; ty < nty; ty++)
                for (int k = 0; k < wb; k++) tiles.push_back(TileRef{(uint16_t)d, (uint16_t)(b0 + k), (uint16_t)ty, 0});
        }
    }
    if (nd > 1 && !h->mlp_deployed) {
        // Workgroup b takes entry xcd_remap(b): one contiguous run of the table
        // per XCD, which keeps shared halos in one L2 on a single domain.  Over
        // uneven domains with empty tiles those runs carry unequal work, so a
        // batch stores the table permuted to give workgroup b tile b (plain
        // order, every XCD a share of every domain): 2 % faster on 64 domains.
        // The tile pass with the learned projection is faster with the runs
        // (5.9 vs 6.4 ms on the compressed batch), so deploying it re-uploads.
        const int nt = (int)tiles.size(), per = nt / 8, rem = nt % 8;
        std::vector<TileRef> perm(tiles.size());
        for (int b = 0; b < nt; b++) {
            const int xcd = b % 8, j = b / 8;   // the device's xcd_remap
            perm[(size_t)(xcd * per + std::min(xcd, rem) + j)] = tiles[(size_t)b];
        }
        tiles.swap(perm);
    }
    if (tiles.size() > h->tiles_cap) {
        if (h->d_tiles) (void)hipFree(h->d_tiles);
        h->d_tiles = nullptr;
        HIPCHK(h, hipMalloc(&h->d_tiles, tiles.size() * sizeof(TileRef)));
        h->tiles_cap = tiles.size();
    }
    if (!h->d_desc) HIPCHK(h, hipMalloc(&h->d_desc, std::max(1, nd) * sizeof(DomainDesc)));
    if (!h->off_pre) HIPCHK(h, hipMalloc(&h->off_pre, (nd + 1) * sizeof(uint32_t)));
    HIPCHK(h, hipMemcpyAsync(h->d_desc, desc.data(), nd * sizeof(DomainDesc), hipMemcpyHostToDevice, h->stream));
    if (!tiles.empty())
        HIPCHK(h, hipMemcpyAsync(h->d_tiles, tiles.data(), tiles.size() * sizeof(TileRef), hipMemcpyHostToDevice,
                                 h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->ntiles = (int)tiles.size();
    h->tables_dirty = false;
    drop_graph(h);
    return ICE_OK;
}

// Two exclusive scans in one three-kernel pass (in_k[0..n_k) -> out_k[0..n_k],
// out_k[n_k] = total), each with its own half of the block-sum scratch.
int scan2_u32(IceHandle *h, const uint32_t *in0, size_t n0, uint32_t *out0, const uint32_t *in1, size_t n1,
              uint32_t *out1) {
    ScanJobs J;
    const uint32_t *in[2] = {in0, in1};
    const size_t n[2] = {n0, n1};
    uint32_t *out[2] = {out0, out1};
    uint32_t gx = 1;
    for (int k = 0; k < 2; k++) {
        const uint32_t nb = (uint32_t)std::max<size_t>(1, (n[k] + kScanTile - 1) / kScanTile);
        if ((size_t)nb + 1 > h->s_bsum_len) {
            h->err = "scan scratch too small";
            return ICE_E_STATE;
        }
        J.j[k] = ScanJob{in[k], n[k], h->s_bsum + (size_t)k * h->s_bsum_len, nb, out[k]};
        gx = std::max(gx, nb);
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(gx, 2), dim3(kScanThreads), 0, h->stream, J);
    LAUNCHCHK(h, "k_scan_reduce");
    hipLaunchKernelGGL(k_scan_blocks, dim3(1, 2), dim3(kScanThreads), 0, h->stream, J);
    LAUNCHCHK(h, "k_scan_blocks");
    hipLaunchKernelGGL(k_scan_apply, dim3(gx, 2), dim3(kScanThreads), 0, h->stream, J);
    LAUNCHCHK(h, "k_scan_apply");
    return ICE_OK;
}

// Sort domain d's points into cell order; keep disabled points behind (drop =
// false) or drop them (drop = true).  Leaves the domain's cs valid and its
// off-tile list empty.
// Domain d's ranges of the sort scratch: cell counters at its cs_base, the
// per-point arrays at its pt_base, the disabled ranks (n + 1 entries) at
// pt_base + d.
struct SortScratch {
    uint32_t *count, *key, *slot, *dflag, *drank, *order;
};
SortScratch sort_scratch(const IceHandle *h, int d) {
    const Domain &D = h->dom[d];
    return SortScratch{h->s_count + D.cs_base, h->s_key + D.pt_base, h->s_slot + D.pt_base,
                       h->s_dflag + D.pt_base, h->s_drank + D.pt_base + d, h->s_order + D.pt_base};
}
// Domain d's sort ranges lie inside the scratch allocations.
bool sort_ranges_ok(const IceHandle *h, int d) {
    const Domain &D = h->dom[d];
    return (size_t)D.pt_base + d + D.n + 1 <= h->s_len + 1 && D.cs_base + D.N + 2 <= h->cs_total && D.n <= D.capacity;
}

int device_sort(IceHandle *h, int d, bool drop, uint32_t *perm_dev) {
    if (!h->pts) { h->err = "no points allocated"; return ICE_E_STATE; }
    Domain &D = h->dom[d];
    const uint32_t n = D.n;
    const uint32_t NC = (uint32_t)D.N;
    uint32_t *cs = h->cs + D.cs_base;
    const SortScratch X = sort_scratch(h, d);
    if (!sort_ranges_ok(h, d)) { h->err = "sort scratch does not cover the domain"; return ICE_E_STATE; }
    int rc;
    // a sort that did not finish may have left counts anywhere: clear them all
    if (!h->s_count_clean && (rc = launch_fill_u32(h, h->s_count, h->cs_total, 0u))) return rc;
    h->s_count_clean = false;
    if (n) {
        hipLaunchKernelGGL(k_sort_keys, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->dlines(d), n, D.gx, D.gy,
                           NC, X.count, X.key, X.slot, X.dflag, h->dcnt(d));
        LAUNCHCHK(h, "k_sort_keys");
    }
    if ((rc = scan2_u32(h, X.count, NC, cs, X.dflag, n, X.drank))) return rc;
    if (n) {
        hipLaunchKernelGGL(k_sort_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, n, NC, X.key,
                           X.slot, cs, X.drank, X.order);
        LAUNCHCHK(h, "k_sort_scatter");
    }
    hipLaunchKernelGGL(k_sort_cells, dim3(blocks_for(NC, 256)), dim3(256), 0, h->stream, NC, cs, X.order,
                       X.count);
    LAUNCHCHK(h, "k_sort_cells");
    h->s_count_clean = true;
    uint32_t n_active = n;   // only a sort that drops the disabled points needs the count on the host
    if (drop) {
        HIPCHK(h, hipMemcpyAsync(&n_active, cs + NC, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    const uint32_t n_out = drop ? n_active : n;
    if (n_out) {
        // into the domain's range of the other half, which then holds the points
        const uint64_t to = h->pt_off(d, D.half ^ 1);
        Lines dst{h->pts + to, h->pitch};
        hipLaunchKernelGGL(k_sort_gather, dim3(blocks_for(n_out, 256)), dim3(256), 0, h->stream, h->dlines(d), dst,
                           h->nlines, n_out, X.order, perm_dev);
        LAUNCHCHK(h, "k_sort_gather");
        D.half ^= 1;
    }
    // the table entry (else the next table upload writes it), the off-tile
    // list, the disabled count of a dropping sort
    DomainDesc *entry = (n_out && h->d_desc && !h->tables_dirty) ? h->d_desc + d : nullptr;
    hipLaunchKernelGGL(k_sort_done, dim3(1), dim3(64), 0, h->stream, entry, h->pt_off(d, D.half), h->dcnt(d),
                       drop ? 1 : 0);
    LAUNCHCHK(h, "k_sort_done");
    // the sorted paths' graphs read the point base, the domain table, the cell
    // starts and the counters from device memory and stay valid; the naive
    // path captures each domain's point range and count by value
    if (h->naive() || drop) drop_graph(h);
    D.n = n_out;
    if (drop) HIPCHK(h, hipStreamSynchronize(h->stream));   // ice_sort's caller reads the result at once
    D.sorted = true;
    D.steps_since_sort = 0;
    D.n_sorts++;
    return ICE_OK;
}

// The sorts of domains ds (disabled points kept) as one set of launches:
// keys, the two scans of every domain, scatter, per-cell ordering, gather and
// the table entries, each launch over all the domains' blocks.  Each domain
// ends exactly as device_sort(h, d, false, nullptr) leaves it.  A single
// domain, or the naive path, goes through device_sort.
int device_sort_many(IceHandle *h, const std::vector<int> &ds) {
    if (ds.empty()) return ICE_OK;
    if (!h->pts) { h->err = "no points allocated"; return ICE_E_STATE; }
    if (ds.size() == 1 || h->naive()) {
        for (int d : ds) {
            int rc = device_sort(h, d, false, nullptr);
            if (rc) return rc;
        }
        return ICE_OK;
    }
    const int nj = (int)ds.size();
    // table layout: SortJob[nj] | ScanJob[2 nj] | pblk[nj + 1] | cblk[nj + 1] | sblk[2 nj + 1]
    const size_t o_scan = round_up(sizeof(SortJob) * nj, 256);
    const size_t o_pblk = o_scan + round_up(sizeof(ScanJob) * 2 * nj, 256);
    const size_t o_cblk = o_pblk + round_up(sizeof(uint32_t) * (nj + 1), 256);
    const size_t o_sblk = o_cblk + round_up(sizeof(uint32_t) * (nj + 1), 256);
    const size_t bytes = o_sblk + round_up(sizeof(uint32_t) * (2 * nj + 1), 256);
    if (bytes > h->sort_tab_bytes) {
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (h->sort_tab) (void)hipFree(h->sort_tab);
        if (h->sort_tab_host) (void)hipHostFree(h->sort_tab_host);
        h->sort_tab = h->sort_tab_host = nullptr;
        h->sort_tab_bytes = 0;
        h->sort_tab_pending = false;
        HIPCHK(h, hipMalloc(&h->sort_tab, bytes));
        HIPCHK(h, hipHostMalloc(&h->sort_tab_host, bytes));
        if (!h->sort_tab_ev) HIPCHK(h, hipEventCreateWithFlags(&h->sort_tab_ev, hipEventDisableTiming));
        h->sort_tab_bytes = bytes;
    }
    if (h->sort_tab_pending) HIPCHK(h, hipEventSynchronize(h->sort_tab_ev));   // the last upload has left the staging
    h->sort_tab_pending = false;
    char *hb = (char *)h->sort_tab_host, *db = (char *)h->sort_tab;
    SortJob *job = (SortJob *)hb;
    ScanJob *sj = (ScanJob *)(hb + o_scan);
    uint32_t *pblk = (uint32_t *)(hb + o_pblk), *cblk = (uint32_t *)(hb + o_cblk), *sblk = (uint32_t *)(hb + o_sblk);
    size_t np = 0, nc = 0, ns = 0, bs = 0;
    for (int j = 0; j < nj; j++) {
        const int d = ds[j];
        const Domain &D = h->dom[d];
        if (!sort_ranges_ok(h, d)) { h->err = "sort scratch does not cover the domain"; return ICE_E_STATE; }
        const SortScratch X = sort_scratch(h, d);
        const uint64_t to = h->pt_off(d, D.half ^ 1);
        SortJob &J = job[j];
        J.src = h->dlines(d);
        J.dst = Lines{h->pts + to, h->pitch};
        J.n = D.n;
        J.NC = (uint32_t)D.N;
        J.gx = D.gx;
        J.gy = D.gy;
        J.count = X.count;
        J.cs = h->cs + D.cs_base;
        J.key = X.key;
        J.slot = X.slot;
        J.dflag = X.dflag;
        J.drank = X.drank;
        J.order = X.order;
        J.cnt = h->dcnt(d);
        J.entry = (D.n && h->d_desc && !h->tables_dirty) ? h->d_desc + d : nullptr;
        J.pt_off = D.n ? to : h->pt_off(d, D.half);
        pblk[j] = (uint32_t)np;
        cblk[j] = (uint32_t)nc;
        np += (D.n + 255) / 256;
        nc += (D.N + 255) / 256;
        // the two scans of the domain: cell counts -> cs, disabled flags -> ranks
        const size_t len[2] = {D.N, D.n};
        const uint32_t *in[2] = {J.count, J.dflag};
        uint32_t *out[2] = {J.cs, J.drank};
        for (int k = 0; k < 2; k++) {
            const uint32_t nb = (uint32_t)std::max<size_t>(1, (len[k] + kScanTile - 1) / kScanTile);
            if (bs + nb + 1 > 2 * h->s_bsum_len) { h->err = "scan scratch too small"; return ICE_E_STATE; }
            sj[2 * j + k] = ScanJob{in[k], len[k], h->s_bsum + bs, nb, out[k]};
            sblk[2 * j + k] = (uint32_t)ns;
            ns += nb;
            bs += nb + 1;
        }
    }
    pblk[nj] = (uint32_t)np;
    cblk[nj] = (uint32_t)nc;
    sblk[2 * nj] = (uint32_t)ns;
    if (np >= (1u << 31) || nc >= (1u << 31) || ns >= (1u << 31)) { h->err = "sort batch too large"; return ICE_E_STATE; }
    HIPCHK(h, hipMemcpyAsync(db, hb, bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipEventRecord(h->sort_tab_ev, h->stream));
    h->sort_tab_pending = true;
    const SortJob *djob = (const SortJob *)db;
    const ScanJob *dsj = (const ScanJob *)(db + o_scan);
    const uint32_t *dp = (const uint32_t *)(db + o_pblk), *dc = (const uint32_t *)(db + o_cblk),
                   *dsb = (const uint32_t *)(db + o_sblk);
    int rc;
    if (!h->s_count_clean && (rc = launch_fill_u32(h, h->s_count, h->cs_total, 0u))) return rc;
    h->s_count_clean = false;
    if (np) {
        hipLaunchKernelGGL(k_sort_keys_many, dim3((unsigned)np), dim3(256), 0, h->stream, djob, dp, nj);
        LAUNCHCHK(h, "k_sort_keys_many");
    }
    hipLaunchKernelGGL(k_scan_reduce_many, dim3((unsigned)ns), dim3(kScanThreads), 0, h->stream, dsj, dsb, 2 * nj);
    LAUNCHCHK(h, "k_scan_reduce_many");
    hipLaunchKernelGGL(k_scan_blocks_many, dim3((unsigned)(2 * nj)), dim3(kScanThreads), 0, h->stream, dsj);
    LAUNCHCHK(h, "k_scan_blocks_many");
    hipLaunchKernelGGL(k_scan_apply_many, dim3((unsigned)ns), dim3(kScanThreads), 0, h->stream, dsj, dsb, 2 * nj);
    LAUNCHCHK(h, "k_scan_apply_many");
    if (np) {
        hipLaunchKernelGGL(k_sort_scatter_many, dim3((unsigned)np), dim3(256), 0, h->stream, djob, dp, nj);
        LAUNCHCHK(h, "k_sort_scatter_many");
    }
    hipLaunchKernelGGL(k_sort_cells_many, dim3((unsigned)nc), dim3(256), 0, h->stream, djob, dc, nj);
    LAUNCHCHK(h, "k_sort_cells_many");
    h->s_count_clean = true;
    if (np) {
        hipLaunchKernelGGL(k_sort_gather_many, dim3((unsigned)np), dim3(256), 0, h->stream, djob, dp, nj, h->nlines);
        LAUNCHCHK(h, "k_sort_gather_many");
    }
    hipLaunchKernelGGL(k_sort_done_many, dim3((unsigned)((nj + 63) / 64)), dim3(64), 0, h->stream, djob, nj);
    LAUNCHCHK(h, "k_sort_done_many");
    for (int j = 0; j < nj; j++) {
        Domain &D = h->dom[ds[j]];
        if (D.n) D.half ^= 1;
        D.sorted = true;
        D.steps_since_sort = 0;
        D.n_sorts++;
    }
    return ICE_OK;
}

// Sort every domain not in cell order (one batch of sorts).
int sort_unsorted(IceHandle *h) {
    std::vector<int> ds;
    for (int d = 0; d < h->ndom(); d++)
        if (!h->dom[d].sorted) ds.push_back(d);
    return device_sort_many(h, ds);
}

TimingSet *cur_timing(IceHandle *h) {
    if (!h->timing) return nullptr;
    if (h->ev_used == h->ev_pool.size()) {
        TimingSet t;
        for (auto &e : t.e)
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
        h->ev_pool.push_back(t);
    }
    return &h->ev_pool[h->ev_used];
}

void rec(IceHandle *h, int k) {
    TimingSet *t = cur_timing(h);
    if (t) (void)hipEventRecord(t->e[k], h->stream);
}

// The off-tile kernels walk every domain's list as one item space (grid-stride
// over a length only the device knows): for a batch, the lists' prefix first.
// Up to one workgroup per 2048 slots of all domains, so a long list (fast
// flows between re-sorts) is spread over the whole GPU; workgroups past the
// end exit after one load.
unsigned offtile_blocks(const IceHandle *h) {
    return std::min<unsigned>(1024, blocks_for(h->total_capacity, 2048));
}

int offtile_prefix(IceHandle *h, const uint32_t *&pre) {
    pre = nullptr;
    if (!h->d_desc || !h->off_pre) { h->err = "domain tables not uploaded"; return ICE_E_STATE; }
    if (h->ndom() > 1) {
        hipLaunchKernelGGL(k_offtile_prefix, dim3(1), dim3(256), 0, h->stream, (const Counters *)h->cnt, h->ndom(),
                           h->off_pre);
        LAUNCHCHK(h, "k_offtile_prefix");
        pre = h->off_pre;
    }
    return ICE_OK;
}

// The pipelined off-tile pass.
int launch_offtile(IceHandle *h, const TileKernels *tk, const StepArg &S, const GridLines &G, const float *Fc,
                   float *Fn, int do_g2p) {
    const uint32_t *pre;
    int rc = offtile_prefix(h, pre);
    if (rc) return rc;
    hipLaunchKernelGGL(tk->offtile, dim3(offtile_blocks(h)), dim3(256), 0, h->stream, h->dp, S,
                       (const DomainDesc *)h->d_desc, h->ndom(), pre, h->lines(), G, h->cnt,
                       (const uint32_t *)h->offtile, Fc, Fn, do_g2p);
    LAUNCHCHK(h, "k_offtile_pipeline");
    return ICE_OK;
}

DevParams dom_params(const IceHandle *h, int d) {
    DevParams P = h->dp;
    P.gx = h->dom[d].gx;
    P.gy = h->dom[d].gy;
    P.has_current = h->dom[d].has_current ? 1 : 0;
    return P;
}

int tile_launch_shape(IceHandle *h, PipeFn &fn, PipeFn persistent_fn, unsigned &grid, uint32_t *&queue);

// Phases of one substep.  Event slots: 0 start, 1 after the main P2G kernel,
// 2 after P2G, 3 before the main G2P kernel, 4 after G2P.
int phase_p2g(IceHandle *h) {
    if (!h->pts) { h->err = "no points allocated"; return ICE_E_STATE; }
    if (h->naive()) {
        for (int d = 0; d < h->ndom(); d++) {
            for (int l = G_M; l <= G_PY; l++) {   // reset_grid (gpu_partition.cpp:300-308)
                int rc = launch_fill(h, h->dgline(l, d), h->dom[d].N);
                if (rc) return rc;
            }
            if (h->dom[d].n) {
                hipLaunchKernelGGL(k_p2g_naive, dim3(blocks_for(h->dom[d].n, 256)), dim3(256), 0, h->stream,
                                   dom_params(h, d), h->dlines(d), h->dom[d].n, h->dgline(G_M, d), h->dgline(G_PX, d),
                                   h->dgline(G_PY, d), h->dcnt(d));
                LAUNCHCHK(h, "k_p2g_naive");
            }
        }
        rec(h, 1);
        rec(h, 2);
        return ICE_OK;
    }
    int rc = sort_unsorted(h);
    if (rc) return rc;
    rc = upload_tables(h);
    if (rc) return rc;
    const TileKernels *tk = find_tile(h->tile_w, h->tile_h);
    const GridLines G = grid_lines(h);
    if (h->det()) {
        // fixed-point pull per tile; the tile kernel also adds the off-tile points
        if (h->ntiles) {
            hipLaunchKernelGGL(tk->p2g_det, dim3(h->ntiles), dim3(kTileThreads), 0, h->stream, h->dp, h->d_desc,
                               h->d_tiles, h->ntiles, h->lines(), h->cs, G, h->cnt, h->offtile, h->fix);
            LAUNCHCHK(h, "k_p2g_tiles_det");
        }
        rec(h, 1);
        rec(h, 2);
        return ICE_OK;
    }
    // the tile pass's P2G (points loaded one per lane, their terms staged in
    // LDS for the cell lanes: each point line read once) into the frames, the
    // off-tile points into the same frames, then every node's frame sum into
    // the grid lines
    float *F = h->fr[h->fr_cur];
    const StepArg S0{0.0, 0, nullptr, 0};
    if (h->ntiles) {
        unsigned grid;
        uint32_t *queue;
        PipeFn fn = tk->fill;
        if ((rc = tile_launch_shape(h, fn, tk->fill_persistent, grid, queue))) return rc;
        hipLaunchKernelGGL(fn, dim3(grid), dim3(kTileThreads), 0, h->stream, h->dp, S0, SurrArg{}, h->d_desc,
                           h->d_tiles, h->ntiles, h->lines(), (const uint32_t *)h->cs, G, h->cnt, h->offtile,
                           (const float *)F, F, queue);
        LAUNCHCHK(h, "k_tile_pipeline (P2G)");
    }
    rec(h, 1);
    if ((rc = launch_offtile(h, tk, S0, G, F, F, 0))) return rc;
    for (int d = 0; d < h->ndom(); d++) {
        hipLaunchKernelGGL(tk->to_grid, dim3((unsigned)std::min<size_t>(blocks_for(h->dom[d].N, 256), 1024)),
                           dim3(256), 0, h->stream, (const DomainDesc *)h->d_desc, d, (const float *)F, G);
        LAUNCHCHK(h, "k_frames_to_grid");
    }
    rec(h, 2);
    return ICE_OK;
}

int phase_update(IceHandle *h, const StepArg &S) {
    if (!h->naive()) return ICE_OK;   // fused into the tile G2P
    for (int d = 0; d < h->ndom(); d++) {
        const size_t N = h->dom[d].N;
        hipLaunchKernelGGL(k_update_naive, dim3(blocks_for(N, 256)), dim3(256), 0, h->stream, dom_params(h, d), S,
                           h->dgline(G_M, d), h->dgline(G_PX, d), h->dgline(G_PY, d), h->status + h->dom[d].node_base,
                           h->dgline(G_CX, d), h->dgline(G_CY, d), h->dgline(G_WX, d), h->dgline(G_WY, d),
                           h->dgline(G_FX, d), h->dgline(G_FY, d), h->dcnt(d));
        LAUNCHCHK(h, "k_update_naive");
    }
    return ICE_OK;
}

int phase_g2p(IceHandle *h, const StepArg &S) {
    if (h->naive()) {
        rec(h, 3);
        for (int d = 0; d < h->ndom(); d++)
            if (h->dom[d].n) {
                hipLaunchKernelGGL(k_g2p_naive, dim3(blocks_for(h->dom[d].n, 256)), dim3(256), 0, h->stream,
                                   dom_params(h, d), S, h->dlines(d), h->dom[d].n, (const float *)h->dgline(G_PX, d),
                                   (const float *)h->dgline(G_PY, d), h->dcnt(d));
                LAUNCHCHK(h, "k_g2p_naive");
            }
        rec(h, 4);
        return ICE_OK;
    }
    if (!h->pts) { h->err = "no points allocated"; return ICE_E_STATE; }
    for (const auto &D : h->dom)
        if (!D.sorted) { h->err = "the sorted path's G2P needs the sorted points of a preceding P2G"; return ICE_E_STATE; }
    int rc = upload_tables(h);   // no work once a P2G has uploaded them
    if (rc) return rc;
    const GridLines G = grid_lines(h);
    const uint32_t *pre;
    if ((rc = offtile_prefix(h, pre))) return rc;
    hipLaunchKernelGGL(k_g2p_offtile, dim3(offtile_blocks(h)), dim3(256), 0, h->stream, h->dp, S,
                       (const DomainDesc *)h->d_desc, h->ndom(), pre, h->lines(), G, h->cnt,
                       (const uint32_t *)h->offtile);
    LAUNCHCHK(h, "k_g2p_offtile");
    rec(h, 3);
    if (h->ntiles) {
        const TileKernels *tk = find_tile(h->tile_w, h->tile_h);
        hipLaunchKernelGGL(tk->g2p, dim3(h->ntiles), dim3(kTileThreads), 0, h->stream, h->dp, S, h->d_desc, h->d_tiles,
                           h->ntiles, h->lines(), (const uint32_t *)h->cs, G, h->cnt, h->offtile);
        LAUNCHCHK(h, "k_g2p_tiles");
    }
    rec(h, 4);
    for (auto &D : h->dom) D.steps_since_sort++;
    return ICE_OK;
}

void end_timing_set(IceHandle *h) {
    if (h->timing && h->ev_used < h->ev_pool.size()) h->ev_used++;
}

// The unfused phases leave the frames stale.
void frames_stale(IceHandle *h) {
    h->frames_valid = false;
    h->last_fused = false;
    h->p2g_called = false;
}

bool surr_infer(const IceHandle *h) { return h->mlp_deployed && h->mlp_have_params && h->mlp_have_norm; }
// The tile pass with the learned projection runs when it is deployed or while
// trial states are recorded.
bool surr_pass(const IceHandle *h) { return surr_infer(h) || h->mlp_rec != nullptr; }

SurrArg surr_arg(const IceHandle *h) {
    SurrArg a{};
    a.infer = surr_infer(h) ? 1 : 0;
    a.rec_x = h->mlp_rec;
    a.rec_count = h->mlp_rec_cnt;
    a.rec_cap = h->mlp_rec_cap;
    a.rec_keep = h->mlp_rec_keep;
    a.M = MlpDev{h->mlp_pk, h->mlp_p, h->mlp_ipk};
    a.N = h->mlp_norm;
    a.tol = h->mlp_tol;
    a.lj0 = (float)h->mlp_samp.logJ_min;
    a.lj1 = (float)h->mlp_samp.logJ_max;
    a.d0 = (float)h->mlp_samp.D_min;
    a.d1 = (float)h->mlp_samp.D_max;
    a.jp0 = (float)h->mlp_samp.jp_min;
    a.jp1 = (float)h->mlp_samp.jp_max;
    return a;
}

// Grid and queue of a k_tile_pipeline launch: one workgroup per tile, or
// (option "persistent") as many workgroups as fit on the device at once, a
// multiple of 8 (one queue per XCD, each with workgroups), pulling tiles from
// the per-XCD queues, zeroed here by a kernel (graph-safe).
int tile_launch_shape(IceHandle *h, PipeFn &fn, PipeFn persistent_fn, unsigned &grid, uint32_t *&queue) {
    grid = (unsigned)h->ntiles;
    queue = nullptr;
    if (!h->persistent || !persistent_fn) return ICE_OK;
    fn = persistent_fn;
    int occ = 0;
    HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)fn, kTileThreads, 0));
    const unsigned fit = (unsigned)std::max(8, h->n_cu * std::max(occ, 1) / 8 * 8);
    grid = std::min<unsigned>(fit, (unsigned)round_up((size_t)h->ntiles, 8));
    int rc = launch_fill_u32(h, h->tile_queue, 8, 0u);
    if (rc) return rc;
    queue = h->tile_queue;
    return ICE_OK;
}

// Pipeline fill: P2G of the current point state into fr[fr_cur].
int pipeline_fill(IceHandle *h) {
    const TileKernels *tk = find_tile(h->tile_w, h->tile_h);
    const GridLines G = grid_lines(h);
    const StepArg S0{0.0, 0, nullptr, 0};
    float *F = h->fr[h->fr_cur];
    if (h->ntiles) {
        unsigned grid;
        uint32_t *queue;
        PipeFn fn = tk->fill;
        int rc = tile_launch_shape(h, fn, tk->fill_persistent, grid, queue);
        if (rc) return rc;
        hipLaunchKernelGGL(fn, dim3(grid), dim3(kTileThreads), 0, h->stream, h->dp, S0, SurrArg{}, h->d_desc,
                           h->d_tiles, h->ntiles, h->lines(), (const uint32_t *)h->cs, G, h->cnt, h->offtile,
                           (const float *)F, F, queue);
        LAUNCHCHK(h, "k_tile_pipeline (fill)");
    }
    int rc = launch_offtile(h, tk, S0, G, F, F, 0);
    if (rc) return rc;
    h->frames_valid = true;
    h->last_fused = false;
    return ICE_OK;
}

// One pipelined substep: update + G2P of S from fr[fr_cur], then the P2G of
// the new point state into fr[fr_cur ^ 1].  Event slots 1-3 mark the start of
// the tile kernel, 4 its end.
int phase_pipelined(IceHandle *h, const StepArg &S) {
    if (!h->pts) { h->err = "no points allocated"; return ICE_E_STATE; }
    int rc;
    if ((rc = sort_unsorted(h))) return rc;
    if ((rc = upload_tables(h))) return rc;
    if (!h->frames_valid && (rc = pipeline_fill(h))) return rc;
    rec(h, 1);
    rec(h, 2);
    rec(h, 3);
    const TileKernels *tk = find_tile(h->tile_w, h->tile_h);
    const GridLines G = grid_lines(h);
    const float *Fc = h->fr[h->fr_cur];
    float *Fn = h->fr[h->fr_cur ^ 1];
    if (h->ntiles) {
        const bool surr = surr_pass(h);
        // 32-bit addressing when every point slot and frame offset fits it
        const bool buf = buf_lines_ok(h->pitch) && h->fr_used < (size_t(1) << 30);
        PipeFn fn = surr ? (buf ? tk->fused_surr_buf : tk->fused_surr) : (buf ? tk->fused_buf : tk->fused);
        unsigned grid;
        uint32_t *queue;
        if ((rc = tile_launch_shape(h, fn, surr ? nullptr : tk->fused_persistent, grid, queue))) return rc;
        hipLaunchKernelGGL(fn, dim3(grid), dim3(kTileThreads), 0, h->stream, h->dp, S,
                           surr ? surr_arg(h) : SurrArg{}, h->d_desc, h->d_tiles, h->ntiles, h->lines(),
                           (const uint32_t *)h->cs, G, h->cnt, h->offtile, Fc, Fn, queue);
        LAUNCHCHK(h, "k_tile_pipeline");
    }
    rec(h, 4);
    if ((rc = launch_offtile(h, tk, S, G, Fc, Fn, 1))) return rc;
    h->fr_cur ^= 1;
    h->last_fused = true;
    h->p2g_called = false;
    for (auto &D : h->dom) D.steps_since_sort++;
    return ICE_OK;
}

// ice_p2g on the pipelined path: this substep's P2G is already in the frames
// (the previous substep's tile pass wrote it from the new point state), or
// the pipeline fill writes it now.  The update and G2P follow in ice_g2p as
// one tile pass, which also writes the next substep's P2G.
int phase_p2g_pipelined(IceHandle *h) {
    if (!h->pts) { h->err = "no points allocated"; return ICE_E_STATE; }
    int rc;
    if ((rc = sort_unsorted(h))) return rc;
    if ((rc = upload_tables(h))) return rc;
    if (!h->frames_valid && (rc = pipeline_fill(h))) return rc;
    h->p2g_called = true;
    return ICE_OK;
}

// The wind-stress lines at the substep's time for every domain with
// time-varying wind, before its node update reads them.
int wind_frames_apply(IceHandle *h, const StepArg &S) {
    for (int d = 0; d < h->ndom(); d++) {
        const Domain &D = h->dom[d];
        if (!D.wind_frames || D.N == 0) continue;
        hipLaunchKernelGGL(k_wind_frames, dim3((unsigned)std::min<size_t>(blocks_for(D.N, 256), 4096)), dim3(256), 0,
                           h->stream, S, h->dp, h->dgline(G_WX, d), h->dgline(G_WY, d),
                           (const float *)h->dgline(G_W0X, d), (const float *)h->dgline(G_W0Y, d),
                           (const float *)h->dgline(G_DWX, d), (const float *)h->dgline(G_DWY, d), D.N, D.wind_t0,
                           D.wind_inv_span);
        LAUNCHCHK(h, "k_wind_frames");
    }
    return ICE_OK;
}

int substep_args(IceHandle *h, const StepArg &S) {
    rec(h, 0);
    int rc;
    if ((rc = wind_frames_apply(h, S))) return rc;
    if (h->pipelined()) {
        if ((rc = phase_pipelined(h, S))) return rc;
    } else {
        frames_stale(h);
        if ((rc = phase_p2g(h))) return rc;
        if ((rc = phase_update(h, S))) return rc;
        if ((rc = phase_g2p(h, S))) return rc;
    }
    rec(h, 5);
    end_timing_set(h);
    return ICE_OK;
}

int read_counters(IceHandle *h) {
    HIPCHK(h, hipMemcpyAsync(h->host_cnt, h->cnt, sizeof(Counters) * (h->dom.size() + 1), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ICE_OK;
}

int build_graph(IceHandle *h, int chunk) {
    drop_graph(h);
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const bool was_timing = h->timing;
    h->timing = false;
    const std::vector<Domain> dom_save = h->dom;   // host-side counters advance at launch, not capture
    const int fr0 = h->fr_cur;
    int rc = ICE_OK;
    for (int k = 0; k < chunk && rc == ICE_OK; k++) {
        StepArg S{0.0, 0, &h->step_cnt()->step, k};
        S.complete = k == chunk - 1;
        rc = substep_args(h, S);
    }
    if (rc == ICE_OK) hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, h->stream, h->step_cnt(), chunk);
    h->timing = was_timing;
    h->dom = dom_save;
    h->graph_fr0 = fr0;
    if (rc == ICE_OK && h->fr_cur != fr0) { rc = ICE_E_STATE; h->err = "graph chunk must hold an even number of substeps"; }
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) { h->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e); return ICE_E_HIP; }
    e = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { h->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e); return ICE_E_HIP; }
    h->graph_chunk = chunk;
    return ICE_OK;
}

// After a graph launch: the host-side state the captured substeps stand for.
void graph_launched(IceHandle *h, int chunk) {
    for (auto &D : h->dom) D.steps_since_sort += chunk;
    if (h->pipelined()) h->last_fused = true;   // chunk is even: fr_cur is unchanged
    h->p2g_called = false;
}

bool grid_ok(int gx, int gy, std::string &why) {
    if (gx < 5 || gy < 5 || gx > 65535 || gy > 65535) {
        why = "grid size out of range (5..65535 nodes per axis; the cell word holds 16 bits per axis)";
        return false;
    }
    if ((uint64_t)gx * gy >= 0xffffffffull) { why = "grid too large"; return false; }
    return true;
}

// Allocate everything that depends only on the domains' grids.
int alloc_grid(IceHandle *h) {
    size_t nb = 0, cb = 0;
    for (auto &D : h->dom) {
        D.N = (size_t)D.gx * D.gy;
        D.node_base = nb;
        nb = round_up(nb + D.N, 64);
        D.cs_base = cb;
        cb = round_up(cb + D.N + 2, 64);
    }
    h->total_nodes = std::max<size_t>(nb, 64);
    h->gpitch = h->total_nodes;
    h->cs_total = std::max<size_t>(cb, 64);
    HIPCHK(h, hipMalloc(&h->grid, sizeof(float) * h->gpitch * kGridLines));
    HIPCHK(h, hipMalloc(&h->status, round_up(h->total_nodes, 256)));
    HIPCHK(h, hipMalloc(&h->cnt, sizeof(Counters) * (h->dom.size() + 1)));
    HIPCHK(h, hipHostMalloc(&h->host_cnt, sizeof(Counters) * (h->dom.size() + 1)));
    HIPCHK(h, hipMalloc(&h->s_count, sizeof(uint32_t) * h->cs_total));   // each domain's counters at its cs_base
    HIPCHK(h, hipMemsetAsync(h->s_count, 0, sizeof(uint32_t) * h->cs_total, h->stream));
    HIPCHK(h, hipMalloc(&h->cs, sizeof(uint32_t) * h->cs_total));
    HIPCHK(h, hipMemsetAsync(h->grid, 0, sizeof(float) * h->gpitch * kGridLines, h->stream));
    HIPCHK(h, hipMemsetAsync(h->status, ICE_GRID_STATUS_WATER, round_up(h->total_nodes, 256), h->stream));
    HIPCHK(h, hipMemsetAsync(h->cnt, 0, sizeof(Counters) * (h->dom.size() + 1), h->stream));
    HIPCHK(h, hipMemsetAsync(h->cs, 0, sizeof(uint32_t) * h->cs_total, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::memset(h->host_cnt, 0, sizeof(Counters) * (h->dom.size() + 1));
    return ICE_OK;
}

// (Re)allocate the point buffers and the scratch that depends on capacity.
int alloc_points(IceHandle *h) {
    size_t base = 0;
    for (const auto &D : h->dom) base = round_up(base + std::max<uint32_t>(D.capacity, 1), 64);
    // kernels address a line slot by a 32-bit byte offset; a rejected
    // capacity leaves the current allocation as it was
    if (base >= (1ull << 30)) { h->err = "total point capacity exceeds 2^30 slots"; return ICE_E_CAPACITY; }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graph(h);
    for (void **p : {(void **)&h->pts, (void **)&h->offtile, (void **)&h->s_key,
                     (void **)&h->s_slot, (void **)&h->s_dflag, (void **)&h->s_drank, (void **)&h->s_order,
                     (void **)&h->s_bsum}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    base = 0;
    for (auto &D : h->dom) {
        D.pt_base = (uint32_t)base;
        D.half = 0;
        D.n = 0;
        D.sorted = false;
        base = round_up(base + std::max<uint32_t>(D.capacity, 1), 64);
    }
    h->total_capacity = (uint32_t)base;
    h->pitch = base;
    h->half_floats = h->pitch * ICE_N_LINES_MAX;
    const size_t bytes = sizeof(float) * h->half_floats;
    // sort scratch: every domain's range at its pt_base (drank: pt_base + d,
    // one more entry per domain), so many domains sort in one set of launches
    const size_t mc = h->total_capacity + h->dom.size();
    h->s_len = mc;
    HIPCHK(h, hipMalloc(&h->pts, 2 * bytes));
    HIPCHK(h, hipMemsetAsync(h->pts, 0, bytes, h->stream));   // half 0, where every domain starts
    HIPCHK(h, hipMalloc(&h->offtile, sizeof(uint32_t) * h->total_capacity));
    HIPCHK(h, hipMalloc(&h->s_key, sizeof(uint32_t) * mc));
    HIPCHK(h, hipMalloc(&h->s_slot, sizeof(uint32_t) * mc));
    HIPCHK(h, hipMalloc(&h->s_dflag, sizeof(uint32_t) * mc));
    HIPCHK(h, hipMalloc(&h->s_drank, sizeof(uint32_t) * (mc + 1)));
    HIPCHK(h, hipMalloc(&h->s_order, sizeof(uint32_t) * mc));
    // block sums: one half per scan of a single sort's pass; a batch of sorts
    // lays out nb + 1 per scan job over both halves (at most the total points
    // plus the total cells over kScanTile, plus 2 per job and domain)
    h->s_bsum_len = std::max<size_t>((mc + h->cs_total) / kScanTile + 4 * h->dom.size(),
                                     std::max<size_t>(h->max_capacity(), h->max_nodes()) / kScanTile) + 8;
    HIPCHK(h, hipMalloc(&h->s_bsum, sizeof(uint32_t) * 2 * h->s_bsum_len));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->tables_dirty = true;
    return ICE_OK;
}

int create_common(const IceParams *p, int device, std::vector<Domain> doms, bool batch, IceHandle **out) {
    if (!p || !out) { g_create_error = "null argument"; return ICE_E_ARG; }
    *out = nullptr;
    for (auto &D : doms)
        if (!grid_ok(D.gx, D.gy, g_create_error)) return ICE_E_ARG;
    if (doms.empty() || doms.size() > 65535) { g_create_error = "1..65535 domains"; return ICE_E_ARG; }
    if (!(p->InitialTimeStep > 0) || !(p->cellsize > 0)) { g_create_error = "dt and cellsize must be > 0"; return ICE_E_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_create_error = "no HIP device"; return ICE_E_DEVICE; }
    if (device < 0 || device >= ndev) { g_create_error = "device index out of range"; return ICE_E_DEVICE; }
    IceHandle *h = new IceHandle;
    h->prm = *p;
    derive_params(h->prm);
    h->device = device;
    h->batch = batch;
    h->dom = std::move(doms);
    auto fail = [&](int code) {
        g_create_error = h->err;
        ice_destroy(h);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) { h->err = "hipSetDevice failed"; return fail(ICE_E_HIP); }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) { h->err = "stream"; return fail(ICE_E_HIP); }
    int rc = alloc_grid(h);
    if (rc) return fail(rc);
    h->dp = make_dev_params(h->prm);
    h->fix = make_fix_scale(h->prm, 1.f);
    *out = h;
    return ICE_OK;
}

int check_domain(IceHandle *h, int d) {
    if (d < 0 || d >= h->ndom()) { h->err = "domain index out of range"; return ICE_E_ARG; }
    return ICE_OK;
}

int upload_points_d(IceHandle *h, int d, const float *soa, size_t host_pitch, uint32_t n, int n_lines) {
    if (check_domain(h, d)) return ICE_E_ARG;
    if (!soa && n) return ICE_E_ARG;
    if (!h->pts) { h->err = "allocate points first"; return ICE_E_STATE; }
    Domain &D = h->dom[d];
    if (n > D.capacity) { h->err = "more points than the allocated capacity"; return ICE_E_CAPACITY; }
    if (n_lines < ICE_N_LINES || n_lines > ICE_N_LINES_MAX || host_pitch < n) { h->err = "bad layout"; return ICE_E_ARG; }
    if ((h->prm.flags & ICE_FLAG_COHESION_LINE) && n_lines < ICE_N_LINES_MAX) {
        h->err = "ICE_FLAG_COHESION_LINE needs 20 lines";
        return ICE_E_ARG;
    }
    HIPCHK(h, hipSetDevice(h->device));
    h->nlines = (h->prm.flags & ICE_FLAG_COHESION_LINE) ? ICE_N_LINES_MAX : ICE_N_LINES;
    if (n) {
        HIPCHK(h, hipMemcpy2DAsync(h->dpts(d), h->pitch * sizeof(float), soa, host_pitch * sizeof(float),
                                   n * sizeof(float), h->nlines, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_fe_minus_identity, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->dpts(d),
                           (size_t)h->pitch, n);
        LAUNCHCHK(h, "k_fe_minus_identity");
    }
    D.n = n;
    // the partition's disabled count starts at 0 (gpu_implementation5.cpp:137, gpu_partition.cpp:101-103)
    static const int32_t zero = 0;   // outlives the asynchronous copy
    HIPCHK(h, hipMemcpyAsync(&h->dcnt(d)->disabled, &zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
    if (n) {
        const float *th = soa + (size_t)ICE_IDX_THICKNESS * host_pitch;
        for (uint32_t i = 0; i < n; i++) h->max_thickness = std::max(h->max_thickness, th[i]);
    }
    h->fix = make_fix_scale(h->prm, h->max_thickness);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    D.sorted = false;
    drop_graph(h);
    frames_stale(h);
    if (!h->naive()) return device_sort(h, d, false, nullptr);
    return ICE_OK;
}

int upload_grid_d(IceHandle *h, int d, const uint8_t *status, const float *normals) {
    if (check_domain(h, d) || !status) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    const Domain &D = h->dom[d];
    HIPCHK(h, hipMemcpyAsync(h->status + D.node_base, status, D.N, hipMemcpyHostToDevice, h->stream));
    if (normals) {
        HIPCHK(h, hipMemcpyAsync(h->dgline(G_NX, d), normals, D.N * sizeof(float), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->dgline(G_NY, d), normals + D.N, D.N * sizeof(float), hipMemcpyHostToDevice,
                                 h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ICE_OK;
}

int upload_forcing_d(IceHandle *h, int d, const float *current, const float *wind) {
    if (check_domain(h, d)) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    Domain &D = h->dom[d];
    int rc;
    if (current) {
        HIPCHK(h, hipMemcpyAsync(h->dgline(G_CX, d), current, D.N * sizeof(float), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->dgline(G_CY, d), current + D.N, D.N * sizeof(float), hipMemcpyHostToDevice,
                                 h->stream));
        bool any = false;
        for (size_t i = 0; i < 2 * D.N && !any; i++) any = current[i] != 0.f;
        D.has_current = any;
    } else {
        if ((rc = launch_fill(h, h->dgline(G_CX, d), D.N))) return rc;
        if ((rc = launch_fill(h, h->dgline(G_CY, d), D.N))) return rc;
        D.has_current = false;
    }
    D.wind_frames = false;   // a static wind field replaces any time-varying one
    if (wind) {
        HIPCHK(h, hipMemcpyAsync(h->dgline(G_WX, d), wind, D.N * sizeof(float), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->dgline(G_WY, d), wind + D.N, D.N * sizeof(float), hipMemcpyHostToDevice, h->stream));
    } else {
        if ((rc = launch_fill(h, h->dgline(G_WX, d), D.N))) return rc;
        if ((rc = launch_fill(h, h->dgline(G_WY, d), D.N))) return rc;
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->tables_dirty = true;
    drop_graph(h);
    return ICE_OK;
}

int download_d(IceHandle *h, int d, float *soa, size_t host_pitch, int n_lines, float *forces, uint32_t *error_code,
               int32_t *disabled_count) {
    if (check_domain(h, d)) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    const Domain &D = h->dom[d];
    if (soa) {
        if (n_lines < ICE_N_LINES || n_lines > ICE_N_LINES_MAX || host_pitch < D.n) { h->err = "bad layout"; return ICE_E_ARG; }
        const int nl = std::min(n_lines, h->nlines);
        if (D.n)
            HIPCHK(h, hipMemcpy2DAsync(soa, host_pitch * sizeof(float), h->dpts(d), h->pitch * sizeof(float),
                                       D.n * sizeof(float), nl, hipMemcpyDeviceToHost, h->stream));
    }
    if (forces) {
        HIPCHK(h, hipMemcpyAsync(forces, h->dgline(G_FX, d), D.N * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(forces + D.N, h->dgline(G_FY, d), D.N * sizeof(float), hipMemcpyDeviceToHost,
                                 h->stream));
    }
    int rc = read_counters(h);
    if (rc) return rc;
    if (soa) {
        uint32_t *u = reinterpret_cast<uint32_t *>(soa + (size_t)ICE_IDX_UTILITY * host_pitch);
        for (uint32_t i = 0; i < D.n; i++) u[i] &= ~kInternalBits;
        // the device holds Fe - I; the reference's layout holds Fe
        float *f00 = soa + (size_t)ICE_IDX_FE00 * host_pitch, *f11 = soa + (size_t)(ICE_IDX_FE00 + 3) * host_pitch;
        for (uint32_t i = 0; i < D.n; i++) {
            f00[i] += 1.f;
            f11[i] += 1.f;
        }
    }
    if (error_code) *error_code = h->host_cnt[d].err;
    if (disabled_count) *disabled_count = h->host_cnt[d].disabled;
    return ICE_OK;
}

void fill_info(IceHandle *h, int d, IceInfo *info) {
    const Domain &D = h->dom[d];
    info->n_points = D.n;
    info->capacity = D.capacity;
    info->point_pitch = h->pitch;
    info->grid_pitch = h->gpitch;
    info->grid_x = D.gx;
    info->grid_y = D.gy;
    info->device = h->device;
    info->sorted = D.sorted ? 1 : 0;
    info->steps_since_sort = D.steps_since_sort;
    info->n_sorts = D.n_sorts;
    info->flags = h->prm.flags;
    info->tile_w = h->tile_w;
    info->tile_h = h->tile_h;
    info->n_offtile = h->host_cnt ? h->host_cnt[d].n_esc : 0;
    info->n_domains = h->ndom();
    info->n_tiles = h->ntiles;
}

}  // namespace

extern "C" {

int ice_params_reset(IceParams *p) {
    if (!p) return ICE_E_ARG;
    std::memset(p, 0, sizeof(*p));
    // SimParams::Reset (parameters_sim.cpp:7-55)
    p->InitialTimeStep = 3.e-5;
    p->YoungsModulus = 5.e8;
    p->sea_water_density = 1030;
    p->AnimationFramePeriod = 200;
    p->PoissonsRatio = 0.3;
    p->IceDensity = 916;
    p->IceCompressiveStrength = 100e6;
    p->IceTensileStrength = 10e6;
    p->IceShearStrength = 1e6;
    p->IceTensileStrength2 = 10e6;
    p->RidgeFormationCoeff = 30;
    p->DP_phi = 62;
    p->DP_threshold_p = 1e4;
    p->waterDragEffectiveLinear = 0.01;
    p->waterDragEffectiveQuadratic = 0.01;
    p->tpb_P2G = 256;
    p->tpb_Upd = 512;
    p->tpb_G2P = 128;
    p->cellsize = 1.0;
    p->windStressToAccel = 1.0 / 916.0;
    derive_params(*p);
    return ICE_OK;
}

int ice_params_derive(IceParams *p) {
    if (!p) return ICE_E_ARG;
    if (!(p->InitialTimeStep > 0) || !(p->cellsize > 0)) return ICE_E_ARG;
    derive_params(*p);
    return ICE_OK;
}

const char *ice_last_error(const IceHandle *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int ice_create(const IceParams *p, int device, IceHandle **out) {
    if (!p) { g_create_error = "null argument"; return ICE_E_ARG; }
    Domain D;
    D.gx = p->GridXTotal;
    D.gy = p->GridYTotal;
    return create_common(p, device, {D}, false, out);
}

int ice_create_batch(const IceParams *p, int device, int n_domains, const int32_t *grid_x, const int32_t *grid_y,
                     const uint32_t *capacity, IceHandle **out) {
    if (!p || !grid_x || !grid_y || !capacity || !out || n_domains <= 0) {
        g_create_error = "bad arguments";
        return ICE_E_ARG;
    }
    std::vector<Domain> doms(n_domains);
    for (int d = 0; d < n_domains; d++) {
        doms[d].gx = grid_x[d];
        doms[d].gy = grid_y[d];
        doms[d].capacity = capacity[d];
    }
    int rc = create_common(p, device, std::move(doms), true, out);
    if (rc) return rc;
    rc = alloc_points(*out);
    if (rc) {
        g_create_error = (*out)->err;
        ice_destroy(*out);
        *out = nullptr;
    }
    return rc;
}

int ice_destroy(IceHandle *h) {
    if (!h) return ICE_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    for (auto &t : h->ev_pool)
        for (auto &e : t.e) (void)hipEventDestroy(e);
    for (void *p : {(void *)h->pts, (void *)h->grid, (void *)h->status, (void *)h->cnt,
                    (void *)h->offtile, (void *)h->s_count, (void *)h->cs, (void *)h->s_key, (void *)h->s_slot,
                    (void *)h->s_dflag, (void *)h->s_drank, (void *)h->s_order, (void *)h->s_bsum,
                    (void *)h->d_desc, (void *)h->off_pre, (void *)h->d_tiles, (void *)h->fr[0], (void *)h->fr[1], (void *)h->mlp_p,
                    (void *)h->mlp_g, (void *)h->mlp_m, (void *)h->mlp_v, (void *)h->mlp_loss, (void *)h->mlp_pk, (void *)h->mlp_ipk, (void *)h->colors, (void *)h->tile_queue,
                    (void *)h->mlp_fail, (void *)h->mlp_x, (void *)h->mlp_t, (void *)h->mlp_rec,
                    (void *)h->mlp_rec_cnt})
        if (p) (void)hipFree(p);
    if (h->host_cnt) (void)hipHostFree(h->host_cnt);
    if (h->sort_tab) (void)hipFree(h->sort_tab);
    if (h->sort_tab_host) (void)hipHostFree(h->sort_tab_host);
    if (h->sort_tab_ev) (void)hipEventDestroy(h->sort_tab_ev);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return ICE_OK;
}

int ice_allocate(IceHandle *h, uint32_t cap) {
    if (!h) return ICE_E_ARG;
    if (h->batch) { h->err = "capacities of a batch are fixed by ice_create_batch"; return ICE_E_STATE; }
    HIPCHK(h, hipSetDevice(h->device));
    const uint32_t old = h->dom[0].capacity;
    h->dom[0].capacity = cap;
    const int rc = alloc_points(h);
    if (rc == ICE_E_CAPACITY) h->dom[0].capacity = old;
    return rc;
}

int ice_upload_points(IceHandle *h, const float *soa, size_t host_pitch, uint32_t n, int n_lines) {
    if (!h) return ICE_E_ARG;
    return upload_points_d(h, 0, soa, host_pitch, n, n_lines);
}

int ice_upload_grid(IceHandle *h, const uint8_t *status, const float *normals) {
    if (!h) return ICE_E_ARG;
    return upload_grid_d(h, 0, status, normals);
}

int ice_upload_forcing(IceHandle *h, const float *current, const float *wind) {
    if (!h) return ICE_E_ARG;
    return upload_forcing_d(h, 0, current, wind);
}

int ice_upload_wind_frames(IceHandle *h, int d, const float *wind0, const float *wind1, double t0, double t1) {
    if (!h || !wind0 || !wind1 || !(t1 > t0) || !std::isfinite(t0) || !std::isfinite(t1)) return ICE_E_ARG;
    if (check_domain(h, d)) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    Domain &D = h->dom[d];
    const size_t N = D.N;
    std::vector<float> dw(2 * N);
    for (size_t i = 0; i < 2 * N; i++) dw[i] = wind1[i] - wind0[i];
    HIPCHK(h, hipStreamSynchronize(h->stream));   // no substep in flight reads the lines
    for (int c = 0; c < 2; c++) {
        HIPCHK(h, hipMemcpy(h->dgline(c ? G_W0Y : G_W0X, d), wind0 + c * N, N * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->dgline(c ? G_WY : G_WX, d), wind0 + c * N, N * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->dgline(c ? G_DWY : G_DWX, d), dw.data() + c * N, N * sizeof(float),
                            hipMemcpyHostToDevice));
    }
    D.wind_frames = true;
    D.wind_t0 = t0;
    D.wind_inv_span = 1.0 / (t1 - t0);
    drop_graph(h);   // captured substeps must now include the wind update
    return ICE_OK;
}

int ice_upload_domain(IceHandle *h, int d, const float *soa, size_t host_pitch, uint32_t n, int n_lines,
                      const uint8_t *status, const float *current, const float *wind_stress) {
    if (!h) return ICE_E_ARG;
    int rc;
    if (status && (rc = upload_grid_d(h, d, status, nullptr))) return rc;
    if ((rc = upload_forcing_d(h, d, current, wind_stress))) return rc;
    return upload_points_d(h, d, soa, host_pitch, n, n_lines);
}

int ice_update_constants(IceHandle *h, const IceParams *p) {
    if (!h || !p) return ICE_E_ARG;
    if (!h->batch && (p->GridXTotal != h->dom[0].gx || p->GridYTotal != h->dom[0].gy)) {
        h->err = "grid size cannot change";
        return ICE_E_ARG;
    }
    const uint32_t old_flags = h->prm.flags;
    h->prm = *p;
    derive_params(h->prm);
    if ((old_flags ^ h->prm.flags) & ICE_FLAG_NAIVE)
        for (auto &D : h->dom) D.sorted = false;
    h->dp = make_dev_params(h->prm);
    h->fix = make_fix_scale(h->prm, h->max_thickness);
    frames_stale(h);
    // GPU_Partition::update_constants zeroes the error word (gpu_partition.cpp:143)
    static const uint32_t zero = 0;   // outlives the asynchronous copies
    for (int d = 0; d < h->ndom(); d++)
        HIPCHK(h, hipMemcpyAsync(&h->dcnt(d)->err, &zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graph(h);
    return ICE_OK;
}

int ice_set_tile(IceHandle *h, int tw, int th) {
    if (!h) return ICE_E_ARG;
    if (!find_tile(tw, th)) { h->err = "unsupported tile shape"; return ICE_E_ARG; }
    h->tile_w = tw;
    h->tile_h = th;
    h->tables_dirty = true;
    frames_stale(h);
    drop_graph(h);
    return ICE_OK;
}

int ice_set_option(IceHandle *h, const char *name, int value) {
    if (!h || !name) return ICE_E_ARG;
    if (std::strcmp(name, "band") == 0) {
        if (value < 1 || value > 4096) { h->err = "band must be 1..4096"; return ICE_E_ARG; }
        h->band = value;
        h->tables_dirty = true;
    } else if (std::strcmp(name, "pipelined") == 0) {
        if (value != 0 && value != 1) { h->err = "pipelined must be 0 or 1"; return ICE_E_ARG; }
        h->fused_opt = value != 0;
        frames_stale(h);
    } else if (std::strcmp(name, "resort_drift") == 0) {
        if (value < 0 || value > 100) { h->err = "resort_drift must be 0..100 (0: off)"; return ICE_E_ARG; }
        h->resort_drift = value;
    } else if (std::strcmp(name, "persistent") == 0) {
        if (value != 0 && value != 1) { h->err = "persistent must be 0 or 1"; return ICE_E_ARG; }
        HIPCHK(h, hipSetDevice(h->device));
        if (value && !h->tile_queue) HIPCHK(h, hipMalloc(&h->tile_queue, 8 * sizeof(uint32_t)));
        if (value && !h->n_cu)
            HIPCHK(h, hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, h->device));
        h->persistent = value != 0;
    } else {
        h->err = std::string("unknown option ") + name;
        return ICE_E_ARG;
    }
    drop_graph(h);
    return ICE_OK;
}

int ice_reset_grid(IceHandle *h) {
    if (!h) return ICE_E_ARG;
    rec(h, 0);
    // The tiled P2G stores every node, so only the naive path clears (in ice_p2g).
    return ICE_OK;
}

// The phase calls, in the order GPU_Implementation5 issues them per substep
// (gpu_implementation5.cpp:18-81).  On the pipelined path they defer: ice_p2g
// makes sure the substep's P2G is in the frames, ice_update_nodes records the
// time, and ice_g2p runs the one tile pass (update + G2P, then the next P2G),
// so a caller driving the phases runs the same kernels as ice_substep.  The
// naive, deterministic and unfused paths launch each phase's kernels.
int ice_p2g(IceHandle *h) {
    if (!h) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    if (h->pipelined()) return phase_p2g_pipelined(h);
    frames_stale(h);
    return phase_p2g(h);
}

int ice_update_nodes(IceHandle *h, float simulation_time, float windSpeed, float windAngle) {
    if (!h) return ICE_E_ARG;
    (void)windSpeed;   // uniform wind of the reference's signature; per-node stress comes from ice_upload_forcing
    (void)windAngle;
    HIPCHK(h, hipSetDevice(h->device));
    h->pending_t = simulation_time;
    h->last_t = simulation_time;
    if (h->pipelined()) return ICE_OK;   // fused into the tile pass of ice_g2p
    StepArg S{(double)simulation_time, 0, nullptr, 0};
    frames_stale(h);
    int rc = wind_frames_apply(h, S);
    if (rc) return rc;
    return phase_update(h, S);
}

int ice_g2p(IceHandle *h, int recordPQ, int enablePointTransfer, int applyGlensLaw) {
    if (!h) return ICE_E_ARG;
    (void)enablePointTransfer;   // single partition: no point transfer
    (void)applyGlensLaw;         // Glen's law is commented out in the reference (kernels.cuh:625-652)
    HIPCHK(h, hipSetDevice(h->device));
    StepArg S{h->pending_t, recordPQ, nullptr, 0};
    h->last_t = h->pending_t;
    int rc;
    if (h->pipelined()) {
        rc = wind_frames_apply(h, S);
        if (!rc) rc = phase_pipelined(h, S);
    } else {
        frames_stale(h);
        rc = phase_g2p(h, S);
    }
    rec(h, 5);
    end_timing_set(h);
    return rc;
}

int ice_substep(IceHandle *h, double t, int recordPQ) {
    if (!h) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    StepArg S{t, recordPQ, nullptr, 0};
    h->last_t = t;
    return substep_args(h, S);
}

int ice_run(IceHandle *h, int first_step, int nsteps, int use_graph) {
    if (!h || nsteps < 0) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    if (!h->naive()) {
        int rc = sort_unsorted(h);
        if (rc) return rc;
    }
    int rc = h->naive() ? ICE_OK : upload_tables(h);
    if (rc) return rc;
    if (h->pipelined() && !h->frames_valid && (rc = pipeline_fill(h))) return rc;   // never inside a graph
    hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, h->stream, h->step_cnt(), first_step);
    LAUNCHCHK(h, "k_step_set");
    int done = 0;
    while (done < nsteps) {
        const int chunk = std::min(kChunk, nsteps - done);
        if (!h->naive()) {   // the chunk's last substep counts the points off their sorted cell
            hipLaunchKernelGGL(k_clear_drift, dim3(1), dim3(256), 0, h->stream, h->cnt, h->ndom());
            LAUNCHCHK(h, "k_clear_drift");
        }
        // the naive path is the reference's launch structure: always eager
        if (use_graph && chunk == kChunk && !h->timing && !h->naive()) {
            if (!h->graph_exec || h->graph_chunk != chunk || h->graph_fr0 != h->fr_cur) {
                if ((rc = build_graph(h, chunk))) return rc;
            }
            HIPCHK(h, hipGraphLaunch(h->graph_exec, h->stream));
            graph_launched(h, chunk);
        } else {
            for (int k = 0; k < chunk; k++) {
                StepArg S{0.0, 0, &h->step_cnt()->step, k};
                S.complete = k == chunk - 1;
                if ((rc = substep_args(h, S))) return rc;
            }
            hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, h->stream, h->step_cnt(), chunk);
            LAUNCHCHK(h, "k_step_advance");
        }
        done += chunk;
        if (!h->naive()) {
            // re-sort domains whose off-tile list grew past max(256, n/256), and
            // (pipelined path) those where resort_drift % of the points sit off
            // their sorted cell: those escape in bulk within the next chunk
            if ((rc = read_counters(h))) return rc;
            std::vector<int> ds;
            for (int d = 0; d < h->ndom(); d++) {
                const Domain &D = h->dom[d];
                const uint32_t limit = std::max<uint32_t>(256, D.n / 256);
                const bool drifted = h->resort_drift > 0 &&
                                     (uint64_t)h->host_cnt[d].n_drift * 100 > (uint64_t)D.n * h->resort_drift;
                if (h->host_cnt[d].n_esc > limit || drifted) ds.push_back(d);
            }
            if ((rc = device_sort_many(h, ds))) return rc;
        }
    }
    h->last_t = h->prm.InitialTimeStep * (first_step + nsteps - 1);
    return ICE_OK;
}

int ice_clear_force_accumulator(IceHandle *h) {
    if (!h) return ICE_E_ARG;
    int rc = launch_fill(h, h->gline(G_FX), h->gpitch);
    if (rc) return rc;
    return launch_fill(h, h->gline(G_FY), h->gpitch);
}

int ice_download(IceHandle *h, float *soa, size_t host_pitch, int n_lines, float *forces, uint32_t *error_code,
                 int32_t *disabled_count) {
    if (!h) return ICE_E_ARG;
    return download_d(h, 0, soa, host_pitch, n_lines, forces, error_code, disabled_count);
}

int ice_download_domain(IceHandle *h, int d, float *soa, size_t host_pitch, int n_lines, float *forces,
                        uint32_t *error_code, int32_t *disabled_count) {
    if (!h) return ICE_E_ARG;
    return download_d(h, d, soa, host_pitch, n_lines, forces, error_code, disabled_count);
}

int ice_upload_point_colors(IceHandle *h, const uint32_t *rgb, uint32_t n) {
    if (!h || (n && !rgb)) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->colors) (void)hipFree(h->colors);
    h->colors = nullptr;
    h->ncolors = 0;
    if (n) {
        HIPCHK(h, hipMalloc(&h->colors, sizeof(uint32_t) * n));
        HIPCHK(h, hipMemcpy(h->colors, rgb, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
        h->ncolors = n;
    }
    return ICE_OK;
}

int ice_prepare_frame_arrays(IceHandle *h, int elapsed_cycles, IceFrameArrays *out) {
    if (!h || !out) return ICE_E_ARG;
    if (h->ndom() != 1) { h->err = "frame arrays of a single-domain handle"; return ICE_E_STATE; }
    if (h->naive()) { h->err = "frame arrays need the sorted path"; return ICE_E_STATE; }
    if (!h->pts) { h->err = "no points allocated"; return ICE_E_STATE; }
    HIPCHK(h, hipSetDevice(h->device));
    int rc = device_sort(h, 0, false, nullptr);   // every point in its cell's range
    if (rc) return rc;
    const Domain &D = h->dom[0];
    const size_t N = D.N;
    // one buffer: count [N] u8, mask [N] u8, rgb [3N] u8, fields [10][N] f32, partials, regions
    const int rblocks = 256;
    const size_t o_mask = round_up(N, 256), o_rgb = o_mask + round_up(N, 256), o_fields = o_rgb + round_up(3 * N, 256);
    const size_t o_part = o_fields + round_up(sizeof(float) * kFrameFields * N, 256);
    const size_t o_reg = o_part + sizeof(double) * 2 * kMaxRegions * rblocks;
    char *buf = nullptr;
    HIPCHK(h, hipMalloc(&buf, o_reg + sizeof(float) * 2 * kMaxRegions));
    FrameOut fo{out->count ? (uint8_t *)buf : nullptr, out->fields ? (float *)(buf + o_fields) : nullptr,
                out->rgb ? (uint8_t *)(buf + o_rgb) : nullptr, out->mass_mask ? (uint8_t *)(buf + o_mask) : nullptr};
    DomainDesc dd{};
    dd.gx = D.gx;
    dd.gy = D.gy;
    hipError_t e = hipSuccess;
    if (fo.count || fo.fields || fo.rgb || fo.mask) {
        hipLaunchKernelGGL(k_frame_fields, dim3((unsigned)std::min<size_t>(blocks_for(N, 256), 8192)), dim3(256), 0,
                           h->stream, dd, h->dlines(0), (const uint32_t *)(h->cs + D.cs_base),
                           (const uint8_t *)(h->status + D.node_base), (const uint32_t *)h->colors, h->ncolors,
                           (float)h->prm.ParticleMass, fo);
        e = hipGetLastError();
    }
    if (e == hipSuccess && out->forces_per_region) {
        const double scale = elapsed_cycles > 0 ? 1.0 / (elapsed_cycles * h->prm.InitialTimeStep) : 1.0;
        hipLaunchKernelGGL(k_region_forces, dim3(rblocks), dim3(256), 0, h->stream, dd,
                           (const float *)h->dgline(G_FX, 0), (const float *)h->dgline(G_FY, 0),
                           (const uint8_t *)(h->status + D.node_base), scale, (double *)(buf + o_part));
        e = hipGetLastError();
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_region_sum, dim3(1), dim3(256), 0, h->stream, (const double *)(buf + o_part), rblocks,
                               (float *)(buf + o_reg));
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(out->forces_per_region, buf + o_reg, sizeof(float) * 2 * kMaxRegions,
                               hipMemcpyDeviceToHost, h->stream);
    }
    if (e == hipSuccess && out->count) e = hipMemcpyAsync(out->count, buf, N, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && out->mass_mask)
        e = hipMemcpyAsync(out->mass_mask, buf + o_mask, N, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && out->rgb) e = hipMemcpyAsync(out->rgb, buf + o_rgb, 3 * N, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && out->fields)
        e = hipMemcpyAsync(out->fields, buf + o_fields, sizeof(float) * kFrameFields * N, hipMemcpyDeviceToHost,
                           h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) { h->err = std::string("ice_prepare_frame_arrays: ") + hipGetErrorString(e); return ICE_E_HIP; }
    return ICE_OK;
}

int ice_download_grid(IceHandle *h, float *mpp, float *velocity) {
    if (!h) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    const Domain &D = h->dom[0];
    if ((h->last_fused || h->p2g_called) && (mpp || velocity)) {
        // the node sums are in the frames: those the last substep read, or
        // after ice_p2g on the pipelined path those the next tile pass reads
        // (a tile change clears both flags, so the frame layout is current)
        const TileKernels *tk = find_tile(h->tile_w, h->tile_h);
        const int src = h->p2g_called ? h->fr_cur : h->fr_cur ^ 1;
        hipLaunchKernelGGL(tk->to_grid, dim3(1024), dim3(256), 0, h->stream, (const DomainDesc *)h->d_desc, 0,
                           (const float *)h->fr[src], grid_lines(h));
        LAUNCHCHK(h, "k_frames_to_grid");
    }
    if (mpp)
        HIPCHK(h, hipMemcpy2DAsync(mpp, D.N * sizeof(float), h->gline(G_M) + D.node_base, h->gpitch * sizeof(float),
                                   D.N * sizeof(float), 3, hipMemcpyDeviceToHost, h->stream));
    if (velocity) {
        if (h->naive()) {
            // the naive update kernel leaves velocities in (px, py)
            HIPCHK(h, hipMemcpy2DAsync(velocity, D.N * sizeof(float), h->gline(G_PX) + D.node_base,
                                       h->gpitch * sizeof(float), D.N * sizeof(float), 2, hipMemcpyDeviceToHost,
                                       h->stream));
        } else {
            int rc = upload_tables(h);
            if (rc) return rc;
            float *buf = nullptr;
            HIPCHK(h, hipMalloc(&buf, 2 * D.N * sizeof(float)));
            hipLaunchKernelGGL(k_node_velocity_out, dim3(1024), dim3(256), 0, h->stream, h->dp, h->last_t, h->d_desc,
                               0, grid_lines(h), buf, buf + D.N);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(velocity, buf, 2 * D.N * sizeof(float), hipMemcpyDeviceToHost, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            (void)hipFree(buf);
            if (e != hipSuccess) { h->err = std::string("node velocity: ") + hipGetErrorString(e); return ICE_E_HIP; }
        }
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ICE_OK;
}

int ice_synchronize(IceHandle *h) {
    if (!h) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ICE_OK;
}

int ice_sort(IceHandle *h, uint32_t *n_out, uint32_t *perm) {
    if (!h) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    uint32_t *perm_dev = nullptr;
    const uint32_t n_before = h->dom[0].n;
    if (perm && n_before) HIPCHK(h, hipMalloc(&perm_dev, sizeof(uint32_t) * n_before));
    int rc = device_sort(h, 0, true, perm_dev);
    if (rc == ICE_OK && perm && h->dom[0].n) {
        hipError_t e = hipMemcpyAsync(perm, perm_dev, sizeof(uint32_t) * h->dom[0].n, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) { h->err = hipGetErrorString(e); rc = ICE_E_HIP; }
    }
    if (perm_dev) (void)hipFree(perm_dev);
    if (n_out) *n_out = h->dom[0].n;
    return rc;
}

int ice_reset_timings(IceHandle *h) {
    if (!h) return ICE_E_ARG;
    (void)hipStreamSynchronize(h->stream);
    h->ev_used = 0;
    std::fill(std::begin(h->acc), std::end(h->acc), 0.0);
    return ICE_OK;
}

int ice_enable_timings(IceHandle *h, int enable) {
    if (!h) return ICE_E_ARG;
    h->timing = enable != 0;
    return ICE_OK;
}

int ice_get_timings(IceHandle *h, double out[8]) {
    if (!h || !out) return ICE_E_ARG;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (size_t k = 0; k < h->ev_used; k++) {
        TimingSet &t = h->ev_pool[k];
        float ms[4];
        HIPCHK(h, hipEventElapsedTime(&ms[0], t.e[0], t.e[1]));   // main P2G kernel
        HIPCHK(h, hipEventElapsedTime(&ms[1], t.e[0], t.e[2]));   // P2G phase
        HIPCHK(h, hipEventElapsedTime(&ms[2], t.e[3], t.e[4]));   // main G2P kernel
        HIPCHK(h, hipEventElapsedTime(&ms[3], t.e[0], t.e[5]));   // substep
        h->acc[0] += ms[0];
        h->acc[1] += ms[1];
        h->acc[2] += ms[3] - ms[1] - ms[2];   // update + off-tile G2P
        h->acc[3] += ms[2];
        h->acc[4] += ms[3];
        h->acc[5] += 1;
    }
    h->ev_used = 0;
    for (int k = 0; k < 8; k++) out[k] = h->acc[k];
    return ICE_OK;
}

int ice_autotune(IceHandle *h, int nsteps, int *tile_w, int *tile_h, double *best_ms) {
    if (!h || nsteps <= 0) return ICE_E_ARG;
    if (h->naive()) { h->err = "the naive path has no tiles"; return ICE_E_STATE; }
    if (!h->pts) { h->err = "no points allocated"; return ICE_E_STATE; }
    HIPCHK(h, hipSetDevice(h->device));
    int rc;
    if ((rc = sort_unsorted(h))) return rc;
    // Save everything a trial changes (points in both halves of the point
    // buffer, cell starts, force lines, counters, off-tile lists) and restore
    // it, with the domains' halves, after each candidate, so tuning leaves the
    // simulation state as it found it.
    const size_t pts_bytes = sizeof(float) * h->pitch * h->nlines;
    const size_t cs_bytes = sizeof(uint32_t) * h->cs_total;
    const size_t f_bytes = sizeof(float) * h->gpitch * 2;
    const size_t c_bytes = sizeof(Counters) * (h->dom.size() + 1);
    const size_t o_bytes = sizeof(uint32_t) * h->total_capacity;
    char *save = nullptr;
    HIPCHK(h, hipMalloc(&save, 2 * pts_bytes + cs_bytes + f_bytes + c_bytes + o_bytes));
    const std::vector<Domain> dom_save = h->dom;
    auto copy_all = [&](bool store) -> hipError_t {
        char *p = save;
        void *dev[] = {h->pts, h->pts + h->half_floats, h->cs, h->gline(G_FX), h->cnt, h->offtile};
        const size_t len[] = {pts_bytes, pts_bytes, cs_bytes, f_bytes, c_bytes, o_bytes};
        for (int k = 0; k < 6; k++) {
            hipError_t e = store ? hipMemcpyAsync(p, dev[k], len[k], hipMemcpyDeviceToDevice, h->stream)
                                 : hipMemcpyAsync(dev[k], p, len[k], hipMemcpyDeviceToDevice, h->stream);
            if (e != hipSuccess) return e;
            p += len[k];
        }
        return hipStreamSynchronize(h->stream);
    };
    hipError_t e = copy_all(true);
    const int old_w = h->tile_w, old_h = h->tile_h;
    int bw = old_w, bh = old_h;
    double bms = 1e30;
    const bool was_timing = h->timing;
    h->timing = false;
    rc = ICE_OK;
    const int cand[][2] = {{8, 8}, {8, 16}, {16, 8}, {16, 16}, {16, 32}, {32, 16}, {32, 32}, {32, 64}, {64, 32}};
    for (auto &c : cand) {
        if (e != hipSuccess || rc != ICE_OK) break;
        h->tile_w = c[0];
        h->tile_h = c[1];
        h->tables_dirty = true;
        if ((rc = upload_tables(h))) break;
        const StepArg S{0.0, 0, &h->step_cnt()->step, 0};
        if ((rc = substep_args(h, S))) break;   // warm-up
        if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) break;
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < nsteps && rc == ICE_OK; k++) rc = substep_args(h, S);
        if (rc) break;
        if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) break;
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / nsteps;
        if (ms < bms) { bms = ms; bw = c[0]; bh = c[1]; }
        h->dom = dom_save;
        e = copy_all(false);
    }
    h->timing = was_timing;
    h->dom = dom_save;
    if (e == hipSuccess) e = copy_all(false);
    (void)hipFree(save);
    const bool ok = rc == ICE_OK && e == hipSuccess;
    h->tile_w = ok ? bw : old_w;
    h->tile_h = ok ? bh : old_h;
    h->tables_dirty = true;
    drop_graph(h);
    if (e != hipSuccess) { h->err = std::string("autotune: ") + hipGetErrorString(e); return ICE_E_HIP; }
    if (rc) return rc;
    if (tile_w) *tile_w = bw;
    if (tile_h) *tile_h = bh;
    if (best_ms) *best_ms = bms;
    return upload_tables(h);
}

int ice_get_info(IceHandle *h, IceInfo *info) {
    if (!h || !info) return ICE_E_ARG;
    int rc = read_counters(h);
    if (rc) return rc;
    fill_info(h, 0, info);
    return ICE_OK;
}

int ice_get_domain_info(IceHandle *h, int d, IceInfo *info) {
    if (!h || !info) return ICE_E_ARG;
    if (check_domain(h, d)) return ICE_E_ARG;
    int rc = read_counters(h);
    if (rc) return rc;
    fill_info(h, d, info);
    return ICE_OK;
}

// ---- learned plastic projection ---------------------------------------------

}  // extern "C"

namespace {

int mlp_alloc(IceHandle *h) {
    if (h->mlp_p) return ICE_OK;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipMalloc(&h->mlp_p, sizeof(float) * kMlpParams));
    HIPCHK(h, hipMalloc(&h->mlp_g, sizeof(float) * kMlpParams));
    HIPCHK(h, hipMalloc(&h->mlp_m, sizeof(float) * kMlpParams));
    HIPCHK(h, hipMalloc(&h->mlp_v, sizeof(float) * kMlpParams));
    HIPCHK(h, hipMalloc(&h->mlp_pk, sizeof(bf16x4) * kPkOps * 64));
    HIPCHK(h, hipMalloc(&h->mlp_ipk, sizeof(bf16x8) * kIpOps * 64));
    HIPCHK(h, hipMalloc(&h->mlp_fail, sizeof(uint32_t)));
    int rc;
    if ((rc = launch_fill(h, h->mlp_p, kMlpParams))) return rc;
    if ((rc = launch_fill(h, h->mlp_m, kMlpParams))) return rc;
    if ((rc = launch_fill(h, h->mlp_v, kMlpParams))) return rc;
    ice_mlp_default_sampling(&h->mlp_samp);
    return ICE_OK;
}

MlpPhys mlp_phys(const IceHandle *h) {
    const IceParams &p = h->prm;
    MlpPhys m;
    m.kappa = p.kappa;
    m.mu = p.mu;
    m.ridge = p.RidgeFormationCoeff * 9.8 * p.IceDensity;
    m.pmax = p.IceCompressiveStrength;
    m.pmin = -p.IceTensileStrength;
    m.ellipse_c = 2 * p.IceShearStrength / (m.pmax - m.pmin);
    m.tan_phi = std::tan(p.DP_phi * 3.14159265358979323846 / 180);
    m.thickness = 1.0;
    return m;
}

// Master parameters -> packed training operands, and with infer the packed
// inference operands too.
int mlp_pack(IceHandle *h, bool infer = true) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(kPkOps), dim3(64), 0, h->stream, (const float *)h->mlp_p, h->mlp_pk);
    LAUNCHCHK(h, "k_mlp_pack");
    if (infer) {
        hipLaunchKernelGGL(k_mlp_pack_infer, dim3(kIpOps), dim3(64), 0, h->stream, (const float *)h->mlp_p,
                           h->mlp_ipk);
        LAUNCHCHK(h, "k_mlp_pack_infer");
    }
    return ICE_OK;
}

int mlp_pairs_alloc(IceHandle *h, uint32_t n);
int mlp_pairs_finish(IceHandle *h, uint32_t n, float *x, float *t, uint32_t *n_failed, bool ranges_from_data);

int mlp_ready(IceHandle *h) {
    if (!h->mlp_have_params || !h->mlp_have_norm) {
        h->err = "surrogate parameters and standardisation not set";
        return ICE_E_STATE;
    }
    return ICE_OK;
}

// Device copy of n rows of `width` floats (freed by the caller).
int mlp_upload(IceHandle *h, const float *src, size_t count, float **dst) {
    *dst = nullptr;
    HIPCHK(h, hipMalloc(dst, sizeof(float) * std::max<size_t>(count, 1)));
    if (count) HIPCHK(h, hipMemcpyAsync(*dst, src, sizeof(float) * count, hipMemcpyHostToDevice, h->stream));
    return ICE_OK;
}

}  // namespace

extern "C" {

int ice_mlp_default_sampling(IceMlpSampling *s) {
    if (!s) return ICE_E_ARG;
    s->logJ_min = -0.02;
    s->logJ_max = 0.0002;
    s->D_min = 1e-4;
    s->D_max = 2e-2;
    s->jp_min = 0.5;
    s->jp_max = 2.0;
    return ICE_OK;
}

int ice_mlp_generate(IceHandle *h, uint32_t n, uint64_t seed, const IceMlpSampling *smp, float *x, float *t,
                     uint32_t *n_failed) {
    if (!h || n == 0) return ICE_E_ARG;
    int rc = mlp_alloc(h);
    if (rc) return rc;
    if (smp) h->mlp_samp = *smp;
    const IceMlpSampling &S = h->mlp_samp;
    if (!(S.D_min > 0 && S.D_max > S.D_min && S.jp_min > 0 && S.jp_max > S.jp_min && S.logJ_max > S.logJ_min)) {
        h->err = "bad sampling ranges";
        return ICE_E_ARG;
    }
    if ((rc = mlp_pairs_alloc(h, n))) return rc;
    const MlpSampling ms{S.logJ_min, S.logJ_max, S.D_min, S.D_max, S.jp_min, S.jp_max};
    hipLaunchKernelGGL(k_mlp_pairs, dim3(std::min<unsigned>(blocks_for(n, 256), 16384)), dim3(256), 0, h->stream,
                       mlp_phys(h), ms, seed, n, h->mlp_x, h->mlp_t, h->mlp_fail);
    LAUNCHCHK(h, "k_mlp_pairs");
    return mlp_pairs_finish(h, n, x, t, n_failed, false);
}

int ice_mlp_record(IceHandle *h, uint32_t capacity, float keep_fraction) {
    if (!h || !(keep_fraction >= 0.f && keep_fraction <= 1.f)) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graph(h);   // graphs hold the tile pass's arguments by value
    if (h->mlp_rec) (void)hipFree(h->mlp_rec);
    if (h->mlp_rec_cnt) (void)hipFree(h->mlp_rec_cnt);
    h->mlp_rec = nullptr;
    h->mlp_rec_cnt = nullptr;
    h->mlp_rec_cap = 0;
    if (capacity == 0) return ICE_OK;
    HIPCHK(h, hipMalloc(&h->mlp_rec, sizeof(float) * kMlpIn * (size_t)capacity));
    HIPCHK(h, hipMalloc(&h->mlp_rec_cnt, 2 * sizeof(unsigned long long)));
    int rc = launch_fill_u32(h, reinterpret_cast<uint32_t *>(h->mlp_rec_cnt), 4, 0u);
    if (rc) return rc;
    // rows start as NaN: a slot counted as stored but never written shows
    if ((rc = launch_fill_u32(h, reinterpret_cast<uint32_t *>(h->mlp_rec), kMlpIn * (size_t)capacity, 0x7fc00000u)))
        return rc;
    h->mlp_rec_cap = capacity;
    h->mlp_rec_keep = (uint32_t)std::min(4294967295.0, std::floor((double)keep_fraction * 4294967296.0));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ICE_OK;
}

int ice_mlp_record_keep(IceHandle *h, float keep_fraction) {
    if (!h || !(keep_fraction >= 0.f && keep_fraction <= 1.f)) return ICE_E_ARG;
    if (!h->mlp_rec) { h->err = "no recording (ice_mlp_record first)"; return ICE_E_STATE; }
    const uint32_t keep = (uint32_t)std::min(4294967295.0, std::floor((double)keep_fraction * 4294967296.0));
    if (keep == h->mlp_rec_keep) return ICE_OK;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graph(h);   // graphs hold the tile pass's arguments by value
    h->mlp_rec_keep = keep;
    return ICE_OK;
}

int ice_mlp_record_count(IceHandle *h, uint64_t *seen, uint32_t *stored) {
    if (!h) return ICE_E_ARG;
    unsigned long long c[2] = {0, 0};
    if (h->mlp_rec_cnt) {
        HIPCHK(h, hipSetDevice(h->device));
        HIPCHK(h, hipMemcpyAsync(c, h->mlp_rec_cnt, sizeof(c), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    if (seen) *seen = c[0];
    if (stored) *stored = (uint32_t)std::min<unsigned long long>(c[1], h->mlp_rec_cap);
    return ICE_OK;
}

int ice_mlp_records(IceHandle *h, float *x, uint32_t n_max, uint32_t *n) {
    if (!h || (!x && n_max)) return ICE_E_ARG;
    uint32_t stored = 0;
    int rc = ice_mlp_record_count(h, nullptr, &stored);
    if (rc) return rc;
    const uint32_t k = std::min(stored, n_max);
    if (k) {
        HIPCHK(h, hipMemcpyAsync(x, h->mlp_rec, sizeof(float) * kMlpIn * (size_t)k, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    if (n) *n = k;
    return ICE_OK;
}

int ice_mlp_generate_recorded(IceHandle *h, uint32_t n, uint64_t seed, float *x, float *t, uint32_t *n_failed) {
    if (!h || n == 0) return ICE_E_ARG;
    uint32_t stored = 0;
    int rc = ice_mlp_record_count(h, nullptr, &stored);
    if (rc) return rc;
    if (stored == 0) { h->err = "no recorded trial states (ice_mlp_record, then run pipelined substeps)"; return ICE_E_STATE; }
    if ((rc = mlp_alloc(h))) return rc;
    if ((rc = mlp_pairs_alloc(h, n))) return rc;
    hipLaunchKernelGGL(k_mlp_pairs_recorded, dim3(std::min<unsigned>(blocks_for(n, 256), 16384)), dim3(256), 0,
                       h->stream, mlp_phys(h), (const float *)h->mlp_rec, stored, seed, n, h->mlp_x, h->mlp_t,
                       h->mlp_fail);
    LAUNCHCHK(h, "k_mlp_pairs_recorded");
    return mlp_pairs_finish(h, n, x, t, n_failed, true);
}

}  // extern "C"

namespace {

// (Re)allocate the training pairs for n rows and clear the failure count.
int mlp_pairs_alloc(IceHandle *h, uint32_t n) {
    if (h->mlp_x) (void)hipFree(h->mlp_x);
    if (h->mlp_t) (void)hipFree(h->mlp_t);
    h->mlp_x = h->mlp_t = nullptr;
    h->mlp_n = 0;
    HIPCHK(h, hipMalloc(&h->mlp_x, sizeof(float) * (size_t)n * kMlpIn));
    HIPCHK(h, hipMalloc(&h->mlp_t, sizeof(float) * (size_t)n * kMlpOut));
    return launch_fill_u32(h, h->mlp_fail, 1, 0u);
}

// After a pairs kernel: the standardisation from (up to) the first 2^18
// pairs, with ranges_from_data also the feature ranges the deployed network
// accepts (the sampled ranges otherwise); copies to x, t when not NULL.
int mlp_pairs_finish(IceHandle *h, uint32_t n, float *x, float *t, uint32_t *n_failed, bool ranges_from_data) {
    h->mlp_n = n;
    // standardisation from (up to) the first 2^18 pairs
    const uint32_t ns = std::min<uint32_t>(n, 1u << 18);
    std::vector<float> hx((size_t)ns * kMlpIn), ht((size_t)ns * kMlpOut);
    HIPCHK(h, hipMemcpyAsync(hx.data(), h->mlp_x, hx.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(ht.data(), h->mlp_t, ht.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    uint32_t failed = 0;
    HIPCHK(h, hipMemcpyAsync(&failed, h->mlp_fail, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    auto stats = [&](const std::vector<float> &v, int w, int k, float &m, float &sd) {
        double a = 0, b = 0;
        for (uint32_t i = 0; i < ns; i++) a += v[(size_t)i * w + k];
        a /= ns;
        for (uint32_t i = 0; i < ns; i++) {
            const double d = v[(size_t)i * w + k] - a;
            b += d * d;
        }
        m = (float)a;
        sd = (float)std::max(std::sqrt(b / ns), 1e-12);
    };
    for (int k = 0; k < kMlpIn; k++) stats(hx, kMlpIn, k, h->mlp_norm.xm[k], h->mlp_norm.xs[k]);
    for (int k = 0; k < kMlpOut; k++) stats(ht, kMlpOut, k, h->mlp_norm.ym[k], h->mlp_norm.ys[k]);
    h->mlp_have_norm = true;
    if (ranges_from_data) {
        // the span of the sampled features (log J, D, Jp_inv), to be accepted at deployment
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (uint32_t i = 0; i < ns; i++)
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], (double)hx[(size_t)i * kMlpIn + k]);
                hi[k] = std::max(hi[k], (double)hx[(size_t)i * kMlpIn + k]);
            }
        h->mlp_samp = IceMlpSampling{lo[0], hi[0], std::max(lo[1], 1e-12), std::max(hi[1], 2e-12), lo[2], hi[2]};
    }
    if (x) HIPCHK(h, hipMemcpy(x, h->mlp_x, sizeof(float) * (size_t)n * kMlpIn, hipMemcpyDeviceToHost));
    if (t) HIPCHK(h, hipMemcpy(t, h->mlp_t, sizeof(float) * (size_t)n * kMlpOut, hipMemcpyDeviceToHost));
    if (n_failed) *n_failed = failed;
    drop_graph(h);
    return ICE_OK;
}

}  // namespace

extern "C" {

int ice_mlp_set_norm(IceHandle *h, const IceMlpNorm *norm) {
    if (!h || !norm) return ICE_E_ARG;
    static_assert(sizeof(IceMlpNorm) == sizeof(MlpNorm), "norm layout");
    std::memcpy(&h->mlp_norm, norm, sizeof(MlpNorm));
    h->mlp_have_norm = true;
    drop_graph(h);
    return ICE_OK;
}

int ice_mlp_get_norm(IceHandle *h, IceMlpNorm *norm) {
    if (!h || !norm) return ICE_E_ARG;
    std::memcpy(norm, &h->mlp_norm, sizeof(MlpNorm));
    return ICE_OK;
}

int ice_mlp_set_params(IceHandle *h, const float *params) {
    if (!h || !params) return ICE_E_ARG;
    int rc = mlp_alloc(h);
    if (rc) return rc;
    HIPCHK(h, hipMemcpyAsync(h->mlp_p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, h->stream));
    if ((rc = launch_fill(h, h->mlp_m, kMlpParams)) || (rc = launch_fill(h, h->mlp_v, kMlpParams))) return rc;
    if ((rc = mlp_pack(h))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->mlp_adam_t = 0;
    h->mlp_have_params = true;
    return ICE_OK;
}

int ice_mlp_get_params(IceHandle *h, float *params) {
    if (!h || !params) return ICE_E_ARG;
    if (!h->mlp_p) { h->err = "no surrogate parameters"; return ICE_E_STATE; }
    HIPCHK(h, hipMemcpyAsync(params, h->mlp_p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ICE_OK;
}

int ice_mlp_init_params(IceHandle *h, uint64_t seed) {
    if (!h) return ICE_E_ARG;
    std::vector<float> p(kMlpParams, 0.f);
    uint64_t st = seed * 0x9e3779b97f4a7c15ull + 1;
    auto uni = [&]() {   // xorshift64*, uniform in [-1, 1)
        st ^= st >> 12;
        st ^= st << 25;
        st ^= st >> 27;
        return (double)((st * 0x2545f4914f6cdd1dull) >> 11) / 4503599627370496.0 - 1.0;
    };
    auto fill = [&](int off, int rows, int cols) {
        const double a = std::sqrt(6.0 / (rows + cols));   // Glorot-uniform
        for (int i = 0; i < rows * cols; i++) p[off + i] = (float)(a * uni());
    };
    fill(kOffW1, kMlpHid, kMlpIn);
    fill(kOffW2, kMlpHid, kMlpHid);
    fill(kOffW3, kMlpOut, kMlpHid);
    return ice_mlp_set_params(h, p.data());
}

int ice_mlp_train(IceHandle *h, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float *loss) {
    if (!h || batch == 0 || !(lr > 0)) return ICE_E_ARG;
    int rc = mlp_ready(h);
    if (rc) return rc;
    if (!h->mlp_n) { h->err = "no training pairs (ice_mlp_generate)"; return ICE_E_STATE; }
    HIPCHK(h, hipSetDevice(h->device));
    if (steps > h->mlp_loss_cap) {
        if (h->mlp_loss) (void)hipFree(h->mlp_loss);
        h->mlp_loss = nullptr;
        HIPCHK(h, hipMalloc(&h->mlp_loss, sizeof(float) * steps));
        h->mlp_loss_cap = steps;
    }
    if (steps && (rc = launch_fill(h, h->mlp_loss, steps))) return rc;
    const MlpDev M{h->mlp_pk, h->mlp_p, h->mlp_ipk};
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    // the gradient starts at zero; each Adam step clears what it consumed
    if (steps && (rc = launch_fill(h, h->mlp_g, kMlpParams))) return rc;
    for (uint32_t k = 0; k < steps; k++) {
        hipLaunchKernelGGL(k_mlp_train, dim3(blocks_for(batch, 64)), dim3(64), 0, h->stream, M, h->mlp_norm,
                           (const float *)h->mlp_x, (const float *)h->mlp_t, h->mlp_n, batch, seed, k, 0, h->mlp_g,
                           h->mlp_loss + k);
        LAUNCHCHK(h, "k_mlp_train");
        h->mlp_adam_t++;
        const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)h->mlp_adam_t)));
        const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)h->mlp_adam_t)));
        hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(kMlpParams, 256)), dim3(256), 0, h->stream, h->mlp_p,
                           h->mlp_g, h->mlp_m, h->mlp_v, lr, b1, b2, eps, c1, c2);
        LAUNCHCHK(h, "k_mlp_adam");
        if ((rc = mlp_pack(h, k + 1 == steps))) return rc;
    }
    if (loss && steps) HIPCHK(h, hipMemcpyAsync(loss, h->mlp_loss, sizeof(float) * steps, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graph(h);
    return ICE_OK;
}

int ice_mlp_forward(IceHandle *h, const float *x, uint32_t n, float *y) {
    if (!h || (!x && n) || (!y && n)) return ICE_E_ARG;
    int rc = mlp_ready(h);
    if (rc || n == 0) return rc;
    float *dx = nullptr, *dy = nullptr;
    if ((rc = mlp_upload(h, x, (size_t)n * kMlpIn, &dx))) return rc;
    hipError_t e = hipMalloc(&dy, sizeof(float) * (size_t)n * kMlpOut);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_infer, dim3(blocks_for(n, 64)), dim3(64), 0, h->stream, MlpDev{h->mlp_pk, h->mlp_p, h->mlp_ipk},
                           h->mlp_norm, (const float *)dx, n, dy);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, dy, sizeof(float) * (size_t)n * kMlpOut, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (e != hipSuccess) { h->err = std::string("ice_mlp_forward: ") + hipGetErrorString(e); return ICE_E_HIP; }
    return ICE_OK;
}

int ice_mlp_grad(IceHandle *h, const float *x, const float *t, uint32_t n, float *grad, float *loss) {
    if (!h || !x || !t || !grad || n == 0) return ICE_E_ARG;
    int rc = mlp_ready(h);
    if (rc) return rc;
    float *dx = nullptr, *dt = nullptr, *dl = nullptr;
    if ((rc = mlp_upload(h, x, (size_t)n * kMlpIn, &dx))) return rc;
    if ((rc = mlp_upload(h, t, (size_t)n * kMlpOut, &dt))) { (void)hipFree(dx); return rc; }
    hipError_t e = hipMalloc(&dl, sizeof(float));
    if (e == hipSuccess && (rc = launch_fill(h, h->mlp_g, kMlpParams)) == ICE_OK && (rc = launch_fill(h, dl, 1)) == ICE_OK) {
        hipLaunchKernelGGL(k_mlp_train, dim3(blocks_for(n, 64)), dim3(64), 0, h->stream, MlpDev{h->mlp_pk, h->mlp_p, h->mlp_ipk},
                           h->mlp_norm, (const float *)dx, (const float *)dt, n, n, (uint64_t)0, 0u, 1, h->mlp_g, dl);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(grad, h->mlp_g, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, dl, sizeof(float), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    (void)hipFree(dx);
    (void)hipFree(dt);
    if (dl) (void)hipFree(dl);
    if (rc) return rc;
    if (e != hipSuccess) { h->err = std::string("ice_mlp_grad: ") + hipGetErrorString(e); return ICE_E_HIP; }
    return ICE_OK;
}

int ice_mlp_eval(IceHandle *h, const float *x, uint32_t n, float tol, float *out) {
    if (!h || (!x && n) || (!out && n)) return ICE_E_ARG;
    int rc = mlp_ready(h);
    if (rc || n == 0) return rc;
    float *dx = nullptr, *dout = nullptr;
    if ((rc = mlp_upload(h, x, (size_t)n * kMlpIn, &dx))) return rc;
    hipError_t e = hipMalloc(&dout, sizeof(float) * (size_t)n * 6);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_eval, dim3(blocks_for(n, 64)), dim3(64), 0, h->stream, MlpDev{h->mlp_pk, h->mlp_p, h->mlp_ipk},
                           h->mlp_norm, h->dp, mlp_phys(h), (const float *)dx, n, tol, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(float) * (size_t)n * 6, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dx);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) { h->err = std::string("ice_mlp_eval: ") + hipGetErrorString(e); return ICE_E_HIP; }
    return ICE_OK;
}

}  // extern "C"

namespace {

// Staging of ice_mlp_eval_noalloc: ICE_MLP_EVAL_NOALLOC_MAX_ROWS rows of
// features (4 floats) then as many rows of results (6 floats), 2.6 MB.  Each
// GPU holds its own copy until the process ends; every call on a GPU shares
// it (calls that overlap on one GPU race).
constexpr size_t kEvalRows = ICE_MLP_EVAL_NOALLOC_MAX_ROWS;
__device__ __align__(256) float scratch[kEvalRows * (kMlpIn + 6)];

// The array's device address on each GPU, looked up once per device with that
// device current (host code never uses the symbol's own name as an address).
constexpr int kMaxDevices = 64;
float *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_for(IceHandle *h, float **out) {
    if (h->device < 0 || h->device >= kMaxDevices) { h->err = "device index beyond the scratch table"; return ICE_E_DEVICE; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[h->device]) {
        HIPCHK(h, hipSetDevice(h->device));
        void *p = nullptr;
        HIPCHK(h, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[h->device] = static_cast<float *>(p);
    }
    *out = g_scratch[h->device];
    return ICE_OK;
}

}  // namespace

extern "C" {

int ice_mlp_eval_noalloc(IceHandle *h, const float *x, uint32_t n, float tol, float *out) {
    if (!h || (!x && n) || (!out && n)) return ICE_E_ARG;
    if (n > kEvalRows) {
        h->err = "ice_mlp_eval_noalloc: more than ICE_MLP_EVAL_NOALLOC_MAX_ROWS (65536) rows";
        return ICE_E_CAPACITY;
    }
    int rc = mlp_ready(h);
    if (rc || n == 0) return rc;
    float *S = nullptr;
    if ((rc = scratch_for(h, &S))) return rc;
    HIPCHK(h, hipSetDevice(h->device));
    float *dx = S, *dout = S + kEvalRows * kMlpIn;
    HIPCHK(h, hipMemcpyAsync(dx, x, sizeof(float) * (size_t)n * kMlpIn, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_mlp_eval, dim3(blocks_for(n, 64)), dim3(64), 0, h->stream,
                       MlpDev{h->mlp_pk, h->mlp_p, h->mlp_ipk}, h->mlp_norm, h->dp, mlp_phys(h), (const float *)dx, n,
                       tol, dout);
    LAUNCHCHK(h, "k_mlp_eval");
    HIPCHK(h, hipMemcpyAsync(out, dout, sizeof(float) * (size_t)n * 6, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ICE_OK;
}

int ice_mlp_deploy(IceHandle *h, int enable, float tol) {
    if (!h || !(tol >= 0)) return ICE_E_ARG;
    if (enable) {
        int rc = mlp_ready(h);
        if (rc) return rc;
    }
    if (h->ndom() > 1 && h->mlp_deployed != (enable != 0)) h->tables_dirty = true;   // tile order
    h->mlp_deployed = enable != 0;
    h->mlp_tol = tol;
    static const uint32_t zero2[2] = {0u, 0u};   // outlives the asynchronous copies
    for (int d = 0; d < h->ndom(); d++)
        HIPCHK(h, hipMemcpyAsync(&h->dcnt(d)->pad[0], zero2, sizeof(zero2), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graph(h);
    return ICE_OK;
}

int ice_mlp_counts(IceHandle *h, uint64_t counts[2]) {
    if (!h || !counts) return ICE_E_ARG;
    int rc = read_counters(h);
    if (rc) return rc;
    counts[0] = counts[1] = 0;
    for (int d = 0; d < h->ndom(); d++) {
        counts[0] += h->host_cnt[d].pad[0];
        counts[1] += h->host_cnt[d].pad[1];
    }
    return ICE_OK;
}

}  // extern "C"
