// icempm_abi.hip — the C-ABI of include/icempm.h on one MI355X (gfx950).
//
// One IceHandle = one GPU_Partition of the reference (gpu_partition.h:63-134):
// it owns the device SoA (two buffers, for the sort), the grid lines, the
// status bytes, the stream and the timing events.  Nothing here throws across
// the ABI; failures return ICE_E_* and leave a message in ice_last_error.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

#include "../../include/icempm.h"
#include "device_math.hpp"
#include "point_ops.hpp"
#include "kernels_naive.hpp"
#include "kernels_tiled.hpp"
#include "kernels_sort.hpp"
#include "kernels_p2g_cells.hpp"

using namespace ice;

namespace {

std::string g_create_error;

constexpr int kGridLines = 11;   // m px py nx ny cur_x cur_y fx fy wind_x wind_y
enum { G_M = 0, G_PX, G_PY, G_NX, G_NY, G_CX, G_CY, G_FX, G_FY, G_WX, G_WY };
constexpr int kChunk = 16;       // substeps per captured graph

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct TimingSet {
    hipEvent_t e[5];
};

}  // namespace

struct IceHandle {
    IceParams prm{};
    DevParams dp{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // points
    float *pts = nullptr, *pts_alt = nullptr;
    size_t pitch = 0;
    uint32_t capacity = 0, n = 0;
    int nlines = ICE_N_LINES;
    // grid
    float *grid = nullptr;
    size_t gpitch = 0;
    uint8_t *status = nullptr;
    bool has_current = false;
    Counters *cnt = nullptr;
    uint32_t *offtile = nullptr;
    // sort scratch
    uint32_t *s_count = nullptr, *s_cs = nullptr, *s_key = nullptr, *s_slot = nullptr, *s_dflag = nullptr,
             *s_drank = nullptr, *s_order = nullptr, *s_bsum = nullptr;
    size_t s_bsum_len = 0;
    bool sorted = false;
    int steps_since_sort = 0, n_sorts = 0;
    // tiles
    int tile_w = 32, tile_h = 32;
    int p2g_point_lanes = 0;       // 1: one lane per point in the tiled P2G (the default is one per cell)
    int band = 4;                  // tile columns per band of the workgroup -> tile order
    FixScale fix{};
    // phase state
    double pending_t = 0.0;
    double last_t = 0.0;
    // timings
    bool timing = false;
    std::vector<TimingSet> ev_pool;
    size_t ev_used = 0;
    double acc[8] = {0};
    // graph
    hipGraphExec_t graph_exec = nullptr;
    int graph_chunk = 0;
    uint32_t *host_cnt = nullptr;   // pinned copy of Counters

    size_t N() const { return (size_t)prm.GridXTotal * prm.GridYTotal; }
    float *gline(int l) const { return grid + (size_t)l * gpitch; }
    Lines lines() const { return Lines{pts, pitch}; }
    bool naive() const { return (prm.flags & ICE_FLAG_NAIVE) != 0; }
    bool det() const { return (prm.flags & ICE_FLAG_DETERMINISTIC) != 0; }
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

DevParams make_dev_params(const IceParams &p, bool has_current) {
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
    d.has_current = has_current ? 1 : 0;
    d.update_every = p.UpdateEveryNthStep;
    d.dt_d = p.InitialTimeStep;
    return d;
}

// Fixed-point scales for the deterministic P2G: each node sum must stay below
// 2^62 for up to 4096 point-equivalents of the heaviest point at |v| = vmax,
// plus a stress term at the compressive strength.
FixScale make_fix_scale(const IceParams &p, float max_thickness) {
    const double mp = p.ParticleMass * std::max(1.0f, max_thickness);
    const double m_bound = 4096.0 * mp;
    const double stress = p.dt_vol_Dpinv * 4.0 * std::max(p.IceCompressiveStrength, p.kappa) * p.cellsize;
    const double p_bound = m_bound * p.vmax * 2.0 + 4096.0 * stress * std::max(1.0f, max_thickness);
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

using P2GFn = void (*)(DevParams, Lines, const uint32_t *, float *, float *, float *, Counters *, int, int, FixScale);
using P2GCellFn = void (*)(DevParams, Lines, const uint32_t *, float *, float *, float *, Counters *, int, int,
                           FixScale, const uint32_t *, int);
using G2PFn = void (*)(DevParams, StepArg, Lines, const uint32_t *, GridIn, float *, float *, Counters *, uint32_t *,
                       int, int, int);

struct TileKernels {
    int w, h;
    P2GFn p2g, p2g_det;            // one lane per point
    P2GCellFn p2gc, p2gc_det;      // one lane per cell
    G2PFn g2p;
};

#define TK(W, H)                                                                                      \
    {W, H, k_p2g_tiled<W, H, false>, k_p2g_tiled<W, H, true>, k_p2g_cells<W, H, false>,             \
     k_p2g_cells<W, H, true>, k_g2p_tiled<W, H>}
const TileKernels kTiles[] = {TK(8, 8),   TK(8, 16),  TK(16, 8),  TK(16, 16), TK(16, 32),
                              TK(32, 16), TK(32, 32), TK(32, 64), TK(64, 32), TK(64, 64)};
#undef TK

const TileKernels *find_tile(int w, int h) {
    for (const auto &t : kTiles)
        if (t.w == w && t.h == h) return &t;
    return nullptr;
}

int scan_u32(IceHandle *h, const uint32_t *in, size_t n, uint32_t *out) {
    const uint32_t nb = (uint32_t)std::max<size_t>(1, (n + kScanTile - 1) / kScanTile);
    if ((size_t)nb + 1 > h->s_bsum_len) {
        h->err = "scan scratch too small";
        return ICE_E_STATE;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanThreads), 0, h->stream, in, n, h->s_bsum);
    LAUNCHCHK(h, "k_scan_reduce");
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanThreads), 0, h->stream, h->s_bsum, nb);
    LAUNCHCHK(h, "k_scan_blocks");
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanThreads), 0, h->stream, in, n, h->s_bsum, nb, out);
    LAUNCHCHK(h, "k_scan_apply");
    return ICE_OK;
}

void drop_graph(IceHandle *h);

// Sort points into cell order; keep disabled points behind (drop = false) or
// drop them (drop = true).  Leaves cs valid and the off-tile list empty.
int device_sort(IceHandle *h, bool drop, uint32_t *perm_dev) {
    const uint32_t n = h->n;
    const uint32_t NC = (uint32_t)h->N();
    int rc;
    if ((rc = launch_fill_u32(h, h->s_count, (size_t)NC + 1, 0u))) return rc;
    if (n) {
        hipLaunchKernelGGL(k_sort_keys, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->lines(), n,
                           h->prm.GridXTotal, h->prm.GridYTotal, NC, h->s_count, h->s_key, h->s_slot, h->s_dflag,
                           h->cnt);
        LAUNCHCHK(h, "k_sort_keys");
    }
    if ((rc = scan_u32(h, h->s_count, NC, h->s_cs))) return rc;
    if ((rc = scan_u32(h, h->s_dflag, n, h->s_drank))) return rc;
    if (n) {
        hipLaunchKernelGGL(k_sort_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, n, NC, h->s_key,
                           h->s_slot, h->s_cs, h->s_drank, h->s_order);
        LAUNCHCHK(h, "k_sort_scatter");
    }
    hipLaunchKernelGGL(k_sort_cells, dim3(blocks_for(NC, 256)), dim3(256), 0, h->stream, NC, h->s_cs, h->s_order);
    LAUNCHCHK(h, "k_sort_cells");
    uint32_t n_active = 0;
    HIPCHK(h, hipMemcpyAsync(&n_active, h->s_cs + NC, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const uint32_t n_out = drop ? n_active : n;
    if (n_out) {
        Lines dst{h->pts_alt, h->pitch};
        hipLaunchKernelGGL(k_sort_gather, dim3(blocks_for(n_out, 256)), dim3(256), 0, h->stream, h->lines(), dst,
                           h->nlines, n_out, h->s_order, perm_dev);
        LAUNCHCHK(h, "k_sort_gather");
    }
    std::swap(h->pts, h->pts_alt);
    h->n = n_out;
    drop_graph(h);   // a captured graph holds the old point buffer
    if (drop) {
        const int32_t zero = 0;
        HIPCHK(h, hipMemcpyAsync(&h->cnt->disabled, &zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    const uint32_t zero_u = 0;
    HIPCHK(h, hipMemcpyAsync(&h->cnt->n_esc, &zero_u, sizeof(zero_u), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->sorted = true;
    h->steps_since_sort = 0;
    h->n_sorts++;
    return ICE_OK;
}

GridIn grid_in(const IceHandle *h) {
    GridIn G;
    G.m = h->gline(G_M);
    G.px = h->gline(G_PX);
    G.py = h->gline(G_PY);
    G.cur_x = h->gline(G_CX);
    G.cur_y = h->gline(G_CY);
    G.wind_x = h->gline(G_WX);
    G.wind_y = h->gline(G_WY);
    G.status = h->status;
    return G;
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

unsigned offtile_blocks(const IceHandle *h) { return std::min<unsigned>(512, blocks_for(std::max<uint32_t>(h->n, 1), 256)); }

// Phases of one substep.  Event slots: 0 start, 1 after the main P2G kernel,
// 2 after P2G, 3 before the main G2P kernel, 4 after G2P.
int phase_p2g(IceHandle *h) {
    if (!h->pts) { h->err = "no points allocated"; return ICE_E_STATE; }
    const DevParams &P = h->dp;
    if (h->naive()) {
        int rc = launch_fill(h, h->grid, 3 * h->gpitch);   // reset_grid (gpu_partition.cpp:300-308)
        if (rc) return rc;
        if (h->n) {
            hipLaunchKernelGGL(k_p2g_naive, dim3(blocks_for(h->n, 256)), dim3(256), 0, h->stream, P, h->lines(), h->n,
                               h->gline(G_M), h->gline(G_PX), h->gline(G_PY), h->cnt);
            LAUNCHCHK(h, "k_p2g_naive");
        }
        rec(h, 1);
        rec(h, 2);
        return ICE_OK;
    }
    if (!h->sorted) {
        int rc = device_sort(h, false, nullptr);
        if (rc) return rc;
    }
    const TileKernels *tk = find_tile(h->tile_w, h->tile_h);
    const int ntx = (P.gx + h->tile_w - 1) / h->tile_w, nty = (P.gy + h->tile_h - 1) / h->tile_h;
    if (h->p2g_point_lanes && !h->det()) {
        hipLaunchKernelGGL(tk->p2g, dim3(ntx * nty), dim3(kTileThreads), 0, h->stream, P, h->lines(), h->s_cs,
                           h->gline(G_M), h->gline(G_PX), h->gline(G_PY), h->cnt, ntx, nty, h->fix);
    } else {
        P2GCellFn fn = h->det() ? tk->p2gc_det : tk->p2gc;
        hipLaunchKernelGGL(fn, dim3(ntx * nty), dim3(kTileThreads), 0, h->stream, P, h->lines(), h->s_cs,
                           h->gline(G_M), h->gline(G_PX), h->gline(G_PY), h->cnt, ntx, nty, h->fix,
                           (const uint32_t *)h->offtile, h->band);
    }
    LAUNCHCHK(h, "k_p2g_tiled");
    rec(h, 1);
    if (!h->det()) {   // the deterministic tile kernel adds the off-tile points itself
        hipLaunchKernelGGL(k_p2g_offtile, dim3(offtile_blocks(h)), dim3(256), 0, h->stream, P, h->lines(),
                           h->offtile, h->gline(G_M), h->gline(G_PX), h->gline(G_PY), h->cnt);
        LAUNCHCHK(h, "k_p2g_offtile");
    }
    rec(h, 2);
    return ICE_OK;
}

int phase_update(IceHandle *h, const StepArg &S) {
    if (h->naive()) {
        const size_t N = h->N();
        hipLaunchKernelGGL(k_update_naive, dim3(blocks_for(N, 256)), dim3(256), 0, h->stream, h->dp, S, h->gline(G_M),
                           h->gline(G_PX), h->gline(G_PY), h->status, h->gline(G_CX), h->gline(G_CY), h->gline(G_WX),
                           h->gline(G_WY), h->gline(G_FX), h->gline(G_FY), h->cnt);
        LAUNCHCHK(h, "k_update_naive");
    }
    return ICE_OK;
}

int phase_g2p(IceHandle *h, const StepArg &S) {
    const DevParams &P = h->dp;
    if (h->naive()) {
        rec(h, 3);
        if (h->n) {
            hipLaunchKernelGGL(k_g2p_naive, dim3(blocks_for(h->n, 256)), dim3(256), 0, h->stream, P, S, h->lines(),
                               h->n, (const float *)h->gline(G_PX), (const float *)h->gline(G_PY), h->cnt);
            LAUNCHCHK(h, "k_g2p_naive");
        }
        rec(h, 4);
        return ICE_OK;
    }
    const GridIn G = grid_in(h);
    hipLaunchKernelGGL(k_g2p_offtile, dim3(offtile_blocks(h)), dim3(256), 0, h->stream, P, S, h->lines(), h->offtile, G,
                       h->cnt);
    LAUNCHCHK(h, "k_g2p_offtile");
    rec(h, 3);
    const TileKernels *tk = find_tile(h->tile_w, h->tile_h);
    const int ntx = (P.gx + h->tile_w - 1) / h->tile_w, nty = (P.gy + h->tile_h - 1) / h->tile_h;
    hipLaunchKernelGGL(tk->g2p, dim3(ntx * nty), dim3(kTileThreads), 0, h->stream, P, S, h->lines(), h->s_cs, G,
                       h->gline(G_FX), h->gline(G_FY), h->cnt, h->offtile, ntx, nty, h->band);
    LAUNCHCHK(h, "k_g2p_tiled");
    rec(h, 4);
    h->steps_since_sort++;
    return ICE_OK;
}

void end_timing_set(IceHandle *h) {
    if (h->timing && h->ev_used < h->ev_pool.size()) h->ev_used++;
}

int substep_args(IceHandle *h, const StepArg &S) {
    rec(h, 0);
    int rc = phase_p2g(h);
    if (rc) return rc;
    if ((rc = phase_update(h, S))) return rc;
    if ((rc = phase_g2p(h, S))) return rc;
    end_timing_set(h);
    return ICE_OK;
}

int read_counters(IceHandle *h, Counters &c) {
    HIPCHK(h, hipMemcpyAsync(h->host_cnt, h->cnt, sizeof(Counters), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::memcpy(&c, h->host_cnt, sizeof(Counters));
    return ICE_OK;
}

void drop_graph(IceHandle *h) {
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    h->graph_exec = nullptr;
    h->graph_chunk = 0;
}

int build_graph(IceHandle *h, int chunk) {
    drop_graph(h);
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const bool was_timing = h->timing;
    h->timing = false;
    int rc = ICE_OK;
    for (int k = 0; k < chunk && rc == ICE_OK; k++) {
        StepArg S{0.0, 0, &h->cnt->step, k};
        rc = substep_args(h, S);
    }
    if (rc == ICE_OK) {
        hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, h->stream, h->cnt, chunk);
    }
    h->timing = was_timing;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) { h->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e); return ICE_E_HIP; }
    e = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { h->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e); return ICE_E_HIP; }
    h->graph_chunk = chunk;
    return ICE_OK;
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
    if (!p || !out) { g_create_error = "null argument"; return ICE_E_ARG; }
    *out = nullptr;
    if (p->GridXTotal < 5 || p->GridYTotal < 5 || p->GridXTotal > 65535 || p->GridYTotal > 65535) {
        g_create_error = "grid size out of range (5..65535 nodes per axis; the cell word holds 16 bits per axis)";
        return ICE_E_ARG;
    }
    if ((uint64_t)p->GridXTotal * p->GridYTotal >= 0xffffffffull) { g_create_error = "grid too large"; return ICE_E_ARG; }
    if (!(p->InitialTimeStep > 0) || !(p->cellsize > 0)) { g_create_error = "dt and cellsize must be > 0"; return ICE_E_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_create_error = "no HIP device"; return ICE_E_DEVICE; }
    if (device < 0 || device >= ndev) { g_create_error = "device index out of range"; return ICE_E_DEVICE; }
    IceHandle *h = new IceHandle;
    h->prm = *p;
    derive_params(h->prm);
    h->device = device;
    auto fail = [&](const std::string &m, int code) {
        g_create_error = m;
        ice_destroy(h);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return fail("hipSetDevice failed", ICE_E_HIP);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail("stream", ICE_E_HIP);
    const size_t N = h->N();
    h->gpitch = round_up(N, 64);
    if (hipMalloc(&h->grid, sizeof(float) * h->gpitch * kGridLines) != hipSuccess) return fail("grid alloc", ICE_E_HIP);
    if (hipMalloc(&h->status, round_up(N, 256)) != hipSuccess) return fail("status alloc", ICE_E_HIP);
    if (hipMalloc(&h->cnt, sizeof(Counters)) != hipSuccess) return fail("counter alloc", ICE_E_HIP);
    if (hipHostMalloc(&h->host_cnt, sizeof(Counters)) != hipSuccess) return fail("pinned alloc", ICE_E_HIP);
    if (hipMalloc(&h->s_count, sizeof(uint32_t) * (N + 1)) != hipSuccess) return fail("sort alloc", ICE_E_HIP);
    if (hipMalloc(&h->s_cs, sizeof(uint32_t) * (N + 2)) != hipSuccess) return fail("sort alloc", ICE_E_HIP);
    if (hipMemsetAsync(h->grid, 0, sizeof(float) * h->gpitch * kGridLines, h->stream) != hipSuccess) return fail("memset", ICE_E_HIP);
    if (hipMemsetAsync(h->status, ICE_GRID_STATUS_WATER, round_up(N, 256), h->stream) != hipSuccess) return fail("memset", ICE_E_HIP);
    if (hipMemsetAsync(h->cnt, 0, sizeof(Counters), h->stream) != hipSuccess) return fail("memset", ICE_E_HIP);
    if (hipMemsetAsync(h->s_cs, 0, sizeof(uint32_t) * (N + 2), h->stream) != hipSuccess) return fail("memset", ICE_E_HIP);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return fail("sync", ICE_E_HIP);
    h->dp = make_dev_params(h->prm, false);
    h->fix = make_fix_scale(h->prm, 1.f);
    *out = h;
    return ICE_OK;
}

int ice_destroy(IceHandle *h) {
    if (!h) return ICE_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    for (auto &t : h->ev_pool)
        for (auto &e : t.e) (void)hipEventDestroy(e);
    for (void *p : {(void *)h->pts, (void *)h->pts_alt, (void *)h->grid, (void *)h->status, (void *)h->cnt,
                    (void *)h->offtile, (void *)h->s_count, (void *)h->s_cs, (void *)h->s_key, (void *)h->s_slot,
                    (void *)h->s_dflag, (void *)h->s_drank, (void *)h->s_order, (void *)h->s_bsum})
        if (p) (void)hipFree(p);
    if (h->host_cnt) (void)hipHostFree(h->host_cnt);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return ICE_OK;
}

int ice_allocate(IceHandle *h, uint32_t cap) {
    if (!h) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graph(h);
    for (void **p : {(void **)&h->pts, (void **)&h->pts_alt, (void **)&h->offtile, (void **)&h->s_key,
                     (void **)&h->s_slot, (void **)&h->s_dflag, (void **)&h->s_drank, (void **)&h->s_order,
                     (void **)&h->s_bsum}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    h->capacity = cap;
    h->n = 0;
    h->sorted = false;
    const size_t c = std::max<uint32_t>(cap, 1);
    h->pitch = round_up(c, 64);
    const size_t bytes = sizeof(float) * h->pitch * ICE_N_LINES_MAX;
    HIPCHK(h, hipMalloc(&h->pts, bytes));
    HIPCHK(h, hipMalloc(&h->pts_alt, bytes));
    HIPCHK(h, hipMemsetAsync(h->pts, 0, bytes, h->stream));
    HIPCHK(h, hipMalloc(&h->offtile, sizeof(uint32_t) * c));
    HIPCHK(h, hipMalloc(&h->s_key, sizeof(uint32_t) * c));
    HIPCHK(h, hipMalloc(&h->s_slot, sizeof(uint32_t) * c));
    HIPCHK(h, hipMalloc(&h->s_dflag, sizeof(uint32_t) * c));
    HIPCHK(h, hipMalloc(&h->s_drank, sizeof(uint32_t) * (c + 1)));
    HIPCHK(h, hipMalloc(&h->s_order, sizeof(uint32_t) * c));
    h->s_bsum_len = std::max(c, h->N()) / kScanTile + 8;
    HIPCHK(h, hipMalloc(&h->s_bsum, sizeof(uint32_t) * h->s_bsum_len));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ICE_OK;
}

int ice_upload_points(IceHandle *h, const float *soa, size_t host_pitch, uint32_t n, int n_lines) {
    if (!h || (!soa && n)) return ICE_E_ARG;
    if (!h->pts) { h->err = "ice_allocate first"; return ICE_E_STATE; }
    if (n > h->capacity) { h->err = "more points than the allocated capacity"; return ICE_E_CAPACITY; }
    if (n_lines < ICE_N_LINES || n_lines > ICE_N_LINES_MAX || host_pitch < n) { h->err = "bad layout"; return ICE_E_ARG; }
    if ((h->prm.flags & ICE_FLAG_COHESION_LINE) && n_lines < ICE_N_LINES_MAX) {
        h->err = "ICE_FLAG_COHESION_LINE needs 20 lines";
        return ICE_E_ARG;
    }
    HIPCHK(h, hipSetDevice(h->device));
    h->nlines = (h->prm.flags & ICE_FLAG_COHESION_LINE) ? ICE_N_LINES_MAX : ICE_N_LINES;
    if (n) HIPCHK(h, hipMemcpy2DAsync(h->pts, h->pitch * sizeof(float), soa, host_pitch * sizeof(float),
                                      n * sizeof(float), h->nlines, hipMemcpyHostToDevice, h->stream));
    h->n = n;
    // the partition's disabled count starts at 0 (gpu_implementation5.cpp:137, gpu_partition.cpp:101-103)
    const int32_t zero = 0;
    HIPCHK(h, hipMemcpyAsync(&h->cnt->disabled, &zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
    float tmax = 1.f;
    if (n) {
        const float *th = soa + (size_t)ICE_IDX_THICKNESS * host_pitch;
        for (uint32_t i = 0; i < n; i++) tmax = std::max(tmax, th[i]);
    }
    h->fix = make_fix_scale(h->prm, tmax);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->sorted = false;
    drop_graph(h);
    if (!h->naive()) return device_sort(h, false, nullptr);
    return ICE_OK;
}

int ice_upload_grid(IceHandle *h, const uint8_t *status, const float *normals) {
    if (!h || !status) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    const size_t N = h->N();
    HIPCHK(h, hipMemcpyAsync(h->status, status, N, hipMemcpyHostToDevice, h->stream));
    if (normals) {
        HIPCHK(h, hipMemcpyAsync(h->gline(G_NX), normals, N * sizeof(float), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->gline(G_NY), normals + N, N * sizeof(float), hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return ICE_OK;
}

int ice_upload_forcing(IceHandle *h, const float *current, const float *wind) {
    if (!h) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    const size_t N = h->N();
    int rc;
    if (current) {
        HIPCHK(h, hipMemcpyAsync(h->gline(G_CX), current, N * sizeof(float), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->gline(G_CY), current + N, N * sizeof(float), hipMemcpyHostToDevice, h->stream));
        bool any = false;
        for (size_t i = 0; i < 2 * N && !any; i++) any = current[i] != 0.f;
        h->has_current = any;
    } else {
        if ((rc = launch_fill(h, h->gline(G_CX), 2 * h->gpitch))) return rc;
        h->has_current = false;
    }
    if (wind) {
        HIPCHK(h, hipMemcpyAsync(h->gline(G_WX), wind, N * sizeof(float), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->gline(G_WY), wind + N, N * sizeof(float), hipMemcpyHostToDevice, h->stream));
    } else {
        if ((rc = launch_fill(h, h->gline(G_WX), 2 * h->gpitch))) return rc;
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->dp = make_dev_params(h->prm, h->has_current);
    drop_graph(h);
    return ICE_OK;
}

int ice_update_constants(IceHandle *h, const IceParams *p) {
    if (!h || !p) return ICE_E_ARG;
    if (p->GridXTotal != h->prm.GridXTotal || p->GridYTotal != h->prm.GridYTotal) {
        h->err = "grid size cannot change";
        return ICE_E_ARG;
    }
    const uint32_t old_flags = h->prm.flags;
    h->prm = *p;
    derive_params(h->prm);
    if ((old_flags ^ h->prm.flags) & ICE_FLAG_NAIVE) h->sorted = false;
    h->dp = make_dev_params(h->prm, h->has_current);
    // GPU_Partition::update_constants zeroes the error word (gpu_partition.cpp:143)
    const uint32_t zero = 0;
    HIPCHK(h, hipMemcpyAsync(&h->cnt->err, &zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    drop_graph(h);
    return ICE_OK;
}

int ice_set_tile(IceHandle *h, int tw, int th) {
    if (!h) return ICE_E_ARG;
    if (!find_tile(tw, th)) { h->err = "unsupported tile shape"; return ICE_E_ARG; }
    h->tile_w = tw;
    h->tile_h = th;
    drop_graph(h);
    return ICE_OK;
}

int ice_set_option(IceHandle *h, const char *name, int value) {
    if (!h || !name) return ICE_E_ARG;
    if (std::strcmp(name, "p2g_point_lanes") == 0) {
        h->p2g_point_lanes = value != 0;
    } else if (std::strcmp(name, "band") == 0) {
        if (value < 1 || value > 4096) { h->err = "band must be 1..4096"; return ICE_E_ARG; }
        h->band = value;
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
    // The tiled P2G stores every node, so only the naive path clears here.
    return ICE_OK;
}

int ice_p2g(IceHandle *h) {
    if (!h) return ICE_E_ARG;
    return phase_p2g(h);
}

int ice_update_nodes(IceHandle *h, float simulation_time, float windSpeed, float windAngle) {
    if (!h) return ICE_E_ARG;
    (void)windSpeed;   // uniform wind of the reference's signature; per-node stress comes from ice_upload_forcing
    (void)windAngle;
    h->pending_t = simulation_time;
    StepArg S{(double)simulation_time, 0, nullptr, 0};
    return phase_update(h, S);
}

int ice_g2p(IceHandle *h, int recordPQ, int enablePointTransfer, int applyGlensLaw) {
    if (!h) return ICE_E_ARG;
    (void)enablePointTransfer;   // single partition: no point transfer
    (void)applyGlensLaw;         // Glen's law is commented out in the reference (kernels.cuh:625-652)
    StepArg S{h->pending_t, recordPQ, nullptr, 0};
    h->last_t = h->pending_t;
    int rc = phase_g2p(h, S);
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
    if (!h->naive() && !h->sorted) {
        int rc = device_sort(h, false, nullptr);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, h->stream, h->cnt, first_step);
    LAUNCHCHK(h, "k_step_set");
    int done = 0;
    const uint32_t resort_at = std::max<uint32_t>(256, h->n / 256);
    while (done < nsteps) {
        const int chunk = std::min(kChunk, nsteps - done);
        if (use_graph && chunk == kChunk && !h->timing) {
            if (!h->graph_exec || h->graph_chunk != chunk) {
                int rc = build_graph(h, chunk);
                if (rc) return rc;
            }
            HIPCHK(h, hipGraphLaunch(h->graph_exec, h->stream));
        } else {
            for (int k = 0; k < chunk; k++) {
                StepArg S{0.0, 0, &h->cnt->step, k};
                int rc = substep_args(h, S);
                if (rc) return rc;
            }
            hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, h->stream, h->cnt, chunk);
            LAUNCHCHK(h, "k_step_advance");
        }
        done += chunk;
        if (!h->naive()) {
            Counters c;
            int rc = read_counters(h, c);
            if (rc) return rc;
            if (c.n_esc > resort_at) {
                if ((rc = device_sort(h, false, nullptr))) return rc;
            }
        }
    }
    h->last_t = h->prm.InitialTimeStep * (first_step + nsteps - 1);
    return ICE_OK;
}

int ice_clear_force_accumulator(IceHandle *h) {
    if (!h) return ICE_E_ARG;
    return launch_fill(h, h->gline(G_FX), 2 * h->gpitch);
}

int ice_download(IceHandle *h, float *soa, size_t host_pitch, int n_lines, float *forces, uint32_t *error_code,
                 int32_t *disabled_count) {
    if (!h) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    const size_t N = h->N();
    if (soa) {
        if (n_lines < ICE_N_LINES || n_lines > ICE_N_LINES_MAX || host_pitch < h->n) { h->err = "bad layout"; return ICE_E_ARG; }
        const int nl = std::min(n_lines, h->nlines);
        if (h->n)
            HIPCHK(h, hipMemcpy2DAsync(soa, host_pitch * sizeof(float), h->pts, h->pitch * sizeof(float),
                                       h->n * sizeof(float), nl, hipMemcpyDeviceToHost, h->stream));
    }
    if (forces) {
        HIPCHK(h, hipMemcpyAsync(forces, h->gline(G_FX), N * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(forces + N, h->gline(G_FY), N * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    }
    Counters c;
    int rc = read_counters(h, c);
    if (rc) return rc;
    if (soa) {
        uint32_t *u = reinterpret_cast<uint32_t *>(soa + (size_t)ICE_IDX_UTILITY * host_pitch);
        for (uint32_t i = 0; i < h->n; i++) u[i] &= ~kInternalBits;
    }
    if (error_code) *error_code = c.err;
    if (disabled_count) *disabled_count = c.disabled;
    return ICE_OK;
}

int ice_download_grid(IceHandle *h, float *mpp, float *velocity) {
    if (!h) return ICE_E_ARG;
    HIPCHK(h, hipSetDevice(h->device));
    const size_t N = h->N();
    if (mpp)
        HIPCHK(h, hipMemcpy2DAsync(mpp, N * sizeof(float), h->grid, h->gpitch * sizeof(float), N * sizeof(float), 3,
                                   hipMemcpyDeviceToHost, h->stream));
    if (velocity) {
        if (h->naive()) {
            // the naive update kernel leaves velocities in (px, py)
            HIPCHK(h, hipMemcpy2DAsync(velocity, N * sizeof(float), h->gline(G_PX), h->gpitch * sizeof(float),
                                       N * sizeof(float), 2, hipMemcpyDeviceToHost, h->stream));
        } else {
            float *buf = nullptr;
            HIPCHK(h, hipMalloc(&buf, 2 * h->gpitch * sizeof(float)));
            hipLaunchKernelGGL(k_node_velocity_out, dim3(1024), dim3(256), 0, h->stream, h->dp, h->last_t, grid_in(h),
                               buf, buf + h->gpitch);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpy2DAsync(velocity, N * sizeof(float), buf, h->gpitch * sizeof(float), N * sizeof(float), 2,
                                     hipMemcpyDeviceToHost, h->stream);
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
    const uint32_t n_before = h->n;
    if (perm && n_before) HIPCHK(h, hipMalloc(&perm_dev, sizeof(uint32_t) * n_before));
    int rc = device_sort(h, true, perm_dev);
    if (rc == ICE_OK && perm && h->n) {
        hipError_t e = hipMemcpy(perm, perm_dev, sizeof(uint32_t) * h->n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) { h->err = hipGetErrorString(e); rc = ICE_E_HIP; }
    }
    if (perm_dev) (void)hipFree(perm_dev);
    drop_graph(h);
    if (n_out) *n_out = h->n;
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
        HIPCHK(h, hipEventElapsedTime(&ms[3], t.e[0], t.e[4]));   // substep
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

int ice_copy_bench(int device, size_t bytes, int reps, double *ms_per_copy) {
    if (!ms_per_copy || reps <= 0 || bytes < 1024) return ICE_E_ARG;
    *ms_per_copy = 0.0;
    if (hipSetDevice(device) != hipSuccess) { g_create_error = "hipSetDevice failed"; return ICE_E_DEVICE; }
    const size_t n4 = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ICE_OK;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && rc == ICE_OK) { g_create_error = hipGetErrorString(e); rc = ICE_E_HIP; }
        return rc == ICE_OK;
    };
    if (ok(hipMalloc(&a, n4 * 16)) && ok(hipMalloc(&b, n4 * 16)) && ok(hipStreamCreate(&st)) &&
        ok(hipEventCreate(&e0)) && ok(hipEventCreate(&e1)) && ok(hipMemsetAsync(a, 0, n4 * 16, st))) {
        const unsigned blocks = 4096;
        hipLaunchKernelGGL(k_copy4, dim3(blocks), dim3(256), 0, st, a, b, n4);   // warm-up
        ok(hipEventRecord(e0, st));
        for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy4, dim3(blocks), dim3(256), 0, st, a, b, n4);
        ok(hipGetLastError());
        ok(hipEventRecord(e1, st));
        float ms = 0.f;
        if (ok(hipEventSynchronize(e1)) && ok(hipEventElapsedTime(&ms, e0, e1))) *ms_per_copy = ms / reps;
    }
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int ice_get_info(IceHandle *h, IceInfo *info) {
    if (!h || !info) return ICE_E_ARG;
    info->n_points = h->n;
    info->capacity = h->capacity;
    info->point_pitch = h->pitch;
    info->grid_pitch = h->gpitch;
    info->grid_x = h->prm.GridXTotal;
    info->grid_y = h->prm.GridYTotal;
    info->device = h->device;
    info->sorted = h->sorted ? 1 : 0;
    info->steps_since_sort = h->steps_since_sort;
    info->n_sorts = h->n_sorts;
    info->flags = h->prm.flags;
    info->tile_w = h->tile_w;
    info->tile_h = h->tile_h;
    Counters c{};
    if (h->cnt) {
        int rc = read_counters(h, c);
        if (rc) return rc;
    }
    info->n_offtile = c.n_esc;
    return ICE_OK;
}

}  // extern "C"
