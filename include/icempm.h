/*
 * icempm.h — C-ABI of the MI355X-native single-partition MPM substep.
 *
 * Plain C: pointers, sizes and ints only, no C++ or torch types.  Every entry
 * point returns 0 on success and a negative ICE_E_* code on failure; it never
 * throws across the ABI.  ice_last_error(h) gives the message of the last
 * failure on a handle (ice_last_error(NULL) the last failure of ice_create).
 * A handle is not thread-safe: drive it from one host thread.  Downloads are
 * synchronous when they return.
 *
 * Each declaration names the reference interface it replaces
 * (igorg520b/iceMPM_2d_sp, file:line).
 */
#ifndef ICEMPM_H
#define ICEMPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- particle SoA layout (hssoa), one line of `pitch` 4-byte slots per index.
 * Same indices as SimParams (simulation/parameters_sim.h:65-76) with
 * t_PointReal = float (parameters_sim.h:26).  Integer lines (utility, cell,
 * point index) are uint32 bit patterns stored in the float slots. */
#define ICE_IDX_UTILITY    0   /* bit16 crushed, bit17 disabled, bit18 error/weakened */
#define ICE_IDX_CELL       1   /* (y << 16) | x of the nearest node */
#define ICE_IDX_POINT      2   /* original point index */
#define ICE_IDX_P          3
#define ICE_IDX_Q          4
#define ICE_IDX_JP_INV     5
#define ICE_IDX_POSX       6   /* cell-local position in [-0.5, 0.5], 6..7 */
#define ICE_IDX_VELX       8   /* 8..9 */
#define ICE_IDX_FE00       10  /* Fe row-major, 10..13 */
#define ICE_IDX_BP00       14  /* Bp (APIC C) row-major, 14..17 */
#define ICE_IDX_THICKNESS  18
#define ICE_N_LINES        19  /* nPtsArrays of the reference */
#define ICE_IDX_COHESION   19  /* optional per-particle cohesion line (Pa), build addition */
#define ICE_N_LINES_MAX    20

#define ICE_STATUS_CRUSHED  0x10000u  /* kernels.cuh:11 */
#define ICE_STATUS_DISABLED 0x20000u  /* kernels.cuh:12 */
#define ICE_STATUS_ERROR    0x40000u  /* kernels.cuh:13 */

/* error bits of the device error word (kernels.cuh:17-26) */
#define ICE_ERR_POINT_POS_NAN     0x0001u
#define ICE_ERR_POINT_VEL_NAN     0x0002u
#define ICE_ERR_POINT_JUMP_CELLS  0x0004u
#define ICE_ERR_POINT_LEFT_AREA   0x0008u
#define ICE_ERR_POINT_BP_NAN      0x0010u
#define ICE_ERR_POINT_FE_NAN      0x0020u
#define ICE_ERR_GRID_P2G_NAN_VEL  0x0100u
#define ICE_ERR_GRID_P2G_NAN_MASS 0x0200u
#define ICE_ERR_GRID_NAN          0x0400u

/* grid status byte: 100 = modelled water, anything else = sticky wall whose
 * value is the boundary region id (snapshotmanager.cpp:129-144). */
#define ICE_GRID_STATUS_WATER 100

/* ---- flags in IceParams.flags */
#define ICE_FLAG_COHESION_LINE  0x1u  /* line 19 holds per-particle cohesion */
#define ICE_FLAG_DETERMINISTIC  0x2u  /* fixed-point P2G accumulation (bit-reproducible) */
#define ICE_FLAG_NAIVE          0x4u  /* unsorted atomic P2G + separate update + G2P kernels */

/* return codes */
#define ICE_OK          0
#define ICE_E_ARG      -1
#define ICE_E_HIP      -2
#define ICE_E_STATE    -3
#define ICE_E_CAPACITY -4
#define ICE_E_DEVICE   -5

/* Model constants: the fields of SimParams (parameters_sim.h:78-117) that the
 * substep reads, with the FP64 types of the reference.  The block marked
 * "derived" is filled by ice_params_derive from the others. */
typedef struct IceParams {
    int32_t GridXTotal, GridYTotal;            /* grid nodes along x and y */
    double  cellsize;                          /* h, m */
    double  InitialTimeStep;                   /* dt, s */
    double  AnimationFramePeriod;              /* s; sets UpdateEveryNthStep */
    double  IceDensity, PoissonsRatio, YoungsModulus;
    double  IceCompressiveStrength, IceTensileStrength, IceShearStrength, IceTensileStrength2;
    double  DP_phi, DP_threshold_p;            /* degrees; Pa */
    double  RidgeFormationCoeff;
    double  ParticleVolume;                    /* m^2 per particle */
    double  sea_water_density;
    double  waterDragEffectiveLinear;          /* 1/s */
    double  waterDragEffectiveQuadratic;       /* 1/m */
    double  windStressToAccel;                 /* 1/(rho_ice*h_ice): wind stress (Pa) -> m/s^2; build addition */
    /* derived (parameters_sim.cpp:122-140) */
    double  dt_vol_Dpinv, vmax, lambda, mu, kappa, ParticleMass, cellsize_inv, Dp_inv;
    int32_t UpdateEveryNthStep;
    /* launch shape and mode */
    int32_t tpb_P2G, tpb_Upd, tpb_G2P;         /* parameters_sim.cpp:46-48; naive path only */
    uint32_t flags;                            /* ICE_FLAG_* */
} IceParams;

typedef struct IceHandle IceHandle;

/* Fill defaults of SimParams::Reset (parameters_sim.cpp:7-55) into *p. */
int ice_params_reset(IceParams *p);
/* SimParams::ComputeHelperVariables + ComputeLame (parameters_sim.cpp:122-140). */
int ice_params_derive(IceParams *p);

/* GPU_Implementation5::initialize + GPU_Partition::initialize
 * (gpu_implementation5.cpp:94-128; gpu_partition.cpp:234-256) and
 * update_constants (gpu_partition.cpp:140-170): one partition on `device`. */
int ice_create(const IceParams *p, int device, IceHandle **out);
/* GPU_Partition destructor (gpu_partition.cpp:214-232). */
int ice_destroy(IceHandle *h);
const char *ice_last_error(const IceHandle *h);

/* GPU_Partition::allocate (gpu_partition.cpp:259-293).  Device pitch is
 * reported by ice_get_info.  A handle holds fewer than 2^30 point slots
 * (ICE_E_CAPACITY beyond). */
int ice_allocate(IceHandle *h, uint32_t n_points_capacity);

/* GPU_Partition::transfer_points_from_soa_to_device (gpu_partition.cpp:81-106):
 * `n` points from a host SoA with `host_pitch` slots per line and `n_lines`
 * lines (19, or 20 with the cohesion line).  Sets the disabled count to 0, as
 * split_hssoa_into_partitions does (gpu_implementation5.cpp:137).  Unless
 * ICE_FLAG_NAIVE is set, the points are then put in cell order on the device;
 * ICE_IDX_POINT identifies them in later downloads.  The device keeps Fe - I
 * in the four Fe lines (small strains keep their FP32 precision); the
 * downloads add I back, so callers read and write Fe. */
int ice_upload_points(IceHandle *h, const float *soa, size_t host_pitch, uint32_t n, int n_lines);
/* GPU_Partition::transfer_grid_data_to_device (gpu_partition.cpp:108-137):
 * status [gx*gy] bytes, node index x*gy+y; normals [2][gx*gy] or NULL. */
int ice_upload_grid(IceHandle *h, const uint8_t *status, const float *normals);
/* GPU_Partition::update_current_field (gpu_partition.cpp:174-196) plus the
 * wind-stress field the reference declares but never applies
 * (gpu_implementation5.cpp:51-71): each [2][gx*gy] (x then y) or NULL = zero. */
int ice_upload_forcing(IceHandle *h, const float *current, const float *wind_stress);
/* Time-varying wind, the forcing the reference leaves unfinished
 * (WindAndCurrentInterpolator::SetTime, windandcurrentinterpolator.cpp:66-70;
 * the commented-out interpolation of kernels.cuh:580-622, (1 - tb) w0 + tb w1):
 * two wind-stress frames [2][gx*gy] (x then y, Pa) of domain d, valid at
 * t0 < t1.  Every later substep's node update uses w0 + s (w1 - w0) with
 * s = clamp((t - t0) / (t1 - t0), 0, 1) at its simulation time t.  A caller
 * stepping through a series uploads the next pair when t passes t1;
 * ice_upload_forcing (a static wind field) ends it. */
int ice_upload_wind_frames(IceHandle *h, int domain, const float *wind0, const float *wind1, double t0, double t1);
/* GPU_Partition::update_constants (gpu_partition.cpp:140-170): new constants,
 * same grid size. */
int ice_update_constants(IceHandle *h, const IceParams *p);

/* ---- per-substep phases (gpu_implementation5.cpp:18-81) */
int ice_reset_grid(IceHandle *h);                                   /* gpu_partition.cpp:300-308 */
int ice_p2g(IceHandle *h);                                          /* gpu_partition.cpp:322-334 */
int ice_update_nodes(IceHandle *h, float simulation_time, float windSpeed, float windAngle); /* :336-348 */
int ice_g2p(IceHandle *h, int recordPQ, int enablePointTransfer, int applyGlensLaw);        /* :350-365 */
/* One whole substep (reset_grid, p2g, update_nodes, g2p) as one call; the
 * sequence Model::Step runs (model.cpp:18-39).  Unless ICE_FLAG_NAIVE or
 * ICE_FLAG_DETERMINISTIC is set (or the "pipelined" option is 0), substeps
 * run pipelined: one tile pass does the update and G2P of this substep and
 * the P2G of the next from the new point state, so the first substep after
 * an upload (or any change to the points or constants) also runs one P2G. */
int ice_substep(IceHandle *h, double simulation_time, int recordPQ);
/* `nsteps` substeps starting at step index `first_step` (t = dt*step),
 * recordPQ on steps that are multiples of UpdateEveryNthStep; replays a
 * captured HIP graph of 16 substeps when use_graph != 0, except on the naive
 * path (ICE_FLAG_NAIVE, the reference's launch structure), which always
 * launches eagerly. */
int ice_run(IceHandle *h, int first_step, int nsteps, int use_graph);

/* ---- frame boundary */
int ice_clear_force_accumulator(IceHandle *h);                      /* gpu_partition.cpp:310-319 */
/* GPU_Partition::transfer_from_device (gpu_partition.cpp:17-59): points into a
 * host SoA (NULL to skip), boundary forces [2][gx*gy] (NULL to skip), the
 * error word and the disabled count.  Synchronous. */
int ice_download(IceHandle *h, float *soa, size_t host_pitch, int n_lines,
                 float *boundary_forces, uint32_t *error_code, int32_t *disabled_count);
/* ---- frame arrays: SnapshotManager::PrepareFrameArrays
 * (data_manager/snapshotmanager.cpp:623-765) on the device.  Point colours
 * (0xRRGGBB per original point index, the reference's point_colors_rgb) are
 * uploaded once; n may be 0 (every point black). */
int ice_upload_point_colors(IceHandle *h, const uint32_t *rgb, uint32_t n);
#define ICE_MAX_REGIONS 100          /* SimParams::MAX_REGIONS, parameters_sim.h:46 */
#define ICE_FRAME_FIELDS 10          /* point density, mass, vx, vy, r, g, b, Jp_inv - 1, P, Q */
typedef struct IceFrameArrays {      /* host arrays, any may be NULL (skipped) */
    uint8_t *count;                  /* [gx*gy] points per cell, modulo 256 (the reference's uint8 counter) */
    float *fields;                   /* [ICE_FRAME_FIELDS][gx*gy]: vis_point_density, vis_mass, vis_vx, vis_vy,
                                        vis_r, vis_g, vis_b, vis_Jpinv, vis_P, vis_Q after normalisation */
    uint8_t *rgb;                    /* [gx*gy][3] colour blended with the water colour */
    uint8_t *mass_mask;              /* [gx*gy] point density > 0.1 */
    float *forces_per_region;        /* [ICE_MAX_REGIONS][2] boundary forces summed per region */
} IceFrameArrays;
/* The frame arrays of a single-domain handle from its current points and
 * force accumulator; boundary forces are divided by elapsed_cycles *
 * InitialTimeStep (as transfer_from_device leaves them,
 * gpu_implementation5.cpp:219-221) when elapsed_cycles > 0.  Sorts the points
 * by cell first (the device order changes; ICE_IDX_POINT identifies them).
 * Regions past ICE_MAX_REGIONS are skipped (the reference writes past its
 * array).  Synchronous. */
int ice_prepare_frame_arrays(IceHandle *h, int elapsed_cycles, IceFrameArrays *out);

/* Grid fields after the last substep: [3][gx*gy] m, px, py as P2G left them
 * (momentum), or NULL; [2][gx*gy] node velocities after the update, or NULL. */
int ice_download_grid(IceHandle *h, float *mpp, float *velocity);
int ice_synchronize(IceHandle *h);                                  /* gpu_implementation5.cpp:227-234 */

/* Cell-binned sort on the device: drop disabled points and stable-sort the
 * rest by cell index x*GridY+y (HostSideSOA::RemoveDisabledAndSort,
 * host_side_soa.cpp:36-47, with std::stable_sort order).  *n_out = points
 * kept.  perm (NULL to skip) receives, per output slot, the input slot. */
int ice_sort(IceHandle *h, uint32_t *n_out, uint32_t *perm);

/* ---- timings (GPU_Partition::reset_timings / record_timings /
 * normalize_timings, gpu_partition.cpp:370-459), HIP events on the handle's
 * stream around each phase, accumulated since the last reset:
 * out[0] ms in the main P2G kernel, out[1] ms in the P2G phase (grid clear,
 * P2G, off-tile P2G), out[2] ms in the node update (naive) or off-tile G2P,
 * out[3] ms in the main G2P kernel, out[4] ms of whole substeps, out[5]
 * substeps recorded, out[6..7] 0.  A pipelined substep counts its tile pass
 * as the main G2P kernel, the pipeline's initial P2G (if any) as the main P2G
 * kernel, and its off-tile pass under out[2].  Graph replays record nothing. */
int ice_reset_timings(IceHandle *h);
int ice_enable_timings(IceHandle *h, int enable);
int ice_get_timings(IceHandle *h, double out[8]);

/* Tile of grid nodes per workgroup of the sorted path (width along x, height
 * along y).  Supported: 8x8, 8x16, 16x8, 16x16, 16x32 (default), 32x16, 32x32,
 * 32x64, 64x32, 64x64.  No reference counterpart (the reference's only launch knobs
 * are tpb_P2G/Upd/G2P, parameters_sim.cpp:111-113). */
int ice_set_tile(IceHandle *h, int tile_w, int tile_h);
/* Tuning switches of the sorted path (no reference counterpart): "band"
 * (1..4096, default 1) is the number of tile columns per band of the
 * workgroup-to-tile order; "pipelined" (0 or 1, default 1) selects the
 * pipelined substep of ice_substep / ice_run; "persistent" (0 or 1, default
 * 0) runs its tile passes (without the surrogate) as persistent launches, as
 * many workgroups as fit on the device pulling tiles from per-XCD queues;
 * "resort_drift" (0..100, default 30; 0 turns it off): between ice_run's
 * 16-substep chunks a domain is re-sorted when its off-tile list exceeds
 * max(256, n/256) points and also, on the pipelined path, when more than this
 * percentage of its points sit in a cell other than the one they were sorted
 * into (they leave the tile pass in bulk once they move again). */
int ice_set_option(IceHandle *h, const char *name, int value);
/* Tile autotuner (no reference counterpart): times `nsteps` substeps for each
 * tile shape from 8x8 to 64x32, restoring the simulation state after each
 * trial, and keeps the fastest.  Outputs the chosen shape and its ms/substep. */
int ice_autotune(IceHandle *h, int nsteps, int *tile_w, int *tile_h, double *best_ms);

/* ---- batched domains (no reference counterpart; the reference runs one
 * partition, gpu_implementation5.cpp:96).  `n_domains` independent domains of
 * grid_x[d] x grid_y[d] nodes with room for capacity[d] points share one
 * handle: every substep is one launch per phase over all tiles of all domains
 * (largest domains first), and ice_run replays it from a HIP graph.  The
 * material constants of `p` are shared; its GridX/YTotal are ignored.  The
 * single-domain calls above act on domain 0. */
int ice_create_batch(const IceParams *p, int device, int n_domains, const int32_t *grid_x, const int32_t *grid_y,
                     const uint32_t *capacity, IceHandle **out);
/* Points, status bytes (NULL keeps the current ones) and forcing of domain d. */
int ice_upload_domain(IceHandle *h, int d, const float *soa, size_t host_pitch, uint32_t n, int n_lines,
                      const uint8_t *status, const float *current, const float *wind_stress);
int ice_download_domain(IceHandle *h, int d, float *soa, size_t host_pitch, int n_lines, float *boundary_forces,
                        uint32_t *error_code, int32_t *disabled_count);

/* ---- introspection */
typedef struct IceInfo {
    uint32_t n_points;           /* nPts_partition */
    uint32_t capacity;           /* points the device buffer holds */
    uint64_t point_pitch;        /* device slots per line (nPtsPitch) */
    uint64_t grid_pitch;         /* device slots per grid line (nGridPitch) */
    int32_t  grid_x, grid_y;
    int32_t  device;
    int32_t  sorted;             /* 1 when points are in cell order */
    int32_t  steps_since_sort;
    int32_t  n_sorts;            /* device sorts done since create */
    uint32_t flags;
    int32_t  tile_w, tile_h;     /* tile of the sorted path */
    uint32_t n_offtile;          /* points handled off-tile until the next sort */
    int32_t  n_domains;          /* domains in the handle */
    int32_t  n_tiles;            /* tiles per launch, all domains */
} IceInfo;
int ice_get_info(IceHandle *h, IceInfo *info);                      /* domain 0 */
int ice_get_domain_info(IceHandle *h, int d, IceInfo *info);

/* ---- learned plastic projection (no reference counterpart; replaces, when
 * deployed, the plastic branch of Wolper_Drucker_Prager, kernels.cuh:431-479).
 * A 4 -> 64 -> 64 -> 3 MLP with SiLU, BF16 weights and activations on MFMA,
 * FP32 accumulation.  Features of a trial state: log J, D = log(s0/s1),
 * Jp_inv, log(c / 1e5) (s0 >= s1 the singular values of the elastic
 * predictor, c the cohesion); targets: log(-log J'), log D' of the projected
 * state (log J' < 0, D' > 0 on the ridge-capped branch) and
 * log(Jp_inv' / Jp_inv).  Parameters: W1[64][4], b1[64], W2[64][64], b2[64],
 * W3[3][64], b3[3] (ICE_MLP_PARAMS floats). */
#define ICE_MLP_PARAMS 4675
typedef struct IceMlpNorm {
    float xm[4], xs[4];          /* standardised feature = (raw - xm) / xs */
    float ym[3], ys[3];          /* raw target = ym + ys * output */
} IceMlpNorm;
typedef struct IceMlpSampling {  /* trial states for training pairs */
    double logJ_min, logJ_max;   /* uniform */
    double D_min, D_max;         /* log-uniform */
    double jp_min, jp_max;       /* log-uniform; c ~ lognormal(log 1e5, 0.5), capped at 1e6 */
} IceMlpSampling;
/* Sampling ranges used when `s` is NULL. */
int ice_mlp_default_sampling(IceMlpSampling *s);
/* n training pairs on the device from the exact projection in FP64 with the
 * handle's constants (thickness 1, ridge cap active), kept for ice_mlp_train; copies to x
 * [n][4] and t [n][3] when not NULL.  Sets the standardisation from the pairs.
 * *n_failed = pairs whose sampling found no plastic state. */
int ice_mlp_generate(IceHandle *h, uint32_t n, uint64_t seed, const IceMlpSampling *s, float *x, float *t,
                     uint32_t *n_failed);
/* Record the trial states the exact map projects on the ridge-capped
 * plastic branch (thickness 1) during the following pipelined substeps, as
 * feature rows (log J, D, Jp_inv, log(c / 1e5)): each state is kept with
 * probability keep_fraction (a hash of its slot and step), up to `capacity`
 * rows.  capacity 0 stops recording and frees the rows.  Recording runs the
 * tile pass of the learned projection (the network itself only when it is
 * deployed). */
int ice_mlp_record(IceHandle *h, uint32_t capacity, float keep_fraction);
/* Change the kept fraction of an open recording; rows stored so far stay. */
int ice_mlp_record_keep(IceHandle *h, float keep_fraction);
/* States seen on that branch since ice_mlp_record, and rows stored. */
int ice_mlp_record_count(IceHandle *h, uint64_t *seen, uint32_t *stored);
/* Copy up to n_max stored rows [n][4] to x; *n = rows copied. */
int ice_mlp_records(IceHandle *h, float *x, uint32_t n_max, uint32_t *n);
/* As ice_mlp_generate, from the recorded states: n pairs, row i from a
 * hash-chosen stored state, its target from the exact projection in FP64.
 * Also sets the feature ranges the deployed network accepts to the span of
 * the pairs. */
int ice_mlp_generate_recorded(IceHandle *h, uint32_t n, uint64_t seed, float *x, float *t, uint32_t *n_failed);
int ice_mlp_set_norm(IceHandle *h, const IceMlpNorm *norm);
int ice_mlp_get_norm(IceHandle *h, IceMlpNorm *norm);
/* FP32 master parameters (ICE_MLP_PARAMS). ice_mlp_init_params draws
 * scaled-uniform initial weights from `seed`. */
int ice_mlp_set_params(IceHandle *h, const float *params);
int ice_mlp_get_params(IceHandle *h, float *params);
int ice_mlp_init_params(IceHandle *h, uint64_t seed);
/* `steps` Adam steps on batches of `batch` pairs drawn from the generated
 * pairs (mean squared error of the standardised targets); loss per step into
 * `loss` when not NULL. */
int ice_mlp_train(IceHandle *h, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float *loss);
/* Forward pass over n raw feature rows x [n][4]: standardised outputs y [n][3],
 * by the deployed inference (BF16 layers 2 and 3, split-precision layer 1). */
int ice_mlp_forward(IceHandle *h, const float *x, uint32_t n, float *y);
/* Gradient of the mean squared error over the n rows (x raw, t raw) into
 * grad (ICE_MLP_PARAMS), loss into *loss; parameters unchanged. */
int ice_mlp_grad(IceHandle *h, const float *x, const float *t, uint32_t n, float *grad, float *loss);
/* Per raw trial state x [n][4]: out [n][6] = surrogate (p', q') of the new
 * state, exact FP64 (p', q'), yield violation |q1 - q_yield(p1)| / q_yield(p1)
 * of the projection the surrogate implies (p1 = -kappa/2 (J'^2 cosh^2 D' - 1),
 * q1 = 2 mu J' sinh D' / J_trial: the point the exact map places on the yield
 * surface, kernels.cuh:441-466, zero for the exact map), 1 if the surrogate
 * is accepted at tolerance `tol`. */
int ice_mlp_eval(IceHandle *h, const float *x, uint32_t n, float tol, float *out);
/* No counterpart in the reference.  ice_mlp_eval for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same kernel and results, with the
 * rows staged through a fixed device array of ICE_MLP_EVAL_NOALLOC_MAX_ROWS
 * rows (2.6 MB per GPU, held until the process ends) instead of buffers
 * allocated per call.  n beyond that returns ICE_E_CAPACITY.  Every call on a
 * GPU shares the array: calls that overlap on one GPU race. */
#define ICE_MLP_EVAL_NOALLOC_MAX_ROWS 65536
int ice_mlp_eval_noalloc(IceHandle *h, const float *x, uint32_t n, float tol, float *out);
/* Use the surrogate for the plastic projection in pipelined substeps
 * (enable != 0), falling back to the exact map when its yield violation
 * exceeds `tol`, its features leave the training ranges or thickness != 1.
 * counts[0] = projections the surrogate made, counts[1] = fallbacks, summed
 * over domains since the last ice_mlp_deploy. */
int ice_mlp_deploy(IceHandle *h, int enable, float tol);
int ice_mlp_counts(IceHandle *h, uint64_t counts[2]);

#ifdef __cplusplus
}
#endif
#endif /* ICEMPM_H */
