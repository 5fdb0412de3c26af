/* icempm_io.h — host-side file formats of the reference around the substep
 * (libicempm_io.so, HDF5 1.10 through its C++ API; no device code):
 *
 *  - snapshots: SnapshotManager::SaveSnapshot / ReadPointsFromSnapshot
 *    (simulation/data_manager/snapshotmanager.cpp:308-396): dataset
 *    "pts_data" [nPtsArrays = 19][n], chunked (19 x min(256 Ki, n)) and
 *    deflated (level 5), with scalar attributes SimulationStep (int),
 *    SimulationTime (double), HSSOA_size (unsigned) and nPtsArrays (int);
 *  - the static current field: WindAndCurrentInterpolator::ReadStaticFlowData
 *    (simulation/data_manager/windandcurrentinterpolator.cpp:18-63).
 *
 * Every function returns 0 on success and a negative code on failure, with a
 * message from ice_io_last_error() (per thread).  Nothing here throws.
 */
#ifndef ICEMPM_IO_H
#define ICEMPM_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ICE_IO_E_ARG    -1
#define ICE_IO_E_FILE   -2   /* HDF5 open / read / write failure */
#define ICE_IO_E_FORMAT -3   /* the file is not in the expected layout */

const char *ice_io_last_error(void);

/* SnapshotManager::SaveSnapshot (snapshotmanager.cpp:353-396): the first 19
 * lines of a host SoA (float lines [n_lines][pitch], the device's layout:
 * utility, cell and point index are bit patterns) as "pts_data".  precision
 * 32 writes NATIVE_FLOAT (the reference built with t_PointReal = float), 64
 * NATIVE_DOUBLE (t_PointReal = double: the integer lines are then stored the
 * reference's way, as the low 32 bits of the double's bit pattern,
 * proxypoint2d.cpp:74-89).  With n_lines = 20 the per-point cohesion (line
 * 19, no reference counterpart) goes into a second dataset "cohesion" [n],
 * which the reference's reader ignores. */
int ice_snapshot_save(const char *path, const float *soa, size_t pitch, uint32_t n, int n_lines, int simulation_step,
                      double simulation_time, int precision);

/* The attributes of a snapshot: point count (HSSOA_size), step, time, the
 * stored precision (32 or 64) and whether a "cohesion" dataset is present. */
int ice_snapshot_info(const char *path, uint32_t *n, int *simulation_step, double *simulation_time, int *precision,
                      int *has_cohesion);

/* SnapshotManager::ReadPointsFromSnapshot (snapshotmanager.cpp:308-350): the
 * points into a host SoA of capacity `capacity` slots per line (pitch), float
 * lines; double files are converted (integer lines taken from the low 32 bits
 * of each double).  With n_lines = 20 line 19 takes the "cohesion" dataset,
 * or 0 when the file has none.  Fails when the dataset is not
 * [19][HSSOA_size] or nPtsArrays != 19, as the reference does. */
int ice_snapshot_load(const char *path, float *soa, size_t pitch, uint32_t capacity, int n_lines, uint32_t *n,
                      int *simulation_step, double *simulation_time);

/* WindAndCurrentInterpolator::ReadStaticFlowData
 * (windandcurrentinterpolator.cpp:18-63): dataset "velocities" [height][width][2]
 * float of an image of width x height pixels; grid node (i, j) takes pixel
 * (i + ox, j + oy).  current [2][gx*gy] (node index i*gy + j), ready for
 * ice_upload_forcing.  Fails on a rank or dimension mismatch, as the
 * reference does, and when the grid does not fit in the image. */
int ice_read_static_flow(const char *path, int width, int height, int ox, int oy, int gx, int gy, float *current);

/* Writes a "velocities" [height][width][2] float dataset (test and
 * preparation helper; the reference's preparer writes this format). */
int ice_write_static_flow(const char *path, int width, int height, const float *velocities);

#ifdef __cplusplus
}
#endif

#endif /* ICEMPM_IO_H */
