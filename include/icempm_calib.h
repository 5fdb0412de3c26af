/* icempm_calib.h — HBM streaming measurements (libicempm_calib.so).
 *
 * No reference counterpart and not part of the simulation boundary: these are
 * the measurement tools that give a device's achievable streaming bandwidth and
 * the scale of the profiler's byte counters (FETCH_SIZE, WRITE_SIZE) against
 * kernels whose traffic is known exactly.  Return codes are the ICE_* codes of
 * icempm.h; ice_calib_last_error() describes the last failure.
 */
#ifndef ICEMPM_CALIB_H
#define ICEMPM_CALIB_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Average time of one pass over `bytes` on `device`, over `reps` launches after
 * one warm-up launch: mode 0 copies (reads and writes `bytes`), 1 reads, 2
 * writes.  Allocates and frees its own buffers. */
int ice_stream_bench(int device, size_t bytes, int reps, int mode, double *ms_per_pass);
/* Mode 0 of ice_stream_bench. */
int ice_copy_bench(int device, size_t bytes, int reps, double *ms_per_copy);
/* Message of the last failed call (empty if none). */
const char *ice_calib_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
