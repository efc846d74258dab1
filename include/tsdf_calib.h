/* Device-rate calibration for the bench's roofline (libtsdf_calib.so).
 *
 * Not part of the integration path: no counterpart in the reference.  It
 * measures what the MI355X's memory system sustains for the global atomic
 * patterns the pass kernels use, so that a kernel's atomic count per launch
 * (from rocprofv3 counters) can be set against a measured ceiling rather
 * than a guessed one.  The reference's atomics these patterns stand in for:
 * the per-voxel spin-lock and counter atomics of tsdf_handler.cu:446,
 * :643-664 and :780.
 */
#ifndef TSDF_CALIB_H
#define TSDF_CALIB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Atomic patterns, by lane:
 * SCATTER_RET64: 64-bit add with return, every lane on its own line of a
 *   128 MiB table, lines hashed per lane and step (k_bin's per-block record
 *   reservations, k_seg's heap pops);
 * SCATTER_NORET32: 32-bit add, result unused, same addressing;
 * COALESCED_RET32: 32-bit add with return, a wave's 64 lanes on 64
 *   consecutive words (256 bytes), waves on hashed places of the table;
 * SAME_RET32: 32-bit add with return on ONE word, by one lane per wave (the
 *   pass's list counters: touched_n, new_n, fold_n). */
#define TSDF_CAL_SCATTER_RET64 0
#define TSDF_CAL_SCATTER_NORET32 1
#define TSDF_CAL_COALESCED_RET32 2
#define TSDF_CAL_SAME_RET32 3

/* Runs pattern `pattern` on device `device` with `workgroups` workgroups of
 * 256 threads, each active lane doing `steps` atomics, and times the launch
 * with HIP events (one untimed launch first).  *ops = lane atomics of the
 * timed launch, *seconds = its duration.  Returns 0, or -1 with a message
 * from tsdf_calib_last_error(). */
int tsdf_calib_atomics(int device, int pattern, int workgroups, int steps, int64_t* ops, double* seconds);

const char* tsdf_calib_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TSDF_CALIB_H */
