/* su2hip_traj.h -- heatbath trajectories of the reference's HB_TRAJECTORY
 * build (src/hb_trajectory.c), part of libsu2hip.so: from the current
 * configuration of a multicanonical lattice, run sweeps with the
 * multicanonical weighting off and the random pass order on until the order
 * parameter leaves [min, max], n_traj times, recording the order parameter
 * along the way.  All functions return 0 or a negative SU2HIP_E* code
 * (su2hip_last_error() has the message).
 */
#ifndef SU2HIP_TRAJ_H
#define SU2HIP_TRAJ_H

#include <stdint.h>

#include "su2hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The reference's `trajectory` struct without its file (src/su2.h:568-575). */
typedef struct su2hip_trajectory_config {
  int32_t n_traj;        /* trajectories per set */
  int32_t interval;      /* order-parameter measurement every `interval` sweeps */
  int32_t mode_interval; /* a set every `mode_interval` iterations of the main loop */
  int32_t reserved;
  double min, max;       /* a trajectory ends when the order parameter leaves [min, max] */
} su2hip_trajectory_config;

/* Read `realtime_config` (read_realtime_config, src/hb_trajectory.c:99-163):
 * lines of "key value" (first two words, at most 198 characters a line,
 * '#' keys skipped, later lines win), keys n_traj, interval, mode_interval
 * (strtol), min, max (strtod), case-insensitive.  A missing file or key is
 * SU2HIP_EIO.  (The reference checks the keys through check_set, a function
 * its sources do not define.) */
int su2hip_realtime_config_read(const char *path, su2hip_trajectory_config *out);

/* Header line of the trajectory file, written once at start-up
 * (src/main.c:233-237): "===== Realtime trajectories: min=%g, max=%g,
 * n_traj=%d ===== \n", appended to `path`. */
int su2hip_trajectory_header(const char *path, const su2hip_trajectory_config *t);

/* One set of trajectories of lattice `slot` (make_realtime_trajectories,
 * src/hb_trajectory.c:11-95), appended to `path` in the reference's format:
 * the set header with the current order parameter, then per trajectory
 * "--- Trajectory %d ---" and "%d %.8lf" lines (sweep, order parameter) at
 * sweep 0 and every `interval` sweeps until it leaves [min, max].  Every
 * trajectory starts from the lattice's configuration at the call; the
 * lattice gets it back at the end.  The sweeps count in the lattice's
 * counters and consume its Philox passes, and the order parameter's parity
 * contributions are left at the last trajectory's values, as the reference
 * leaves them.  Multicanonical acceptance and the random pass order are
 * restored to their state at the call.  The lattice must have a weight
 * installed (the reference requires multicanonical, src/main.c:228-231). */
int su2hip_trajectories(su2hip_batch_t *b, int slot, const su2hip_trajectory_config *t, int id,
                        const char *path);

#ifdef __cplusplus
}
#endif

#endif
