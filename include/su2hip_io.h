/* su2hip_io.h -- host-side file formats of the sweep, part of libsu2hip.so:
 * the reference's `config` key-value file, the binary `lattice` checkpoint,
 * the multicanonical `weight` file and the `labels` / `measure` /
 * `measure_flow` text outputs, byte for byte.  Pure C, no GPU calls.
 * All functions return 0 or a negative SU2HIP_E* code
 * (su2hip_last_error() has the message); none calls exit().
 */
#ifndef SU2HIP_IO_H
#define SU2HIP_IO_H

#include <stdint.h>
#include <stdio.h>

#include "su2hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SU2HIP_PATHLEN 200

/* Everything a run reads from `config` (src/parameters.c:157-402), plus the
 * keys the reference does not know (model, seed, start, link_overrelax,
 * precision, link_format, flow_scheme), which it ignores when it reads the
 * same file. */
typedef struct su2hip_run_config {
  su2hip_params p;
  int32_t gradflow;          /* field content includes the gradient-flow keys */
  int32_t reset, run_checks, multicanonical, measure_local;
  int64_t iterations, interval, checkpoint, n_thermalize;
  double phi0, sigma0, singlet0;
  int32_t hot_start;         /* start hot (no reference counterpart) */
  int32_t do_flow, flow_interval, flow_meas_interval, flow_scheme;
  double flow_dt, flow_t_max;
  su2hip_muca muca;          /* bins/min/max/mode/checks/orderparam/delta from config */
  char resultsfile[SU2HIP_PATHLEN];
  char latticefile[SU2HIP_PATHLEN];
  char weightfile[SU2HIP_PATHLEN];
  int32_t correlators;       /* field content includes the correlator keys */
  int32_t do_correlators, correlator_interval;
  int32_t measure_z;         /* field content includes the z-profile keys */
  int32_t do_z_meas, setup_wall;
  int64_t interval_z;
  int32_t blocking;          /* field content includes the blocking key */
  int32_t blocks;            /* blocking levels of the blocked correlators */
  int32_t trajectory;        /* field content includes the heatbath-trajectory key */
  int32_t do_trajectory;     /* heatbath trajectories every mode_interval iterations */
} su2hip_run_config;

/* Parse `path` the way get_parameters + get_weight_parameters do
 * (case-insensitive labels, `#` comments, first match searching on from the
 * previous match and wrapping once, strtol/strtod conversion, unknown
 * algorithm names -> metropolis).  `model` overrides the config's `model` key
 * when non-NULL: '+'-separated tokens among higgs, 2hdm (two doublets, the
 * -DNHIGGS=2 build), u1, triplet, singlet, gradflow, correlators, measure_z,
 * blocking (implies correlators), trajectory (-DHB_TRAJECTORY); the default
 * is the reference Makefile's build, higgs+singlet+u1. */
int su2hip_config_read(const char *path, const char *model, su2hip_run_config *out);
/* Re-read iterations, interval and checkpoint (read_updated_parameters,
 * src/parameters.c:460-494). */
int su2hip_config_reread(const char *path, su2hip_run_config *cfg);

/* Binary checkpoint (save_lattice/load_lattice, src/checkpoint.c:84-214):
 * int size(=1), int dim(=3), int L[3], long iter, double total_time,
 * double comms_time, then su2link, u1link, doublet(s), triplet, singlet for
 * the fields of p, each in the reference host layout (with two doublets
 * `doublet` holds both, 2 V 4 doubles, the first one first). */
int su2hip_lattice_write(const char *path, const su2hip_params *p, int64_t iter,
                         double total_time, double comms_time, const double *su2link,
                         const double *u1link, const double *doublet,
                         const double *triplet, const double *singlet);
/* Checks dim and extents against p (the reference dies on a mismatch,
 * src/checkpoint.c:154-179; here SU2HIP_EIO).  A file written by an MPI run
 * (size > 1, rank after rank, src/checkpoint.c:226-269) is read too: its
 * sites are put back into the single-lattice order (su2hip_mpi_site_order).
 * The reference itself only reads a file of its own rank count. */
int su2hip_lattice_read(const char *path, const su2hip_params *p, int64_t *iter,
                        double *total_time, double *comms_time, double *su2link,
                        double *u1link, double *doublet, double *triplet, double *singlet);
/* Node grid (nslices) and node box (slice) of an MPI run of `size` ranks on
 * extents L (make_slices, src/layout.c:118-237). */
int su2hip_mpi_slices(const int32_t *L, int size, int32_t *nslices, int32_t *slice);
/* order[q] = single-lattice site index of the q-th site of an MPI-written
 * field block (V entries): ranks in turn, each its node box's sites
 * x-fastest, EVEN before ODD (sitemap and paritymap, src/layout.c:245-402,
 * :826-851). */
int su2hip_mpi_site_order(const int32_t *L, int size, int64_t *order);

/* Weight file (load_weight / save_weight, src/multicanonical.c:82-288).
 * On input m holds the config's bins/min/max/mode/delta/orderparam/
 * checks_per_sweep; pos/W must have room for max_bins+2 values.  A missing
 * file gives a flat weight (an error for READONLY, as in the reference). */
int su2hip_weight_load(const char *path, su2hip_muca *m, double *pos, double *W,
                       int max_bins);
int su2hip_weight_save(const char *path, const su2hip_muca *m, const double *pos,
                       const double *W);

/* `labels` (print_labels, src/measure.c:15-64) for the field content of p. */
int su2hip_labels_write(const char *path, const su2hip_params *p);
/* One `measure` line (src/measure.c:252-306): "%g " per column, newline. */
int su2hip_measure_format(const su2hip_params *p, const double *sums, double weight,
                          double muca_param, char *buf, int buflen);

/* `labels_correl` (print_labels_correlators, src/correlation.c:313-337). */
int su2hip_correl_labels_write(const char *path, const su2hip_params *p);
/* Append one measurement to `path` as measure_correlators writes it
 * (src/correlation.c:339-391): a "Measurement id" header, then per distance
 * "%d " and "%g " for the model's columns of su2hip_correlators' rows. */
int su2hip_correl_append(const char *path, const su2hip_params *p, int meas_id,
                         const double *corr, int ndist);

/* `labels_z` (print_z_labels, src/z_coord.c:78-106). */
int su2hip_zprof_labels_write(const char *path, const su2hip_params *p);
/* Append one profile to `path` as measure_along_z writes it
 * (src/z_coord.c:265-292): per plane "%ld " and "%g " for the model's
 * columns of su2hip_zprofile's rows. */
int su2hip_zprof_append(const char *path, const su2hip_params *p, const double *prof, int nz);

/* gsum/nsum of the SLOW weight recursion, bins+2 entries, from a
 * `weight_params` file (load_weight_params, src/multicanonical.c:291-333):
 * a missing file leaves every entry at gsum 5, nsum 1 and sets *found = 0;
 * a file for another bin count is an error.  Entries 0 and bins+1 are not
 * in the file and stay 5 and 1. */
int su2hip_weight_params_load(const char *path, int bins, int64_t *gsum, int64_t *nsum,
                              int *found);
/* Write `weight_params` as save_weight does for SLOW runs
 * (src/multicanonical.c:99-109); nothing for other modes. */
int su2hip_weight_params_save(const char *path, const su2hip_muca *m, const double *pos,
                              const int64_t *gsum, const int64_t *nsum);

/* Acceptance summary (print_acceptance, src/checkpoint.c:17-78) to `out`. */
int su2hip_print_acceptance(FILE *out, const su2hip_run_config *cfg,
                            const su2hip_counters *c);

#ifdef __cplusplus
}
#endif

#endif
