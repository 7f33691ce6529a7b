/* su2oracle.h -- CPU FP64 restatement of su2higgs's lattice update sweep.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or the
 * timed CPU baseline; the product path (su2higgs_amd/) never calls it.
 *
 * Everything here follows the reference's per-site algorithms in FP64 (with
 * the reference's long double where it has one), on the reference's host
 * layout and site order, with one difference: random numbers come from the
 * counter-based Philox4x32-10 stream defined below instead of the global
 * MT19937-64 stream (src/generic/mersenne.c), so that each site's draws are
 * addressable.  The stream is keyed by the lattice seed and addressed by
 * (pass, site, k): `pass` is a 64-bit counter incremented once per
 * checkerboard pass (field, parity[, direction]) in update_lattice order,
 * `site` the reference site index, `k` the draw number at that site.
 * Multicanonical accept draws use site = ORC_MUCA_SITE and k = check number.
 */
#ifndef SU2ORACLE_H
#define SU2ORACLE_H

#include <stdint.h>
#include "../include/su2hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MUCA_SITE 0xFFFFFFFFu
#define ORC_HOTSTART_PASS 0xFFFFFFFFFFFFFFFFull
#define ORC_WALL_PASS 0xFFFFFFFFFFFFFFFEull
/* site of a sweep's pass-order draws (random order) */
#define ORC_SCHED_SITE 0xFFFFFFFEu

typedef struct orc orc_t;

/* Philox4x32-10 block function (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* 64-bit word of draw k at (seed, pass, site): the reference's iran(). */
uint64_t orc_word(uint64_t seed, uint64_t pass, uint32_t site, uint32_t k);
/* 53-bit uniform in [0,1) for draw k at (seed, pass, site): (word >> 11) 2^-53. */
double orc_uniform(uint64_t seed, uint64_t pass, uint32_t site, uint32_t k);

orc_t *orc_create(const su2hip_params *p);
void orc_destroy(orc_t *o);
long orc_volume(const orc_t *o);
long orc_evensites(const orc_t *o);
/* next/prev tables [V][3] and the natural index of every reference site. */
void orc_tables(const orc_t *o, long *next, long *prev, long *natural);

void orc_set_fields(orc_t *o, const double *su2link, const double *u1link,
                    const double *doublet, const double *triplet,
                    const double *singlet);
void orc_get_fields(const orc_t *o, double *su2link, double *u1link,
                    double *doublet, double *triplet, double *singlet);
void orc_cold_start(orc_t *o, double phi0, double sigma0, double singlet0);
void orc_hot_start(orc_t *o);

void orc_rng_set(orc_t *o, uint64_t seed, uint64_t pass);
void orc_rng_get(const orc_t *o, uint64_t *seed, uint64_t *pass);
void orc_get_counters(const orc_t *o, su2hip_counters *c);
void orc_reset_counters(orc_t *o);

/* Per-site pieces (deterministic). */
void orc_link_staple(const orc_t *o, long i, int dir, double V[4]);
void orc_wilson_staple(const orc_t *o, long i, int dir, double V[4]);
void orc_doublet_staple(const orc_t *o, long i, double s[4]);
/* doublet h (0 or 1) of the two-doublet model; h = 0 is the one above */
void orc_doublet_staple_h(const orc_t *o, long i, int h, double s[4]);
double orc_localact_doublet_h(const orc_t *o, long i, int h);
void orc_triplet_staple(const orc_t *o, long i, double s[3]);
double orc_localact_su2link(const orc_t *o, long i, int dir);
double orc_localact_u1link(const orc_t *o, long i, int dir);
double orc_localact_doublet(const orc_t *o, long i);
double orc_localact_triplet(const orc_t *o, long i);
double orc_localact_singlet(const orc_t *o, long i);
double orc_action_local(const orc_t *o, long i);
double orc_polysolve3(double a, double b, double c, double d);

/* One per-site update of `kind` (SU2HIP_K_*) at site i with draws from
 * (seed, pass, i, 0..); for the doublet kinds dir is the doublet.  Returns
 * 1 if accepted. */
int orc_site_update(orc_t *o, int kind, long i, int dir, uint64_t pass);
/* One checkerboard pass over all sites of a parity (consumes one pass
 * number; multicanonical checks run inside scalar passes). */
void orc_pass(orc_t *o, int kind, int parity, int dir);
/* n compound sweeps (update_lattice). */
void orc_update(orc_t *o, int n);
/* The reference's random_sweeps: shuffled link pass order and a random
 * first parity of every scalar round, drawn per sweep. */
void orc_set_random_order(orc_t *o, int on);
/* multicanonical global accept/reject on (the default after orc_muca_set)
 * or off (w->do_acceptance, src/su2.h:244) */
void orc_muca_set_acceptance(orc_t *o, int on);

/* FP64 volume sums in the SU2HIP_SUM_* layout. */
void orc_measure(const orc_t *o, double *sums);

/* Multicanonical: install a weight whose pos/W arrays (bins+2 entries)
 * already carry the virtual end bins, and compute param_value[]. */
void orc_muca_set(orc_t *o, const su2hip_muca *m, const double *pos,
                  const double *W);
void orc_muca_get(const orc_t *o, su2hip_muca *m, double *pos, double *W,
                  int32_t *hits);
/* Correlators along dir of the reference's CORRELATORS build
 * (measure_correlators, src/correlation.c:339-391): for d = 0..L[dir]/2-1,
 * 10 values: H(d), Sigma^2 correlator, photon (mode 0, p=1) re im, (mode 0,
 * p=2) re im, (mode 1, p=1) re im, (mode 1, p=2) re im; absent fields 0.
 * Returns the number of distances. */
int orc_correlators(const orc_t *o, int dir, double *out);
int orc_longest_dir(const orc_t *o);
/* Two-phase start of the MEASURE_Z build (prepare_wall) and its plane
 * averages (measure_along_z); dir -1 = the longest direction. */
void orc_wall_start(orc_t *o, int zdir);
/* Correlators on the blocked lattices of the BLOCKING build, out[level][
 * L[longest]/2][10] as orc_correlators; returns the number of levels (0
 * with a doublet, where the reference measures nothing). */
int orc_blocked_correlators(const orc_t *o, int levels, double *out);
int orc_zprofile(const orc_t *o, int dir, double *out);

/* gsum/nsum of the SLOW recursion (bins+2 entries; load_weight_params,
 * src/multicanonical.c:291-333, initialises them to 5 and 1). */
void orc_muca_set_sums(orc_t *o, const int64_t *gsum, const int64_t *nsum);
void orc_muca_get_sums(const orc_t *o, int64_t *gsum, int64_t *nsum);
double orc_muca_weight(const orc_t *o, double val);
int orc_muca_whichbin(const orc_t *o, double val);
double orc_muca_orderparam(orc_t *o, int parity);

/* Gradient flow. */
void orc_flow_force(const orc_t *o, double *force /* [V][3][4] */,
                    double *tforce /* [V][3] or NULL */);
/* One flow step on the lattice's own fields (scheme EULER or RK3). */
void orc_flow_step(orc_t *o, int scheme, double dt);
/* grad_flow on a copy of the fields; rows of {t, sums[NSUMS]}. */
int orc_flow(orc_t *o, int scheme, double dt, double t_max, int meas_interval,
             double *rows, int max_rows);

#ifdef __cplusplus
}
#endif

#endif
