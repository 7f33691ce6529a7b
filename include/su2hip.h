/* su2hip.h -- C ABI of the MI355X (gfx950) lattice update sweep for 3d SU(2)
 * (+U(1)) gauge theory with fundamental / adjoint / singlet scalars.
 *
 * One handle (su2hip_batch) owns one or more independent lattices that are
 * swept together on one HIP stream: every kernel launch covers every active
 * lattice of the batch.  A batch of one is the plain single-lattice run.
 * Lattices of one batch share the field content, precision and pass schedule
 * (algorithms and per-sweep update counts); they may differ in extents,
 * couplings and RNG seed.
 *
 * Fields cross this boundary in the reference's host layout and site order
 * (src/alloc.c:30-75, src/layout.c:826-851): su2link double[V][3][4],
 * u1link double[V][3], doublet double[nhiggs][V][4], triplet double[V][3],
 * singlet double[V], with the V sites ordered "even parity first, odd
 * second, natural x-fastest order inside each parity".  All functions return
 * 0 on success and a negative SU2HIP_E* code otherwise; su2hip_error_string()
 * and su2hip_last_error() describe the failure.  No function calls exit().
 *
 * Threading: a handle is used by one host thread at a time.
 */
#ifndef SU2HIP_H
#define SU2HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Identifiers with the reference's values (src/generic/stddefs.h:46-65). */
#define SU2HIP_METROPOLIS 1
#define SU2HIP_HEATBATH 2
#define SU2HIP_OVERRELAX 3

#define SU2HIP_EVEN 0
#define SU2HIP_ODD 1

#define SU2HIP_PHISQ 1
#define SU2HIP_SIGMASQ 2
#define SU2HIP_PHI2MINUSSIGMA2 3
#define SU2HIP_PHI2SQ 4

#define SU2HIP_READONLY 0
#define SU2HIP_FAST 1
#define SU2HIP_SLOW 2

/* Arithmetic of the sweep kernels.  Reductions, multicanonical weights and
 * the global accept test are FP64 in both modes. */
#define SU2HIP_FP64 64
#define SU2HIP_FP32 32

/* FP32 link storage (su2hip_params.link_format; FP64 always stores four
 * components).  FULL: four floats per link (16 bytes).  PACKED: three
 * floats (12 bytes), the stereographic projection of the unit quaternion
 * from the pole opposite the sign of its real part, that sign in the lowest
 * mantissa bit; a load rebuilds the link with one reciprocal, which costs a
 * little arithmetic and saves a quarter of the link bytes.  Every FP32 link
 * is renormalised when it is stored, in both forms.  AUTO, the default: a
 * handle created with its lattices takes PACKED when their FP32 fields in
 * the full form exceed SU2HIP_LINKS_PACKED_ABOVE bytes (above it the saved
 * bytes outweigh the rebuild, below it the rebuild does), FULL otherwise; a
 * handle created empty (lattices added later) takes PACKED.  A handle uses
 * one format for all its lattices: the first lattice's setting decides, or
 * the measured one when it is AUTO. */
#define SU2HIP_LINKS_AUTO 0
#define SU2HIP_LINKS_FULL 1
#define SU2HIP_LINKS_PACKED 2
#define SU2HIP_LINKS_PACKED_ABOVE (32ull << 20)

/* Gradient-flow integrators. */
#define SU2HIP_FLOW_EULER 1 /* the reference's scheme, src/gradflow.c:14-18 */
#define SU2HIP_FLOW_RK3 3   /* Luscher's 3rd-order Runge-Kutta, arXiv:1006.4518 */

/* Error codes. */
#define SU2HIP_OK 0
#define SU2HIP_EINVAL -1      /* bad argument */
#define SU2HIP_ENOMEM -2      /* device or host allocation failed / budget exceeded */
#define SU2HIP_EDEVICE -3     /* HIP runtime error */
#define SU2HIP_ERANGE -4      /* index or size out of range */
#define SU2HIP_EUNSUPPORTED -5/* model combination not supported */
#define SU2HIP_ESTATE -6      /* call not valid in the current state */
#define SU2HIP_EIO -7         /* file I/O or parse error (host I/O layer) */

/* Run parameters of one lattice.  The scalar couplings keep the reference's
 * names and meaning (src/su2.h:76-181, parsed at src/parameters.c:157-338).
 * Field content is a runtime choice here: nhiggs (0/1/2), u1, triplet and
 * singlet replace the reference's -DNHIGGS/-DU1/-DTRIPLET/-DSINGLET flags.
 * nhiggs = 2 is the two-Higgs-doublet model (-DNHIGGS=2), with or without
 * U(1); its couplings are the last fields (src/su2.h:110-114). */
typedef struct su2hip_params {
  int32_t L[3];              /* extents; all even */
  int32_t nhiggs;            /* fundamental doublets: 0, 1 or 2 */
  int32_t u1;                /* compact U(1) hypercharge links */
  int32_t triplet;           /* adjoint scalar */
  int32_t singlet;           /* real singlet scalar */
  int32_t precision;         /* SU2HIP_FP64 or SU2HIP_FP32 */
  int32_t algorithm_su2link; /* HEATBATH or METROPOLIS */
  int32_t algorithm_u1link;  /* METROPOLIS */
  int32_t algorithm_su2doublet; /* OVERRELAX or METROPOLIS */
  int32_t algorithm_su2triplet; /* OVERRELAX or METROPOLIS */
  int32_t algorithm_singlet;    /* OVERRELAX or METROPOLIS */
  int32_t update_links;      /* link rounds per sweep (update_links) */
  int32_t link_overrelax;    /* SU(2) link overrelaxation rounds after each link round (no reference counterpart) */
  int32_t scalar_sweeps;     /* scalar rounds per sweep (scalar_sweeps) */
  int32_t update_su2doublet; /* doublet passes per scalar round (update_doublet) */
  int32_t update_su2triplet; /* triplet passes per scalar round (update_triplet) */
  int32_t update_singlet;    /* singlet passes per scalar round (update_singlet) */
  int32_t reunitarize;       /* FP32: links are stored unit (SU2HIP_LINKS_*); kept for
                                the ABI, no further effect */
  int32_t link_format;       /* FP32 link storage: SU2HIP_LINKS_AUTO / _FULL / _PACKED */
  double betasu2;
  double betau1, r_u1;
  double msq_phi, lambda_phi;
  double msq_triplet, a2, b4;
  double b1_s, msq_s, b3_s, b4_s, a1_s, a2_s;
  uint64_t seed;             /* Philox key of this lattice */
  /* two-doublet potential (nhiggs = 2, src/su2u1.c:537-554): msq_phi and
   * lambda_phi above are m1^2 and lam1 */
  double msq_phi2;
  double m12sq_re, m12sq_im;
  double lam2, lam3, lam4;
  double lam5_re, lam5_im, lam6_re, lam6_im, lam7_re, lam7_im;
} su2hip_params;

/* Accept/total counters, the reference's `counters` (src/su2.h:208-231) for
 * one lattice, plus the link overrelaxation counts.  The *_doublet members
 * count the first doublet, the *_doublet2 members the second. */
typedef struct su2hip_counters {
  int64_t total_su2link, accepted_su2link;
  int64_t total_u1link, accepted_u1link;
  int64_t total_doublet, accepted_doublet;
  int64_t total_overrelax_doublet, acc_overrelax_doublet;
  int64_t total_triplet, accepted_triplet;
  int64_t total_overrelax_triplet, acc_overrelax_triplet;
  int64_t total_singlet, accepted_singlet;
  int64_t total_overrelax_singlet, acc_overrelax_singlet;
  int64_t total_muca, accepted_muca;
  int64_t total_link_or, accepted_link_or;
  int64_t total_doublet2, accepted_doublet2;
  int64_t total_overrelax_doublet2, acc_overrelax_doublet2;
} su2hip_counters;

/* Multicanonical state of one lattice (the scalar part of the reference's
 * `weight`, src/su2.h:235-274).  The bins+2 arrays pos[] and W[] travel
 * separately; index 0 and bins+1 are the reference's virtual end bins. */
typedef struct su2hip_muca {
  int32_t enabled;           /* multicanonical != 0 */
  int32_t orderparam;        /* SU2HIP_PHISQ / SIGMASQ / PHI2MINUSSIGMA2 / PHI2SQ */
  int32_t mode;              /* SU2HIP_READONLY / FAST / SLOW */
  int32_t checks_per_sweep;
  int32_t bins;
  int32_t last_max;
  int32_t do_acceptance;
  int32_t muca_count;
  int32_t update_interval;
  int32_t reserved;
  double delta, min, max, wrk_min, wrk_max;
  double param_value[2];     /* order parameter contribution of EVEN / ODD sites */
} su2hip_muca;

/* Layout of the FP64 volume sums returned by su2hip_measure().  Every model
 * uses the same slots; absent fields leave theirs at 0.  The column values
 * of the reference's `measure` file (src/measure.c:252-306) are these sums
 * divided as su2hip_measure_columns() does. */
enum {
  SU2HIP_SUM_ACTION = 0,
  SU2HIP_SUM_WILSON,        /* sum of local_su2wilson (includes beta) */
  SU2HIP_SUM_U1WILSON,
  SU2HIP_SUM_COV_PHI,
  SU2HIP_SUM_PHI2,
  SU2HIP_SUM_PHI4,
  SU2HIP_SUM_HOP_SIGMA,
  SU2HIP_SUM_SIGMA2,
  SU2HIP_SUM_SIGMA4,
  SU2HIP_SUM_PHI2SIGMA2,
  SU2HIP_SUM_MAGCHARGE,
  SU2HIP_SUM_MAGCHARGE_ABS,
  SU2HIP_SUM_S_COV,
  SU2HIP_SUM_S,
  SU2HIP_SUM_S2,
  SU2HIP_SUM_S3,
  SU2HIP_SUM_S4,
  SU2HIP_SUM_SPHISQ,
  SU2HIP_SUM_S2PHISQ,
  SU2HIP_SUM_COV_PHI_2,     /* second doublet (nhiggs = 2): covariant, phi^2, phi^4 */
  SU2HIP_SUM_PHI2_2,
  SU2HIP_SUM_PHI4_2,
  SU2HIP_SUM_R12,           /* Re phi1^+ phi2 */
  SU2HIP_SUM_I12,           /* Im phi1^+ phi2 */
  SU2HIP_NSUMS
};

/* Kinds of kernel launches, used by su2hip_pass() and the timing API. */
enum {
  SU2HIP_K_LINK_HEATBATH = 0,
  SU2HIP_K_LINK_METROPOLIS,
  SU2HIP_K_LINK_OVERRELAX,
  SU2HIP_K_U1_METROPOLIS,
  SU2HIP_K_DOUBLET_OVERRELAX,
  SU2HIP_K_DOUBLET_METROPOLIS,
  SU2HIP_K_TRIPLET_OVERRELAX,
  SU2HIP_K_TRIPLET_METROPOLIS,
  SU2HIP_K_SINGLET_OVERRELAX,
  SU2HIP_K_SINGLET_METROPOLIS,
  SU2HIP_K_MEASURE,
  SU2HIP_K_MUCA,
  SU2HIP_K_FLOW,
  SU2HIP_K_AUX,
  SU2HIP_NKINDS
};

/* Opaque handle: one or more lattices on one GPU stream. */
typedef struct su2hip_batch su2hip_batch_t;

/* ---- errors and device ---------------------------------------------- */

/* Text for an error code. */
const char *su2hip_error_string(int code);
/* Detailed message of the last failure on this host thread. */
const char *su2hip_last_error(void);

/* ---- handle lifecycle --------------------------------------------------
 * Replaces the reference's layout()+alloc_fields() (src/su2.h:310,335;
 * src/layout.c:657-704, src/alloc.c:103-125) for n lattices at once.
 * max_slots >= n reserves room for lattices added later (serving); n = 0
 * makes an empty handle whose template is params[0];
 * hbm_budget (bytes, 0 = unlimited) caps the field memory of all slots. */
int su2hip_create(int n, const su2hip_params *params, int max_slots,
                  uint64_t hbm_budget, su2hip_batch_t **out);
int su2hip_destroy(su2hip_batch_t *b);

/* Device bytes one lattice with these params occupies (fields, multicanonical
 * backup, and, with with_flow, the flow scratch). */
int su2hip_footprint(const su2hip_params *p, int with_flow, uint64_t *bytes);

/* Add a lattice in a free slot (fields cold-started); fails with
 * SU2HIP_ENOMEM when the HBM budget or the slots are exhausted. */
int su2hip_add(su2hip_batch_t *b, const su2hip_params *p, int *slot);
/* Free a slot and its device memory. */
int su2hip_remove(su2hip_batch_t *b, int slot);
/* Number of slots / whether a slot is in use. */
int su2hip_slots(su2hip_batch_t *b, int *max_slots, int *active);
int su2hip_slot_active(su2hip_batch_t *b, int slot, int *active);
/* The parameters a slot was created with. */
int su2hip_slot_params(su2hip_batch_t *b, int slot, su2hip_params *p);

/* ---- fields ------------------------------------------------------------
 * Host <-> device in the reference layout (src/checkpoint.c:101-115 order of
 * fields; a NULL pointer skips a field).  Replaces load_lattice/save_lattice's
 * field transfer (src/su2.h:458-459). */
int su2hip_set_fields(su2hip_batch_t *b, int slot, const double *su2link,
                      const double *u1link, const double *doublet,
                      const double *triplet, const double *singlet);
int su2hip_get_fields(su2hip_batch_t *b, int slot, double *su2link,
                      double *u1link, double *doublet, double *triplet,
                      double *singlet);
/* Cold start, the reference's setfields() (src/init.c:92-112). */
int su2hip_cold_start(su2hip_batch_t *b, int slot, double phi0, double sigma0,
                      double singlet0);
/* Hot start from the lattice's Philox key: links normalised N(0,1)
 * quaternions, scalar components N(0,1), U(1) angles uniform in [-pi,pi).
 * No reference counterpart (the reference only cold-starts). */
int su2hip_hot_start(su2hip_batch_t *b, int slot);

/* Two-phase start with an interface across direction dir (-1: the
 * reference's z direction, the longest, of equal ones the last): the
 * structure and value ranges of prepare_wall (src/z_coord.c:304-376, the
 * MEASURE_Z build's setup_wall) with the small deviations drawn from the
 * lattice's Philox key. */
int su2hip_wall_start(su2hip_batch_t *b, int slot, int dir);

/* ---- sweep ---------------------------------------------------------------
 * n_sweeps compound sweeps of every active lattice whose sweep budget is not
 * exhausted; one sweep is the reference's update_lattice() (src/su2.h:426,
 * src/update.c:270-430) with link_overrelax extra link rounds.
 * Asynchronous: returns once the work is queued. */
int su2hip_update(su2hip_batch_t *b, int n_sweeps);
/* Per-lattice sweep budget: a lattice with budget 0 is skipped by
 * su2hip_update(); each completed sweep decrements it.  -1 = unlimited. */
int su2hip_set_budget(su2hip_batch_t *b, int slot, int64_t sweeps);
/* Budgets of all slots (max_slots entries, -2 for an empty slot). */
int su2hip_get_budgets(su2hip_batch_t *b, int64_t *budgets);
/* One checkerboard pass of one kind over every active lattice (the
 * reference's checkerboard_sweep_* for one parity and direction,
 * src/update.c:35-259).  For the doublet kinds of a two-doublet model dir
 * is the doublet (0 or 1, the reference's higgs_id, > 1 taken as 1); other
 * scalar passes ignore it. */
int su2hip_pass(su2hip_batch_t *b, int kind, int parity, int dir);
/* Capture each compound sweep into a hipGraph that is replayed (1) or launch
 * kernels one by one (0, default). */
int su2hip_use_graph(su2hip_batch_t *b, int enable);
/* Wait for all queued work. */
int su2hip_sync(su2hip_batch_t *b);
/* The reference's random_sweeps (src/update.c:292-300, :341-345): every
 * sweep of every lattice takes its link passes in a shuffled (parity,
 * direction) order (parities and directions shuffled separately, as
 * shuffle() does, src/update.c:461-473) and each scalar round's parities in
 * a random order.  The draws come from the lattice's Philox stream at (pass
 * counter at the sweep start, site 0xFFFFFFFE), on the device, so lattices
 * of one batch take their own orders in the same launches.  Off by default,
 * as the reference's main sets it (src/main.c:187); its heatbath
 * trajectories switch it on. */
int su2hip_set_random_order(su2hip_batch_t *b, int enable);
int su2hip_get_random_order(su2hip_batch_t *b, int *enabled);

/* Counters of a lattice since the last reset (src/init.c:167-199). */
int su2hip_get_counters(su2hip_batch_t *b, int slot, su2hip_counters *c);
int su2hip_reset_counters(su2hip_batch_t *b, int slot);
/* RNG position: Philox key and the 64-bit pass counter (for exact resume). */
int su2hip_rng_get(su2hip_batch_t *b, int slot, uint64_t *seed, uint64_t *pass);
int su2hip_rng_set(su2hip_batch_t *b, int slot, uint64_t seed, uint64_t pass);

/* ---- measurement -----------------------------------------------------------
 * FP64 volume sums of the reference's measure() site loop
 * (src/su2.h:475, src/measure.c:110-174), deterministic tree order.
 * sums has SU2HIP_NSUMS entries. */
int su2hip_measure(su2hip_batch_t *b, int slot, double *sums);
/* Sums of every slot at once: sums[max_slots][SU2HIP_NSUMS] (empty slots 0). */
int su2hip_measure_all(su2hip_batch_t *b, double *sums);
/* The sums of su2hip_measure for one lattice, allocating nothing: no
 * reference counterpart; for callers that cannot allocate device memory
 * (HIP graph capture, tight serving loops).  sums_device is device memory
 * of SU2HIP_NSUMS doubles; the kernels are queued on the handle's stream
 * after the work already there, and the call returns without waiting (the
 * values are bitwise those of su2hip_measure).  The block partial sums go
 * to a scratch array the library holds on every GPU it runs on: lattices of
 * at most 2 * 256 * SU2HIP_NOALLOC_MAX_BLOCKS sites (SU2HIP_ERANGE beyond).
 * The handle's device must be current.  Calls on one GPU share that scratch:
 * calls whose kernels overlap (other streams, other threads) race. */
#define SU2HIP_NOALLOC_MAX_BLOCKS 8192
int su2hip_measure_noalloc(su2hip_batch_t *b, int slot, double *sums_device);
/* Column values as written to `measure` (src/measure.c:252-306) from sums,
 * for the field content of p; weight/muca_param fill the first two columns.
 * Returns the number of columns in *ncols. */
int su2hip_measure_columns(const su2hip_params *p, const double *sums,
                           double weight, double muca_param, double *cols,
                           int *ncols);

/* Correlators of the reference's CORRELATORS build (measure_correlators,
 * src/correlation.c:339-391) along direction dir (-1: the longest, as the
 * reference picks it, src/layout.c:771-777).  For d = 0 .. L[dir]/2-1
 * (at most max_dist rows), out[d][10] holds H(d) = 4/V sum_z h(z) h(z+d)
 * with h the plane sums of Tr Phi^+ Phi / 2, the Sigma^2 correlator, and
 * the photon correlators of photon_correl as re, im pairs: mode 0 at
 * transverse momentum 1 and 2, mode 1 at 1 and 2.  Entries of absent fields
 * are 0.  Plane sums in FP64.  *ndist = L[dir]/2. */
int su2hip_correlators(su2hip_batch_t *b, int slot, int dir, double *out, int max_dist,
                       int *ndist);

/* Correlators on blocked lattices (the BLOCKING build: main.c:106-152,
 * measure_blocked_correlators, src/correlation.c:393-421): every direction
 * but the longest is halved per level, the links of those directions and the
 * triplet smeared (smear_link, smear_triplet, src/su2u1.c:752-857), up to
 * `levels` levels (capped as max_block_level, src/blocking.c:19-38).
 * out[level][L/2][10] in su2hip_correlators' column order with L the
 * longest extent.  Triplet-only models; with a doublet the reference
 * measures nothing and *nlevels = 0. */
int su2hip_blocked_correlators(su2hip_batch_t *b, int slot, int levels, double *out,
                               int *nlevels, int *ndist);

/* Plane averages along direction dir (-1: the longest, of equal ones the
 * last) of measure_along_z (src/z_coord.c:206-298, the MEASURE_Z build):
 * out[z][4] = Tr Phi^+Phi/2, Tr Sigma^2/2, S, S^2 averaged over the plane
 * (0 for absent fields).  *nrows = L[dir]. */
int su2hip_zprofile(su2hip_batch_t *b, int slot, int dir, double *out, int max_rows,
                    int *nrows);

/* ---- multicanonical --------------------------------------------------------
 * Install the weight of a lattice (load_weight's result, src/su2.h:483,
 * src/multicanonical.c:118-288) and compute both parity contributions of
 * the order parameter (calc_orderparam, src/multicanonical.c:709-751).
 * pos/W have m->bins+2 entries.  The global accept/reject then runs on the
 * device inside every scalar pass of the order-parameter field. */
int su2hip_muca_set(su2hip_batch_t *b, int slot, const su2hip_muca *m,
                    const double *pos, const double *W);
/* Read back the weight state (pos/W/hits may be NULL). */
int su2hip_muca_get(su2hip_batch_t *b, int slot, su2hip_muca *m, double *pos,
                    double *W, int32_t *hits);
/* gsum/nsum of the SLOW recursion (bins+2 entries each; either may be NULL).
 * su2hip_muca_set starts them at 5 and 1 as load_weight_params does without
 * a file (src/multicanonical.c:296-303). */
int su2hip_muca_set_sums(su2hip_batch_t *b, int slot, const int64_t *gsum,
                         const int64_t *nsum);
int su2hip_muca_get_sums(su2hip_batch_t *b, int slot, int64_t *gsum, int64_t *nsum);
/* Multicanonical accept/reject of a lattice on (the default once a weight
 * is installed) or off (w->do_acceptance, src/su2.h:244): off, sweeps run
 * without checks or backups and the weight is not updated, as in the
 * reference's heatbath trajectories (src/hb_trajectory.c:15-19). */
int su2hip_muca_acceptance(su2hip_batch_t *b, int slot, int enable);
/* The order parameter's parity contributions (param_value[EVEN], [ODD]) and
 * their sum; with recompute = 1 both are first recomputed from the fields
 * (calc_orderparam of EVEN and ODD, src/multicanonical.c:709-751).  pv or
 * value may be NULL. */
int su2hip_muca_orderparam(su2hip_batch_t *b, int slot, int recompute, double *pv,
                           double *value);
/* Overwrite the parity contributions, e.g. with the values the reference
 * leaves after its trajectories (computed on their copy of the fields). */
int su2hip_muca_set_orderparam(su2hip_batch_t *b, int slot, const double *pv);
/* W(R) of the installed weight (get_weight, src/multicanonical.c:350-379). */
int su2hip_muca_weight_of(su2hip_batch_t *b, int slot, double value,
                          double *weight);

/* ---- gradient flow ---------------------------------------------------------
 * Flow a copy of the lattice's links (+ triplet) from t=0 to t_max
 * (grad_flow, src/su2.h:543, src/gradflow.c:46-173); the lattice itself is
 * not changed.  Every meas_interval steps (and at the end) a row of
 * 1 + SU2HIP_NSUMS doubles {t, sums...} is stored in rows (row 0 is t=0).
 * *nrows returns the number of rows written (at most max_rows). */
int su2hip_flow(su2hip_batch_t *b, int slot, int scheme, double dt,
                double t_max, int meas_interval, double *rows, int max_rows,
                int *nrows);
/* The same flow, asynchronous: enqueued on the handle's second stream after
 * the work already queued on its main stream, so the other lattices keep
 * sweeping meanwhile.  The lattice must have sweep budget 0 (it must not
 * change while its copy flows); its budget, fields and starts are refused
 * (SU2HIP_ESTATE) until su2hip_flow_finish.  One flow per slot at a time. */
int su2hip_flow_start(su2hip_batch_t *b, int slot, int scheme, double dt, double t_max,
                      int meas_interval);
/* *done = 1 once the flow of `slot` has finished on the device. */
int su2hip_flow_poll(su2hip_batch_t *b, int slot, int *done);
/* Wait for the flow of `slot`; rows as su2hip_flow returns them. */
int su2hip_flow_finish(su2hip_batch_t *b, int slot, double *rows, int max_rows, int *nrows);
/* Gradient force on every link, double[V][3][4] with component 0 = 0
 * (calc_gradient's link part, src/gradflow.c:335-354). */
int su2hip_flow_force(su2hip_batch_t *b, int slot, double *force);

/* ---- timing ----------------------------------------------------------------
 * With timing on, every launch is bracketed by HIP events on the handle's
 * stream; su2hip_kernel_times() returns the summed milliseconds and the
 * launch count per SU2HIP_K_* kind since the last reset. */
int su2hip_set_timing(su2hip_batch_t *b, int enable);
int su2hip_kernel_times(su2hip_batch_t *b, double *ms, int64_t *launches);
int su2hip_reset_timing(su2hip_batch_t *b);
/* Durations of `repeats` back-to-back launches of one pass kind over every
 * active lattice on the handle's stream: *ms_kernel is the average kernel
 * duration (start and end stamped on each dispatch by hipExtLaunchKernel's
 * events, the interval a kernel profiler reports), *ms_span the average
 * share of the whole span including the gaps between dependent launches.
 * Directions and parities cycle as in a sweep; the fields are updated and
 * each lattice consumes `repeats` pass numbers and one sweep of its budget.
 * Either output may be NULL. */
int su2hip_time_pass(su2hip_batch_t *b, int kind, int repeats, double *ms_kernel,
                     double *ms_span);
/* ---- launch shape and autotuning -----------------------------------------
 * Pass kernels run one thread per site of one parity in workgroups of
 * 64, 128 or 256 threads, each over a contiguous chunk of that many sites
 * of one lattice (chunks dealt to the 8 XCDs in contiguous runs).  The
 * chunk of a pass kind is the one set here, or with 0 (default) the widest
 * that still gives 4 workgroups per CU for the batch.  Results do not depend
 * on the chunk.  Multicanonical steps always use 256. */
int su2hip_set_chunk(su2hip_batch_t *b, int kind, int chunk);
/* The chunk a pass kind launches with now. */
int su2hip_get_chunk(su2hip_batch_t *b, int kind, int *chunk);
/* Occupancy target of a pass kind's kernel: 0 = the compiler's own register
 * allocation, 5 = registers capped so that 5 waves fit per SIMD.  The link
 * overrelaxation is built both ways (default 5); every other kind only
 * with 0, and asking it for 5 fails with SU2HIP_EINVAL.  Which one is
 * faster depends on the lattice; results do not depend on it. */
int su2hip_set_waves(su2hip_batch_t *b, int kind, int waves);
/* The occupancy target a pass kind launches with now. */
int su2hip_get_waves(su2hip_batch_t *b, int kind, int *waves);
/* Hopping-term cache (no reference counterpart; the doublet part of the
 * link staple, src/staples.c:150-190).  With one doublet the handle can keep,
 * per lattice, -1/2 Phi(x+mu) exp(-i alpha sigma_3) Phi(x)^+ for every link
 * (4 values per link, 12 per site, inside the HBM budget), filled before
 * each round of link passes: the link passes then read it instead of two
 * doublets and a U(1) link.  The arithmetic is the same either way; FP64
 * results agree to rounding, FP32 ones to FP32 rounding.  Default: kept,
 * except when the handle is created with lattices whose fields fit in
 * SU2HIP_INFINITY_CACHE_BYTES and whose fields and caches together do not
 * (there the other passes lose more to the footprint than the link passes
 * gain).  A lattice added to a handle that keeps it gets one if it fits the
 * budget, and otherwise runs without.  Enabling it for a model without
 * exactly one doublet fails with SU2HIP_EINVAL; enabling it beyond the
 * budget fails with SU2HIP_ENOMEM and leaves it off for every lattice. */
#define SU2HIP_INFINITY_CACHE_BYTES (256ull << 20)
int su2hip_set_hop_cache(su2hip_batch_t *b, int enable);
/* 1 if the handle keeps the hopping-term cache for every one of its
 * lattices, else 0. */
int su2hip_get_hop_cache(su2hip_batch_t *b, int *enabled);
/* Time every pass kind of the batch's sweep with each chunk, and with each
 * occupancy target it is built with (kernel durations of `repeats`
 * back-to-back launches, as su2hip_time_pass), and keep the fastest; 256
 * stays unless a smaller chunk is 2% faster, and the default target unless
 * the other is 1% faster.  Fields, RNG positions, budgets and counters are
 * restored afterwards (the handle needs one more copy of its lattices'
 * memory meanwhile).  table, if not NULL, receives the milliseconds as
 * table[kind * 3 + {0,1,2}] for chunks 64, 128, 256, the faster target's
 * for each chunk (entries of kinds not in the sweep are left alone). */
int su2hip_autotune(su2hip_batch_t *b, int repeats, double *table);

/* Copy `bytes` of device memory with the library's streaming copy kernel and
 * return its duration (ms): the calibration kernel for HBM counters. */
int su2hip_copy_probe(uint64_t bytes, int repeats, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* SU2HIP_H */
