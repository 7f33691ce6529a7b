/* su2hip_bind.c -- drive su2higgs's sweep on an MI355X through libsu2hip. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "su2.h"
#include "generic/mersenne.h"
#include "su2hip.h"

static su2hip_batch_t *H;

static su2hip_params to_su2hip(lattice const *l, params const *p) {
  su2hip_params q = {0};
  for (int d = 0; d < 3; d++) q.L[d] = (int)l->L[d];
  q.nhiggs = NHIGGS;
#ifdef U1
  q.u1 = 1; q.betau1 = p->betau1; q.r_u1 = p->r_u1;
#endif
#ifdef TRIPLET
  q.triplet = 1; q.msq_triplet = p->msq_triplet; q.b4 = p->b4; q.a2 = p->a2;
  q.algorithm_su2triplet = p->algorithm_su2triplet;
  q.update_su2triplet = p->update_su2triplet;
#endif
#ifdef SINGLET
  q.singlet = 1; q.update_singlet = p->update_singlet;
  q.algorithm_singlet = p->algorithm_singlet;
  q.b1_s = p->b1_s; q.msq_s = p->msq_s; q.b3_s = p->b3_s; q.b4_s = p->b4_s;
  q.a1_s = p->a1_s; q.a2_s = p->a2_s;
#endif
  q.precision = SU2HIP_FP64;          /* or SU2HIP_FP32 */
  q.algorithm_su2link = p->algorithm_su2link;
  q.algorithm_u1link = SU2HIP_METROPOLIS;
  q.algorithm_su2doublet = p->algorithm_su2doublet;
  q.update_links = p->update_links;
  q.scalar_sweeps = p->scalar_sweeps;
  q.update_su2doublet = p->update_su2doublet;
  q.betasu2 = p->betasu2;
  q.msq_phi = p->msq_phi;
  q.lambda_phi = p->lambda_phi;
#if (NHIGGS == 2)
  q.msq_phi2 = p->msq_phi2;
  q.m12sq_re = p->m12sq.re; q.m12sq_im = p->m12sq.im;
  q.lam2 = p->lam2; q.lam3 = p->lam3; q.lam4 = p->lam4;
  q.lam5_re = p->lam5.re; q.lam5_im = p->lam5.im;
  q.lam6_re = p->lam6.re; q.lam6_im = p->lam6.im;
  q.lam7_re = p->lam7.re; q.lam7_im = p->lam7.im;
#endif
  /* Philox key: the next word of the reference's own generator, which
   * main() seeds from the clock (src/main.c:59-78), so every run is a new
   * chain.  A resumed run takes its key back from <latticefile>.rng. */
  q.seed = iran();
  return q;
}

/* the field blocks of `f` in su2hip_set_fields/get_fields order; NULL for
 * a field this build does not have */
#ifdef U1
#define U1_OF(f) ((f)->u1link[0])
#else
#define U1_OF(f) NULL
#endif
#if (NHIGGS == 2)
/* the two doublets are two allocations (src/alloc.c:109); su2hip takes them
 * as one array, the first doublet first: staged through Dstage */
static double *Dstage;
static long Vsites;
#define DOUBLET_OF(f) Dstage
#elif (NHIGGS == 1)
#define DOUBLET_OF(f) ((f)->su2doublet[0][0])
#else
#define DOUBLET_OF(f) NULL
#endif
#ifdef TRIPLET
#define TRIPLET_OF(f) ((f)->su2triplet[0])
#else
#define TRIPLET_OF(f) NULL
#endif
#ifdef SINGLET
#define SINGLET_OF(f) ((f)->singlet[0])
#else
#define SINGLET_OF(f) NULL
#endif

/* <latticefile>.rng: the Philox key and pass counter beside the lattice file */
static void rng_path(char *out, size_t n, const char *latticefile) {
  snprintf(out, n, "%s.rng", latticefile);
}

/* once, after alloc_fields/setfields or load_lattice (src/main.c:160-200).
 * A run that continues from a lattice file (reset 0, src/main.c:280-282)
 * continues its Philox stream from <latticefile>.rng when that exists. */
int su2hip_bind_init(lattice const *l, fields const *f, params const *p) {
  su2hip_params q = to_su2hip(l, p);
  int rc = su2hip_create(1, &q, 1, 0, &H);
  if (rc) return rc;
  su2hip_use_graph(H, 1);
  char rp[sizeof(p->latticefile) + 8];
  rng_path(rp, sizeof(rp), p->latticefile);
  FILE *fr = p->reset ? NULL : fopen(rp, "r");
  unsigned long long key = q.seed, pass = 0;
  if (fr) {
    int got = fscanf(fr, "%llu %llu", &key, &pass);
    fclose(fr);
    if (got != 2) return SU2HIP_EIO;
    rc = su2hip_rng_set(H, 0, key, pass);
    if (rc) return rc;
  }
  printf("su2hip: Philox key %llu, pass %llu\n", key, pass);
#if (NHIGGS == 2)
  Vsites = l->sites;
  Dstage = malloc(sizeof(double) * 2 * SU2DB * Vsites);
  if (!Dstage) return SU2HIP_ENOMEM;
  for (int db = 0; db < 2; db++)
    memcpy(Dstage + db * SU2DB * Vsites, f->su2doublet[db][0], sizeof(double) * SU2DB * Vsites);
#endif
  return su2hip_set_fields(H, 0, f->su2link[0][0], U1_OF(f), DOUBLET_OF(f), TRIPLET_OF(f),
                           SINGLET_OF(f));
}

/* after save_lattice(l, f, c, fname) (src/su2.h:458, src/main.c:372,391):
 * the Philox position, so that a resumed run replays nothing */
int su2hip_bind_save_rng(const char *latticefile) {
  if (!H) return SU2HIP_OK; /* no sweep yet: nothing drawn */
  uint64_t key, pass;
  int rc = su2hip_rng_get(H, 0, &key, &pass);
  if (rc) return rc;
  char rp[256];
  rng_path(rp, sizeof(rp), latticefile);
  FILE *fw = fopen(rp, "w");
  if (!fw) return SU2HIP_EIO;
  fprintf(fw, "%llu %llu\n", (unsigned long long)key, (unsigned long long)pass);
  return fclose(fw) ? SU2HIP_EIO : SU2HIP_OK;
}

/* replaces update_lattice(&l, f, &p, &c, &w) (src/su2.h:426, src/main.c:353) */
int su2hip_bind_update(void) { return su2hip_update(H, 1); }

/* before measure() or save_lattice(): bring the fields back into `f`
 * (or call su2hip_measure + su2hip_measure_columns and print them instead). */
int su2hip_bind_download(fields *f) {
  int rc = su2hip_get_fields(H, 0, f->su2link[0][0], U1_OF(f), DOUBLET_OF(f), TRIPLET_OF(f),
                             SINGLET_OF(f));
#if (NHIGGS == 2)
  for (int db = 0; rc == 0 && db < 2; db++)
    memcpy(f->su2doublet[db][0], Dstage + db * SU2DB * Vsites, sizeof(double) * SU2DB * Vsites);
#endif
  return rc;
}

/* multicanonical runs: after load_weight and calc_orderparam
 * (src/main.c:89, :216-217), the weight and its state move to the device,
 * which then does the global accept/reject inside every scalar pass
 * (multicanonical_acceptance, src/update.c:13-30) and, in FAST and SLOW
 * modes, the weight recursion.  Call again when the main loop changes
 * w->mode (src/main.c:255-278). */
int su2hip_bind_muca(weight const *w) {
  su2hip_muca m = {0};
  m.enabled = 1;
  m.orderparam = w->orderparam;      /* PHISQ .. PHI2SQ: the same codes */
  m.mode = w->mode;                  /* READONLY, FAST, SLOW: the same codes */
  m.checks_per_sweep = w->checks_per_sweep;
  m.bins = w->bins;
  m.last_max = w->last_max;
  m.do_acceptance = w->do_acceptance;
  m.muca_count = w->muca_count;
  m.update_interval = w->update_interval;
  m.delta = w->delta;
  m.min = w->min;
  m.max = w->max;
  m.wrk_min = w->wrk_min;
  m.wrk_max = w->wrk_max;
  m.param_value[0] = w->param_value[EVEN];
  m.param_value[1] = w->param_value[ODD];
  int rc = su2hip_muca_set(H, 0, &m, w->pos, w->W);
  if (rc == 0 && w->mode == SLOW) rc = su2hip_muca_set_sums(H, 0, w->gsum, w->nsum);
  return rc;
}

/* before measure() or save_weight(): the device's weight state back into w */
int su2hip_bind_muca_download(weight *w) {
  su2hip_muca m;
  int rc = su2hip_muca_get(H, 0, &m, w->pos, w->W, w->hits);
  if (rc) return rc;
  if (w->mode == SLOW) rc = su2hip_muca_get_sums(H, 0, w->gsum, w->nsum);
  w->param_value[EVEN] = m.param_value[0];
  w->param_value[ODD] = m.param_value[1];
  w->last_max = m.last_max;
  w->muca_count = m.muca_count;
  linearize_weight(w);
  return rc;
}

/* after each su2hip_bind_update: the device's acceptance counts since the
 * previous call, added into the reference's counters (which main() resets
 * after thermalization, src/main.c:273), for print_acceptance */
int su2hip_bind_counters(counters *c) {
  static su2hip_counters last;
  su2hip_counters now;
  int rc = su2hip_get_counters(H, 0, &now);
  if (rc) return rc;
#define ADD(dst, member) ((dst) += now.member - last.member)
  ADD(c->total_su2link, total_su2link);
  ADD(c->accepted_su2link, accepted_su2link);
  ADD(c->total_u1link, total_u1link);
  ADD(c->accepted_u1link, accepted_u1link);
#if (NHIGGS > 0)
  ADD(c->total_doublet[0], total_doublet);
  ADD(c->accepted_doublet[0], accepted_doublet);
  ADD(c->total_overrelax_doublet[0], total_overrelax_doublet);
  ADD(c->acc_overrelax_doublet[0], acc_overrelax_doublet);
#endif
#if (NHIGGS == 2)
  ADD(c->total_doublet[1], total_doublet2);
  ADD(c->accepted_doublet[1], accepted_doublet2);
  ADD(c->total_overrelax_doublet[1], total_overrelax_doublet2);
  ADD(c->acc_overrelax_doublet[1], acc_overrelax_doublet2);
#endif
#ifdef SINGLET
  ADD(c->total_singlet, total_singlet);
  ADD(c->accepted_singlet, accepted_singlet);
  ADD(c->total_overrelax_singlet, total_overrelax_singlet);
  ADD(c->acc_overrelax_singlet, acc_overrelax_singlet);
#endif
  ADD(c->total_triplet, total_triplet);
  ADD(c->accepted_triplet, accepted_triplet);
  ADD(c->total_overrelax_triplet, total_overrelax_triplet);
  ADD(c->acc_overrelax_triplet, acc_overrelax_triplet);
  ADD(c->total_muca, total_muca);
  ADD(c->accepted_muca, accepted_muca);
#undef ADD
  last = now;
  return SU2HIP_OK;
}
