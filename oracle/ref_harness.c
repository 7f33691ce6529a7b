/* ref_harness.c -- drives the reference su2higgs C code compiled from its own
 * sources (oracle/Makefile builds it into oracle/_ref/), for checking the
 * oracle restatement.  TEST INFRASTRUCTURE: loaded only by tests/ and by
 * bench.py's cpu_baseline leg.
 *
 * The reference's RNG entry points (src/generic/mersenne.h:9-11) are
 * defined here instead of mersenne.o: dran() returns the oracle's Philox
 * uniform for the current (pass, site, k).  The per-site update functions
 * and multicanonical_acceptance are linked with --wrap, so the wrappers see
 * every call the reference's own update_lattice (src/update.c:270-430)
 * makes: a call at the first site of a parity opens a new pass, and each
 * call sets the site context before running the reference function.
 * update_lattice itself is wrapped too: draws it takes outside a per-site
 * function (the random pass order, src/update.c:292-300, :341-345) come
 * from (pass at the start of the sweep, ORC_SCHED_SITE, k).  iran() is the
 * 64-bit Philox word and dran() its top 53 bits, as genrand64_real2 is
 * built from genrand64_int64 (src/generic/mersenne.c:141-150).
 */
#include "su2.h"

#include <stdint.h>
#include <string.h>

#include "su2oracle.h"

MPI_Comm MPI_COMM_WORLD;

static lattice l;
static params p;
static fields f;
static counters c;
static weight w;

static uint64_t h_seed, h_next_pass, h_cur_pass, h_sched_pass;
static uint32_t h_site, h_k, h_check, h_sched_k;
static int h_in_site;

unsigned long long genrand64_int64(void) {
  if (h_in_site) return orc_word(h_seed, h_cur_pass, h_site, h_k++);
  return orc_word(h_seed, h_sched_pass, ORC_SCHED_SITE, h_sched_k++);
}
double genrand64_real2(void) {
  return (double)(genrand64_int64() >> 11) * (1.0 / 9007199254740992.0);
}
void init_genrand64(unsigned long long s) { (void)s; }

static void site_ctx(long i) {
  if (i == 0 || i == l.evensites) {
    h_cur_pass = h_next_pass++;
    h_check = 0;
  }
  h_site = (uint32_t)i;
  h_k = 0;
}

#define WRAP(ret, name, decl, args) \
  ret __real_##name decl;          \
  ret __wrap_##name decl {         \
    site_ctx(i);                   \
    h_in_site = 1;                 \
    ret r_ = __real_##name args;   \
    h_in_site = 0;                 \
    return r_;                     \
  }

void __real_update_lattice(lattice *ll, fields *ff, params const *pp, counters *cc, weight *ww);
void __wrap_update_lattice(lattice *ll, fields *ff, params const *pp, counters *cc, weight *ww) {
  h_sched_pass = h_next_pass;
  h_sched_k = 0;
  __real_update_lattice(ll, ff, pp, cc, ww);
}

WRAP(int, heatbath_su2link, (lattice const *ll, fields *ff, params const *pp, long i, int dir), (ll, ff, pp, i, dir))
WRAP(int, metro_su2link, (lattice const *ll, fields *ff, params const *pp, long i, int dir), (ll, ff, pp, i, dir))
#ifdef U1
WRAP(int, metro_u1link, (lattice const *ll, fields *ff, params const *pp, long i, int dir), (ll, ff, pp, i, dir))
#endif
#if (NHIGGS > 0)
WRAP(int, metro_doublet, (lattice const *ll, fields *ff, params const *pp, long i, int h), (ll, ff, pp, i, h))
WRAP(int, overrelax_doublet, (lattice const *ll, fields *ff, params const *pp, long i), (ll, ff, pp, i))
#endif
#if (NHIGGS == 2)
WRAP(int, overrelax_higgs2, (lattice const *ll, fields *ff, params const *pp, long i, int h), (ll, ff, pp, i, h))
#endif
#ifdef TRIPLET
WRAP(int, metro_triplet, (lattice const *ll, fields *ff, params const *pp, long i), (ll, ff, pp, i))
WRAP(int, overrelax_triplet, (lattice const *ll, fields *ff, params const *pp, long i), (ll, ff, pp, i))
#endif
#ifdef SINGLET
WRAP(int, metro_singlet, (lattice const *ll, fields *ff, params const *pp, long i), (ll, ff, pp, i))
WRAP(int, overrelax_singlet, (lattice const *ll, fields *ff, params const *pp, long i), (ll, ff, pp, i))
#endif

int __real_multicanonical_acceptance(lattice const *ll, weight *ww, double o, double n);
int __wrap_multicanonical_acceptance(lattice const *ll, weight *ww, double o, double n) {
  h_site = ORC_MUCA_SITE;
  h_k = h_check++;
  h_in_site = 1;
  int r = __real_multicanonical_acceptance(ll, ww, o, n);
  h_in_site = 0;
  return r;
}

/* which model flags this build has: bit0 NHIGGS, bit1 U1, bit2 TRIPLET,
 * bit3 SINGLET, bit4 GRADFLOW, bit5 CORRELATORS, bit6 MEASURE_Z,
 * bit7 BLOCKING, bit8 NHIGGS == 2, bit9 HB_TRAJECTORY */
int ref_model(void) {
  int m = 0;
#if (NHIGGS > 0)
  m |= 1;
#endif
#ifdef U1
  m |= 2;
#endif
#ifdef TRIPLET
  m |= 4;
#endif
#ifdef SINGLET
  m |= 8;
#endif
#ifdef GRADFLOW
  m |= 16;
#endif
#ifdef CORRELATORS
  m |= 32;
#endif
#ifdef MEASURE_Z
  m |= 64;
#endif
#ifdef BLOCKING
  m |= 128;
#endif
#if (NHIGGS == 2)
  m |= 256;
#endif
#ifdef HB_TRAJECTORY
  m |= 512;
#endif
  return m;
}

int ref_setup(const su2hip_params *P) {
  memset(&l, 0, sizeof(l));
  memset(&p, 0, sizeof(p));
  memset(&w, 0, sizeof(w));
  myRank = 0;
  MPISize = 1;
  l.rank = 0;
  l.size = 1;
  l.dim = 3;
  l.L = malloc(3 * sizeof(int));
  l.vol = 1;
  for (int d = 0; d < 3; d++) {
    l.L[d] = P->L[d];
    l.vol *= P->L[d];
  }
  layout(&l, 0, 0);
  alloc_fields(&l, &f);
  p.betasu2 = P->betasu2;
  p.algorithm_su2link = P->algorithm_su2link;
  p.algorithm_su2doublet = P->algorithm_su2doublet;
  p.update_links = P->update_links;
  p.scalar_sweeps = P->scalar_sweeps;
  p.update_su2doublet = P->update_su2doublet;
  p.random_sweeps = 0;
  p.multicanonical = 0;
  p.msq_phi = P->msq_phi;
  p.lambda_phi = P->lambda_phi;
  p.msq_triplet = P->msq_triplet;
  p.a2 = P->a2;
  p.b4 = P->b4;
  p.update_su2triplet = P->update_su2triplet;
  p.algorithm_su2triplet = P->algorithm_su2triplet;
#ifdef U1
  p.betau1 = P->betau1;
  p.r_u1 = P->r_u1;
  p.algorithm_u1link = P->algorithm_u1link;
#endif
#ifdef SINGLET
  p.update_singlet = P->update_singlet;
  p.algorithm_singlet = P->algorithm_singlet;
  p.b1_s = P->b1_s;
  p.msq_s = P->msq_s;
  p.b3_s = P->b3_s;
  p.b4_s = P->b4_s;
#if (NHIGGS == 1)
  p.a1_s = P->a1_s;
  p.a2_s = P->a2_s;
#endif
#endif
#if (NHIGGS == 2)
  p.msq_phi2 = P->msq_phi2;
  p.m12sq.re = P->m12sq_re;
  p.m12sq.im = P->m12sq_im;
  p.lam2 = P->lam2;
  p.lam3 = P->lam3;
  p.lam4 = P->lam4;
  p.lam5.re = P->lam5_re;
  p.lam5.im = P->lam5_im;
  p.lam6.re = P->lam6_re;
  p.lam6.im = P->lam6_im;
  p.lam7.re = P->lam7_re;
  p.lam7.im = P->lam7_im;
#endif
  w.do_acceptance = 0;
  w.orderparam = -1;
  init_counters(&c);
  h_seed = P->seed;
  h_next_pass = 0;
  return 0;
}

long ref_evensites(void) { return l.evensites; }

void ref_tables(long *next, long *prev) {
  for (long i = 0; i < l.sites; i++)
    for (int d = 0; d < 3; d++) {
      next[i * 3 + d] = l.next[i][d];
      prev[i * 3 + d] = l.prev[i][d];
    }
}

void ref_put_fields(const double *link, const double *u1, const double *db,
                    const double *tr, const double *sg) {
  long V = l.sites;
  if (link) memcpy(&f.su2link[0][0][0], link, sizeof(double) * V * 12);
#ifdef U1
  if (u1) memcpy(&f.u1link[0][0], u1, sizeof(double) * V * 3);
#endif
#if (NHIGGS > 0)
  for (int h = 0; h < NHIGGS; h++)
    if (db) memcpy(&f.su2doublet[h][0][0], db + h * V * 4, sizeof(double) * V * 4);
#endif
#ifdef TRIPLET
  if (tr) memcpy(&f.su2triplet[0][0], tr, sizeof(double) * V * 3);
#endif
#ifdef SINGLET
  if (sg) memcpy(&f.singlet[0][0], sg, sizeof(double) * V);
#endif
  (void)u1;
  (void)db;
  (void)tr;
  (void)sg;
}

void ref_get_fields(double *link, double *u1, double *db, double *tr, double *sg) {
  long V = l.sites;
  if (link) memcpy(link, &f.su2link[0][0][0], sizeof(double) * V * 12);
#ifdef U1
  if (u1) memcpy(u1, &f.u1link[0][0], sizeof(double) * V * 3);
#endif
#if (NHIGGS > 0)
  for (int h = 0; h < NHIGGS; h++)
    if (db) memcpy(db + h * V * 4, &f.su2doublet[h][0][0], sizeof(double) * V * 4);
#endif
#ifdef TRIPLET
  if (tr) memcpy(tr, &f.su2triplet[0][0], sizeof(double) * V * 3);
#endif
#ifdef SINGLET
  if (sg) memcpy(sg, &f.singlet[0][0], sizeof(double) * V);
#endif
  (void)u1;
  (void)db;
  (void)tr;
  (void)sg;
}

void ref_rng_set(uint64_t seed, uint64_t pass) {
  h_seed = seed;
  h_next_pass = pass;
}
uint64_t ref_rng_pass(void) { return h_next_pass; }

void ref_counters(su2hip_counters *o) {
  memset(o, 0, sizeof(*o));
  o->total_su2link = c.total_su2link;
  o->accepted_su2link = c.accepted_su2link;
  o->total_u1link = c.total_u1link;
  o->accepted_u1link = c.accepted_u1link;
#if (NHIGGS > 0)
  o->total_doublet = c.total_doublet[0];
  o->accepted_doublet = c.accepted_doublet[0];
  o->total_overrelax_doublet = c.total_overrelax_doublet[0];
  o->acc_overrelax_doublet = c.acc_overrelax_doublet[0];
#endif
#if (NHIGGS == 2)
  o->total_doublet2 = c.total_doublet[1];
  o->accepted_doublet2 = c.accepted_doublet[1];
  o->total_overrelax_doublet2 = c.total_overrelax_doublet[1];
  o->acc_overrelax_doublet2 = c.acc_overrelax_doublet[1];
#endif
  o->total_triplet = c.total_triplet;
  o->accepted_triplet = c.accepted_triplet;
  o->total_overrelax_triplet = c.total_overrelax_triplet;
  o->acc_overrelax_triplet = c.acc_overrelax_triplet;
#ifdef SINGLET
  o->total_singlet = c.total_singlet;
  o->accepted_singlet = c.accepted_singlet;
  o->total_overrelax_singlet = c.total_overrelax_singlet;
  o->acc_overrelax_singlet = c.acc_overrelax_singlet;
#endif
  o->total_muca = c.total_muca;
  o->accepted_muca = c.accepted_muca;
}

/* the reference's own update_lattice, n times */
void ref_update(int n) {
  for (int k = 0; k < n; k++) update_lattice(&l, &f, &p, &c, &w);
}

/* the random pass order of update_lattice (p->random_sweeps) */
void ref_set_random_sweeps(int on) { p.random_sweeps = on; }

/* one per-site update with an explicit pass number; for the doublet kinds
 * dir is the doublet (higgs_id) */
int ref_site(int kind, long i, int dir, uint64_t pass) {
  h_cur_pass = pass;
  h_site = (uint32_t)i;
  h_k = 0;
  h_in_site = 1;
  int r = -1;
  switch (kind) {
    case SU2HIP_K_LINK_HEATBATH: r = __real_heatbath_su2link(&l, &f, &p, i, dir); break;
    case SU2HIP_K_LINK_METROPOLIS: r = __real_metro_su2link(&l, &f, &p, i, dir); break;
#ifdef U1
    case SU2HIP_K_U1_METROPOLIS: r = __real_metro_u1link(&l, &f, &p, i, dir); break;
#endif
#if (NHIGGS == 2)
    case SU2HIP_K_DOUBLET_OVERRELAX: r = __real_overrelax_higgs2(&l, &f, &p, i, dir); break;
    case SU2HIP_K_DOUBLET_METROPOLIS: r = __real_metro_doublet(&l, &f, &p, i, dir); break;
#elif (NHIGGS > 0)
    case SU2HIP_K_DOUBLET_OVERRELAX: r = __real_overrelax_doublet(&l, &f, &p, i); break;
    case SU2HIP_K_DOUBLET_METROPOLIS: r = __real_metro_doublet(&l, &f, &p, i, 0); break;
#endif
#ifdef TRIPLET
    case SU2HIP_K_TRIPLET_OVERRELAX: r = __real_overrelax_triplet(&l, &f, &p, i); break;
    case SU2HIP_K_TRIPLET_METROPOLIS: r = __real_metro_triplet(&l, &f, &p, i); break;
#endif
#ifdef SINGLET
    case SU2HIP_K_SINGLET_OVERRELAX: r = __real_overrelax_singlet(&l, &f, &p, i); break;
    case SU2HIP_K_SINGLET_METROPOLIS: r = __real_metro_singlet(&l, &f, &p, i); break;
#endif
  }
  h_in_site = 0;
  return r;
}

void ref_link_staple(long i, int dir, double *V) { su2link_staple(&l, &f, &p, i, dir, V); }
void ref_wilson_staple(long i, int dir, double *V) { su2staple_wilson(&l, &f, i, dir, V); }
double ref_localact_su2link(long i, int dir) { return localact_su2link(&l, &f, &p, i, dir); }
double ref_action_local(long i) { return action_local(&l, &f, &p, i); }
double ref_polysolve3(double a, double b, double cc, double d) { return polysolve3(a, b, cc, d); }
#if (NHIGGS > 0)
void ref_doublet_staple(long i, double *s) { staple_doublet(s, &l, &f, &p, i, 0); }
double ref_localact_doublet(long i) { return localact_doublet(&l, &f, &p, i, 0); }
void ref_doublet_staple_h(long i, int h, double *s) { staple_doublet(s, &l, &f, &p, i, h); }
double ref_localact_doublet_h(long i, int h) { return localact_doublet(&l, &f, &p, i, h); }
#endif
#ifdef U1
double ref_localact_u1link(long i, int dir) { return localact_u1link(&l, &f, &p, i, dir); }
#endif
#ifdef TRIPLET
double ref_localact_triplet(long i) { return localact_triplet(&l, &f, &p, i); }
double ref_magcharge_cube(long i) { return magcharge_cube(&l, &f, &p, i); }
#endif
#ifdef SINGLET
double ref_localact_singlet(long i) { return localact_singlet(&l, &f, &p, i); }
#endif

/* measure()'s site loop through the reference's own per-site functions,
 * in the SU2HIP_SUM_* layout */
void ref_measure_sums(double *s) {
  memset(s, 0, sizeof(double) * SU2HIP_NSUMS);
  for (long i = 0; i < l.sites; i++) {
    s[SU2HIP_SUM_ACTION] += action_local(&l, &f, &p, i);
    s[SU2HIP_SUM_WILSON] += local_su2wilson(&l, &f, &p, i);
#ifdef U1
    s[SU2HIP_SUM_U1WILSON] += local_u1wilson(&l, &f, &p, i);
#endif
#if (NHIGGS > 0)
    double mod = doubletsq(f.su2doublet[0][i]);
    s[SU2HIP_SUM_COV_PHI] += covariant_doublet(&l, &f, i, 0);
    s[SU2HIP_SUM_PHI2] += mod;
    s[SU2HIP_SUM_PHI4] += mod * mod;
#ifdef TRIPLET
    s[SU2HIP_SUM_PHI2SIGMA2] += mod * tripletsq(f.su2triplet[i]);
#endif
#endif
#if (NHIGGS == 2)
    double mod2 = doubletsq(f.su2doublet[1][i]);
    s[SU2HIP_SUM_COV_PHI_2] += covariant_doublet(&l, &f, i, 1);
    s[SU2HIP_SUM_PHI2_2] += mod2;
    s[SU2HIP_SUM_PHI4_2] += mod2 * mod2;
    complex f12 = get_phi12(f.su2doublet[0][i], f.su2doublet[1][i]);
    s[SU2HIP_SUM_R12] += f12.re;
    s[SU2HIP_SUM_I12] += f12.im;
#endif
#ifdef TRIPLET
    double tm = tripletsq(f.su2triplet[i]);
    s[SU2HIP_SUM_SIGMA2] += tm;
    s[SU2HIP_SUM_SIGMA4] += tm * tm;
    for (int d = 0; d < 3; d++) s[SU2HIP_SUM_HOP_SIGMA] += hopping_triplet_forward(&l, &f, &p, i, d);
    double ch = magcharge_cube(&l, &f, &p, i);
    s[SU2HIP_SUM_MAGCHARGE] += ch;
    s[SU2HIP_SUM_MAGCHARGE_ABS] += fabs(ch);
#endif
#ifdef SINGLET
    double S = f.singlet[i][0];
    for (int d = 0; d < 3; d++) s[SU2HIP_SUM_S_COV] += S * S - S * f.singlet[l.next[i][d]][0];
    s[SU2HIP_SUM_S] += S;
    s[SU2HIP_SUM_S2] += S * S;
    s[SU2HIP_SUM_S3] += S * S * S;
    s[SU2HIP_SUM_S4] += S * S * S * S;
#if (NHIGGS == 1)
    s[SU2HIP_SUM_SPHISQ] += doubletsq(f.su2doublet[0][i]) * S;
    s[SU2HIP_SUM_S2PHISQ] += doubletsq(f.su2doublet[0][i]) * S * S;
#endif
#endif
  }
}

/* the reference's measure() line, as text */
int ref_measure_text(char *buf, int n) {
  char *mem = NULL;
  size_t sz = 0;
  FILE *fp = open_memstream(&mem, &sz);
  measure(fp, &l, &f, &p, &w);
  fclose(fp);
  int len = (int)sz;
  if (len >= n) len = n - 1;
  memcpy(buf, mem, len);
  buf[len] = 0;
  free(mem);
  return (int)sz;
}

/* labels file of this build into the current directory */
void ref_print_labels(void) { print_labels(); }

/* multicanonical: load a weight file the reference's way and compute the
 * order parameter of the current fields */
int ref_muca_load(const char *weightfile, const su2hip_muca *m) {
  p.multicanonical = 1;
  w.orderparam = m->orderparam;
  w.mode = m->mode;
  w.checks_per_sweep = m->checks_per_sweep;
  w.bins = m->bins;
  w.min = m->min;
  w.max = m->max;
  w.delta = m->delta;
  strncpy(w.weightfile, weightfile, sizeof(w.weightfile) - 1);
  load_weight(&w);
  alloc_muca_backups(&l, &w);
  calc_orderparam(&l, &f, &p, &w, EVEN);
  calc_orderparam(&l, &f, &p, &w, ODD);
  return 0;
}

void ref_muca_get(su2hip_muca *m, double *pos, double *W, int32_t *hits) {
  m->enabled = p.multicanonical;
  m->orderparam = w.orderparam;
  m->mode = w.mode;
  m->checks_per_sweep = w.checks_per_sweep;
  m->bins = w.bins;
  m->last_max = w.last_max;
  m->do_acceptance = w.do_acceptance;
  m->muca_count = w.muca_count;
  m->update_interval = w.update_interval;
  m->delta = w.delta;
  m->min = w.min;
  m->max = w.max;
  m->wrk_min = w.wrk_min;
  m->wrk_max = w.wrk_max;
  m->param_value[0] = w.param_value[0];
  m->param_value[1] = w.param_value[1];
  for (int i = 0; i < w.bins + 2; i++) {
    if (pos) pos[i] = w.pos[i];
    if (W) W[i] = w.W[i];
    if (hits) hits[i] = w.hits[i];
  }
}

void ref_muca_set_mode(int mode) { w.mode = mode; }
void ref_muca_set_interval(int n) { w.update_interval = n; }
/* gsum/nsum exist only in SLOW mode (src/multicanonical.c:134-138); zeros otherwise */
void ref_muca_sums(int64_t *gsum, int64_t *nsum) {
  for (int i = 0; i < w.bins + 2; i++) {
    gsum[i] = w.mode == SLOW && w.gsum ? w.gsum[i] : 0;
    nsum[i] = w.mode == SLOW && w.nsum ? w.nsum[i] : 0;
  }
}
double ref_muca_weight(double v) { return get_weight(&w, v); }
int ref_muca_whichbin(double v) { return whichbin(&w, v); }
double ref_muca_orderparam(int par) { return calc_orderparam(&l, &f, &p, &w, (char)par); }

#ifdef GRADFLOW
void ref_flow_force(double *force, double *tforce) {
  fields fo;
  alloc_fields(&l, &fo);
  calc_gradient(&l, &f, &p, &fo);
  memcpy(force, &fo.su2link[0][0][0], sizeof(double) * l.sites * 12);
#ifdef TRIPLET
  if (tforce) memcpy(tforce, &fo.su2triplet[0][0], sizeof(double) * l.sites * 3);
#endif
  (void)tforce;
  free_fields(&l, &fo);
}

void ref_flow_euler_step(double dt) {
  fields fo;
  alloc_fields(&l, &fo);
  calc_gradient(&l, &f, &p, &fo);
  flow_gauge(&l, &f, &fo, dt);
#ifdef TRIPLET
  flow_triplet(&l, &f, &fo, dt);
#endif
  free_fields(&l, &fo);
}

/* grad_flow() writing measure_flow into the current directory */
void ref_grad_flow(double t_max, double dt, int meas_interval, int id) {
  p.flow_meas_interval = meas_interval;
  grad_flow(&l, &f, &p, &w, t_max, dt, id);
}
#endif

/* multicanonical accept/reject on or off (w->do_acceptance) */
void ref_muca_set_acceptance(int on) { w.do_acceptance = on; }

#ifdef HB_TRAJECTORY
/* one set of heatbath trajectories (make_realtime_trajectories,
 * src/hb_trajectory.c:11-95), appending to ./trajectory */
void ref_trajectories(int n_traj, int interval, double min, double max, int id) {
  trajectory tr;
  memset(&tr, 0, sizeof(tr));
  tr.n_traj = n_traj;
  tr.interval = interval;
  tr.mode_interval = 1;
  tr.min = min;
  tr.max = max;
  make_realtime_trajectories(&l, &f, &p, &c, &w, &tr, id);
}
/* read_realtime_config (src/hb_trajectory.c:99-163) calls check_set,
 * which the reference's sources do not define: this library is loaded with
 * lazy binding and that function is never called. */
#endif

/* checkpoint I/O through the reference (src/checkpoint.c:84-214) */
void ref_save_lattice(const char *path, long iter) {
  c.iter = iter;
  Global_total_time = 0.0;
  Global_comms_time = 0.0;
  save_lattice(&l, f, c, (char *)path);
}
long ref_load_lattice(const char *path) {
  load_lattice(&l, &f, &c, (char *)path);
  return c.iter;
}

/* config parsing through the reference (src/parameters.c:157-402) */
int ref_get_parameters(const char *path, su2hip_params *out, long *longs,
                       su2hip_muca *m, char *latticefile, char *resultsfile) {
  lattice l2;
  params p2;
  weight w2;
  memset(&l2, 0, sizeof(l2));
  memset(&p2, 0, sizeof(p2));
  memset(&w2, 0, sizeof(w2));
  get_parameters((char *)path, &l2, &p2);
  get_weight_parameters((char *)path, &p2, &w2);
  memset(out, 0, sizeof(*out));
  for (int d = 0; d < 3 && d < l2.dim; d++) out->L[d] = l2.L[d];
  out->algorithm_su2link = p2.algorithm_su2link;
  out->algorithm_su2doublet = p2.algorithm_su2doublet;
  out->algorithm_su2triplet = p2.algorithm_su2triplet;
  out->update_links = p2.update_links;
  out->scalar_sweeps = p2.scalar_sweeps;
  out->update_su2doublet = p2.update_su2doublet;
  out->update_su2triplet = p2.update_su2triplet;
  out->betasu2 = p2.betasu2;
  out->msq_phi = p2.msq_phi;
  out->lambda_phi = p2.lambda_phi;
  out->msq_triplet = p2.msq_triplet;
  out->a2 = p2.a2;
  out->b4 = p2.b4;
#ifdef U1
  out->betau1 = p2.betau1;
  out->r_u1 = p2.r_u1;
  out->algorithm_u1link = p2.algorithm_u1link;
#endif
#ifdef SINGLET
  out->update_singlet = p2.update_singlet;
  out->algorithm_singlet = p2.algorithm_singlet;
  out->b1_s = p2.b1_s;
  out->msq_s = p2.msq_s;
  out->b3_s = p2.b3_s;
  out->b4_s = p2.b4_s;
#if (NHIGGS == 1)
  out->a1_s = p2.a1_s;
  out->a2_s = p2.a2_s;
#endif
#endif
#if (NHIGGS == 2)
  out->msq_phi2 = p2.msq_phi2;
  out->m12sq_re = p2.m12sq.re;
  out->m12sq_im = p2.m12sq.im;
  out->lam2 = p2.lam2;
  out->lam3 = p2.lam3;
  out->lam4 = p2.lam4;
  out->lam5_re = p2.lam5.re;
  out->lam5_im = p2.lam5.im;
  out->lam6_re = p2.lam6.re;
  out->lam6_im = p2.lam6.im;
  out->lam7_re = p2.lam7.re;
  out->lam7_im = p2.lam7.im;
#endif
  longs[0] = p2.iterations;
  longs[1] = p2.interval;
  longs[2] = p2.checkpoint;
  longs[3] = p2.n_thermalize;
  longs[4] = p2.reset;
  longs[5] = p2.multicanonical;
  longs[6] = p2.run_checks;
  longs[7] = p2.do_local_meas;
#ifdef GRADFLOW
  longs[8] = p2.do_flow;
  longs[9] = p2.flow_interval;
  longs[10] = p2.flow_meas_interval;
#endif
  m->enabled = p2.multicanonical;
  m->orderparam = w2.orderparam;
  m->mode = w2.mode;
  m->checks_per_sweep = w2.checks_per_sweep;
  m->bins = w2.bins;
  m->min = w2.min;
  m->max = w2.max;
  m->delta = w2.delta;
  strcpy(latticefile, p2.latticefile);
  (void)resultsfile;
  if (p2.resultsfile) fclose(p2.resultsfile);
  return 0;
}

double ref_phi0(const char *path) {
  (void)path;
  return p.phi0;
}

#ifdef CORRELATORS
/* defined in src/correlation.c without a prototype in su2.h */
complex photon_correl(lattice *l, fields const *f, params const *p, int d, int dir, int transmom,
                      int mode);
#if (NHIGGS > 0)
double higgs_correlator(lattice *l, fields const *f, int d, int dir, int higgs_id);
#endif
#ifdef TRIPLET
double triplet_correlator(lattice *l, fields const *f, int d, int dir);
#endif

/* measure_correlators' columns (src/correlation.c:339-391) for d = 0 ..
 * L[dir]/2-1, 9 per d: H(d), Sigma^2 correlator, photon_correl(mode 0,
 * p = 1, 2) and (mode 1, p = 1, 2) as re, im pairs; absent fields 0 */
int ref_correlators(int dir, double *out) {
  int maxd = l.L[dir] / 2;
  for (int d = 0; d < maxd; d++) {
    double *o = out + (size_t)d * 10;
    for (int k = 0; k < 10; k++) o[k] = 0.0;
#if (NHIGGS > 0)
    o[0] = higgs_correlator(&l, &f, d, dir, 0);
#endif
#ifdef TRIPLET
    o[1] = triplet_correlator(&l, &f, d, dir);
    complex c;
    c = photon_correl(&l, &f, &p, d, dir, 1, 0);
    o[2] = c.re;
    o[3] = c.im;
    c = photon_correl(&l, &f, &p, d, dir, 2, 0);
    o[4] = c.re;
    o[5] = c.im;
    c = photon_correl(&l, &f, &p, d, dir, 1, 1);
    o[6] = c.re;
    o[7] = c.im;
    c = photon_correl(&l, &f, &p, d, dir, 2, 1);
    o[8] = c.re;
    o[9] = c.im;
#endif
  }
  return maxd;
}

void ref_measure_correlators(const char *fname, int dir, int id) {
  measure_correlators((char *)fname, &l, &f, &p, dir, id);
}
void ref_print_labels_correlators(void) { print_labels_correlators(); }
#endif

int ref_longest_dir(void) { return l.longest_dir; }

#ifdef MEASURE_Z
/* print_z_labels (sets p.n_meas_z) and measure_along_z into ./measure_z
 * (src/z_coord.c:78-298); the z direction (init_z_coord, :18-24) */
void ref_print_z_labels(void) { print_z_labels(&l, &p); }
void ref_measure_along_z(long id) { measure_along_z(&l, &f, &p, id); }
int ref_z_dir(void) { return l.z_dir; }
/* prepare_wall (src/z_coord.c:304-376), its draws from the interposed
 * uniform stream */
void ref_prepare_wall(void) { prepare_wall(&l, &f, &p); }
#endif

#ifdef BLOCKING
/* The blocked lattices of main() (src/main.c:106-152): blocking in every
 * direction but the longest, up to `levels` (capped at max_block_level). */
static lattice *g_bl;
static fields *g_bf;
static int g_nbl, g_bdir[3];

int ref_blocked_setup(int levels) {
  for (int d = 0; d < l.dim; d++) g_bdir[d] = d != l.longest_dir;
  l.blocklist.sends = 0;
  l.blocklist.recvs = 0;
  realloc_comlist(&l.blocklist, RECV);
  l.standby = 0;
  l.blocking_level = 0;
  int max = max_block_level(&l, g_bdir);
  if (levels > max) levels = max;
  g_bl = calloc(levels > 0 ? levels : 1, sizeof(lattice));
  g_bf = calloc(levels > 0 ? levels : 1, sizeof(fields));
  for (int k = 0; k < levels; k++) {
    lattice *base = k ? &g_bl[k - 1] : &l;
    g_bl[k].blocklist.sends = 0;
    g_bl[k].blocklist.recvs = 0;
    block_lattice(base, &g_bl[k], g_bdir);
    alloc_fields(&g_bl[k], &g_bf[k]);
    g_bl[k].blocking_level = k + 1;
  }
  g_nbl = levels;
  return levels;
}

/* the numbers measure_blocked_correlators prints (src/correlation.c:393-421):
 * smear_fields + make_blocked_fields into each level, then the correlator
 * columns of ref_correlators on the blocked lattice, out[level][L/2][10] */
int ref_blocked_values(double *out) {
  int dir = l.longest_dir, maxd = l.L[dir] / 2;
  for (int k = 0; k < g_nbl; k++) {
    lattice *base = k ? &g_bl[k - 1] : &l;
    fields *fb = k ? &g_bf[k - 1] : &f;
    fields sm;
    alloc_fields(base, &sm);
    smear_fields(base, fb, &sm, g_bdir);
    make_blocked_fields(base, &g_bl[k], &sm, &g_bf[k]);
    free_fields(base, &sm);
    lattice *b = &g_bl[k];
    for (int d = 0; d < maxd; d++) {
      double *o = out + ((size_t)k * maxd + d) * 10;
      for (int q = 0; q < 10; q++) o[q] = 0.0;
      o[1] = triplet_correlator(b, &g_bf[k], d, dir);
      complex c;
      c = photon_correl(b, &g_bf[k], &p, d, dir, 1, 0);
      o[2] = c.re;
      o[3] = c.im;
      c = photon_correl(b, &g_bf[k], &p, d, dir, 2, 0);
      o[4] = c.re;
      o[5] = c.im;
      c = photon_correl(b, &g_bf[k], &p, d, dir, 1, 1);
      o[6] = c.re;
      o[7] = c.im;
      c = photon_correl(b, &g_bf[k], &p, d, dir, 2, 1);
      o[8] = c.re;
      o[9] = c.im;
    }
  }
  return g_nbl;
}

/* measure_blocked_correlators for every level (src/main.c:334-347): files
 * correl1, correl2, ... in the working directory */
void ref_blocked_correlators(int id) {
  for (int k = 0; k < g_nbl; k++)
    measure_blocked_correlators(k ? &g_bl[k - 1] : &l, &g_bl[k], k ? &g_bf[k - 1] : &f, &g_bf[k],
                                &p, g_bdir, l.longest_dir, id);
}
#endif
