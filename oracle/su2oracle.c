/* su2oracle.c -- CPU FP64 restatement of su2higgs's lattice update sweep.
 *
 * TEST INFRASTRUCTURE (see su2oracle.h): the checker and the CPU baseline,
 * never part of the product path.
 *
 * Each function names the reference function it restates (file:line under
 * the reference's src/).  The algebra is written with quaternion products
 * (U = u0 + i u.sigma, product rule of su2rot, src/su2u1.c:60-70) instead of
 * the reference's expanded polynomials; the results agree to rounding.
 * Where the reference computes in long double (polysolve3, local_su2wilson)
 * this file does too.
 *
 * Parity status: every per-site function (the two-doublet model's
 * included), the sweep, measure sums, the
 * multicanonical steps and the Euler flow are pinned against the reference
 * compiled from source (oracle/_ref, driven by oracle/ref_harness.c) and the
 * committed golden fixtures.  Functions with no reference counterpart are
 * pinned by invariants in tests/: link overrelaxation (action preserved,
 * involution, det 1), the RK3 flow (third-order convergence, agreement with
 * small-step Euler) and the hot start (moments of N(0,1)).
 */
#include "su2oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define D 3

struct orc {
  su2hip_params p;
  int L[D];
  long vol, even;
  long *next, *prev, *nat;
  double *U, *A, *phi, *sig, *S;
  double *bu_phi, *bu_sig;
  double *phi2, *bu_phi2; /* second doublet (nhiggs = 2) */
  uint64_t seed, pass;
  su2hip_counters c;
  su2hip_muca m;
  double *pos, *W, *slope, *b, *hgram;
  int32_t *hits;
  long *gsum, *nsum;
  int nb;
  int random_order; /* the reference's random_sweeps */
  uint64_t r_pass;
  uint32_t r_site, r_k;
};

/* ------------------------------------------------------------------ */
/* Philox4x32-10: Salmon et al., "Parallel random numbers: as easy as
 * 1, 2, 3", SC'11.  Round: (hi,lo)=M0*c0, (hi',lo')=M1*c2,
 * c' = (hi'^c1^k0, lo', hi^c3^k1, lo); key bumped by the Weyl constants
 * between rounds. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c0 = n0;
    c1 = (uint32_t)p1;
    c2 = n2;
    c3 = (uint32_t)p0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

uint64_t orc_word(uint64_t seed, uint64_t pass, uint32_t site, uint32_t k) {
  uint32_t ctr[4] = {k >> 1, site, (uint32_t)pass, (uint32_t)(pass >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  orc_philox4x32_10(ctr, key, o);
  return (k & 1) ? (((uint64_t)o[3] << 32) | o[2]) : (((uint64_t)o[1] << 32) | o[0]);
}

double orc_uniform(uint64_t seed, uint64_t pass, uint32_t site, uint32_t k) {
  /* dran = (iran >> 11) 2^-53, as genrand64_real2 from genrand64_int64
   * (src/generic/mersenne.c:141-150) */
  return (double)(orc_word(seed, pass, site, k) >> 11) * (1.0 / 9007199254740992.0);
}

static inline double dran(orc_t *o) {
  return orc_uniform(o->seed, o->r_pass, o->r_site, o->r_k++);
}

static inline void ctx(orc_t *o, uint64_t pass, uint32_t site) {
  o->r_pass = pass;
  o->r_site = site;
  o->r_k = 0;
}

/* ------------------------------------------------------------------ */
/* quaternion algebra, U = u0 + i u.sigma */
static inline void qmul(const double *a, const double *b, double *r) {
  double r0 = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  double r1 = a[0] * b[1] + a[1] * b[0] - (a[2] * b[3] - a[3] * b[2]);
  double r2 = a[0] * b[2] + a[2] * b[0] - (a[3] * b[1] - a[1] * b[3]);
  double r3 = a[0] * b[3] + a[3] * b[0] - (a[1] * b[2] - a[2] * b[1]);
  r[0] = r0;
  r[1] = r1;
  r[2] = r2;
  r[3] = r3;
}
static inline void qconj(const double *a, double *r) {
  r[0] = a[0];
  r[1] = -a[1];
  r[2] = -a[2];
  r[3] = -a[3];
}
/* r = a * b^dagger */
static inline void qmul_nc(const double *a, const double *b, double *r) {
  double bc[4];
  qconj(b, bc);
  qmul(a, bc, r);
}
/* r = a^dagger * b */
static inline void qmul_cn(const double *a, const double *b, double *r) {
  double ac[4];
  qconj(a, ac);
  qmul(ac, b, r);
}
static inline double qsq(const double *a) {
  return a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
}
/* hypercharge factor exp(-i alpha sigma_3) as a quaternion (src/staples.c:161-177) */
static inline void u1twist(double alpha, double *e) {
  e[0] = cos(alpha);
  e[1] = 0.0;
  e[2] = 0.0;
  e[3] = -sin(alpha);
}
/* R(U) v = vector part of U (0,v) U^dagger */
static inline void rotate(const double *u, const double *v, double *r) {
  double q[4] = {0.0, v[0], v[1], v[2]}, t[4];
  qmul(u, q, t);
  qmul_nc(t, u, q);
  r[0] = q[1];
  r[1] = q[2];
  r[2] = q[3];
}

#define LNK(o, i, d) ((o)->U + ((size_t)(i) * D + (d)) * 4)
#define PHI(o, i) ((o)->phi + (size_t)(i) * 4)
/* doublet h (the reference's higgs_id) at site i */
#define PHIH(o, h, i) (((h) ? (o)->phi2 : (o)->phi) + (size_t)(i) * 4)
#define SIG(o, i) ((o)->sig + (size_t)(i) * 3)
#define ALP(o, i, d) ((o)->A[(size_t)(i) * D + (d)])
#define NX(o, i, d) ((o)->next[(size_t)(i) * D + (d)])
#define PV(o, i, d) ((o)->prev[(size_t)(i) * D + (d)])

/* ------------------------------------------------------------------ */
/* geometry: serial layout (src/layout.c:657-756, paritymap :826-851,
 * indexToCoords/coordsToIndex :945-972) */
static void build_layout(orc_t *o) {
  long V = o->vol;
  int *L = o->L;
  long *newidx = (long *)malloc(sizeof(long) * V);
  long ne = 0;
  for (long i = 0; i < V; i++) {
    long x = i % L[0], y = (i / L[0]) % L[1], z = i / ((long)L[0] * L[1]);
    if ((x + y + z) % 2 == 0) ne++;
  }
  o->even = ne;
  long e = 0, od = 0;
  for (long i = 0; i < V; i++) {
    long x = i % L[0], y = (i / L[0]) % L[1], z = i / ((long)L[0] * L[1]);
    if ((x + y + z) % 2 == 0)
      newidx[i] = e++;
    else
      newidx[i] = ne + od++;
  }
  for (long i = 0; i < V; i++) {
    long c[D] = {i % L[0], (i / L[0]) % L[1], i / ((long)L[0] * L[1])};
    long ni = newidx[i];
    o->nat[ni] = i;
    for (int d = 0; d < D; d++) {
      long cp[D], cm[D];
      for (int k = 0; k < D; k++) cp[k] = cm[k] = c[k];
      cp[d] = (c[d] + 1) % L[d];
      cm[d] = (c[d] - 1 + L[d]) % L[d];
      long ip = cp[0] + L[0] * (cp[1] + (long)L[1] * cp[2]);
      long im = cm[0] + L[0] * (cm[1] + (long)L[1] * cm[2]);
      o->next[ni * D + d] = newidx[ip];
      o->prev[ni * D + d] = newidx[im];
    }
  }
  free(newidx);
}

orc_t *orc_create(const su2hip_params *p) {
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->p = *p;
  for (int d = 0; d < D; d++) o->L[d] = p->L[d];
  o->vol = (long)p->L[0] * p->L[1] * p->L[2];
  long V = o->vol;
  o->next = (long *)malloc(sizeof(long) * V * D);
  o->prev = (long *)malloc(sizeof(long) * V * D);
  o->nat = (long *)malloc(sizeof(long) * V);
  build_layout(o);
  o->U = (double *)calloc((size_t)V * D * 4, sizeof(double));
  o->A = (double *)calloc((size_t)V * D, sizeof(double));
  o->phi = (double *)calloc((size_t)V * 4, sizeof(double));
  o->sig = (double *)calloc((size_t)V * 3, sizeof(double));
  o->S = (double *)calloc((size_t)V, sizeof(double));
  o->bu_phi = (double *)calloc((size_t)V * 4, sizeof(double));
  o->bu_sig = (double *)calloc((size_t)V * 3, sizeof(double));
  if (p->nhiggs == 2) {
    o->phi2 = (double *)calloc((size_t)V * 4, sizeof(double));
    o->bu_phi2 = (double *)calloc((size_t)V * 4, sizeof(double));
  }
  o->seed = p->seed;
  o->pass = 0;
  return o;
}

static void free_muca(orc_t *o) {
  free(o->pos);
  free(o->W);
  free(o->slope);
  free(o->b);
  free(o->hgram);
  free(o->hits);
  free(o->gsum);
  free(o->nsum);
  o->pos = o->W = o->slope = o->b = o->hgram = NULL;
  o->hits = NULL;
  o->gsum = o->nsum = NULL;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  free(o->next);
  free(o->prev);
  free(o->nat);
  free(o->U);
  free(o->A);
  free(o->phi);
  free(o->sig);
  free(o->S);
  free(o->bu_phi);
  free(o->bu_sig);
  free(o->phi2);
  free(o->bu_phi2);
  free_muca(o);
  free(o);
}

long orc_volume(const orc_t *o) { return o->vol; }
long orc_evensites(const orc_t *o) { return o->even; }

void orc_tables(const orc_t *o, long *next, long *prev, long *natural) {
  if (next) memcpy(next, o->next, sizeof(long) * o->vol * D);
  if (prev) memcpy(prev, o->prev, sizeof(long) * o->vol * D);
  if (natural) memcpy(natural, o->nat, sizeof(long) * o->vol);
}

void orc_set_fields(orc_t *o, const double *su2link, const double *u1link,
                    const double *doublet, const double *triplet,
                    const double *singlet) {
  size_t V = (size_t)o->vol;
  if (su2link) memcpy(o->U, su2link, sizeof(double) * V * D * 4);
  if (u1link) memcpy(o->A, u1link, sizeof(double) * V * D);
  if (doublet) memcpy(o->phi, doublet, sizeof(double) * V * 4);
  if (doublet && o->phi2) memcpy(o->phi2, doublet + V * 4, sizeof(double) * V * 4);
  if (triplet) memcpy(o->sig, triplet, sizeof(double) * V * 3);
  if (singlet) memcpy(o->S, singlet, sizeof(double) * V);
}

void orc_get_fields(const orc_t *o, double *su2link, double *u1link,
                    double *doublet, double *triplet, double *singlet) {
  size_t V = (size_t)o->vol;
  if (su2link) memcpy(su2link, o->U, sizeof(double) * V * D * 4);
  if (u1link) memcpy(u1link, o->A, sizeof(double) * V * D);
  if (doublet) memcpy(doublet, o->phi, sizeof(double) * V * 4);
  if (doublet && o->phi2) memcpy(doublet + V * 4, o->phi2, sizeof(double) * V * 4);
  if (triplet) memcpy(triplet, o->sig, sizeof(double) * V * 3);
  if (singlet) memcpy(singlet, o->S, sizeof(double) * V);
}

/* setfields (src/init.c:92-112) */
void orc_cold_start(orc_t *o, double phi0, double sigma0, double singlet0) {
  for (long i = 0; i < o->vol; i++) {
    for (int d = 0; d < D; d++) {
      double *u = LNK(o, i, d);
      u[0] = 1.0;
      u[1] = u[2] = u[3] = 0.0;
      ALP(o, i, d) = 0.0;
    }
    for (int h = 0; h < (o->p.nhiggs == 2 ? 2 : 1); h++) {
      double *f = PHIH(o, h, i);
      f[0] = phi0;
      f[1] = f[2] = f[3] = 0.0;
    }
    double *s = SIG(o, i);
    s[0] = sigma0;
    s[1] = s[2] = 0.0;
    o->S[i] = singlet0;
  }
}

/* Box-Muller pair from two uniforms */
static inline void gauss2(double u1, double u2, double *z0, double *z1) {
  double r = sqrt(-2.0 * log(1.0 - u1));
  double t = 2.0 * M_PI * u2;
  *z0 = r * cos(t);
  *z1 = r * sin(t);
}

/* Hot start (no reference counterpart).  Draw order at each site: links
 * dir 0..2 (4 Gaussians each), U(1) angles, doublets (4 Gaussians each), triplet
 * (2 pairs, last Gaussian unused), singlet (1 pair, second unused). */
void orc_hot_start(orc_t *o) {
  for (long i = 0; i < o->vol; i++) {
    ctx(o, ORC_HOTSTART_PASS, (uint32_t)i);
    for (int d = 0; d < D; d++) {
      double z[4], u0, u1;
      u0 = dran(o);
      u1 = dran(o);
      gauss2(u0, u1, &z[0], &z[1]);
      u0 = dran(o);
      u1 = dran(o);
      gauss2(u0, u1, &z[2], &z[3]);
      double n = sqrt(qsq(z));
      double *u = LNK(o, i, d);
      for (int k = 0; k < 4; k++) u[k] = z[k] / n;
    }
    if (o->p.u1)
      for (int d = 0; d < D; d++) ALP(o, i, d) = M_PI * (2.0 * dran(o) - 1.0);
    for (int h = 0; h < o->p.nhiggs; h++) {
      double z[4], u0, u1;
      u0 = dran(o);
      u1 = dran(o);
      gauss2(u0, u1, &z[0], &z[1]);
      u0 = dran(o);
      u1 = dran(o);
      gauss2(u0, u1, &z[2], &z[3]);
      for (int k = 0; k < 4; k++) PHIH(o, h, i)[k] = z[k];
    }
    if (o->p.triplet) {
      double z[4], u0, u1;
      u0 = dran(o);
      u1 = dran(o);
      gauss2(u0, u1, &z[0], &z[1]);
      u0 = dran(o);
      u1 = dran(o);
      gauss2(u0, u1, &z[2], &z[3]);
      for (int k = 0; k < 3; k++) SIG(o, i)[k] = z[k];
    }
    if (o->p.singlet) {
      double z0, z1, u0, u1;
      u0 = dran(o);
      u1 = dran(o);
      gauss2(u0, u1, &z0, &z1);
      o->S[i] = z0;
    }
  }
}

void orc_rng_set(orc_t *o, uint64_t seed, uint64_t pass) {
  o->seed = seed;
  o->pass = pass;
}
void orc_rng_get(const orc_t *o, uint64_t *seed, uint64_t *pass) {
  *seed = o->seed;
  *pass = o->pass;
}
void orc_get_counters(const orc_t *o, su2hip_counters *c) { *c = o->c; }
void orc_reset_counters(orc_t *o) { memset(&o->c, 0, sizeof(o->c)); }

/* ------------------------------------------------------------------ */
/* local action pieces (src/su2u1.c) */

/* su2ptrace (src/su2u1.c:108-116): Re Tr U_a(x) U_b(x+a) U_a(x+b)^+ U_b(x)^+ */
static double ptrace(const orc_t *o, long i, int a, int b) {
  double t[4], s[4];
  qmul(LNK(o, i, a), LNK(o, NX(o, i, a), b), t);
  qmul_nc(t, LNK(o, NX(o, i, b), a), s);
  qmul_nc(s, LNK(o, i, b), t);
  return 2.0 * t[0];
}

/* local_su2wilson (src/su2u1.c:143-157), long double accumulation */
static long double local_su2wilson(const orc_t *o, long i) {
  long double res = 0.0;
  for (int d1 = 0; d1 < D; d1++)
    for (int d2 = 0; d2 < d1; d2++) res += (1.0 - 0.5 * ptrace(o, i, d2, d1));
  return o->p.betasu2 * res;
}

/* u1ptrace (src/su2u1.c:300-309) */
static double u1ptrace(const orc_t *o, long i, int a, int b) {
  return ALP(o, i, a) + ALP(o, NX(o, i, a), b) - ALP(o, NX(o, i, b), a) -
         ALP(o, i, b);
}

/* local_u1wilson (src/su2u1.c:316-328) */
static double local_u1wilson(const orc_t *o, long i) {
  double res = 0.0;
  for (int d1 = 0; d1 < D; d1++)
    for (int d2 = 0; d2 < d1; d2++)
      res += (1.0 - cos(o->p.r_u1 * u1ptrace(o, i, d2, d1)));
  return o->p.betau1 * res;
}

static inline double doubletsq(const double *a) {
  return 0.5 * (a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3]);
}
static inline double tripletsq(const double *a) {
  return 0.5 * (a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}

/* hopping_doublet_forward (src/su2u1.c:428-445) of doublet h:
 * -Re Tr Phi(x)^+ U_j(x) Phi(x+j) exp(-i alpha_j(x) sigma_3) */
static double hop_doublet_fwd(const orc_t *o, long i, int d, int h) {
  double t[4], s[4];
  qmul_cn(PHIH(o, h, i), LNK(o, i, d), t);
  qmul(t, PHIH(o, h, NX(o, i, d)), s);
  if (o->p.u1) {
    double e[4];
    u1twist(ALP(o, i, d), e);
    qmul(s, e, t);
    return -t[0];
  }
  return -s[0];
}

/* hopping_doublet_backward (src/su2u1.c:451-472) */
static double hop_doublet_bwd(const orc_t *o, long i, int d, int h) {
  long pv = PV(o, i, d);
  double t[4], s[4];
  qmul_cn(PHIH(o, h, pv), LNK(o, pv, d), t);
  qmul(t, PHIH(o, h, i), s);
  if (o->p.u1) {
    double e[4];
    u1twist(ALP(o, pv, d), e);
    qmul(s, e, t);
    return -t[0];
  }
  return -s[0];
}

/* covariant_doublet (src/su2u1.c:477-488) */
static double covariant_doublet(const orc_t *o, long i, int h) {
  double tot = 0.0, mod = doubletsq(PHIH(o, h, i));
  for (int d = 0; d < D; d++) tot += 2.0 * mod + hop_doublet_fwd(o, i, d, h);
  return tot;
}

/* get_phi12 (src/su2u1.c:513-520): R + i I = phi1^+ phi2 */
static inline void phi12(const double *h1, const double *h2, double *R, double *I) {
  *R = 0.5 * (h1[0] * h2[0] + h1[1] * h2[1] + h1[2] * h2[2] + h1[3] * h2[3]);
  *I = 0.5 * (h1[3] * h2[0] + h1[2] * h2[1] - h1[1] * h2[2] - h1[0] * h2[3]);
}

/* hopping_triplet_forward (src/su2u1.c:605-617): -2 Tr A(x) U A(x+j) U^+ */
static double hop_triplet_fwd(const orc_t *o, long i, int d) {
  double r[3];
  rotate(LNK(o, i, d), SIG(o, NX(o, i, d)), r);
  const double *a = SIG(o, i);
  return -(a[0] * r[0] + a[1] * r[1] + a[2] * r[2]);
}

/* hopping_triplet_backward (src/su2u1.c:623-637) */
static double hop_triplet_bwd(const orc_t *o, long i, int d) {
  long pv = PV(o, i, d);
  double r[3];
  rotate(LNK(o, pv, d), SIG(o, i), r);
  const double *a = SIG(o, pv);
  return -(a[0] * r[0] + a[1] * r[1] + a[2] * r[2]);
}

/* covariant_triplet (src/su2u1.c:643-651) */
static double covariant_triplet(const orc_t *o, long i) {
  double tot = 0.0, mod = tripletsq(SIG(o, i));
  for (int d = 0; d < D; d++) tot += 2.0 * mod + hop_triplet_fwd(o, i, d);
  return tot;
}

/* potential_singlet (src/su2u1.c:708-723) */
static double potential_singlet(const orc_t *o, long i) {
  const su2hip_params *p = &o->p;
  double S = o->S[i];
  double pot = p->b1_s * S + 0.5 * p->msq_s * S * S +
               1.0 / 3.0 * p->b3_s * S * S * S + 0.25 * p->b4_s * S * S * S * S;
  if (p->nhiggs == 1) {
    double mod = doubletsq(PHI(o, i));
    pot += 0.5 * p->a1_s * S * mod + 0.5 * p->a2_s * S * S * mod;
  }
  return pot;
}

/* higgspotential (src/su2u1.c:528-572) */
static double higgspotential(const orc_t *o, long i) {
  const su2hip_params *p = &o->p;
  double pot = 0.0, mod = 0.0;
  if (p->nhiggs) {
    mod = doubletsq(PHI(o, i));
    pot += p->msq_phi * mod + p->lambda_phi * mod * mod;
  }
  if (p->nhiggs == 2) {
    /* two-doublet terms (src/su2u1.c:537-554) */
    double f11 = mod, f22 = doubletsq(PHIH(o, 1, i)), R, I;
    phi12(PHI(o, i), PHIH(o, 1, i), &R, &I);
    pot += p->msq_phi2 * f22 + p->m12sq_re * R - p->m12sq_im * I + p->lam2 * f22 * f22 +
           p->lam3 * f11 * f22 + p->lam4 * (R * R + I * I) + p->lam5_re * (R * R - I * I) -
           2.0 * p->lam5_im * R * I + f11 * (p->lam6_re * R - p->lam6_im * I) +
           f22 * (p->lam7_re * R + p->lam7_im * I);
  }
  if (p->triplet) {
    double mt = tripletsq(SIG(o, i));
    pot += p->msq_triplet * mt + p->b4 * mt * mt;
    if (p->nhiggs) pot += p->a2 * mod * mt;
  }
  if (p->singlet) pot += potential_singlet(o, i);
  return pot;
}

/* localact_su2link (src/su2u1.c:166-198) */
double orc_localact_su2link(const orc_t *o, long i, int dir) {
  double tot = 0.0;
  for (int d2 = 0; d2 < D; d2++) {
    if (d2 == dir) continue;
    tot += (1.0 - 0.5 * ptrace(o, i, dir, d2));
    tot += (1.0 - 0.5 * ptrace(o, PV(o, i, d2), dir, d2));
  }
  tot *= o->p.betasu2;
  for (int h = 0; h < o->p.nhiggs; h++) tot += hop_doublet_fwd(o, i, dir, h);
  if (o->p.triplet) tot += hop_triplet_fwd(o, i, dir);
  return tot;
}

/* localact_u1link (src/su2u1.c:336-354) */
double orc_localact_u1link(const orc_t *o, long i, int dir) {
  double tot = 0.0, r = o->p.r_u1;
  for (int d2 = 0; d2 < D; d2++) {
    if (d2 == dir) continue;
    tot += (1.0 - cos(r * u1ptrace(o, i, dir, d2)));
    tot += (1.0 - cos(r * u1ptrace(o, PV(o, i, d2), dir, d2)));
  }
  tot *= o->p.betau1;
  for (int h = 0; h < o->p.nhiggs; h++) tot += hop_doublet_fwd(o, i, dir, h);
  return tot;
}

/* localact_doublet (src/su2u1.c:495-510) of doublet h */
double orc_localact_doublet_h(const orc_t *o, long i, int h) {
  double tot = covariant_doublet(o, i, h);
  for (int d = 0; d < D; d++) tot += hop_doublet_bwd(o, i, d, h);
  tot += higgspotential(o, i);
  return tot;
}
double orc_localact_doublet(const orc_t *o, long i) { return orc_localact_doublet_h(o, i, 0); }

/* localact_triplet (src/su2u1.c:657-679) */
double orc_localact_triplet(const orc_t *o, long i) {
  double tot = covariant_triplet(o, i);
  for (int d = 0; d < D; d++) tot += hop_triplet_bwd(o, i, d);
  double mod = tripletsq(SIG(o, i));
  tot += o->p.msq_triplet * mod + o->p.b4 * mod * mod;
  if (o->p.nhiggs) tot += o->p.a2 * doubletsq(PHI(o, i)) * mod;
  return tot;
}

/* localact_singlet (src/su2u1.c:688-705) */
double orc_localact_singlet(const orc_t *o, long i) {
  double S = o->S[i];
  double res = D * S * S;
  for (int d = 0; d < D; d++)
    res -= S * (o->S[NX(o, i, d)] + o->S[PV(o, i, d)]);
  res += potential_singlet(o, i);
  return res;
}

/* action_local (src/measure.c:314-343) */
double orc_action_local(const orc_t *o, long i) {
  double tot = 0.0;
  tot += local_su2wilson(o, i);
  if (o->p.u1) tot += local_u1wilson(o, i);
  tot += higgspotential(o, i);
  for (int h = 0; h < o->p.nhiggs; h++) tot += covariant_doublet(o, i, h);
  if (o->p.triplet) tot += covariant_triplet(o, i);
  if (o->p.singlet) {
    double S = o->S[i];
    tot += D * S * S;
    for (int d = 0; d < D; d++) tot -= S * o->S[NX(o, i, d)];
  }
  return tot;
}

/* ------------------------------------------------------------------ */
/* staples (src/staples.c) */

/* su2staple_wilson (src/staples.c:84-115):
 * sum_{nu!=mu} U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
 *            + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) */
void orc_wilson_staple(const orc_t *o, long i, int mu, double V[4]) {
  double tot[4] = {0.0, 0.0, 0.0, 0.0}, t[4], s[4];
  long xp = NX(o, i, mu);
  for (int nu = 0; nu < D; nu++) {
    if (nu == mu) continue;
    qmul_nc(LNK(o, xp, nu), LNK(o, NX(o, i, nu), mu), t);
    qmul_nc(t, LNK(o, i, nu), s);
    for (int k = 0; k < 4; k++) tot[k] += s[k];
    long xm = PV(o, i, nu);
    double ac[4], bc[4];
    qconj(LNK(o, PV(o, xp, nu), nu), ac);
    qconj(LNK(o, xm, mu), bc);
    qmul(ac, bc, t);
    qmul(t, LNK(o, xm, nu), s);
    for (int k = 0; k < 4; k++) tot[k] += s[k];
  }
  for (int k = 0; k < 4; k++) V[k] = tot[k];
}

/* su2link_staple (src/staples.c:134-193): -beta/2 * Wilson staple
 * - 1/2 Phi(x+mu) exp(-i alpha sigma_3) Phi(x)^+ */
void orc_link_staple(const orc_t *o, long i, int dir, double V[4]) {
  orc_wilson_staple(o, i, dir, V);
  for (int k = 0; k < 4; k++) V[k] *= -0.5 * o->p.betasu2;
  for (int h = 0; h < o->p.nhiggs; h++) {
    double b[4], t[4];
    const double *nb = PHIH(o, h, NX(o, i, dir));
    if (o->p.u1) {
      double e[4];
      u1twist(ALP(o, i, dir), e);
      qmul(nb, e, b);
    } else {
      memcpy(b, nb, sizeof(b));
    }
    qmul_nc(b, PHIH(o, h, i), t);
    for (int k = 0; k < 4; k++) V[k] -= 0.5 * t[k];
  }
}

/* staple_doublet (src/staples.c:243-302) of doublet h: s_a with
 * S = phi_a s_a + ... */
void orc_doublet_staple_h(const orc_t *o, long i, int h, double s[4]) {
  for (int k = 0; k < 4; k++) s[k] = 0.0;
  for (int d = 0; d < D; d++) {
    double t[4], q[4];
    /* forward: -(U_d(x) Phi(x+d) E_d(x)) */
    qmul(LNK(o, i, d), PHIH(o, h, NX(o, i, d)), t);
    if (o->p.u1) {
      double e[4];
      u1twist(ALP(o, i, d), e);
      qmul(t, e, q);
    } else {
      memcpy(q, t, sizeof(q));
    }
    for (int k = 0; k < 4; k++) s[k] -= q[k];
    /* backward: -(U_d(x-d)^+ Phi(x-d) E_d(x-d)^+) */
    long pv = PV(o, i, d);
    qmul_cn(LNK(o, pv, d), PHIH(o, h, pv), t);
    if (o->p.u1) {
      double e[4], ec[4];
      u1twist(ALP(o, pv, d), e);
      qconj(e, ec);
      qmul(t, ec, q);
    } else {
      memcpy(q, t, sizeof(q));
    }
    for (int k = 0; k < 4; k++) s[k] -= q[k];
  }
}
void orc_doublet_staple(const orc_t *o, long i, double s[4]) { orc_doublet_staple_h(o, i, 0, s); }

/* hopping staple of the triplet (src/overrelax.c:143-180):
 * s = -sum_d [ R(U_d(x)) Sigma(x+d) + R(U_d(x-d))^T Sigma(x-d) ] */
void orc_triplet_staple(const orc_t *o, long i, double s[3]) {
  s[0] = s[1] = s[2] = 0.0;
  for (int d = 0; d < D; d++) {
    double r[3], uc[4];
    rotate(LNK(o, i, d), SIG(o, NX(o, i, d)), r);
    for (int k = 0; k < 3; k++) s[k] -= r[k];
    long pv = PV(o, i, d);
    qconj(LNK(o, pv, d), uc);
    rotate(uc, SIG(o, pv), r);
    for (int k = 0; k < 3; k++) s[k] -= r[k];
  }
}

/* ------------------------------------------------------------------ */
/* polysolve3 (src/overrelax.c:14-34): the real root of a x^3+b x^2+c x+d
 * by Cardano's formula when the discriminant is negative; otherwise the
 * reference returns b/a (with a warning).  long double like the reference,
 * with its double-precision sqrt and cbrt. */
static double polysolve3_ld(long double a, long double b, long double c,
                            long double d) {
  long double disc = 18.0 * a * b * c * d - 4.0 * b * b * b * d + b * b * c * c -
                     4.0 * a * c * c * c - 27.0 * a * a * d * d;
  long double d1 = 2.0 * b * b * b - 9.0 * a * b * c + 27.0 * a * a * d;
  double x;
  if (disc < 0.0) {
    long double A = 0.5 * (d1 + a * sqrt((double)(-27.0 * disc)));
    A = cbrt((double)A);
    x = -1.0 * (b + A + (b * b - 3.0 * a * c) / A) / (3.0 * a);
  } else {
    x = b / a;
  }
  return x;
}

double orc_polysolve3(double a, double b, double c, double d) {
  return polysolve3_ld(a, b, c, d);
}

/* ------------------------------------------------------------------ */
/* per-site updates */

/* heatbath_su2link (src/heatbath.c:15-124), Kennedy-Pendleton */
static int site_heatbath(orc_t *o, long i, int dir) {
  double old[4], oldact = 0.0;
  double *u = LNK(o, i, dir);
  if (o->p.triplet) {
    memcpy(old, u, sizeof(old));
    oldact += hop_triplet_fwd(o, i, dir);
  }
  double V[4];
  orc_link_staple(o, i, dir, V);
  double a = sqrt(qsq(V));
  for (int k = 0; k < 4; k++) V[k] *= -1.0 / a;
  a *= 2.0;
  double r1, r2, r3, d, a0;
  int loop = 1;
  do {
    r1 = -1.0 * log(1.0 - dran(o)) / a;
    r2 = -1.0 * log(1.0 - dran(o)) / a;
    r3 = cos(2 * M_PI * dran(o));
    r3 *= r3;
    d = r1 * r3 + r2;
    r3 = 1.0 - dran(o);
    loop++;
  } while (r3 * r3 > 1 - 0.5 * d && loop <= 200);
  a0 = (loop > 200) ? 1e-9 : 1.0 - d;
  double rad = 1.0 - a0 * a0;
  if (rad < 0.0) return 0;
  rad = sqrt(rad);
  d = 2.0;
  while (d > 1.0 || d == 0.0) {
    r1 = 2.0 * dran(o) - 1.0;
    r2 = 2.0 * dran(o) - 1.0;
    r3 = 2.0 * dran(o) - 1.0;
    d = r1 * r1 + r2 * r2 + r3 * r3;
  }
  d = sqrt(d);
  r1 *= rad / d;
  r2 *= rad / d;
  r3 *= rad / d;
  double A[4] = {a0, r1, r2, r3};
  qmul_nc(A, V, u);
  if (o->p.triplet) {
    double diff = oldact - hop_triplet_fwd(o, i, dir);
    if (dran(o) < exp(diff)) return 1;
    memcpy(u, old, sizeof(old));
    return 0;
  }
  return 1;
}

/* accept rule of every Metropolis update in src/metropolis.c */
static inline int metro_accept(orc_t *o, double diff) {
  if (diff < 0) return 1;
  if (diff > 0 && exp(-diff) > dran(o)) return 1;
  return 0;
}

/* metro_su2link (src/metropolis.c:15-41) + random_su2link (src/init.c:43-58) */
static int site_metro_link(orc_t *o, long i, int dir) {
  double *u = LNK(o, i, dir), old[4];
  memcpy(old, u, sizeof(old));
  double act_old = orc_localact_su2link(o, i, dir);
  double r[4];
  r[0] = 10.0 * (dran(o) - 0.5);
  r[1] = 1.0 * (dran(o) - 0.5);
  r[2] = 1.0 * (dran(o) - 0.5);
  r[3] = 1.0 * (dran(o) - 0.5);
  double n = sqrt(qsq(r));
  for (int k = 0; k < 4; k++) r[k] = r[k] / n;
  qmul(u, r, u);
  double diff = orc_localact_su2link(o, i, dir) - act_old;
  if (metro_accept(o, diff)) return 1;
  memcpy(u, old, sizeof(old));
  return 0;
}

/* SU(2) link overrelaxation (no reference counterpart):
 * U' = W^+ U^+ W^+ with W = staple/|staple| keeps Re Tr(U staple); with an
 * adjoint scalar the hopping term quadratic in U gets a Metropolis test
 * like the heatbath's (src/heatbath.c:107-120). */
static int site_link_or(orc_t *o, long i, int dir) {
  double *u = LNK(o, i, dir), old[4], oldact = 0.0;
  if (o->p.triplet) {
    memcpy(old, u, sizeof(old));
    oldact = hop_triplet_fwd(o, i, dir);
  }
  double V[4];
  orc_link_staple(o, i, dir, V);
  double a = sqrt(qsq(V));
  if (a == 0.0) return 1;
  double w[4], t[4];
  for (int k = 0; k < 4; k++) w[k] = V[k] / a;
  double wc[4], uc[4];
  qconj(w, wc);
  qconj(u, uc);
  qmul(wc, uc, t);
  qmul(t, wc, u);
  if (o->p.triplet) {
    double diff = oldact - hop_triplet_fwd(o, i, dir);
    if (dran(o) < exp(diff)) return 1;
    memcpy(u, old, sizeof(old));
    return 0;
  }
  return 1;
}

/* metro_u1link (src/metropolis.c:48-70) */
static int site_metro_u1(orc_t *o, long i, int dir) {
  double old = ALP(o, i, dir);
  double act_old = orc_localact_u1link(o, i, dir);
  ALP(o, i, dir) += 0.95 * (dran(o) - 0.5);
  double diff = orc_localact_u1link(o, i, dir) - act_old;
  if (metro_accept(o, diff)) return 1;
  ALP(o, i, dir) = old;
  return 0;
}

/* overrelax_doublet (src/overrelax.c:54-128), Cartesian (X,Y) overrelax */
static int site_or_doublet(orc_t *o, long i) {
  const su2hip_params *p = &o->p;
  double *f = PHI(o, i);
  double s[4];
  orc_doublet_staple(o, i, s);
  double F = 0.0;
  for (int k = 0; k < 4; k++) F += s[k] * s[k];
  F = sqrt(F);
  double X = 0.0;
  for (int k = 0; k < 4; k++) X += f[k] * s[k];
  X /= F;
  double Ysq = 0.0;
  for (int k = 0; k < 4; k++) {
    double y = f[k] - X * s[k] / F;
    Ysq += y * y;
  }
  double B = 0.5 * p->msq_phi + 1.0 * D;
  if (p->triplet) B += 0.5 * p->a2 * tripletsq(SIG(o, i));
  if (p->singlet) {
    double S = o->S[i];
    B += 0.25 * p->a1_s * S + 0.25 * p->a2_s * S * S;
  }
  double C = 0.25 * p->lambda_phi;
  long double al = C;
  long double be = C * X;
  long double cc = B + 2.0 * Ysq * C;
  long double ga = (cc + C * X * X);
  long double de = (F + cc * X + C * X * X * X);
  double nX = polysolve3_ld(al, be, ga, de);
  double dV = F + 2.0 * B * X + 4.0 * C * (X * X * X + Ysq * X);
  double dVn = F + 2.0 * B * nX + 4.0 * C * (nX * nX * nX + Ysq * nX);
  double ratio = fabs(dV / dVn);
  if (ratio >= dran(o)) {
    for (int k = 0; k < 4; k++) f[k] = -1.0 * f[k] + (nX + X) * s[k] / F;
    return 1;
  }
  return 0;
}

/* overrelax_higgs2 (src/overrelax.c:295-417): the (X,Y) overrelaxation of
 * doublet h of two, Y kept fixed, the local potential a quartic in X whose
 * coefficients come from the other doublet H and G = -+ i sigma_2 H */
static int site_or_higgs2(orc_t *o, long i, int h) {
  const su2hip_params *p = &o->p;
  double *f = PHIH(o, h, i);
  const double *H = PHIH(o, 1 - h, i);
  double s[4];
  orc_doublet_staple_h(o, i, h, s);
  double sg = h ? 1.0 : -1.0;
  double G[4] = {sg * H[3], sg * H[2], -sg * H[1], -sg * H[0]};
  double lam12, msq, l67re, l67im, o67re, o67im;
  if (h == 0) {
    lam12 = p->lambda_phi;
    msq = p->msq_phi;
    l67re = p->lam7_re;
    l67im = p->lam7_im;
    o67re = p->lam6_re;
    o67im = -p->lam6_im;
  } else {
    lam12 = p->lam2;
    msq = p->msq_phi2;
    l67re = p->lam6_re;
    l67im = -p->lam6_im;
    o67re = p->lam7_re;
    o67im = p->lam7_im;
  }
  double Hsq = doubletsq(H);
  for (int k = 0; k < 4; k++) {
    s[k] += 0.5 * (p->m12sq_re + Hsq * l67re) * H[k];
    s[k] += 0.5 * (-1.0 * p->m12sq_im + Hsq * l67im) * G[k];
    s[k] = -1.0 * s[k];
  }
  double F = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + s[3] * s[3]);
  double X = 0.0;
  for (int k = 0; k < 4; k++) {
    s[k] /= F;
    X += f[k] * s[k];
  }
  double Y[4], Ysq = 0.0;
  for (int k = 0; k < 4; k++) {
    Y[k] = f[k] - X * s[k];
    Ysq += Y[k] * Y[k];
  }
  double Hf = 0.0, Gf = 0.0, HY = 0.0, GY = 0.0;
  for (int k = 0; k < 4; k++) {
    Hf += H[k] * s[k];
    Gf += G[k] * s[k];
    HY += H[k] * Y[k];
    GY += G[k] * Y[k];
  }
  double b4 = 0.25 * lam12;
  double b3 = 0.25 * (Hf * o67re + Gf * o67im);
  double b2 = 0.5 * msq + 1.0 * D + 0.5 * Ysq * lam12 + 0.25 * (HY * o67re + GY * o67im);
  double b1 = -1.0 * F + 0.25 * Ysq * (Hf * o67re + Gf * o67im);
  b2 += 0.5 * Hsq * p->lam3 + 0.25 * p->lam4 * (Hf * Hf + Gf * Gf) +
        0.25 * p->lam5_re * (Hf * Hf - Gf * Gf) - 0.5 * p->lam5_im * Gf * Hf;
  b1 += 0.5 * Hf * (HY * p->lam4 + HY * p->lam5_re - GY * p->lam5_im) +
        0.5 * Gf * (GY * p->lam4 - GY * p->lam5_re - HY * p->lam5_im);
  long double aa = b4;
  long double bb = b3 + b4 * X;
  long double cc = b2 + b3 * X + b4 * X * X;
  long double dd = b1 + b2 * X + b3 * X * X + b4 * X * X * X;
  double nX = polysolve3_ld(aa, bb, cc, dd);
  double dV = 4.0 * b4 * X * X * X + 3.0 * b3 * X * X + 2.0 * b2 * X + b1;
  double dVn = 4.0 * b4 * nX * nX * nX + 3.0 * b3 * nX * nX + 2.0 * b2 * nX + b1;
  if (fabs(dV / dVn) >= dran(o)) {
    for (int k = 0; k < 4; k++) f[k] = f[k] + s[k] * (nX - X);
    return 1;
  }
  return 0;
}

/* metro_doublet (src/metropolis.c:77-106) of doublet h */
static int site_metro_doublet(orc_t *o, long i, int h) {
  double *f = PHIH(o, h, i), old[4];
  memcpy(old, f, sizeof(old));
  double act_old = orc_localact_doublet_h(o, i, h);
  for (int k = 0; k < 4; k++) f[k] += 1.0 * (dran(o) - 0.5);
  double diff = orc_localact_doublet_h(o, i, h) - act_old;
  if (metro_accept(o, diff)) return 1;
  memcpy(f, old, sizeof(old));
  return 0;
}

/* overrelax_triplet (src/overrelax.c:137-236).  The reference's portal term
 * sits under `#ifdef HIGGS` (src/overrelax.c:202), a flag no build defines,
 * so B never contains it; kept that way. */
static int site_or_triplet(orc_t *o, long i) {
  const su2hip_params *p = &o->p;
  double *f = SIG(o, i);
  double s[3];
  orc_triplet_staple(o, i, s);
  double F = 0.0;
  for (int k = 0; k < 3; k++) F += s[k] * s[k];
  F = sqrt(F);
  double X = 0.0;
  for (int k = 0; k < 3; k++) X += f[k] * s[k];
  X /= F;
  double Ysq = 0.0;
  for (int k = 0; k < 3; k++) {
    double y = f[k] - X * s[k] / F;
    Ysq += y * y;
  }
  double B = 0.5 * p->msq_triplet + 1.0 * D;
  double C = 0.25 * p->b4;
  long double al = C;
  long double be = C * X;
  long double cc = B + 2.0 * Ysq * C;
  long double ga = (cc + C * X * X);
  long double de = (F + cc * X + C * X * X * X);
  double nX = polysolve3_ld(al, be, ga, de);
  double dV = F + 2.0 * B * X + 4.0 * C * (X * X * X + Ysq * X);
  double dVn = F + 2.0 * B * nX + 4.0 * C * (nX * nX * nX + Ysq * nX);
  double ratio = fabs(dV / dVn);
  if (ratio >= dran(o)) {
    for (int k = 0; k < 3; k++) f[k] = -1.0 * f[k] + (nX + X) * s[k] / F;
    return 1;
  }
  return 0;
}

/* metro_triplet (src/metropolis.c:113-143) */
static int site_metro_triplet(orc_t *o, long i) {
  double *f = SIG(o, i), old[3];
  memcpy(old, f, sizeof(old));
  double act_old = orc_localact_triplet(o, i);
  for (int k = 0; k < 3; k++) f[k] += 1.0 * (dran(o) - 0.5);
  double diff = orc_localact_triplet(o, i) - act_old;
  if (metro_accept(o, diff)) return 1;
  memcpy(f, old, sizeof(old));
  return 0;
}

/* overrelax_singlet (src/overrelax.c:243-285) */
static int site_or_singlet(orc_t *o, long i) {
  const su2hip_params *p = &o->p;
  double S = o->S[i];
  double c1 = p->b1_s;
  for (int d = 0; d < D; d++) c1 -= (o->S[NX(o, i, d)] + o->S[PV(o, i, d)]);
  double c2 = D + 0.5 * p->msq_s;
  double c3 = 1.0 / 3.0 * p->b3_s;
  double c4 = 0.25 * p->b4_s;
  if (p->nhiggs == 1) {
    double ph = doubletsq(PHI(o, i));
    c1 += 0.5 * p->a1_s * ph;
    c2 += 0.5 * p->a2_s * ph;
  }
  double d0 = c1 + c2 * S + c3 * S * S + c4 * S * S * S;
  double d1 = c2 + c3 * S + c4 * S * S;
  double d2 = c3 + c4 * S;
  double d3 = c4;
  double Y = polysolve3_ld(d3, d2, d1, d0);
  double dV = c1 + 2.0 * c2 * S + 3.0 * c3 * S * S + 4.0 * c4 * S * S * S;
  double dVn = c1 + 2.0 * c2 * Y + 3.0 * c3 * Y * Y + 4.0 * c4 * Y * Y * Y;
  if (fabs(dV / dVn) >= dran(o)) {
    o->S[i] = Y;
    return 1;
  }
  return 0;
}

/* metro_singlet (src/metropolis.c:148-167) */
static int site_metro_singlet(orc_t *o, long i) {
  double old = o->S[i];
  double act_old = orc_localact_singlet(o, i);
  o->S[i] += 1.0 * (dran(o) - 0.5);
  double diff = orc_localact_singlet(o, i) - act_old;
  if (metro_accept(o, diff)) return 1;
  o->S[i] = old;
  return 0;
}

static int site_dispatch(orc_t *o, int kind, long i, int dir) {
  switch (kind) {
    case SU2HIP_K_LINK_HEATBATH: return site_heatbath(o, i, dir);
    case SU2HIP_K_LINK_METROPOLIS: return site_metro_link(o, i, dir);
    case SU2HIP_K_LINK_OVERRELAX: return site_link_or(o, i, dir);
    case SU2HIP_K_U1_METROPOLIS: return site_metro_u1(o, i, dir);
    case SU2HIP_K_DOUBLET_OVERRELAX:
      return o->p.nhiggs == 2 ? site_or_higgs2(o, i, dir != 0) : site_or_doublet(o, i);
    case SU2HIP_K_DOUBLET_METROPOLIS: return site_metro_doublet(o, i, o->p.nhiggs == 2 && dir);
    case SU2HIP_K_TRIPLET_OVERRELAX: return site_or_triplet(o, i);
    case SU2HIP_K_TRIPLET_METROPOLIS: return site_metro_triplet(o, i);
    case SU2HIP_K_SINGLET_OVERRELAX: return site_or_singlet(o, i);
    case SU2HIP_K_SINGLET_METROPOLIS: return site_metro_singlet(o, i);
  }
  return 0;
}

int orc_site_update(orc_t *o, int kind, long i, int dir, uint64_t pass) {
  ctx(o, pass, (uint32_t)i);
  return site_dispatch(o, kind, i, dir);
}

/* ------------------------------------------------------------------ */
/* multicanonical (src/multicanonical.c) */

/* linearize_weight (src/multicanonical.c:337-346) */
static void linearize(orc_t *o) {
  int n = o->m.bins;
  o->slope[0] = 0;
  o->slope[n + 1] = 0;
  o->b[0] = o->W[0];
  o->b[n + 1] = o->W[n + 1];
  for (int i = 1; i < n + 1; i++) {
    o->slope[i] = (o->W[i + 1] - o->W[i]) / (o->pos[i + 1] - o->pos[i]);
    o->b[i] = o->W[i] - o->slope[i] * o->pos[i];
  }
}

/* whichbin (src/multicanonical.c:671-704) */
int orc_muca_whichbin(const orc_t *o, double val) {
  if (val < o->m.min) return 0;
  if (val >= o->m.max) return o->m.bins + 1;
  int lo = 1, hi = o->m.bins;
  while (lo <= hi) {
    int bin = lo + (hi - lo) / 2;
    double cur = o->pos[bin], nxt = o->pos[bin + 1];
    if ((val < nxt && val > cur) || (fabs(val - cur) < 1e-12)) return bin;
    if (val > cur)
      lo = bin + 1;
    else
      hi = bin - 1;
  }
  return o->m.bins;
}

/* get_weight (src/multicanonical.c:350-379) */
double orc_muca_weight(const orc_t *o, double val) {
  int n = o->m.bins;
  if (val >= o->m.max) return o->b[n + 1];
  if (val < o->m.min) return o->b[0];
  int bin = orc_muca_whichbin(o, val);
  if (bin >= n + 1) return o->b[n + 1];
  return o->b[bin] + o->slope[bin] * val;
}

/* muca_accumulate_hits (src/multicanonical.c:451-498) */
static void accumulate_hits(orc_t *o, double val) {
  int n = o->m.bins;
  int bin = orc_muca_whichbin(o, val);
  if (bin == 0) {
    double dmin = o->pos[2] - o->pos[1];
    if (val < o->pos[1] - 0.5 * dmin) return;
    bin += 1;
  } else if (bin < n + 1) {
    if (val - o->pos[bin] > o->pos[bin + 1] - val) bin = bin + 1;
  } else {
    double dmax = o->pos[n + 1] - o->pos[n];
    if (val > o->pos[n + 1] + 0.5 * dmax) return;
  }
  if (o->m.mode == SU2HIP_FAST) {
    o->hits[bin] += 5;
    if (bin > 1) o->hits[bin - 1] += 3;
    if (bin > 2) o->hits[bin - 2] += 1;
    if (bin < n) o->hits[bin + 1] += 3;
    if (bin < n - 1) o->hits[bin + 2] += 1;
  } else if (o->m.mode == SU2HIP_SLOW) {
    o->hits[bin]++;
    o->hgram[bin] += exp(orc_muca_weight(o, val));
  }
}

/* update_weight (src/multicanonical.c:504-557), FAST recursion */
static int update_weight_fast(orc_t *o) {
  int n = o->m.bins, first = -1;
  double delta = 0.0;
  for (int i = 0; i < n + 2; i++) {
    if (o->pos[i] >= o->m.wrk_min && o->m.wrk_max >= o->pos[i]) {
      if (first < 0) first = i;
      delta = (o->hits[i] - o->hits[first]) * o->m.delta / ((double)n);
      o->W[i] += delta;
    } else if (o->m.wrk_max < o->pos[i]) {
      o->W[i] += delta;
    }
  }
  o->W[0] = o->W[1];
  o->W[n + 1] = o->W[n];
  linearize(o);
  int tunnel = 0;
  if (o->m.last_max && o->hits[1] > 0) {
    o->m.last_max = 0;
    tunnel = 1;
  } else if (!o->m.last_max && o->hits[n] > 0) {
    o->m.last_max = 1;
    tunnel = 1;
  }
  if (tunnel) o->m.delta /= 1.5;
  for (int i = 0; i < n + 2; i++) o->hits[i] = 0;
  return tunnel;
}

/* update_weight_slow (src/multicanonical.c:561-666), hep-lat/9804019 */
static void update_weight_slow(orc_t *o) {
  int n = o->m.bins;
  double *wold = (double *)malloc(sizeof(double) * (n + 2));
  memcpy(wold, o->W, sizeof(double) * (n + 2));
  double dmin = o->pos[2] - o->pos[1];
  o->hgram[1] /= dmin;
  for (int i = 2; i < n + 1; i++) o->hgram[i] /= (o->pos[i + 1] - o->pos[i - 1]);
  int first = -1, last = -1;
  const int hits_min = 8;
  for (int i = 1; i < n + 1; i++) {
    if (o->pos[i] >= o->m.wrk_min && o->m.wrk_max >= o->pos[i]) {
      if (first < 0) {
        first = i;
        continue;
      }
      o->nsum[i] += o->hits[i];
      int g;
      if (o->hits[i] >= hits_min && o->hits[i - 1] >= hits_min) {
        g = o->hits[i] + o->hits[i - 1];
      } else {
        o->W[i] = wold[i] - wold[i - 1] + o->W[i - 1];
        continue;
      }
      long gnew = o->gsum[i] + g;
      double ln = o->hgram[i - 1] / o->hgram[i];
      ln = -1.0 * g * log(ln);
      o->W[i] = o->W[i - 1] + ((wold[i] - wold[i - 1]) * o->gsum[i] + ln) / ((double)gnew);
      o->gsum[i] = gnew;
    } else if (o->m.wrk_max < o->pos[i]) {
      if (last < 0) last = i - 1;
      o->W[i] += o->W[last] - wold[last];
    }
  }
  for (int i = 0; i < n + 2; i++) {
    o->hits[i] = 0;
    o->hgram[i] = 0.0;
  }
  o->W[0] = o->W[1];
  o->W[n + 1] = o->W[n];
  memcpy(wold, o->W, sizeof(double) * (n + 2));
  double corr = 0.0;
  if (first == 1)
    dmin = o->pos[first + 1] - o->pos[first];
  else
    dmin = o->pos[first + 1] - o->pos[first - 1];
  for (int i = first + 1; i < n + 1; i++) {
    if (i <= last) {
      double d1 = o->pos[i + 1] - o->pos[i - 1];
      corr = 2.0 * log((dmin / d1) * o->nsum[i] / o->nsum[first]);
    }
    o->W[i] += corr;
  }
  o->W[0] = o->W[1];
  o->W[n + 1] = o->W[n];
  linearize(o);
  memcpy(o->W, wold, sizeof(double) * (n + 2));
  free(wold);
}

/* calc_orderparam (src/multicanonical.c:709-751) */
double orc_muca_orderparam(orc_t *o, int par) {
  long off = par == SU2HIP_EVEN ? 0 : o->even;
  long max = par == SU2HIP_EVEN ? o->even : o->vol;
  double tot = 0.0;
  switch (o->m.orderparam) {
    case SU2HIP_SIGMASQ:
      for (long i = off; i < max; i++) tot += tripletsq(SIG(o, i));
      break;
    case SU2HIP_PHI2MINUSSIGMA2:
      for (long i = off; i < max; i++)
        tot += doubletsq(PHI(o, i)) - tripletsq(SIG(o, i));
      break;
    case SU2HIP_PHISQ:
      for (long i = off; i < max; i++) tot += doubletsq(PHI(o, i));
      break;
    case SU2HIP_PHI2SQ:
      for (long i = off; i < max; i++) tot += doubletsq(PHIH(o, 1, i));
      break;
  }
  tot = tot / o->vol;
  o->m.param_value[par] = tot;
  return tot + o->m.param_value[1 - par];
}

/* multicanonical_acceptance (src/multicanonical.c:385-445); one draw */
static int muca_acceptance(orc_t *o, double oldval, double newval) {
  double diff = orc_muca_weight(o, newval) - orc_muca_weight(o, oldval);
  int accept = exp(-(diff)) > dran(o);
  if (o->m.mode != SU2HIP_READONLY) {
    if (!accept) newval = oldval;
    accumulate_hits(o, newval);
    o->m.muca_count++;
    if (o->m.muca_count % o->m.update_interval == 0) {
      if (o->m.mode == SU2HIP_FAST)
        update_weight_fast(o);
      else if (o->m.mode == SU2HIP_SLOW)
        update_weight_slow(o);
      o->m.muca_count = 0;
    }
  }
  return accept;
}

/* muca_check (src/update.c:13-30) */
static int muca_check(orc_t *o, int par, uint64_t pass, uint32_t check) {
  double old = o->m.param_value[0] + o->m.param_value[1];
  double nw = orc_muca_orderparam(o, par);
  ctx(o, pass, ORC_MUCA_SITE);
  o->r_k = check;
  int acc = muca_acceptance(o, old, nw);
  if (!acc) o->m.param_value[par] = old - o->m.param_value[1 - par];
  o->c.accepted_muca += acc;
  o->c.total_muca++;
  return acc;
}

void orc_muca_set(orc_t *o, const su2hip_muca *m, const double *pos,
                  const double *W) {
  free_muca(o);
  o->m = *m;
  int n = m->bins;
  o->nb = n;
  o->pos = (double *)malloc(sizeof(double) * (n + 2));
  o->W = (double *)malloc(sizeof(double) * (n + 2));
  o->slope = (double *)malloc(sizeof(double) * (n + 2));
  o->b = (double *)malloc(sizeof(double) * (n + 2));
  o->hgram = (double *)calloc(n + 2, sizeof(double));
  o->hits = (int32_t *)calloc(n + 2, sizeof(int32_t));
  o->gsum = (long *)malloc(sizeof(long) * (n + 2));
  o->nsum = (long *)malloc(sizeof(long) * (n + 2));
  for (int i = 0; i < n + 2; i++) {
    o->gsum[i] = 5;
    o->nsum[i] = 1;
  }
  memcpy(o->pos, pos, sizeof(double) * (n + 2));
  memcpy(o->W, W, sizeof(double) * (n + 2));
  if (o->m.update_interval <= 0)
    o->m.update_interval = (o->m.mode == SU2HIP_SLOW) ? 5000 : 100;
  if (o->m.mode == SU2HIP_SLOW)
    for (int i = 0; i < n + 2; i++) o->hgram[i] = exp(o->W[i]);
  linearize(o);
  if (o->m.enabled) {
    orc_muca_orderparam(o, SU2HIP_EVEN);
    orc_muca_orderparam(o, SU2HIP_ODD);
  }
}

void orc_muca_set_acceptance(orc_t *o, int on) { o->m.do_acceptance = on != 0; }

void orc_muca_get(const orc_t *o, su2hip_muca *m, double *pos, double *W,
                  int32_t *hits) {
  *m = o->m;
  int n = o->m.bins;
  if (pos && o->pos) memcpy(pos, o->pos, sizeof(double) * (n + 2));
  if (W && o->W) memcpy(W, o->W, sizeof(double) * (n + 2));
  if (hits && o->hits) memcpy(hits, o->hits, sizeof(int32_t) * (n + 2));
}

void orc_muca_set_sums(orc_t *o, const int64_t *gsum, const int64_t *nsum) {
  for (int i = 0; i < o->m.bins + 2 && o->gsum; i++) {
    if (gsum) o->gsum[i] = (long)gsum[i];
    if (nsum) o->nsum[i] = (long)nsum[i];
  }
}

void orc_muca_get_sums(const orc_t *o, int64_t *gsum, int64_t *nsum) {
  for (int i = 0; i < o->m.bins + 2 && o->gsum; i++) {
    if (gsum) gsum[i] = o->gsum[i];
    if (nsum) nsum[i] = o->nsum[i];
  }
}

/* ------------------------------------------------------------------ */
/* passes and the sweep (src/update.c) */

static void count(orc_t *o, int kind, int acc, int h) {
  su2hip_counters *c = &o->c;
  if (h && kind == SU2HIP_K_DOUBLET_OVERRELAX) {
    c->acc_overrelax_doublet2 += acc;
    c->total_overrelax_doublet2++;
    return;
  }
  if (h && kind == SU2HIP_K_DOUBLET_METROPOLIS) {
    c->accepted_doublet2 += acc;
    c->total_doublet2++;
    return;
  }
  switch (kind) {
    case SU2HIP_K_LINK_HEATBATH:
    case SU2HIP_K_LINK_METROPOLIS:
      c->accepted_su2link += acc;
      c->total_su2link++;
      break;
    case SU2HIP_K_LINK_OVERRELAX:
      c->accepted_link_or += acc;
      c->total_link_or++;
      break;
    case SU2HIP_K_U1_METROPOLIS:
      c->accepted_u1link += acc;
      c->total_u1link++;
      break;
    case SU2HIP_K_DOUBLET_OVERRELAX:
      c->acc_overrelax_doublet += acc;
      c->total_overrelax_doublet++;
      break;
    case SU2HIP_K_DOUBLET_METROPOLIS:
      c->accepted_doublet += acc;
      c->total_doublet++;
      break;
    case SU2HIP_K_TRIPLET_OVERRELAX:
      c->acc_overrelax_triplet += acc;
      c->total_overrelax_triplet++;
      break;
    case SU2HIP_K_TRIPLET_METROPOLIS:
      c->accepted_triplet += acc;
      c->total_triplet++;
      break;
    case SU2HIP_K_SINGLET_OVERRELAX:
      c->acc_overrelax_singlet += acc;
      c->total_overrelax_singlet++;
      break;
    case SU2HIP_K_SINGLET_METROPOLIS:
      c->accepted_singlet += acc;
      c->total_singlet++;
      break;
  }
}

/* cp_field over one parity (src/init.c:152-163) */
static void cp_parity(const orc_t *o, const double *src, double *dst, int dofs,
                      int par) {
  long off = par == SU2HIP_EVEN ? 0 : o->even;
  long max = par == SU2HIP_EVEN ? o->even : o->vol;
  memcpy(dst + off * dofs, src + off * dofs, sizeof(double) * dofs * (max - off));
}

/* checkerboard_sweep_* (src/update.c:35-259) for one kind/parity/dir (for
 * the doublet kinds dir is the doublet), including the multicanonical
 * checks of src/update.c:95-153 */
void orc_pass(orc_t *o, int kind, int par, int dir) {
  uint64_t pass = o->pass++;
  long off = par == SU2HIP_EVEN ? 0 : o->even;
  long max = par == SU2HIP_EVEN ? o->even : o->vol;
  int field = 0; /* 1 doublet, 2 triplet */
  if (kind == SU2HIP_K_DOUBLET_OVERRELAX || kind == SU2HIP_K_DOUBLET_METROPOLIS) field = 1;
  if (kind == SU2HIP_K_TRIPLET_OVERRELAX || kind == SU2HIP_K_TRIPLET_METROPOLIS) field = 2;
  int h = field == 1 && o->p.nhiggs == 2 && dir;
  int do_muca = 0;
  if (field && o->m.enabled && o->m.do_acceptance) {
    int op = o->m.orderparam;
    if (field == 1 && h == 0 && (op == SU2HIP_PHISQ || op == SU2HIP_PHI2MINUSSIGMA2)) do_muca = 1;
    if (field == 1 && h == 1 && op == SU2HIP_PHI2SQ) do_muca = 1;
    if (field == 2 && (op == SU2HIP_SIGMASQ || op == SU2HIP_PHI2MINUSSIGMA2)) do_muca = 1;
  }
  double *fld = field == 1 ? (h ? o->phi2 : o->phi) : o->sig;
  double *bu = field == 1 ? (h ? o->bu_phi2 : o->bu_phi) : o->bu_sig;
  int dofs = field == 1 ? 4 : 3;
  long interval = o->vol, cnt = 0;
  uint32_t check = 0;
  if (do_muca) {
    cp_parity(o, fld, bu, dofs, par);
    interval = (max - off) / o->m.checks_per_sweep;
    if (interval <= 0) interval = 1;
  }
  for (long i = off; i < max; i++) {
    ctx(o, pass, (uint32_t)i);
    int acc = site_dispatch(o, kind, i, dir);
    count(o, kind, acc, h);
    if (do_muca) {
      cnt++;
      if (cnt % interval == 0) {
        int more = (i < max - 1);
        int a = muca_check(o, par, pass, check++);
        if (!a)
          cp_parity(o, bu, fld, dofs, par);
        else if (more)
          cp_parity(o, fld, bu, dofs, par);
      }
    }
  }
  if (do_muca && (max - off) % o->m.checks_per_sweep != 0) {
    int a = muca_check(o, par, pass, check++);
    if (!a) cp_parity(o, bu, fld, dofs, par);
  }
}

/* scalar kind of this round: OR unless forced Metropolis (src/update.c:333-360) */
static int scalar_round(int alg, int last, int kor, int kme) {
  if (last) return alg == SU2HIP_OVERRELAX ? kme : -1;
  return alg == SU2HIP_OVERRELAX ? kor : kme;
}

void orc_set_random_order(orc_t *o, int on) { o->random_order = on != 0; }

/* shuffle (src/update.c:461-473) with iran() = the 64-bit word of draw k */
static void shuffle(int *a, int len, uint64_t seed, uint64_t pass, uint32_t *k) {
  for (int i = len - 1; i > 0; i--) {
    unsigned int j = orc_word(seed, pass, ORC_SCHED_SITE, (*k)++) % (unsigned long long)(i + 1);
    int t = a[i];
    a[i] = a[j];
    a[j] = t;
  }
}

/* update_lattice (src/update.c:270-430) with link_overrelax rounds after
 * the heatbath/Metropolis link passes (in the same order).  With the random
 * order (random_sweeps, src/update.c:292-300, :341-345) the sweep's
 * shuffles and parity draws come from pass = the first pass of the sweep,
 * site ORC_SCHED_SITE, in the reference's order: per link round a shuffle
 * of the parities then of the directions (5 draws each), per scalar round
 * one uniform for the first parity. */
void orc_update(orc_t *o, int n) {
  const su2hip_params *p = &o->p;
  for (int it = 0; it < n; it++) {
    uint64_t sp = o->pass;
    uint32_t sk = 0;
    for (int k = 0; k < p->update_links; k++) {
      int lk = p->algorithm_su2link == SU2HIP_HEATBATH ? SU2HIP_K_LINK_HEATBATH
                                                        : SU2HIP_K_LINK_METROPOLIS;
      int par_a[2 * D], dir_a[2 * D];
      for (int j = 0; j < 2 * D; j++) {
        par_a[j] = j & 1;
        dir_a[j] = j >> 1;
      }
      if (o->random_order) {
        shuffle(par_a, 2 * D, o->seed, sp, &sk);
        shuffle(dir_a, 2 * D, o->seed, sp, &sk);
      }
      for (int j = 0; j < 2 * D; j++) orc_pass(o, lk, par_a[j], dir_a[j]);
      for (int r = 0; r < p->link_overrelax; r++)
        for (int j = 0; j < 2 * D; j++) orc_pass(o, SU2HIP_K_LINK_OVERRELAX, par_a[j], dir_a[j]);
      if (p->u1)
        for (int j = 0; j < 2 * D; j++) orc_pass(o, SU2HIP_K_U1_METROPOLIS, par_a[j], dir_a[j]);
    }
    int ns = p->scalar_sweeps + 1;
    for (int s = 0; s < ns; s++) {
      int last = (s >= ns - 1);
      int p0 = 0;
      if (o->random_order)
        p0 = orc_uniform(o->seed, sp, ORC_SCHED_SITE, sk++) < 0.5 ? SU2HIP_EVEN : SU2HIP_ODD;
      int pa[2] = {p0, 1 - p0};
      if (p->nhiggs) {
        int kd = scalar_round(p->algorithm_su2doublet, last, SU2HIP_K_DOUBLET_OVERRELAX,
                              SU2HIP_K_DOUBLET_METROPOLIS);
        if (kd >= 0)
          for (int k = 0; k < p->update_su2doublet; k++)
            for (int h = 0; h < p->nhiggs; h++)
              for (int j = 0; j < 2; j++) orc_pass(o, kd, pa[j], h);
      }
      if (p->triplet) {
        int kt = scalar_round(p->algorithm_su2triplet, last, SU2HIP_K_TRIPLET_OVERRELAX,
                              SU2HIP_K_TRIPLET_METROPOLIS);
        if (kt >= 0)
          for (int k = 0; k < p->update_su2triplet; k++)
            for (int j = 0; j < 2; j++) orc_pass(o, kt, pa[j], 0);
      }
      if (p->singlet) {
        int ks = scalar_round(p->algorithm_singlet, last, SU2HIP_K_SINGLET_OVERRELAX,
                              SU2HIP_K_SINGLET_METROPOLIS);
        if (ks >= 0)
          for (int k = 0; k < p->update_singlet; k++)
            for (int j = 0; j < 2; j++) orc_pass(o, ks, pa[j], 0);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* magnetic charge (src/magfield.c, TRIPLET measure columns) */

typedef struct {
  double re[4], im[4]; /* row-major 2x2 */
} cmat;

static void cmat_mul(const cmat *a, const cmat *b, int dag, cmat *r) {
  cmat bb = *b;
  if (dag) {
    bb.re[0] = b->re[0];
    bb.im[0] = -b->im[0];
    bb.re[1] = b->re[2];
    bb.im[1] = -b->im[2];
    bb.re[2] = b->re[1];
    bb.im[2] = -b->im[1];
    bb.re[3] = b->re[3];
    bb.im[3] = -b->im[3];
  }
  cmat t;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      double sr = 0, si = 0;
      for (int k = 0; k < 2; k++) {
        double ar = a->re[2 * i + k], ai = a->im[2 * i + k];
        double br = bb.re[2 * k + j], bi = bb.im[2 * k + j];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      t.re[2 * i + j] = sr;
      t.im[2 * i + j] = si;
    }
  *r = t;
}

/* (1 + n.sigma)/2 with n = Sigma/|Sigma| (projector, src/magfield.c:90-104) */
static void halfproj(const double *a, cmat *m) {
  double md = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  double n0 = a[0] / md, n1 = a[1] / md, n2 = a[2] / md;
  m->re[0] = 0.5 * (1 + n2);
  m->im[0] = 0.0;
  m->re[1] = 0.5 * n0;
  m->im[1] = -0.5 * n1;
  m->re[2] = 0.5 * n0;
  m->im[2] = 0.5 * n1;
  m->re[3] = 0.5 * (1 - n2);
  m->im[3] = 0.0;
}

/* project_u1 (src/magfield.c:115-184): Pi(x) U_d(x) Pi(x+d) */
static void project_u1(const orc_t *o, long i, int d, cmat *r) {
  cmat pl, pr, u, t;
  halfproj(SIG(o, i), &pl);
  halfproj(SIG(o, NX(o, i, d)), &pr);
  const double *q = LNK(o, i, d);
  u.re[0] = q[0];
  u.im[0] = q[3];
  u.re[1] = q[2];
  u.im[1] = q[1];
  u.re[2] = -q[2];
  u.im[2] = q[1];
  u.re[3] = q[0];
  u.im[3] = -q[3];
  cmat_mul(&pl, &u, 0, &t);
  cmat_mul(&t, &pr, 0, r);
}

/* alpha_proj (src/magfield.c:189-216) */
static double alpha_proj(const orc_t *o, long i, int d1, int d2) {
  cmat u1, u2, u3, u4;
  project_u1(o, i, d1, &u1);
  project_u1(o, NX(o, i, d1), d2, &u2);
  project_u1(o, NX(o, i, d2), d1, &u3);
  project_u1(o, i, d2, &u4);
  cmat_mul(&u1, &u2, 0, &u1);
  cmat_mul(&u1, &u3, 1, &u1);
  cmat_mul(&u1, &u4, 1, &u1);
  double al = atan2(u1.im[0] + u1.im[3], u1.re[0] + u1.re[3]);
  return al * sqrt(o->p.betasu2);
}

/* magfield (src/magfield.c:221-257) */
static double magfield(const orc_t *o, long i, int dir) {
  double res = 0.0;
  for (int d1 = 0; d1 < D; d1++) {
    if (d1 == dir) continue;
    for (int d2 = d1 + 1; d2 < D; d2++) {
      if (d2 == dir) continue;
      double al = alpha_proj(o, i, d1, d2);
      if (d1 > dir)
        res += al;
      else if (d2 > dir)
        res -= al;
      else
        res += al;
    }
  }
  return res;
}

/* ------------------------------------------------------------------ */
/* correlators (the reference's CORRELATORS build, src/correlation.c) */

/* a.(b.(c.d)) */
static void qmul4(const double *a, const double *b, const double *c, const double *d, double *r) {
  double t[4], u[4];
  qmul(c, d, t);
  qmul(b, t, u);
  qmul(a, u, r);
}

/* clover_su2 (src/su2u1.c:207-275): the four plaquettes of the (d1, d2)
 * plane with a corner at site i, added */
static void clover_su2(const orc_t *o, long i, int d1, int d2, double *cl) {
  double a[4], b[4], c[4], d[4], t[4];
  long s;
  /* su2plaquette (src/su2u1.c:119-136) */
  qconj(LNK(o, NX(o, i, d2), d1), c);
  qconj(LNK(o, i, d2), d);
  qmul4(LNK(o, i, d1), LNK(o, NX(o, i, d1), d2), c, d, cl);
  /* U_d2(x) U_d1(x+d2-d1)^+ U_d2(x-d1)^+ U_d1(x-d1) */
  s = PV(o, NX(o, i, d2), d1);
  qconj(LNK(o, s, d1), b);
  s = PV(o, i, d1);
  qconj(LNK(o, s, d2), c);
  qmul4(LNK(o, i, d2), b, c, LNK(o, s, d1), t);
  for (int k = 0; k < 4; k++) cl[k] += t[k];
  /* U_d1(x-d1)^+ U_d2(x-d1-d2)^+ U_d1(x-d1-d2) U_d2(x-d2) */
  s = PV(o, i, d1);
  qconj(LNK(o, s, d1), a);
  s = PV(o, s, d2);
  qconj(LNK(o, s, d2), b);
  memcpy(c, LNK(o, s, d1), sizeof(c));
  s = PV(o, i, d2);
  qmul4(a, b, c, LNK(o, s, d2), t);
  for (int k = 0; k < 4; k++) cl[k] += t[k];
  /* U_d2(x-d2)^+ U_d1(x-d2) U_d2(x+d1-d2) U_d1(x)^+ */
  s = PV(o, i, d2);
  qconj(LNK(o, s, d2), a);
  memcpy(b, LNK(o, s, d1), sizeof(b));
  s = NX(o, s, d1);
  qconj(LNK(o, i, d1), d);
  qmul4(a, b, LNK(o, s, d2), d, t);
  for (int k = 0; k < 4; k++) cl[k] += t[k];
}

static double alpha_proj(const orc_t *o, long i, int d1, int d2);

/* measure_correlators (src/correlation.c:339-391) without the file:
 * plane sums along `dir` (plane_sum, :16-34) of doubletsq, tripletsq and
 * the photon operators of projected_photon_operator (:91-156, mode 0) and
 * photon_projection_full (:207-255, mode 1) at transverse momenta 1 and 2
 * (photon_correl, :261-305, whose momentum loop stops after dir0 = 0), then
 * the correlators at distances 0 .. L[dir]/2-1, 10 values each. */
int orc_correlators(const orc_t *o, int dir, double *out) {
  int Ld = o->L[dir];
  int d1 = dir == 0 ? 1 : 0, d2 = dir == 2 ? 1 : 2;
  double *pl = (double *)calloc((size_t)Ld * 10, sizeof(double));
  for (long i = 0; i < o->vol; i++) {
    long nat = o->nat[i];
    long c[D] = {nat % o->L[0], (nat / o->L[0]) % o->L[1], nat / ((long)o->L[0] * o->L[1])};
    double *q = pl + (size_t)c[dir] * 10;
    if (o->p.nhiggs) q[0] += doubletsq(PHI(o, i));
    if (o->p.triplet) {
      q[1] += tripletsq(SIG(o, i));
      double cl[4];
      clover_su2(o, i, d1, d2, cl);
      const double *a = SIG(o, i);
      double tr = a[0] * (cl[1] / 4.0) + a[1] * (cl[2] / 4.0) + a[2] * (cl[3] / 4.0);
      double al = 2.0 * alpha_proj(o, i, d1, d2);
      for (int m = 1; m <= 2; m++) {
        double px = dir != 0 ? (double)(m * c[0]) / (double)o->L[0] : 0.0;
        double cs = cos(2.0 * M_PI * px), sn = sin(2.0 * M_PI * px);
        q[2 * m] += tr * cs;
        q[2 * m + 1] += tr * sn;
        q[4 + 2 * m] += al * cs;
        q[4 + 2 * m + 1] += al * sn;
      }
    }
  }
  int maxd = Ld / 2;
  for (int d = 0; d < maxd; d++) {
    double *r = out + (size_t)d * 10;
    for (int k = 0; k < 10; k++) r[k] = 0.0;
    for (int z = 0; z < Ld; z++) {
      const double *a = pl + (size_t)z * 10, *b = pl + (size_t)((z + d) % Ld) * 10;
      r[0] += a[0] * b[0];
      r[1] += a[1] * b[1];
      for (int k = 2; k < 10; k += 2) {
        /* cmult(op1, conj(op2)) */
        double br = b[k], bi = -b[k + 1];
        r[k] += a[k] * br - a[k + 1] * bi;
        r[k + 1] += br * a[k + 1] + a[k] * bi;
      }
    }
    r[0] *= 4.0 / (double)o->vol;
    for (int k = 1; k < 10; k++) r[k] /= (double)o->vol;
  }
  free(pl);
  return maxd;
}

/* longest direction; of equal ones the last (src/layout.c:771-777) */
int orc_longest_dir(const orc_t *o) {
  int l = D - 1;
  for (int d = D - 1; d >= 0; d--)
    if (o->L[d] > o->L[l]) l = d;
  return l;
}

/* magcharge_cube (src/magfield.c:262-280) */
static double magcharge_cube(const orc_t *o, long i) {
  double res = 0.0;
  for (int d = 0; d < D; d++)
    res += magfield(o, NX(o, i, d), d) - magfield(o, i, d);
  return res;
}

/* ------------------------------------------------------------------ */
/* measure() site loop (src/measure.c:110-174), raw sums */
void orc_measure(const orc_t *o, double *sums) {
  const su2hip_params *p = &o->p;
  double s[SU2HIP_NSUMS];
  memset(s, 0, sizeof(s));
  for (long i = 0; i < o->vol; i++) {
    s[SU2HIP_SUM_ACTION] += orc_action_local(o, i);
    s[SU2HIP_SUM_WILSON] += local_su2wilson(o, i);
    if (p->u1) s[SU2HIP_SUM_U1WILSON] += local_u1wilson(o, i);
    if (p->nhiggs) {
      double mod = doubletsq(PHI(o, i));
      s[SU2HIP_SUM_COV_PHI] += covariant_doublet(o, i, 0);
      s[SU2HIP_SUM_PHI2] += mod;
      s[SU2HIP_SUM_PHI4] += mod * mod;
      if (p->triplet) s[SU2HIP_SUM_PHI2SIGMA2] += mod * tripletsq(SIG(o, i));
    }
    if (p->nhiggs == 2) {
      double mod = doubletsq(PHIH(o, 1, i)), R, I;
      s[SU2HIP_SUM_COV_PHI_2] += covariant_doublet(o, i, 1);
      s[SU2HIP_SUM_PHI2_2] += mod;
      s[SU2HIP_SUM_PHI4_2] += mod * mod;
      phi12(PHI(o, i), PHIH(o, 1, i), &R, &I);
      s[SU2HIP_SUM_R12] += R;
      s[SU2HIP_SUM_I12] += I;
    }
    if (p->triplet) {
      double tm = tripletsq(SIG(o, i));
      s[SU2HIP_SUM_SIGMA2] += tm;
      s[SU2HIP_SUM_SIGMA4] += tm * tm;
      for (int d = 0; d < D; d++) s[SU2HIP_SUM_HOP_SIGMA] += hop_triplet_fwd(o, i, d);
      double ch = magcharge_cube(o, i);
      s[SU2HIP_SUM_MAGCHARGE] += ch;
      s[SU2HIP_SUM_MAGCHARGE_ABS] += fabs(ch);
    }
    if (p->singlet) {
      double S = o->S[i];
      for (int d = 0; d < D; d++) {
        double Sn = o->S[NX(o, i, d)];
        s[SU2HIP_SUM_S_COV] += S * S - S * Sn;
      }
      s[SU2HIP_SUM_S] += S;
      s[SU2HIP_SUM_S2] += S * S;
      s[SU2HIP_SUM_S3] += S * S * S;
      s[SU2HIP_SUM_S4] += S * S * S * S;
      if (p->nhiggs == 1) {
        double mod = doubletsq(PHI(o, i));
        s[SU2HIP_SUM_SPHISQ] += mod * S;
        s[SU2HIP_SUM_S2PHISQ] += mod * S * S;
      }
    }
  }
  memcpy(sums, s, sizeof(s));
}

/* ------------------------------------------------------------------ */
/* gradient flow (src/gradflow.c) */

/* grad_force_link (src/gradflow.c:190-258) and grad_force_triplet
 * (:267-330) for every site (calc_gradient, :335-354) */
void orc_flow_force(const orc_t *o, double *force, double *tforce) {
  const su2hip_params *p = &o->p;
  for (long i = 0; i < o->vol; i++) {
    for (int d = 0; d < D; d++) {
      double s[4], m[4];
      orc_wilson_staple(o, i, d, s);
      for (int a = 0; a < 4; a++) s[a] *= -p->betasu2 / (8.0);
      if (p->triplet) {
        /* s <- s - Sigma(x+d) U^+ Sigma(x), Sigma = a.sigma/2 */
        double a2[4] = {0.0, SIG(o, NX(o, i, d))[0], SIG(o, NX(o, i, d))[1],
                        SIG(o, NX(o, i, d))[2]};
        double a1[4] = {0.0, SIG(o, i)[0], SIG(o, i)[1], SIG(o, i)[2]};
        double t[4], r[4];
        qmul_nc(a2, LNK(o, i, d), t);
        qmul(t, a1, r);
        for (int a = 0; a < 4; a++) s[a] += r[a] / 4.0;
      }
      qmul(LNK(o, i, d), s, m);
      double *f = force + ((size_t)i * D + d) * 4;
      f[0] = 0.0;
      for (int a = 1; a < 4; a++) {
        f[a] = 2.0 * m[a];
        f[a] *= 4.0 / p->betasu2;
      }
    }
    if (p->triplet && tforce) {
      double res[3];
      orc_triplet_staple(o, i, res);
      const double *tr = SIG(o, i);
      double tsq = tripletsq(tr);
      for (int a = 0; a < 3; a++) {
        res[a] += (2.0 * D + p->msq_triplet + 2.0 * p->b4 * tsq) * tr[a];
        if (p->nhiggs) res[a] += p->a2 * doubletsq(PHI(o, i)) * tr[a];
      }
      for (int a = 0; a < 3; a++) tforce[(size_t)i * 3 + a] = -1.0 * res[a];
    }
  }
}

/* V <- exp(i z.sigma) V (flow_gauge's per-link step, src/gradflow.c:370-395);
 * a zero vector leaves the link unchanged (the reference returns from the
 * whole update there, src/gradflow.c:377-381: a deliberate divergence). */
static void expmul(const double *z, double *v) {
  double mod = z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
  if (mod <= 0.0) return;
  mod = sqrt(mod);
  double u[4];
  u[0] = cos(mod);
  for (int a = 0; a < 3; a++) u[a + 1] = sin(mod) * z[a] / mod;
  qmul(u, v, v);
}

void orc_flow_step(orc_t *o, int scheme, double dt) {
  size_t V = (size_t)o->vol;
  double *F = (double *)malloc(sizeof(double) * V * D * 4);
  double *TF = o->p.triplet ? (double *)malloc(sizeof(double) * V * 3) : NULL;
  if (scheme == SU2HIP_FLOW_EULER) {
    orc_flow_force(o, F, TF);
    for (size_t l = 0; l < V * D; l++) {
      double z[3];
      for (int a = 0; a < 3; a++) z[a] = dt * F[l * 4 + a + 1];
      expmul(z, o->U + l * 4);
    }
    if (TF)
      for (size_t l = 0; l < V * 3; l++) o->sig[l] += dt * TF[l];
  } else {
    /* Luscher's RK3 (arXiv:1006.4518 appendix C):
     *   W1 = exp(Z0/4) W0, W2 = exp(8/9 Z1 - 17/36 Z0) W1,
     *   W3 = exp(3/4 Z2 - 8/9 Z1 + 17/36 Z0) W2,  Zk = dt F(Wk),
     * in 2N-storage form X <- c_k Zk - d_k X, W <- exp(X) W with
     * (c,d) = (1/4,0), (8/9,17/9), (3/4,1). */
    static const double c[3] = {0.25, 8.0 / 9.0, 0.75};
    static const double dd[3] = {0.0, 17.0 / 9.0, 1.0};
    double *Z = (double *)calloc(V * D * 3, sizeof(double));
    double *ZT = TF ? (double *)calloc(V * 3, sizeof(double)) : NULL;
    for (int st = 0; st < 3; st++) {
      orc_flow_force(o, F, TF);
      for (size_t l = 0; l < V * D; l++) {
        double *z = Z + l * 3;
        for (int a = 0; a < 3; a++) z[a] = c[st] * dt * F[l * 4 + a + 1] - dd[st] * z[a];
        expmul(z, o->U + l * 4);
      }
      if (TF)
        for (size_t l = 0; l < V * 3; l++) {
          ZT[l] = c[st] * dt * TF[l] - dd[st] * ZT[l];
          o->sig[l] += ZT[l];
        }
    }
    free(Z);
    free(ZT);
  }
  free(F);
  free(TF);
}

/* remove_counterterms (src/gradflow.c:424-459): the triplet's tree-level
 * mass for the flow (2-loop lattice counterterm constants of the reference) */
static double triplet_flow_msq(const su2hip_params *p) {
  double Sg = 3.17591153625, zeta = 0.08849, delta = 1.942130, rho = -0.313964;
  double k1 = 0.958382, k2 = 0.25 * Sg * Sg - 0.5 * delta - 0.25;
  double k3 = 0.751498, k4 = 1.204295;
  double gsq = 4.0 / p->betasu2, RG = gsq, b4 = p->b4;
  double ct = -(4.0 * gsq + 5.0 * b4) * Sg / (4.0 * M_PI);
  ct += -1.0 / (16.0 * M_PI * M_PI) *
        ((20.0 * b4 * gsq - 10.0 * b4 * b4) * (log(6.0 / (RG)) + zeta) +
         20.0 * b4 * gsq * (0.25 * Sg * Sg - delta) +
         2.0 * gsq * gsq *
             (1.25 * Sg * Sg + M_PI / 3.0 * Sg - 6.0 * delta - 6.0 * rho + 4.0 * k1 - k2 -
              k3 - 3.0 * k4));
  return p->msq_triplet - ct;
}

/* grad_flow (src/gradflow.c:46-173) on a copy; returns the number of rows */
int orc_flow(orc_t *o, int scheme, double dt, double t_max, int meas_interval,
             double *rows, int max_rows) {
  size_t V = (size_t)o->vol;
  double *U0 = (double *)malloc(sizeof(double) * V * D * 4);
  double *S0 = (double *)malloc(sizeof(double) * V * 3);
  memcpy(U0, o->U, sizeof(double) * V * D * 4);
  memcpy(S0, o->sig, sizeof(double) * V * 3);
  double msq = o->p.msq_triplet;
  if (o->p.triplet) o->p.msq_triplet = triplet_flow_msq(&o->p);
  int nr = 0;
  double t = 0.0;
#define ROW()                                                \
  do {                                                       \
    if (nr < max_rows) {                                     \
      rows[(size_t)nr * (1 + SU2HIP_NSUMS)] = t;             \
      orc_measure(o, rows + (size_t)nr * (1 + SU2HIP_NSUMS) + 1); \
    }                                                        \
    nr++;                                                    \
  } while (0)
  ROW();
  int iter = 1;
  while (t < t_max) {
    if (t + dt > t_max) dt = t_max - t;
    orc_flow_step(o, scheme, dt);
    t += dt;
    if (iter % meas_interval == 0) ROW();
    iter++;
  }
  if ((iter - 1) % meas_interval != 0) ROW();
#undef ROW
  o->p.msq_triplet = msq;
  memcpy(o->U, U0, sizeof(double) * V * D * 4);
  memcpy(o->sig, S0, sizeof(double) * V * 3);
  free(U0);
  free(S0);
  return nr;
}

/* prepare_wall (src/z_coord.c:304-376) with the deviations drawn from the
 * Philox stream (pass ORC_WALL_PASS) in the reference's per-site order:
 * doublets, triplet, singlet, then the links of every direction */
void orc_wall_start(orc_t *o, int zdir) {
  if (zdir < 0) zdir = orc_longest_dir(o);
  for (long i = 0; i < o->vol; i++) {
    long nat = o->nat[i];
    long c[D] = {nat % o->L[0], (nat / o->L[0]) % o->L[1], nat / ((long)o->L[0] * o->L[1])};
    int lo = c[zdir] < 0.5 * o->L[zdir];
    ctx(o, ORC_WALL_PASS, (uint32_t)i);
    for (int h = 0; h < o->p.nhiggs; h++) {
      double a = lo ? 0.2 : 1.2, e = lo ? 0.01 : 0.05;
      double *f = PHIH(o, h, i);
      f[0] = a + e * dran(o);
      f[1] = e * dran(o);
      f[2] = e * dran(o);
      f[3] = e * dran(o);
    }
    if (o->p.triplet) {
      double a = lo ? 1.5 : 0.2, e = lo ? 0.05 : 0.01;
      double *g = SIG(o, i);
      g[0] = a + e * dran(o);
      g[1] = e * dran(o);
      g[2] = e * dran(o);
    }
    if (o->p.singlet) o->S[i] = (lo ? 0.8 : 0.1) + 0.01 * dran(o);
    for (int d = 0; d < D; d++) {
      double u = 1.0 - (lo ? 0.03 : 0.4) * dran(o);
      double *l = LNK(o, i, d);
      l[0] = u;
      l[1] = sqrt((double)(1.0 - u * u));
      l[2] = 0.0;
      l[3] = 0.0;
    }
  }
}

/* measure_along_z (src/z_coord.c:206-298): plane averages along dir of
 * doubletsq, tripletsq, S and S^2, out[z][4]; returns L[dir] */
int orc_zprofile(const orc_t *o, int dir, double *out) {
  if (dir < 0) dir = orc_longest_dir(o);
  int Ld = o->L[dir];
  double area = (double)(o->vol / Ld);
  for (int k = 0; k < Ld * 4; k++) out[k] = 0.0;
  for (long i = 0; i < o->vol; i++) {
    long nat = o->nat[i];
    long c[D] = {nat % o->L[0], (nat / o->L[0]) % o->L[1], nat / ((long)o->L[0] * o->L[1])};
    double *r = out + (size_t)c[dir] * 4;
    if (o->p.nhiggs) r[0] += doubletsq(PHI(o, i));
    if (o->p.triplet) r[1] += tripletsq(SIG(o, i));
    if (o->p.singlet) {
      r[2] += o->S[i];
      r[3] += o->S[i] * o->S[i];
    }
  }
  for (int k = 0; k < Ld * 4; k++) out[k] /= area;
  return Ld;
}


/* ------------------------------------------------------------------ */
/* blocking (the reference's BLOCKING build, src/blocking.c, su2u1.c:736-895) */

/* max_block_level (src/blocking.c:19-38) */
static int max_block_level(const int *L, const int *bdir) {
  int count = 0, len[D];
  for (int d = 0; d < D; d++) len[d] = L[d];
  for (;;) {
    for (int d = 0; d < D; d++) {
      if (bdir[d] && (len[d] % 2 != 0 || len[d] <= 2)) return count;
      if (bdir[d]) len[d] /= 2;
    }
    count++;
  }
}

/* su2staple_wilson_onedir with dagger = 1 (src/staples.c:197-234): the
 * conjugates of the upper and lower staples of U_mu(x) in the nu plane */
static void staple_onedir_dag(const orc_t *o, long i, int mu, int nu, double *res) {
  double a[4], b[4], t[4], up[4], lo[4];
  /* U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ */
  qconj(LNK(o, NX(o, i, nu), mu), a);
  qconj(LNK(o, i, nu), b);
  qmul(a, b, t);
  qmul(LNK(o, NX(o, i, mu), nu), t, up);
  /* U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) */
  long s = PV(o, NX(o, i, mu), nu);
  qconj(LNK(o, s, nu), a);
  qconj(LNK(o, PV(o, i, nu), mu), b);
  qmul(b, LNK(o, PV(o, i, nu), nu), t);
  qmul(a, t, lo);
  res[0] = up[0] + lo[0];
  for (int k = 1; k < 4; k++) res[k] = -up[k] + -lo[k];
}

/* smear_link (src/su2u1.c:752-797) */
static void smear_link(const orc_t *o, const int *bdir, long i, int dir, double *res) {
  double v2[4], st[4];
  memcpy(res, LNK(o, i, dir), sizeof(double) * 4);
  memcpy(v2, LNK(o, NX(o, i, dir), dir), sizeof(double) * 4);
  int paths = 1;
  for (int j = 0; j < D; j++) {
    if (j == dir || !bdir[j]) continue;
    staple_onedir_dag(o, i, dir, j, st);
    for (int k = 0; k < 4; k++) res[k] += st[k];
    staple_onedir_dag(o, NX(o, i, dir), dir, j, st);
    for (int k = 0; k < 4; k++) v2[k] += st[k];
    paths += 2;
  }
  for (int k = 0; k < 4; k++) {
    res[k] /= (double)paths;
    v2[k] /= (double)paths;
  }
  double r[4];
  qmul(res, v2, r);
  double det = sqrt(qsq(r));
  for (int k = 0; k < 4; k++) res[k] = r[k] / det;
}

/* smear_triplet (src/su2u1.c:810-857): Sigma(x) + sum over the smearing
 * directions of U_j(x) Sigma(x+j) U_j(x)^+ + U_j(x-j)^+ Sigma(x-j) U_j(x-j),
 * over the number of sites */
static void smear_triplet(const orc_t *o, const int *bdir, long i, double *res) {
  int sites = 1;
  double cov[3] = {0, 0, 0}, r[3], uc[4];
  for (int d = 0; d < D; d++) {
    if (!bdir[d]) continue;
    rotate(LNK(o, i, d), SIG(o, NX(o, i, d)), r);
    for (int k = 0; k < 3; k++) cov[k] += r[k];
    long pv = PV(o, i, d);
    qconj(LNK(o, pv, d), uc);
    rotate(uc, SIG(o, pv), r);
    for (int k = 0; k < 3; k++) cov[k] += r[k];
    sites += 2;
  }
  for (int k = 0; k < 3; k++) res[k] = (SIG(o, i)[k] + cov[k]) / (double)sites;
}

/* Correlators on the blocked lattices (measure_blocked_correlators,
 * src/correlation.c:393-421, with the blocking of main, src/main.c:106-152
 * and smear_fields/make_blocked_fields): block every direction but the
 * longest, level after level, smearing the triplet and the links of the
 * blocked directions.  out[level][L[dir]/2][10]; returns the number of
 * levels (at most `levels`, capped like max_block_level).  Only triplet
 * models: with a doublet the reference measures nothing (returns 0). */
int orc_blocked_correlators(const orc_t *o, int levels, double *out) {
  if (o->p.nhiggs || !o->p.triplet) return 0;
  int dir = orc_longest_dir(o);
  int bdir[D];
  for (int d = 0; d < D; d++) bdir[d] = d != dir;
  int max = max_block_level(o->L, bdir);
  if (levels > max) levels = max;
  int maxd = o->L[dir] / 2;
  const orc_t *base = o;
  orc_t *prev = NULL;
  for (int k = 0; k < levels; k++) {
    su2hip_params pb = base->p;
    for (int d = 0; d < D; d++) pb.L[d] = bdir[d] ? base->L[d] / 2 : base->L[d];
    orc_t *b = orc_create(&pb);
    for (long ib = 0; ib < b->vol; ib++) {
      long nb = b->nat[ib];
      long cb[D] = {nb % b->L[0], (nb / b->L[0]) % b->L[1], nb / ((long)b->L[0] * b->L[1])};
      long c[D];
      for (int d = 0; d < D; d++) c[d] = bdir[d] ? 2 * cb[d] : cb[d];
      long nat = c[0] + base->L[0] * (c[1] + (long)base->L[1] * c[2]);
      long i = 0;
      while (base->nat[i] != nat) i++; /* the fine site (small lattices only) */
      for (int d = 0; d < D; d++) {
        if (bdir[d])
          smear_link(base, bdir, i, d, LNK(b, ib, d));
        else
          memcpy(LNK(b, ib, d), LNK(base, i, d), sizeof(double) * 4);
      }
      smear_triplet(base, bdir, i, SIG(b, ib));
    }
    orc_correlators(b, dir, out + (size_t)k * maxd * 10);
    if (prev) orc_destroy(prev);
    prev = b;
    base = b;
  }
  if (prev) orc_destroy(prev);
  return levels;
}
