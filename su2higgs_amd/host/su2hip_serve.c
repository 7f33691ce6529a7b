/* su2hip_serve.c -- request loop over the su2hip ABI (include/su2hip_serve.h).
 *
 * One su2hip batch holds every resident chain.  A scheduling tick
 *   1. moves requests whose arrival time has passed into a FIFO queue,
 *   2. admits queued requests head first while their device footprint fits
 *      under the HBM cap and a slot is free (idle kept chains are evicted
 *      oldest first to make room); while the head waits for room, up to
 *      `backfill` later requests for new chains that fit now go ahead of it,
 *      until the head has waited `max_head_wait` seconds,
 *   3. sweeps every running chain by min(tick_sweeps, smallest remaining
 *      budget) compound sweeps -- lattices whose budget is spent sit out --
 *      and waits for the stream,
 *   4. measures the chains that finished (one FP64 reduction launch for the
 *      whole batch), runs their gradient flow, and retires them.
 * A request's result depends only on its chain: Philox key and pass counter
 * belong to the lattice, and the volume sums reduce each lattice in a fixed
 * tree order, so the same request gives the same bits alone or in any mix.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/su2hip.h"
#include "../../include/su2hip_serve.h"

extern int su2hip_internal_fail(int code, const char *msg);

static int sfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return su2hip_internal_fail(code, buf);
}

void su2hip_serve_defaults(su2hip_serve_config *c) {
  if (!c) return;
  memset(c, 0, sizeof(*c));
  su2hip_params *p = &c->tmpl;
  p->L[0] = p->L[1] = p->L[2] = 24;
  p->nhiggs = 1;
  p->precision = SU2HIP_FP32;
  p->algorithm_su2link = SU2HIP_HEATBATH;
  p->algorithm_u1link = SU2HIP_METROPOLIS;
  p->algorithm_su2doublet = SU2HIP_OVERRELAX;
  p->algorithm_su2triplet = SU2HIP_OVERRELAX;
  p->algorithm_singlet = SU2HIP_OVERRELAX;
  p->update_links = 1;
  p->link_overrelax = 4;
  p->scalar_sweeps = 4;
  p->update_su2doublet = 1;
  p->update_su2triplet = 1;
  p->update_singlet = 1;
  p->reunitarize = 1;
  p->link_format = SU2HIP_LINKS_PACKED; /* many resident chains: bytes dominate */
  p->betasu2 = 8.0;
  p->betau1 = 8.0;
  p->r_u1 = 1.0;
  p->msq_phi = -0.25;
  p->lambda_phi = 0.02;
  p->seed = 1;
  c->hbm_budget = 8ull << 30;
  c->max_slots = 64;
  c->tick_sweeps = 16;
  c->use_graph = 1;
  c->realtime = 1;
  c->keep_chains = 0;
  c->flow_scheme = SU2HIP_FLOW_RK3;
  c->flow_dt = 0.01;
  c->flow_t_max = 4.0;
  c->phi0 = 1.0;
  c->backfill = 32;
  c->max_head_wait = 1.0;
  c->sigma0 = 0.0;
  c->singlet0 = 0.0;
}

/* scalar rounds of one field per sweep: update_lattice runs scalar_sweeps
 * rounds of the chosen algorithm plus a Metropolis round after
 * overrelaxation (src/update.c:333-360) */
static int scalar_rounds(int alg, int sweeps) {
  return sweeps + (alg == SU2HIP_OVERRELAX ? 1 : 0);
}

double su2hip_serve_updates_per_sweep(const su2hip_params *p) {
  if (!p) return 0.0;
  double u = (double)p->update_links * (1 + p->link_overrelax + (p->u1 ? 1 : 0)) * 3;
  if (p->nhiggs) u += (double)scalar_rounds(p->algorithm_su2doublet, p->scalar_sweeps) * p->update_su2doublet;
  if (p->triplet) u += (double)scalar_rounds(p->algorithm_su2triplet, p->scalar_sweeps) * p->update_su2triplet;
  if (p->singlet) u += (double)scalar_rounds(p->algorithm_singlet, p->scalar_sweeps) * p->update_singlet;
  return u;
}

/* ---- synthetic stream ---------------------------------------------------- */

static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double unif(uint64_t *s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

int su2hip_serve_stream(uint64_t seed, double rate, int n, int flow_every, su2hip_request *out) {
  static const int Ls[6] = {24, 32, 40, 48, 56, 64};
  if (n < 0 || (n > 0 && !out) || !(rate > 0)) return sfail(SU2HIP_EINVAL, "stream: n/out/rate");
  uint64_t s = seed;
  double t = 0.0;
  for (int i = 0; i < n; i++) {
    su2hip_request *r = &out[i];
    memset(r, 0, sizeof(*r));
    t += -log1p(-unif(&s)) / rate; /* exponential inter-arrival times */
    r->id = i;
    r->chain = -1;
    r->arrival = t;
    r->L = Ls[(int)(unif(&s) * 6)];
    r->betasu2 = 6.0 + 6.0 * unif(&s);
    r->msq_phi = -0.3 + 0.2 * unif(&s);
    r->lambda_phi = 0.005 + 0.025 * unif(&s);
    r->sweeps = 100 + (int64_t)(unif(&s) * 901);
    r->seed = splitmix64(&s);
    if (flow_every > 0 && (i + 1) % flow_every == 0) {
      r->L = 64;
      r->flow = 1;
    }
  }
  return SU2HIP_OK;
}

/* ---- the loop ------------------------------------------------------------ */

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void sleep_s(double s) {
  if (s <= 0) return;
  struct timespec ts;
  ts.tv_sec = (time_t)s;
  ts.tv_nsec = (long)((s - (double)ts.tv_sec) * 1e9);
  nanosleep(&ts, NULL);
}

typedef struct {
  int used;        /* a chain lives in this slot */
  int req;         /* running request index, -1 when idle */
  int64_t chain;   /* id of the request that created the chain */
  uint64_t bytes;  /* reserved footprint */
  int with_flow;   /* footprint includes the flow scratch */
  int64_t lru;     /* tick of the last use */
  int32_t L;
  int flowing;     /* its request's gradient flow is in flight */
  double vol;
} chain_slot;

typedef struct {
  const su2hip_serve_config *cfg;
  const su2hip_request *req;
  su2hip_result *res;
  su2hip_serve_stats st;
  su2hip_batch_t *b;
  chain_slot *cs;
  uint64_t reserved;
  int running, resident, flowing;
  double t0;
  int64_t tick;
  double upd;
} server;

static double elapsed(server *S) { return now_s() - S->t0; }

static void retire_slot(server *S, int slot) {
  chain_slot *c = &S->cs[slot];
  (void)su2hip_remove(S->b, slot);
  S->reserved -= c->bytes;
  S->resident--;
  memset(c, 0, sizeof(*c));
  c->req = -1;
}

/* evict the least recently used idle chain; 0 if there was none */
static int evict_one(server *S) {
  int best = -1;
  for (int i = 0; i < S->cfg->max_slots; i++)
    if (S->cs[i].used && S->cs[i].req < 0 && (best < 0 || S->cs[i].lru < S->cs[best].lru)) best = i;
  if (best < 0) return 0;
  retire_slot(S, best);
  S->st.evictions++;
  return 1;
}

static int free_slots(server *S) { return S->cfg->max_slots - S->resident; }

static void fail_request(server *S, int r, int code) {
  S->res[r].status = code;
  S->res[r].t_done = elapsed(S);
  S->st.failed++;
}

/* 1 admitted (or failed for good), 0 wait for room, <0 device error */
static int admit(server *S, int r) {
  const su2hip_request *q = &S->req[r];
  su2hip_result *o = &S->res[r];
  const su2hip_serve_config *cfg = S->cfg;
  if (q->sweeps < 0) {
    fail_request(S, r, sfail(SU2HIP_EINVAL, "request %lld: negative sweep count", (long long)q->id));
    return 1;
  }
  if (q->chain >= 0) {
    int slot = -1;
    for (int i = 0; i < cfg->max_slots; i++)
      if (S->cs[i].used && S->cs[i].chain == q->chain) slot = i;
    if (slot < 0) {
      fail_request(S, r, sfail(SU2HIP_ESTATE, "request %lld: chain %lld is not resident",
                               (long long)q->id, (long long)q->chain));
      return 1;
    }
    chain_slot *c = &S->cs[slot];
    if (c->req >= 0) return 0; /* the chain's previous request still runs */
    if (q->flow && !c->with_flow) {
      /* the chain now also needs the flow scratch */
      su2hip_params p = cfg->tmpl;
      uint64_t with = 0, without = 0;
      p.L[0] = p.L[1] = p.L[2] = c->L;
      if (su2hip_footprint(&p, 1, &with) || su2hip_footprint(&p, 0, &without)) {
        fail_request(S, r, SU2HIP_EINVAL);
        return 1;
      }
      uint64_t extra = with - without;
      c->req = r; /* not evictable while room is made */
      while (cfg->hbm_budget && S->reserved + extra > cfg->hbm_budget && evict_one(S)) {
      }
      c->req = -1;
      if (cfg->hbm_budget && S->reserved + extra > cfg->hbm_budget) {
        if (S->running) return 0;
        fail_request(S, r, sfail(SU2HIP_ENOMEM, "request %lld: flow scratch exceeds the HBM cap",
                                 (long long)q->id));
        return 1;
      }
      c->bytes += extra;
      c->with_flow = 1;
      S->reserved += extra;
    }
    int rc = su2hip_set_budget(S->b, slot, q->sweeps);
    if (rc) return rc;
    rc = su2hip_reset_counters(S->b, slot);
    if (rc) return rc;
    c->req = r;
    c->lru = S->tick;
    o->slot = slot;
    o->t_admit = elapsed(S);
    S->running++;
    return 1;
  }
  su2hip_params p = cfg->tmpl;
  p.L[0] = p.L[1] = p.L[2] = q->L;
  p.betasu2 = q->betasu2;
  p.msq_phi = q->msq_phi;
  p.lambda_phi = q->lambda_phi;
  p.seed = q->seed;
  uint64_t bytes = 0;
  int rc = su2hip_footprint(&p, q->flow ? 1 : 0, &bytes);
  if (rc) {
    fail_request(S, r, rc);
    return 1;
  }
  if (cfg->hbm_budget && bytes > cfg->hbm_budget) {
    fail_request(S, r, sfail(SU2HIP_ENOMEM, "request %lld: %llu bytes exceed the HBM cap",
                             (long long)q->id, (unsigned long long)bytes));
    return 1;
  }
  while ((cfg->hbm_budget && S->reserved + bytes > cfg->hbm_budget) || free_slots(S) == 0)
    if (!evict_one(S)) break;
  if ((cfg->hbm_budget && S->reserved + bytes > cfg->hbm_budget) || free_slots(S) == 0) return 0;
  int slot = -1;
  rc = su2hip_add(S->b, &p, &slot);
  if (rc == SU2HIP_ENOMEM && S->running) return 0; /* device full: wait for a chain to leave */
  if (rc == SU2HIP_EINVAL || rc == SU2HIP_ENOMEM) {
    fail_request(S, r, rc);
    return 1;
  }
  if (rc) return rc;
  if ((rc = su2hip_cold_start(S->b, slot, cfg->phi0, cfg->sigma0, cfg->singlet0)) ||
      (rc = su2hip_set_budget(S->b, slot, q->sweeps)) || (rc = su2hip_reset_counters(S->b, slot)))
    return rc;
  chain_slot *c = &S->cs[slot];
  c->used = 1;
  c->req = r;
  c->chain = q->id;
  c->bytes = bytes;
  c->with_flow = q->flow ? 1 : 0;
  c->lru = S->tick;
  c->L = q->L;
  c->vol = (double)q->L * q->L * q->L;
  S->reserved += bytes;
  S->resident++;
  S->running++;
  if (S->reserved > S->st.peak_bytes) S->st.peak_bytes = S->reserved;
  if (S->resident > S->st.peak_resident) S->st.peak_resident = S->resident;
  o->slot = slot;
  o->t_admit = elapsed(S);
  return 1;
}

/* a request is done: record it and retire (or park) its chain */
static void complete(server *S, int slot) {
  chain_slot *c = &S->cs[slot];
  int r = c->req;
  const su2hip_request *q = &S->req[r];
  su2hip_result *o = &S->res[r];
  o->status = 0;
  o->t_done = elapsed(S);
  S->st.completed++;
  S->st.site_updates += (double)q->sweeps * c->vol * S->upd;
  c->req = -1;
  c->lru = S->tick;
  if (!S->cfg->keep_chains) retire_slot(S, slot);
}

/* measure the chains whose budget is spent; start the gradient flow of those
 * that asked for one (on the second stream, while the others keep
 * sweeping) and complete the rest */
static int finish(server *S, const int64_t *budgets, double *sums) {
  const su2hip_serve_config *cfg = S->cfg;
  int any = 0, rc;
  for (int i = 0; i < cfg->max_slots; i++)
    if (S->cs[i].used && S->cs[i].req >= 0 && !S->cs[i].flowing && budgets[i] == 0) any = 1;
  if (!any) return SU2HIP_OK;
  if ((rc = su2hip_measure_all(S->b, sums))) return rc;
  for (int i = 0; i < cfg->max_slots; i++) {
    chain_slot *c = &S->cs[i];
    if (!c->used || c->req < 0 || c->flowing || budgets[i] != 0) continue;
    int r = c->req;
    const su2hip_request *q = &S->req[r];
    su2hip_result *o = &S->res[r];
    memcpy(o->sums, sums + (size_t)i * SU2HIP_NSUMS, sizeof(o->sums));
    uint64_t seed = 0;
    if ((rc = su2hip_rng_get(S->b, i, &seed, &o->pass))) return rc;
    o->sweeps_done = q->sweeps;
    S->running--;
    if (q->flow) {
      /* rows at t = 0 and t = t_max only (one measurement interval longer
       * than the flow) */
      if ((rc = su2hip_flow_start(S->b, i, cfg->flow_scheme, cfg->flow_dt, cfg->flow_t_max, 1 << 30)))
        return rc;
      c->flowing = 1;
      S->flowing++;
      continue;
    }
    complete(S, i);
  }
  return SU2HIP_OK;
}

/* complete the requests whose flow has finished (wait: block on them) */
static int collect_flows(server *S, int wait) {
  int rc;
  for (int i = 0; i < S->cfg->max_slots && S->flowing; i++) {
    chain_slot *c = &S->cs[i];
    if (!c->used || !c->flowing) continue;
    int done = 1;
    if (!wait && (rc = su2hip_flow_poll(S->b, i, &done))) return rc;
    if (!done) continue;
    double rows[2 * (1 + SU2HIP_NSUMS)];
    int nr = 0;
    if ((rc = su2hip_flow_finish(S->b, i, rows, 2, &nr))) return rc;
    su2hip_result *o = &S->res[c->req];
    const double *last = rows + (size_t)(nr > 1 ? 1 : 0) * (1 + SU2HIP_NSUMS);
    o->flow_t = last[0];
    memcpy(o->flow_sums, last + 1, sizeof(o->flow_sums));
    S->st.flows++;
    c->flowing = 0;
    S->flowing--;
    complete(S, i);
  }
  return SU2HIP_OK;
}

static int cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* nearest-rank percentile of sorted values */
static double pct(const double *v, int n, double q) {
  if (n <= 0) return 0.0;
  int k = (int)ceil(q * n) - 1;
  if (k < 0) k = 0;
  if (k >= n) k = n - 1;
  return v[k];
}

int su2hip_serve(const su2hip_serve_config *cfg, const su2hip_request *req, int n,
                 su2hip_result *res, su2hip_serve_stats *stats) {
  if (!cfg || n < 0 || (n > 0 && (!req || !res))) return sfail(SU2HIP_EINVAL, "serve: NULL argument");
  if (cfg->max_slots < 1 || cfg->tick_sweeps < 1) return sfail(SU2HIP_EINVAL, "serve: max_slots/tick_sweeps");
  for (int i = 1; i < n; i++)
    if (req[i].arrival < req[i - 1].arrival) return sfail(SU2HIP_EINVAL, "serve: requests out of arrival order");
  server S;
  memset(&S, 0, sizeof(S));
  S.cfg = cfg;
  S.req = req;
  S.res = res;
  S.upd = su2hip_serve_updates_per_sweep(&cfg->tmpl);
  int ms = cfg->max_slots;
  S.cs = calloc((size_t)ms, sizeof(chain_slot));
  int64_t *budgets = malloc(sizeof(int64_t) * (size_t)ms);
  double *sums = malloc(sizeof(double) * SU2HIP_NSUMS * (size_t)ms);
  int *queue = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!S.cs || !budgets || !sums || !queue) {
    free(S.cs);
    free(budgets);
    free(sums);
    free(queue);
    return sfail(SU2HIP_ENOMEM, "serve: host allocation");
  }
  for (int i = 0; i < ms; i++) S.cs[i].req = -1;
  for (int i = 0; i < n; i++) {
    memset(&res[i], 0, sizeof(res[i]));
    res[i].id = req[i].id;
    res[i].slot = -1;
    res[i].status = SU2HIP_ESTATE; /* until it completes */
  }
  int rc = su2hip_create(0, &cfg->tmpl, ms, cfg->hbm_budget, &S.b);
  if (!rc) rc = su2hip_use_graph(S.b, cfg->use_graph);
  int qh = 0, qt = 0, next = 0;
  S.t0 = now_s();
  while (!rc) {
    if (S.flowing && (rc = collect_flows(&S, 0))) break;
    double t = elapsed(&S);
    while (next < n && (!cfg->realtime || req[next].arrival <= t)) {
      res[next].t_arrival = cfg->realtime ? req[next].arrival : 0.0;
      queue[qt++] = next++;
    }
    while (qh < qt) {
      int a = admit(&S, queue[qh]);
      if (a < 0) {
        rc = a;
        break;
      }
      if (a == 0) break;
      qh++;
      while (qh < qt && queue[qh] < 0) qh++; /* admitted earlier by backfill */
    }
    if (rc) break;
    if (cfg->backfill > 0 && qh < qt &&
        elapsed(&S) - res[queue[qh]].t_arrival < cfg->max_head_wait) {
      /* the head waits for room: later new chains that fit now go ahead
       * (a continuation keeps the order of its chain) */
      for (int i = qh + 1, looked = 0; i < qt && looked < cfg->backfill; i++) {
        if (queue[i] < 0 || req[queue[i]].chain >= 0) continue;
        looked++;
        int a = admit(&S, queue[i]);
        if (a < 0) {
          rc = a;
          break;
        }
        if (a == 1) queue[i] = -1;
      }
      if (rc) break;
    }
    if (S.running == 0 && S.flowing) {
      /* only flows in flight: wait for the next one, or for an arrival */
      if (next >= n && qh == qt) {
        if ((rc = collect_flows(&S, 1))) break;
      } else {
        sleep_s(0.0005);
      }
      continue;
    }
    if (S.running == 0) {
      if (qh < qt) {
        /* nothing runs, nothing more can be evicted, and the head still
         * does not fit: it never will */
        fail_request(&S, queue[qh], sfail(SU2HIP_ENOMEM, "request %lld does not fit",
                                          (long long)req[queue[qh]].id));
        qh++;
        continue;
      }
      if (next >= n) break;
      double wait = req[next].arrival - elapsed(&S);
      sleep_s(wait < 0.05 ? wait : 0.05);
      continue;
    }
    if ((rc = su2hip_get_budgets(S.b, budgets))) break;
    int64_t k = cfg->tick_sweeps;
    for (int i = 0; i < ms; i++)
      if (S.cs[i].used && S.cs[i].req >= 0 && budgets[i] > 0 && budgets[i] < k) k = budgets[i];
    double ta = now_s();
    if ((rc = su2hip_update(S.b, (int)k))) break;
    if ((rc = su2hip_get_budgets(S.b, budgets))) break; /* waits for the stream */
    S.st.busy += now_s() - ta;
    S.st.ticks++;
    S.tick++;
    if ((rc = finish(&S, budgets, sums))) break;
  }
  double end = elapsed(&S);
  if (S.b) su2hip_destroy(S.b);
  /* statistics over completed requests */
  double *lat = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  int nl = 0;
  double sum = 0.0, maxw = 0.0, last = 0.0;
  for (int i = 0; i < n && lat; i++) {
    if (res[i].status != 0) continue;
    double l = res[i].t_done - res[i].t_arrival;
    lat[nl++] = l;
    sum += l;
    if (res[i].t_admit - res[i].t_arrival > maxw) maxw = res[i].t_admit - res[i].t_arrival;
    if (res[i].t_done > last) last = res[i].t_done;
  }
  if (lat) qsort(lat, (size_t)nl, sizeof(double), cmp_double);
  S.st.p50 = pct(lat, nl, 0.50);
  S.st.p99 = pct(lat, nl, 0.99);
  S.st.mean = nl ? sum / nl : 0.0;
  S.st.max_wait = maxw;
  S.st.makespan = last > 0 ? last : end;
  S.st.site_updates_per_s = S.st.makespan > 0 ? S.st.site_updates / S.st.makespan : 0.0;
  if (stats) *stats = S.st;
  free(lat);
  free(S.cs);
  free(budgets);
  free(sums);
  free(queue);
  return rc;
}
