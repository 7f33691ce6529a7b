/* su2hip_serve.h -- request loop over the su2hip ABI: scan points (extent,
 * couplings, sweep count) arrive over time, are admitted by HBM footprint
 * into one su2hip batch, are swept together (continuous batching: lattices
 * join and leave between scheduling ticks while the others keep sweeping),
 * and leave with their FP64 measure sums and, on request, a gradient-flow
 * measurement of the final configuration.  Chain state never leaves the
 * device between ticks; with keep_chains it stays resident after a request
 * ends so that a later request naming the same chain continues it.
 *
 * The reference has no serving layer: its driver runs one chain per process
 * (src/main.c:251-379, iterations of update_lattice then measure).  A
 * request's result is the state of that loop after `sweeps` iterations from
 * a cold start (src/init.c:92-112), measured once at the end
 * (src/measure.c:71-309), and grad_flow (src/gradflow.c:46-173) with the
 * RK3 integrator on the final configuration.
 *
 * Every function returns 0 or a negative SU2HIP_E* code (su2hip.h) and
 * sets su2hip_last_error().
 */
#ifndef SU2HIP_SERVE_H
#define SU2HIP_SERVE_H

#include <stdint.h>

#include "su2hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One scan point.  Extents are cubic (L^3, L even). */
typedef struct su2hip_request {
  int64_t id;          /* caller's tag, copied to the result */
  int64_t chain;       /* -1: new chain; else the id of an earlier request whose
                          chain (still resident) this request continues */
  int64_t sweeps;      /* compound sweeps (update_lattice calls) to run */
  double arrival;      /* seconds after the start of the stream */
  double betasu2, msq_phi, lambda_phi;
  uint64_t seed;       /* Philox key of a new chain */
  int32_t L;           /* cubic extent of a new chain */
  int32_t flow;        /* 1: gradient flow of the final configuration */
} su2hip_request;

/* What a request returns. */
typedef struct su2hip_result {
  int64_t id;
  int32_t status;      /* 0 ok, else the SU2HIP_E* code that failed it */
  int32_t slot;        /* batch slot it ran in */
  double t_arrival, t_admit, t_done; /* seconds after the start of the stream */
  double sums[SU2HIP_NSUMS];         /* su2hip_measure sums after the sweeps */
  double flow_t;                     /* flow time of flow_sums (0 without flow) */
  double flow_sums[SU2HIP_NSUMS];    /* measure sums at flow time t_max */
  uint64_t pass;                     /* Philox pass counter at the end */
  int64_t sweeps_done;
} su2hip_result;

/* Server settings.  tmpl fixes what every chain of the batch shares: field
 * content, precision, algorithms and per-sweep update counts (extents,
 * couplings and seeds come from the requests). */
typedef struct su2hip_serve_config {
  su2hip_params tmpl;
  uint64_t hbm_budget;   /* admission cap on the device bytes of resident chains
                            (fields + flow scratch of flow requests); 0 = none */
  int32_t max_slots;     /* lattices resident at once */
  int32_t tick_sweeps;   /* most sweeps between two scheduling decisions */
  int32_t use_graph;     /* replay each compound sweep from a hipGraph */
  int32_t realtime;      /* 1: requests become visible at their arrival time on
                            the wall clock; 0: all are queued at t = 0 */
  int32_t keep_chains;   /* chains stay resident (idle) after their request,
                            evicted oldest first when room is needed */
  int32_t flow_scheme;   /* SU2HIP_FLOW_RK3 or SU2HIP_FLOW_EULER */
  double flow_dt, flow_t_max;
  double phi0, sigma0, singlet0; /* cold start of new chains (src/init.c:92-112) */
  int32_t backfill;      /* while the queue's head waits for room, admit up to this
                            many later new-chain requests that fit now (0: FIFO) */
  int32_t pad;
  double max_head_wait;  /* seconds: once the head has waited this long, nothing
                            more goes ahead of it (room drains to the head) */
} su2hip_serve_config;

/* Totals of one su2hip_serve() call. */
typedef struct su2hip_serve_stats {
  int64_t completed, failed;
  double makespan;         /* seconds from the start to the last result */
  double p50, p99, mean;   /* time-to-result t_done - t_arrival, seconds */
  double max_wait;         /* longest t_admit - t_arrival */
  double site_updates;     /* variable updates done (20 per site per sweep for
                              the doublet mix, see su2hip_serve_updates_per_sweep) */
  double site_updates_per_s; /* site_updates / makespan */
  double busy;             /* seconds with at least one chain sweeping */
  int64_t ticks, peak_resident;
  uint64_t peak_bytes;     /* largest admitted footprint */
  int64_t flows, evictions;
} su2hip_serve_stats;

/* Defaults: FP32 SU(2)+doublet with packed links, heatbath + 4 link overrelaxation rounds,
 * 4 doublet overrelaxation rounds + the final Metropolis round, 64 slots,
 * 8 GiB cap, tick 16, graphs on, realtime, RK3 flow dt 0.01 to t 4.0,
 * backfill 32 with max_head_wait 1 s. */
void su2hip_serve_defaults(su2hip_serve_config *cfg);

/* Variable updates per site in one compound sweep of p's schedule. */
double su2hip_serve_updates_per_sweep(const su2hip_params *p);

/* Synthetic request stream: Poisson arrivals of the given rate (seeded),
 * L uniform in {24,32,40,48,56,64}, betasu2 in [6,12], msq in [-0.3,-0.1],
 * lambda in [0.005,0.03], sweeps in [100,1000]; every flow_every-th request
 * (1-based; 0 = never) is 64^3 with flow = 1.  Deterministic in seed. */
int su2hip_serve_stream(uint64_t seed, double rate, int n, int flow_every,
                        su2hip_request *out);

/* Run the requests to completion.  Requests must be in arrival order.
 * res has n entries (in the order of req).  stats may be NULL. */
int su2hip_serve(const su2hip_serve_config *cfg, const su2hip_request *req,
                 int n, su2hip_result *res, su2hip_serve_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* SU2HIP_SERVE_H */
