/* su2hip_main.c -- the `su2hip <config>` driver: the reference's main()
 * (src/main.c:10-421) over the C ABI of libsu2hip.so.  Reads the same
 * config, writes the same labels / measure / lattice / weight /
 * measure_flow files; the sweeps run on the GPU.
 *
 *   su2hip [--model higgs+u1] [--fp32] [--no-graph] <config>
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "../../include/su2hip.h"
#include "../../include/su2hip_io.h"
#include "../../include/su2hip_traj.h"

#define MAXBINS 100000

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void die(const char *what, int rc) {
  fprintf(stderr, "su2hip: %s failed: %s (%s)\n", what, su2hip_error_string(rc), su2hip_last_error());
  exit(1);
}

#define TRY(what, x)          \
  do {                        \
    int rc_ = (x);            \
    if (rc_) die(what, rc_);  \
  } while (0)

typedef struct {
  double *link, *u1, *db, *tr, *sg;
} hostfields;

static void alloc_fields(const su2hip_params *p, hostfields *h) {
  long V = (long)p->L[0] * p->L[1] * p->L[2];
  h->link = calloc(V * 12, sizeof(double));
  h->u1 = p->u1 ? calloc(V * 3, sizeof(double)) : NULL;
  h->db = p->nhiggs ? calloc(V * 4 * p->nhiggs, sizeof(double)) : NULL;
  h->tr = p->triplet ? calloc(V * 3, sizeof(double)) : NULL;
  h->sg = p->singlet ? calloc(V, sizeof(double)) : NULL;
  if (!h->link || (p->u1 && !h->u1) || (p->nhiggs && !h->db) || (p->triplet && !h->tr) ||
      (p->singlet && !h->sg)) {
    fprintf(stderr, "su2hip: host allocation failed\n");
    exit(1);
  }
}

/* save_lattice (src/checkpoint.c:84-122) plus the Philox position in
 * <latticefile>.rng for an exact resume */
static void checkpoint(su2hip_batch_t *b, const su2hip_run_config *cfg, hostfields *h, long iter,
                       double total) {
  TRY("get_fields", su2hip_get_fields(b, 0, h->link, h->u1, h->db, h->tr, h->sg));
  TRY("lattice_write", su2hip_lattice_write(cfg->latticefile, &cfg->p, iter, total, 0.0, h->link,
                                            h->u1, h->db, h->tr, h->sg));
  uint64_t seed, pass;
  TRY("rng_get", su2hip_rng_get(b, 0, &seed, &pass));
  char rp[SU2HIP_PATHLEN + 8];
  snprintf(rp, sizeof(rp), "%s.rng", cfg->latticefile);
  FILE *f = fopen(rp, "w");
  if (f) {
    fprintf(f, "%llu %llu\n", (unsigned long long)seed, (unsigned long long)pass);
    fclose(f);
  }
  printf("Wrote fields to %s.\n", cfg->latticefile);
}

/* gsum/nsum of the SLOW recursion, kept on the host between muca_set calls */
static int64_t *g_gsum, *g_nsum;

static void save_weight(su2hip_batch_t *b, const su2hip_run_config *cfg, double *pos, double *W) {
  if (!cfg->multicanonical) return;
  su2hip_muca m;
  TRY("muca_get", su2hip_muca_get(b, 0, &m, pos, W, NULL));
  TRY("weight_save", su2hip_weight_save(cfg->weightfile, &m, pos, W));
  if (m.mode == SU2HIP_SLOW) {
    TRY("muca_get_sums", su2hip_muca_get_sums(b, 0, g_gsum, g_nsum));
    TRY("weight_params_save", su2hip_weight_params_save("weight_params", &m, pos, g_gsum, g_nsum));
  }
}

/* install the weight; SLOW runs keep their gsum/nsum across re-installs */
static void muca_install(su2hip_batch_t *b, const su2hip_muca *m, const double *pos, const double *W) {
  TRY("muca_set", su2hip_muca_set(b, 0, m, pos, W));
  if (g_gsum) TRY("muca_set_sums", su2hip_muca_set_sums(b, 0, g_gsum, g_nsum));
}

/* weight and order parameter columns of a measure line (src/measure.c:180-188) */
static void muca_columns(su2hip_batch_t *b, const su2hip_run_config *cfg, double *weight,
                         double *param) {
  double w = 0.0, r = 0.0;
  if (cfg->multicanonical) {
    su2hip_muca m;
    TRY("muca_get", su2hip_muca_get(b, 0, &m, NULL, NULL, NULL));
    r = m.param_value[0] + m.param_value[1];
    TRY("muca_weight_of", su2hip_muca_weight_of(b, 0, r, &w));
  }
  *weight = -1.0 * w;
  *param = r;
}

/* grad_flow (src/gradflow.c:46-173): rows appended to measure_flow */
static void flow(su2hip_batch_t *b, const su2hip_run_config *cfg, int flow_id) {
  int maxrows = (int)(cfg->flow_t_max / cfg->flow_dt) / (cfg->flow_meas_interval > 0 ? cfg->flow_meas_interval : 1) + 8;
  double *rows = malloc(sizeof(double) * (size_t)maxrows * (1 + SU2HIP_NSUMS));
  int n = 0;
  TRY("flow", su2hip_flow(b, 0, cfg->flow_scheme, cfg->flow_dt, cfg->flow_t_max, cfg->flow_meas_interval,
                          rows, maxrows, &n));
  double weight, param;
  muca_columns(b, cfg, &weight, &param);
  FILE *f = fopen("measure_flow", "a");
  if (!f) {
    fprintf(stderr, "su2hip: cannot open measure_flow\n");
    exit(1);
  }
  fprintf(f, "\n =========== Flow id: %d ===========\n", flow_id);
  /* grad_flow's time loop, to know which row is the unaligned last one */
  double t = 0.0, dt = cfg->flow_dt;
  int iter = 1, r = 0;
  char line[4096];
  TRY("measure_format", su2hip_measure_format(&cfg->p, rows + 1, weight, param, line, sizeof(line)));
  fprintf(f, "%.6lf %s", rows[0], line);
  r = 1;
  while (t < cfg->flow_t_max) {
    if (t + dt > cfg->flow_t_max) dt = cfg->flow_t_max - t;
    t += dt;
    if (iter % cfg->flow_meas_interval == 0 && r < n) {
      double *row = rows + (size_t)r * (1 + SU2HIP_NSUMS);
      TRY("measure_format", su2hip_measure_format(&cfg->p, row + 1, weight, param, line, sizeof(line)));
      fprintf(f, "%.6lf %s", row[0], line);
      r++;
    }
    iter++;
  }
  if ((iter - 1) % cfg->flow_meas_interval != 0 && r < n) {
    double *row = rows + (size_t)r * (1 + SU2HIP_NSUMS);
    TRY("measure_format", su2hip_measure_format(&cfg->p, row + 1, weight, param, line, sizeof(line)));
    fprintf(f, "%g %s", row[0], line);
  }
  fclose(f);
  free(rows);
}

int main(int argc, char **argv) {
  const char *model = NULL, *path = NULL;
  int fp32 = 0, graph = 1;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "--model") && i + 1 < argc)
      model = argv[++i];
    else if (!strcmp(argv[i], "--fp32"))
      fp32 = 1;
    else if (!strcmp(argv[i], "--no-graph"))
      graph = 0;
    else
      path = argv[i];
  }
  if (!path) {
    printf("Usage: su2hip [--model higgs|2hdm+u1+triplet+singlet+gradflow+correlators+blocking+measure_z]"
           " [--fp32] [--no-graph] <config file>\n");
    return 0;
  }
  su2hip_run_config cfg;
  TRY("config", su2hip_config_read(path, model, &cfg));
  su2hip_params *p = &cfg.p;
  if (fp32) {
    p->precision = SU2HIP_FP32;
    p->reunitarize = 1;
  }
  if (p->seed == 0) {
    /* the reference seeds from the clock (src/main.c:64-68) */
    long seed = time(NULL);
    seed = seed ^ (seed << 26) ^ (seed << 9);
    p->seed = (uint64_t)seed;
  }
  printf("Seed: %llu\n", (unsigned long long)p->seed);
  printf("Volume %d x %d x %d, total %ld\n", p->L[0], p->L[1], p->L[2], (long)p->L[0] * p->L[1] * p->L[2]);
  printf("iterations %ld, measurement interval %ld, checkpoint interval %ld\n", (long)cfg.iterations,
         (long)cfg.interval, (long)cfg.checkpoint);
  printf("Will perform %d gauge sweeps (+%d overrelaxation rounds), %d scalar sweeps per iteration, %s\n",
         p->update_links, p->link_overrelax, p->scalar_sweeps, p->precision == SU2HIP_FP32 ? "FP32-mixed" : "FP64");

  su2hip_batch_t *b = NULL;
  TRY("create", su2hip_create(1, p, 1, 0, &b));
  TRY("use_graph", su2hip_use_graph(b, graph));
  hostfields h;
  alloc_fields(p, &h);
  long iter = 1;
  int64_t file_iter = 0;
  double total_time = 0.0;
  if (access(cfg.latticefile, R_OK) == 0) {
    printf("\nLoading latticefile: %s\n", cfg.latticefile);
    double tt, ct;
    TRY("lattice_read", su2hip_lattice_read(cfg.latticefile, p, &file_iter, &tt, &ct, h.link, h.u1, h.db,
                                            h.tr, h.sg));
    TRY("set_fields", su2hip_set_fields(b, 0, h.link, h.u1, h.db, h.tr, h.sg));
    char rp[SU2HIP_PATHLEN + 8];
    snprintf(rp, sizeof(rp), "%s.rng", cfg.latticefile);
    FILE *f = fopen(rp, "r");
    unsigned long long seed, pass;
    if (f && fscanf(f, "%llu %llu", &seed, &pass) == 2 && !cfg.reset) {
      TRY("rng_set", su2hip_rng_set(b, 0, seed, pass));
      printf("Resuming the Philox stream of %s\n", rp);
    }
    if (f) fclose(f);
    printf("Fields loaded succesfully.\n");
  } else {
    if (cfg.hot_start) {
      printf("No latticefile found; starting with a hot configuration.\n");
      TRY("hot_start", su2hip_hot_start(b, 0));
    } else {
      printf("No latticefile found; starting with cold configuration.\n");
      TRY("cold_start", su2hip_cold_start(b, 0, cfg.phi0, cfg.sigma0, cfg.singlet0));
    }
    cfg.reset = 1;
  }
  double *zprof = NULL;
  if (cfg.measure_z) {
    /* src/main.c:188-199: a two-phase wall overrides the start; labels_z */
    int zd = 2;
    for (int d = 2; d >= 0; d--)
      if (p->L[d] > p->L[zd]) zd = d;
    if (cfg.reset && cfg.setup_wall) {
      TRY("wall_start", su2hip_wall_start(b, 0, zd));
      printf("Wall profile initialized.\n");
    }
    if (cfg.do_z_meas) {
      TRY("zprof_labels", su2hip_zprof_labels_write("labels_z", p));
      printf("Measuring profiles along direction %d every %ld iterations\n", zd + 1, (long)cfg.interval_z);
      zprof = calloc((size_t)p->L[zd] * 4, sizeof(double));
    }
  }
  TRY("labels", su2hip_labels_write("labels", p));

  double *pos = NULL, *W = NULL;
  int mode = SU2HIP_READONLY;
  if (cfg.multicanonical) {
    pos = calloc(MAXBINS + 2, sizeof(double));
    W = calloc(MAXBINS + 2, sizeof(double));
    printf("\nLoading weightfile: %s\n", cfg.weightfile);
    TRY("weight_load", su2hip_weight_load(cfg.weightfile, &cfg.muca, pos, W, MAXBINS));
    mode = cfg.muca.mode;
    if (mode == SU2HIP_SLOW) {
      /* load_weight reads weight_params for SLOW runs and save_weight writes
       * it back at once (src/multicanonical.c:272-286) */
      g_gsum = calloc(MAXBINS + 2, sizeof(int64_t));
      g_nsum = calloc(MAXBINS + 2, sizeof(int64_t));
      int found = 0;
      printf("Using slow but safe weight updating: parameter file 'weight_params'\n");
      TRY("weight_params_load", su2hip_weight_params_load("weight_params", cfg.muca.bins, g_gsum, g_nsum, &found));
      if (!found) printf("!!! Unable to access weight parameter file. Initializing new\n");
      TRY("weight_params_save", su2hip_weight_params_save("weight_params", &cfg.muca, pos, g_gsum, g_nsum));
    }
    muca_install(b, &cfg.muca, pos, W);
    printf("Using weight function with %d bins in range %lf, %lf\n", cfg.muca.bins, cfg.muca.min, cfg.muca.max);
  }

  /* heatbath trajectories of the HB_TRAJECTORY build (src/main.c:223-238) */
  su2hip_trajectory_config traj;
  memset(&traj, 0, sizeof(traj));
  int do_traj = cfg.trajectory && cfg.do_trajectory;
  if (do_traj) {
    printf("\nReal time simulation: reading file \"realtime_config\"\n");
    TRY("realtime_config", su2hip_realtime_config_read("realtime_config", &traj));
    /* print_realtime_params (src/hb_trajectory.c:166-178) */
    printf("Heatbath trajectory mode every %d iterations, %d trajectories each\n", traj.mode_interval,
           traj.n_traj);
    printf("--- min %lf, max %lf, measure interval %d ---\n", traj.min, traj.max, traj.interval);
    if (p->algorithm_su2link != SU2HIP_HEATBATH) printf("\n--- WARNING: SU(2) update not using heatbath!\n");
    if (p->u1 && p->algorithm_u1link != SU2HIP_HEATBATH)
      printf("\n--- WARNING: U(1) update not using heatbath!\n");
    fflush(stdout);
    if (!cfg.multicanonical) {
      printf("\nError: Need multicanonical for heatbath trajectories!! Exiting...\n");
      return 1;
    }
    TRY("trajectory_header", su2hip_trajectory_header("trajectory", &traj));
  }

  if (cfg.reset) {
    /* thermalise without modifying the weight (src/main.c:253-282) */
    su2hip_muca m;
    if (cfg.multicanonical && mode != SU2HIP_READONLY) {
      TRY("muca_get", su2hip_muca_get(b, 0, &m, pos, W, NULL));
      m.mode = SU2HIP_READONLY;
      muca_install(b, &m, pos, W);
    }
    printf("\nThermalizing %ld iterations\n", (long)cfg.n_thermalize);
    double t0 = now();
    TRY("update", su2hip_update(b, (int)cfg.n_thermalize));
    TRY("sync", su2hip_sync(b));
    printf("Thermalization done, took %lf seconds.\n", now() - t0);
    TRY("reset_counters", su2hip_reset_counters(b, 0));
    if (cfg.multicanonical && mode != SU2HIP_READONLY) {
      TRY("muca_get", su2hip_muca_get(b, 0, &m, pos, W, NULL));
      m.mode = mode;
      /* init_last_max (src/multicanonical.c:826-831) */
      double val = m.param_value[0] + m.param_value[1];
      m.last_max = fabs(val - m.wrk_min) >= fabs(val - m.wrk_max);
      m.muca_count = 0;
      muca_install(b, &m, pos, W);
    }
    iter = 1;
    printf("\nStarting new simulation!\n");
  } else {
    iter = file_iter + 1;
    printf("\nContinuing from iteration %ld!\n", iter - 1);
  }

  FILE *res = fopen(cfg.resultsfile, "a");
  if (!res) {
    fprintf(stderr, "su2hip: cannot open %s\n", cfg.resultsfile);
    return 1;
  }
  double start = now();
  int flow_id = 1;
  int correlator_id = 1;
  int traj_id = 1;
  double *corr = NULL;
  if (cfg.correlators && cfg.do_correlators) {
    /* main.c:98-104: labels, the longest direction */
    TRY("correl_labels", su2hip_correl_labels_write("labels_correl", p));
    int ld = 2;
    for (int d = 2; d >= 0; d--)
      if (p->L[d] > p->L[ld]) ld = d;
    printf("\nMeasuring correlation functions along direction %d every %d iterations\n", ld + 1,
           cfg.correlator_interval);
    corr = calloc((size_t)(p->L[ld] / 2 + 1) * 10 * (size_t)(cfg.blocks > 1 ? cfg.blocks : 1),
                  sizeof(double));
  }
  char line[4096];
  while (iter <= cfg.iterations) {
    if (iter % cfg.interval == 0) {
      double sums[SU2HIP_NSUMS], weight, param;
      TRY("measure", su2hip_measure(b, 0, sums));
      muca_columns(b, &cfg, &weight, &param);
      TRY("measure_format", su2hip_measure_format(p, sums, weight, param, line, sizeof(line)));
      fputs(line, res);
      fflush(res);
    }
    if (zprof && iter % cfg.interval_z == 0) {
      /* measure_along_z (src/main.c:308-312) */
      int nz = 0;
      TRY("zprofile", su2hip_zprofile(b, 0, -1, zprof, 1 << 20, &nz));
      TRY("zprof_append", su2hip_zprof_append("measure_z", p, zprof, nz));
    }
    if (cfg.gradflow && cfg.do_flow && iter % cfg.flow_interval == 0) flow(b, &cfg, flow_id++);
    if (do_traj && iter % traj.mode_interval == 0)
      TRY("trajectories", su2hip_trajectories(b, 0, &traj, traj_id++, "trajectory"));
    if (corr && iter % cfg.correlator_interval == 0) {
      /* measure_correlators("correl0", ...) (src/main.c:330-349) */
      int nd = 0;
      TRY("correlators", su2hip_correlators(b, 0, -1, corr, 1 << 20, &nd));
      TRY("correl_append", su2hip_correl_append("correl0", p, correlator_id, corr, nd));
      if (cfg.blocking && cfg.blocks > 0) {
        /* the blocked lattices (src/main.c:334-347), files correl1, correl2, ... */
        if (p->nhiggs) {
          for (int k = 0; k < cfg.blocks; k++)
            printf("Error in correlation.c: Higgs smearing not yet implemented!!!\n");
        } else {
          int nl = 0;
          TRY("blocked_correlators", su2hip_blocked_correlators(b, 0, cfg.blocks, corr, &nl, &nd));
          for (int k = 0; k < nl; k++) {
            char fn[32];
            snprintf(fn, sizeof(fn), "correl%d", k + 1);
            TRY("correl_append", su2hip_correl_append(fn, p, correlator_id, corr + (size_t)k * nd * 10, nd));
          }
        }
      }
      correlator_id++;
    }
    TRY("update", su2hip_update(b, 1));
    if (iter % cfg.checkpoint == 0) {
      TRY("sync", su2hip_sync(b));
      total_time += now() - start;
      start = now();
      su2hip_counters c;
      TRY("counters", su2hip_get_counters(b, 0, &c));
      printf("\nCheckpointing at iteration %lu. Total time: %.1lfs, %.2lf%% comms.\n", (unsigned long)iter,
             total_time, 0.0);
      su2hip_print_acceptance(stdout, &cfg, &c);
      fflush(stdout);
      checkpoint(b, &cfg, &h, iter, total_time);
      save_weight(b, &cfg, pos, W);
      TRY("config_reread", su2hip_config_reread(path, &cfg));
    }
    iter++;
  }
  fclose(res);
  TRY("sync", su2hip_sync(b));
  total_time += now() - start;
  checkpoint(b, &cfg, &h, iter, total_time);
  save_weight(b, &cfg, pos, W);
  printf("\nReached end! Total time taken: %.1lfs, of which %.2lf%% comms. time per iteration: %.6lfs \n",
         total_time, 0.0, cfg.iterations > 0 ? total_time / cfg.iterations : 0.0);
  su2hip_destroy(b);
  return 0;
}
