/* asan_io_harness.c -- TEST INFRASTRUCTURE.  Drives the host readers of
 * libsu2hip's untrusted inputs (config, lattice checkpoint, weight,
 * weight_params, realtime_config, MPI slicing) under AddressSanitizer and
 * UndefinedBehaviorSanitizer, with no GPU: tests/test_host_asan.py builds it
 * with gcc from su2higgs_amd/host/su2hip_io.c and su2hip_traj.c and feeds it
 * truncated and corrupted files.
 *
 * Each stdin line is "<kind> <path> [arg]"; the harness prints the return
 * code of each call and exits 0 unless a sanitizer stops it.
 *   config <path> <model>     su2hip_config_read (+ su2hip_config_reread)
 *   lattice <path> <model>    su2hip_lattice_read for extents 4 x 6 x 8
 *   weight <path> <bins>      su2hip_weight_load (READONLY and FAST)
 *   wparams <path> <bins>     su2hip_weight_params_load
 *   realtime <path>           su2hip_realtime_config_read
 *   slices <size>             su2hip_mpi_slices + su2hip_mpi_site_order, 4 x 6 x 8
 *   mklattice <path> <model>  writes a valid checkpoint (input for mutation)
 *   mkweight <path> <bins>    writes a valid weight file
 *
 * The GPU entry points su2hip_traj.c links against are stand-ins that fail:
 * only its file reader runs here. */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "su2hip.h"
#include "su2hip_io.h"
#include "su2hip_traj.h"

/* ---- stand-ins for the device half of the library ---------------------- */
static char g_msg[512];
int su2hip_internal_fail(int code, const char *msg) {
  snprintf(g_msg, sizeof(g_msg), "%s", msg ? msg : "");
  return code;
}
const char *su2hip_last_error(void) { return g_msg; }
int su2hip_measure_columns(const su2hip_params *p, const double *sums, double weight,
                           double muca_param, double *cols, int *ncols) {
  (void)p, (void)sums, (void)weight, (void)muca_param, (void)cols;
  *ncols = 0;
  return SU2HIP_OK;
}
#define NOGPU(...) { return SU2HIP_EUNSUPPORTED; }
int su2hip_get_budgets(su2hip_batch_t *b, int64_t *x) NOGPU()
int su2hip_get_fields(su2hip_batch_t *b, int s, double *a, double *c, double *d, double *e,
                      double *f) NOGPU()
int su2hip_set_fields(su2hip_batch_t *b, int s, const double *a, const double *c,
                      const double *d, const double *e, const double *f) NOGPU()
int su2hip_get_random_order(su2hip_batch_t *b, int *e) NOGPU()
int su2hip_set_random_order(su2hip_batch_t *b, int e) NOGPU()
int su2hip_muca_acceptance(su2hip_batch_t *b, int s, int e) NOGPU()
int su2hip_muca_get(su2hip_batch_t *b, int s, su2hip_muca *m, double *p, double *w,
                    int32_t *h) NOGPU()
int su2hip_muca_orderparam(su2hip_batch_t *b, int s, int r, double *pv, double *v) NOGPU()
int su2hip_muca_set_orderparam(su2hip_batch_t *b, int s, const double *pv) NOGPU()
int su2hip_set_budget(su2hip_batch_t *b, int s, int64_t n) NOGPU()
int su2hip_slot_params(su2hip_batch_t *b, int s, su2hip_params *p) NOGPU()
int su2hip_slots(su2hip_batch_t *b, int *m, int *a) NOGPU()
int su2hip_update(su2hip_batch_t *b, int n) NOGPU()

/* ---- fixed lattice of the checkpoint cases ----------------------------- */
static void model_params(const char *model, su2hip_params *p) {
  memset(p, 0, sizeof(*p));
  p->L[0] = 4;
  p->L[1] = 6;
  p->L[2] = 8;
  p->precision = SU2HIP_FP64;
  p->nhiggs = strstr(model, "2hdm") ? 2 : (strstr(model, "higgs") ? 1 : 0);
  p->u1 = strstr(model, "u1") != NULL;
  p->triplet = strstr(model, "triplet") != NULL;
  p->singlet = strstr(model, "singlet") != NULL;
}

typedef struct {
  double *link, *u1, *db, *tr, *sg;
} fields;

static fields alloc_fields(const su2hip_params *p) {
  long V = (long)p->L[0] * p->L[1] * p->L[2];
  fields f = {0};
  f.link = calloc((size_t)V * 12, sizeof(double));
  if (p->u1) f.u1 = calloc((size_t)V * 3, sizeof(double));
  if (p->nhiggs) f.db = calloc((size_t)V * 4 * p->nhiggs, sizeof(double));
  if (p->triplet) f.tr = calloc((size_t)V * 3, sizeof(double));
  if (p->singlet) f.sg = calloc((size_t)V, sizeof(double));
  for (long i = 0; i < V; i++) {
    f.link[i * 12] = f.link[i * 12 + 4] = f.link[i * 12 + 8] = 1.0;
    if (f.db) f.db[i * 4] = 0.5 + 1e-3 * i;
  }
  return f;
}

static void free_fields(fields *f) {
  free(f->link);
  free(f->u1);
  free(f->db);
  free(f->tr);
  free(f->sg);
}

static void muca_for(su2hip_muca *m, int bins, int mode) {
  memset(m, 0, sizeof(*m));
  m->enabled = 1;
  m->orderparam = SU2HIP_PHISQ;
  m->mode = mode;
  m->checks_per_sweep = 1;
  m->bins = bins;
  m->min = 0.0;
  m->max = 2.0;
  m->wrk_min = 0.0;
  m->wrk_max = 2.0;
  m->delta = 0.3;
}

int main(void) {
  char line[4096];
  while (fgets(line, sizeof(line), stdin)) {
    char kind[32] = "", path[2048] = "", arg[256] = "";
    if (sscanf(line, "%31s %2047s %255s", kind, path, arg) < 2) continue;
    int rc = 0;
    if (!strcmp(kind, "config")) {
      su2hip_run_config *c = malloc(sizeof(*c));
      rc = su2hip_config_read(path, arg[0] ? arg : NULL, c);
      if (rc == 0) rc = su2hip_config_reread(path, c);
      free(c);
    } else if (!strcmp(kind, "lattice") || !strcmp(kind, "mklattice")) {
      su2hip_params p;
      model_params(arg, &p);
      fields f = alloc_fields(&p);
      int64_t it = 0;
      double tt = 0, ct = 0;
      if (kind[0] == 'm')
        rc = su2hip_lattice_write(path, &p, 7, 1.5, 0.25, f.link, f.u1, f.db, f.tr, f.sg);
      else
        rc = su2hip_lattice_read(path, &p, &it, &tt, &ct, f.link, f.u1, f.db, f.tr, f.sg);
      free_fields(&f);
    } else if (!strcmp(kind, "weight") || !strcmp(kind, "mkweight") || !strcmp(kind, "wparams")) {
      int bins = atoi(arg);
      if (bins < 2 || bins > 100000) bins = 10;
      double *pos = calloc((size_t)bins + 2, sizeof(double));
      double *W = calloc((size_t)bins + 2, sizeof(double));
      int64_t *g = calloc((size_t)bins + 2, sizeof(int64_t));
      int64_t *n = calloc((size_t)bins + 2, sizeof(int64_t));
      su2hip_muca m;
      if (!strcmp(kind, "mkweight")) {
        muca_for(&m, bins, SU2HIP_FAST); /* save_weight writes nothing for READONLY */
        for (int i = 0; i < bins + 2; i++) {
          pos[i] = 2.0 * (i - 1) / bins; /* pos[1] = min, pos[bins + 1] = max */
          W[i] = 0.1 * i;
        }
        rc = su2hip_weight_save(path, &m, pos, W);
      } else if (!strcmp(kind, "weight")) {
        for (int mode = SU2HIP_READONLY; mode <= SU2HIP_FAST; mode++) {
          muca_for(&m, bins, mode);
          int r = su2hip_weight_load(path, &m, pos, W, bins);
          rc = rc ? rc : r;
        }
      } else {
        int found = 0;
        rc = su2hip_weight_params_load(path, bins, g, n, &found);
      }
      free(pos);
      free(W);
      free(g);
      free(n);
    } else if (!strcmp(kind, "realtime")) {
      su2hip_trajectory_config t;
      rc = su2hip_realtime_config_read(path, &t);
    } else if (!strcmp(kind, "slices")) {
      int32_t L[3] = {4, 6, 8}, ns[3], sl[3];
      int size = atoi(path);
      rc = su2hip_mpi_slices(L, size, ns, sl);
      int64_t *order = malloc(sizeof(int64_t) * 4 * 6 * 8);
      int r = su2hip_mpi_site_order(L, size, order);
      rc = rc ? rc : r;
      free(order);
    } else {
      fprintf(stderr, "unknown kind %s\n", kind);
      return 2;
    }
    printf("%s %s %d\n", kind, path, rc);
  }
  return 0;
}
