/* su2hip_traj.c -- heatbath trajectories (the reference's HB_TRAJECTORY
 * build, src/hb_trajectory.c) on top of the sweep ABI: the realtime_config
 * reader, the trajectory file and the trajectory loop. */
#include "su2hip_traj.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include "su2hip_io.h"

/* the library's error slot (su2hip_last_error) */
extern int su2hip_internal_fail(int code, const char *msg);

static int terr(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return su2hip_internal_fail(code, buf);
}

/* read_realtime_config (src/hb_trajectory.c:99-163) */
int su2hip_realtime_config_read(const char *path, su2hip_trajectory_config *out) {
  if (!path || !out) return terr(SU2HIP_EINVAL, "NULL argument");
  FILE *f = fopen(path, "r");
  if (!f) return terr(SU2HIP_EIO, "Could not open realtime file \"%s\"! Exiting...", path);
  su2hip_trajectory_config t;
  memset(&t, 0, sizeof(t));
  int set[5] = {0, 0, 0, 0, 0};
  char total[200], key[100], value[100];
  while (!feof(f)) {
    if (fgets(total, 198, f) == NULL) break;
    int ret = sscanf(total, "%99s%99s", key, value);
    if (ret == EOF) continue;
    if (key[0] == '#') continue;
    if (ret < 2) value[0] = 0;
    if (!strcasecmp(key, "n_traj")) {
      t.n_traj = (int)strtol(value, NULL, 10);
      set[0] = 1;
    } else if (!strcasecmp(key, "interval")) {
      t.interval = (int)strtol(value, NULL, 10);
      set[1] = 1;
    } else if (!strcasecmp(key, "mode_interval")) {
      t.mode_interval = (int)strtol(value, NULL, 10);
      set[2] = 1;
    } else if (!strcasecmp(key, "min")) {
      t.min = strtod(value, NULL);
      set[3] = 1;
    } else if (!strcasecmp(key, "max")) {
      t.max = strtod(value, NULL);
      set[4] = 1;
    }
  }
  fclose(f);
  /* the order of the reference's checks (src/hb_trajectory.c:156-160) */
  static const char *names[5] = {"n_traj", "interval", "mode_interval", "min", "max"};
  static const int order[5] = {1, 2, 3, 4, 0};
  for (int k = 0; k < 5; k++)
    if (!set[order[k]]) return terr(SU2HIP_EIO, "%s: %s not set", path, names[order[k]]);
  if (t.interval < 1 || t.mode_interval < 1)
    return terr(SU2HIP_EIO, "%s: interval and mode_interval must be positive", path);
  *out = t;
  return SU2HIP_OK;
}

int su2hip_trajectory_header(const char *path, const su2hip_trajectory_config *t) {
  if (!path || !t) return terr(SU2HIP_EINVAL, "NULL argument");
  FILE *f = fopen(path, "a");
  if (!f) return terr(SU2HIP_EIO, "cannot append to %s", path);
  fprintf(f, "===== Realtime trajectories: min=%g, max=%g, n_traj=%d ===== \n", t->min, t->max,
          t->n_traj);
  fclose(f);
  return SU2HIP_OK;
}

#define TRY(x)             \
  do {                     \
    int rc_ = (x);         \
    if (rc_) {             \
      rc = rc_;            \
      goto done;           \
    }                      \
  } while (0)

/* make_realtime_trajectories (src/hb_trajectory.c:11-95) */
int su2hip_trajectories(su2hip_batch_t *b, int slot, const su2hip_trajectory_config *t, int id,
                        const char *path) {
  if (!b || !t || !path) return terr(SU2HIP_EINVAL, "NULL argument");
  if (t->interval < 1) return terr(SU2HIP_EINVAL, "trajectory interval %d", t->interval);
  su2hip_muca m;
  int rc = su2hip_muca_get(b, slot, &m, NULL, NULL, NULL);
  if (rc) return rc;
  if (!m.enabled) return terr(SU2HIP_ESTATE, "heatbath trajectories need a multicanonical weight");
  int max_slots = 0, nact = 0, was_random = 0;
  rc = su2hip_slots(b, &max_slots, &nact);
  if (rc) return rc;
  rc = su2hip_get_random_order(b, &was_random);
  if (rc) return rc;
  /* the lattice's configuration (f, which every trajectory starts from) */
  su2hip_params p;
  rc = su2hip_slot_params(b, slot, &p);
  if (rc) return rc;
  long V = (long)p.L[0] * p.L[1] * p.L[2];
  double *link = malloc(sizeof(double) * V * 12);
  double *u1 = p.u1 ? malloc(sizeof(double) * V * 3) : NULL;
  double *db = p.nhiggs ? malloc(sizeof(double) * V * 4 * p.nhiggs) : NULL;
  double *tr = p.triplet ? malloc(sizeof(double) * V * 3) : NULL;
  double *sg = p.singlet ? malloc(sizeof(double) * V) : NULL;
  int64_t *budgets = malloc(sizeof(int64_t) * (max_slots > 0 ? max_slots : 1));
  FILE *f = NULL;
  int budgets_taken = 0;
  if (!link || (p.u1 && !u1) || (p.nhiggs && !db) || (p.triplet && !tr) || (p.singlet && !sg) ||
      !budgets) {
    rc = terr(SU2HIP_ENOMEM, "host allocation for trajectories");
    goto done;
  }
  TRY(su2hip_get_fields(b, slot, link, u1, db, tr, sg));
  /* only this lattice sweeps meanwhile */
  TRY(su2hip_get_budgets(b, budgets));
  budgets_taken = 1;
  for (int i = 0; i < max_slots; i++)
    if (budgets[i] != -2) TRY(su2hip_set_budget(b, i, i == slot ? -1 : 0));
  /* "heatbath trajectory mode": no multicanonical weighting, random order */
  TRY(su2hip_muca_acceptance(b, slot, 0));
  TRY(su2hip_set_random_order(b, 1));
  double pv[2] = {m.param_value[0], m.param_value[1]};
  double val = pv[0] + pv[1];
  f = fopen(path, "a");
  if (!f) {
    rc = terr(SU2HIP_EIO, "cannot append to %s", path);
    goto done;
  }
  fprintf(f, "\n----- Begin set %d, start value %lf ----- \n", id, val);
  fprintf(f, "\n--- Trajectory %d --- \n", 1);
  int iter = 1, cur = 1;
  while (cur <= t->n_traj) {
    if (iter % t->interval == 0 || iter == 1) {
      /* calc_orderparam of EVEN then ODD on the trajectory's fields */
      TRY(su2hip_muca_orderparam(b, slot, 1, pv, &val));
      fprintf(f, "%d %.8lf\n", iter - 1, val);
      if (val > t->max || val < t->min) {
        cur++;
        if (cur <= t->n_traj) {
          TRY(su2hip_set_fields(b, slot, link, u1, db, tr, sg));
          fflush(f);
          fprintf(f, "\n--- Trajectory %d --- \n", cur);
          iter = 1;
          continue;
        }
      }
    }
    TRY(su2hip_update(b, 1));
    if (iter % 1000 == 0) fflush(f);
    iter++;
  }
  /* back to f, with the order parameter the reference leaves behind (the
   * trajectory's last values) */
  TRY(su2hip_set_fields(b, slot, link, u1, db, tr, sg));
  TRY(su2hip_muca_set_orderparam(b, slot, pv));
done:
  if (f) fclose(f);
  {
    int r2 = su2hip_set_random_order(b, was_random);
    if (!rc) rc = r2;
    r2 = su2hip_muca_acceptance(b, slot, m.do_acceptance);
    if (!rc) rc = r2;
    if (budgets_taken)
      for (int i = 0; i < max_slots; i++)
        if (budgets[i] != -2) {
          r2 = su2hip_set_budget(b, i, budgets[i]);
          if (!rc) rc = r2;
        }
  }
  free(link);
  free(u1);
  free(db);
  free(tr);
  free(sg);
  free(budgets);
  return rc;
}
