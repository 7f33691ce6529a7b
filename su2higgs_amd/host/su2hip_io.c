/* su2hip_io.c -- host-side file formats (include/su2hip_io.h): the
 * reference's config reader semantics, lattice checkpoint, weight file,
 * labels and measure lines.  Each function names the reference function
 * whose behaviour it reproduces. */
#define _GNU_SOURCE
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <unistd.h>

#include "../../include/su2hip_io.h"

/* error reporting shared with the C ABI layer (su2hip_api.cpp) */
extern int su2hip_internal_fail(int code, const char *msg);

static int ioerr(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return su2hip_internal_fail(code, buf);
}

/* ------------------------------------------------------------------------
 * config file: the file as fgets() with a 200-byte buffer sees it, and a
 * read position that persists between lookups (FindFromFile,
 * src/parameters.c:30-92) */

#define MAXLINE 200

typedef struct {
  char **lines;
  int n;
  int pos;     /* next line to read */
  int at_eof;  /* the EOF indicator is set (last line had no newline) */
  int last_nl; /* the file ends with a newline */
} cfgfile;

static void cfg_free(cfgfile *f) {
  for (int i = 0; i < f->n; i++) free(f->lines[i]);
  free(f->lines);
  memset(f, 0, sizeof(*f));
}

static int cfg_open(const char *path, cfgfile *f) {
  memset(f, 0, sizeof(*f));
  if (access(path, R_OK) != 0) return ioerr(SU2HIP_EIO, "Unable to access file %s", path);
  FILE *fp = fopen(path, "r");
  if (!fp) return ioerr(SU2HIP_EIO, "Unable to open file %s", path);
  char buf[MAXLINE];
  int cap = 64;
  f->lines = (char **)malloc(sizeof(char *) * cap);
  f->last_nl = 1;
  if (!f->lines) goto oom;
  while (fgets(buf, MAXLINE, fp)) {
    if (f->n == cap) {
      char **grown = (char **)realloc(f->lines, sizeof(char *) * cap * 2);
      if (!grown) goto oom;
      f->lines = grown;
      cap *= 2;
    }
    char *line = strdup(buf);
    if (!line) goto oom;
    f->lines[f->n++] = line;
    size_t l = strlen(buf);
    f->last_nl = l > 0 && buf[l - 1] == '\n';
  }
  fclose(fp);
  return SU2HIP_OK;
oom:
  fclose(fp);
  cfg_free(f);
  return ioerr(SU2HIP_ENOMEM, "host allocation reading %s", path);
}

/* One lookup: search on from the current position to the end, then once
 * more from the top; the first line whose first token equals `label`
 * (case-insensitively, `#` lines skipped) gives its second token. */
static int cfg_find(cfgfile *f, const char *label, char *result) {
  char key[MAXLINE], value[MAXLINE];
  result[0] = 0;
  value[0] = 0;
  int wrapped = 0;
  if (!f->at_eof) {
    int i = f->pos;
    for (;;) {
      if (i >= f->n) {
        if (wrapped) break;
        wrapped = 1;
        i = 0;
        continue;
      }
      const char *line = f->lines[i++];
      int ret = sscanf(line, "%99s%99s", key, value);
      if (ret == EOF || ret < 1) continue;
      if (key[0] == '#') continue;
      if (!strcasecmp(key, label)) {
        strcpy(result, value);
        f->pos = i;
        /* fgets() of a last line without newline sets the EOF indicator */
        if (i == f->n && !f->last_nl) f->at_eof = 1;
        break;
      }
    }
    if (!result[0]) {
      f->pos = f->n;
      f->at_eof = 1;
    }
  }
  if (!result[0]) return ioerr(SU2HIP_EIO, "Unable to find parameter %s in file", label);
  return SU2HIP_OK;
}

/* optional key of this build: first match from the top, no state */
static int cfg_opt(const cfgfile *f, const char *label, char *result) {
  char key[MAXLINE], value[MAXLINE];
  for (int i = 0; i < f->n; i++) {
    value[0] = 0;
    int ret = sscanf(f->lines[i], "%99s%99s", key, value);
    if (ret == EOF || ret < 1 || key[0] == '#') continue;
    if (!strcasecmp(key, label) && value[0]) {
      strcpy(result, value);
      return 1;
    }
  }
  return 0;
}

#define GETS(f, label, dst)                \
  do {                                     \
    int rc_ = cfg_find(f, label, dst);     \
    if (rc_) {                             \
      cfg_free(f);                         \
      return rc_;                          \
    }                                      \
  } while (0)

#define GETI(f, label, dst)                \
  do {                                     \
    char v_[MAXLINE];                      \
    GETS(f, label, v_);                    \
    dst = (int)strtol(v_, NULL, 10);       \
  } while (0)

#define GETL(f, label, dst)                \
  do {                                     \
    char v_[MAXLINE];                      \
    GETS(f, label, v_);                    \
    dst = strtol(v_, NULL, 10);            \
  } while (0)

#define GETD(f, label, dst)                \
  do {                                     \
    char v_[MAXLINE];                      \
    GETS(f, label, v_);                    \
    dst = strtod(v_, NULL);                \
  } while (0)

/* GetUpdateAlgorithm (src/parameters.c:132-149) */
static int algorithm_of(const char *s, const char *label) {
  if (!strcasecmp(s, "metropolis")) return SU2HIP_METROPOLIS;
  if (!strcasecmp(s, "heatbath")) return SU2HIP_HEATBATH;
  if (!strcasecmp(s, "overrelax")) return SU2HIP_OVERRELAX;
  fprintf(stderr, "WARNING: Unknown update algorithm for '%s'; using metropolis instead.", label);
  return SU2HIP_METROPOLIS;
}

#define GETALG(f, label, dst)              \
  do {                                     \
    char v_[MAXLINE];                      \
    GETS(f, label, v_);                    \
    dst = algorithm_of(v_, label);         \
  } while (0)

static int parse_model(const char *s, su2hip_run_config *c) {
  c->p.nhiggs = c->p.u1 = c->p.triplet = c->p.singlet = 0;
  c->gradflow = 0;
  c->correlators = 0;
  c->measure_z = 0;
  c->blocking = 0;
  c->trajectory = 0;
  char buf[MAXLINE];
  snprintf(buf, sizeof(buf), "%s", s);
  for (char *tok = strtok(buf, "+,"); tok; tok = strtok(NULL, "+,")) {
    if (!strcasecmp(tok, "higgs"))
      c->p.nhiggs = c->p.nhiggs == 2 ? 2 : 1;
    else if (!strcasecmp(tok, "2hdm"))
      c->p.nhiggs = 2;
    else if (!strcasecmp(tok, "u1"))
      c->p.u1 = 1;
    else if (!strcasecmp(tok, "triplet"))
      c->p.triplet = 1;
    else if (!strcasecmp(tok, "singlet"))
      c->p.singlet = 1;
    else if (!strcasecmp(tok, "gradflow"))
      c->gradflow = 1;
    else if (!strcasecmp(tok, "correlators"))
      c->correlators = 1;
    else if (!strcasecmp(tok, "measure_z"))
      c->measure_z = 1;
    else if (!strcasecmp(tok, "blocking"))
      c->blocking = c->correlators = 1;
    else if (!strcasecmp(tok, "trajectory"))
      c->trajectory = 1;
    else if (!strcasecmp(tok, "su2") || !strcasecmp(tok, "gauge"))
      ;
    else
      return ioerr(SU2HIP_EINVAL, "unknown model token '%s'", tok);
  }
  return SU2HIP_OK;
}

/* get_parameters + get_weight_parameters (src/parameters.c:157-402), the
 * lookups in the reference's order for the chosen field content */
int su2hip_config_read(const char *path, const char *model, su2hip_run_config *out) {
  if (!path || !out) return ioerr(SU2HIP_EINVAL, "NULL argument");
  su2hip_run_config c;
  memset(&c, 0, sizeof(c));
  cfgfile f;
  int rc = cfg_open(path, &f);
  if (rc) return rc;
  char v[MAXLINE];
  /* field content: argument, else the `model` key, else the reference
   * Makefile's default build (-DNHIGGS=1 -DSINGLET -DU1) */
  const char *m = model;
  if (!m && cfg_opt(&f, "model", v)) m = v;
  rc = parse_model(m ? m : "higgs+singlet+u1", &c);
  if (rc) {
    cfg_free(&f);
    return rc;
  }
  su2hip_params *p = &c.p;
  int dim;
  GETI(&f, "dim", dim);
  if (dim != 3) {
    cfg_free(&f);
    return ioerr(SU2HIP_EUNSUPPORTED, "dim %d: only 3 dimensions are supported", dim);
  }
  for (int d = 0; d < 3; d++) {
    char lab[16];
    snprintf(lab, sizeof(lab), "L%d", d + 1);
    GETI(&f, lab, p->L[d]);
  }
  GETS(&f, "resultsfile", c.resultsfile);
  GETALG(&f, "algorithm_su2link", p->algorithm_su2link);
  if (p->algorithm_su2link != SU2HIP_METROPOLIS && p->algorithm_su2link != SU2HIP_HEATBATH) {
    fprintf(stderr, "Cannot use algorithm %d for SU2 links, changing to metropolis...", p->algorithm_su2link);
    p->algorithm_su2link = SU2HIP_METROPOLIS;
  }
  p->algorithm_u1link = SU2HIP_METROPOLIS;
  if (p->u1) {
    GETALG(&f, "algorithm_u1link", p->algorithm_u1link);
    if (p->algorithm_u1link != SU2HIP_METROPOLIS) p->algorithm_u1link = SU2HIP_METROPOLIS;
  }
  p->algorithm_su2doublet = p->algorithm_su2triplet = p->algorithm_singlet = SU2HIP_METROPOLIS;
  if (p->nhiggs) {
    GETALG(&f, "algorithm_su2doublet", p->algorithm_su2doublet);
    if (p->algorithm_su2doublet == SU2HIP_HEATBATH) p->algorithm_su2doublet = SU2HIP_METROPOLIS;
  }
  if (p->triplet) {
    GETALG(&f, "algorithm_su2triplet", p->algorithm_su2triplet);
    if (p->algorithm_su2triplet == SU2HIP_HEATBATH) p->algorithm_su2triplet = SU2HIP_METROPOLIS;
  }
  if (p->singlet) {
    GETALG(&f, "algorithm_singlet", p->algorithm_singlet);
    if (p->algorithm_singlet == SU2HIP_HEATBATH) p->algorithm_singlet = SU2HIP_METROPOLIS;
  }
  GETS(&f, "latticefile", c.latticefile);
  GETL(&f, "iterations", c.iterations);
  GETL(&f, "interval", c.interval);
  GETL(&f, "checkpoint", c.checkpoint);
  GETL(&f, "n_thermalize", c.n_thermalize);
  GETI(&f, "run_checks", c.run_checks);
  GETI(&f, "reset", c.reset);
  GETI(&f, "multicanonical", c.multicanonical);
  GETI(&f, "measure_local", c.measure_local);
  if (c.measure_z) GETL(&f, "interval_z", c.interval_z);
  GETI(&f, "update_links", p->update_links);
  GETI(&f, "update_doublet", p->update_su2doublet);
  GETI(&f, "scalar_sweeps", p->scalar_sweeps);
  GETD(&f, "betasu2", p->betasu2);
  if (p->nhiggs) {
    GETD(&f, "msq", p->msq_phi);
    GETD(&f, "lambda", p->lambda_phi);
    GETD(&f, "phi0", c.phi0);
  }
  if (p->u1) {
    GETD(&f, "betau1", p->betau1);
    GETD(&f, "r_u1", p->r_u1);
  }
  if (p->triplet) {
    GETI(&f, "update_triplet", p->update_su2triplet);
    GETD(&f, "sigma0", c.sigma0);
    GETD(&f, "msq_triplet", p->msq_triplet);
    GETD(&f, "b4", p->b4);
    if (p->nhiggs) GETD(&f, "a2", p->a2);
  }
  if (p->singlet) {
    GETI(&f, "update_singlet", p->update_singlet);
    GETD(&f, "singlet0", c.singlet0);
    GETD(&f, "b1_s", p->b1_s);
    GETD(&f, "msq_s", p->msq_s);
    GETD(&f, "b3_s", p->b3_s);
    GETD(&f, "b4_s", p->b4_s);
    if (p->nhiggs) {
      GETD(&f, "a1_s", p->a1_s);
      GETD(&f, "a2_s", p->a2_s);
    }
  }
  if (p->nhiggs == 2) {
    GETD(&f, "msq_phi2", p->msq_phi2);
    GETD(&f, "m12sq_re", p->m12sq_re);
    GETD(&f, "m12sq_im", p->m12sq_im);
    GETD(&f, "lam2", p->lam2);
    GETD(&f, "lam3", p->lam3);
    GETD(&f, "lam4", p->lam4);
    GETD(&f, "lam5_re", p->lam5_re);
    GETD(&f, "lam5_im", p->lam5_im);
    GETD(&f, "lam6_re", p->lam6_re);
    GETD(&f, "lam6_im", p->lam6_im);
    GETD(&f, "lam7_re", p->lam7_re);
    GETD(&f, "lam7_im", p->lam7_im);
  }
  if (c.correlators) {
    GETI(&f, "do_correlators", c.do_correlators);
    GETI(&f, "correlator_interval", c.correlator_interval);
  }
  if (c.blocking) GETI(&f, "blocks", c.blocks);
  if (c.trajectory) GETI(&f, "do_trajectory", c.do_trajectory);
  c.flow_scheme = SU2HIP_FLOW_EULER;
  if (c.gradflow) {
    GETI(&f, "do_flow", c.do_flow);
    GETI(&f, "flow_interval", c.flow_interval);
    GETI(&f, "flow_meas_interval", c.flow_meas_interval);
    GETD(&f, "flow_dt", c.flow_dt);
    GETD(&f, "flow_t_max", c.flow_t_max);
  }
  if (c.measure_z) {
    GETI(&f, "do_z_meas", c.do_z_meas);
    GETI(&f, "setup_wall", c.setup_wall);
  }
  /* keys of this build only (the reference ignores them) */
  p->precision = SU2HIP_FP64;
  p->seed = 0;
  p->link_overrelax = 0;
  if (cfg_opt(&f, "precision", v)) {
    if (!strcasecmp(v, "fp32") || !strcmp(v, "32"))
      p->precision = SU2HIP_FP32;
    else if (strcasecmp(v, "fp64") && strcmp(v, "64")) {
      cfg_free(&f);
      return ioerr(SU2HIP_EINVAL, "precision '%s' (fp64 or fp32)", v);
    }
  }
  p->reunitarize = p->precision == SU2HIP_FP32;
  p->link_format = SU2HIP_LINKS_AUTO;
  if (cfg_opt(&f, "link_format", v)) {
    if (!strcasecmp(v, "full"))
      p->link_format = SU2HIP_LINKS_FULL;
    else if (!strcasecmp(v, "packed"))
      p->link_format = SU2HIP_LINKS_PACKED;
    else if (strcasecmp(v, "auto")) {
      cfg_free(&f);
      return ioerr(SU2HIP_EINVAL, "link_format '%s' (auto, full or packed)", v);
    }
  }
  if (cfg_opt(&f, "link_overrelax", v)) p->link_overrelax = (int)strtol(v, NULL, 10);
  if (cfg_opt(&f, "seed", v)) p->seed = strtoull(v, NULL, 10);
  if (cfg_opt(&f, "start", v)) c.hot_start = !strcasecmp(v, "hot");
  if (cfg_opt(&f, "flow_scheme", v)) c.flow_scheme = !strcasecmp(v, "rk3") ? SU2HIP_FLOW_RK3 : SU2HIP_FLOW_EULER;
  cfg_free(&f);
  for (int d = 0; d < 3; d++)
    if (p->L[d] % 2 != 0)
      return ioerr(SU2HIP_EINVAL, "Lattice side length L%d = %d is odd; checkerboard updating needs even sides", d + 1, p->L[d]);
  /* get_weight_parameters: a fresh read of the file */
  memset(&c.muca, 0, sizeof(c.muca));
  strcpy(c.weightfile, "weight");
  if (c.multicanonical) {
    rc = cfg_open(path, &f);
    if (rc) return rc;
    GETS(&f, "weightfile", c.weightfile);
    GETI(&f, "muca_mode", c.muca.mode);
    GETI(&f, "checks_per_sweep", c.muca.checks_per_sweep);
    GETI(&f, "bins", c.muca.bins);
    GETD(&f, "min", c.muca.min);
    GETD(&f, "max", c.muca.max);
    GETD(&f, "muca_delta", c.muca.delta);
    GETS(&f, "orderparam", v);
    cfg_free(&f);
    if (!strcasecmp(v, "phisq"))
      c.muca.orderparam = SU2HIP_PHISQ;
    else if (!strcasecmp(v, "Sigmasq"))
      c.muca.orderparam = SU2HIP_SIGMASQ;
    else if (!strcasecmp(v, "phi2minusSigma2"))
      c.muca.orderparam = SU2HIP_PHI2MINUSSIGMA2;
    else if (!strcasecmp(v, "phi2sq"))
      c.muca.orderparam = SU2HIP_PHI2SQ;
    else {
      fprintf(stderr, "WARNING: Unknown multicanonical order parameter!! using phisq");
      c.muca.orderparam = SU2HIP_PHISQ;
    }
    if (c.muca.mode != SU2HIP_READONLY && c.muca.mode != SU2HIP_FAST && c.muca.mode != SU2HIP_SLOW)
      return ioerr(SU2HIP_EINVAL, "Invalid mode specification for multicanonical!! got mode=%d", c.muca.mode);
    c.muca.enabled = 1;
    c.muca.do_acceptance = 1;
  }
  *out = c;
  return SU2HIP_OK;
}

/* read_updated_parameters (src/parameters.c:460-494) */
int su2hip_config_reread(const char *path, su2hip_run_config *cfg) {
  cfgfile f;
  int rc = cfg_open(path, &f);
  if (rc) {
    fprintf(stderr, "!!! Warning: Unable to access config file %s at checkpoint; skipping read...\n", path);
    return SU2HIP_OK;
  }
  long it, iv, ck;
  GETL(&f, "iterations", it);
  GETL(&f, "interval", iv);
  GETL(&f, "checkpoint", ck);
  cfg_free(&f);
  if (cfg->iterations != it) printf("Updated max iterations to %ld\n", it);
  if (cfg->interval != iv) printf("Updated measurement interval to %ld\n", iv);
  if (cfg->checkpoint != ck) printf("Updated checkpoint interval to %ld\n", ck);
  cfg->iterations = it;
  cfg->interval = iv;
  cfg->checkpoint = ck;
  return SU2HIP_OK;
}

/* ------------------------------------------------------------------------
 * lattice checkpoint (save_lattice / load_lattice, src/checkpoint.c:84-214,
 * serial write_field/read_field :337-357) */

static long vol_of(const su2hip_params *p) { return (long)p->L[0] * p->L[1] * p->L[2]; }

int su2hip_lattice_write(const char *path, const su2hip_params *p, int64_t iter,
                         double total_time, double comms_time, const double *su2link,
                         const double *u1link, const double *doublet,
                         const double *triplet, const double *singlet) {
  if (!path || !p || !su2link) return ioerr(SU2HIP_EINVAL, "NULL argument");
  FILE *fp = fopen(path, "wb");
  if (!fp) return ioerr(SU2HIP_EIO, "cannot write %s", path);
  int size = 1, dim = 3;
  long it = (long)iter, V = vol_of(p);
  int ok = fwrite(&size, sizeof(size), 1, fp) == 1 && fwrite(&dim, sizeof(dim), 1, fp) == 1 &&
           fwrite(p->L, sizeof(int), 3, fp) == 3 && fwrite(&it, sizeof(it), 1, fp) == 1 &&
           fwrite(&total_time, sizeof(double), 1, fp) == 1 &&
           fwrite(&comms_time, sizeof(double), 1, fp) == 1;
  ok = ok && fwrite(su2link, sizeof(double), V * 12, fp) == (size_t)(V * 12);
  if (ok && p->u1) ok = u1link && fwrite(u1link, sizeof(double), V * 3, fp) == (size_t)(V * 3);
  /* doublets one after the other (src/checkpoint.c:105-107) */
  long nd = V * 4 * p->nhiggs;
  if (ok && p->nhiggs) ok = doublet && fwrite(doublet, sizeof(double), nd, fp) == (size_t)nd;
  if (ok && p->triplet) ok = triplet && fwrite(triplet, sizeof(double), V * 3, fp) == (size_t)(V * 3);
  if (ok && p->singlet) ok = singlet && fwrite(singlet, sizeof(double), V, fp) == (size_t)V;
  if (fclose(fp) != 0) ok = 0;
  if (!ok) return ioerr(SU2HIP_EIO, "short write to %s", path);
  return SU2HIP_OK;
}

/* Node grid of an MPI run (make_slices, src/layout.c:118-237): factor the
 * rank count into the primes 2..29 and, from the largest prime down, cut the
 * longest node side that the prime divides (the first such side when equal).
 * The rank count and the lattice must allow it, as the reference dies
 * otherwise (src/layout.c:122-151). */
int su2hip_mpi_slices(const int32_t *L, int size, int32_t *nslices, int32_t *slice) {
  if (!L || !nslices || !slice) return ioerr(SU2HIP_EINVAL, "NULL argument");
  static const int prime[10] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29};
  long vol = (long)L[0] * L[1] * L[2];
  if (size < 1 || vol % size != 0)
    return ioerr(SU2HIP_EINVAL, "Can't lay out %ld sites using %d nodes!", vol, size);
  int nf[10], i = size;
  for (int n = 0; n < 10; n++) {
    nf[n] = 0;
    while (i % prime[n] == 0) {
      i /= prime[n];
      nf[n]++;
    }
  }
  if (i != 1) return ioerr(SU2HIP_EINVAL, "Can't factorize %d nodes with primes up to 29", size);
  long sl[3] = {L[0], L[1], L[2]};
  int ns[3] = {1, 1, 1};
  for (int n = 9; n >= 0; n--)
    for (int f = 0; f < nf[n]; f++) {
      long largest = 1;
      for (int d = 0; d < 3; d++)
        if (sl[d] > largest && sl[d] % prime[n] == 0) largest = sl[d];
      if (largest == 1)
        return ioerr(SU2HIP_EINVAL, "no side of %d x %d x %d divides into %d nodes", L[0], L[1],
                     L[2], size);
      for (int d = 0; d < 3; d++)
        if (sl[d] == largest) {
          sl[d] /= prime[n];
          ns[d] *= prime[n];
          break;
        }
    }
  for (int d = 0; d < 3; d++) {
    nslices[d] = ns[d];
    slice[d] = (int32_t)sl[d];
  }
  return SU2HIP_OK;
}

/* Where the sites of an MPI-written checkpoint go.  write_field
 * (src/checkpoint.c:226-269) writes each field rank after rank; a rank's
 * sites are its node box (ranks numbered x-fastest over the node grid, the
 * sites' own indexing logic, doc/doc.tex:653; indexToCoords,
 * src/layout.c:945-963) in the node-local order: the box's sites x-fastest
 * (sitemap puts real sites before halos in that order, src/layout.c:
 * 245-402), then EVEN before ODD with the order kept (paritymap,
 * src/layout.c:826-851), parity from the global coordinates
 * (src/layout.c:730-756).  order[q] is the single-lattice site index
 * (EVEN first, natural order) of the q-th site of the file's field blocks. */
int su2hip_mpi_site_order(const int32_t *L, int size, int64_t *order) {
  if (!order || !L) return ioerr(SU2HIP_EINVAL, "NULL argument");
  for (int d = 0; d < 3; d++)
    if (L[d] < 2 || (L[d] & 1)) return ioerr(SU2HIP_EINVAL, "extent L%d = %d must be even", d + 1, L[d]);
  int32_t ns[3], sl[3];
  int rc = su2hip_mpi_slices(L, size, ns, sl);
  if (rc) return rc;
  long V = (long)L[0] * L[1] * L[2], Vh = V / 2;
  long per = (long)sl[0] * sl[1] * sl[2];
  long q = 0;
  for (int r = 0; r < size; r++) {
    long n0 = r % ns[0], n1 = (r / ns[0]) % ns[1], n2 = r / ((long)ns[0] * ns[1]);
    long o0 = n0 * sl[0], o1 = n1 * sl[1], o2 = n2 * sl[2];
    for (int par = 0; par < 2; par++)
      for (long k = 0; k < per; k++) {
        long x = o0 + k % sl[0], y = o1 + (k / sl[0]) % sl[1], z = o2 + k / ((long)sl[0] * sl[1]);
        if ((int)((x + y + z) & 1) != par) continue;
        long nat = x + (long)L[0] * (y + (long)L[1] * z);
        /* single-lattice order: parity block, then natural order inside it
         * (L[0] even: position nat >> 1) */
        order[q++] = par * Vh + (nat >> 1);
      }
  }
  return q == V ? SU2HIP_OK : ioerr(SU2HIP_EINVAL, "MPI site order: %ld of %ld sites", q, V);
}

int su2hip_lattice_read(const char *path, const su2hip_params *p, int64_t *iter,
                        double *total_time, double *comms_time, double *su2link,
                        double *u1link, double *doublet, double *triplet, double *singlet) {
  if (!path || !p) return ioerr(SU2HIP_EINVAL, "NULL argument");
  FILE *fp = fopen(path, "rb");
  if (!fp) return ioerr(SU2HIP_EIO, "cannot read %s", path);
  int size = 0, dim = 0, L[8] = {0};
  if (fread(&size, sizeof(int), 1, fp) != 1 || fread(&dim, sizeof(int), 1, fp) != 1 || dim < 1 ||
      dim > 8 || fread(L, sizeof(int), dim, fp) != (size_t)dim) {
    fclose(fp);
    return ioerr(SU2HIP_EIO, "Error reading latticefile %s", path);
  }
  int ok = (size >= 1 && dim == 3);
  for (int d = 0; ok && d < 3; d++) ok = L[d] == p->L[d];
  if (!ok) {
    fclose(fp);
    return ioerr(SU2HIP_EIO,
                 "Dimensions in latticefile do not match! Got MPI size %d, dimension %d, volume "
                 "%d x %d x %d; expected dimension 3, volume %d x %d x %d",
                 size, dim, L[0], L[1], L[2], p->L[0], p->L[1], p->L[2]);
  }
  long it = 0;
  double tt = 0, ct = 0;
  if (fread(&it, sizeof(long), 1, fp) != 1 || fread(&tt, sizeof(double), 1, fp) != 1 ||
      fread(&ct, sizeof(double), 1, fp) != 1) {
    fclose(fp);
    return ioerr(SU2HIP_EIO, "Error reading latticefile header of %s", path);
  }
  long V = vol_of(p);
  /* a file of an MPI run: the rank-major site order, mapped back */
  int64_t *order = NULL;
  double *buf = NULL;
  if (size > 1) {
    order = malloc(sizeof(int64_t) * V);
    buf = malloc(sizeof(double) * V * 12);
    if (!order || !buf) {
      free(order);
      free(buf);
      fclose(fp);
      return ioerr(SU2HIP_ENOMEM, "host allocation");
    }
    int32_t L32[3] = {L[0], L[1], L[2]};
    int rc = su2hip_mpi_site_order(L32, size, order);
    if (rc) {
      free(order);
      free(buf);
      fclose(fp);
      return rc;
    }
  }
  double *dst[5] = {su2link, u1link, doublet, triplet, singlet};
  int have[5] = {1, p->u1, p->nhiggs, p->triplet, p->singlet};
  int comps[5] = {12, 3, 4, 3, 1};
  int err = 0;
  for (int k = 0; k < 5 && !err; k++) {
    if (!have[k]) continue;
    /* doublets one after the other, each its own field block */
    int blocks = k == 2 ? p->nhiggs : 1;
    for (int bl = 0; bl < blocks && !err; bl++) {
      long n = V * comps[k];
      double *d = dst[k] ? dst[k] + (long)bl * n : NULL;
      if (!d) {
        err = fseek(fp, n * (long)sizeof(double), SEEK_CUR) != 0;
      } else if (!order) {
        err = fread(d, sizeof(double), n, fp) != (size_t)n;
      } else {
        err = fread(buf, sizeof(double), n, fp) != (size_t)n;
        for (long q = 0; q < V && !err; q++)
          memcpy(d + order[q] * comps[k], buf + q * comps[k], sizeof(double) * comps[k]);
      }
    }
  }
  fclose(fp);
  free(order);
  free(buf);
  if (err) return ioerr(SU2HIP_EIO, "Error reading field!");
  if (iter) *iter = it;
  if (total_time) *total_time = tt;
  if (comms_time) *comms_time = ct;
  return SU2HIP_OK;
}

/* ------------------------------------------------------------------------
 * weight file (src/multicanonical.c:82-288) */

int su2hip_weight_save(const char *path, const su2hip_muca *m, const double *pos, const double *W) {
  if (m->mode == SU2HIP_READONLY) return SU2HIP_OK; /* save_weight: nothing for read-only runs */
  FILE *fp = fopen(path, "w");
  if (!fp) return ioerr(SU2HIP_EIO, "cannot write %s", path);
  fprintf(fp, "%d %.12lf %d %lf %lf\n", m->bins, m->delta, m->last_max, m->wrk_min, m->wrk_max);
  for (long i = 1; i < m->bins + 2; i++) fprintf(fp, "%lf %lf\n", pos[i], W[i]);
  fclose(fp);
  return SU2HIP_OK;
}

int su2hip_weight_load(const char *path, su2hip_muca *m, double *pos, double *W, int max_bins) {
  FILE *fp = NULL;
  if (access(path, R_OK) == 0) fp = fopen(path, "r");
  if (!fp) {
    if (m->mode == SU2HIP_READONLY)
      return ioerr(SU2HIP_EIO, "No multicanonical weight given for a read-only run!");
    if (m->bins <= 0) return ioerr(SU2HIP_EINVAL, "Multicanonical error: bins <= 0");
    if (m->bins > max_bins) return ioerr(SU2HIP_ERANGE, "%d bins exceed room for %d", m->bins, max_bins);
    /* a flat weight with equal bins; the extra last bin starts at max */
    double dbin = (m->max - m->min) / ((double)(m->bins));
    for (int i = 0; i < m->bins + 1; i++) {
      pos[i + 1] = m->min + ((double)i) * dbin;
      W[i + 1] = 0.0;
    }
    m->last_max = 0;
    m->wrk_max = m->max;
    m->wrk_min = m->min;
    m->max += 0.001 * dbin;
    pos[m->bins + 1] = m->max;
  } else {
    int bins;
    if (fscanf(fp, "%d %lf %d %lf %lf", &bins, &m->delta, &m->last_max, &m->wrk_min,
               &m->wrk_max) != 5) {
      fclose(fp);
      return ioerr(SU2HIP_EIO, "Error reading first line of weightfile!");
    }
    if (bins <= 0) {
      fclose(fp);
      return ioerr(SU2HIP_EINVAL, "Multicanonical error: bins <= 0");
    }
    if (bins > max_bins) {
      fclose(fp);
      return ioerr(SU2HIP_ERANGE, "%d bins exceed room for %d", bins, max_bins);
    }
    m->bins = bins;
    for (int i = 1; i < bins + 2; i++) {
      int r = fscanf(fp, "%lf %lf", &pos[i], &W[i]);
      if (r != 2) {
        fclose(fp);
        return ioerr(SU2HIP_EIO, "Error reading weightfile! Got %d values at line %d", r, i + 1);
      }
    }
    fclose(fp);
    m->min = pos[1];
    m->max = pos[bins + 1];
    /* The reference tests `!w->mode != READONLY` (src/multicanonical.c:243,
     * :250): true exactly for READONLY runs, so the range check and the
     * shift of the last bin apply to read-only runs only. */
    int quirk = (!m->mode) != SU2HIP_READONLY;
    if (quirk && (m->wrk_min < m->min || m->wrk_max > m->max))
      return ioerr(SU2HIP_EINVAL, "Multicanonical error: weight update range [%lf, %lf] larger than binning range [%lf, %lf]",
                   m->wrk_min, m->wrk_max, m->min, m->max);
    if (quirk && fabs(m->wrk_max - m->max) < 1e-10) {
      double dbin = m->max - pos[bins];
      m->max += 0.001 * dbin;
      pos[bins + 1] = m->max;
    }
  }
  /* virtual end bins */
  pos[0] = m->min - 1000.0 * (m->max - m->min);
  W[0] = W[1];
  W[m->bins + 1] = W[m->bins];
  m->update_interval = m->mode != SU2HIP_SLOW ? 100 : 5000;
  m->muca_count = 0;
  m->enabled = 1;
  m->do_acceptance = 1;
  return su2hip_weight_save(path, m, pos, W);
}

/* print_labels_correlators (src/correlation.c:313-337) */
int su2hip_correl_labels_write(const char *path, const su2hip_params *p) {
  FILE *f = fopen(path, "w");
  if (!f) return ioerr(SU2HIP_EIO, "cannot write %s", path);
  int k = 1;
  fprintf(f, "%d distance\n", k++);
  if (p->nhiggs) fprintf(f, "%d H(l) = sum_z h(z) h(z+l) / V\n", k++);
  if (p->triplet) {
    fprintf(f, "%d Sigma^2 correlator\n", k++);
    fprintf(f, "%d projected photon correlator RE\n", k++);
    fprintf(f, "%d projected photon correlator IM\n", k++);
    fprintf(f, "%d projected photon correlator RE (p=2)\n", k++);
    fprintf(f, "%d projected photon correlator IM (p=2)\n", k++);
    fprintf(f, "%d alpha_munu RE (p=1)\n", k++);
    fprintf(f, "%d alpha_munu IM (p=1)\n", k++);
    fprintf(f, "%d alpha_munu RE (p=2)\n", k++);
    fprintf(f, "%d alpha_munu IM (p=2)\n", k++);
  }
  fclose(f);
  return SU2HIP_OK;
}

/* measure_correlators' file output (src/correlation.c:339-391) */
int su2hip_correl_append(const char *path, const su2hip_params *p, int meas_id,
                         const double *corr, int ndist) {
  if (!path || !p || (ndist > 0 && !corr)) return ioerr(SU2HIP_EINVAL, "correl_append arguments");
  FILE *f = fopen(path, "a");
  if (!f) return ioerr(SU2HIP_EIO, "cannot append to %s", path);
  fprintf(f, "\n =========== Measurement id: %d ===========\n", meas_id);
  for (int d = 0; d < ndist; d++) {
    const double *r = corr + (size_t)d * 10;
    fprintf(f, "%d ", d);
    if (p->nhiggs) fprintf(f, "%g ", r[0]);
    if (p->triplet) {
      fprintf(f, "%g ", r[1]);
      fprintf(f, "%g %g %g %g ", r[2], r[3], r[4], r[5]);
      fprintf(f, "%g %g %g %g ", r[6], r[7], r[8], r[9]);
    }
    fprintf(f, "\n");
  }
  fclose(f);
  return SU2HIP_OK;
}

/* print_z_labels (src/z_coord.c:78-106) */
int su2hip_zprof_labels_write(const char *path, const su2hip_params *p) {
  FILE *f = fopen(path, "w+");
  if (!f) return ioerr(SU2HIP_EIO, "cannot write %s", path);
  int k = 1;
  fprintf(f, "%d z\n", k++);
  if (p->nhiggs) fprintf(f, "%d phi^2\n", k++);
  if (p->triplet) fprintf(f, "%d Sigma^2\n", k++);
  if (p->singlet) {
    fprintf(f, "%d S\n", k++);
    fprintf(f, "%d S^2\n", k++);
  }
  fclose(f);
  return SU2HIP_OK;
}

/* measure_along_z's file output (src/z_coord.c:265-292) */
int su2hip_zprof_append(const char *path, const su2hip_params *p, const double *prof, int nz) {
  if (!path || !p || (nz > 0 && !prof)) return ioerr(SU2HIP_EINVAL, "zprof_append arguments");
  FILE *f = fopen(path, "a");
  if (!f) return ioerr(SU2HIP_EIO, "cannot append to %s", path);
  for (long z = 0; z < nz; z++) {
    const double *r = prof + (size_t)z * 4;
    fprintf(f, "%ld ", z);
    if (p->nhiggs) fprintf(f, "%g ", r[0]);
    if (p->triplet) fprintf(f, "%g ", r[1]);
    if (p->singlet) fprintf(f, "%g %g ", r[2], r[3]);
    fprintf(f, "\n");
  }
  fclose(f);
  return SU2HIP_OK;
}

/* load_weight_params (src/multicanonical.c:291-333) */
int su2hip_weight_params_load(const char *path, int bins, int64_t *gsum, int64_t *nsum, int *found) {
  if (!path || bins <= 0 || !gsum || !nsum) return ioerr(SU2HIP_EINVAL, "weight_params_load arguments");
  for (int i = 0; i < bins + 2; i++) {
    gsum[i] = 5; /* cannot be 0 */
    nsum[i] = 1;
  }
  FILE *fp = NULL;
  if (access(path, R_OK) == 0) fp = fopen(path, "r");
  if (found) *found = fp != NULL;
  if (!fp) return SU2HIP_OK; /* "Unable to access weight parameter file. Initializing new" */
  int bins_read = 0;
  if (fscanf(fp, "%d", &bins_read) != 1 || bins_read != bins) {
    fclose(fp);
    return ioerr(SU2HIP_EIO, "Error reading parameter file!! Expected %d bins, got %d", bins, bins_read);
  }
  for (int i = 1; i < bins + 1; i++) {
    double temp;
    long g, n;
    int r = fscanf(fp, "%lf %ld %ld", &temp, &g, &n);
    if (r != 3) {
      fclose(fp);
      return ioerr(SU2HIP_EIO, "Error reading 'weight_params'!! Got %d values at line %d", r, i + 1);
    }
    gsum[i] = g;
    nsum[i] = n;
  }
  fclose(fp);
  return SU2HIP_OK;
}

/* the SLOW part of save_weight (src/multicanonical.c:99-109) */
int su2hip_weight_params_save(const char *path, const su2hip_muca *m, const double *pos,
                              const int64_t *gsum, const int64_t *nsum) {
  if (m->mode != SU2HIP_SLOW) return SU2HIP_OK;
  FILE *fp = fopen(path, "w");
  if (!fp) return ioerr(SU2HIP_EIO, "cannot write %s", path);
  fprintf(fp, "%d \n", m->bins);
  for (long i = 1; i < m->bins + 1; i++)
    fprintf(fp, "%lf %ld %ld\n", pos[i], (long)gsum[i], (long)nsum[i]);
  fclose(fp);
  return SU2HIP_OK;
}

/* ------------------------------------------------------------------------
 * labels and measure lines (src/measure.c) */

int su2hip_labels_write(const char *path, const su2hip_params *p) {
  FILE *f = fopen(path, "w+");
  if (!f) return ioerr(SU2HIP_EIO, "cannot write %s", path);
  int k = 1;
  fprintf(f, "%d weight\n", k++);
  fprintf(f, "%d muca param\n", k++);
  fprintf(f, "%d action\n", k++);
  fprintf(f, "%d SU(2) Wilson (divided by beta)\n", k++);
  if (p->u1) fprintf(f, "%d U(1) Wilson (divided by beta)\n", k++);
  if (p->nhiggs) {
    fprintf(f, "%d 2*(phi^2 - phi^+(x) U_i(x) phi(x+i)) (avg over directions)\n", k++);
    fprintf(f, "%d phi^2\n", k++);
    fprintf(f, "%d phi^4\n", k++);
  }
  if (p->nhiggs == 2) {
    fprintf(f, "%d 2*(phi^2 - phi^+(x) U_i(x) phi(x+i)) for phi_2 (avg over directions)\n", k++);
    fprintf(f, "%d phi2^2\n", k++);
    fprintf(f, "%d phi2^4\n", k++);
    fprintf(f, "%d R = Re phi1^+ phi2\n", k++);
    fprintf(f, "%d I = Im phi1^+ phi2\n", k++);
  }
  if (p->triplet) {
    fprintf(f, "%d hopping_Sigma (avg over directions)\n", k++);
    fprintf(f, "%d Sigma^2\n", k++);
    fprintf(f, "%d Sigma^4\n", k++);
  }
  if (p->nhiggs && p->triplet) fprintf(f, "%d phi^2 Sigma^2\n", k++);
  if (p->triplet) {
    fprintf(f, "%d total magnetic charge density\n", k++);
    fprintf(f, "%d number of magnetic monopoles\n", k++);
  }
  if (p->singlet) {
    fprintf(f, "%d S^2 - S(x)S(x+i) (avg over directions)\n", k++);
    fprintf(f, "%d S\n", k++);
    fprintf(f, "%d S^2\n", k++);
    fprintf(f, "%d S^3\n", k++);
    fprintf(f, "%d S^4\n", k++);
    if (p->nhiggs == 1) {
      fprintf(f, "%d S phi^2\n", k++);
      fprintf(f, "%d S^2 phi^2\n", k++);
    }
  }
  fclose(f);
  return SU2HIP_OK;
}

extern int su2hip_measure_columns(const su2hip_params *p, const double *sums, double weight,
                                  double muca_param, double *cols, int *ncols);

int su2hip_measure_format(const su2hip_params *p, const double *sums, double weight,
                          double muca_param, char *buf, int buflen) {
  double cols[64];
  int n = 0;
  int rc = su2hip_measure_columns(p, sums, weight, muca_param, cols, &n);
  if (rc) return rc;
  int off = 0;
  for (int i = 0; i < n; i++) {
    int w = snprintf(buf + off, buflen - off, "%g ", cols[i]);
    if (w < 0 || w >= buflen - off) return ioerr(SU2HIP_ERANGE, "measure line buffer too small");
    off += w;
  }
  if (off + 2 > buflen) return ioerr(SU2HIP_ERANGE, "measure line buffer too small");
  buf[off++] = '\n';
  buf[off] = 0;
  return SU2HIP_OK;
}

/* print_acceptance (src/checkpoint.c:17-78) */
int su2hip_print_acceptance(FILE *out, const su2hip_run_config *cfg, const su2hip_counters *c) {
  const su2hip_params *p = &cfg->p;
  fprintf(out, "\n-- Acceptance rates --\n\t");
  if (p->algorithm_su2link == SU2HIP_METROPOLIS || p->triplet)
    fprintf(out, "SU(2) link %.2lf%%, ", 100.0 * c->accepted_su2link / c->total_su2link);
  if (p->u1 && p->algorithm_u1link == SU2HIP_METROPOLIS)
    fprintf(out, "U(1) link %.2lf%%, ", 100.0 * c->accepted_u1link / c->total_u1link);
  for (int db = 0; db < p->nhiggs; db++) {
    double acc = db ? c->accepted_doublet2 : c->accepted_doublet;
    double tot = db ? c->total_doublet2 : c->total_doublet;
    double oacc = db ? c->acc_overrelax_doublet2 : c->acc_overrelax_doublet;
    double otot = db ? c->total_overrelax_doublet2 : c->total_overrelax_doublet;
    if (p->algorithm_su2doublet == SU2HIP_METROPOLIS)
      fprintf(out, "Higgs#%d %.2lf%%, ", db + 1, 100.0 * acc / tot);
    else if (p->algorithm_su2doublet == SU2HIP_OVERRELAX)
      fprintf(out, "Higgs#%d overrelax %.2lf%%, Higgs#%d Metropolis %.2lf%%, ", db + 1,
              100.0 * oacc / otot, db + 1, 100.0 * acc / tot);
  }
  if (p->triplet) {
    if (p->algorithm_su2triplet == SU2HIP_METROPOLIS)
      fprintf(out, "Triplet %.2lf%%, ", 100.0 * c->accepted_triplet / c->total_triplet);
    else if (p->algorithm_su2triplet == SU2HIP_OVERRELAX)
      fprintf(out, "Triplet overrelax %.2lf%%, Triplet Metropolis %.2lf%%, ",
              100.0 * c->acc_overrelax_triplet / c->total_overrelax_triplet,
              100.0 * c->accepted_triplet / c->total_triplet);
  }
  if (p->singlet) {
    if (p->algorithm_singlet == SU2HIP_METROPOLIS)
      fprintf(out, "Singlet %.2lf%%, ", 100.0 * c->accepted_singlet / c->total_singlet);
    else if (p->algorithm_singlet == SU2HIP_OVERRELAX)
      fprintf(out, "Singlet overrelax %.2lf%%, Singlet Metropolis %.2lf%%, ",
              100.0 * c->acc_overrelax_singlet / c->total_overrelax_singlet,
              100.0 * c->accepted_singlet / c->total_singlet);
  }
  if (p->link_overrelax > 0 && c->total_link_or > 0)
    fprintf(out, "link overrelax %.2lf%%, ", 100.0 * c->accepted_link_or / c->total_link_or);
  if (cfg->multicanonical) fprintf(out, "multicanonical %.2lf%%", 100.0 * c->accepted_muca / c->total_muca);
  fprintf(out, "\n");
  return SU2HIP_OK;
}
