/* bind_shim.c -- TEST INFRASTRUCTURE.  Runs the reference's own, unmodified
 * main loop (src/main.c:251-379) with its sweep replaced by the GPU one
 * through the reference-side binding su2higgs_amd/bind/su2hip_bind.c.
 *
 * Linked with -Wl,--wrap=update_lattice: every call of update_lattice
 * (src/update.c:270, from src/main.c:265 and :353) lands here.  The first
 * call uploads the fields the reference initialised; every call then runs
 * one compound sweep on the GPU and brings the fields back, so that the
 * reference's measure() and save_lattice() see them.  In a multicanonical
 * run the weight moves to the device at the first call, again whenever the
 * main loop changes its mode (READONLY during thermalization,
 * src/main.c:255-278), and its state comes back after every sweep, so that
 * measure() prints the device's order parameter and weight.  The
 * acceptance counts since the previous call are added into the reference's
 * counters, for print_acceptance at checkpoints.
 *
 * Linked with -Wl,--wrap=save_lattice too: after each checkpoint the
 * binding writes the Philox position beside the lattice file.
 *
 * The binding takes its Philox key from the reference's generator, which
 * main() seeds from the clock.  With SU2HIP_SHIM_MT_SEED set, the shim
 * seeds that generator with it first, so a test can ask for two different
 * keys within one second.
 */
#include <stdio.h>
#include <stdlib.h>

#include "su2.h"
#include "generic/mersenne.h"
#include "su2hip.h"

int su2hip_bind_init(lattice const *l, fields const *f, params const *p);
int su2hip_bind_update(void);
int su2hip_bind_download(fields *f);
int su2hip_bind_muca(weight const *w);
int su2hip_bind_muca_download(weight *w);
int su2hip_bind_counters(counters *c);
int su2hip_bind_save_rng(const char *latticefile);
void __real_save_lattice(lattice const *l, fields f, counters c, char *fname);

static void die_with(const char *what) {
  fprintf(stderr, "bind_shim: %s: %s\n", what, su2hip_last_error());
  exit(1);
}

void __wrap_update_lattice(lattice *l, fields *f, params const *p, counters *c, weight *w) {
  static int ready = 0, mode = -1;
  if (!ready) {
    const char *mt = getenv("SU2HIP_SHIM_MT_SEED");
    if (mt) seed_mersenne(strtoull(mt, NULL, 10));
    if (su2hip_bind_init(l, f, p)) die_with("su2hip_bind_init");
    ready = 1;
  }
  if (p->multicanonical && w->mode != mode) {
    if (su2hip_bind_muca(w)) die_with("su2hip_bind_muca");
    mode = w->mode;
  }
  if (su2hip_bind_update()) die_with("su2hip_bind_update");
  if (su2hip_bind_download(f)) die_with("su2hip_bind_download");
  if (p->multicanonical && su2hip_bind_muca_download(w)) die_with("su2hip_bind_muca_download");
  if (su2hip_bind_counters(c)) die_with("su2hip_bind_counters");
}

void __wrap_save_lattice(lattice const *l, fields f, counters c, char *fname) {
  __real_save_lattice(l, f, c, fname);
  if (su2hip_bind_save_rng(fname)) die_with("su2hip_bind_save_rng");
}
