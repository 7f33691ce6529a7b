/* bind_shim.c -- TEST INFRASTRUCTURE.  Runs the reference's own, unmodified
 * main loop (src/main.c:251-379) with its sweep replaced by the GPU one
 * through the reference-side binding su2higgs_amd/bind/su2hip_bind.c.
 *
 * Linked with -Wl,--wrap=update_lattice: every call of update_lattice
 * (src/update.c:270, from src/main.c:265 and :353) lands here.  The first
 * call uploads the fields the reference initialised; every call then runs
 * one compound sweep on the GPU and brings the fields back, so that the
 * reference's measure() and save_lattice() see them.  Multicanonical runs
 * are not routed (the binding does not pass the weight), and the counters
 * are not filled in.
 */
#include <stdio.h>
#include <stdlib.h>

#include "su2.h"
#include "su2hip.h"

int su2hip_bind_init(lattice const *l, fields const *f, params const *p);
int su2hip_bind_update(void);
int su2hip_bind_download(fields *f);

void __wrap_update_lattice(lattice *l, fields *f, params const *p, counters *c, weight *w) {
  static int ready = 0;
  (void)c;
  (void)w;
  if (p->multicanonical) {
    fprintf(stderr, "bind_shim: multicanonical runs are not routed to the GPU\n");
    exit(1);
  }
  if (!ready) {
    if (su2hip_bind_init(l, f, p)) {
      fprintf(stderr, "bind_shim: su2hip_bind_init: %s\n", su2hip_last_error());
      exit(1);
    }
    ready = 1;
  }
  if (su2hip_bind_update() || su2hip_bind_download(f)) {
    fprintf(stderr, "bind_shim: %s\n", su2hip_last_error());
    exit(1);
  }
}
