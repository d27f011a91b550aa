/*
 * Runs the FP64 oracle (oracle/fargo_oracle.c, compiled into this program
 * with AddressSanitizer and UndefinedBehaviorSanitizer) for a few substeps
 * of a deck, touching every output path: the substep loop, the outputs,
 * the diagnostics.  tests/test_sanitizers.py builds and runs it.
 *
 * usage: sanitize_driver DECK OUTDIR NSUBSTEPS [DIAGNOSTICS]
 */
#include <stdio.h>
#include <stdlib.h>

#include "fargo_oracle.h"

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s DECK OUTDIR NSUBSTEPS [DIAGNOSTICS]\n", argv[0]);
    return 2;
  }
  or_sim *s = or_create(argv[1], "", argv[2], 1);
  if (!s) {
    fprintf(stderr, "or_create: %s\n", or_last_error());
    return 2;
  }
  if (argc > 4 && atoi(argv[4])) or_enable_diagnostics(s);
  const int n = atoi(argv[3]);
  or_write_outputs(s, 0);
  for (int k = 0; k < n; k++) or_substep(s);
  or_write_outputs(s, 1);
  if (argc > 4 && atoi(argv[4])) {
    or_update_log(s, 1);
    or_write_bigplanet(s, 1);
  }
  double m = or_circumplanetary_mass(s);
  printf("ok %d substeps, t = %.6f, circumplanetary mass %.6e\n", n, or_get(s, "PhysicalTime"), m);
  or_destroy(s);
  return 0;
}
