// Self-gravity entry points (SgMain.cu).  Filled in by fg_sg.hip.
#include "fg_internal.h"

fg_status fg_selfgravity_kick(fg_ctx *c, double dt) {
  (void)c;
  (void)dt;
  fg_set_error("self-gravity is not available in this build");
  return FG_E_UNSUPPORTED;
}

extern "C" fg_status fg_selfgravity(fg_ctx *c, double dt, int update) {
  (void)update;
  return fg_selfgravity_kick(c, dt);
}
