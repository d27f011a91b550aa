// Fused substep path (filled in below the per-routine path).
#include "fg_internal.h"

extern "C" fg_status fg_step_fused(fg_ctx *c, double dt, double omega_frame, double domega,
                                   const fg_planet_pot *planets, int nb, double ix, double iy) {
  (void)c; (void)dt; (void)omega_frame; (void)domega; (void)planets; (void)nb; (void)ix; (void)iy;
  fg_set_error("fused path not built");
  return FG_E_UNSUPPORTED;
}
extern "C" fg_status fg_fused_prepare(fg_ctx *c, double DeltaT) {
  (void)c; (void)DeltaT;
  fg_set_error("fused path not built");
  return FG_E_UNSUPPORTED;
}
extern "C" fg_status fg_fused_results(fg_ctx *c, int *ncfl, double *force2, double *lost) {
  (void)c; (void)ncfl; (void)force2; (void)lost;
  fg_set_error("fused path not built");
  return FG_E_UNSUPPORTED;
}
