/*
 * The reference-side binding of INTEGRATION.md section 2, compiled and run.
 *
 * The block between the "snippet" markers is the AlgoGas replacement a
 * maintainer puts into the reference's SourceEuler.cu (AlgoGas,
 * SourceEuler.cu:199-307): the same host sequence, with every gas routine a
 * call into libfargo_hip.so.  tests/test_integration_snippet.py checks that
 * INTEGRATION.md shows exactly this text, compiles this file against
 * include/fargo_hip.h and, on the GPU, runs it.
 *
 * Around the snippet, this file supplies what the reference's program
 * would: its globals (Param.cuh, Interpret.cu) and its host planet routines
 * (AdvanceSystemRK5, GetPsysInfo, RotatePsys, Psys.cu / Pframeforce.cu).
 * Those routines are the reference's own and are not part of the binding,
 * so here they take the planets from a second run of the same deck driven
 * by fg_sim_substep (the library's host driver), stepped in lockstep.  The
 * program then checks, after every AlgoGas call, that the snippet's state,
 * time and lost mass equal that run's bit for bit, and that each substep
 * used the same dt.
 *
 * usage: integration_snippet DECK NCALLS MASSTAPER THICKNESSSMOOTHING ASPECTRATIO FLARINGINDEX
 * (the deck must use a fixed frame: Frame FIXED)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fargo_hip.h"

#define YES 1
#define NO 0
#define MARK 1 /* Types.cuh:43-45 */
#define GET 0

/* the reference's globals used by AlgoGas (Param.cuh, Interpret.cu) */
static double DT, MASSTAPER, PhysicalTime, MassTaper, OmegaFrame, LostMass;
static double ROCHESMOOTHING, THICKNESSSMOOTHING, ASPECTRATIO, FLARINGINDEX;
static int Adiabatic, IsDisk = YES, SloppyCFL = NO, Corotating = NO, Indirect_Term = YES, RocheSmoothing = NO;

/* Types.cuh:2-16 (the fields AlgoGas touches) */
typedef struct {
  int nb;
  double mass[FG_MAX_PLANETS], x[FG_MAX_PLANETS], y[FG_MAX_PLANETS], vx[FG_MAX_PLANETS], vy[FG_MAX_PLANETS];
  int FeelDisk[FG_MAX_PLANETS];
} PlanetarySystem;

/* Force.cu:137-142 (no transition zone in the test decks) */
static double compute_smoothing(double r) { return THICKNESSSMOOTHING * ASPECTRATIO * pow(r, 1.0 + FLARINGINDEX); }

/* ---- stand-ins for the reference's host planet routines -------------- */
static fg_sim *g_run;   /* the same deck, stepped by fg_sim_substep */
static double g_run_dt; /* dt of the run's last substep */
static int g_fail;

static void load_planets(PlanetarySystem *sys) {
  double p[5];
  fg_sim_nplanets(g_run, &sys->nb);
  for (int k = 0; k < sys->nb; k++) {
    fg_sim_planet(g_run, k, p);
    sys->x[k] = p[0];
    sys->y[k] = p[1];
    sys->vx[k] = p[2];
    sys->vy[k] = p[3];
    sys->mass[k] = p[4];
  }
}

/* AdvanceSystemRK5 (Pframeforce.cu:124-187): the run takes the same
 * substep, and the planets become its planets after it */
static void AdvanceSystemRK5(PlanetarySystem *sys, double dt) {
  int closed = 0;
  if (fg_sim_substep(g_run, &closed) != FG_OK) {
    fprintf(stderr, "fg_sim_substep: %s\n", fg_last_error());
    exit(2);
  }
  fg_sim_get(g_run, "dt", &g_run_dt);
  if (g_run_dt != dt) {
    fprintf(stderr, "substep dt differs: snippet %.17g, fg_sim_substep %.17g\n", dt, g_run_dt);
    g_fail = 1;
  }
  load_planets(sys);
}
static double GetPsysInfo(PlanetarySystem *sys, int action) {
  (void)sys;
  (void)action;
  return 0.0; /* fixed frame only */
}
static void RotatePsys(PlanetarySystem *sys, double angle) {
  (void)sys;
  (void)angle; /* the planets loaded above are already in the run's frame */
}

/* ---- snippet ---------------------------------------------------------- */
/* AlgoGas (SourceEuler.cu:199-307) with the gas on the GPU.  g_ctx comes
 * from fg_create(&params, &tables, device, &g_ctx) after InitEuler, with
 * fg_upload_state(g_ctx, Dens, Vrad, Vtheta, Energy). */
static fg_status AlgoGas_gpu(fg_ctx *g_ctx, PlanetarySystem *sys) {
  double dt, dtemp = 0.0, OmegaNew, domega, acc[2], ix = 0.0, iy = 0.0, lost;
  int gastimestepcfl = 1;
  fg_planet_pot pot[FG_MAX_PLANETS];
  fg_status st;
#define FG_TRY(call)                   \
  do {                                 \
    if ((st = (call)) != FG_OK) return st; \
  } while (0)
  if (Adiabatic) FG_TRY(fg_sound_speed(g_ctx)); /* ComputeSoundSpeed, SourceEuler.cu:208 */
  if (IsDisk == YES && SloppyCFL == YES) FG_TRY(fg_cfl(g_ctx, DT - dtemp, &gastimestepcfl));
  dt = DT / gastimestepcfl;
  while (dtemp < 0.999999999 * DT) {
    MassTaper = PhysicalTime / (MASSTAPER * 2.0 * M_PI);
    MassTaper = (MassTaper > 1.0 ? 1.0 : pow(sin(MassTaper * M_PI / 2.0), 2.0));
    if (IsDisk == YES && SloppyCFL == NO) {
      FG_TRY(fg_cfl(g_ctx, DT - dtemp, &gastimestepcfl)); /* ConditionCFL, SourceEuler.cu:228 */
      dt = (DT - dtemp) / (double)gastimestepcfl;
    }
    dtemp += dt;
    if (Corotating == YES) GetPsysInfo(sys, MARK);
    if (IsDisk == YES) {
      /* ComputeAccel (SideEuler.cu:135) -> indirect term (Pframeforce.cu:86-95) */
      FG_TRY(fg_disk_force(g_ctx, 0.0, 0.0, 0.0, acc));
      ix = (Indirect_Term == YES) ? -acc[0] : 0.0;
      iy = (Indirect_Term == YES) ? -acc[1] : 0.0;
      /* FillForcesArrays (Pframeforce.cu:56-83): per-planet constants here,
       * the potential field on the GPU */
      for (int k = 0; k < sys->nb; k++) {
        double xp = sys->x[k], yp = sys->y[k], m = sys->mass[k] * MassTaper;
        double d = sqrt(xp * xp + yp * yp);
        double sm = RocheSmoothing ? d * pow((1.0 / 3.0 * m), 1.0 / 3.0) * ROCHESMOOTHING : compute_smoothing(d);
        pot[k].x = xp;
        pot[k].y = yp;
        pot[k].mass = m;
        pot[k].smooth = sm * sm;
        pot[k].inv_d3 = 1.0 / d / d / d;
      }
      FG_TRY(fg_potential(g_ctx, pot, sys->nb, ix, iy));
      /* AdvanceSystemFromDisk (Pframeforce.cu:97-121) */
      for (int k = 0; k < sys->nb; k++) {
        if (sys->FeelDisk[k] == NO) continue;
        double x = sys->x[k], y = sys->y[k], m = sys->mass[k], r = sqrt(x * x + y * y), g[2];
        double sm = RocheSmoothing ? r * pow(m / 3., 1. / 3.) * ROCHESMOOTHING : compute_smoothing(r);
        FG_TRY(fg_disk_force(g_ctx, x, y, sm, g));
        sys->vx[k] += dt * g[0];
        sys->vy[k] += dt * g[1];
        sys->vx[k] += dt * ix;
        sys->vy[k] += dt * iy;
      }
    }
    AdvanceSystemRK5(sys, dt);
    if (Corotating == YES) {
      OmegaNew = GetPsysInfo(sys, GET) / dt;
      domega = OmegaNew - OmegaFrame;
      if (IsDisk == YES) FG_TRY(fg_correct_vtheta(g_ctx, domega)); /* CorrectVtheta, SideEuler.cu:181 */
      OmegaFrame = OmegaNew;
    }
    RotatePsys(sys, OmegaFrame * dt);
    if (IsDisk == YES) {                                       /* SourceEuler.cu:264-295 */
      FG_TRY(fg_boundary(g_ctx, dt, OmegaFrame));              /* ApplyBoundaryCondition */
      FG_TRY(fg_pressure(g_ctx));                              /* ComputePressureField */
      FG_TRY(fg_substep1(g_ctx, dt, OmegaFrame));              /* Substep1 (+ SG, viscosity) */
      FG_TRY(fg_substep2(g_ctx, dt));                          /* Substep2, host, ActualiseGas* */
      FG_TRY(fg_boundary(g_ctx, dt, OmegaFrame));              /* ApplyBoundaryCondition */
      if (Adiabatic) FG_TRY(fg_substep3(g_ctx, dt));           /* viscous terms, Substep3 */
      FG_TRY(fg_transport(g_ctx, dt, OmegaFrame, &lost));      /* Transport */
      LostMass += lost;
      FG_TRY(fg_boundary(g_ctx, dt, OmegaFrame));              /* ApplyBoundaryCondition */
      FG_TRY(fg_temperature(g_ctx));                           /* ComputeTemperatureField */
    }
    PhysicalTime += dt;
  }
  return FG_OK;
}
/* ---- end snippet ------------------------------------------------------ */

static int compare_state(fg_sim *snippet_owner, long S, int call) {
  double *a = malloc(8 * S * sizeof(double)), *b = a + 4 * S;
  int bad = 0;
  fg_download_state(fg_sim_ctx(snippet_owner), a, a + S, a + 2 * S, Adiabatic ? a + 3 * S : NULL, NULL, NULL);
  fg_sim_get_state(g_run, b, b + S, b + 2 * S, Adiabatic ? b + 3 * S : NULL);
  const char *names[4] = {"dens", "vrad", "vtheta", "energy"};
  for (int f = 0; f < (Adiabatic ? 4 : 3); f++)
    for (long l = 0; l < S; l++)
      if (memcmp(&a[f * S + l], &b[f * S + l], sizeof(double)) != 0) {
        fprintf(stderr, "AlgoGas call %d: %s[%ld] snippet %.17g, fg_sim_substep %.17g\n", call, names[f], l,
                a[f * S + l], b[f * S + l]);
        bad = 1;
        break;
      }
  free(a);
  return bad;
}

int main(int argc, char **argv) {
  if (argc != 7) {
    fprintf(stderr, "usage: %s DECK NCALLS MASSTAPER THICKNESSSMOOTHING ASPECTRATIO FLARINGINDEX\n", argv[0]);
    return 2;
  }
  const int ncalls = atoi(argv[2]);
  MASSTAPER = atof(argv[3]);
  THICKNESSSMOOTHING = atof(argv[4]);
  ASPECTRATIO = atof(argv[5]);
  FLARINGINDEX = atof(argv[6]);
  fg_sim_options opt;
  memset(&opt, 0, sizeof opt);
  fg_sim *owner = NULL;
  /* the snippet's context: created with the deck's parameters and tables
   * (fg_sim_create fills fg_params/fg_tables the way Interpret.cu and
   * FillPolar1DArrays fill the reference's globals); only fg_sim_ctx of it
   * is used below */
  if (fg_sim_create(argv[1], &opt, &owner) != FG_OK || fg_sim_create(argv[1], &opt, &g_run) != FG_OK) {
    fprintf(stderr, "fg_sim_create: %s\n", fg_last_error());
    return 2;
  }
  double v;
  fg_sim_get(g_run, "DT", &v);
  DT = v;
  fg_sim_get(g_run, "OmegaFrame", &OmegaFrame);
  fg_sim_get(g_run, "PhysicalTime", &PhysicalTime);
  fg_sim_get(g_run, "Adiabatic", &v);
  Adiabatic = v != 0.0;
  PlanetarySystem sys;
  memset(&sys, 0, sizeof sys);
  load_planets(&sys);
  for (int k = 0; k < sys.nb; k++) sys.FeelDisk[k] = YES;
  fg_ctx *ctx = fg_sim_ctx(owner);
  double nrad, nsec;
  fg_sim_get(g_run, "NRAD", &nrad);
  fg_sim_get(g_run, "NSEC", &nsec);
  const long S = ((long)nrad + 1) * (long)nsec;
  for (int call = 0; call < ncalls; call++) {
    fg_status st = AlgoGas_gpu(ctx, &sys);
    if (st != FG_OK) {
      fprintf(stderr, "AlgoGas_gpu: status %d: %s\n", st, fg_last_error());
      return 1;
    }
    double t, lost;
    fg_sim_get(g_run, "PhysicalTime", &t);
    fg_sim_get(g_run, "LostMass", &lost);
    if (t != PhysicalTime || lost != LostMass) {
      fprintf(stderr, "call %d: time %.17g vs %.17g, lost mass %.17g vs %.17g\n", call, PhysicalTime, t, LostMass,
              lost);
      g_fail = 1;
    }
    g_fail |= compare_state(owner, S, call);
    if (g_fail) return 1;
  }
  printf("snippet matches fg_sim_substep bit for bit over %d AlgoGas calls\n", ncalls);
  fg_sim_destroy(owner);
  fg_sim_destroy(g_run);
  return 0;
}
