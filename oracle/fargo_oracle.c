/*
 * FP64 serial oracle of the FARGO / FARGO-ADSG hydro substep.
 *
 * TEST INFRASTRUCTURE ONLY (see fargo_oracle.h).  A plain-C restatement of
 * the reference routines of MatiasFloresS/CudaProtoplanetaryDisks, written
 * from reading src/*.cu as text.  Every CUDA kernel becomes a loop nest
 * over the same index range, with the same expression in the same
 * left-to-right evaluation order, so that a build with
 * -ffp-contract=off gives the reference's FP64 rounding for every +,-,*,/
 * and sqrt.  Reductions follow the reference's own order: sequential ring
 * sums (Kernels.cu:618-621, 1076-1079, 1276-1279) and DeviceReduce's
 * 32-lane tree + sequential host sum over blocks (Kernels.cu:253-394).
 * OpenMP pragmas parallelise only per-cell loops (no cross-cell
 * accumulation), so the -fopenmp build gives bit-identical results.
 *
 * Parity pins: every routine below has a check the mathematics fixes,
 * chosen so that a dropped term, a wrong sign or factor or a shifted index
 * fails (tests/test_oracle_pins.py, test_oracle_physics_pins.py,
 * test_oracle_diagnostics.py):
 *   parser            known answers on the reference decks in/*.par, *.cfg
 *   grid              closed forms: sum Surf*Ns = pi(Rmax^2-Rmin^2), Rmed
 *   initial state     closed-form Sigma (with cavity), vtheta, vrad, energy
 *   EOS               c_s, P, T closed forms (isothermal and adiabatic)
 *   aspect ratio      transition-zone closed form (through the smoothing)
 *   DeviceReduce      exact (Fraction) sums; order replayed by hand
 *   CFL               brute-force min over cells
 *   potential         softened point mass + both indirect terms at every
 *                     cell, planet on a cell (d = 0), star alone, Roche
 *   Substep1          gradients of linear P and Phi (radial, azimuthal),
 *                     centrifugal term in a rotating frame, imposed drift
 *   viscous stress    power-law shear flows (tau_rphi = rho nu r dOmega/dr),
 *                     radial compression v_r ~ r^2 (tau_rr, div v)
 *   viscous update    their divergences: nu A (q^2-1) r^(q-2), 4 nu v0
 *   alpha viscosity   nu = alpha c_s^2 / Omega_K, ring mean, cavity taper
 *   artificial visc.  uniform compression: CVNR^2 rho dv^2, -grad q / rho,
 *                     the compressional heating; expansion gives zero
 *   Substep3          backward-Euler compression and cooling relaxation,
 *                     equilibrium fixed point, shear-flow heating, ring-0
 *                     power-law extrapolation
 *   radial transport  brute-force van Leer column, divergence-free flow
 *                     keeps uniform Sigma, mass conservation
 *   azimuthal transp. brute-force FARGO ring (mean shift, uniform and
 *                     residual sweeps), integer-Courant shift, TVD
 *   ComputeVelocities identity for a vanishing step (rotating frame)
 *   boundaries        evanescent closed form; open edge ghost values;
 *                     non-reflecting shift, mean reset and characteristic
 *                     face velocity; sub-Keplerian equilibrium rotation
 *   frames            corotating (planet fixed), guiding centre (frame at
 *                     the mean motion, mean longitude fixed); CorrectVtheta
 *   planets (RK5)     circular Kepler orbit; forced circular rotation;
 *                     release migration; planet-planet term (quirk)
 *   AlgoGas           MassTaper 0 on the first substep, substeps add up to
 *                     DT, NTOT+1 calls
 *   FFT               numpy.fft on random data; brute-force convolution
 *   self-gravity      the finite-difference gradient of the direct sum of
 *                     softened log-polar cell potentials; the velocity kick
 *                     (linear interpolation) and the ring mean
 *   disk force, Hill  point-mass closed forms, split radius, circumplanetary
 *                     mass (test_oracle_diagnostics.py)
 * Golden fixtures (tests/golden/) are written by this oracle: they guard
 * against regressions only.
 * Reference behaviour no check here can reach (the reference needs nvcc
 * and an NVIDIA GPU, and ships no outputs), decided as documented:
 *   ulp behaviour of CUDA's device libm inside kernels; the layout-
 *   dependent ViscosityAlpha pun (read here as "alpha viscosity on" in
 *   FViscosity, "off" in the initial-velocity kernel, as the reference's
 *   globals make it); the self-gravity path, which cannot run in the
 *   reference (Make1Dprofile reads SoundSpeed_d before it is allocated),
 *   is restated from the code's evident intent (ring mean of SG_Accr).
 */
#define _GNU_SOURCE
#include "fargo_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

/* Fondam.cuh:1-10 */
#define G 1.0
#define PI 3.14159265358979323844
#define MU 1.0
#define RGAS 1.0
#define CFLSECURITY 0.5
#define CVNR 1.41
#define YES 1
#define NO 0
/* Types.cuh:43-45 */
#define GET 0
#define MARK 1
#define FREQUENCY 2

#define NSTR 300

typedef struct {
  double x, y;
} pair_t;

struct or_sim {
  /* ---- Param.cuh:9-62 (defaults) ---- */
  double DT, SIGMA0;
  int NINTERM, NTOT;
  char OUTPUTDIR[1024], INNERBOUNDARY[NSTR], LABELADVECTION[NSTR];
  char TRANSPORT[NSTR], PLANETCONFIG[NSTR], RADIALSPACING[NSTR];
  double MASSTAPER;
  int NRAD, NSEC;
  double RMIN, RMAX, THICKNESSSMOOTHING, ROCHESMOOTHING, ASPECTRATIO;
  double VISCOSITY, ALPHAVISCOSITY, SIGMASLOPE, RELEASERADIUS, RELEASEDATE;
  double OMEGAFRAME;
  char DISK[NSTR], FRAME[NSTR], OUTERSOURCEMASS[NSTR], WRITEDENSITY[NSTR];
  char WRITEVELOCITY[NSTR], WRITEENERGY[NSTR], WRITETEMPERATURE[NSTR];
  char WRITEDIVV[NSTR], WRITEQPLUS[NSTR], INDIRECTTERM[NSTR], EXCLUDEHILL[NSTR];
  double IMPOSEDDISKDRIFT, FLARINGINDEX, ECCENTRICITY, CAVITYRADIUS;
  double CAVITYRATIO, CAVITYWIDTH, TRANSITIONRADIUS, TRANSITIONRATIO;
  double TRANSITIONWIDTH, LAMBDADOUBLING;
  char SELFGRAVITY[NSTR], CICPLANET[NSTR], FORCEDCIRCULAR[NSTR], ZMPLUS[NSTR];
  char ADIABATIC[NSTR], COOLING[NSTR];
  double ADIABATICINDEX, COOLINGTIME0;
  /* extension keys with no reference counterpart (default off):
   * COOLINGBETA > 0 gives beta-cooling t_cool = beta / Omega_K(r). */
  double COOLINGBETA;

  /* ---- derived flags, Interpret.cu:22-31, 130-206 ---- */
  int ExcludeHill, ViscosityAlpha, RocheSmoothing, OpenInner, AdvecteLabel;
  int LogGrid, FastTransport, GuidingCenter, Indirect_Term, IsDisk;
  int NonReflecting, Corotating, OuterSourceMass, Evanescent;
  int Write_Density, Write_Velocity, Write_Energy, Write_Temperature;
  int Write_DivV, Write_Qplus, SelfGravity, SGZeroMode, ZMPluss;
  int Adiabatic, Cooling, CICPlanet, ForcedCircular;
  /* CLI, Main.cu:113-179 */
  int SloppyCFL;
  double ScalingFactor;
  int write_files;

  /* ---- grid ---- */
  int nrad, nsec;
  long size_grid;
  double *Radii, *Rinf, *Rsup, *Rmed, *Surf, *invRinf, *invSurf, *invRmed;
  double *invdiffRsup, *invdiffRmed, *powRmed;
  double *SigmaMed, *SigmaInf, *EnergyMed, *CoolingTimeMed, *QplusMed;
  double *viscosity_array, *GLOBAL_bufarray, *GLOBAL_AxiSGAccr;

  /* ---- 2D fields (size_grid) ---- */
  double *Dens, *Vrad, *Vtheta, *Energy;
  double *Pressure, *SoundSpeed, *Temperature, *Potential;
  double *VradInt, *VthetaInt, *VradNew, *VthetaNew, *EnergyInt, *EnergyNew;
  double *DensInt, *TemperInt, *Qplus;
  double *DRR, *DPP, *DRP, *DivergenceVelocity, *TAURR, *TAUPP, *TAURP;
  double *RadMomP, *RadMomM, *ThetaMomP, *ThetaMomM, *Work, *QRStar;
  double *DensStar, *dq, *VthetaRes, *TempShift;
  double *CellAbscissa, *CellOrdinate;
  double *fxi, *fxo, *fyi, *fyo;
  double *Vresidual, *DT2D, *Vmoy, *newDT, *DT1D, *VMed, *LostByDisk;
  double *Vradial, *Vazimutal;
  int *Nshift, *NoSplitAdvection;
  /* self-gravity (SgMain.cu): complex arrays of 2*size_grid */
  double *Kr_aux, *Kt_aux;
  double *SGP_Kr, *SGP_Kt, *SGP_Sr, *SGP_St;       /* interleaved complex */
  double *SGP_Krc, *SGP_Ktc, *SGP_Src, *SGP_Stc;
  double *Gr, *Gphi, *Grc, *Gphic;
  double *SG_Accr, *SG_Acct;
  double SGP_eps, SGP_rstep, SGP_tstep;

  /* ---- scalars ---- */
  double PhysicalTime, PhysicalTimeInitial, OmegaFrame, MassTaper;
  double LostMass, mdcp, mdcp1, exces_mdcp, HillRadius;
  int Diagnostics; /* the commented-out Main.cu:302,314-315 / SourceEuler.cu:297-298 calls */
  pair_t DiskOnPrimaryAcceleration, IndirectTerm;
  double Xplanet_ps, Yplanet_ps; /* Psys.cu:4 statics */
  int init_counter;              /* SourceEuler.cu:58 */
  double globalforce[4 * 64];    /* Force.cu:93, dimfxy <= 64 (the reference uses 11) */
  double fx_inner, fy_inner, fx_ex_inner, fy_ex_inner;
  double fx_outer, fy_outer, fx_ex_outer, fy_ex_outer;
  int dimfxy;
  /* AlgoGas loop state (SourceEuler.cu:202-222) */
  int in_dt, gastimestepcfl;
  double dt, dtemp;
  int last_cfl;

  /* ---- planets (Types.cuh:2-12) ---- */
  int nb;
  double *px, *py, *pvx, *pvy, *pmass, *pacc;
  int *feeldisk, *feelothers;
  /* RungeKutta.cu:3-4 statics and Main.cu:653-655 buffers */
  double k1[400], k2[400], k3[400], k4[400], k5[400], k6[400], Dist[100];
  double q0[400], q1[400], PlanetMasses[100];
};

static char g_err[512];
const char *or_last_error(void) { return g_err; }
static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

static double *dalloc(long n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

/* ------------------------------------------------------------------ */
/* parameters: Param.cuh:9-62, Interpret.cu:36-212                      */
/* ------------------------------------------------------------------ */
static void set_defaults(or_sim *s) {
  s->DT = 1.0;
  s->SIGMA0 = 173.0;
  s->NINTERM = 10;
  s->NTOT = 1501;
  strcpy(s->OUTPUTDIR, "~masset");
  strcpy(s->INNERBOUNDARY, "WALL");
  strcpy(s->LABELADVECTION, "NO");
  strcpy(s->TRANSPORT, "FAST");
  strcpy(s->PLANETCONFIG, "Systems/SolarSystem.cfg");
  s->MASSTAPER = 0.0000001;
  strcpy(s->RADIALSPACING, "ARITHMETIC");
  s->NRAD = 64;
  s->NSEC = 64;
  s->RMIN = 1.0;
  s->RMAX = 1.0;
  s->THICKNESSSMOOTHING = 0.0;
  s->ROCHESMOOTHING = 0.0;
  s->ASPECTRATIO = 0.05;
  s->VISCOSITY = 0.0;
  s->ALPHAVISCOSITY = 0.0;
  s->SIGMASLOPE = 0.0;
  s->RELEASERADIUS = 0.0;
  s->RELEASEDATE = 0.0;
  s->OMEGAFRAME = 0.0;
  strcpy(s->DISK, "YES");
  strcpy(s->FRAME, "FIXED");
  strcpy(s->OUTERSOURCEMASS, "NO");
  strcpy(s->WRITEDENSITY, "YES");
  strcpy(s->WRITEVELOCITY, "YES");
  strcpy(s->WRITEENERGY, "NO");
  strcpy(s->WRITETEMPERATURE, "NO");
  strcpy(s->WRITEDIVV, "NO");
  strcpy(s->WRITEQPLUS, "NO");
  strcpy(s->INDIRECTTERM, "YES");
  strcpy(s->EXCLUDEHILL, "NO");
  s->IMPOSEDDISKDRIFT = 0.0;
  s->FLARINGINDEX = 0.0;
  s->ECCENTRICITY = 0.0;
  s->CAVITYRADIUS = 0.0;
  s->CAVITYRATIO = 1.0;
  s->CAVITYWIDTH = 1.0;
  s->TRANSITIONRADIUS = 0.0;
  s->TRANSITIONRATIO = 1.0;
  s->TRANSITIONWIDTH = 1.0;
  s->LAMBDADOUBLING = 0.0;
  strcpy(s->SELFGRAVITY, "NO");
  strcpy(s->CICPLANET, "NO");
  strcpy(s->FORCEDCIRCULAR, "NO");
  strcpy(s->ZMPLUS, "NO");
  strcpy(s->ADIABATIC, "NO");
  s->ADIABATICINDEX = 1.4;
  strcpy(s->COOLING, "NO");
  s->COOLINGTIME0 = 6.28;
  s->COOLINGBETA = 0.0;
  /* Interpret.cu:25-31 initial flag values */
  s->FastTransport = YES;
  s->GuidingCenter = NO;
  s->Indirect_Term = YES;
  s->IsDisk = YES;
  s->Write_Density = YES;
  s->Write_Velocity = YES;
  s->ScalingFactor = 1.0;
  s->dimfxy = 11;
}

/* Interpret.cu:36-96 */
static void search_variable(or_sim *s, const char *name, const char *sv) {
  double temp = 0.0;
  sscanf(sv, "%lf", &temp);
#define NUM(key, field, cast)            \
  if (strcmp(name, key) == 0) s->field = (cast)(temp);
#define STR(key, field)                                    \
  if (strcmp(name, key) == 0) {                            \
    strncpy(s->field, sv, sizeof(s->field) - 1);           \
    s->field[sizeof(s->field) - 1] = 0;                    \
  }
  NUM("DT", DT, double)
  NUM("SIGMA0", SIGMA0, double)
  NUM("NINTERM", NINTERM, int)
  NUM("NTOT", NTOT, int)
  STR("OUTPUTDIR", OUTPUTDIR)
  STR("INNERBOUNDARY", INNERBOUNDARY)
  STR("LABELADVECTION", LABELADVECTION)
  STR("TRANSPORT", TRANSPORT)
  STR("PLANETCONFIG", PLANETCONFIG)
  NUM("MASSTAPER", MASSTAPER, double)
  STR("RADIALSPACING", RADIALSPACING)
  NUM("NRAD", NRAD, int)
  NUM("NSEC", NSEC, int)
  NUM("RMIN", RMIN, double)
  NUM("RMAX", RMAX, double)
  NUM("THICKNESSSMOOTHING", THICKNESSSMOOTHING, double)
  NUM("ROCHESMOOTHING", ROCHESMOOTHING, double)
  NUM("ASPECTRATIO", ASPECTRATIO, double)
  NUM("VISCOSITY", VISCOSITY, double)
  NUM("ALPHAVISCOSITY", ALPHAVISCOSITY, double)
  NUM("SIGMASLOPE", SIGMASLOPE, double)
  NUM("RELEASERADIUS", RELEASERADIUS, double)
  NUM("RELEASEDATE", RELEASEDATE, double)
  NUM("OMEGAFRAME", OMEGAFRAME, double)
  STR("DISK", DISK)
  STR("FRAME", FRAME)
  STR("OUTERSOURCEMASS", OUTERSOURCEMASS)
  STR("WRITEDENSITY", WRITEDENSITY)
  STR("WRITEVELOCITY", WRITEVELOCITY)
  STR("WRITEENERGY", WRITEENERGY)
  STR("WRITETEMPERATURE", WRITETEMPERATURE)
  /* Interpret.cu:74 assigns a double to a std::string (one char); the
   * key has no effect on the substep path.  Kept as a string here. */
  STR("WRITEDIVV", WRITEDIVV)
  STR("WRITEQPLUS", WRITEQPLUS)
  STR("INDIRECTTERM", INDIRECTTERM)
  STR("EXCLUDEHILL", EXCLUDEHILL)
  NUM("IMPOSEDDISKDRIFT", IMPOSEDDISKDRIFT, double)
  NUM("FLARINGINDEX", FLARINGINDEX, double)
  NUM("ECCENTRICITY", ECCENTRICITY, double)
  NUM("CAVITYRADIUS", CAVITYRADIUS, double)
  NUM("CAVITYRATIO", CAVITYRATIO, double)
  NUM("CAVITYWIDTH", CAVITYWIDTH, double)
  NUM("TRANSITIONRADIUS", TRANSITIONRADIUS, double)
  NUM("TRANSITIONRATIO", TRANSITIONRATIO, double)
  NUM("TRANSITIONWIDTH", TRANSITIONWIDTH, double)
  NUM("LAMBDADOUBLING", LAMBDADOUBLING, double)
  STR("SELFGRAVITY", SELFGRAVITY)
  STR("CICPLANET", CICPLANET)
  STR("FORCEDCIRCULAR", FORCEDCIRCULAR)
  STR("ZMPLUS", ZMPLUS)
  STR("ADIABATIC", ADIABATIC)
  NUM("ADIABATICINDEX", ADIABATICINDEX, double)
  STR("COOLING", COOLING)
  NUM("COOLINGTIME0", COOLINGTIME0, double)
  NUM("COOLINGBETA", COOLINGBETA, double)
#undef NUM
#undef STR
}

/* Interpret.cu:100-212.  Returns 0 on success, the reference's exit code
 * otherwise (the reference calls exit(1)). */
static int read_variables(or_sim *s, const char *filename, const char *outdir) {
  char nm[300], line[350], stringval[290];
  double temp = 0.0;
  FILE *in = fopen(filename, "r");
  stringval[0] = 0;
  if (!in) {
    set_err("Unable to read %s . Program stopped.", filename);
    return 1;
  }
  while (fgets(line, 349, in) != NULL) {
    int success = sscanf(line, "%299s ", nm);
    if ((nm[0] != '#') && (success == 1)) {
      const char *s1 = line + strlen(nm);
      s1 += strspn(s1, "\t :=>_");
      sscanf(s1, "%lf", &temp);
      sscanf(s1, "%289s ", stringval);
      for (size_t i = 0; i < strlen(nm); i++) nm[i] = (char)toupper((unsigned char)nm[i]);
      search_variable(s, nm, stringval);
    }
  }
  fclose(in);
  (void)temp;
  if (strcmp(s->LABELADVECTION, "YES") == 0) s->AdvecteLabel = YES;
  if (strcmp(s->OUTERSOURCEMASS, "YES") == 0) s->OuterSourceMass = YES;
  if (strcmp(s->TRANSPORT, "STANDARD") == 0) s->FastTransport = NO;
  if (strcmp(s->INNERBOUNDARY, "OPEN") == 0) s->OpenInner = YES;
  if (strcmp(s->INNERBOUNDARY, "NONREFLECTING") == 0) s->NonReflecting = YES;
  if (strcmp(s->INNERBOUNDARY, "EVANESCENT") == 0) s->Evanescent = YES;
  if (strcmp(s->RADIALSPACING, "LOGARITHM") == 0) s->LogGrid = YES;
  if (strcmp(s->DISK, "NO") == 0) s->IsDisk = NO;
  if (strcmp(s->FRAME, "COROTATING") == 0) s->Corotating = YES;
  if (strcmp(s->FRAME, "GUIDING-CENTER") == 0) {
    s->GuidingCenter = YES;
    s->Corotating = YES;
  }
  if (strcmp(s->WRITEVELOCITY, "NO") == 0) s->Write_Velocity = NO;
  if (strcmp(s->WRITEDENSITY, "NO") == 0) s->Write_Density = NO;
  if (strcmp(s->WRITEENERGY, "YES") == 0) s->Write_Energy = YES;
  if (strcmp(s->WRITETEMPERATURE, "YES") == 0) s->Write_Temperature = YES;
  if (strcmp(s->WRITEDIVV, "YES") == 0) s->Write_DivV = YES;
  if (strcmp(s->WRITEQPLUS, "YES") == 0) s->Write_Qplus = YES;
  if (strcmp(s->INDIRECTTERM, "NO") == 0) s->Indirect_Term = NO;
  if (strcmp(s->SELFGRAVITY, "YES") == 0) s->SelfGravity = YES;
  if (strcmp(s->SELFGRAVITY, "Z") == 0) {
    s->SelfGravity = YES;
    s->SGZeroMode = YES;
  }
  if (strcmp(s->ZMPLUS, "YES") == 0) s->ZMPluss = YES;
  if (s->ZMPluss && !s->SGZeroMode) s->ZMPluss = NO;
  if (strcmp(s->ADIABATIC, "YES") == 0) {
    s->Adiabatic = YES;
    s->Write_Temperature = YES;
  }
  if (strcmp(s->COOLING, "YES") == 0) s->Cooling = YES;
  if (s->Adiabatic && s->ADIABATICINDEX == 1.0) s->Adiabatic = NO;
  if (strcmp(s->WRITEENERGY, "NO") == 0) s->Write_Energy = NO;
  if (strcmp(s->EXCLUDEHILL, "YES") == 0) s->ExcludeHill = YES;
  if (strcmp(s->CICPLANET, "YES") == 0) s->CICPlanet = YES;
  if (strcmp(s->FORCEDCIRCULAR, "YES") == 0) s->ForcedCircular = YES;
  if ((s->ALPHAVISCOSITY != 0.0) && (s->VISCOSITY != 0.0)) {
    set_err("You cannot use at the same time VISCOSITY and ALPHAVISCOSITY.");
    return 1;
  }
  if (s->ALPHAVISCOSITY != 0.0) s->ViscosityAlpha = YES;
  if ((s->THICKNESSSMOOTHING != 0.0) && (s->ROCHESMOOTHING != 0.0)) {
    set_err("You cannot use at the same time ThicknessSmoothing and RocheSmoothing.");
    return 1;
  }
  if ((s->THICKNESSSMOOTHING <= 0.0) && (s->ROCHESMOOTHING <= 0.0)) {
    set_err("A non-vanishing potential smoothing length is required.");
    return 1;
  }
  if (s->ROCHESMOOTHING != 0.0) s->RocheSmoothing = YES;
  if (outdir) {
    strncpy(s->OUTPUTDIR, outdir, sizeof(s->OUTPUTDIR) - 1);
    s->OUTPUTDIR[sizeof(s->OUTPUTDIR) - 1] = 0;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* profiles: Viscosity.cu:93-107 / 33-50, Theo.cu                       */
/* ------------------------------------------------------------------ */
static double aspect_ratio(const or_sim *s, double r) {
  double aspectratio = s->ASPECTRATIO;
  double rmin = s->TRANSITIONRADIUS - s->TRANSITIONWIDTH * s->ASPECTRATIO;
  double rmax = s->TRANSITIONRADIUS + s->TRANSITIONWIDTH * s->ASPECTRATIO;
  double scale = 1.0 + (s->PhysicalTime - s->PhysicalTimeInitial) * s->LAMBDADOUBLING;
  rmin *= scale;
  rmax *= scale;
  if (r < rmin) aspectratio *= s->TRANSITIONRATIO;
  if ((r >= rmin) && (r <= rmax)) aspectratio *= exp((rmax - r) / (rmax - rmin) * log(s->TRANSITIONRATIO));
  return aspectratio;
}

/* Viscosity.cu:33-50.  ViscosityAlpha is tested as the double pun of the
 * reference (non-zero whenever alpha viscosity is on). */
static double fviscosity(const or_sim *s, double r) {
  double viscosity = s->VISCOSITY, rmin, rmax, scale;
  int i = 0;
  if (s->ViscosityAlpha) {
    while (s->Rmed[i] < r) i++;
    viscosity = s->ALPHAVISCOSITY * s->GLOBAL_bufarray[i] * s->GLOBAL_bufarray[i] * pow(r, 1.5);
  }
  rmin = s->CAVITYRADIUS - s->CAVITYWIDTH * s->ASPECTRATIO;
  rmax = s->CAVITYRADIUS + s->CAVITYWIDTH * s->ASPECTRATIO;
  scale = 1.0 + (s->PhysicalTime - s->PhysicalTimeInitial) * s->LAMBDADOUBLING;
  rmin *= scale;
  rmax *= scale;
  if (r < rmin) viscosity *= s->CAVITYRATIO;
  if ((r >= rmin) && (r <= rmax)) viscosity *= exp((rmax - r) / (rmax - rmin) * log(s->CAVITYRATIO));
  return viscosity;
}

/* Theo.cu:13-23 */
static double sigma_profile(const or_sim *s, double r) {
  double cavity = 1.0;
  if (r < s->CAVITYRADIUS) cavity = 1.0 / s->CAVITYRATIO;
  return cavity * s->ScalingFactor * s->SIGMA0 * pow(r, -s->SIGMASLOPE);
}
/* Theo.cu:37-50 */
static double energy_profile(const or_sim *s, double r) {
  return RGAS / MU / (s->ADIABATICINDEX - 1.0) * s->SIGMA0 * pow(s->ASPECTRATIO, 2.0) *
         pow(r, -s->SIGMASLOPE - 1.0 + 2.0 * s->FLARINGINDEX);
}
/* Theo.cu:62-66; beta-cooling extension: t = beta * r^1.5 */
static double cooling_time(const or_sim *s, double r) {
  if (s->COOLINGBETA > 0.0) return s->COOLINGBETA * pow(r, 1.5);
  return s->COOLINGTIME0 * pow(r, 2.0 + 2.0 * s->FLARINGINDEX);
}
/* Theo.cu:79-85 */
static double qplus_init(const or_sim *s, double r) {
  double viscosity = fviscosity(s, r);
  return 2.25 * viscosity * s->SIGMA0 * pow(r, -s->SIGMASLOPE - 3.0);
}
/* Force.cu:137-142 */
static double compute_smoothing(const or_sim *s, double r) {
  return s->THICKNESSSMOOTHING * aspect_ratio(s, r) * pow(r, 1.0 + s->FLARINGINDEX);
}

/* ------------------------------------------------------------------ */
/* DeviceReduce (Kernels.cu:253-394): per-thread strided pair sums,     */
/* 32-lane shuffle-down tree, sequential host sum over blocks.          */
/* ------------------------------------------------------------------ */
static long nearest_pow2(long n) {
  long x = 1;
  if (!n) return n;
  while (x < n) x <<= 1;
  return x;
}

double or_device_reduce(const double *in, int n) {
  const int threads = 32;
  int blocks = (int)((nearest_pow2(n) + threads - 1) / threads);
  if (blocks > 1024) blocks = 1024;
  /* IsPow2 (Kernels.cu:358-361) is true only for n == 1, where the
   * reference then reads in[32], past the end of the array.  That read is
   * skipped here (the element counts as absent); DeviceReduce is only
   * called on NRAD*NSEC or NSEC >= 2 values, where nIsPow2 is false. */
  const int nIsPow2 = 0;
  const unsigned gridSize = (unsigned)(threads * 2 * blocks);
  double *bsum = dalloc(blocks);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < blocks; b++) {
    double lane[32];
    for (int t = 0; t < 32; t++) {
      unsigned i = (unsigned)b * threads * 2 + (unsigned)t;
      double mySum = 0.0;
      while (i < (unsigned)n) {
        mySum += in[i];
        if (nIsPow2 || i + threads < (unsigned)n) mySum += in[i + threads];
        i += gridSize;
      }
      lane[t] = mySum;
    }
    for (int offset = 16; offset > 0; offset /= 2) {
      double nxt[32];
      for (int t = 0; t < 32; t++) nxt[t] = lane[t] + ((t + offset < 32) ? lane[t + offset] : lane[t]);
      memcpy(lane, nxt, sizeof lane);
    }
    bsum[b] = lane[0];
  }
  double sum = 0.0;
  for (int b = 0; b < blocks; b++) sum += bsum[b];
  free(bsum);
  return sum;
}

/* ------------------------------------------------------------------ */
/* grid: SourceEuler.cu:66-162                                          */
/* ------------------------------------------------------------------ */
static int fill_polar_1d(or_sim *s) {
  int nr = s->NRAD;
  double drrsep = (s->RMAX - s->RMIN) / (double)nr;
  double *Radii2 = dalloc(nr + 1), *Rmed2 = dalloc(nr + 1);
  for (int i = 0; i <= nr; i++) {
    if (s->LogGrid == YES)
      Radii2[i] = (double)s->RMIN * exp((double)i / (double)nr * log((double)s->RMAX / (double)s->RMIN));
    else
      Radii2[i] = s->RMIN + drrsep * (double)i;
    s->Radii[i] = Radii2[i];
  }
  for (int i = 0; i < nr; i++) {
    s->Rinf[i] = Radii2[i];
    s->Rsup[i] = Radii2[i + 1];
    Rmed2[i] = 2.0 / 3.0 * (Radii2[i + 1] * Radii2[i + 1] * Radii2[i + 1] - Radii2[i] * Radii2[i] * Radii2[i]);
    Rmed2[i] = Rmed2[i] / (Radii2[i + 1] * Radii2[i + 1] - Radii2[i] * Radii2[i]);
    s->Rmed[i] = Rmed2[i];
    s->Surf[i] = PI * (Radii2[i + 1] * Radii2[i + 1] - Radii2[i] * Radii2[i]) / (double)s->NSEC;
    s->invRmed[i] = 1.0 / Rmed2[i];
    s->invSurf[i] = 1.0 / s->Surf[i];
    s->invdiffRsup[i] = 1.0 / (Radii2[i + 1] - Radii2[i]);
    s->invRinf[i] = 1.0 / Radii2[i];
  }
  s->Rinf[nr] = Radii2[nr];
  for (int i = 0; i < nr; i++) {
    if (i > 0) s->invdiffRmed[i] = 1.0 / (Rmed2[i] - Rmed2[i - 1]);
    s->powRmed[i] = pow(Rmed2[i], -2.5 + s->SIGMASLOPE);
  }
  if (s->write_files) {
    char name[1100];
    snprintf(name, sizeof name, "%sused_rad.dat", s->OUTPUTDIR);
    FILE *out = fopen(name, "w");
    if (!out) {
      set_err("Can't write %s.", name);
      free(Radii2);
      free(Rmed2);
      return 1;
    }
    for (int i = 0; i <= nr; i++) fprintf(out, "%.18g\n", Radii2[i]);
    fclose(out);
  }
  free(Radii2);
  free(Rmed2);
  return 0;
}

/* ------------------------------------------------------------------ */
/* planets: Psys.cu                                                     */
/* ------------------------------------------------------------------ */
static int init_planetary_system(or_sim *s, const char *parfile) {
  char line[512], nm[512], test1[512], test2[512];
  FILE *in = fopen(s->PLANETCONFIG, "r");
  int count = 0;
  /* extension: a relative planet file not found from the working
   * directory is looked up beside the deck */
  if (!in && s->PLANETCONFIG[0] != '/') {
    char alt[1400];
    const char *slash = strrchr(parfile, '/');
    if (slash) {
      snprintf(alt, sizeof alt, "%.*s/%s", (int)(slash - parfile), parfile, s->PLANETCONFIG);
      in = fopen(alt, "r");
    }
  }
  if (!in) {
    set_err("Error : can't find '%s'.", s->PLANETCONFIG);
    return 1;
  }
  while (fgets(line, 510, in) != NULL)
    if (isalpha((unsigned char)line[0])) count++;
  rewind(in);
  s->nb = count;
  int cap = count + 1;
  s->px = dalloc(cap);
  s->py = dalloc(cap);
  s->pvx = dalloc(cap);
  s->pvy = dalloc(cap);
  s->pmass = dalloc(cap);
  s->pacc = dalloc(cap);
  s->feeldisk = (int *)calloc(cap, sizeof(int));
  s->feelothers = (int *)calloc(cap, sizeof(int));
  int i = 0;
  nm[0] = 0;
  while (fgets(line, 510, in) != NULL) {
    sscanf(line, "%511s ", nm);
    if (isalpha((unsigned char)line[0])) {
      double dist = 0, mass = 0, accret = 0;
      const char *s1 = line + strlen(nm);
      test1[0] = test2[0] = 0;
      sscanf(s1 + strspn(s1, "\t :=>_"), "%lf %lf %lf %511s %511s", &dist, &mass, &accret, test1, test2);
      if (s->CICPlanet) {
        int j = 0;
        while (s->Rmed[j] < dist) j++;
        dist = s->Radii[j + 1];
      }
      s->pmass[i] = mass;
      int feeldis = YES, feeloth = YES;
      if (tolower((unsigned char)*test1) == 'n') feeldis = NO;
      if (tolower((unsigned char)*test2) == 'n') feeloth = NO;
      s->px[i] = dist * (1.0 + s->ECCENTRICITY);
      s->py[i] = 0.0;
      s->pvy[i] = sqrt(G * (1.0 + mass) / dist) * sqrt((1.0 - s->ECCENTRICITY) / (1.0 + s->ECCENTRICITY));
      s->pvx[i] = -0.0000000001 * s->pvy[i];
      s->pacc[i] = accret;
      s->feeldisk[i] = feeldis;
      s->feelothers[i] = feeloth;
      i++;
    }
  }
  fclose(in);
  if (s->nb > 0) s->HillRadius = s->px[0] * pow(s->pmass[0] / 3., 1. / 3.);
  return 0;
}

/* Psys.cu:174-244 */
static double get_psys_info(or_sim *s, int action) {
  double d1, d2, cross, x, y, vx, vy, m, h, d, Ax, Ay, e, a, E, M;
  double xc, yc, vxc, vyc, omega, arg, PerihelionPA;
  xc = x = s->px[0];
  yc = y = s->py[0];
  vxc = vx = s->pvx[0];
  vyc = vy = s->pvy[0];
  m = s->pmass[0] + 1.;
  h = x * vy - y * vx;
  d = sqrt(x * x + y * y);
  Ax = x * vy * vy - y * vx * vy - G * m * x / d;
  Ay = y * vx * vx - x * vx * vy - G * m * y / d;
  e = sqrt(Ax * Ax + Ay * Ay) / m;
  a = h * h / G / m / (1. - e * e);
  if (e == 0.0)
    arg = 1.0;
  else
    arg = (1.0 - d / a) / e;
  if (fabs(arg) >= 1.0)
    E = PI * (1. - arg / fabs(arg)) / 2.;
  else
    E = acos((1.0 - d / a) / e);
  if ((x * y * (vy * vy - vx * vx) + vx * vy * (x * x - y * y)) < 0) E = -E;
  M = E - e * sin(E);
  PerihelionPA = atan2(Ay, Ax);
  omega = sqrt(m / a / a / a);
  if (s->GuidingCenter == YES) {
    xc = a * cos(M + PerihelionPA);
    yc = a * sin(M + PerihelionPA);
    vxc = -a * omega * sin(M + PerihelionPA);
    vyc = a * omega * cos(M + PerihelionPA);
  }
  if (e < 1e-08) {
    xc = x;
    yc = y;
    vxc = vx;
    vyc = vy;
  }
  (void)vxc;
  (void)vyc;
  switch (action) {
    case MARK:
      s->Xplanet_ps = xc;
      s->Yplanet_ps = yc;
      return 0.;
    case GET:
      x = xc;
      y = yc;
      d2 = sqrt(x * x + y * y);
      d1 = sqrt(s->Xplanet_ps * s->Xplanet_ps + s->Yplanet_ps * s->Yplanet_ps);
      cross = s->Xplanet_ps * y - x * s->Yplanet_ps;
      s->Xplanet_ps = x;
      s->Yplanet_ps = y;
      return asin(cross / (d1 * d2));
    case FREQUENCY:
      return omega;
  }
  return 0.0;
}

/* Psys.cu:248-265 */
static void rotate_psys(or_sim *s, double angle) {
  double sint = sin(angle), cost = cos(angle), xt, yt;
  for (int i = 0; i < s->nb; i++) {
    xt = s->px[i];
    yt = s->py[i];
    s->px[i] = xt * cost + yt * sint;
    s->py[i] = -xt * sint + yt * cost;
    xt = s->pvx[i];
    yt = s->pvy[i];
    s->pvx[i] = xt * cost + yt * sint;
    s->pvy[i] = -xt * sint + yt * cost;
  }
}

/* RungeKutta.cu:30-67 */
static void deriv_motion_rk5(or_sim *s, const double *q_init, const double *masses, double *deriv, int nb,
                             double dt, const int *feelothers) {
  const double *x = q_init, *y = x + nb, *vx = y + nb, *vy = vx + nb;
  double *derivx = deriv, *derivy = derivx + nb, *derivvx = derivy + nb, *derivvy = derivvx + nb;
  double dist;
  for (int i = 0; i < nb; i++) s->Dist[i] = sqrt(x[i] * x[i] + y[i] * y[i]);
  for (int i = 0; i < nb; i++) {
    derivx[i] = vx[i];
    derivy[i] = vy[i];
    derivvx[i] = -G * 1.0 / s->Dist[i] / s->Dist[i] / s->Dist[i] * x[i];
    derivvy[i] = -G * 1.0 / s->Dist[i] / s->Dist[i] / s->Dist[i] * y[i];
    for (int j = 0; j < nb; j++) {
      if (s->Indirect_Term == YES) {
        derivvx[i] -= G * masses[j] / s->Dist[j] / s->Dist[j] / s->Dist[j] * x[j];
        derivvy[i] -= G * masses[j] / s->Dist[j] / s->Dist[j] / s->Dist[j] * y[j];
      }
      if ((j != i) && (feelothers[i] == YES)) {
        dist = (x[i] - x[j]) * (x[i] - x[j]) + (y[i] - y[j]) * (y[i] - y[j]);
        dist = sqrt(dist);
        (void)dist;
        derivvx[i] += G * masses[j] / s->Dist[j] / s->Dist[j] / s->Dist[j] * (x[j] - x[i]);
        derivvy[i] += G * masses[j] / s->Dist[j] / s->Dist[j] / s->Dist[j] * (y[j] - y[i]);
      }
    }
  }
  for (int i = 0; i < 4 * nb; i++) deriv[i] *= dt;
}

/* RungeKutta.cu:69-75 */
static void translate_planet_rk5(or_sim *s, const double *qold, double c1, double c2, double c3, double c4,
                                 double c5, double *qnew, int nb) {
  for (int i = 0; i < 4 * nb; i++)
    qnew[i] = qold[i] + c1 * s->k1[i] + c2 * s->k2[i] + c3 * s->k3[i] + c4 * s->k4[i] + c5 * s->k5[i];
}

/* RungeKutta.cu:8-28 (Cash-Karp 6-stage, 5th-order weights) */
static void runge_kutta(or_sim *s, double *q0, double timestep, double *masses, double *q1, int nb,
                        int *feelothers) {
  for (int i = 0; i < 4 * nb; i++) s->k1[i] = s->k2[i] = s->k3[i] = s->k4[i] = s->k5[i] = s->k6[i] = 0;
  deriv_motion_rk5(s, q0, masses, s->k1, nb, timestep, feelothers);
  translate_planet_rk5(s, q0, 0.2, 0.0, 0.0, 0.0, 0.0, q1, nb);
  deriv_motion_rk5(s, q1, masses, s->k2, nb, timestep, feelothers);
  translate_planet_rk5(s, q0, 0.075, 0.225, 0.0, 0.0, 0.0, q1, nb);
  deriv_motion_rk5(s, q1, masses, s->k3, nb, timestep, feelothers);
  translate_planet_rk5(s, q0, 0.3, -0.9, 1.2, 0.0, 0.0, q1, nb);
  deriv_motion_rk5(s, q1, masses, s->k4, nb, timestep, feelothers);
  translate_planet_rk5(s, q0, -11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0.0, q1, nb);
  deriv_motion_rk5(s, q1, masses, s->k5, nb, timestep, feelothers);
  translate_planet_rk5(s, q0, 1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0,
                       253.0 / 4096.0, q1, nb);
  deriv_motion_rk5(s, q1, masses, s->k6, nb, timestep, feelothers);
  for (int i = 0; i < 4 * nb; i++)
    q1[i] = q0[i] + 37.0 / 378.0 * s->k1[i] + 250.0 / 621.0 * s->k3[i] + 125.0 / 594.0 * s->k4[i] +
            512.0 / 1771.0 * s->k6[i];
}

/* Pframeforce.cu:124-187 */
static void advance_system_rk5(or_sim *s, double dt) {
  int nb = s->nb;
  double dtheta, omega, rdot, x, y, r, new_r, vx, vy, theta, denom;
  if (!s->ForcedCircular) {
    for (int k = 0; k < nb; k++) {
      s->q0[k] = s->px[k];
      s->q0[k + nb] = s->py[k];
      s->q0[k + 2 * nb] = s->pvx[k];
      s->q0[k + 3 * nb] = s->pvy[k];
      s->PlanetMasses[k] = s->pmass[k];
    }
    runge_kutta(s, s->q0, dt, s->PlanetMasses, s->q1, nb, s->feelothers);
  }
  for (int i = 1 - (s->PhysicalTime >= s->RELEASEDATE); i < nb; i++) {
    if (!s->ForcedCircular) {
      s->px[i] = s->q1[i];
      s->py[i] = s->q1[i + nb];
      s->pvx[i] = s->q1[i + 2 * nb];
      s->pvy[i] = s->q1[i + 3 * nb];
    } else {
      x = s->px[i];
      y = s->py[i];
      theta = atan2(y, x);
      vx = s->pvx[i];
      vy = s->pvy[i];
      r = sqrt(x * x + y * y);
      omega = (-y * vx + x * vy) / r / r;
      dtheta = omega * dt;
      s->px[i] = r * cos(theta + dtheta);
      s->py[i] = r * sin(theta + dtheta);
      s->pvx[i] = vx * cos(dtheta + theta) - vy * sin(dtheta + theta);
      s->pvy[i] = vx * sin(dtheta + theta) + vy * cos(dtheta + theta);
    }
  }
  if (s->PhysicalTime < s->RELEASEDATE) {
    x = s->px[0];
    y = s->py[0];
    r = sqrt(x * x + y * y);
    theta = atan2(y, x);
    rdot = (s->RELEASERADIUS - r) / (s->RELEASEDATE - s->PhysicalTime);
    omega = sqrt((1. + s->pmass[0]) / r / r / r);
    new_r = r + rdot * dt;
    denom = r - new_r;
    if (denom != 0.0)
      dtheta = 2. * dt * r * omega / denom * (sqrt(r / new_r) - 1.);
    else
      dtheta = omega * dt;
    vx = rdot;
    vy = new_r * sqrt((1. + s->pmass[0]) / new_r / new_r / new_r);
    s->px[0] = new_r * cos(dtheta + theta);
    s->py[0] = new_r * sin(dtheta + theta);
    s->pvx[0] = vx * cos(dtheta + theta) - vy * sin(dtheta + theta);
    s->pvy[0] = vx * sin(dtheta + theta) + vx * cos(dtheta + theta); /* sic, Pframeforce.cu:184 */
  }
}

/* ------------------------------------------------------------------ */
/* equation of state: Kernels.cu:167-213                               */
/* ------------------------------------------------------------------ */
/* Kernels.cu:167-184 (with AspectRatioDevice, Kernels.cu:1377-1392) */
void or_sound_speed(or_sim *s) {
  const int nr = s->nrad, ns = s->nsec;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      if (!s->Adiabatic) {
        double ar = aspect_ratio(s, s->Rmed[i]);
        s->SoundSpeed[l] = ar * sqrt(G * 1.0 / s->Rmed[i]) * pow(s->Rmed[i], s->FLARINGINDEX);
      } else
        s->SoundSpeed[l] =
            sqrt(s->ADIABATICINDEX * (s->ADIABATICINDEX - 1.0) * s->Energy[l] / s->Dens[l]);
    }
}

/* Kernels.cu:187-200 */
void or_pressure(or_sim *s) {
  const int nr = s->nrad, ns = s->nsec;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      if (!s->Adiabatic)
        s->Pressure[l] = s->Dens[l] * s->SoundSpeed[l] * s->SoundSpeed[l];
      else
        s->Pressure[l] = (s->ADIABATICINDEX - 1.0) * s->Energy[l];
    }
}

/* Kernels.cu:203-213 */
void or_temperature(or_sim *s) {
  const int nr = s->nrad, ns = s->nsec;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      if (!s->Adiabatic)
        s->Temperature[l] = MU / RGAS * s->Pressure[l] / s->Dens[l];
      else
        s->Temperature[l] = MU / RGAS * (s->ADIABATICINDEX - 1.0) * s->Energy[l] / s->Dens[l];
    }
}

/* Kernels.cu:610-645, option 1: ring mean of SoundSpeed into
 * GLOBAL_bufarray (sequential j sum, then / nsec). */
void or_make1d_cs(or_sim *s) {
  const int nr = s->nrad, ns = s->nsec;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++) {
    double sum = 0.0;
    for (int j = 0; j < ns; j++) sum += s->SoundSpeed[(long)i * ns + j];
    s->GLOBAL_bufarray[i] = sum / (double)ns;
  }
}

/* same ring mean applied to SG_Accr: the self-gravity calls of
 * Make1Dprofile(1) (SgMain.cu:169, SourceEuler.cu:445, SideEuler.cu:200)
 * by their evident intent. */
static void make1d_sgaccr(or_sim *s) {
  const int nr = s->nrad, ns = s->nsec;
  for (int i = 0; i < nr; i++) {
    double sum = 0.0;
    for (int j = 0; j < ns; j++) sum += s->SG_Accr[(long)i * ns + j];
    s->GLOBAL_AxiSGAccr[i] = sum / (double)ns;
  }
}

static void fill_viscosity_array(or_sim *s) {
  for (int i = 0; i < s->nrad; i++) s->viscosity_array[i] = fviscosity(s, s->Rmed[i]);
}

/* ------------------------------------------------------------------ */
/* CFL: SourceEuler.cu:520-546, Kernels.cu:1267-1374                    */
/* ------------------------------------------------------------------ */
static double max2(double a, double b) { return (b > a) ? b : a; }
static double min2(double a, double b) { return (b < a) ? b : a; }

int or_condition_cfl(or_sim *s, double DeltaT) {
  const int nr = s->nrad, ns = s->nsec;
  /* ConditionCFLKernel1D */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++) {
    s->Vmoy[i] = 0.0;
    for (int j = 0; j < ns; j++) s->Vmoy[i] += s->Vtheta[(long)i * ns + j];
    s->Vmoy[i] /= (double)ns;
  }
  memset(s->DT2D, 0, sizeof(double) * (size_t)nr * ns);
  memset(s->DT1D, 0, sizeof(double) * (size_t)nr);
  /* ConditionCFLKernel2D1 */
#pragma omp parallel for schedule(static)
  for (int i = 1; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      double dxrad = s->Rsup[i] - s->Rinf[i];
      double dxtheta = s->Rmed[i] * 2.0 * PI / (double)ns;
      if (s->FastTransport)
        s->Vresidual[l] = s->Vtheta[l] - s->Vmoy[i];
      else
        s->Vresidual[l] = s->Vtheta[l];
      double invdt1 = s->SoundSpeed[l] / (min2(dxrad, dxtheta));
      double invdt2 = fabs(s->Vrad[l]) / dxrad;
      double invdt3 = fabs(s->Vresidual[l]) / dxtheta;
      double dvr = s->Vrad[l + ns] - s->Vrad[l];
      double dvt = s->Vtheta[(long)i * ns + (j + 1) % ns] - s->Vtheta[l];
      if (dvr >= 0.0)
        dvr = 1e-10;
      else
        dvr = -dvr;
      if (dvt >= 0.0)
        dvt = 1e-10;
      else
        dvt = -dvt;
      double invdt4 = max2(dvr / dxrad, dvt / dxtheta);
      invdt4 *= 4.0 * CVNR * CVNR;
      s->DT2D[l] = CFLSECURITY / sqrt(invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 + invdt4 * invdt4);
    }
  /* ConditionCFLKernel2D2 */
  for (int i = 0; i < nr; i++) {
    if (i > 0) {
      s->newDT[i] = 1e30;
      for (int k = 0; k < ns; k++)
        if (s->DT2D[(long)i * ns + k] < s->newDT[i]) s->newDT[i] = s->DT2D[(long)i * ns + k];
    }
    if (i < nr - 1)
      s->DT1D[i] = 2.0 * PI * CFLSECURITY / (double)ns /
                   fabs(s->Vmoy[i] * s->invRmed[i] - s->Vmoy[i + 1] * s->invRmed[i + 1]);
  }
  /* ConditionCFLKernel2D3 */
  double newdt = s->newDT[1];
  for (int i = 2; i < nr; i++)
    if (s->newDT[i] < newdt) newdt = s->newDT[i];
  for (int i = 0; i < nr - 1; i++)
    if (s->DT1D[i] < newdt) newdt = s->DT1D[i];
  if (DeltaT < newdt) newdt = DeltaT;
  s->last_cfl = (int)(ceil(DeltaT / newdt));
  return s->last_cfl;
}

/* ------------------------------------------------------------------ */
/* disk forces: Force.cu:102-135, Kernels.cu:693-729, SideEuler.cu:135  */
/* ------------------------------------------------------------------ */
static void compute_force(or_sim *s, double x, double y, double rsmoothing, double mass, int dimfxy, double a,
                          double rh) {
  const int nr = s->nrad, ns = s->nsec;
  const long n = (long)nr * ns;
  (void)mass;
  for (int k = 0; k < dimfxy; k++) {
    memset(s->fxi, 0, sizeof(double) * n);
    memset(s->fxo, 0, sizeof(double) * n);
    memset(s->fyi, 0, sizeof(double) * n);
    memset(s->fyo, 0, sizeof(double) * n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nr; i++)
      for (int j = 0; j < ns; j++) {
        long l = (long)i * ns + j;
        double cellmass = s->Surf[i] * s->Dens[l];
        double dx = s->CellAbscissa[l] - x;
        double dy = s->CellOrdinate[l] - y;
        double d2 = dx * dx + dy * dy;
        double dist2 = d2 + rsmoothing * rsmoothing;
        double distance = sqrt(dist2);
        double InvDist3 = 1.0 / dist2 / distance;
        double hillcutfactor = (double)k / (double)(dimfxy - 1);
        double hill_cut, rhk = rh;
        if (k != 0) {
          rhk *= hillcutfactor;
          hill_cut = 1. - exp(-d2 / (rhk * rhk));
        } else
          hill_cut = 1.;
        if (s->Rmed[i] < a) {
          s->fxi[l] = G * cellmass * dx * InvDist3 * hill_cut;
          s->fyi[l] = G * cellmass * dy * InvDist3 * hill_cut;
        } else {
          s->fxo[l] = G * cellmass * dx * InvDist3 * hill_cut;
          s->fyo[l] = G * cellmass * dy * InvDist3 * hill_cut;
        }
      }
    s->globalforce[k] = or_device_reduce(s->fxi, (int)n);
    s->globalforce[k + dimfxy] = or_device_reduce(s->fxo, (int)n);
    s->globalforce[k + 2 * dimfxy] = or_device_reduce(s->fyi, (int)n);
    s->globalforce[k + 3 * dimfxy] = or_device_reduce(s->fyo, (int)n);
  }
  s->fx_inner = s->globalforce[0];
  s->fx_ex_inner = s->globalforce[dimfxy - 1];
  s->fx_outer = s->globalforce[dimfxy];
  s->fx_ex_outer = s->globalforce[2 * dimfxy - 1];
  s->fy_inner = s->globalforce[2 * dimfxy];
  s->fy_ex_inner = s->globalforce[3 * dimfxy - 1];
  s->fy_outer = s->globalforce[3 * dimfxy];
  s->fy_ex_outer = s->globalforce[4 * dimfxy - 1];
}

/* SideEuler.cu:135-148.  dimfxy passes are identical when a = rh = 0
 * (every hill_cut is 1), so one pass is computed; the k = dimfxy-1 slot
 * equals the k = 0 slot. */
void or_compute_accel(or_sim *s, double x, double y, double rsmoothing, double mass, double *out2) {
  compute_force(s, x, y, rsmoothing, mass, 1, 0.0, 0.0);
  s->fx_ex_inner = s->fx_inner;
  s->fx_ex_outer = s->fx_outer;
  s->fy_ex_inner = s->fy_inner;
  s->fy_ex_outer = s->fy_outer;
  if (s->ExcludeHill) {
    out2[0] = s->fx_ex_inner + s->fx_ex_outer;
    out2[1] = s->fy_ex_inner + s->fy_ex_outer;
  } else {
    out2[0] = s->fx_inner + s->fx_outer;
    out2[1] = s->fy_inner + s->fy_outer;
  }
}

/* Force.cu:102-135 for any dimfxy (1..64): the GlobalForce array */
void or_compute_force_hill(or_sim *s, double x, double y, double rsmoothing, int dimfxy, double a, double rh,
                           double *globalforce) {
  if (dimfxy < 1 || dimfxy > 64) return;
  compute_force(s, x, y, rsmoothing, 0.0, dimfxy, a, rh);
  memcpy(globalforce, s->globalforce, sizeof(double) * 4 * dimfxy);
}

/* CircumPlanetaryMassKernel (Kernels.cu:236-251) + DeviceReduce, as
 * CircumPlanetaryMass (SourceEuler.cu:549-563) with a given centre and
 * radius */
double or_circumplanetary_mass_at(or_sim *s, double xpl, double ypl, double hill) {
  const int nr = s->nrad, ns = s->nsec;
  const long n = (long)nr * ns;
  double *m = dalloc(n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      double dist = sqrt((s->CellAbscissa[l] - xpl) * (s->CellAbscissa[l] - xpl) +
                         (s->CellOrdinate[l] - ypl) * (s->CellOrdinate[l] - ypl));
      m[l] = dist < hill ? s->Surf[i] * s->Dens[l] : 0.0;
    }
  double r = or_device_reduce(m, (int)n);
  free(m);
  return r;
}

/* SourceEuler.cu:549-563: planet 0, HillRadius of Psys.cu:140 */
double or_circumplanetary_mass(or_sim *s) {
  if (s->nb < 1) return 0.0;
  return or_circumplanetary_mass_at(s, s->px[0], s->py[0], s->HillRadius);
}

/* Main.cu:302 (commented out in the reference) */
void or_enable_diagnostics(or_sim *s) {
  s->Diagnostics = 1;
  s->mdcp = or_circumplanetary_mass(s);
}

/* UpdateLog, Force.cu:19-84 (dimfxy = 11, Main.cu:74) */
void or_update_log(or_sim *s, int TimeStep) {
  const int dimfxy = 11;
  for (int i = 0; i < s->nb; i++) {
    double x = s->px[i], y = s->py[i], vx = s->pvx[i], vy = s->pvy[i], m = s->pmass[i];
    double r = sqrt(x * x + y * y);
    double a = sqrt(x * x + y * y);
    double rh = pow(m / 3., 1. / 3.) * a + 1e-15;
    double smoothing = s->RocheSmoothing ? r * pow(m / 3., 1. / 3.) * s->ROCHESMOOTHING : compute_smoothing(s, r);
    compute_force(s, x, y, smoothing, m, dimfxy, a, rh);
    const double *gf = s->globalforce;
    char path[1400];
    snprintf(path, sizeof path, "%stqwk%d.dat", s->OUTPUTDIR, i);
    FILE *f = fopen(path, "a");
    if (!f) continue;
    fprintf(f, "%d\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\n", TimeStep,
            x * s->fy_inner - y * s->fx_inner, x * s->fy_outer - y * s->fx_outer,
            x * s->fy_ex_inner - y * s->fx_ex_inner, x * s->fy_ex_outer - y * s->fx_ex_outer,
            vx * s->fx_inner + vy * s->fy_inner, vx * s->fx_outer + vy * s->fy_outer,
            vx * s->fx_ex_inner + vy * s->fy_ex_inner, vx * s->fx_ex_outer + vy * s->fy_ex_outer, s->PhysicalTime);
    fclose(f);
    if (s->SelfGravity) continue;
    for (int k = 0; k < dimfxy; k++) {
      snprintf(path, sizeof path, "%stqwk%d_%d.dat", s->OUTPUTDIR, i, k);
      f = fopen(path, "a");
      if (!f) continue;
      fprintf(f, "%d\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\n", TimeStep, x * gf[2 * dimfxy + k] - y * gf[k],
              x * gf[3 * dimfxy + k] - y * gf[dimfxy + k], vx * gf[k] + vy * gf[2 * dimfxy + k],
              vx * gf[dimfxy + k] + vy * gf[3 * dimfxy + k], s->PhysicalTime);
      fclose(f);
    }
  }
}

/* WriteBigPlanetSystemFile, Output.cu:62-107 */
void or_write_bigplanet(or_sim *s, int TimeStep) {
  for (int k = 0; k < s->nb; k++) {
    char path[1400];
    snprintf(path, sizeof path, "%sbigplanet%d.dat", s->OUTPUTDIR, k);
    FILE *f = fopen(path, "a");
    if (!f) continue;
    fprintf(f, "%d\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\n", TimeStep, s->px[k],
            s->py[k], s->pvx[k], s->pvy[k], s->pmass[k], s->LostMass, s->PhysicalTime, s->OmegaFrame, s->mdcp1,
            s->exces_mdcp);
    fclose(f);
  }
}

/* Pframeforce.cu:86-95 */
static void compute_indirect_term(or_sim *s) {
  s->IndirectTerm.x = -s->DiskOnPrimaryAcceleration.x;
  s->IndirectTerm.y = -s->DiskOnPrimaryAcceleration.y;
  if (!s->Indirect_Term) {
    s->IndirectTerm.x = 0.0;
    s->IndirectTerm.y = 0.0;
  }
}

/* Pframeforce.cu:56-83 + Kernels.cu:1226-1253 */
void or_fill_forces_arrays(or_sim *s) {
  const int nr = s->nrad, ns = s->nsec;
  compute_indirect_term(s);
  memset(s->Potential, 0, sizeof(double) * (size_t)s->size_grid);
  for (int k = 0; k < s->nb; k++) {
    double xplanet = s->px[k], yplanet = s->py[k];
    double mplanet = s->pmass[k] * s->MassTaper;
    double PlanetDistance = sqrt(xplanet * xplanet + yplanet * yplanet);
    double InvPlanetDistance3 = 1.0 / PlanetDistance / PlanetDistance / PlanetDistance;
    double RRoche = PlanetDistance * pow((1.0 / 3.0 * mplanet), 1.0 / 3.0);
    double smoothing;
    if (s->RocheSmoothing)
      smoothing = RRoche * s->ROCHESMOOTHING;
    else
      smoothing = compute_smoothing(s, PlanetDistance);
    double smooth = smoothing * smoothing;
    pair_t it = s->IndirectTerm;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nr; i++)
      for (int j = 0; j < ns; j++) {
        long l = (long)i * ns + j;
        double InvDistance = 1.0 / s->Rmed[i];
        double angle = (double)j / (double)ns * 2.0 * PI;
        double x = s->Rmed[i] * cos(angle);
        double y = s->Rmed[i] * sin(angle);
        double distance = (x - xplanet) * (x - xplanet) + (y - yplanet) * (y - yplanet);
        double distancesmooth = sqrt(distance + smooth);
        double pot = -G * mplanet / distancesmooth;
        if (s->Indirect_Term == YES) pot += G * mplanet * InvPlanetDistance3 * (x * xplanet + y * yplanet);
        s->Potential[l] += pot;
        if (k == 0) {
          pot = -G * 1.0 * InvDistance;
          pot -= it.x * x + it.y * y;
          s->Potential[l] += pot;
        }
      }
  }
}

/* Pframeforce.cu:97-121 */
static void advance_system_from_disk(or_sim *s, double dt) {
  for (int k = 0; k < s->nb; k++) {
    if (s->feeldisk[k]) {
      double m = s->pmass[k], x = s->px[k], y = s->py[k];
      double r = sqrt(x * x + y * y), smoothing, gamma[2];
      if (s->RocheSmoothing)
        smoothing = r * pow(m / 3., 1. / 3.) * s->ROCHESMOOTHING;
      else
        smoothing = compute_smoothing(s, r);
      or_compute_accel(s, x, y, smoothing, m, gamma);
      s->pvx[k] += dt * gamma[0];
      s->pvy[k] += dt * gamma[1];
      s->pvx[k] += dt * s->IndirectTerm.x;
      s->pvy[k] += dt * s->IndirectTerm.y;
    }
  }
}

/* ------------------------------------------------------------------ */
/* boundaries: SideEuler.cu:31-213, Kernels.cu:482-607, 967-1015        */
/* ------------------------------------------------------------------ */
static void open_boundary(or_sim *s) {
  const int ns = s->nsec;
  for (int j = 0; j < ns; j++) {
    int i = 1;
    s->Dens[(i - 1) * ns + j] = s->Dens[i * ns + j];
    s->Energy[(i - 1) * ns + j] = s->Energy[i * ns + j];
    if (s->Vrad[(i + 1) * ns + j] > 0.0 || (s->Dens[i * ns + j] < s->SigmaMed[0]))
      s->Vrad[i * ns + j] = 0.0;
    else
      s->Vrad[i * ns + j] = s->Vrad[(i + 1) * ns + j];
  }
}

/* SideEuler.cu:45-100.  NonReflectingBoundaryKernel2 runs here over
 * j < NSEC only (the reference launches it without a bound on j). */
static void non_reflecting_boundary(or_sim *s) {
  const int nr = s->nrad, ns = s->nsec;
  double *buf = dalloc(ns);
  double cs0_r, cs1_r, csnrm1_r, csnrm2_r;
  for (int j = 0; j < ns; j++) buf[j] = s->SoundSpeed[j];
  cs0_r = or_device_reduce(buf, ns) / ns;
  for (int j = 0; j < ns; j++) buf[j] = s->SoundSpeed[ns + j];
  cs1_r = or_device_reduce(buf, ns) / ns;
  for (int j = 0; j < ns; j++) buf[j] = s->SoundSpeed[(long)(nr - 1) * ns + j];
  csnrm1_r = or_device_reduce(buf, ns) / ns;
  for (int j = 0; j < ns; j++) buf[j] = s->SoundSpeed[(long)(nr - 2) * ns + j];
  csnrm2_r = or_device_reduce(buf, ns) / ns;

  int i = 1;
  double dangle = (pow(s->Rinf[i], -1.5) - 1.0) / (.5 * (cs0_r + cs1_r));
  dangle *= (s->Rmed[i] - s->Rmed[i - 1]);
  int i_angle = (int)(dangle / 2.0 / PI * (double)ns + .5);
  i = nr - 1;
  double dangle2 = (pow(s->Rinf[i - 1], -1.5) - 1.0) / (.5 * (csnrm1_r + csnrm2_r));
  dangle2 *= (s->Rmed[i] - s->Rmed[i - 1]);
  int i_angle2 = (int)(dangle2 / 2.0 / PI * (double)ns + .5);
  double SigmaMed1 = s->SigmaMed[1], SigmaMed2 = s->SigmaMed[nr - 2];

  /* NonReflectingBoundaryKernel (Kernels.cu:534-573): ring 0 receives
   * ring 1 shifted by i_angle sectors, ring NRAD-1 receives ring NRAD-2
   * shifted by -i_angle2.  The reference wraps once by +-NSEC; a true
   * modulo is used so that shifts beyond one turn stay in the ring. */
  for (int j = 0; j < ns; j++) {
    int ii = 1;
    long dst = ((j + i_angle) % ns + ns) % ns;
    s->Dens[dst] = s->Dens[ii * ns + j];
    s->Energy[dst] = s->Energy[ii * ns + j];
  }
  for (int j = 0; j < ns; j++) {
    int ii = nr - 1;
    long src = (long)(ii - 1) * ns + ((j - i_angle2) % ns + ns) % ns;
    s->Dens[(long)ii * ns + j] = s->Dens[src];
    s->Energy[(long)ii * ns + j] = s->Energy[src];
  }
  /* NonReflectingBoundaryKernel2 */
  for (int j = 0; j < ns; j++) {
    int ii = 1;
    double Vrad_med = -s->SoundSpeed[ii * ns + j] * (s->Dens[ii * ns + j] - SigmaMed1) / SigmaMed1;
    s->Vrad[ii * ns + j] = 2.0 * Vrad_med - s->Vrad[(ii + 1) * ns + j];
    ii = nr - 1;
    Vrad_med = s->SoundSpeed[(long)ii * ns + j] * (s->Dens[(long)(ii - 1) * ns + j] - SigmaMed2) / SigmaMed2;
    s->Vrad[(long)ii * ns + j] = 2. * Vrad_med - s->Vrad[(long)(ii - 1) * ns + j];
  }
  /* ReduceMean + MinusMeanKernel */
  double mean_dens_r, mean_dens_r2, mean_energy_r, mean_energy_r2;
  for (int j = 0; j < ns; j++) buf[j] = s->Dens[j];
  mean_dens_r = or_device_reduce(buf, ns) / ns;
  for (int j = 0; j < ns; j++) buf[j] = s->Dens[(long)(nr - 1) * ns + j];
  mean_dens_r2 = or_device_reduce(buf, ns) / ns;
  for (int j = 0; j < ns; j++) buf[j] = s->Energy[j];
  mean_energy_r = or_device_reduce(buf, ns) / ns;
  for (int j = 0; j < ns; j++) buf[j] = s->Energy[(long)(nr - 1) * ns + j];
  mean_energy_r2 = or_device_reduce(buf, ns) / ns;
  for (int j = 0; j < ns; j++) {
    s->Dens[j] += s->SigmaMed[0] - mean_dens_r;
    s->Energy[j] += s->EnergyMed[0] - mean_energy_r;
    long l = (long)(nr - 1) * ns + j;
    s->Dens[l] += s->SigmaMed[nr - 1] - mean_dens_r2;
    s->Energy[l] += s->EnergyMed[nr - 1] - mean_energy_r2;
  }
  free(buf);
}

/* SideEuler.cu:104-123 + Kernels.cu:967-1015.  With self-gravity the
 * axisymmetric radial acceleration is GLOBAL_AxiSGAccr (the reference
 * passes a device array that is never filled). */
static void evanescent_boundary(or_sim *s, double step) {
  const int nr = s->nrad, ns = s->nsec;
  double Tin = 2.0 * PI * pow(s->Rmed[0], 3. / 2);
  double Tout = 2.0 * PI * pow(s->Rmed[nr - 1], 3. / 2);
  double DRMIN = s->Rmed[0] * 1.25;
  double DRMAX = s->Rmed[nr - 1] * 0.84;
  fill_viscosity_array(s);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++) {
    double Rm = s->Rmed[i];
    if ((Rm < DRMIN) || (Rm > DRMAX)) {
      double damping, lambda = 0.0, vtheta0, vrad0, energy0, dens0, AspectRatio;
      if (Rm < DRMIN) {
        damping = (Rm - DRMIN) / (s->Rmed[0] - DRMIN);
        lambda = damping * damping * 10.0 * step / Tin;
      }
      if (Rm > DRMAX) {
        damping = (Rm - DRMAX) / (s->Rmed[nr - 1] - DRMAX);
        lambda = damping * damping * 10.0 * step / Tout;
      }
      AspectRatio = aspect_ratio(s, Rm);
      if (!s->SelfGravity)
        vtheta0 = sqrt(G * 1.0 / Rm *
                       (1.0 - (1.0 + s->SIGMASLOPE - 2.0 * s->FLARINGINDEX) * pow(AspectRatio, 2.0) *
                                  pow(Rm, 2.0 * s->FLARINGINDEX)));
      else
        vtheta0 = sqrt(G * 1.0 / Rm *
                           (1.0 - (1.0 + s->SIGMASLOPE - 2.0 * s->FLARINGINDEX) * pow(AspectRatio, 2.0) *
                                      pow(Rm, 2.0 * s->FLARINGINDEX)) -
                       Rm * s->GLOBAL_AxiSGAccr[i]);
      vtheta0 -= Rm * s->OmegaFrame;
      vrad0 = -3.0 * s->viscosity_array[i] / Rm * (-s->SIGMASLOPE + .5);
      dens0 = s->SigmaMed[i];
      energy0 = s->EnergyMed[i];
      for (int j = 0; j < ns; j++) {
        long l = (long)i * ns + j;
        s->Vrad[l] = (s->Vrad[l] + lambda * vrad0) / (1.0 + lambda);
        s->Vtheta[l] = (s->Vtheta[l] + lambda * vtheta0) / (1.0 + lambda);
        s->Dens[l] = (s->Dens[l] + lambda * dens0) / (1.0 + lambda);
        if (s->Adiabatic) s->Energy[l] = (s->Energy[l] + lambda * energy0) / (1.0 + lambda);
      }
    }
  }
}

/* SideEuler.cu:31-41 */
void or_apply_boundary(or_sim *s, double dt) {
  if (s->OpenInner == YES) open_boundary(s);
  if (s->NonReflecting == YES) {
    if (s->Adiabatic) or_sound_speed(s);
    non_reflecting_boundary(s);
  }
  if (s->Evanescent == YES) evanescent_boundary(s, dt);
}

/* SideEuler.cu:181-185 + Kernels.cu:1256-1263 */
void or_correct_vtheta(or_sim *s, double domega) {
  const int nr = s->nrad, ns = s->nsec;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) s->Vtheta[(long)i * ns + j] = s->Vtheta[(long)i * ns + j] - domega * s->Rmed[i];
}

/* SideEuler.cu:189-213 */
static void apply_sub_keplerian_boundary(or_sim *s) {
  const int nr = s->nrad, ns = s->nsec;
  double VKepIn, VKepOut, f = 1.0 + s->SIGMASLOPE - 2.0 * s->FLARINGINDEX;
  if (!s->SelfGravity) {
    VKepIn = sqrt(G * 1.0 / s->Rmed[0] *
                  (1.0 - f * pow(aspect_ratio(s, s->Rmed[0]), 2.0) * pow(s->Rmed[0], 2.0 * s->FLARINGINDEX)));
    VKepOut = sqrt(G * 1.0 / s->Rmed[nr - 1] *
                   (1.0 - f * pow(aspect_ratio(s, s->Rmed[nr - 1]), 2.0) *
                              pow(s->Rmed[nr - 1], 2.0 * s->FLARINGINDEX)));
  } else {
    VKepIn = sqrt(G * 1.0 / s->Rmed[0] *
                      (1.0 - f * pow(aspect_ratio(s, s->Rmed[0]), 2.0) * pow(s->Rmed[0], 2.0 * s->FLARINGINDEX)) -
                  s->Rmed[0] * s->GLOBAL_AxiSGAccr[0]);
    VKepOut = sqrt(G * 1.0 / s->Rmed[nr - 1] *
                       (1.0 - f * pow(aspect_ratio(s, s->Rmed[nr - 1]), 2.0) *
                                  pow(s->Rmed[nr - 1], 2.0 * s->FLARINGINDEX)) -
                   s->Rmed[nr - 1] * s->GLOBAL_AxiSGAccr[nr - 1]);
  }
  for (int j = 0; j < ns; j++) {
    s->VthetaInt[j] = VKepIn - s->Rmed[0] * s->OmegaFrame;
    s->VthetaInt[(long)(nr - 1) * ns + j] = VKepOut - s->Rmed[nr - 1] * s->OmegaFrame;
  }
}

/* ------------------------------------------------------------------ */
/* viscosity: Viscosity.cu:25-67, Kernels.cu:118-150, 733-771           */
/* ------------------------------------------------------------------ */
static void viscous_terms_kernel(or_sim *s, const double *Vradial, const double *Vazimutal) {
  const int nr = s->nrad, ns = s->nsec;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      long jp = (long)i * ns + (j + 1) % ns, jm = (long)i * ns + ((j - 1) + ns) % ns;
      double dphi = 2.0 * M_PI / (double)ns;
      double invdphi = 1.0 / dphi;
      double onethird = 1.0 / 3.0;
      s->DRR[l] = (Vradial[l + ns] - Vradial[l]) * s->invdiffRsup[i];
      s->DPP[l] = (Vazimutal[jp] - Vazimutal[l]) * invdphi * s->invRmed[i] +
                  0.5 * (Vradial[l + ns] + Vradial[l]) * s->invRmed[i];
      s->DivergenceVelocity[l] =
          (Vradial[l + ns] * s->Rsup[i] - Vradial[l] * s->Rinf[i]) * s->invdiffRsup[i] * s->invRmed[i];
      s->DivergenceVelocity[l] += (Vazimutal[jp] - Vazimutal[l]) * invdphi * s->invRmed[i];
      if (i > 0)
        s->DRP[l] = 0.5 * (s->Rinf[i] * (Vazimutal[l] * s->invRmed[i] - Vazimutal[l - ns] * s->invRmed[i - 1]) *
                               s->invdiffRmed[i] +
                           (Vradial[l] - Vradial[jm]) * invdphi * s->invRinf[i]);
      s->TAURR[l] = 2.0 * s->Dens[l] * s->viscosity_array[i] * (s->DRR[l] - onethird * s->DivergenceVelocity[l]);
      s->TAUPP[l] = 2.0 * s->Dens[l] * s->viscosity_array[i] * (s->DPP[l] - onethird * s->DivergenceVelocity[l]);
      if (i > 0)
        s->TAURP[l] = 2.0 * 0.25 * (s->Dens[l] + s->Dens[l - ns] + s->Dens[jm] + s->Dens[jm - ns]) *
                      s->viscosity_array[i] * s->DRP[l];
    }
}

/* Viscosity.cu:52-67: use_intermediate selects VradInt/VthetaInt
 * (Substep1) or Vrad/Vtheta (adiabatic Substep3). */
void or_viscous_terms(or_sim *s, int use_intermediate) {
  if (s->ViscosityAlpha) or_make1d_cs(s);
  fill_viscosity_array(s);
  if (use_intermediate)
    viscous_terms_kernel(s, s->VradInt, s->VthetaInt);
  else
    viscous_terms_kernel(s, s->Vrad, s->Vtheta);
}

/* Kernels.cu:118-150 */
static void update_velocities_with_viscosity(or_sim *s, double DeltaT) {
  const int nr = s->nrad, ns = s->nsec;
#pragma omp parallel for schedule(static)
  for (int i = 1; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      long jp = (long)i * ns + (j + 1) % ns, jm = (long)i * ns + ((j - 1) + ns) % ns;
      double dphi = 2.0 * M_PI / (double)ns;
      double invdphi = 1.0 / dphi;
      if (i < nr - 1)
        s->VthetaInt[l] += DeltaT * s->invRmed[i] *
                           ((s->Rsup[i] * s->TAURP[l + ns] - s->Rinf[i] * s->TAURP[l]) * s->invdiffRsup[i] +
                            (s->TAUPP[l] - s->TAUPP[jm]) * invdphi + 0.5 * (s->TAURP[l] + s->TAURP[l + ns])) /
                           (0.5 * (s->Dens[l] + s->Dens[jm]));
      s->VradInt[l] += DeltaT * s->invRinf[i] *
                       ((s->Rmed[i] * s->TAURR[l] - s->Rmed[i - 1] * s->TAURR[l - ns]) * s->invdiffRmed[i] +
                        (s->TAURP[jp] - s->TAURP[l]) * invdphi - 0.5 * (s->TAUPP[l] + s->TAUPP[l - ns])) /
                       (0.5 * (s->Dens[l] + s->Dens[l - ns]));
    }
}

/* ------------------------------------------------------------------ */
/* source steps: SourceEuler.cu:311-375, Kernels.cu:13-115, 409-479     */
/* ------------------------------------------------------------------ */
static void substep1_kernel(or_sim *s, double dt) {
  const int nr = s->nrad, ns = s->nsec;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      long jm = (long)i * ns + ((j - 1) + ns) % ns;
      double gradp, gradphi, vradint, vradint2, supp_torque, dxtheta, invdxtheta, vt2;
      if (i > 0) {
        long jp = (long)i * ns + (j + 1) % ns;
        gradp = (s->Pressure[l] - s->Pressure[l - ns]) * 2.0 / (s->Dens[l] + s->Dens[l - ns]) * s->invdiffRmed[i];
        gradphi = (s->Potential[l] - s->Potential[l - ns]) * s->invdiffRmed[i];
        vt2 = s->Vtheta[l] + s->Vtheta[jp] + s->Vtheta[l - ns] + s->Vtheta[jp - ns];
        vt2 = vt2 / 4.0 + s->OmegaFrame * s->Rinf[i];
        vt2 = vt2 * vt2;
        vradint = -gradp - gradphi;
        vradint2 = vradint + vt2 * s->invRinf[i];
        s->VradInt[l] = s->Vrad[l] + dt * vradint2;
      }
      supp_torque = s->IMPOSEDDISKDRIFT * 0.5 * pow(s->Rmed[i], -2.5 + s->SIGMASLOPE);
      dxtheta = 2.0 * PI / (double)ns * s->Rmed[i];
      invdxtheta = 1.0 / dxtheta;
      gradp = (s->Pressure[l] - s->Pressure[jm]) * 2.0 / (s->Dens[l] + s->Dens[jm]) * invdxtheta;
      gradphi = (s->Potential[l] - s->Potential[jm]) * invdxtheta;
      s->VthetaInt[l] = s->Vtheta[l] - dt * (gradp + gradphi);
      s->VthetaInt[l] += dt * supp_torque;
    }
}

void or_substep1(or_sim *s, double dt) {
  substep1_kernel(s, dt);
  if (s->SelfGravity) {
    memcpy(s->Vradial, s->VradInt, sizeof(double) * (size_t)s->size_grid);
    memcpy(s->Vazimutal, s->VthetaInt, sizeof(double) * (size_t)s->size_grid);
    or_selfgravity(s, dt, YES);
    memcpy(s->VradInt, s->Vradial, sizeof(double) * (size_t)s->size_grid);
    memcpy(s->VthetaInt, s->Vazimutal, sizeof(double) * (size_t)s->size_grid);
  }
  or_viscous_terms(s, 1);
  update_velocities_with_viscosity(s, dt);
  if (!s->Evanescent) apply_sub_keplerian_boundary(s);
}

/* Kernels.cu:409-434 */
void or_substep2(or_sim *s, double dt) {
  const int nr = s->nrad, ns = s->nsec;
  (void)dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      double dv = s->VradInt[l + ns] - s->VradInt[l];
      if (dv < 0.0)
        s->DensInt[l] = CVNR * CVNR * s->Dens[l] * dv * dv;
      else
        s->DensInt[l] = 0.0;
      dv = s->VthetaInt[(long)i * ns + (j + 1) % ns] - s->VthetaInt[l];
      if (dv < 0.0)
        s->TemperInt[l] = CVNR * CVNR * s->Dens[l] * dv * dv;
      else
        s->TemperInt[l] = 0.0;
    }
}

/* Kernels.cu:438-479 ("kernel", launched by host()) */
void or_host_kernel(or_sim *s, double dt) {
  const int nr = s->nrad, ns = s->nsec;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      long jm = (long)i * ns + ((j - 1) + ns) % ns, jp = (long)i * ns + (j + 1) % ns;
      double dens, densint, dxtheta, invdxtheta, dens2, tempint;
      if (i > 0) {
        dens = s->Dens[l] + s->Dens[l - ns];
        densint = s->DensInt[l] - s->DensInt[l - ns];
        s->VradNew[l] = s->VradInt[l] - dt * 2.0 / dens * densint * s->invdiffRmed[i];
      }
      dxtheta = 2.0 * PI / (double)ns * s->Rmed[i];
      invdxtheta = 1.0 / dxtheta;
      dens2 = s->Dens[l] + s->Dens[jm];
      tempint = (s->TemperInt[l] - s->TemperInt[jm]);
      s->VthetaNew[l] = s->VthetaInt[l] - dt * 2.0 / dens2 * tempint * invdxtheta;
      if (s->Adiabatic) {
        s->EnergyInt[l] = s->Energy[l] - dt * s->DensInt[l] * (s->VradInt[l + ns] - s->VradInt[l]) * s->invdiffRsup[i] -
                          dt * s->TemperInt[l] * (s->VthetaInt[jp] - s->VthetaInt[l]) * invdxtheta;
      }
    }
}

/* SourceEuler.cu:492-503 */
void or_actualise_velocities(or_sim *s) {
  memcpy(s->Vrad, s->VradNew, sizeof(double) * (size_t)s->size_grid);
  memcpy(s->Vtheta, s->VthetaNew, sizeof(double) * (size_t)s->size_grid);
}

/* SourceEuler.cu:287-291: ComputeViscousTerms on the updated velocities,
 * Substep3 (Kernels.cu:56-115), ActualiseGasEnergy. */
void or_substep3(or_sim *s, double dt) {
  const int nr = s->nrad, ns = s->nsec;
  or_viscous_terms(s, 0);
  fill_viscosity_array(s);
#pragma omp parallel for schedule(static)
  for (int i = 1; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      if (s->viscosity_array[i] != 0.0) {
        s->Qplus[l] = 0.5 / s->viscosity_array[i] / s->Dens[l] *
                      (s->TAURR[l] * s->TAURR[l] + s->TAURP[l] * s->TAURP[l] + s->TAUPP[l] * s->TAUPP[l]);
        s->Qplus[l] += (2.0 / 9.0) * s->viscosity_array[i] * s->Dens[l] * s->DivergenceVelocity[l] *
                       s->DivergenceVelocity[l];
      } else
        s->Qplus[l] = 0.0;
    }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      double num, den;
      if (i == 0) {
        if (s->viscosity_array[nr - 1] != 0.0)
          s->Qplus[l] = s->Qplus[l + ns] * exp(log(s->Qplus[l + ns] / s->Qplus[l + 2 * ns]) *
                                               log(s->Rmed[0] / s->Rmed[1]) / log(s->Rmed[1] / s->Rmed[2]));
        else
          s->Qplus[l] = 0.0;
      }
      if (!s->Cooling) {
        num = dt * s->Qplus[l] + s->EnergyInt[l];
        den = 1.0 + (s->ADIABATICINDEX - 1.0) * dt * s->DivergenceVelocity[l];
        s->EnergyNew[l] = num / den;
      } else {
        num = s->EnergyMed[i] * dt * s->Dens[l] / s->SigmaMed[i] + s->CoolingTimeMed[i] * s->EnergyInt[l] +
              dt * s->CoolingTimeMed[i] * (s->Qplus[l] - s->QplusMed[i] * s->Dens[l] / s->SigmaMed[i]);
        den = dt + s->CoolingTimeMed[i] + (s->ADIABATICINDEX - 1.0) * dt * s->CoolingTimeMed[i] *
                                              s->DivergenceVelocity[l];
        s->EnergyNew[l] = num / den;
      }
    }
  memcpy(s->Energy, s->EnergyNew, sizeof(double) * (size_t)s->size_grid);
}

/* ------------------------------------------------------------------ */
/* transport: TransportEuler.cu, Kernels.cu:774-838, 1018-1210          */
/* ------------------------------------------------------------------ */
/* Kernels.cu:1018-1025: an element-wise quotient over the flat range the
 * reference's transposed launch covers (all interior cells). */
static void divise_polar_grid(or_sim *s, const double *Qbase, double *Work) {
  const long n = (long)s->nrad * s->nsec;
#pragma omp parallel for schedule(static)
  for (long l = 0; l < n; l++) Work[l] = Qbase[l] / (s->DensInt[l] + 1e-20);
}

/* Kernels.cu:800-838 */
static void star_rad(or_sim *s, const double *Qbase2, double *QStar, double dt) {
  const int nr = s->nrad, ns = s->nsec;
  double *dq = s->dq; /* layout is immaterial; kept row-major */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      if (i == 0 || i == nr - 1)
        dq[l] = 0.0;
      else {
        double dqm = (Qbase2[l] - Qbase2[l - ns]) * s->invdiffRmed[i];
        double dqp = (Qbase2[l + ns] - Qbase2[l]) * s->invdiffRmed[i + 1];
        if (dqp * dqm > 0.0)
          dq[l] = 2.0 * dqp * dqm / (dqp + dqm);
        else
          dq[l] = 0.0;
      }
    }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      if (s->Vrad[l] > 0.0 && i > 0)
        QStar[l] = Qbase2[l - ns] + (s->Rmed[i] - s->Rmed[i - 1] - s->Vrad[l] * dt) * 0.5 * dq[l - ns];
      else
        QStar[l] = Qbase2[l] - (s->Rmed[i + 1] - s->Rmed[i] + s->Vrad[l] * dt) * 0.5 * dq[l];
    }
  for (int j = 0; j < ns; j++) QStar[j] = QStar[(long)ns * nr + j] = 0.0;
}

/* Kernels.cu:1028-1046 */
static void van_leer_radial_kernel(or_sim *s, double *Qbase, double dt) {
  const int nr = s->nrad, ns = s->nsec;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      double dtheta = 2.0 * PI / (double)ns;
      double varq = dt * dtheta * s->Rinf[i] * s->QRStar[l] * s->DensStar[l] * s->Vrad[l];
      varq -= dt * dtheta * s->Rsup[i] * s->QRStar[l + ns] * s->DensStar[l + ns] * s->Vrad[l + ns];
      Qbase[l] += varq * s->invSurf[i];
      if (i == 0 && s->OpenInner) s->LostByDisk[j] = varq;
    }
}

/* TransportEuler.cu:106-156 */
static double van_leer_radial(or_sim *s, double *Qbase, double dt, int ReturnLost) {
  divise_polar_grid(s, Qbase, s->Work);
  star_rad(s, s->Work, s->QRStar, dt);
  van_leer_radial_kernel(s, Qbase, dt);
  if (ReturnLost) return or_device_reduce(s->LostByDisk, s->nsec);
  return 0.0;
}

/* Kernels.cu:1129-1170 */
static void star_theta(or_sim *s, const double *Qbase, const double *Vaz, double *QStar, double dt) {
  const int nr = s->nrad, ns = s->nsec;
  double *dq = s->dq;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++) {
    double dxtheta = 2.0 * PI / (double)ns * s->Rmed[i];
    double invdxtheta = 1.0 / dxtheta;
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      double dqm = (Qbase[l] - Qbase[(long)i * ns + ((j - 1) + ns) % ns]);
      double dqp = (Qbase[(long)i * ns + (j + 1) % ns] - Qbase[l]);
      if (dqp * dqm > 0.0)
        dq[l] = dqp * dqm / (dqp + dqm) * invdxtheta;
      else
        dq[l] = 0.0;
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++) {
    double dxtheta = 2.0 * PI / (double)ns * s->Rmed[i];
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j, lm = (long)i * ns + ((j - 1) + ns) % ns;
      double ksi = Vaz[l] * dt;
      if (ksi > 0.0)
        QStar[l] = Qbase[lm] + (dxtheta - ksi) * dq[lm];
      else
        QStar[l] = Qbase[l] - (dxtheta + ksi) * dq[l];
    }
  }
}

/* TransportEuler.cu:249-257 + Kernels.cu:1049-1066 (UniformTransport is
 * passed as the literal 0, so every ring is advected) */
static void van_leer_theta(or_sim *s, const double *Vaz, double *Qbase, double dt) {
  const int nr = s->nrad, ns = s->nsec;
  divise_polar_grid(s, Qbase, s->Work);
  star_theta(s, s->Work, Vaz, s->QRStar, dt);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j, lp = (long)i * ns + (j + 1) % ns;
      double dxrad = (s->Rsup[i] - s->Rinf[i]) * dt;
      double invsurf = 1.0 / s->Surf[i];
      double varq = dxrad * s->QRStar[l] * s->DensStar[l] * Vaz[l];
      varq -= dxrad * s->QRStar[lp] * s->DensStar[lp] * Vaz[lp];
      Qbase[l] += varq * invsurf;
    }
}

/* TransportEuler.cu:229-245 */
static void quantities_advection(or_sim *s, const double *Vaz, double dt) {
  star_theta(s, s->Dens, Vaz, s->DensStar, dt);
  memcpy(s->DensInt, s->Dens, sizeof(double) * (size_t)s->size_grid);
  van_leer_theta(s, Vaz, s->RadMomP, dt);
  van_leer_theta(s, Vaz, s->RadMomM, dt);
  van_leer_theta(s, Vaz, s->ThetaMomP, dt);
  van_leer_theta(s, Vaz, s->ThetaMomM, dt);
  if (s->Adiabatic) van_leer_theta(s, Vaz, s->Energy, dt);
  van_leer_theta(s, Vaz, s->Dens, dt);
}

/* TransportEuler.cu:210-216 + Kernels.cu:1173-1190 */
static void advect_shift(or_sim *s, double *array) {
  const int nr = s->nrad, ns = s->nsec;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      int ji = j - s->Nshift[i];
      while (ji < 0) ji += ns;
      while (ji >= ns) ji -= ns;
      s->TempShift[(long)i * ns + j] = array[(long)i * ns + ji];
    }
  memcpy(array, s->TempShift, sizeof(double) * (size_t)s->size_grid);
}

/* TransportEuler.cu:30-41 */
void or_transport(or_sim *s, double dt) {
  const int nr = s->nrad, ns = s->nsec;
  /* ComputeLRMomenta, Kernels.cu:774-787 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j, lp = (long)i * ns + (j + 1) % ns;
      s->RadMomP[l] = s->Dens[l] * s->Vrad[l + ns];
      s->RadMomM[l] = s->Dens[l] * s->Vrad[l];
      s->ThetaMomP[l] = s->Dens[l] * (s->Vtheta[lp] + s->Rmed[i] * s->OmegaFrame) * s->Rmed[i];
      s->ThetaMomM[l] = s->Dens[l] * (s->Vtheta[l] + s->Rmed[i] * s->OmegaFrame) * s->Rmed[i];
    }
  /* OneWindRad, TransportEuler.cu:60-79 */
  star_rad(s, s->Dens, s->DensStar, dt);
  memcpy(s->DensInt, s->Dens, sizeof(double) * (size_t)s->size_grid);
  van_leer_radial(s, s->RadMomP, dt, 0);
  van_leer_radial(s, s->RadMomM, dt, 0);
  van_leer_radial(s, s->ThetaMomP, dt, 0);
  van_leer_radial(s, s->ThetaMomM, dt, 0);
  if (s->Adiabatic) van_leer_radial(s, s->Energy, dt, 0);
  s->LostMass += van_leer_radial(s, s->Dens, dt, 1);

  /* OneWindTheta, TransportEuler.cu:177-194 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++) {
    double moy = 0.0;
    for (int j = 0; j < ns; j++) moy += s->Vtheta[(long)i * ns + j];
    s->VMed[i] = moy / (double)ns;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) s->VthetaRes[(long)i * ns + j] = s->Vtheta[(long)i * ns + j] - s->VMed[i];
  /* ComputeConstantResidualKernel, Kernels.cu:1094-1126 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++) {
    double maxfrac = s->FastTransport ? 1.0 : 0.0;
    double invdt = 1.0 / dt;
    double dpinvns = 2.0 * PI / (double)ns;
    double Ntilde = s->VMed[i] * s->invRmed[i] * dt * (double)ns / 2.0 / PI;
    double Nround = floor(Ntilde + 0.5);
    long nitemp = (long)Nround;
    s->Nshift[i] = (int)(long)nitemp;
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      s->Vtheta[l] = (Ntilde - Nround) * s->Rmed[i] * invdt * dpinvns;
      if (maxfrac < 0.5) {
        s->NoSplitAdvection[i] = YES;
        s->VthetaRes[l] += s->Vtheta[l];
        s->Vtheta[l] = 0.0;
      } else
        s->NoSplitAdvection[i] = NO;
    }
  }
  quantities_advection(s, s->VthetaRes, dt);
  quantities_advection(s, s->Vtheta, dt);
  advect_shift(s, s->RadMomP);
  advect_shift(s, s->RadMomM);
  advect_shift(s, s->ThetaMomP);
  advect_shift(s, s->ThetaMomM);
  if (s->Adiabatic) advect_shift(s, s->Energy);
  advect_shift(s, s->Dens);

  /* ComputeVelocities, Kernels.cu:1193-1210 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j, lm = (long)i * ns + ((j - 1) + ns) % ns;
      if (i == 0)
        s->Vrad[l] = 0.0;
      else
        s->Vrad[l] = (s->RadMomP[l - ns] + s->RadMomM[l]) / (s->Dens[l] + s->Dens[l - ns] + 1e-20);
      s->Vtheta[l] = (s->ThetaMomP[lm] + s->ThetaMomM[l]) / (s->Dens[l] + s->Dens[lm] + 1e-15) / s->Rmed[i] -
                     s->Rmed[i] * s->OmegaFrame;
    }
}

/* ------------------------------------------------------------------ */
/* self-gravity: SgMain.cu, Kernels.cu:840-932                          */
/* ------------------------------------------------------------------ */
/* Recursive mixed-radix Cooley-Tukey DFT of length n with stride, for
 * factors 2, 3, 5 (direct DFT for any other prime factor).  Twiddles are
 * cos/sin of 2*pi*k/n evaluated directly. */
static void fft1d_rec(const double *in, double *out, int n, int stride, int sign, double *scratch) {
  if (n == 1) {
    out[0] = in[0];
    out[1] = in[1];
    return;
  }
  int p = 0;
  if (n % 4 == 0 || n % 2 == 0)
    p = 2;
  else if (n % 3 == 0)
    p = 3;
  else if (n % 5 == 0)
    p = 5;
  if (p == 0) { /* direct DFT */
    for (int k = 0; k < n; k++) {
      double re = 0.0, im = 0.0;
      for (int t = 0; t < n; t++) {
        double ang = sign * 2.0 * PI * (double)(((long)k * t) % n) / (double)n;
        double c = cos(ang), sn = sin(ang);
        double xr = in[2 * (long)t * stride], xi = in[2 * (long)t * stride + 1];
        re += xr * c - xi * sn;
        im += xr * sn + xi * c;
      }
      out[2 * k] = re;
      out[2 * k + 1] = im;
    }
    return;
  }
  int m = n / p;
  /* sub-transforms of the p decimated sequences into scratch blocks */
  double *sub = scratch;
  double *next = scratch + 2 * (long)n;
  for (int r = 0; r < p; r++) fft1d_rec(in + 2 * (long)r * stride, sub + 2 * (long)r * m, m, stride * p, sign, next);
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      int kk = k + q * m;
      double re = 0.0, im = 0.0;
      for (int r = 0; r < p; r++) {
        double ang = sign * 2.0 * PI * (double)(((long)r * kk) % n) / (double)n;
        double c = cos(ang), sn = sin(ang);
        double xr = sub[2 * ((long)r * m + k)], xi = sub[2 * ((long)r * m + k) + 1];
        re += xr * c - xi * sn;
        im += xr * sn + xi * c;
      }
      out[2 * kk] = re;
      out[2 * kk + 1] = im;
    }
  }
}

void or_fft2d(double *data, int n1, int n2, int sign) {
  int nmax = n1 > n2 ? n1 : n2;
  /* scratch: each recursion level needs 2n doubles; depth <= log2 n */
  long scr = 2L * nmax * 64;
  double *scratch = dalloc(scr);
  double *line = dalloc(2L * nmax), *res = dalloc(2L * nmax);
  for (int a = 0; a < n1; a++) { /* rows (contiguous, length n2) */
    memcpy(line, data + 2L * a * n2, sizeof(double) * 2 * n2);
    fft1d_rec(line, res, n2, 1, sign, scratch);
    memcpy(data + 2L * a * n2, res, sizeof(double) * 2 * n2);
  }
  for (int b = 0; b < n2; b++) { /* columns (length n1) */
    for (int a = 0; a < n1; a++) {
      line[2 * a] = data[2 * ((long)a * n2 + b)];
      line[2 * a + 1] = data[2 * ((long)a * n2 + b) + 1];
    }
    fft1d_rec(line, res, n1, 1, sign, scratch);
    for (int a = 0; a < n1; a++) {
      data[2 * ((long)a * n2 + b)] = res[2 * a];
      data[2 * ((long)a * n2 + b) + 1] = res[2 * a + 1];
    }
  }
  free(scratch);
  free(line);
  free(res);
}

/* SgMain.cu:49-81 (kernel tables, built once; coshf as in SgMain.cu:68) */
static void sg_compute_kernel(or_sim *s) {
  const int nr = s->nrad, ns = s->nsec;
  double eps = s->SGP_eps;
  for (int i = 0; i < 2 * nr; i++) {
    double u;
    if (i < nr)
      u = log(s->Radii[i] / s->Radii[0]);
    else
      u = -log(s->Radii[2 * nr - i] / s->Radii[0]);
    for (int j = 0; j < ns; j++) {
      double theta = 2.0 * PI * (double)j / (double)ns;
      double Kr = 1.0 + eps * eps - cos(theta) * exp(-u);
      double base = eps * eps * exp(u) + 2.0 * (coshf((float)u) - cos(theta));
      Kr *= pow(base, -1.5);
      double Kt = sin(theta);
      Kt *= pow(base, -1.5);
      s->Kr_aux[(long)i * ns + j] = Kr;
      s->Kt_aux[(long)i * ns + j] = Kt;
    }
  }
}

/* SgMain.cu:23-45.  The kernel spectra are transformed on every call, as
 * in the reference. */
void or_selfgravity(or_sim *s, double dt, int update) {
  const int nr = s->nrad, ns = s->nsec;
  const long n2 = 2L * nr * ns;
  /* ComputeFFTKernel, Kernels.cu:840-868 */
  for (int i = 0; i < 2 * nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      s->SGP_Kr[2 * l] = s->Kr_aux[l];
      s->SGP_Kr[2 * l + 1] = 0.;
      s->SGP_Kt[2 * l] = s->Kt_aux[l];
      s->SGP_Kt[2 * l + 1] = 0.;
      s->SGP_Sr[2 * l + 1] = 0.;
      s->SGP_St[2 * l + 1] = 0.;
      if (i < nr) {
        double var = s->Dens[l] * sqrt(s->Rmed[i] / s->Rmed[0]);
        s->SGP_Sr[2 * l] = var;
        s->SGP_St[2 * l] = var * s->Rmed[i] / s->Rmed[0];
      } else {
        s->SGP_Sr[2 * l] = 0.;
        s->SGP_St[2 * l] = 0.;
      }
    }
  memcpy(s->SGP_Stc, s->SGP_St, sizeof(double) * 2 * n2);
  memcpy(s->SGP_Ktc, s->SGP_Kt, sizeof(double) * 2 * n2);
  memcpy(s->SGP_Src, s->SGP_Sr, sizeof(double) * 2 * n2);
  memcpy(s->SGP_Krc, s->SGP_Kr, sizeof(double) * 2 * n2);
  or_fft2d(s->SGP_Stc, 2 * nr, ns, -1);
  or_fft2d(s->SGP_Ktc, 2 * nr, ns, -1);
  or_fft2d(s->SGP_Src, 2 * nr, ns, -1);
  or_fft2d(s->SGP_Krc, 2 * nr, ns, -1);
  /* ComputeConvolutionKernel, Kernels.cu:871-890 */
  for (long l = 0; l < n2; l++) {
    double ktx = s->SGP_Ktc[2 * l], kty = s->SGP_Ktc[2 * l + 1];
    double stx = s->SGP_Stc[2 * l], sty = s->SGP_Stc[2 * l + 1];
    double krx = s->SGP_Krc[2 * l], kry = s->SGP_Krc[2 * l + 1];
    double srx = s->SGP_Src[2 * l], sry = s->SGP_Src[2 * l + 1];
    s->Gphic[2 * l] = -G * (ktx * stx - kty * sty);
    s->Gphic[2 * l + 1] = -G * (ktx * sty + kty * stx);
    s->Grc[2 * l] = -G * (krx * srx - kry * sry);
    s->Grc[2 * l + 1] = -G * (krx * sry + kry * srx);
  }
  memcpy(s->Gphi, s->Gphic, sizeof(double) * 2 * n2);
  memcpy(s->Gr, s->Grc, sizeof(double) * 2 * n2);
  or_fft2d(s->Gphi, 2 * nr, ns, +1);
  or_fft2d(s->Gr, 2 * nr, ns, +1);
  /* ComputeSgAccKernel, Kernels.cu:893-912 */
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) {
      long l = (long)i * ns + j;
      double divRmed = s->Rmed[i] / s->Rmed[0];
      double normaccr = s->SGP_rstep * s->SGP_tstep / ((double)(2 * nr) * (double)ns);
      double normacct = normaccr;
      normaccr /= sqrt(divRmed);
      normacct /= (divRmed * sqrt(divRmed));
      s->SG_Acct[l] = s->Gphi[2 * l] * normacct;
      s->SG_Accr[l] = s->Gr[2 * l] * normaccr;
      s->SG_Accr[l] += G * s->Dens[l] * s->SGP_rstep * s->SGP_tstep / s->SGP_eps;
    }
  make1d_sgaccr(s);
  if (update) {
    /* Update_sgvelocityKernel, Kernels.cu:915-932; the azimuthal
     * neighbour wraps periodically ((j-1)%nsec reads the previous ring
     * at j = 0 in the reference). */
    for (int i = 0; i < nr; i++)
      for (int j = 0; j < ns; j++) {
        long l = (long)i * ns + j, lm = (long)i * ns + ((j - 1) + ns) % ns;
        if (i > 0)
          s->Vradial[l] += dt * ((s->Rinf[i] - s->Rmed[i - 1]) * s->SG_Accr[l] +
                                 (s->Rmed[i] - s->Rinf[i]) * s->SG_Accr[l - ns]) *
                           s->invdiffRmed[i];
        s->Vazimutal[l] += 0.5 * dt * (s->SG_Acct[l] + s->SG_Acct[lm]);
      }
  }
}

/* ------------------------------------------------------------------ */
/* AlgoGas: SourceEuler.cu:199-307                                      */
/* ------------------------------------------------------------------ */
static void begin_dt(or_sim *s) {
  s->dtemp = 0.0;
  s->gastimestepcfl = 1;
  if (s->Adiabatic) or_sound_speed(s);
  if (s->IsDisk == YES && s->SloppyCFL == YES) s->gastimestepcfl = or_condition_cfl(s, s->DT - s->dtemp);
  s->dt = s->DT / s->gastimestepcfl;
  s->in_dt = 1;
}

int or_substep(or_sim *s) {
  if (!s->in_dt) begin_dt(s);
  double dt = s->dt, OmegaNew, domega;
  s->MassTaper = s->PhysicalTime / (s->MASSTAPER * 2.0 * M_PI);
  s->MassTaper = (s->MassTaper > 1.0 ? 1.0 : pow(sin(s->MassTaper * M_PI / 2.0), 2.0));
  if (s->IsDisk == YES) {
    if (s->SloppyCFL == NO) {
      s->gastimestepcfl = 1;
      s->gastimestepcfl = or_condition_cfl(s, s->DT - s->dtemp);
      dt = (s->DT - s->dtemp) / (double)s->gastimestepcfl;
      s->dt = dt;
    }
    /* AccreteOntoPlanets is a no-op (Planet.cu:43) */
  }
  s->dtemp += dt;
  s->DiskOnPrimaryAcceleration.x = 0.0;
  s->DiskOnPrimaryAcceleration.y = 0.0;
  if (s->Corotating == YES) get_psys_info(s, MARK);
  if (s->IsDisk == YES) {
    double acc[2];
    or_compute_accel(s, 0.0, 0.0, 0.0, 0.0, acc);
    s->DiskOnPrimaryAcceleration.x = acc[0];
    s->DiskOnPrimaryAcceleration.y = acc[1];
    or_fill_forces_arrays(s);
    advance_system_from_disk(s, dt);
  }
  advance_system_rk5(s, dt);
  if (s->Corotating == YES) {
    OmegaNew = get_psys_info(s, GET) / dt;
    domega = OmegaNew - s->OmegaFrame;
    if (s->IsDisk == YES) or_correct_vtheta(s, domega);
    s->OmegaFrame = OmegaNew;
  }
  rotate_psys(s, s->OmegaFrame * dt);
  if (s->IsDisk == YES) {
    or_apply_boundary(s, dt);
    or_pressure(s);
    or_substep1(s, dt);
    or_substep2(s, dt);
    or_host_kernel(s, dt);
    or_actualise_velocities(s);
    or_apply_boundary(s, dt);
    if (s->Adiabatic) or_substep3(s, dt);
    or_transport(s, dt);
    or_apply_boundary(s, dt);
    or_temperature(s);
    if (s->Diagnostics) { /* SourceEuler.cu:297-298 */
      s->mdcp1 = or_circumplanetary_mass(s);
      s->exces_mdcp = s->mdcp1 - s->mdcp;
    }
  }
  s->init_counter++;
  s->PhysicalTime += dt;
  if (!(s->dtemp < 0.999999999 * s->DT)) {
    s->in_dt = 0;
    return 1;
  }
  return 0;
}

void or_algogas(or_sim *s) {
  begin_dt(s);
  while (s->dtemp < 0.999999999 * s->DT) or_substep(s);
  s->in_dt = 0;
}

/* ------------------------------------------------------------------ */
/* outputs: Output.cu                                                   */
/* ------------------------------------------------------------------ */
static void write_disk_polar(or_sim *s, const double *array, const char *name, int number) {
  char dir[1200], path[1300];
  snprintf(dir, sizeof dir, "%s%s", s->OUTPUTDIR, name);
  struct stat st;
  if (stat(dir, &st) == -1) mkdir(dir, 0700);
  snprintf(path, sizeof path, "%s%s/%s%d.dat", s->OUTPUTDIR, name, name, number);
  FILE *f = fopen(path, "w");
  if (!f) return;
  fwrite(array, sizeof(double), (size_t)s->nrad * s->nsec, f);
  fclose(f);
}

void or_write_outputs(or_sim *s, int idx) {
  if (s->IsDisk == YES) {
    if (s->Write_Density == YES) write_disk_polar(s, s->Dens, "dens", idx);
    if (s->Write_Velocity == YES) {
      write_disk_polar(s, s->Vrad, "vrad", idx);
      write_disk_polar(s, s->Vtheta, "vtheta", idx);
    }
    if (s->Write_Energy == YES) write_disk_polar(s, s->Energy, "energy", idx);
    if (s->Write_Temperature == YES) write_disk_polar(s, s->Temperature, "temperature", idx);
    if (s->Write_Qplus == YES) write_disk_polar(s, s->Qplus, "qplus", idx);
  }
  for (int k = 0; k < s->nb; k++) {
    char path[1300];
    snprintf(path, sizeof path, "%splanet%d.dat", s->OUTPUTDIR, k);
    FILE *f = fopen(path, "a");
    if (!f) continue;
    fprintf(f, "%d\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\n", idx, s->px[k], s->py[k],
            s->pvx[k], s->pvy[k], s->pmass[k], s->LostMass, s->PhysicalTime, s->OmegaFrame, s->mdcp, s->exces_mdcp);
    fclose(f);
  }
}

void or_run(or_sim *s, int it0, int it1) {
  for (int i = it0; i <= it1; i++) {
    if (s->write_files && s->Diagnostics) { /* Main.cu:312-316: TimeStep of iteration i-1 */
      int prev = i > 0 ? (i - 1) / s->NINTERM : 0;
      or_write_bigplanet(s, prev);
      or_update_log(s, prev);
    }
    int TimeStep = i / s->NINTERM;
    if (s->write_files && s->NINTERM * TimeStep == i) or_write_outputs(s, TimeStep);
    or_algogas(s);
  }
}

/* ------------------------------------------------------------------ */
/* creation: Main.cu:89-305, SourceEuler.cu:166-195, Pframeforce.cu     */
/* ------------------------------------------------------------------ */
static void alloc_arrays(or_sim *s) {
  long S = s->size_grid, R = s->nrad + 1, N = (long)s->nrad * s->nsec;
  double **ring[] = {&s->Radii,    &s->Rinf,        &s->Rsup,           &s->Rmed,           &s->Surf,
                     &s->invRinf,  &s->invSurf,     &s->invRmed,        &s->invdiffRsup,    &s->invdiffRmed,
                     &s->powRmed,  &s->SigmaMed,    &s->SigmaInf,       &s->EnergyMed,      &s->CoolingTimeMed,
                     &s->QplusMed, &s->viscosity_array, &s->GLOBAL_bufarray, &s->GLOBAL_AxiSGAccr, &s->Vmoy,
                     &s->newDT,    &s->DT1D,        &s->VMed};
  for (size_t k = 0; k < sizeof ring / sizeof ring[0]; k++) *ring[k] = dalloc(R);
  double **grid[] = {&s->Dens,      &s->Vrad,       &s->Vtheta,    &s->Energy,    &s->Pressure,
                     &s->SoundSpeed, &s->Temperature, &s->Potential, &s->VradInt,  &s->VthetaInt,
                     &s->VradNew,   &s->VthetaNew,  &s->EnergyInt, &s->EnergyNew, &s->DensInt,
                     &s->TemperInt, &s->Qplus,      &s->DRR,       &s->DPP,       &s->DRP,
                     &s->DivergenceVelocity, &s->TAURR, &s->TAUPP,  &s->TAURP,     &s->RadMomP,
                     &s->RadMomM,   &s->ThetaMomP,  &s->ThetaMomM, &s->Work,      &s->QRStar,
                     &s->DensStar,  &s->dq,         &s->VthetaRes, &s->TempShift, &s->CellAbscissa,
                     &s->CellOrdinate, &s->Vresidual, &s->Vradial, &s->Vazimutal};
  for (size_t k = 0; k < sizeof grid / sizeof grid[0]; k++) *grid[k] = dalloc(S);
  s->fxi = dalloc(N);
  s->fxo = dalloc(N);
  s->fyi = dalloc(N);
  s->fyo = dalloc(N);
  s->DT2D = dalloc(N);
  s->LostByDisk = dalloc(s->nsec);
  s->Nshift = (int *)calloc((size_t)R, sizeof(int));
  s->NoSplitAdvection = (int *)calloc((size_t)R, sizeof(int));
  if (s->SelfGravity) {
    s->Kr_aux = dalloc(2 * S);
    s->Kt_aux = dalloc(2 * S);
    double **cx[] = {&s->SGP_Kr, &s->SGP_Kt, &s->SGP_Sr,  &s->SGP_St, &s->SGP_Krc, &s->SGP_Ktc,
                     &s->SGP_Src, &s->SGP_Stc, &s->Gr,     &s->Gphi,   &s->Grc,     &s->Gphic};
    for (size_t k = 0; k < sizeof cx / sizeof cx[0]; k++) *cx[k] = dalloc(4 * S);
    s->SG_Accr = dalloc(S);
    s->SG_Acct = dalloc(S);
  }
}

or_sim *or_create(const char *parfile, const char *cli_flags, const char *outdir, int write_files) {
  or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
  set_defaults(s);
  s->write_files = write_files;
  if (cli_flags && strchr(cli_flags, 'c')) s->SloppyCFL = YES;
  if (read_variables(s, parfile, outdir)) {
    free(s);
    return NULL;
  }
  s->nrad = s->NRAD;
  s->nsec = s->NSEC;
  if (s->nrad < 3 || s->nsec < 2) {
    set_err("grid too small (NRAD >= 3, NSEC >= 2 required)");
    free(s);
    return NULL;
  }
  s->size_grid = (long)(s->nrad + 1) * s->nsec;
  alloc_arrays(s);
  if (fill_polar_1d(s)) return NULL;
  if (init_planetary_system(s, parfile)) return NULL;
  /* InitGasDensity (Pframeforce.cu:31-40) + FillSigma (Theo.cu:25-34) */
  for (int i = 0; i < s->nrad; i++) {
    s->SigmaMed[i] = sigma_profile(s, s->Rmed[i]);
    s->SigmaInf[i] = sigma_profile(s, s->Rinf[i]);
  }
  for (int i = 0; i < s->nrad; i++)
    for (int j = 0; j < s->nsec; j++) s->Dens[j + i * s->nsec] = s->SigmaMed[i];
  if (s->Adiabatic) {
    for (int i = 0; i < s->nrad; i++) s->EnergyMed[i] = energy_profile(s, s->Rmed[i]);
    for (int i = 0; i < s->nrad; i++)
      for (int j = 0; j < s->nsec; j++) s->Energy[i * s->nsec + j] = s->EnergyMed[i];
  }
  if (s->SelfGravity) {
    /* Main.cu:268-290 */
    s->SGP_eps = s->THICKNESSSMOOTHING * s->ASPECTRATIO;
    s->SGP_rstep = log(s->Radii[s->nrad] / s->Radii[0]) / (double)s->nrad;
    s->SGP_tstep = 2.0 * PI / (double)s->nsec;
    sg_compute_kernel(s);
    or_selfgravity(s, 0.0, NO);
    /* Init_planetarysys_withSG, SgMain.cu:166-194 */
    for (int k = 0; k < s->nb; k++) {
      double r = s->px[k];
      double dist = (double)(r / (1. + s->ECCENTRICITY));
      int ipl = 0;
      while (s->Rmed[ipl] <= dist) ipl++;
      double ri = s->Rmed[ipl], rip1 = s->Rmed[ipl + 1], dr = rip1 - ri;
      double sgacc = (dist - ri) * s->GLOBAL_AxiSGAccr[ipl + 1] + (rip1 - dist) * s->GLOBAL_AxiSGAccr[ipl];
      sgacc /= dr;
      s->pvy[k] *= (double)sqrt(1. - dist * dist * sgacc);
    }
  }
  s->OmegaFrame = s->OMEGAFRAME;
  if (s->Corotating) s->OmegaFrame = get_psys_info(s, FREQUENCY);
  /* InitEuler (SourceEuler.cu:166-195) */
  for (int i = 0; i < s->nrad; i++)
    for (int j = 0; j < s->nsec; j++) {
      /* InitComputeAccelKernel, Kernels.cu:154-164 */
      s->CellAbscissa[(long)i * s->nsec + j] = s->Rmed[i] * cos((2.0 * PI * (double)j) / (double)s->nsec);
      s->CellOrdinate[(long)i * s->nsec + j] = s->Rmed[i] * sin((2.0 * PI * (double)j) / (double)s->nsec);
    }
  or_sound_speed(s);
  or_pressure(s);
  or_temperature(s);
  /* InitGasVelocities (Pframeforce.cu:189-253) */
  if (s->SelfGravity) {
    /* Init_azimutalvelocity_withSG + Azimutalvelocity_withSGKernel
     * (SourceEuler.cu:441-450, Kernels.cu:935-950) */
    for (int i = 0; i < s->nrad; i++) {
      double invr = 1. / s->Rmed[i];
      double omegakep = sqrt(G * 1.0 * invr * invr * invr);
      double omega = sqrt(omegakep * omegakep *
                              (1.0 - (1. + s->SIGMASLOPE - 2.0 * s->FLARINGINDEX) * pow(s->ASPECTRATIO, 2.0) *
                                         pow(s->Rmed[i], 2.0 * s->FLARINGINDEX)) -
                          invr * s->GLOBAL_AxiSGAccr[i]);
      for (int j = 0; j < s->nsec; j++) s->Vtheta[(long)i * s->nsec + j] = s->Rmed[i] * omega;
    }
  }
  if (s->ViscosityAlpha) or_make1d_cs(s);
  if (s->Cooling) {
    for (int i = 0; i < s->nrad; i++) s->CoolingTimeMed[i] = cooling_time(s, s->Rmed[i]);
    for (int i = 0; i < s->nrad; i++) s->QplusMed[i] = qplus_init(s, s->Rmed[i]);
  }
  /* InitVelocities (Pframeforce.cu:256-270) + InitGasVelocitiesKernel
   * (Kernels.cu:648-689).  The kernel's int ViscosityAlpha argument is
   * the truncation of a denormal double, i.e. 0. */
  fill_viscosity_array(s);
  s->viscosity_array[s->nrad] = fviscosity(s, s->Rmed[s->nrad - 1]);
  for (int i = 0; i <= s->nrad; i++)
    for (int j = 0; j < s->nsec; j++) {
      long l = (long)i * s->nsec + j;
      double r, ri;
      if (i == s->nrad) {
        r = s->Rmed[s->nrad - 1];
        ri = s->Rinf[s->nrad - 1];
      } else {
        r = s->Rmed[i];
        ri = s->Rinf[i];
      }
      if (!s->SelfGravity) {
        double omega = sqrt(G * 1.0 / r / r / r);
        s->Vtheta[l] = omega * r *
                       sqrt(1.0 - pow(s->ASPECTRATIO, 2.0) * pow(r, 2.0 * s->FLARINGINDEX) *
                                      (1. + s->SIGMASLOPE - 2.0 * s->FLARINGINDEX));
      }
      s->Vtheta[l] -= s->OmegaFrame * r;
      if (i == s->nrad)
        s->Vrad[l] = 0.0;
      else {
        s->Vrad[l] = s->IMPOSEDDISKDRIFT * s->SIGMA0 / s->SigmaInf[i] / ri;
        s->Vrad[l] -= 3.0 * s->viscosity_array[i] / r * (-s->SIGMASLOPE + .5);
      }
    }
  for (int j = 0; j < s->nsec; j++) s->Vrad[j] = s->Vrad[(long)s->nrad * s->nsec + j] = 0.0;
  if (s->write_files) {
    /* WriteDim (Output.cu:19-35) and EmptyPlanetSystemFile (Output.cu:37) */
    char path[1300];
    snprintf(path, sizeof path, "%sdims.dat", s->OUTPUTDIR);
    FILE *f = fopen(path, "w");
    if (f) {
      fprintf(f, "%d\t%d\t\t%d\t%d\t%f\t%d\t%d\t%d\n", 0, 0, 0, 0, s->RMAX, s->NTOT / s->NINTERM, s->nrad, s->nsec);
      fclose(f);
    }
    for (int k = 0; k < s->nb; k++) {
      snprintf(path, sizeof path, "%splanet%d.dat", s->OUTPUTDIR, k);
      f = fopen(path, "w");
      if (f) fclose(f);
    }
  }
  s->PhysicalTimeInitial = s->PhysicalTime;
  return s;
}

void or_destroy(or_sim *s) {
  if (!s) return;
  void *ptrs[] = {s->Radii,   s->Rinf,      s->Rsup,      s->Rmed,       s->Surf,        s->invRinf,
                  s->invSurf, s->invRmed,   s->invdiffRsup, s->invdiffRmed, s->powRmed,  s->SigmaMed,
                  s->SigmaInf, s->EnergyMed, s->CoolingTimeMed, s->QplusMed, s->viscosity_array,
                  s->GLOBAL_bufarray, s->GLOBAL_AxiSGAccr, s->Vmoy, s->newDT, s->DT1D, s->VMed,
                  s->Dens,    s->Vrad,      s->Vtheta,    s->Energy,     s->Pressure,    s->SoundSpeed,
                  s->Temperature, s->Potential, s->VradInt, s->VthetaInt, s->VradNew,  s->VthetaNew,
                  s->EnergyInt, s->EnergyNew, s->DensInt,  s->TemperInt,  s->Qplus,       s->DRR,
                  s->DPP,     s->DRP,       s->DivergenceVelocity, s->TAURR, s->TAUPP,   s->TAURP,
                  s->RadMomP, s->RadMomM,   s->ThetaMomP, s->ThetaMomM,  s->Work,        s->QRStar,
                  s->DensStar, s->dq,       s->VthetaRes, s->TempShift,  s->CellAbscissa, s->CellOrdinate,
                  s->Vresidual, s->Vradial, s->Vazimutal, s->fxi,        s->fxo,         s->fyi,
                  s->fyo,     s->DT2D,      s->LostByDisk, s->Nshift,    s->NoSplitAdvection,
                  s->Kr_aux,  s->Kt_aux,    s->SGP_Kr,    s->SGP_Kt,     s->SGP_Sr,      s->SGP_St,
                  s->SGP_Krc, s->SGP_Ktc,   s->SGP_Src,   s->SGP_Stc,    s->Gr,          s->Gphi,
                  s->Grc,     s->Gphic,     s->SG_Accr,   s->SG_Acct,    s->px,          s->py,
                  s->pvx,     s->pvy,       s->pmass,     s->pacc,       s->feeldisk,    s->feelothers};
  for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; k++) free(ptrs[k]);
  free(s);
}

/* ------------------------------------------------------------------ */
/* accessors                                                            */
/* ------------------------------------------------------------------ */
int or_nrad(const or_sim *s) { return s->nrad; }
int or_nsec(const or_sim *s) { return s->nsec; }

typedef struct {
  const char *name;
  size_t off;
  int kind; /* 0 = size_grid, 1 = ring (NRAD+1), 2 = NRAD*NSEC, 3 = NSEC, 4 = 2*size_grid, 5 = complex */
} field_desc;

#define FD(n, k) {#n, offsetof(or_sim, n), k}
#include <stddef.h>
static const field_desc FIELDS[] = {
    FD(Dens, 0),        FD(Vrad, 0),        FD(Vtheta, 0),       FD(Energy, 0),     FD(Pressure, 0),
    FD(SoundSpeed, 0),  FD(Temperature, 0), FD(Potential, 0),    FD(VradInt, 0),    FD(VthetaInt, 0),
    FD(VradNew, 0),     FD(VthetaNew, 0),   FD(EnergyInt, 0),    FD(EnergyNew, 0),  FD(DensInt, 0),
    FD(TemperInt, 0),   FD(Qplus, 0),       FD(DRR, 0),          FD(DPP, 0),        FD(DRP, 0),
    FD(DivergenceVelocity, 0), FD(TAURR, 0), FD(TAUPP, 0),       FD(TAURP, 0),      FD(RadMomP, 0),
    FD(RadMomM, 0),     FD(ThetaMomP, 0),   FD(ThetaMomM, 0),    FD(Work, 0),       FD(QRStar, 0),
    FD(DensStar, 0),    FD(dq, 0),          FD(VthetaRes, 0),    FD(TempShift, 0),  FD(CellAbscissa, 0),
    FD(CellOrdinate, 0), FD(Vresidual, 0),  FD(Vradial, 0),      FD(Vazimutal, 0),  FD(SG_Accr, 0),
    FD(SG_Acct, 0),     FD(Radii, 1),       FD(Rinf, 1),         FD(Rsup, 1),       FD(Rmed, 1),
    FD(Surf, 1),        FD(invRinf, 1),     FD(invSurf, 1),      FD(invRmed, 1),    FD(invdiffRsup, 1),
    FD(invdiffRmed, 1), FD(powRmed, 1),     FD(SigmaMed, 1),     FD(SigmaInf, 1),   FD(EnergyMed, 1),
    FD(CoolingTimeMed, 1), FD(QplusMed, 1), FD(viscosity_array, 1), FD(GLOBAL_bufarray, 1),
    FD(GLOBAL_AxiSGAccr, 1), FD(Vmoy, 1),   FD(newDT, 1),        FD(DT1D, 1),       FD(VMed, 1),
    FD(fxi, 2),         FD(fxo, 2),         FD(fyi, 2),          FD(fyo, 2),        FD(DT2D, 2),
    FD(LostByDisk, 3),  FD(Kr_aux, 4),      FD(Kt_aux, 4),       FD(Gr, 5),         FD(Gphi, 5),
    FD(SGP_Sr, 5),      FD(SGP_St, 5),      FD(SGP_Kr, 5),       FD(SGP_Kt, 5),     FD(Grc, 5),
    FD(Gphic, 5),       FD(SGP_Src, 5),     FD(SGP_Stc, 5),      FD(SGP_Krc, 5),    FD(SGP_Ktc, 5),
};

static const field_desc *find_field(const char *name) {
  for (size_t k = 0; k < sizeof FIELDS / sizeof FIELDS[0]; k++)
    if (strcmp(FIELDS[k].name, name) == 0) return &FIELDS[k];
  return NULL;
}

double *or_field(or_sim *s, const char *name) {
  const field_desc *f = find_field(name);
  if (!f) return NULL;
  return *(double **)((char *)s + f->off);
}

long or_field_len(or_sim *s, const char *name) {
  const field_desc *f = find_field(name);
  if (!f) return -1;
  switch (f->kind) {
    case 0: return s->size_grid;
    case 1: return s->nrad + 1;
    case 2: return (long)s->nrad * s->nsec;
    case 3: return s->nsec;
    case 4: return 2 * s->size_grid;
    case 5: return 4 * s->size_grid;
  }
  return -1;
}

int *or_ifield(or_sim *s, const char *name) {
  if (strcmp(name, "Nshift") == 0) return s->Nshift;
  if (strcmp(name, "NoSplitAdvection") == 0) return s->NoSplitAdvection;
  return NULL;
}

typedef struct {
  const char *name;
  size_t off;
} scalar_desc;
#define SD(n) {#n, offsetof(or_sim, n)}
static const scalar_desc SCALARS[] = {
    SD(PhysicalTime), SD(PhysicalTimeInitial), SD(OmegaFrame), SD(MassTaper), SD(LostMass),
    SD(dt),           SD(dtemp),               SD(DT),         SD(SIGMA0),    SD(ASPECTRATIO),
    SD(VISCOSITY),    SD(ALPHAVISCOSITY),      SD(SIGMASLOPE), SD(FLARINGINDEX), SD(RMIN),
    SD(RMAX),         SD(THICKNESSSMOOTHING),  SD(ROCHESMOOTHING), SD(ADIABATICINDEX), SD(COOLINGTIME0),
    SD(OMEGAFRAME),   SD(MASSTAPER),           SD(IMPOSEDDISKDRIFT), SD(ECCENTRICITY), SD(CAVITYRADIUS),
    SD(CAVITYRATIO),  SD(CAVITYWIDTH),         SD(TRANSITIONRADIUS), SD(TRANSITIONRATIO),
    SD(TRANSITIONWIDTH), SD(LAMBDADOUBLING),   SD(RELEASEDATE), SD(RELEASERADIUS), SD(HillRadius),
    SD(SGP_eps),      SD(SGP_rstep),           SD(SGP_tstep),  SD(COOLINGBETA), SD(ScalingFactor),
    SD(fx_inner),     SD(fx_outer),            SD(fy_inner),   SD(fy_outer),   SD(mdcp),
    SD(mdcp1),        SD(exces_mdcp),
};

double or_get(or_sim *s, const char *name) {
  for (size_t k = 0; k < sizeof SCALARS / sizeof SCALARS[0]; k++)
    if (strcmp(SCALARS[k].name, name) == 0) return *(double *)((char *)s + SCALARS[k].off);
  if (strcmp(name, "DiskOnPrimaryAcceleration.x") == 0) return s->DiskOnPrimaryAcceleration.x;
  if (strcmp(name, "DiskOnPrimaryAcceleration.y") == 0) return s->DiskOnPrimaryAcceleration.y;
  if (strcmp(name, "IndirectTerm.x") == 0) return s->IndirectTerm.x;
  if (strcmp(name, "IndirectTerm.y") == 0) return s->IndirectTerm.y;
  return NAN;
}

int or_set(or_sim *s, const char *name, double v) {
  for (size_t k = 0; k < sizeof SCALARS / sizeof SCALARS[0]; k++)
    if (strcmp(SCALARS[k].name, name) == 0) {
      *(double *)((char *)s + SCALARS[k].off) = v;
      return 0;
    }
  if (strcmp(name, "DiskOnPrimaryAcceleration.x") == 0) { s->DiskOnPrimaryAcceleration.x = v; return 0; }
  if (strcmp(name, "DiskOnPrimaryAcceleration.y") == 0) { s->DiskOnPrimaryAcceleration.y = v; return 0; }
  return -1;
}

int or_iget(or_sim *s, const char *name) {
#define IG(n) if (strcmp(name, #n) == 0) return s->n;
  IG(NRAD) IG(NSEC) IG(NINTERM) IG(NTOT) IG(Adiabatic) IG(Cooling) IG(SelfGravity) IG(ViscosityAlpha)
  IG(Evanescent) IG(OpenInner) IG(NonReflecting) IG(LogGrid) IG(FastTransport) IG(Corotating)
  IG(GuidingCenter) IG(Indirect_Term) IG(IsDisk) IG(ExcludeHill) IG(RocheSmoothing) IG(ForcedCircular)
  IG(CICPlanet) IG(Write_Density) IG(Write_Velocity) IG(Write_Energy) IG(Write_Temperature)
  IG(Write_Qplus) IG(SloppyCFL) IG(last_cfl) IG(gastimestepcfl) IG(in_dt) IG(nb) IG(AdvecteLabel)
  IG(OuterSourceMass) IG(SGZeroMode)
#undef IG
  return -1;
}

int or_nplanets(const or_sim *s) { return s->nb; }
void or_planet(const or_sim *s, int k, double *out) {
  out[0] = s->px[k];
  out[1] = s->py[k];
  out[2] = s->pvx[k];
  out[3] = s->pvy[k];
  out[4] = s->pmass[k];
}
void or_set_planet(or_sim *s, int k, const double *in) {
  s->px[k] = in[0];
  s->py[k] = in[1];
  s->pvx[k] = in[2];
  s->pvy[k] = in[3];
  s->pmass[k] = in[4];
}
