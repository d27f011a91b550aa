// Host side of libfargo_hip.so: deck parser, polar grid, initial state,
// planetary system (FP64), output writers and the AlgoGas driver that
// calls the device entry points of fargo_hip.h.
//
// Built with g++ -O2 -ffp-contract=off: every host-side transcendental
// (pow/exp/log/sin/cos/sqrt of ring and sector positions, planet orbits)
// is evaluated with glibc, and the derived 1D tables are uploaded so the
// kernels never call a device transcendental on these paths.
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <string>
#include <vector>

#include "fargo_hip.h"

// error channel of fg_last_error (fg_context.hip)
void fg_set_error(const char *fmt, ...);
#define host_error fg_set_error
// a planet's disk force queued behind a fused substep, and its result once
// the fused results are in (fg_fused.hip)
fg_status fg_fused_planet_force_async(fg_ctx *c, int k, double x, double y, double rs);
void fg_fused_planet_force(const fg_ctx *c, int k, double out2[2]);

namespace {

constexpr double PI = 3.14159265358979323844;  // Fondam.cuh:3
constexpr double G = 1.0;
constexpr double CFLSECURITY = 0.5;


// ------------------------------------------------------------ deck ----
// Param.cuh:9-62 defaults; Interpret.cu:36-212 semantics.
struct Deck {
  double DT = 1.0, SIGMA0 = 173.0;
  int NINTERM = 10, NTOT = 1501;
  std::string OUTPUTDIR = "~masset", INNERBOUNDARY = "WALL", LABELADVECTION = "NO", TRANSPORT = "FAST",
              PLANETCONFIG = "Systems/SolarSystem.cfg", RADIALSPACING = "ARITHMETIC";
  double MASSTAPER = 0.0000001;
  int NRAD = 64, NSEC = 64;
  double RMIN = 1.0, RMAX = 1.0, THICKNESSSMOOTHING = 0.0, ROCHESMOOTHING = 0.0, ASPECTRATIO = 0.05;
  double VISCOSITY = 0.0, ALPHAVISCOSITY = 0.0, SIGMASLOPE = 0.0, RELEASERADIUS = 0.0, RELEASEDATE = 0.0;
  double OMEGAFRAME = 0.0;
  std::string DISK = "YES", FRAME = "FIXED", OUTERSOURCEMASS = "NO", WRITEDENSITY = "YES", WRITEVELOCITY = "YES",
              WRITEENERGY = "NO", WRITETEMPERATURE = "NO", WRITEDIVV = "NO", WRITEQPLUS = "NO",
              INDIRECTTERM = "YES", EXCLUDEHILL = "NO";
  double IMPOSEDDISKDRIFT = 0.0, FLARINGINDEX = 0.0, ECCENTRICITY = 0.0, CAVITYRADIUS = 0.0, CAVITYRATIO = 1.0,
         CAVITYWIDTH = 1.0, TRANSITIONRADIUS = 0.0, TRANSITIONRATIO = 1.0, TRANSITIONWIDTH = 1.0,
         LAMBDADOUBLING = 0.0;
  std::string SELFGRAVITY = "NO", CICPLANET = "NO", FORCEDCIRCULAR = "NO", ZMPLUS = "NO", ADIABATIC = "NO",
              COOLING = "NO";
  double ADIABATICINDEX = 1.4, COOLINGTIME0 = 6.28;
  double COOLINGBETA = 0.0;  // extension: beta-cooling, t_cool = beta / Omega_K

  // derived flags, Interpret.cu:22-31 initial values
  int ExcludeHill = 0, ViscosityAlpha = 0, RocheSmoothing = 0, OpenInner = 0, AdvecteLabel = 0, LogGrid = 0;
  int FastTransport = 1, GuidingCenter = 0, Indirect_Term = 1, IsDisk = 1, NonReflecting = 0, Corotating = 0;
  int OuterSourceMass = 0, Evanescent = 0, Write_Density = 1, Write_Velocity = 1, Write_Energy = 0;
  int Write_Temperature = 0, Write_DivV = 0, Write_Qplus = 0, SelfGravity = 0, SGZeroMode = 0, ZMPluss = 0;
  int Adiabatic = 0, Cooling = 0, CICPlanet = 0, ForcedCircular = 0;
  int SloppyCFL = 0;
  double ScalingFactor = 1.0;

  void set(const char *name, const char *sv) {
    double temp = 0.0;
    sscanf(sv, "%lf", &temp);
    struct NumKey {
      const char *k;
      double *d;
      int *i;
    };
    const NumKey nums[] = {{"DT", &DT, nullptr},
                           {"SIGMA0", &SIGMA0, nullptr},
                           {"NINTERM", nullptr, &NINTERM},
                           {"NTOT", nullptr, &NTOT},
                           {"MASSTAPER", &MASSTAPER, nullptr},
                           {"NRAD", nullptr, &NRAD},
                           {"NSEC", nullptr, &NSEC},
                           {"RMIN", &RMIN, nullptr},
                           {"RMAX", &RMAX, nullptr},
                           {"THICKNESSSMOOTHING", &THICKNESSSMOOTHING, nullptr},
                           {"ROCHESMOOTHING", &ROCHESMOOTHING, nullptr},
                           {"ASPECTRATIO", &ASPECTRATIO, nullptr},
                           {"VISCOSITY", &VISCOSITY, nullptr},
                           {"ALPHAVISCOSITY", &ALPHAVISCOSITY, nullptr},
                           {"SIGMASLOPE", &SIGMASLOPE, nullptr},
                           {"RELEASERADIUS", &RELEASERADIUS, nullptr},
                           {"RELEASEDATE", &RELEASEDATE, nullptr},
                           {"OMEGAFRAME", &OMEGAFRAME, nullptr},
                           {"IMPOSEDDISKDRIFT", &IMPOSEDDISKDRIFT, nullptr},
                           {"FLARINGINDEX", &FLARINGINDEX, nullptr},
                           {"ECCENTRICITY", &ECCENTRICITY, nullptr},
                           {"CAVITYRADIUS", &CAVITYRADIUS, nullptr},
                           {"CAVITYRATIO", &CAVITYRATIO, nullptr},
                           {"CAVITYWIDTH", &CAVITYWIDTH, nullptr},
                           {"TRANSITIONRADIUS", &TRANSITIONRADIUS, nullptr},
                           {"TRANSITIONRATIO", &TRANSITIONRATIO, nullptr},
                           {"TRANSITIONWIDTH", &TRANSITIONWIDTH, nullptr},
                           {"LAMBDADOUBLING", &LAMBDADOUBLING, nullptr},
                           {"ADIABATICINDEX", &ADIABATICINDEX, nullptr},
                           {"COOLINGTIME0", &COOLINGTIME0, nullptr},
                           {"COOLINGBETA", &COOLINGBETA, nullptr}};
    for (const NumKey &k : nums)
      if (strcmp(name, k.k) == 0) {
        if (k.d) *k.d = temp;
        if (k.i) *k.i = (int)temp;
      }
    struct StrKey {
      const char *k;
      std::string *s;
    };
    const StrKey strs[] = {{"OUTPUTDIR", &OUTPUTDIR},         {"INNERBOUNDARY", &INNERBOUNDARY},
                           {"LABELADVECTION", &LABELADVECTION}, {"TRANSPORT", &TRANSPORT},
                           {"PLANETCONFIG", &PLANETCONFIG},   {"RADIALSPACING", &RADIALSPACING},
                           {"DISK", &DISK},                   {"FRAME", &FRAME},
                           {"OUTERSOURCEMASS", &OUTERSOURCEMASS}, {"WRITEDENSITY", &WRITEDENSITY},
                           {"WRITEVELOCITY", &WRITEVELOCITY}, {"WRITEENERGY", &WRITEENERGY},
                           {"WRITETEMPERATURE", &WRITETEMPERATURE}, {"WRITEDIVV", &WRITEDIVV},
                           {"WRITEQPLUS", &WRITEQPLUS},       {"INDIRECTTERM", &INDIRECTTERM},
                           {"EXCLUDEHILL", &EXCLUDEHILL},     {"SELFGRAVITY", &SELFGRAVITY},
                           {"CICPLANET", &CICPLANET},         {"FORCEDCIRCULAR", &FORCEDCIRCULAR},
                           {"ZMPLUS", &ZMPLUS},               {"ADIABATIC", &ADIABATIC},
                           {"COOLING", &COOLING}};
    for (const StrKey &k : strs)
      if (strcmp(name, k.k) == 0) *k.s = sv;
  }

  // ReadVariables, Interpret.cu:100-212; returns FG_OK or FG_E_PARAM/IO
  fg_status read(const char *filename, const char *outdir) {
    FILE *in = fopen(filename, "r");
    if (!in) {
      host_error("Unable to read %s . Program stopped.", filename);
      return FG_E_IO;
    }
    char line[350], nm[300], stringval[290] = "";
    while (fgets(line, 349, in) != NULL) {
      int success = sscanf(line, "%299s ", nm);
      if (success == 1 && nm[0] != '#') {
        const char *s1 = line + strlen(nm);
        s1 += strspn(s1, "\t :=>_");
        sscanf(s1, "%289s ", stringval);
        for (char *q = nm; *q; q++) *q = (char)toupper((unsigned char)*q);
        set(nm, stringval);
      }
    }
    fclose(in);
    if (LABELADVECTION == "YES") AdvecteLabel = 1;
    if (OUTERSOURCEMASS == "YES") OuterSourceMass = 1;
    if (TRANSPORT == "STANDARD") FastTransport = 0;
    if (INNERBOUNDARY == "OPEN") OpenInner = 1;
    if (INNERBOUNDARY == "NONREFLECTING") NonReflecting = 1;
    if (INNERBOUNDARY == "EVANESCENT") Evanescent = 1;
    if (RADIALSPACING == "LOGARITHM") LogGrid = 1;
    if (DISK == "NO") IsDisk = 0;
    if (FRAME == "COROTATING") Corotating = 1;
    if (FRAME == "GUIDING-CENTER") GuidingCenter = Corotating = 1;
    if (WRITEVELOCITY == "NO") Write_Velocity = 0;
    if (WRITEDENSITY == "NO") Write_Density = 0;
    if (WRITEENERGY == "YES") Write_Energy = 1;
    if (WRITETEMPERATURE == "YES") Write_Temperature = 1;
    if (WRITEDIVV == "YES") Write_DivV = 1;
    if (WRITEQPLUS == "YES") Write_Qplus = 1;
    if (INDIRECTTERM == "NO") Indirect_Term = 0;
    if (SELFGRAVITY == "YES") SelfGravity = 1;
    if (SELFGRAVITY == "Z") SelfGravity = SGZeroMode = 1;
    if (ZMPLUS == "YES") ZMPluss = 1;
    if (ZMPluss && !SGZeroMode) ZMPluss = 0;
    if (ADIABATIC == "YES") Adiabatic = Write_Temperature = 1;
    if (COOLING == "YES") Cooling = 1;
    if (Adiabatic && ADIABATICINDEX == 1.0) Adiabatic = 0;
    if (WRITEENERGY == "NO") Write_Energy = 0;
    if (EXCLUDEHILL == "YES") ExcludeHill = 1;
    if (CICPLANET == "YES") CICPlanet = 1;
    if (FORCEDCIRCULAR == "YES") ForcedCircular = 1;
    if (ALPHAVISCOSITY != 0.0 && VISCOSITY != 0.0) {
      host_error("You cannot use at the same time VISCOSITY and ALPHAVISCOSITY.");
      return FG_E_PARAM;
    }
    if (ALPHAVISCOSITY != 0.0) ViscosityAlpha = 1;
    if (THICKNESSSMOOTHING != 0.0 && ROCHESMOOTHING != 0.0) {
      host_error("You cannot use at the same time ThicknessSmoothing and RocheSmoothing.");
      return FG_E_PARAM;
    }
    if (THICKNESSSMOOTHING <= 0.0 && ROCHESMOOTHING <= 0.0) {
      host_error("A non-vanishing potential smoothing length is required.");
      return FG_E_PARAM;
    }
    if (ROCHESMOOTHING != 0.0) RocheSmoothing = 1;
    if (outdir) OUTPUTDIR = outdir;
    return FG_OK;
  }
};

// -------------------------------------------------------- profiles ----
struct Profiles {
  const Deck *d;
  double time = 0.0, time0 = 0.0;  // PhysicalTime, PhysicalTimeInitial
  // AspectRatioHost, Viscosity.cu:93-107
  double aspect_ratio(double r) const {
    double aspectratio = d->ASPECTRATIO;
    double rmin = d->TRANSITIONRADIUS - d->TRANSITIONWIDTH * d->ASPECTRATIO;
    double rmax = d->TRANSITIONRADIUS + d->TRANSITIONWIDTH * d->ASPECTRATIO;
    double scale = 1.0 + (time - time0) * d->LAMBDADOUBLING;
    rmin *= scale;
    rmax *= scale;
    if (r < rmin) aspectratio *= d->TRANSITIONRATIO;
    if ((r >= rmin) && (r <= rmax)) aspectratio *= exp((rmax - r) / (rmax - rmin) * log(d->TRANSITIONRATIO));
    return aspectratio;
  }
  // cavity factor of FViscosity (Viscosity.cu:42-47): 0 = none
  int cavity(double r, double *factor) const {
    double rmin = d->CAVITYRADIUS - d->CAVITYWIDTH * d->ASPECTRATIO;
    double rmax = d->CAVITYRADIUS + d->CAVITYWIDTH * d->ASPECTRATIO;
    double scale = 1.0 + (time - time0) * d->LAMBDADOUBLING;
    rmin *= scale;
    rmax *= scale;
    if (r < rmin) {
      *factor = d->CAVITYRATIO;
      return 1;
    }
    if ((r >= rmin) && (r <= rmax)) {
      *factor = exp((rmax - r) / (rmax - rmin) * log(d->CAVITYRATIO));
      return 1;
    }
    *factor = 1.0;
    return 0;
  }
  // FViscosity with the ring mean of c_s in cs_mean (index found as in
  // Viscosity.cu:39)
  double viscosity(double r, const std::vector<double> &Rmed, const std::vector<double> &cs_mean) const {
    double viscosity = d->VISCOSITY;
    if (d->ViscosityAlpha) {
      int i = 0;
      while (Rmed[i] < r) i++;
      viscosity = d->ALPHAVISCOSITY * cs_mean[i] * cs_mean[i] * pow(r, 1.5);
    }
    double f;
    if (cavity(r, &f)) viscosity *= f;
    return viscosity;
  }
  double sigma(double r) const {  // Theo.cu:13-23
    double cav = 1.0;
    if (r < d->CAVITYRADIUS) cav = 1.0 / d->CAVITYRATIO;
    return cav * d->ScalingFactor * d->SIGMA0 * pow(r, -d->SIGMASLOPE);
  }
  double energy(double r) const {  // Theo.cu:37-50 (R/MU = 1)
    return 1.0 / 1.0 / (d->ADIABATICINDEX - 1.0) * d->SIGMA0 * pow(d->ASPECTRATIO, 2.0) *
           pow(r, -d->SIGMASLOPE - 1.0 + 2.0 * d->FLARINGINDEX);
  }
  double cooling_time(double r) const {  // Theo.cu:62-66; beta-cooling extension
    if (d->COOLINGBETA > 0.0) return d->COOLINGBETA * pow(r, 1.5);
    return d->COOLINGTIME0 * pow(r, 2.0 + 2.0 * d->FLARINGINDEX);
  }
  double smoothing(double r) const {  // Compute_smoothing, Force.cu:137-142
    return d->THICKNESSSMOOTHING * aspect_ratio(r) * pow(r, 1.0 + d->FLARINGINDEX);
  }
};

// ---------------------------------------------------------- planets ---
struct Planets {
  int nb = 0;
  std::vector<double> x, y, vx, vy, mass, acc;
  std::vector<int> feeldisk, feelothers;
  double Xplanet = 0.0, Yplanet = 0.0;  // Psys.cu:4
  // RungeKutta.cu statics
  double k1[400], k2[400], k3[400], k4[400], k5[400], k6[400], Dist[100];
  double q0[400], q1[400], masses[100];
};

// --------------------------------------------------------------- sim ---
}  // namespace

struct fg_sim {
  Deck d;
  Profiles prof;
  Planets ps;
  fg_sim_options opt;
  fg_ctx *ctx = nullptr;
  int nr = 0, ns = 0;
  long S = 0;
  // FillPolar1DArrays outputs
  std::vector<double> Radii, Rinf, Rsup, Rmed, Surf, invRinf, invSurf, invRmed, invdiffRsup, invdiffRmed;
  std::vector<double> SigmaMed, SigmaInf, EnergyMed, CoolingTimeMed, QplusMed, cs_mean;
  // scalars of the run
  double PhysicalTime = 0.0, PhysicalTimeInitial = 0.0, OmegaFrame = 0.0, MassTaper = 0.0, LostMass = 0.0;
  double HillRadius = 0.0;
  double mdcp = 0.0, mdcp1 = 0.0, exces_mdcp = 0.0;  // circumplanetary mass (diagnostics)
  double dt = 0.0, dtemp = 0.0;
  int in_dt = 0, last_cfl = 0;
  double IndirectX = 0.0, IndirectY = 0.0;
  int fused_active = 0;   // substeps run through fg_step_fused
  long nsub = 0;          // substeps taken since creation
  int planets_fresh = 0;  // planetK.dat emptied (first output) or kept (restart)
  int results_ready = 0;  // fused CFL count / star force are current
  int pf_ready = 0;       // the planets' disk forces for the next fused substep are in the result block
  // bumped by every change of the state outside a batch launch (set_state,
  // restart, a single-disk substep); a batch re-derives the CFL count and
  // star force of a disk whose generation it has not seen
  unsigned long state_gen = 0;
  // a batch found this disk's state blown up (FG_E_CRASH) and stopped
  // advancing it; a new state (fg_sim_set_state, fg_sim_restart) clears it
  int crashed = 0;
};

namespace {

// FillPolar1DArrays, SourceEuler.cu:66-162 (radii.dat is not read)
fg_status fill_polar(fg_sim *s) {
  const Deck &d = s->d;
  const int nr = d.NRAD;
  const size_t R = (size_t)nr + 1;
  for (auto *v : {&s->Radii, &s->Rinf, &s->Rsup, &s->Rmed, &s->Surf, &s->invRinf, &s->invSurf, &s->invRmed,
                  &s->invdiffRsup, &s->invdiffRmed, &s->SigmaMed, &s->SigmaInf, &s->EnergyMed,
                  &s->CoolingTimeMed, &s->QplusMed, &s->cs_mean})
    v->assign(R, 0.0);
  double drrsep = (d.RMAX - d.RMIN) / (double)nr;
  std::vector<double> &r = s->Radii;
  for (int i = 0; i <= nr; i++) {
    if (d.LogGrid)
      r[i] = (double)d.RMIN * exp((double)i / (double)nr * log((double)d.RMAX / (double)d.RMIN));
    else
      r[i] = d.RMIN + drrsep * (double)i;
  }
  for (int i = 0; i < nr; i++) {
    s->Rinf[i] = r[i];
    s->Rsup[i] = r[i + 1];
    double rm = 2.0 / 3.0 * (r[i + 1] * r[i + 1] * r[i + 1] - r[i] * r[i] * r[i]);
    rm = rm / (r[i + 1] * r[i + 1] - r[i] * r[i]);
    s->Rmed[i] = rm;
    s->Surf[i] = PI * (r[i + 1] * r[i + 1] - r[i] * r[i]) / (double)d.NSEC;
    s->invRmed[i] = 1.0 / rm;
    s->invSurf[i] = 1.0 / s->Surf[i];
    s->invdiffRsup[i] = 1.0 / (r[i + 1] - r[i]);
    s->invRinf[i] = 1.0 / r[i];
  }
  s->Rinf[nr] = r[nr];
  for (int i = 1; i < nr; i++) s->invdiffRmed[i] = 1.0 / (s->Rmed[i] - s->Rmed[i - 1]);
  if (s->opt.write_files) {
    std::string name = d.OUTPUTDIR + "used_rad.dat";
    FILE *out = fopen(name.c_str(), "w");
    if (!out) {
      host_error("Can't write %s.", name.c_str());
      return FG_E_IO;
    }
    for (int i = 0; i <= nr; i++) fprintf(out, "%.18g\n", r[i]);
    fclose(out);
  }
  return FG_OK;
}

// InitPlanetarySystem, Psys.cu:93-142
fg_status init_planets(fg_sim *s, const char *parfile) {
  const Deck &d = s->d;
  FILE *in = fopen(d.PLANETCONFIG.c_str(), "r");
  if (!in && !d.PLANETCONFIG.empty() && d.PLANETCONFIG[0] != '/') {
    // a relative planet file not found from the working directory is
    // looked up beside the deck
    std::string deck(parfile);
    size_t slash = deck.rfind('/');
    if (slash != std::string::npos) in = fopen((deck.substr(0, slash) + "/" + d.PLANETCONFIG).c_str(), "r");
  }
  if (!in) {
    host_error("Error : can't find '%s'.", d.PLANETCONFIG.c_str());
    return FG_E_IO;
  }
  char line[512], nm[512] = "", t1[512], t2[512];
  int count = 0;
  while (fgets(line, 510, in)) count += isalpha((unsigned char)line[0]) ? 1 : 0;
  rewind(in);
  Planets &p = s->ps;
  p.nb = count;
  for (auto *v : {&p.x, &p.y, &p.vx, &p.vy, &p.mass, &p.acc}) v->assign(count + 1, 0.0);
  p.feeldisk.assign(count + 1, 1);
  p.feelothers.assign(count + 1, 1);
  int i = 0;
  while (fgets(line, 510, in)) {
    sscanf(line, "%511s ", nm);
    if (!isalpha((unsigned char)line[0])) continue;
    double dist = 0, mass = 0, accret = 0;
    const char *s1 = line + strlen(nm);
    t1[0] = t2[0] = 0;
    sscanf(s1 + strspn(s1, "\t :=>_"), "%lf %lf %lf %511s %511s", &dist, &mass, &accret, t1, t2);
    if (d.CICPlanet) {
      int j = 0;
      while (s->Rmed[j] < dist) j++;
      dist = s->Radii[j + 1];
    }
    p.mass[i] = mass;
    p.feeldisk[i] = (tolower((unsigned char)t1[0]) == 'n') ? 0 : 1;
    p.feelothers[i] = (tolower((unsigned char)t2[0]) == 'n') ? 0 : 1;
    p.x[i] = dist * (1.0 + d.ECCENTRICITY);
    p.y[i] = 0.0;
    p.vy[i] = sqrt(G * (1.0 + mass) / dist) * sqrt((1.0 - d.ECCENTRICITY) / (1.0 + d.ECCENTRICITY));
    p.vx[i] = -0.0000000001 * p.vy[i];
    p.acc[i] = accret;
    i++;
  }
  fclose(in);
  if (p.nb > FG_MAX_PLANETS) {
    host_error("%d planets: at most %d are supported", p.nb, FG_MAX_PLANETS);
    return FG_E_UNSUPPORTED;
  }
  if (p.nb > 0) s->HillRadius = p.x[0] * pow(p.mass[0] / 3., 1. / 3.);
  return FG_OK;
}

// GetPsysInfo, Psys.cu:174-244 (action: 0 GET, 1 MARK, 2 FREQUENCY)
double psys_info(fg_sim *s, int action) {
  Planets &p = s->ps;
  double x, y, vx, vy, xc, yc;
  xc = x = p.x[0];
  yc = y = p.y[0];
  vx = p.vx[0];
  vy = p.vy[0];
  double m = p.mass[0] + 1.;
  double h = x * vy - y * vx;
  double d = sqrt(x * x + y * y);
  double Ax = x * vy * vy - y * vx * vy - G * m * x / d;
  double Ay = y * vx * vx - x * vx * vy - G * m * y / d;
  double e = sqrt(Ax * Ax + Ay * Ay) / m;
  double a = h * h / G / m / (1. - e * e);
  double arg = (e == 0.0) ? 1.0 : (1.0 - d / a) / e;
  double E;
  if (fabs(arg) >= 1.0)
    E = PI * (1. - arg / fabs(arg)) / 2.;
  else
    E = acos((1.0 - d / a) / e);
  if ((x * y * (vy * vy - vx * vx) + vx * vy * (x * x - y * y)) < 0) E = -E;
  double M = E - e * sin(E);
  double PerihelionPA = atan2(Ay, Ax);
  double omega = sqrt(m / a / a / a);
  if (s->d.GuidingCenter) {
    xc = a * cos(M + PerihelionPA);
    yc = a * sin(M + PerihelionPA);
  }
  if (e < 1e-08) {
    xc = x;
    yc = y;
  }
  if (action == 1) {
    p.Xplanet = xc;
    p.Yplanet = yc;
    return 0.;
  }
  if (action == 0) {
    double d2 = sqrt(xc * xc + yc * yc);
    double d1 = sqrt(p.Xplanet * p.Xplanet + p.Yplanet * p.Yplanet);
    double cross = p.Xplanet * yc - xc * p.Yplanet;
    p.Xplanet = xc;
    p.Yplanet = yc;
    return asin(cross / (d1 * d2));
  }
  if (action == 2) return omega;
  return 0.0;
}

// RotatePsys, Psys.cu:248-265
void rotate_psys(fg_sim *s, double angle) {
  Planets &p = s->ps;
  double sint = sin(angle), cost = cos(angle);
  for (int i = 0; i < p.nb; i++) {
    double xt = p.x[i], yt = p.y[i];
    p.x[i] = xt * cost + yt * sint;
    p.y[i] = -xt * sint + yt * cost;
    xt = p.vx[i];
    yt = p.vy[i];
    p.vx[i] = xt * cost + yt * sint;
    p.vy[i] = -xt * sint + yt * cost;
  }
}

// DerivMotionRK5, RungeKutta.cu:30-67 (the "feel others" term uses the
// stellar distance of planet j, as written there)
void deriv_rk5(fg_sim *s, const double *q, double *deriv, int nb, double dt) {
  Planets &p = s->ps;
  const double *x = q, *y = q + nb, *vx = q + 2 * nb, *vy = q + 3 * nb;
  double *dx = deriv, *dy = deriv + nb, *dvx = deriv + 2 * nb, *dvy = deriv + 3 * nb;
  for (int i = 0; i < nb; i++) p.Dist[i] = sqrt(x[i] * x[i] + y[i] * y[i]);
  for (int i = 0; i < nb; i++) {
    dx[i] = vx[i];
    dy[i] = vy[i];
    dvx[i] = -G * 1.0 / p.Dist[i] / p.Dist[i] / p.Dist[i] * x[i];
    dvy[i] = -G * 1.0 / p.Dist[i] / p.Dist[i] / p.Dist[i] * y[i];
    for (int j = 0; j < nb; j++) {
      if (s->d.Indirect_Term) {
        dvx[i] -= G * p.masses[j] / p.Dist[j] / p.Dist[j] / p.Dist[j] * x[j];
        dvy[i] -= G * p.masses[j] / p.Dist[j] / p.Dist[j] / p.Dist[j] * y[j];
      }
      if ((j != i) && p.feelothers[i]) {
        dvx[i] += G * p.masses[j] / p.Dist[j] / p.Dist[j] / p.Dist[j] * (x[j] - x[i]);
        dvy[i] += G * p.masses[j] / p.Dist[j] / p.Dist[j] / p.Dist[j] * (y[j] - y[i]);
      }
    }
  }
  for (int i = 0; i < 4 * nb; i++) deriv[i] *= dt;
}

void translate_rk5(fg_sim *s, const double *qold, double c1, double c2, double c3, double c4, double c5,
                   double *qnew, int nb) {
  Planets &p = s->ps;
  for (int i = 0; i < 4 * nb; i++)
    qnew[i] = qold[i] + c1 * p.k1[i] + c2 * p.k2[i] + c3 * p.k3[i] + c4 * p.k4[i] + c5 * p.k5[i];
}

// RungeKutta, RungeKutta.cu:8-28 (Cash-Karp tableau)
void runge_kutta(fg_sim *s, double dt) {
  Planets &p = s->ps;
  const int nb = p.nb;
  for (int i = 0; i < 4 * nb; i++) p.k1[i] = p.k2[i] = p.k3[i] = p.k4[i] = p.k5[i] = p.k6[i] = 0;
  deriv_rk5(s, p.q0, p.k1, nb, dt);
  translate_rk5(s, p.q0, 0.2, 0.0, 0.0, 0.0, 0.0, p.q1, nb);
  deriv_rk5(s, p.q1, p.k2, nb, dt);
  translate_rk5(s, p.q0, 0.075, 0.225, 0.0, 0.0, 0.0, p.q1, nb);
  deriv_rk5(s, p.q1, p.k3, nb, dt);
  translate_rk5(s, p.q0, 0.3, -0.9, 1.2, 0.0, 0.0, p.q1, nb);
  deriv_rk5(s, p.q1, p.k4, nb, dt);
  translate_rk5(s, p.q0, -11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0.0, p.q1, nb);
  deriv_rk5(s, p.q1, p.k5, nb, dt);
  translate_rk5(s, p.q0, 1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0,
                p.q1, nb);
  deriv_rk5(s, p.q1, p.k6, nb, dt);
  for (int i = 0; i < 4 * nb; i++)
    p.q1[i] = p.q0[i] + 37.0 / 378.0 * p.k1[i] + 250.0 / 621.0 * p.k3[i] + 125.0 / 594.0 * p.k4[i] +
              512.0 / 1771.0 * p.k6[i];
}

// AdvanceSystemRK5, Pframeforce.cu:124-187
void advance_rk5(fg_sim *s, double dt) {
  Planets &p = s->ps;
  const Deck &d = s->d;
  const int nb = p.nb;
  if (!d.ForcedCircular) {
    for (int k = 0; k < nb; k++) {
      p.q0[k] = p.x[k];
      p.q0[k + nb] = p.y[k];
      p.q0[k + 2 * nb] = p.vx[k];
      p.q0[k + 3 * nb] = p.vy[k];
      p.masses[k] = p.mass[k];
    }
    runge_kutta(s, dt);
  }
  for (int i = 1 - (s->PhysicalTime >= d.RELEASEDATE); i < nb; i++) {
    if (!d.ForcedCircular) {
      p.x[i] = p.q1[i];
      p.y[i] = p.q1[i + nb];
      p.vx[i] = p.q1[i + 2 * nb];
      p.vy[i] = p.q1[i + 3 * nb];
    } else {
      double x = p.x[i], y = p.y[i], theta = atan2(y, x), vx = p.vx[i], vy = p.vy[i];
      double r = sqrt(x * x + y * y);
      double omega = (-y * vx + x * vy) / r / r;
      double dtheta = omega * dt;
      p.x[i] = r * cos(theta + dtheta);
      p.y[i] = r * sin(theta + dtheta);
      p.vx[i] = vx * cos(dtheta + theta) - vy * sin(dtheta + theta);
      p.vy[i] = vx * sin(dtheta + theta) + vy * cos(dtheta + theta);
    }
  }
  if (s->PhysicalTime < d.RELEASEDATE) {
    double x = p.x[0], y = p.y[0], r = sqrt(x * x + y * y), theta = atan2(y, x);
    double rdot = (d.RELEASERADIUS - r) / (d.RELEASEDATE - s->PhysicalTime);
    double omega = sqrt((1. + p.mass[0]) / r / r / r);
    double new_r = r + rdot * dt, denom = r - new_r, dtheta;
    if (denom != 0.0)
      dtheta = 2. * dt * r * omega / denom * (sqrt(r / new_r) - 1.);
    else
      dtheta = omega * dt;
    double vx = rdot, vy = new_r * sqrt((1. + p.mass[0]) / new_r / new_r / new_r);
    p.x[0] = new_r * cos(dtheta + theta);
    p.y[0] = new_r * sin(dtheta + theta);
    p.vx[0] = vx * cos(dtheta + theta) - vy * sin(dtheta + theta);
    p.vy[0] = vx * sin(dtheta + theta) + vx * cos(dtheta + theta);  // as Pframeforce.cu:184
  }
}

// CircumPlanetaryMass, SourceEuler.cu:549-563: around planet 0, with the
// Hill radius of the initial planet system (Psys.cu:140)
static fg_status circumplanetary_mass(fg_sim *s, double *m) {
  *m = 0.0;
  if (s->ps.nb < 1) return FG_OK;
  return fg_circumplanetary_mass(s->ctx, s->ps.x[0], s->ps.y[0], s->HillRadius, m);
}

// end of a substep with diagnostics on, SourceEuler.cu:297-298
static fg_status substep_diagnostics(fg_sim *s) {
  if (!s->opt.diagnostics || !s->d.IsDisk) return FG_OK;
  fg_status st = circumplanetary_mass(s, &s->mdcp1);
  s->exces_mdcp = s->mdcp1 - s->mdcp;
  return st;
}

// smoothing length of a planet's potential / force, Pframeforce.cu:74-76
double planet_smoothing(fg_sim *s, double r, double mass, bool from_disk) {
  const Deck &d = s->d;
  if (d.RocheSmoothing) {
    if (from_disk) return r * pow(mass / 3., 1. / 3.) * d.ROCHESMOOTHING;  // Pframeforce.cu:110
    return r * pow((1.0 / 3.0 * mass), 1.0 / 3.0) * d.ROCHESMOOTHING;   // Pframeforce.cu:74-75
  }
  return s->prof.smoothing(r);
}

// FillForcesArrays's host half, Pframeforce.cu:56-83
int potential_planets(fg_sim *s, fg_planet_pot *out) {
  Planets &p = s->ps;
  for (int k = 0; k < p.nb; k++) {
    double xp = p.x[k], yp = p.y[k], mp = p.mass[k] * s->MassTaper;
    double PlanetDistance = sqrt(xp * xp + yp * yp);
    double InvPlanetDistance3 = 1.0 / PlanetDistance / PlanetDistance / PlanetDistance;
    double RRoche = PlanetDistance * pow((1.0 / 3.0 * mp), 1.0 / 3.0);
    double smoothing = s->d.RocheSmoothing ? RRoche * s->d.ROCHESMOOTHING : s->prof.smoothing(PlanetDistance);
    out[k].x = xp;
    out[k].y = yp;
    out[k].mass = mp;
    out[k].smooth = smoothing * smoothing;
    out[k].inv_d3 = InvPlanetDistance3;
  }
  return p.nb;
}

uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// host tables of fargo_hip.h, evaluated with the kernels' expressions
struct TableStore {
  std::vector<double> cs_iso, supp_torque, dxtheta, invdxtheta, dxtheta_cfl, ev_coef, ev_period, ev_vtheta_arg,
      visc_const, visc_pow15, visc_cav, cos_cell, sin_cell, cos_pot, sin_pot, sg_Kr, sg_Kt;
  std::vector<int> ev_zone, visc_cav_on;
};

void build_tables(fg_sim *s, TableStore &T, fg_tables &t) {
  const Deck &d = s->d;
  const int nr = s->nr, ns = s->ns;
  const size_t R = (size_t)nr + 1;
  for (auto *v : {&T.cs_iso, &T.supp_torque, &T.dxtheta, &T.invdxtheta, &T.dxtheta_cfl, &T.ev_coef, &T.ev_period,
                  &T.ev_vtheta_arg, &T.visc_const, &T.visc_pow15, &T.visc_cav})
    v->assign(R, 0.0);
  T.ev_zone.assign(R, 0);
  T.visc_cav_on.assign(R, 0);
  const double Tin = 2.0 * PI * pow(s->Rmed[0], 3. / 2);
  const double Tout = 2.0 * PI * pow(s->Rmed[nr - 1], 3. / 2);
  const double DRMIN = s->Rmed[0] * 1.25, DRMAX = s->Rmed[nr - 1] * 0.84;
  for (int i = 0; i < nr; i++) {
    const double Rm = s->Rmed[i];
    // ComputeSoundSpeedKernel isothermal branch, Kernels.cu:177-179
    T.cs_iso[i] = s->prof.aspect_ratio(Rm) * sqrt(G * 1.0 / Rm) * pow(Rm, d.FLARINGINDEX);
    T.supp_torque[i] = d.IMPOSEDDISKDRIFT * 0.5 * pow(Rm, -2.5 + d.SIGMASLOPE);  // Kernels.cu:40
    T.dxtheta[i] = 2.0 * PI / (double)ns * Rm;                                     // Kernels.cu:41
    T.invdxtheta[i] = 1.0 / T.dxtheta[i];
    T.dxtheta_cfl[i] = Rm * 2.0 * PI / (double)ns;  // Kernels.cu:1295
    // EvanescentBoundaryKernel constants, Kernels.cu:978-992
    if ((Rm < DRMIN) || (Rm > DRMAX)) {
      double damping;
      T.ev_zone[i] = 1;
      if (Rm < DRMIN) {
        damping = (Rm - DRMIN) / (s->Rmed[0] - DRMIN);
        T.ev_coef[i] = damping * damping * 10.0;
        T.ev_period[i] = Tin;
      }
      if (Rm > DRMAX) {
        damping = (Rm - DRMAX) / (s->Rmed[nr - 1] - DRMAX);
        T.ev_coef[i] = damping * damping * 10.0;
        T.ev_period[i] = Tout;
      }
    }
    double ar = s->prof.aspect_ratio(Rm);
    T.ev_vtheta_arg[i] =
        G * 1.0 / Rm * (1.0 - (1.0 + d.SIGMASLOPE - 2.0 * d.FLARINGINDEX) * pow(ar, 2.0) * pow(Rm, 2.0 * d.FLARINGINDEX));
    T.visc_const[i] = d.VISCOSITY;
    T.visc_pow15[i] = pow(Rm, 1.5);
    double f;
    T.visc_cav_on[i] = s->prof.cavity(Rm, &f);
    T.visc_cav[i] = f;
  }
  for (auto *v : {&T.cos_cell, &T.sin_cell, &T.cos_pot, &T.sin_pot}) v->assign(ns, 0.0);
  for (int j = 0; j < ns; j++) {
    T.cos_cell[j] = cos((2.0 * PI * (double)j) / (double)ns);  // Kernels.cu:161-162
    T.sin_cell[j] = sin((2.0 * PI * (double)j) / (double)ns);
    double angle = (double)j / (double)ns * 2.0 * PI;  // Kernels.cu:1236
    T.cos_pot[j] = cos(angle);
    T.sin_pot[j] = sin(angle);
  }
  if (d.SelfGravity) {
    // compute_kernel, SgMain.cu:49-81 (coshf in FP32 as written there)
    const double eps = d.THICKNESSSMOOTHING * d.ASPECTRATIO;
    T.sg_Kr.assign(2 * (size_t)nr * ns, 0.0);
    T.sg_Kt.assign(2 * (size_t)nr * ns, 0.0);
    for (int i = 0; i < 2 * nr; i++) {
      double u = (i < nr) ? log(s->Radii[i] / s->Radii[0]) : -log(s->Radii[2 * nr - i] / s->Radii[0]);
      for (int j = 0; j < ns; j++) {
        double theta = 2.0 * PI * (double)j / (double)ns;
        double Kr = 1.0 + eps * eps - cos(theta) * exp(-u);
        double base = eps * eps * exp(u) + 2.0 * (coshf((float)u) - cos(theta));
        Kr *= pow(base, -1.5);
        double Kt = sin(theta);
        Kt *= pow(base, -1.5);
        T.sg_Kr[(size_t)i * ns + j] = Kr;
        T.sg_Kt[(size_t)i * ns + j] = Kt;
      }
    }
  }
  t = fg_tables{};
  t.Radii = s->Radii.data();
  t.Rinf = s->Rinf.data();
  t.Rsup = s->Rsup.data();
  t.Rmed = s->Rmed.data();
  t.Surf = s->Surf.data();
  t.invRinf = s->invRinf.data();
  t.invSurf = s->invSurf.data();
  t.invRmed = s->invRmed.data();
  t.invdiffRsup = s->invdiffRsup.data();
  t.invdiffRmed = s->invdiffRmed.data();
  t.SigmaMed = s->SigmaMed.data();
  t.SigmaInf = s->SigmaInf.data();
  t.EnergyMed = s->EnergyMed.data();
  t.CoolingTimeMed = s->CoolingTimeMed.data();
  t.QplusMed = s->QplusMed.data();
  t.cs_iso = T.cs_iso.data();
  t.supp_torque = T.supp_torque.data();
  t.dxtheta = T.dxtheta.data();
  t.invdxtheta = T.invdxtheta.data();
  t.dxtheta_cfl = T.dxtheta_cfl.data();
  t.ev_zone = T.ev_zone.data();
  t.ev_coef = T.ev_coef.data();
  t.ev_period = T.ev_period.data();
  t.ev_vtheta_arg = T.ev_vtheta_arg.data();
  t.visc_const = T.visc_const.data();
  t.visc_pow15 = T.visc_pow15.data();
  t.visc_cav_on = T.visc_cav_on.data();
  t.visc_cav = T.visc_cav.data();
  t.cos_cell = T.cos_cell.data();
  t.sin_cell = T.sin_cell.data();
  t.cos_pot = T.cos_pot.data();
  t.sin_pot = T.sin_pot.data();
  t.sg_Kr = T.sg_Kr.empty() ? nullptr : T.sg_Kr.data();
  t.sg_Kt = T.sg_Kt.empty() ? nullptr : T.sg_Kt.data();
  t.dtheta = 2.0 * PI / (double)ns;                   // Kernels.cu:1037
  t.invdphi = 1.0 / (2.0 * M_PI / (double)ns);        // Kernels.cu:742-743
  t.dt1d_coef = 2.0 * PI * CFLSECURITY / (double)ns;  // Kernels.cu:1332
  t.vrad0_slope = -d.SIGMASLOPE + .5;                 // Kernels.cu:1002
  t.visc_value = d.VISCOSITY;
}

#define CHECK(x)                 \
  do {                           \
    fg_status st_ = (x);         \
    if (st_) return st_;         \
  } while (0)

// the DT-opening part of AlgoGas, SourceEuler.cu:202-221
fg_status begin_dt(fg_sim *s) {
  s->dtemp = 0.0;
  int cfl = 1;
  if (s->d.Adiabatic) CHECK(fg_sound_speed(s->ctx));
  if (s->d.IsDisk && s->d.SloppyCFL) CHECK(fg_cfl(s->ctx, s->d.DT - s->dtemp, &cfl));
  s->dt = s->d.DT / cfl;
  s->last_cfl = cfl;
  s->in_dt = 1;
  return FG_OK;
}

}  // namespace

// Self-gravity of the initial disk (Main.cu:268-290): the field of the
// uploaded density, the planets' orbital speed corrected by the ring mean
// of the radial SG acceleration (Init_planetarysys_withSG,
// SgMain.cu:166-194), the frame rate recomputed from them, and the gas
// rotation in equilibrium with it (Init_azimutalvelocity_withSG,
// SourceEuler.cu:441-450, Kernels.cu:935-950) minus the frame rotation
// (Kernels.cu:648-689).
static fg_status sg_initial_state(fg_sim *s, std::vector<double> &vtheta) {
  const Deck &d = s->d;
  const int nr = s->nr, ns = s->ns;
  CHECK(fg_selfgravity(s->ctx, 0.0, 0));
  std::vector<double> axi(nr + 1, 0.0);
  CHECK(fg_get_rings(s->ctx, nullptr, nullptr, nullptr, nullptr, axi.data()));
  Planets &p = s->ps;
  for (int k = 0; k < p.nb; k++) {
    double r = p.x[k];
    double dist = (double)(r / (1. + d.ECCENTRICITY));
    int ipl = 0;
    while (ipl < nr - 2 && s->Rmed[ipl] <= dist) ipl++;
    double ri = s->Rmed[ipl], rip1 = s->Rmed[ipl + 1], dr = rip1 - ri;
    double sgacc = (dist - ri) * axi[ipl + 1] + (rip1 - dist) * axi[ipl];
    sgacc /= dr;
    p.vy[k] *= (double)sqrt(1. - dist * dist * sgacc);
  }
  s->OmegaFrame = d.OMEGAFRAME;
  if (d.Corotating) s->OmegaFrame = psys_info(s, 2);
  for (int i = 0; i <= nr; i++) {
    double v = 0.0, r = (i == nr) ? s->Rmed[nr - 1] : s->Rmed[i];
    if (i < nr) {
      double invr = 1. / s->Rmed[i];
      double omegakep = sqrt(G * 1.0 * invr * invr * invr);
      double omega = sqrt(omegakep * omegakep *
                              (1.0 - (1. + d.SIGMASLOPE - 2.0 * d.FLARINGINDEX) * pow(d.ASPECTRATIO, 2.0) *
                                         pow(s->Rmed[i], 2.0 * d.FLARINGINDEX)) -
                          invr * axi[i]);
      v = s->Rmed[i] * omega;
    }
    v -= s->OmegaFrame * r;
    for (int j = 0; j < ns; j++) vtheta[(size_t)i * ns + j] = v;
  }
  return fg_set_field(s->ctx, FG_VTHETA, vtheta.data());
}

// -------------------------------------------------------------- C-ABI --
extern "C" fg_status fg_sim_create(const char *parfile, const fg_sim_options *opt, fg_sim **out) {
  if (!parfile || !out) {
    host_error("fg_sim_create: NULL argument");
    return FG_E_INVALID;
  }
  *out = nullptr;
  fg_sim *s = new fg_sim();
  if (opt)
    s->opt = *opt;
  else
    s->opt = fg_sim_options{};
  Deck &d = s->d;
  const char *flags = s->opt.cli_flags ? s->opt.cli_flags : "";
  if (strchr(flags, 'c')) d.SloppyCFL = 1;
  fg_status st = d.read(parfile, s->opt.outdir);
  if (st) {
    delete s;
    return st;
  }
  if (d.LAMBDADOUBLING != 0.0) {
    host_error("LambdaDoubling != 0 (time-dependent aspect ratio) is not supported");
    delete s;
    return FG_E_UNSUPPORTED;
  }
  if (d.AdvecteLabel) {
    host_error("LabelAdvection is disabled in the reference (TransportEuler.cu:34) and not supported");
    delete s;
    return FG_E_UNSUPPORTED;
  }
  if (d.NRAD < 3 || d.NSEC < 2) {
    host_error("grid %d x %d too small", d.NRAD, d.NSEC);
    delete s;
    return FG_E_PARAM;
  }
  s->nr = d.NRAD;
  s->ns = d.NSEC;
  s->S = (long)(s->nr + 1) * s->ns;
  s->prof.d = &s->d;
  const int nr = s->nr, ns = s->ns;
  if ((st = fill_polar(s)) || (st = init_planets(s, parfile))) {
    delete s;
    return st;
  }
  // InitGasDensity / InitGasEnergy (Pframeforce.cu:31-52, Theo.cu)
  std::vector<double> dens(s->S, 0.0), vrad(s->S, 0.0), vtheta(s->S, 0.0), energy(s->S, 0.0);
  for (int i = 0; i < nr; i++) {
    s->SigmaMed[i] = s->prof.sigma(s->Rmed[i]);
    s->SigmaInf[i] = s->prof.sigma(s->Rinf[i]);
  }
  for (int i = 0; i < nr; i++)
    for (int j = 0; j < ns; j++) dens[(size_t)i * ns + j] = s->SigmaMed[i];
  if (s->opt.noise_seed) {
    // seeded white noise on the initial surface density: xi uniform in
    // [-1, 1) from splitmix64(seed ^ (i*NSEC + j))
    for (int i = 0; i < nr; i++)
      for (int j = 0; j < ns; j++) {
        uint64_t h = splitmix64(s->opt.noise_seed ^ (uint64_t)((uint64_t)i * ns + j));
        double xi = (double)(h >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
        dens[(size_t)i * ns + j] = s->SigmaMed[i] * (1.0 + s->opt.noise_amp * xi);
      }
  }
  if (d.Adiabatic) {
    for (int i = 0; i < nr; i++) s->EnergyMed[i] = s->prof.energy(s->Rmed[i]);
    for (int i = 0; i < nr; i++)
      for (int j = 0; j < ns; j++) energy[(size_t)i * ns + j] = s->EnergyMed[i];
  }
  s->OmegaFrame = d.OMEGAFRAME;
  if (d.Corotating) s->OmegaFrame = psys_info(s, 2);
  // ring mean of the initial c_s (Make1Dprofile(1) in InitGasVelocities,
  // Pframeforce.cu:237-240): sequential sum over the sectors
  TableStore T;
  fg_tables tables;
  build_tables(s, T, tables);
  if (d.ViscosityAlpha) {
    for (int i = 0; i < nr; i++) {
      double sum = 0.0;
      for (int j = 0; j < ns; j++) {
        size_t l = (size_t)i * ns + j;
        double cs = d.Adiabatic ? sqrt(d.ADIABATICINDEX * (d.ADIABATICINDEX - 1.0) * energy[l] / dens[l]) : T.cs_iso[i];
        sum += cs;
      }
      s->cs_mean[i] = sum / (double)ns;
    }
  }
  if (d.Cooling) {
    for (int i = 0; i < nr; i++) s->CoolingTimeMed[i] = s->prof.cooling_time(s->Rmed[i]);
    for (int i = 0; i < nr; i++)  // Qplusinit, Theo.cu:79-85
      s->QplusMed[i] = 2.25 * s->prof.viscosity(s->Rmed[i], s->Rmed, s->cs_mean) * d.SIGMA0 *
                       pow(s->Rmed[i], -d.SIGMASLOPE - 3.0);
  }
  // InitVelocities + InitGasVelocitiesKernel (Pframeforce.cu:256-270,
  // Kernels.cu:648-689); the kernel receives ViscosityAlpha as 0, so the
  // initial drift always uses the constant-viscosity formula.
  std::vector<double> visc(nr + 1, 0.0);
  for (int i = 0; i < nr; i++) visc[i] = s->prof.viscosity(s->Rmed[i], s->Rmed, s->cs_mean);
  visc[nr] = s->prof.viscosity(s->Rmed[nr - 1], s->Rmed, s->cs_mean);
  for (int i = 0; i <= nr; i++) {
    double r = (i == nr) ? s->Rmed[nr - 1] : s->Rmed[i];
    double ri = (i == nr) ? s->Rinf[nr - 1] : s->Rinf[i];
    for (int j = 0; j < ns; j++) {
      size_t l = (size_t)i * ns + j;
      double omega = sqrt(G * 1.0 / r / r / r);
      vtheta[l] = omega * r *
                  sqrt(1.0 - pow(d.ASPECTRATIO, 2.0) * pow(r, 2.0 * d.FLARINGINDEX) *
                                 (1. + d.SIGMASLOPE - 2.0 * d.FLARINGINDEX));
      vtheta[l] -= s->OmegaFrame * r;
      if (i == nr)
        vrad[l] = 0.0;
      else {
        vrad[l] = d.IMPOSEDDISKDRIFT * d.SIGMA0 / s->SigmaInf[i] / ri;
        vrad[l] -= 3.0 * visc[i] / r * (-d.SIGMASLOPE + .5);
      }
    }
  }
  for (int j = 0; j < ns; j++) vrad[j] = vrad[(size_t)nr * ns + j] = 0.0;

  fg_params p{};
  p.nrad = nr;
  p.nsec = ns;
  p.precision = s->opt.precision;
  p.adiabatic = d.Adiabatic;
  p.cooling = d.Cooling;
  p.self_gravity = d.SelfGravity;
  p.viscosity_alpha = d.ViscosityAlpha;
  p.fast_transport = d.FastTransport;
  p.open_inner = d.OpenInner;
  p.non_reflecting = d.NonReflecting;
  p.evanescent = d.Evanescent;
  p.indirect_term = d.Indirect_Term;
  p.exclude_hill = d.ExcludeHill;
  p.adiabatic_index = d.ADIABATICINDEX;
  p.alphaviscosity = d.ALPHAVISCOSITY;
  p.sigmaslope = d.SIGMASLOPE;
  p.flaringindex = d.FLARINGINDEX;
  // self-gravity constants, Main.cu:269-271
  p.sg_eps = d.THICKNESSSMOOTHING * d.ASPECTRATIO;
  p.sg_rstep = log(s->Radii[nr] / s->Radii[0]) / (double)nr;
  p.sg_tstep = 2.0 * PI / (double)ns;
  if ((st = fg_create(&p, &tables, s->opt.device, &s->ctx))) {
    delete s;
    return st;
  }
  if ((st = fg_upload_state(s->ctx, dens.data(), vrad.data(), vtheta.data(), d.Adiabatic ? energy.data() : nullptr))) {
    fg_sim_destroy(s);
    return st;
  }
  if (d.SelfGravity && (st = sg_initial_state(s, vtheta))) {
    fg_sim_destroy(s);
    return st;
  }
  // device-side EOS and c_s profile of the initial state (InitEuler,
  // SourceEuler.cu:189-191)
  if ((st = fg_sound_speed(s->ctx)) || (st = fg_pressure(s->ctx)) || (st = fg_temperature(s->ctx)) ||
      (d.ViscosityAlpha && (st = fg_ring_mean_cs(s->ctx))) || (st = fg_sync(s->ctx))) {
    fg_sim_destroy(s);
    return st;
  }
  if (s->opt.write_files) {
    // WriteDim (Output.cu:19-35); the planet files are emptied by the first
    // output (EmptyPlanetSystemFile, Output.cu:37), not after a restart
    std::string path = d.OUTPUTDIR + "dims.dat";
    FILE *f = fopen(path.c_str(), "w");
    if (f) {
      fprintf(f, "%d\t%d\t\t%d\t%d\t%f\t%d\t%d\t%d\n", 0, 0, 0, 0, d.RMAX, d.NTOT / d.NINTERM, nr, ns);
      fclose(f);
    }
  }
  s->PhysicalTimeInitial = s->PhysicalTime;
  s->prof.time = s->PhysicalTime;
  s->prof.time0 = s->PhysicalTimeInitial;
  // the fused path covers decks with evanescent edges and a per-substep
  // CFL, isothermal or adiabatic, with or without self-gravity; anything
  // else stays on the per-routine path
  const bool fusable = d.Evanescent && !d.OpenInner && !d.NonReflecting && !d.SloppyCFL && d.IsDisk && ns <= 4096 &&
                       (!d.Adiabatic || nr >= 3);
  // the FP32 bulk mode exists only as fused kernels, isothermal without
  // self-gravity
  if (s->opt.precision == FG_PREC_FP32 && (!fusable || d.Adiabatic || d.SelfGravity)) {
    host_error("FP32 bulk mode needs a locally isothermal deck with evanescent edges, per-substep CFL and "
               "NSEC <= 4096");
    fg_sim_destroy(s);
    return FG_E_UNSUPPORTED;
  }
  s->fused_active = (s->opt.fused || s->opt.precision == FG_PREC_FP32) && fusable;
  // Main.cu:302: circumplanetary mass of the initial density
  if (s->opt.diagnostics && (st = circumplanetary_mass(s, &s->mdcp))) {
    fg_sim_destroy(s);
    return st;
  }
  *out = s;
  return FG_OK;
}

extern "C" fg_status fg_sim_destroy(fg_sim *s) {
  if (!s) return FG_OK;
  if (s->ctx) fg_destroy(s->ctx);
  delete s;
  return FG_OK;
}

// AlgoGas's loop body on the fused path: the same host-side sequence, the
// GPU part in one fg_step_fused call whose reductions already hold the
// next substep's CFL count and star force.  Split in three so that a
// batch of disks can run the GPU part of all of them in one launch:
//   fused_open   start of a DT sub-cycle, mass taper
//   fused_plan   dt from the CFL count, indirect term from the star force,
//                planets (disk kick, RK5, frame) and their potential terms
//   fused_close  time bookkeeping after the GPU part
struct fused_launch {
  double dt, domega, DeltaT_next;
  fg_planet_pot pl[FG_MAX_PLANETS];
  int nb, done;
};

static void fused_open(fg_sim *s) {
  const Deck &d = s->d;
  if (!s->in_dt) {
    s->dtemp = 0.0;
    s->in_dt = 1;
  }
  s->MassTaper = s->PhysicalTime / (d.MASSTAPER * 2.0 * M_PI);
  s->MassTaper = (s->MassTaper > 1.0 ? 1.0 : pow(sin(s->MassTaper * M_PI / 2.0), 2.0));
}

static fg_status fused_plan(fg_sim *s, int cfl, const double force[2], fused_launch &L) {
  const Deck &d = s->d;
  fg_ctx *c = s->ctx;
  s->last_cfl = cfl;
  const double dt = (d.DT - s->dtemp) / (double)cfl;
  s->dt = dt;
  s->dtemp += dt;
  if (d.Corotating) psys_info(s, 1);
  s->IndirectX = d.Indirect_Term ? -force[0] : 0.0;
  s->IndirectY = d.Indirect_Term ? -force[1] : 0.0;
  L.nb = potential_planets(s, L.pl);
  Planets &p = s->ps;
  const bool pf = s->pf_ready;
  s->pf_ready = 0;
  for (int k = 0; k < p.nb; k++) {  // AdvanceSystemFromDisk, Pframeforce.cu:97-121
    if (!p.feeldisk[k]) continue;
    double m = p.mass[k], x = p.x[k], y = p.y[k], r = sqrt(x * x + y * y);
    double gamma[2];
    if (pf)  // launched behind the last substep, for these positions and this density
      fg_fused_planet_force(c, k, gamma);
    else
      CHECK(fg_disk_force(c, x, y, planet_smoothing(s, r, m, true), gamma));
    p.vx[k] += dt * gamma[0];
    p.vy[k] += dt * gamma[1];
    p.vx[k] += dt * s->IndirectX;
    p.vy[k] += dt * s->IndirectY;
  }
  advance_rk5(s, dt);
  double domega = 0.0;
  if (d.Corotating) {
    double OmegaNew = psys_info(s, 0) / dt;
    domega = OmegaNew - s->OmegaFrame;
    s->OmegaFrame = OmegaNew;
  }
  rotate_psys(s, s->OmegaFrame * dt);
  L.done = !(s->dtemp < 0.999999999 * d.DT);
  L.DeltaT_next = L.done ? d.DT : d.DT - s->dtemp;
  L.dt = dt;
  L.domega = domega;
  return FG_OK;
}

static void fused_close(fg_sim *s, const fused_launch &L, int *closed) {
  s->nsub++;
  s->PhysicalTime += L.dt;
  s->prof.time = s->PhysicalTime;
  if (L.done) s->in_dt = 0;
  if (closed) *closed = L.done;
}

static fg_status fused_substep(fg_sim *s, int *closed) {
  const Deck &d = s->d;
  fg_ctx *c = s->ctx;
  s->state_gen++;
  fused_open(s);
  if (!s->results_ready) {
    CHECK(fg_fused_prepare(c, d.DT - s->dtemp));
    s->results_ready = 1;
    s->pf_ready = 0;
  }
  int cfl = 1;
  double force[2];
  CHECK(fg_fused_results(c, &cfl, force, nullptr));
  fused_launch L;
  CHECK(fused_plan(s, cfl, force, L));
  CHECK(fg_step_fused(c, L.dt, s->OmegaFrame, L.domega, L.pl, L.nb, s->IndirectX, s->IndirectY, L.DeltaT_next,
                      (d.Write_Temperature ? FG_STEP_TEMPERATURE : 0) | (L.done ? FG_STEP_CLOSES_DT : 0)));
  // the next substep's AdvanceSystemFromDisk reads the density this one
  // leaves at the planets' positions now (fused_plan moved them): its
  // reductions queue behind this substep and come back with its results,
  // so the next substep waits for the GPU once
  {
    const Planets &p = s->ps;
    int any = 0;
    for (int k = 0; k < p.nb; k++) {
      if (!p.feeldisk[k]) continue;
      const double m = p.mass[k], x = p.x[k], y = p.y[k], r = sqrt(x * x + y * y);
      CHECK(fg_fused_planet_force_async(c, k, x, y, planet_smoothing(s, r, m, true)));
      any = 1;
    }
    s->pf_ready = any;
  }
  CHECK(substep_diagnostics(s));
  fused_close(s, L, closed);
  return FG_OK;
}

// one iteration of the while loop of AlgoGas, SourceEuler.cu:222-305
extern "C" fg_status fg_sim_substep(fg_sim *s, int *closed) {
  if (!s) return FG_E_INVALID;
  if (s->fused_active) return fused_substep(s, closed);
  s->state_gen++;
  const Deck &d = s->d;
  fg_ctx *c = s->ctx;
  if (!s->in_dt) CHECK(begin_dt(s));
  double dt = s->dt;
  s->MassTaper = s->PhysicalTime / (d.MASSTAPER * 2.0 * M_PI);
  s->MassTaper = (s->MassTaper > 1.0 ? 1.0 : pow(sin(s->MassTaper * M_PI / 2.0), 2.0));
  if (d.IsDisk && !d.SloppyCFL) {
    int cfl = 1;
    CHECK(fg_cfl(c, d.DT - s->dtemp, &cfl));
    s->last_cfl = cfl;
    dt = (d.DT - s->dtemp) / (double)cfl;
    s->dt = dt;
  }
  s->dtemp += dt;
  double dopa[2] = {0.0, 0.0};
  if (d.Corotating) psys_info(s, 1);
  if (d.IsDisk) {
    CHECK(fg_disk_force(c, 0.0, 0.0, 0.0, dopa));
    s->IndirectX = d.Indirect_Term ? -dopa[0] : 0.0;
    s->IndirectY = d.Indirect_Term ? -dopa[1] : 0.0;
    fg_planet_pot pl[FG_MAX_PLANETS];
    int nb = potential_planets(s, pl);
    CHECK(fg_potential(c, pl, nb, s->IndirectX, s->IndirectY));
    // AdvanceSystemFromDisk, Pframeforce.cu:97-121
    Planets &p = s->ps;
    for (int k = 0; k < p.nb; k++) {
      if (!p.feeldisk[k]) continue;
      double m = p.mass[k], x = p.x[k], y = p.y[k], r = sqrt(x * x + y * y);
      double smoothing = planet_smoothing(s, r, m, true);
      double gamma[2];
      CHECK(fg_disk_force(c, x, y, smoothing, gamma));
      p.vx[k] += dt * gamma[0];
      p.vy[k] += dt * gamma[1];
      p.vx[k] += dt * s->IndirectX;
      p.vy[k] += dt * s->IndirectY;
    }
  }
  advance_rk5(s, dt);
  if (d.Corotating) {
    double OmegaNew = psys_info(s, 0) / dt;
    double domega = OmegaNew - s->OmegaFrame;
    if (d.IsDisk) CHECK(fg_correct_vtheta(c, domega));
    s->OmegaFrame = OmegaNew;
  }
  rotate_psys(s, s->OmegaFrame * dt);
  if (d.IsDisk) {
    double lost = 0.0;
    CHECK(fg_boundary(c, dt, s->OmegaFrame));
    CHECK(fg_pressure(c));
    CHECK(fg_substep1(c, dt, s->OmegaFrame));
    CHECK(fg_substep2(c, dt));
    CHECK(fg_boundary(c, dt, s->OmegaFrame));
    if (d.Adiabatic) CHECK(fg_substep3(c, dt));
    CHECK(fg_transport(c, dt, s->OmegaFrame, &lost));
    s->LostMass += lost;
    CHECK(fg_boundary(c, dt, s->OmegaFrame));
    CHECK(fg_temperature(c));
    CHECK(substep_diagnostics(s));
  }
  s->nsub++;
  s->PhysicalTime += dt;
  s->prof.time = s->PhysicalTime;
  int done = !(s->dtemp < 0.999999999 * d.DT);
  if (done) s->in_dt = 0;
  if (closed) *closed = done;
  return FG_OK;
}

extern "C" fg_status fg_sim_algogas(fg_sim *s) {
  if (!s) return FG_E_INVALID;
  if (s->fused_active) {
    int closed = 0;
    do CHECK(fused_substep(s, &closed));
    while (!closed);
    return FG_OK;
  }
  CHECK(begin_dt(s));
  while (s->dtemp < 0.999999999 * s->d.DT) CHECK(fg_sim_substep(s, nullptr));
  s->in_dt = 0;
  return FG_OK;
}

static fg_status write_polar(fg_sim *s, const std::vector<double> &a, const char *name, int number) {
  std::string dir = s->d.OUTPUTDIR + name;
  struct stat st;
  if (stat(dir.c_str(), &st) == -1) mkdir(dir.c_str(), 0700);
  std::string path = dir + "/" + name + std::to_string(number) + ".dat";
  FILE *f = fopen(path.c_str(), "w");
  if (!f) {
    host_error("Unable to open '%s'", path.c_str());
    return FG_E_IO;
  }
  fwrite(a.data(), sizeof(double), (size_t)s->nr * s->ns, f);
  fclose(f);
  return FG_OK;
}

// SendOutput + WritePlanetSystemFile, Output.cu:76-181
extern "C" fg_status fg_sim_write_outputs(fg_sim *s, int idx) {
  if (!s) return FG_E_INVALID;
  const Deck &d = s->d;
  if (d.IsDisk) {
    std::vector<double> dens(s->S), vr(s->S), vt(s->S), en(s->S, 0.0), temp(s->S), qp(s->S, 0.0);
    CHECK(fg_download_state(s->ctx, dens.data(), vr.data(), vt.data(), d.Adiabatic ? en.data() : nullptr,
                            temp.data(), d.Cooling ? qp.data() : nullptr));
    if (d.Write_Density) CHECK(write_polar(s, dens, "dens", idx));
    if (d.Write_Velocity) {
      CHECK(write_polar(s, vr, "vrad", idx));
      CHECK(write_polar(s, vt, "vtheta", idx));
    }
    if (d.Write_Energy) CHECK(write_polar(s, en, "energy", idx));
    if (d.Write_Temperature) CHECK(write_polar(s, temp, "temperature", idx));
    if (d.Write_Qplus) CHECK(write_polar(s, qp, "qplus", idx));
  }
  if (!s->planets_fresh) {  // EmptyPlanetSystemFile, Output.cu:37-59
    for (int k = 0; k < s->ps.nb; k++) {
      std::string path = d.OUTPUTDIR + "planet" + std::to_string(k) + ".dat";
      FILE *f = fopen(path.c_str(), "w");
      if (f) fclose(f);
    }
    s->planets_fresh = 1;
  }
  for (int k = 0; k < s->ps.nb; k++) {
    std::string path = d.OUTPUTDIR + "planet" + std::to_string(k) + ".dat";
    FILE *f = fopen(path.c_str(), "a");
    if (!f) {
      host_error("Can't write 'planet%d.dat' file.", k);
      return FG_E_IO;
    }
    const Planets &p = s->ps;
    fprintf(f, "%d\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\n", idx, p.x[k], p.y[k],
            p.vx[k], p.vy[k], p.mass[k], s->LostMass, s->PhysicalTime, s->OmegaFrame, s->mdcp, s->exces_mdcp);
    fclose(f);
  }
  return FG_OK;
}

// WriteBigPlanetSystemFile, Output.cu:62-107: appended every main-loop
// iteration (Main.cu:314), with the latest circumplanetary mass
static fg_status write_bigplanet(fg_sim *s, int TimeStep) {
  const Planets &p = s->ps;
  for (int k = 0; k < p.nb; k++) {
    std::string path = s->d.OUTPUTDIR + "bigplanet" + std::to_string(k) + ".dat";
    FILE *f = fopen(path.c_str(), "a");
    if (!f) {
      host_error("Can't write 'bigplanet.dat' file.");
      return FG_E_IO;
    }
    fprintf(f, "%d\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\n", TimeStep, p.x[k],
            p.y[k], p.vx[k], p.vy[k], p.mass[k], s->LostMass, s->PhysicalTime, s->OmegaFrame, s->mdcp1,
            s->exces_mdcp);
    fclose(f);
  }
  return FG_OK;
}

// UpdateLog, Force.cu:19-84: per planet, the disk torque and power split
// inside / outside the planet's orbit, with (ex_) and without the Hill cut,
// into tqwkK.dat; without self-gravity also each of the dimfxy Hill-cut
// variants into tqwkK_k.dat.  Appended every main-loop iteration
// (Main.cu:315).
static fg_status update_log(fg_sim *s, int TimeStep) {
  const int dimfxy = 11;  // Main.cu:74
  const Planets &p = s->ps;
  double gf[4 * dimfxy];
  for (int i = 0; i < p.nb; i++) {
    double x = p.x[i], y = p.y[i], vx = p.vx[i], vy = p.vy[i], m = p.mass[i];
    double r = sqrt(x * x + y * y);
    double a = sqrt(x * x + y * y);
    double rh = pow(m / 3., 1. / 3.) * a + 1e-15;
    double smoothing = planet_smoothing(s, r, m, true);
    CHECK(fg_disk_force_hill(s->ctx, x, y, smoothing, dimfxy, a, rh, gf));
    const double fx_inner = gf[0], fx_ex_inner = gf[dimfxy - 1], fx_outer = gf[dimfxy],
                 fx_ex_outer = gf[2 * dimfxy - 1], fy_inner = gf[2 * dimfxy], fy_ex_inner = gf[3 * dimfxy - 1],
                 fy_outer = gf[3 * dimfxy], fy_ex_outer = gf[4 * dimfxy - 1];
    std::string base = s->d.OUTPUTDIR + "tqwk" + std::to_string(i);
    FILE *f = fopen((base + ".dat").c_str(), "a");
    if (!f) {
      host_error("Can't open %s.dat", base.c_str());
      return FG_E_IO;
    }
    fprintf(f, "%d\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\n", TimeStep,
            x * fy_inner - y * fx_inner, x * fy_outer - y * fx_outer, x * fy_ex_inner - y * fx_ex_inner,
            x * fy_ex_outer - y * fx_ex_outer, vx * fx_inner + vy * fy_inner, vx * fx_outer + vy * fy_outer,
            vx * fx_ex_inner + vy * fy_ex_inner, vx * fx_ex_outer + vy * fy_ex_outer, s->PhysicalTime);
    fclose(f);
    if (s->d.SelfGravity) continue;
    for (int k = 0; k < dimfxy; k++) {
      std::string path = base + "_" + std::to_string(k) + ".dat";
      f = fopen(path.c_str(), "a");
      if (!f) {
        host_error("Can't open %s", path.c_str());
        return FG_E_IO;
      }
      fprintf(f, "%d\t%.18g\t%.18g\t%.18g\t%.18g\t%.18g\n", TimeStep, x * gf[2 * dimfxy + k] - y * gf[k],
              x * gf[3 * dimfxy + k] - y * gf[dimfxy + k], vx * gf[k] + vy * gf[2 * dimfxy + k],
              vx * gf[dimfxy + k] + vy * gf[3 * dimfxy + k], s->PhysicalTime);
      fclose(f);
    }
  }
  return FG_OK;
}

// Restart from output set `idx`: the reference parses -s N (Main.cu:152-160)
// but never reads the files back.  The fields are the raw NRAD*NSEC FP64
// dumps of WriteDiskPolar (Output.cu:150-181), the planets, time, frame
// rate and lost mass the line with TimeStep == idx of each planetK.dat
// (Output.cu:110-131, %.18g: an exact round trip).  Outputs are written at
// DT boundaries, so no sub-cycle state is carried.
static fg_status read_polar(fg_sim *s, std::vector<double> &a, const char *name, int idx) {
  std::string path = s->d.OUTPUTDIR + name + "/" + name + std::to_string(idx) + ".dat";
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) {
    host_error("restart: cannot open '%s'", path.c_str());
    return FG_E_IO;
  }
  const size_t n = (size_t)s->nr * s->ns;
  const size_t got = fread(a.data(), sizeof(double), n, f);
  fclose(f);
  if (got != n) {
    host_error("restart: '%s' holds %zu of %zu values", path.c_str(), got, n);
    return FG_E_IO;
  }
  return FG_OK;
}

extern "C" fg_status fg_sim_restart(fg_sim *s, int idx) {
  if (!s || idx < 0) return FG_E_INVALID;
  const Deck &d = s->d;
  std::vector<double> dens(s->S), vr(s->S), vt(s->S), en(s->S, 0.0);
  // current state first: the ghost ring is not in the files
  CHECK(fg_download_state(s->ctx, dens.data(), vr.data(), vt.data(), d.Adiabatic ? en.data() : nullptr, nullptr,
                          nullptr));
  CHECK(read_polar(s, dens, "dens", idx));
  CHECK(read_polar(s, vr, "vrad", idx));
  CHECK(read_polar(s, vt, "vtheta", idx));
  if (d.Adiabatic) CHECK(read_polar(s, en, "energy", idx));
  double t = idx * d.NINTERM * d.DT, omega = s->OmegaFrame, lost = s->LostMass, mdcp = s->mdcp;
  Planets &p = s->ps;
  for (int k = 0; k < p.nb; k++) {
    std::string path = d.OUTPUTDIR + "planet" + std::to_string(k) + ".dat";
    FILE *f = fopen(path.c_str(), "r");
    if (!f) {
      host_error("restart: cannot open '%s'", path.c_str());
      return FG_E_IO;
    }
    char line[1024];
    bool found = false;
    while (fgets(line, sizeof line, f)) {
      int n = 0;
      double v[10] = {0.0};
      if (sscanf(line, "%d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &n, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5],
                 &v[6], &v[7], &v[8], &v[9]) >= 9 &&
          n == idx) {
        p.x[k] = v[0];
        p.y[k] = v[1];
        p.vx[k] = v[2];
        p.vy[k] = v[3];
        p.mass[k] = v[4];
        lost = v[5];
        t = v[6];
        omega = v[7];
        if (k == 0) mdcp = v[8];
        found = true;
      }
    }
    fclose(f);
    if (!found) {
      host_error("restart: no line %d in '%s'", idx, path.c_str());
      return FG_E_IO;
    }
  }
  s->PhysicalTime = t;
  s->prof.time = t;
  s->OmegaFrame = omega;
  s->LostMass = lost;
  s->in_dt = 0;
  s->results_ready = 0;
  s->state_gen++;
  s->crashed = 0;
  s->planets_fresh = 1;  // keep the planet files: their lines up to idx are this run's history
  CHECK(fg_upload_state(s->ctx, dens.data(), vr.data(), vt.data(), d.Adiabatic ? en.data() : nullptr));
  CHECK(fg_sound_speed(s->ctx));
  CHECK(fg_pressure(s->ctx));
  CHECK(fg_temperature(s->ctx));
  if (d.ViscosityAlpha) CHECK(fg_ring_mean_cs(s->ctx));
  if (s->opt.diagnostics) {
    // the initial circumplanetary mass comes back from the planet file;
    // the latest one is that of the state just restored (the last substep
    // before output idx ended there), 0 before the first substep
    s->mdcp = mdcp;
    s->mdcp1 = 0.0;
    if (idx > 0 && d.IsDisk) CHECK(circumplanetary_mass(s, &s->mdcp1));
    s->exces_mdcp = s->mdcp1 - s->mdcp;
    if (idx == 0) s->exces_mdcp = 0.0;
  }
  return fg_sync(s->ctx);
}

// main()'s loop, Main.cu:309-333
extern "C" fg_status fg_sim_run(fg_sim *s, int it0, int it1) {
  if (!s) return FG_E_INVALID;
  for (int i = it0; i <= it1; i++) {
    if (s->opt.write_files && s->opt.diagnostics) {
      // Main.cu:312-316 run before TimeStep is updated for iteration i, so
      // they carry the output number of iteration i - 1 (0 at the start)
      int prev = i > 0 ? (i - 1) / s->d.NINTERM : 0;
      CHECK(write_bigplanet(s, prev));
      CHECK(update_log(s, prev));
    }
    int TimeStep = i / s->d.NINTERM;
    if (s->opt.write_files && s->d.NINTERM * TimeStep == i) CHECK(fg_sim_write_outputs(s, TimeStep));
    CHECK(fg_sim_algogas(s));
  }
  return FG_OK;
}

extern "C" fg_status fg_sim_get_state(fg_sim *s, double *dens, double *vrad, double *vtheta, double *energy) {
  if (!s) return FG_E_INVALID;
  return fg_download_state(s->ctx, dens, vrad, vtheta, energy, nullptr, nullptr);
}

extern "C" fg_status fg_sim_set_state(fg_sim *s, const double *dens, const double *vrad, const double *vtheta,
                                      const double *energy) {
  if (!s) return FG_E_INVALID;
  s->results_ready = 0;
  s->state_gen++;
  s->crashed = 0;
  return fg_upload_state(s->ctx, dens, vrad, vtheta, energy);
}

extern "C" fg_status fg_sim_get(fg_sim *s, const char *name, double *v) {
  if (!s || !name || !v) return FG_E_INVALID;
  struct {
    const char *k;
    double val;
  } tab[] = {{"PhysicalTime", s->PhysicalTime}, {"OmegaFrame", s->OmegaFrame},
             {"LostMass", s->LostMass},         {"dt", s->dt},
             {"dtemp", s->dtemp},               {"MassTaper", s->MassTaper},
             {"DT", s->d.DT},                   {"NTOT", (double)s->d.NTOT},
             {"NINTERM", (double)s->d.NINTERM}, {"NRAD", (double)s->nr},
             {"NSEC", (double)s->ns},           {"last_cfl", (double)s->last_cfl},
             {"in_dt", (double)s->in_dt},       {"IndirectTerm.x", s->IndirectX},
             {"IndirectTerm.y", s->IndirectY},  {"HillRadius", s->HillRadius},
             {"Adiabatic", (double)s->d.Adiabatic}, {"substeps", (double)s->nsub}, {"SelfGravity", (double)s->d.SelfGravity},
             {"fused_active", (double)s->fused_active}, {"mdcp", s->mdcp}, {"mdcp1", s->mdcp1},
             {"crashed", (double)s->crashed},
             {"exces_mdcp", s->exces_mdcp}};
  for (auto &e : tab)
    if (strcmp(e.k, name) == 0) {
      *v = e.val;
      return FG_OK;
    }
  host_error("unknown scalar '%s'", name);
  return FG_E_INVALID;
}

extern "C" fg_status fg_sim_nplanets(fg_sim *s, int *nb) {
  if (!s || !nb) return FG_E_INVALID;
  *nb = s->ps.nb;
  return FG_OK;
}

extern "C" fg_status fg_sim_planet(fg_sim *s, int k, double *o) {
  if (!s || !o || k < 0 || k >= s->ps.nb) return FG_E_INVALID;
  o[0] = s->ps.x[k];
  o[1] = s->ps.y[k];
  o[2] = s->ps.vx[k];
  o[3] = s->ps.vy[k];
  o[4] = s->ps.mass[k];
  return FG_OK;
}

extern "C" fg_ctx *fg_sim_ctx(fg_sim *s) { return s ? s->ctx : nullptr; }

extern "C" fg_status fg_sim_ring(fg_sim *s, const char *name, double *out) {
  if (!s || !name || !out) return FG_E_INVALID;
  const std::vector<double> *v = nullptr;
  if (!strcmp(name, "Radii")) v = &s->Radii;
  if (!strcmp(name, "Rinf")) v = &s->Rinf;
  if (!strcmp(name, "Rsup")) v = &s->Rsup;
  if (!strcmp(name, "Rmed")) v = &s->Rmed;
  if (!strcmp(name, "Surf")) v = &s->Surf;
  if (!strcmp(name, "SigmaMed")) v = &s->SigmaMed;
  if (!strcmp(name, "EnergyMed")) v = &s->EnergyMed;
  if (!v) {
    host_error("unknown ring array '%s'", name);
    return FG_E_INVALID;
  }
  memcpy(out, v->data(), v->size() * sizeof(double));
  return FG_OK;
}

// ------------------------------------------------------ ensemble batch --
// Independent disks (each its own deck, FP32 bulk mode) advanced together:
// per batch substep every disk short of the target time takes one substep
// of its own sub-cycle (its own CFL-limited dt); the host part of each
// (dt, planets, frame) runs here in FP64 and the GPU part of all of them
// is one captured graph (fg_fused.hip).  A disk that has reached the
// target idles.
struct fg_batch_dev;
fg_status fg_batch_dev_create(fg_ctx *const *ctx, int n, fg_batch_dev **out);
void fg_batch_dev_destroy(fg_batch_dev *b);
fg_status fg_batch_dev_set(fg_batch_dev *b, int k, int active, double dt, double omega_frame, double domega,
                           const fg_planet_pot *planets, int nb, double ix, double iy, double DeltaT_next);
fg_status fg_batch_dev_launch(fg_batch_dev *b);
fg_status fg_batch_dev_wait(fg_batch_dev *b);
void fg_batch_dev_result(const fg_batch_dev *b, int k, int *cfl, double *force2);
void fg_batch_dev_invalidate(fg_batch_dev *b);
fg_status fg_fused_check(fg_ctx *c, double cfl, const double *force2);

struct fg_batch {
  std::vector<fg_sim *> sims;
  fg_batch_dev *dev = nullptr;
  std::vector<int> cfl;
  std::vector<double> force;
  std::vector<fused_launch> L;
  std::vector<char> active;
  std::vector<unsigned long> seen_gen;  // state generation of each disk at its last batch step
  int started = 0;
};

extern "C" fg_status fg_batch_destroy(fg_batch *b) {
  if (!b) return FG_OK;
  fg_batch_dev_destroy(b->dev);
  for (fg_sim *s : b->sims) fg_sim_destroy(s);
  delete b;
  return FG_OK;
}

extern "C" fg_status fg_batch_create(int n, const char *const *parfiles, const fg_sim_options *opt,
                                     fg_batch **out) {
  if (!out || !parfiles || n <= 0) {
    host_error("fg_batch_create: bad argument");
    return FG_E_INVALID;
  }
  *out = nullptr;
  if (opt && opt->diagnostics) {
    host_error("fg_batch_create: diagnostics (circumplanetary mass, torque logs) are per-disk runs only");
    return FG_E_UNSUPPORTED;
  }
  fg_sim_options o = opt ? *opt : fg_sim_options{};
  o.precision = FG_PREC_FP32;
  o.fused = 1;
  fg_batch *b = new fg_batch();
  for (int k = 0; k < n; k++) {
    fg_sim *s = nullptr;
    fg_status st = fg_sim_create(parfiles[k], &o, &s);
    if (st) {
      fg_batch_destroy(b);
      return st;
    }
    b->sims.push_back(s);
  }
  std::vector<fg_ctx *> ctx(n);
  for (int k = 0; k < n; k++) ctx[k] = b->sims[k]->ctx;
  fg_status st = fg_batch_dev_create(ctx.data(), n, &b->dev);
  if (st) {
    fg_batch_destroy(b);
    return st;
  }
  b->cfl.assign(n, 1);
  b->force.assign(2 * (size_t)n, 0.0);
  b->L.resize(n);
  b->active.assign(n, 0);
  b->seen_gen.assign(n, 0);
  *out = b;
  return FG_OK;
}

extern "C" fg_status fg_batch_set_tiles(fg_batch *b, int k, int src_rows, int trad_band) {
  if (!b || k < 0 || k >= (int)b->sims.size()) return FG_E_INVALID;
  CHECK(fg_set_tiles(b->sims[k]->ctx, src_rows, trad_band));
  fg_batch_dev_invalidate(b->dev);
  return FG_OK;
}

// A disk whose state blows up (FG_E_CRASH: a negative or non-finite
// density, or a CFL count or star force showing a time step gone to 0,
// infinity or NaN) is marked crashed and left alone; the other disks go
// on.  The batch substep in which that happens returns FG_E_CRASH naming
// the disks; fg_sim_get(disk, "crashed") tells them apart afterwards.
static void batch_crash(fg_batch *b, int k, std::string &msg, int &ncrash) {
  b->sims[k]->crashed = 1;
  b->active[k] = 0;
  if (ncrash++ < 8) {
    if (!msg.empty()) msg += "; ";
    msg += "batch disk " + std::to_string(k) + ": " + fg_last_error();
  }
}

static fg_status batch_crash_status(const std::string &msg, int ncrash) {
  if (!ncrash) return FG_OK;
  host_error("%s%s", msg.c_str(), ncrash > 8 ? " (and more)" : "");
  return FG_E_CRASH;
}

extern "C" fg_status fg_batch_substep(fg_batch *b, double t_end, int *n_active) {
  if (!b) return FG_E_INVALID;
  const int n = (int)b->sims.size();
  std::string crash_msg;
  int ncrash = 0;
  // CFL count and star force of every initial state, and again for a disk
  // whose state changed since its last batch step (fg_sim_set_state,
  // fg_sim_restart or single-disk substeps through fg_batch_sim)
  for (int k = 0; k < n; k++) {
    fg_sim *s = b->sims[k];
    if (s->crashed || (b->started && s->state_gen == b->seen_gen[k])) continue;
    fused_open(s);
    CHECK(fg_fused_prepare(s->ctx, s->d.DT - s->dtemp));
    fg_status st = fg_fused_results(s->ctx, &b->cfl[k], &b->force[2 * k], nullptr);
    if (st == FG_E_CRASH) {
      batch_crash(b, k, crash_msg, ncrash);
    } else if (st) {
      const std::string why = fg_last_error();  // the message buffer is rewritten below
      host_error("batch disk %d: %s", k, why.c_str());
      return st;
    }
    b->seen_gen[k] = s->state_gen;
  }
  b->started = 1;
  int na = 0;
  for (int k = 0; k < n; k++) {
    fg_sim *s = b->sims[k];
    // a disk stops at the first DT boundary at or after t_end
    const bool act = !s->crashed && (s->in_dt || s->PhysicalTime < t_end - 1e-9 * s->d.DT);
    b->active[k] = act;
    if (!act) {
      CHECK(fg_batch_dev_set(b->dev, k, 0, 0, 0, 0, nullptr, 0, 0, 0, 0));
      continue;
    }
    na++;
    fused_open(s);
    fused_launch &L = b->L[k];
    s->pf_ready = 0;  // batch disks take their planets' forces synchronously
    CHECK(fused_plan(s, b->cfl[k], &b->force[2 * k], L));
    CHECK(fg_batch_dev_set(b->dev, k, 1, L.dt, s->OmegaFrame, L.domega, L.pl, L.nb, s->IndirectX, s->IndirectY,
                           L.DeltaT_next));
  }
  if (n_active) *n_active = na;
  if (!na) return batch_crash_status(crash_msg, ncrash);
  CHECK(fg_batch_dev_launch(b->dev));
  CHECK(fg_batch_dev_wait(b->dev));
  for (int k = 0; k < n; k++) {
    if (!b->active[k]) continue;
    fg_batch_dev_result(b->dev, k, &b->cfl[k], &b->force[2 * k]);
    fused_close(b->sims[k], b->L[k], nullptr);
    b->sims[k]->results_ready = 0;  // single-disk stepping re-derives them
    b->seen_gen[k] = b->sims[k]->state_gen;
  }
  for (int k = 0; k < n; k++) {
    if (!b->active[k]) continue;
    fg_status st = fg_fused_check(b->sims[k]->ctx, (double)b->cfl[k], &b->force[2 * k]);
    if (st) batch_crash(b, k, crash_msg, ncrash);
  }
  return batch_crash_status(crash_msg, ncrash);
}

// batch substeps until every disk that has not crashed reached t_end;
// FG_E_CRASH at the end when disks crashed on the way (the others finished)
extern "C" fg_status fg_batch_run(fg_batch *b, double t_end, long *substeps) {
  if (!b) return FG_E_INVALID;
  long count = 0;
  int na = 1;
  std::string crash_msg;
  while (na) {
    fg_status st = fg_batch_substep(b, t_end, &na);
    if (st == FG_E_CRASH) {
      if (!crash_msg.empty()) crash_msg += "; ";
      crash_msg += fg_last_error();
    } else if (st) {
      return st;
    }
    if (na) count++;
  }
  if (substeps) *substeps = count;
  if (!crash_msg.empty()) {
    host_error("%s", crash_msg.substr(0, 900).c_str());
    return FG_E_CRASH;
  }
  return FG_OK;
}

extern "C" fg_sim *fg_batch_sim(fg_batch *b, int k) {
  if (!b || k < 0 || k >= (int)b->sims.size()) return nullptr;
  return b->sims[k];
}
