/*
 * fargo_hip.h -- C-ABI of the MI355X (gfx950) FARGO / FARGO-ADSG hydro
 * substep library (libfargo_hip.so).
 *
 * The reference (MatiasFloresS/CudaProtoplanetaryDisks) has no plugin or
 * FFI layer: its substep path is a set of host functions operating on
 * file-scope device globals (include/SourceEuler.cuh:1-22,
 * TransportEuler.cuh:1-19, SideEuler.cuh:1-15, Viscosity.cuh:1-7,
 * SgMain.cuh:1-9, Pframeforce.cuh:1-8, Force.cuh:1-7).  Each entry point
 * below replaces one of those functions; the comment names it (file:line).
 * Differences from the reference boundary:
 *   - state lives in an opaque fg_ctx (one per host thread, one HIP stream
 *     each) instead of globals; the library owns all device memory;
 *   - every call returns an fg_status; nothing calls exit();
 *   - calls are stream-ordered; only downloads and fg_sync block.
 * Layout of every 2D field: FP64 row-major [(NRAD+1) x NSEC], index
 * i*NSEC + j with i the ring (i = NRAD is the ghost ring) and j the sector,
 * exactly the reference's layout (Main.cu:203).
 */
#ifndef FARGO_HIP_H
#define FARGO_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int fg_status;
#define FG_OK 0
#define FG_E_INVALID 1     /* bad argument or NULL context */
#define FG_E_HIP 2         /* HIP runtime error (see fg_last_error) */
#define FG_E_UNSUPPORTED 3 /* option not implemented on this path */
#define FG_E_PARAM 4       /* deck rejected (the reference exit(1)s) */
#define FG_E_IO 5          /* file could not be read or written */
#define FG_E_CRASH 6       /* negative or non-finite density/energy */
#define FG_E_NOMEM 7       /* device or host allocation failed */

/* planets per context (fg_potential) */
#define FG_MAX_PLANETS 16

#define FG_PREC_FP64 0 /* all fields and arithmetic in FP64 */
#define FG_PREC_FP32 1 /* FP32 bulk fields; FP64 planets, residual, sums */

typedef struct fg_ctx fg_ctx;

/* Physics switches and constants a context needs.  Filled by the host
 * driver from a parsed deck (fg_sim_create does this). */
typedef struct fg_params {
  int nrad, nsec;
  int precision;        /* FG_PREC_* */
  int adiabatic;        /* Interpret.cu:163-172 */
  int cooling;          /* Interpret.cu:167 */
  int self_gravity;     /* Interpret.cu:150-154 */
  int viscosity_alpha;  /* Interpret.cu:184-187 */
  int fast_transport;   /* Interpret.cu:132 (FARGO = 1) */
  int open_inner;       /* Interpret.cu:133 */
  int non_reflecting;   /* Interpret.cu:134 */
  int evanescent;       /* Interpret.cu:135 */
  int indirect_term;    /* Interpret.cu:149 */
  int exclude_hill;     /* Interpret.cu:174 */
  double adiabatic_index;
  double alphaviscosity;
  double sigmaslope;
  double flaringindex;
  double sg_eps, sg_rstep, sg_tstep; /* Main.cu:269-271 */
} fg_params;

/* Host-computed 1D tables.  Every transcendental the kernels need is
 * evaluated once on the host (glibc) and uploaded, so device and host
 * arithmetic see identical inputs.  Ring arrays have NRAD+1 entries,
 * sector arrays NSEC entries. */
typedef struct fg_tables {
  /* FillPolar1DArrays, SourceEuler.cu:66-162 */
  const double *Radii, *Rinf, *Rsup, *Rmed, *Surf;
  const double *invRinf, *invSurf, *invRmed, *invdiffRsup, *invdiffRmed;
  /* Theo.cu profiles */
  const double *SigmaMed, *SigmaInf, *EnergyMed, *CoolingTimeMed, *QplusMed;
  /* per-ring kernel constants (same expressions as the kernels) */
  const double *cs_iso;      /* isothermal c_s, Kernels.cu:179 */
  const double *supp_torque; /* Kernels.cu:40 */
  const double *dxtheta;     /* 2 pi / NSEC * Rmed, Kernels.cu:41 */
  const double *invdxtheta;  /* 1 / dxtheta */
  const double *dxtheta_cfl; /* Rmed * 2 pi / NSEC, Kernels.cu:1295 */
  /* evanescent damping, Kernels.cu:978-1004 */
  const int *ev_zone;          /* 1 inside a wave-killing zone */
  const double *ev_coef;       /* damping^2 * 10 */
  const double *ev_period;     /* Tin or Tout */
  const double *ev_vtheta_arg; /* G/r (1 - ...) argument of the sqrt */
  /* FViscosity, Viscosity.cu:33-50 */
  const double *visc_const; /* nu for constant viscosity, before cavity */
  const double *visc_pow15; /* r^1.5 for alpha viscosity */
  const int *visc_cav_on;   /* cavity factor applies */
  const double *visc_cav;   /* cavity factor */
  /* sector tables */
  const double *cos_cell, *sin_cell; /* cos/sin((2 pi j)/NSEC), Kernels.cu:161 */
  const double *cos_pot, *sin_pot;   /* cos/sin(j/NSEC*2 pi), Kernels.cu:1236 */
  /* self-gravity kernel tables, 2*NRAD x NSEC (SgMain.cu:49-81) */
  const double *sg_Kr, *sg_Kt;
  /* scalars */
  double dtheta;      /* 2 pi / NSEC */
  double invdphi;     /* 1 / (2 M_PI / NSEC), Kernels.cu:742-743 */
  double dt1d_coef;   /* 2 pi CFLSECURITY / NSEC, Kernels.cu:1332 */
  double vrad0_slope; /* -SIGMASLOPE + .5, Kernels.cu:1002 */
  double visc_value;  /* VISCOSITY */
} fg_tables;

/* one planet's contribution to the potential, as prepared by
 * FillForcesArrays (Pframeforce.cu:68-78) */
typedef struct fg_planet_pot {
  double x, y;
  double mass;   /* mass * MassTaper */
  double smooth; /* smoothing length squared */
  double inv_d3; /* 1 / |r_p|^3 */
} fg_planet_pot;

/* field identifiers for fg_get_field / fg_set_field */
enum fg_field {
  FG_DENS = 0, FG_VRAD, FG_VTHETA, FG_ENERGY, FG_PRESSURE, FG_SOUNDSPEED, FG_TEMPERATURE,
  FG_POTENTIAL, FG_VRADINT, FG_VTHETAINT, FG_VRADNEW, FG_VTHETANEW, FG_ENERGYINT,
  FG_ENERGYNEW, FG_DENSINT, FG_TEMPERINT, FG_QPLUS, FG_DIVV, FG_TAURR, FG_TAUPP, FG_TAURP,
  FG_RADMOMP, FG_RADMOMM, FG_THETAMOMP, FG_THETAMOMM, FG_VTHETARES, FG_SG_ACCR, FG_SG_ACCT,
  FG_NFIELDS
};

/* ---- context ---------------------------------------------------------- */
/* Allocates device memory (Cudamalloc Main.cu:660, Computecudamalloc
 * SourceEuler.cu:378, InitTransportDevice TransportEuler.cu:291,
 * InitViscosityDevice Viscosity.cu:81, InitComputeAccelDevice
 * SideEuler.cu:173) and uploads the tables. */
fg_status fg_create(const fg_params *p, const fg_tables *t, int device, fg_ctx **out);
/* FreeCuda, Main.cu:367 */
fg_status fg_destroy(fg_ctx *ctx);
const char *fg_last_error(void);
fg_status fg_sync(fg_ctx *ctx);

/* Host <-> device state.  NULL pointers are skipped.  Arrays are
 * (NRAD+1)*NSEC doubles (Main.cu:811-825 / DeviceToHostcudaMemcpy
 * Main.cu:605-616). */
fg_status fg_upload_state(fg_ctx *ctx, const double *dens, const double *vrad, const double *vtheta,
                          const double *energy);
fg_status fg_download_state(fg_ctx *ctx, double *dens, double *vrad, double *vtheta, double *energy,
                            double *temperature, double *qplus);
fg_status fg_get_field(fg_ctx *ctx, int field, double *host);
fg_status fg_set_field(fg_ctx *ctx, int field, const double *host);
/* ring arrays (NRAD entries): Vmoy, VMed, viscosity, ring mean of c_s,
 * ring mean of SG_Accr */
fg_status fg_get_rings(fg_ctx *ctx, double *vmoy, double *vmed, double *viscosity, double *cs_mean,
                       double *sgacc_mean);
/* Nshift of the last transport (NRAD ints), ComputeConstantResidualKernel
 * Kernels.cu:1114 */
fg_status fg_get_nshift(fg_ctx *ctx, int *nshift);

/* ---- the substep path, one entry per reference routine --------------- */
/* ComputeSoundSpeed, SourceEuler.cu:475 */
fg_status fg_sound_speed(fg_ctx *ctx);
/* ComputePressureField, SourceEuler.cu:467 */
fg_status fg_pressure(fg_ctx *ctx);
/* ComputeTemperatureField, SourceEuler.cu:484 */
fg_status fg_temperature(fg_ctx *ctx);
/* Make1Dprofile(1) + FViscosity, Kernels.cu:626 / Viscosity.cu:33: ring
 * mean of c_s and the viscosity profile derived from it */
fg_status fg_ring_mean_cs(fg_ctx *ctx);
/* ConditionCFL, SourceEuler.cu:520: writes the sub-cycle count */
fg_status fg_cfl(fg_ctx *ctx, double DeltaT, int *ncfl);
/* ComputeAccel, SideEuler.cu:135 (Force.cu:102): disk gravity on a point
 * mass at (x, y) with smoothing rsmoothing, summed in DeviceReduce order
 * (Kernels.cu:364).  out2 = {ax, ay}. */
fg_status fg_disk_force(fg_ctx *ctx, double x, double y, double rsmoothing, double *out2);
/* ComputeForce, Force.cu:102-135 (ComputeForceKernel Kernels.cu:693-729):
 * disk gravity on a point mass at (x, y), split into cells inside
 * (Rmed < a) and outside r = a, for the dimfxy Hill-cut variants
 * k = 0..dimfxy-1 (weight 1 - exp(-d^2 / (rh k/(dimfxy-1))^2), none for
 * k = 0), each of the 4 dimfxy sums in DeviceReduce order.  globalforce
 * receives the reference's GlobalForce layout: [k] fx inner,
 * [dimfxy + k] fx outer, [2 dimfxy + k] fy inner, [3 dimfxy + k] fy outer.
 * 1 <= dimfxy <= 64.  Used by UpdateLog (Force.cu:19). */
fg_status fg_disk_force_hill(fg_ctx *ctx, double x, double y, double rsmoothing, int dimfxy, double a, double rh,
                             double *globalforce);
/* No counterpart in the reference: fg_disk_force_hill for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It allocates nothing:
 * the block partial sums go to a 2 MiB scratch array built into the
 * library (one per GPU, kept until the process ends), and the
 * 4 * dimfxy results, same layout and bits as fg_disk_force_hill, go to
 * globalforce_dev, a device pointer the caller owns.  Asynchronous on the
 * context's stream (fg_sync waits).  dimfxy <= FG_HILL_MAX (64); every grid
 * fits.  Calls that overlap on one GPU share the scratch array and race. */
fg_status fg_disk_force_hill_noalloc(fg_ctx *ctx, double x, double y, double rsmoothing, int dimfxy, double a,
                                     double rh, double *globalforce_dev);
/* CircumPlanetaryMass, SourceEuler.cu:549-563 (Kernels.cu:236-251): disk
 * mass within hill_radius of (xpl, ypl), in DeviceReduce order */
fg_status fg_circumplanetary_mass(fg_ctx *ctx, double xpl, double ypl, double hill_radius, double *mass);
/* FillForcesArrays, Pframeforce.cu:56 (+ ComputeIndirectTerm :86) */
fg_status fg_potential(fg_ctx *ctx, const fg_planet_pot *planets, int nb, double indirect_x,
                       double indirect_y);
/* CorrectVtheta, SideEuler.cu:181 */
fg_status fg_correct_vtheta(fg_ctx *ctx, double domega);
/* ApplyBoundaryCondition, SideEuler.cu:31 */
fg_status fg_boundary(fg_ctx *ctx, double dt, double omega_frame);
/* Substep1, SourceEuler.cu:311 (incl. self-gravity kick, viscous stress
 * Viscosity.cu:52 and its velocity update Viscosity.cu:25, and the
 * sub-Keplerian boundary SideEuler.cu:189) */
fg_status fg_substep1(fg_ctx *ctx, double dt, double omega_frame);
/* Substep2 + host + ActualiseGasVrad/Vtheta, SourceEuler.cu:345-358,
 * 492-503 */
fg_status fg_substep2(fg_ctx *ctx, double dt);
/* ComputeViscousTerms + Substep3 + ActualiseGasEnergy,
 * SourceEuler.cu:287-291, 362 */
fg_status fg_substep3(fg_ctx *ctx, double dt);
/* Transport, TransportEuler.cu:30; lost_mass (may be NULL) receives the
 * mass that left through the open inner edge (TransportEuler.cu:77) */
fg_status fg_transport(fg_ctx *ctx, double dt, double omega_frame, double *lost_mass);
/* compute_selfgravity, SgMain.cu:23 (update = 0 evaluates only) */
fg_status fg_selfgravity(fg_ctx *ctx, double dt, int update);

/* ---- fused substep (same arithmetic, fewer passes over HBM) ---------- */
/* Everything AlgoGas does on the GPU for one substep after the planets
 * are advanced (SourceEuler.cu:255-295): CorrectVtheta (domega, 0 when
 * the frame does not co-rotate), the potential (FillForcesArrays), the
 * three boundary passes, pressure, Substep1/2, the self-gravity kick
 * (compute_selfgravity inside Substep1, SourceEuler.cu:322-330), Substep3
 * of an adiabatic deck, transport and, with FG_STEP_TEMPERATURE in flags,
 * the temperature.  Also leaves the CFL sub-cycle count for DeltaT_next
 * (the next substep's ConditionCFL, SourceEuler.cu:228) and the disk force
 * on the star for it (SourceEuler.cu:241), fetched by fg_fused_results.
 * FG_STEP_CLOSES_DT marks the last substep of a DT: an adiabatic deck then
 * recomputes the sound speed from the new state (AlgoGas does it when the
 * next DT opens, SourceEuler.cu:207-213) for that CFL count and the next
 * DT's alpha viscosity.  Decks with evanescent edges; isothermal or
 * adiabatic, with or without self-gravity (FP32 bulk mode: isothermal
 * without self-gravity). */
enum { FG_STEP_TEMPERATURE = 1, FG_STEP_CLOSES_DT = 2 };
fg_status fg_step_fused(fg_ctx *ctx, double dt, double omega_frame, double domega,
                        const fg_planet_pot *planets, int nb, double indirect_x, double indirect_y,
                        double DeltaT_next, int flags);
/* CFL count and star force of the current state (fused reductions) */
fg_status fg_fused_prepare(fg_ctx *ctx, double DeltaT);
fg_status fg_fused_results(fg_ctx *ctx, int *ncfl, double *force2, double *lost_mass);
/* Tile shapes of the fused kernels: rings per FP32 source tile (8, 16
 * or 32) and rings per radial-sweep band.  Results do not depend on them; the
 * autotuner (cudaprotoplanetarydisks_amd/autotune.py) picks them per grid.
 * No reference counterpart (the reference's launch shape is fixed,
 * Main.cu:204-226). */
fg_status fg_set_tiles(fg_ctx *ctx, int src_rows, int trad_band);
/* FP32 bulk mode: Sigma, between the fused kernels, is stored as a pair of
 * floats (high part and rounding residual) on the rings within hill_radii
 * Hill radii (radially) of a planet, and as one float elsewhere; 1 by
 * default, 0 stores single floats everywhere.  The pairs keep the rounding
 * of the stored Sigma from adding up in circumplanetary material.  No
 * reference counterpart (the reference computes in FP64 only). */
fg_status fg_set_fp32_pairs(fg_ctx *ctx, double hill_radii);
/* No counterpart in the reference: a self-check of the FP64 division with a
 * shared denominator that the fused kernels use (one reciprocal refinement
 * of d for several numerators).  Computes num[k] / den[k] both ways on the
 * GPU for k < n (host arrays) and stores the number of results whose bits
 * differ in *mismatches.  The two agree for d and 1/d normal and numerators
 * that are zero, infinite, NaN or finite with 2^-969 <= |n| and
 * |n / d| < 2^767.  For tests. */
fg_status fg_check_shared_division(const double *num, const double *den, long n, long *mismatches);

/* ---- tracing ---------------------------------------------------------- */
/* Kernel timing of the fused path with HIP events on the context's
 * stream: enable = 1 starts (and clears) it, 0 stops it.
 * fg_profile_read fills ms[k] (total milliseconds) and calls[k] for the
 * kernels k = 0 k_src, 1 k_trad, 2 k_taz, 3 k_fin, 4 disk-force reduction,
 * 5 CFL count; returns the number of kernels.  No reference counterpart
 * (the reference only prints the wall time of the whole loop,
 * Main.cu:308-337). */
fg_status fg_profile(fg_ctx *ctx, int enable);
int fg_profile_read(fg_ctx *ctx, double *ms, long *calls, int max);

/* ---- host driver: the reference's main()/AlgoGas over this library ---- */
typedef struct fg_sim fg_sim;
typedef struct fg_sim_options {
  int device;      /* HIP device ordinal */
  int precision;   /* FG_PREC_* */
  int fused;       /* 1 = fg_step_fused path, 0 = one call per routine */
  int write_files; /* used_rad.dat, dims.dat, planet files, outputs */
  const char *outdir;    /* overrides OUTPUTDIR when not NULL ("-o") */
  const char *cli_flags; /* single-letter switches of Main.cu:113-179 */
  unsigned long long noise_seed; /* 0 = none; else Sigma *= 1 + amp*xi */
  double noise_amp;
  /* 1 = the diagnostics the reference ships commented out: the initial
   * circumplanetary mass (Main.cu:302), its value after every substep
   * (SourceEuler.cu:297-298, written to the planet files), and per main
   * loop iteration bigplanetK.dat (Main.cu:314, Output.cu:90-107) and the
   * torque/power logs tqwkK.dat, tqwkK_k.dat (Main.cu:315, Force.cu:19-84) */
  int diagnostics;
} fg_sim_options;

/* Main.cu:187-305 */
fg_status fg_sim_create(const char *parfile, const fg_sim_options *opt, fg_sim **out);
fg_status fg_sim_destroy(fg_sim *sim);
/* one iteration of AlgoGas's sub-cycle loop (SourceEuler.cu:222-305);
 * *closed = 1 when it completed the DT */
fg_status fg_sim_substep(fg_sim *sim, int *closed);
/* AlgoGas, SourceEuler.cu:199 */
fg_status fg_sim_algogas(fg_sim *sim);
/* main()'s loop for it = it0..it1 (Main.cu:309-333), with outputs */
fg_status fg_sim_run(fg_sim *sim, int it0, int it1);
fg_status fg_sim_write_outputs(fg_sim *sim, int index);
/* restart from output set `index` (dens/densN.dat, vrad, vtheta, energy
 * when adiabatic, and line N of every planetK.dat): what -s N
 * (Main.cu:152-160) announces; the reference never reads them back */
fg_status fg_sim_restart(fg_sim *sim, int index);
fg_status fg_sim_get_state(fg_sim *sim, double *dens, double *vrad, double *vtheta, double *energy);
fg_status fg_sim_set_state(fg_sim *sim, const double *dens, const double *vrad, const double *vtheta,
                           const double *energy);
/* scalars: "PhysicalTime", "OmegaFrame", "LostMass", "dt", "dtemp",
 * "MassTaper", "DT", "NTOT", "NINTERM", "NRAD", "NSEC", "last_cfl",
 * "mdcp", "mdcp1", "exces_mdcp" (diagnostics), "crashed" (batch disks) */
fg_status fg_sim_get(fg_sim *sim, const char *name, double *value);
fg_status fg_sim_nplanets(fg_sim *sim, int *nb);
/* out = {x, y, vx, vy, mass} */
fg_status fg_sim_planet(fg_sim *sim, int k, double *out5);
fg_ctx *fg_sim_ctx(fg_sim *sim);
/* 1D grid of the run (NRAD+1 entries): "Radii", "Rinf", "Rsup", "Rmed",
 * "Surf", "SigmaMed", "EnergyMed" */
fg_status fg_sim_ring(fg_sim *sim, const char *name, double *out);

/* ---- ensemble: many independent disks in one launch per phase ------ */
/* A parameter-sweep ensemble of FP32 bulk-mode disks (uneven grids).  Per
 * batch substep every disk that has not reached t_end takes one substep
 * of its own AlgoGas sub-cycle (SourceEuler.cu:222-305) with its own dt;
 * the GPU part of all disks is one replay of a captured hipGraph (each
 * fused phase one launch over a (disk, tile) work list).  A disk stops at
 * the first DT boundary at or after t_end.  Each disk gives the same
 * result as the same deck run alone with precision FG_PREC_FP32.  The
 * reference runs one disk per process (Main.cu:89); no counterpart. */
typedef struct fg_batch fg_batch;
fg_status fg_batch_create(int n, const char *const *parfiles, const fg_sim_options *opt, fg_batch **out);
fg_status fg_batch_destroy(fg_batch *b);
fg_status fg_batch_set_tiles(fg_batch *b, int k, int src_rows, int trad_band);
/* one batch substep; *n_active = disks that moved (0: all reached t_end).
 * A disk whose state blows up is marked crashed (fg_sim_get "crashed") and
 * no longer advanced; the substep in which that happens returns
 * FG_E_CRASH naming it, and the other disks go on. */
fg_status fg_batch_substep(fg_batch *b, double t_end, int *n_active);
/* batch substeps until every disk that has not crashed reached t_end;
 * FG_E_CRASH at the end when some crashed on the way */
fg_status fg_batch_run(fg_batch *b, double t_end, long *substeps);
/* disk k (state, scalars, planets through the fg_sim_* calls) */
fg_sim *fg_batch_sim(fg_batch *b, int k);

/* ---- learned emulator: MLP on the matrix cores ---------------------- */
/* Dense layers widths[0] -> widths[1] -> ... -> widths[nlayers], GELU on
 * the hidden layers, BF16 weights/activations with FP32 accumulation
 * (v_mfma_f32_16x16x32_bf16), FP32 master weights and Adam moments, MSE
 * loss.  Parameters are one flat FP32 vector: for each layer W (out x in,
 * row-major) then b.  Maps disk parameters to ring-averaged profiles for
 * warm starts; no reference counterpart. */
typedef struct fg_mlp fg_mlp;
fg_status fg_mlp_create(const int *widths, int nlayers, int max_batch, unsigned long long seed, int device,
                        fg_mlp **out);
fg_status fg_mlp_destroy(fg_mlp *mlp);
long fg_mlp_nparam(const fg_mlp *mlp);
fg_status fg_mlp_get_params(fg_mlp *mlp, float *params);
fg_status fg_mlp_set_params(fg_mlp *mlp, const float *params);
/* y (batch x outputs) = MLP(x (batch x inputs)), row-major FP32 */
fg_status fg_mlp_forward(fg_mlp *mlp, const float *x, int batch, float *y);
/* loss = mean((y - t)^2) and its gradient (flat, parameter order) */
fg_status fg_mlp_gradient(fg_mlp *mlp, const float *x, const float *t, int batch, float *grad, float *loss);
/* `steps` Adam steps (lr, beta 0.9/0.999) over n samples in batches of
 * `batch` rows; losses[s] = loss of step s (may be NULL) */
fg_status fg_mlp_train(fg_mlp *mlp, const float *x, const float *t, int n, int batch, int steps, float lr,
                       float *losses);
/* C = A B with A, B rounded to BF16, FP32 accumulation (the layer GEMM) */
fg_status fg_bf16_gemm(const float *A, const float *B, int M, int N, int K, float *C);

#ifdef __cplusplus
}
#endif
#endif
