/*
 * FP64 serial oracle of the FARGO / FARGO-ADSG hydro substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a CPU restatement of the reference
 * routines (MatiasFloresS/CudaProtoplanetaryDisks, src/*.cu) in their own
 * operation order.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline;
 * the shipped library never calls it.
 *
 * Every routine cites the reference file:line it restates.  Parity status
 * of each part is listed in fargo_oracle.c's header comment.
 */
#ifndef FARGO_ORACLE_H
#define FARGO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_sim or_sim;

/* Main.cu:89-305: parse flags + deck, build grid, planets and initial
 * state.  cli_flags holds the single-letter switches of Main.cu:113-179
 * ("c" = SloppyCFL, ...); outdir (may be NULL) overrides OUTPUTDIR like
 * "-o"; write_files = 0 suppresses used_rad.dat / dims.dat / planet files. */
or_sim *or_create(const char *parfile, const char *cli_flags,
                  const char *outdir, int write_files);
void or_destroy(or_sim *s);
const char *or_last_error(void);

/* One iteration of AlgoGas's CFL sub-cycle (SourceEuler.cu:222-305).
 * A new DT is opened automatically (SourceEuler.cu:202-221).  Returns 1
 * when the substep closed the DT, 0 otherwise. */
int or_substep(or_sim *s);
/* One full AlgoGas call (SourceEuler.cu:199-307). */
void or_algogas(or_sim *s);
/* main() time loop (Main.cu:309-333) for it = it0 .. it1 inclusive, with
 * outputs every NINTERM. */
void or_run(or_sim *s, int it0, int it1);
/* SendOutput + WritePlanetSystemFile for output number idx. */
void or_write_outputs(or_sim *s, int idx);

/* accessors: internal arrays of the simulation, size (NRAD+1)*NSEC for 2D
 * fields, NRAD+1 for ring arrays (see or_field_len). */
int or_nrad(const or_sim *s);
int or_nsec(const or_sim *s);
double *or_field(or_sim *s, const char *name);
long or_field_len(or_sim *s, const char *name);
int *or_ifield(or_sim *s, const char *name);
double or_get(or_sim *s, const char *name);
int or_set(or_sim *s, const char *name, double v);
int or_iget(or_sim *s, const char *name);
int or_nplanets(const or_sim *s);
/* out = {x, y, vx, vy, mass} */
void or_planet(const or_sim *s, int k, double *out);
void or_set_planet(or_sim *s, int k, const double *in);

/* Single reference routines, for per-step parity checks. */
int or_condition_cfl(or_sim *s, double DeltaT);              /* SourceEuler.cu:520 */
void or_compute_accel(or_sim *s, double x, double y, double rsmoothing,
                      double mass, double *out2);            /* SideEuler.cu:135 */
double or_device_reduce(const double *in, int n);            /* Kernels.cu:364 */
/* ComputeForce, Force.cu:102-135: 4*dimfxy sums, GlobalForce layout */
void or_compute_force_hill(or_sim *s, double x, double y, double rsmoothing, int dimfxy,
                           double a, double rh, double *globalforce);
double or_circumplanetary_mass_at(or_sim *s, double xpl, double ypl,
                                  double hill);              /* Kernels.cu:236 */
double or_circumplanetary_mass(or_sim *s);                   /* SourceEuler.cu:549 */
/* turn on the diagnostics the reference ships commented out (Main.cu:302,
 * 314-315; SourceEuler.cu:297-298); call right after or_create */
void or_enable_diagnostics(or_sim *s);
void or_update_log(or_sim *s, int TimeStep);                 /* Force.cu:19 */
void or_write_bigplanet(or_sim *s, int TimeStep);            /* Output.cu:62 */
void or_fill_forces_arrays(or_sim *s);                       /* Pframeforce.cu:56 */
void or_apply_boundary(or_sim *s, double dt);                /* SideEuler.cu:31 */
void or_correct_vtheta(or_sim *s, double domega);            /* SideEuler.cu:181 */
void or_sound_speed(or_sim *s);                              /* SourceEuler.cu:475 */
void or_pressure(or_sim *s);                                 /* SourceEuler.cu:467 */
void or_temperature(or_sim *s);                              /* SourceEuler.cu:484 */
void or_substep1(or_sim *s, double dt);                      /* SourceEuler.cu:311 */
void or_substep2(or_sim *s, double dt);                      /* SourceEuler.cu:345 */
void or_host_kernel(or_sim *s, double dt);                   /* SourceEuler.cu:353 */
void or_actualise_velocities(or_sim *s);                     /* SourceEuler.cu:492-503 */
void or_substep3(or_sim *s, double dt);                      /* SourceEuler.cu:287-291 */
void or_transport(or_sim *s, double dt);                     /* TransportEuler.cu:30 */
void or_viscous_terms(or_sim *s, int use_intermediate);      /* Viscosity.cu:52 */
void or_selfgravity(or_sim *s, double dt, int update);       /* SgMain.cu:23 */
void or_make1d_cs(or_sim *s);                                /* Kernels.cu:626 */

/* 2D complex DFT used in place of cuFFT Z2Z (unnormalised; sign -1 =
 * forward, +1 = inverse), data interleaved re/im, row-major n1 x n2. */
void or_fft2d(double *data, int n1, int n2, int sign);

#ifdef __cplusplus
}
#endif
#endif
