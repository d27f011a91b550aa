// fargo_mi355x: command-line drop-in for the reference's bin/fargoGPU
// (Main.cu:89-363).  Same usage: [-flags] parameter-file; the GPU ordinal
// is read from standard input like Main.cu:92-93, unless FARGO_DEVICE is
// set in the environment.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <string>

#include "fargo_hip.h"

static void usage(const char *exe) {  // PrintUsage, Interpret.cu:216-239
  printf("Usage : %s [-abcdeimnptvz] [-(0-9)] [-s number] [-f scaling] parameters file\n", exe);
  printf("\n-c : Sloppy CFL condition (checked at each DT, not at each timestep)\n");
  printf("-n : Disable simulation. The program just reads parameters file\n");
  printf("-o : Overrides output directory of input file.\n");
  printf("-s N : Restart from output set N of the output directory\n");
  printf("-F : (this build) use the fused substep kernels\n");
  printf("-S : (this build) FP32 bulk precision mode\n");
  printf("-L : (this build) circumplanetary mass, bigplanetK.dat and tqwkK*.dat torque logs\n");
  exit(1);
}

int main(int argc, char **argv) {
  int device = 0;
  const char *env = getenv("FARGO_DEVICE");
  if (env)
    device = atoi(env);
  else {
    printf("enter gpu device: ");
    fflush(stdout);
    if (scanf("%d", &device) != 1) device = 0;
  }
  if (argc == 1) usage(argv[0]);
  std::string flags, outdir, parfile;
  int disable = 0, fused = 0, fp32 = 0, restart = -1, diagnostics = 0;
  for (int i = 1; i < argc; i++) {
    if (argv[i][0] == '-') {
      if (strspn(argv[i], "-secndovtpfamzib0123456789FSL") != strlen(argv[i])) usage(argv[0]);
      flags += argv[i] + 1;
      if (strchr(argv[i], 'n')) disable = 1;
      if (strchr(argv[i], 'F')) fused = 1;
      if (strchr(argv[i], 'S')) fp32 = 1;
      if (strchr(argv[i], 'L')) diagnostics = 1;
      if (strchr(argv[i], 's')) {  // Main.cu:152-160
        if (++i >= argc) usage(argv[0]);
        restart = atoi(argv[i]);
        if (restart < 0) usage(argv[0]);
      } else if (strchr(argv[i], 'o')) {
        if (++i >= argc) usage(argv[0]);
        outdir = argv[i];
      } else if (strchr(argv[i], 'f')) {
        if (++i >= argc) usage(argv[0]);
      }
    } else
      parfile = argv[i];
  }
  if (parfile.empty()) usage(argv[0]);
  fg_sim_options opt{};
  opt.device = device;
  opt.precision = fp32 ? FG_PREC_FP32 : FG_PREC_FP64;
  opt.fused = fused;
  opt.write_files = disable ? 0 : 1;
  opt.outdir = outdir.empty() ? nullptr : outdir.c_str();
  opt.cli_flags = flags.c_str();
  opt.diagnostics = diagnostics;
  fg_sim *sim = nullptr;
  fg_status st = fg_sim_create(parfile.c_str(), &opt, &sim);
  if (st) {
    fprintf(stderr, "fargo_mi355x: %s\n", fg_last_error());
    return st == FG_E_PARAM || st == FG_E_IO ? 1 : 2;
  }
  if (disable) {
    fg_sim_destroy(sim);
    return 0;
  }
  double ntot = 0;
  fg_sim_get(sim, "NTOT", &ntot);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  double ninterm = 1;
  fg_sim_get(sim, "NINTERM", &ninterm);
  int it0 = 0;
  if (restart >= 0) {
    st = fg_sim_restart(sim, restart);
    if (st) {
      fprintf(stderr, "fargo_mi355x: %s\n", fg_last_error());
      fg_sim_destroy(sim);
      return 1;
    }
    it0 = restart * (int)ninterm;
  }
  st = fg_sim_run(sim, it0, (int)ntot);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (st) {
    fprintf(stderr, "fargo_mi355x: %s\n", fg_last_error());
    fg_sim_destroy(sim);
    return 2;
  }
  printf("%g\n", (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec));  // Main.cu:337
  fg_sim_destroy(sim);
  return 0;
}
