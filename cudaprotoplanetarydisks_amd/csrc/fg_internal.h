// Internal state of an fg_ctx: device buffers, tables and launch helpers.
// Shared by the HIP translation units of libfargo_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "fargo_hip.h"

#define FG_PI 3.14159265358979323844
#define FG_CVNR 1.41
#define FG_CFLSECURITY 0.5
// DeviceReduce (Kernels.cu:364-394) never uses more than 1024 blocks.
#define FG_RED_BLOCKS 1024
// most Hill-cut variants of one fg_disk_force_hill call (the reference uses
// dimfxy = 11, Main.cu:74)
#define FG_HILL_MAX 64

void fg_set_error(const char *fmt, ...);
// once per device, with that device current: locate the device's copy of
// the scratch array of fg_disk_force_hill_noalloc (fg_reduce.hip)
fg_status fg_noalloc_setup(int device);
// self-gravity part of a fused substep (fg_sg.hip)
fg_status fg_sg_fused(fg_ctx *c, double dt, double *kick_r, double *kick_t);
// alpha viscosity of the current sound speed into nu_out (fg_steps.hip)
fg_status fg_viscosity_into(fg_ctx *c, double *nu_out);

// Result block of the fused path (fg_ctx::h_fused, coherent pinned host
// memory the final kernels write straight into): [0..1] star force,
// FG_FUSED_CFL_SLOT the CFL count (as a double), FG_FUSED_CRASH_SLOT an int
// that k_fin and the CFL ring pass set to 1 when they meet a negative or
// non-finite density, and from FG_FUSED_PF_SLOT the disk force on each
// planet that feels the disk (two doubles per planet), launched behind a
// substep for the next one.  FG_FUSED_SEQ_SLOT: the final kernel writes the
// sequence number of its launch there last, after a system-scope fence,
// so the host can wait for the results by watching that slot.
#define FG_FUSED_PF_SLOT 8
#define FG_FUSED_SEQ_SLOT 5
#define FG_FUSED_SLOTS (FG_FUSED_PF_SLOT + 2 * FG_MAX_PLANETS)
#define FG_FUSED_CFL_SLOT 6
#define FG_FUSED_CRASH_SLOT 7
// disk force on planet k at (x, y) with smoothing rs from the current
// density into the result block (fg_fused.hip), read once the fused
// results are in (fg_fused_results)
fg_status fg_fused_planet_force_async(fg_ctx *c, int k, double x, double y, double rs);
void fg_fused_planet_force(const fg_ctx *c, int k, double out2[2]);

// Largest CFL sub-cycle count accepted per DT.  The count is
// ceil(DeltaT / dt) (SourceEuler.cu:228); a count outside 1..FG_CFL_MAX
// (a time step gone to 0, infinity or NaN) or a non-finite disk force
// means the state has blown up.
#define FG_CFL_MAX 10000000
// FG_OK, or FG_E_CRASH with a message; force2 may be NULL
static inline fg_status fg_check_substep_results(double cfl, const double *force2) {
  if (!(cfl >= 1.0 && cfl <= (double)FG_CFL_MAX)) {
    fg_set_error("crash: CFL sub-cycle count %g outside 1..%d (time step 0, infinite or NaN)", cfl, FG_CFL_MAX);
    return FG_E_CRASH;
  }
  if (force2 && !(__builtin_isfinite(force2[0]) && __builtin_isfinite(force2[1]))) {
    fg_set_error("crash: non-finite disk force (%g, %g)", force2[0], force2[1]);
    return FG_E_CRASH;
  }
  return FG_OK;
}

#define FG_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      fg_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return FG_E_HIP;                                                               \
    }                                                                                \
  } while (0)

#define FG_LAUNCH_CHECK()                                                                 \
  do {                                                                                    \
    hipError_t e_ = hipGetLastError();                                                    \
    if (e_ != hipSuccess) {                                                               \
      fg_set_error("%s:%d kernel launch: %s", __FILE__, __LINE__, hipGetErrorString(e_)); \
      return FG_E_HIP;                                                                    \
    }                                                                                     \
  } while (0)

// Device copies of the host tables (fg_tables).
struct fg_dev_tables {
  double *Radii, *Rinf, *Rsup, *Rmed, *Surf, *invRinf, *invSurf, *invRmed, *invdiffRsup, *invdiffRmed;
  double *SigmaMed, *SigmaInf, *EnergyMed, *CoolingTimeMed, *QplusMed;
  double *cs_iso, *supp_torque, *dxtheta, *invdxtheta, *dxtheta_cfl;
  int *ev_zone;
  double *ev_coef, *ev_period, *ev_vtheta_arg;
  double *visc_const, *visc_pow15, *visc_cav;
  int *visc_cav_on;
  double *cos_cell, *sin_cell, *cos_pot, *sin_pot;
  double *sg_Kr, *sg_Kt;
  double *sg_norm_r, *sg_norm_t, *sg_sqrt_ratio, *sg_ratio;
  double dtheta, invdphi, dt1d_coef, vrad0_slope, visc_value;
};

// mixed-radix FFT plan of one transform length (fg_sg.hip)
#define FG_FFT_MAXST 40
struct fg_fft_plan;
void fg_sg_release(struct fg_ctx *c);

struct fg_planets_arg {
  int nb;
  double x[FG_MAX_PLANETS], y[FG_MAX_PLANETS], m[FG_MAX_PLANETS], smooth[FG_MAX_PLANETS],
      invd3[FG_MAX_PLANETS];
  double ix, iy;  // indirect term of the star (IndirectTerm.x/.y)
};

struct fg_ctx {
  fg_params p;
  int nr, ns;
  long S;  // (nr+1)*ns
  int device;
  hipStream_t stream;
  fg_dev_tables t;

  // evolved state and per-substep fields, (nr+1)*ns doubles each
  double *f[FG_NFIELDS];
  // remaining work arrays of the per-routine path
  double *DRR, *DPP, *DRP, *Work, *QRStar, *DensStar, *dq, *TempShift, *Vradial, *Vazimutal;
  // ring arrays (nr+1)
  double *Vmoy, *newDT, *DT1D, *VMed, *nu, *cs_mean, *sgacc_mean;
  // fused path: the previous self-gravity evaluation's ring means (the
  // first damping pass of a substep uses them), and the viscosity of an
  // adiabatic alpha-viscosity deck refreshed at the close of a DT, which
  // takes over from nu after the next substep's first damping pass
  double *sgacc_prev, *nu_next;
  int nu_pending;
  int *Nshift, *NoSplit;
  double *LostByDisk;  // ns
  // reduction scratch
  double *red_partial;  // 2 * FG_RED_BLOCKS
  double *red_out;      // [0..1] force, [2] lost mass, [3] circumplanetary mass, [4..5] ring means,
                        // [6] fused path's CFL count
  int *cfl_out;         // [0] CFL count
  double *h_out;        // pinned host mirror of red_out (4 doubles) + CFL
  int *h_cfl;
  // self-gravity complex buffers: 2*nr x ns interleaved complex
  double *sg_Sr, *sg_St, *sg_Krc, *sg_Ktc, *sg_Gr, *sg_Gt, *sg_tmp;
  int sg_kernel_ready;
  fg_fft_plan *fft_row, *fft_col;  // FFT plans of the NSEC rows and 2*NRAD columns
  // FP32 bulk mirrors (precision mode)
  float *f32[8];
  // host copies of ring values the host-side parts of a routine need
  std::vector<double> h_Rmed, h_Rinf, h_SigmaMed, h_EnergyMed, h_vtheta_arg;
  double lr01, lr12;  // log(Rmed0/Rmed1), log(Rmed1/Rmed2), Kernels.cu:93-94
  // kernel timing of the fused path (fg_profile): events around each
  // launch of one substep, folded into per-kernel totals once complete
  int prof_on;
  hipEvent_t prof_ev[16];
  int prof_nev, prof_pending;
  double prof_ms[16];
  long prof_calls[16];
  // tile shapes of the FP32 fused kernels (rings per k_src tile, rings
  // per radial-sweep band), set by fg_set_tiles
  int tile_src_th, tile_trad_band;
  // FP32 bulk mode: Sigma stored as float pairs on the rings within this
  // many Hill radii of a planet (fg_set_fp32_pairs)
  double fp32_pair_hill;
  // fused-path scratch
  double *fu_mom[4];  // transported radial momenta per ring band
  double *fu_ring;    // per-ring partial results
  double *fu_fpart;   // FP32 mode: per-ring star-force partials (4 * NRAD)
  // FP32 mode: per-ring background of Sigma the passes' FP32 offsets are
  // taken from (ring means of the last state), double-buffered: fu_sbar_cur
  // is read by this substep, the other written by its k_fin for the next
  double *fu_sbar[2];
  int fu_sbar_cur;
  // fused path's force and CFL count: written by the final reduction kernel
  // straight into coherent pinned host memory (no copy); device alias
  double *h_fused, *d_fused;
  unsigned long long fu_seq;  // sequence number of the last final-kernel launch
  int fu_seq_last;            // nothing the results depend on is queued behind it
  // Hill-cut force variants (fg_disk_force_hill), allocated on first use
  double *hill_partial;  // 4 * FG_HILL_MAX * FG_RED_BLOCKS block sums
  double *hill_out;      // 4 * FG_HILL_MAX results
};

static inline int fg_blocks(long n, int threads) { return (int)((n + threads - 1) / threads); }
