// Single-context reductions: DeviceReduce-order sums and the disk force
// (the device functions live in fg_reduce.h), and the non-reflecting
// boundary, whose ring means are such sums.
#include <atomic>
#include <mutex>

#include "fg_reduce.h"

template <class Src, int CH>
__global__ __launch_bounds__(RED_TPB) void k_dreduce_partial(Src src, long n, int B, double *partial) {
  dreduce_block<Src, CH>(src, n, B, blockIdx.x, partial);
}

__global__ void k_dreduce_final(const double *partial, int B, double *out, int two, int add_inner) {
  dreduce_final(partial, B, out, two, add_inner);
}

template <class Src>
static fg_status dreduce(fg_ctx *c, Src src, long n, double *out_dev, int two, int add_inner) {
  int B = red_blocks(n);
  // 16 summand rows per round: 8, 24 or 48 measured slower (more rounds, or
  // fewer resident workgroups)
  hipLaunchKernelGGL((k_dreduce_partial<Src, RED_CH>), dim3(B), dim3(RED_TPB), 0, c->stream, src, n, B,
                     c->red_partial);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dreduce_final, dim3(1), dim3(256), 0, c->stream, c->red_partial, B, out_dev, two, add_inner);
  FG_LAUNCH_CHECK();
  return FG_OK;
}

fg_status fg_reduce_array_internal(fg_ctx *c, const double *in, int n, double *result_dev) {
  array_src src{in, nullptr};
  return dreduce(c, src, n, result_dev, 0, 0);
}

fg_status fg_disk_force_async(fg_ctx *c, const double *dens, double x, double y, double rs, double *out_dev) {
  force_src src{dens, c->t.Surf, c->t.Rmed, c->t.cos_cell, c->t.sin_cell, x, y, rs, c->ns};
  return dreduce(c, src, (long)c->nr * c->ns, out_dev, 1, 1);
}

extern "C" fg_status fg_disk_force(fg_ctx *c, double x, double y, double rsmoothing, double *out2) {
  if (!c || !out2) {
    fg_set_error("fg_disk_force: NULL argument");
    return FG_E_INVALID;
  }
  if (fg_status s = fg_disk_force_async(c, c->f[FG_DENS], x, y, rsmoothing, c->red_out)) return s;
  FG_HIP(hipMemcpyAsync(c->h_out, c->red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FG_HIP(hipStreamSynchronize(c->stream));
  out2[0] = c->h_out[0];
  out2[1] = c->h_out[1];
  return FG_OK;
}

// ---- ComputeForce with the Hill cut (Force.cu:102-135) ---------------
// All dimfxy variants in one launch (blockIdx.y = k), each replaying the
// four DeviceReduce calls of Force.cu:120-123; the final pass adds each of
// the 4 * dimfxy block-sum rows in order into the GlobalForce layout
// [k + o * dimfxy], o = fxi, fxo, fyi, fyo.
__global__ __launch_bounds__(RED_TPB) void k_hill_partial(hill_force_src src, long n, int B, double *partial) {
  const int k = blockIdx.y;
  dreduce_block4(src, k, n, B, blockIdx.x, partial + (long)4 * k * B);
}

__global__ void k_hill_final(const double *partial, int B, int dimfxy, double *out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;  // row k * 4 + o
  if (r >= 4 * dimfxy) return;
  const double *p = partial + (long)r * B;
  double s = 0.0;
  for (int b = 0; b < B; b++) s += p[b];
  out[(r & 3) * dimfxy + (r >> 2)] = s;
}

extern "C" fg_status fg_disk_force_hill(fg_ctx *c, double x, double y, double rsmoothing, int dimfxy, double a,
                                        double rh, double *globalforce) {
  if (!c || !globalforce) {
    fg_set_error("fg_disk_force_hill: NULL argument");
    return FG_E_INVALID;
  }
  if (dimfxy < 1 || dimfxy > FG_HILL_MAX) {
    fg_set_error("fg_disk_force_hill: dimfxy = %d outside 1..%d", dimfxy, FG_HILL_MAX);
    return FG_E_INVALID;
  }
  const long n = (long)c->nr * c->ns;
  const int B = red_blocks(n);
  if (!c->hill_partial) {
    FG_HIP(hipMalloc((void **)&c->hill_partial, sizeof(double) * 4 * FG_HILL_MAX * FG_RED_BLOCKS));
    FG_HIP(hipMalloc((void **)&c->hill_out, sizeof(double) * 4 * FG_HILL_MAX));
  }
  hill_force_src src{c->f[FG_DENS], c->t.Surf, c->t.Rmed, c->t.cos_cell, c->t.sin_cell, x, y, rsmoothing, a, rh,
                     c->ns, dimfxy};
  hipLaunchKernelGGL(k_hill_partial, dim3(B, dimfxy), dim3(RED_TPB), 0, c->stream, src, n, B, c->hill_partial);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_hill_final, dim3((4 * dimfxy + 63) / 64), dim3(64), 0, c->stream, c->hill_partial, B, dimfxy,
                     c->hill_out);
  FG_LAUNCH_CHECK();
  FG_HIP(hipMemcpyAsync(globalforce, c->hill_out, sizeof(double) * 4 * dimfxy, hipMemcpyDeviceToHost, c->stream));
  FG_HIP(hipStreamSynchronize(c->stream));
  return FG_OK;
}

// ---- fg_disk_force_hill without allocation ----------------------------
// The block partials of fg_disk_force_hill_noalloc live in a file-scope
// device array: 4 * FG_HILL_MAX rows of FG_RED_BLOCKS block sums (2 MiB),
// enough for every grid (DeviceReduce never uses more than FG_RED_BLOCKS
// blocks) and every dimfxy <= FG_HILL_MAX.  Each GPU holds its own copy
// until the process ends.  On the host the array's name is not a device
// address: fg_noalloc_setup (called by fg_create with the context's device
// current) asks hipGetSymbolAddress for this device's copy once and keeps
// the pointer; kernels only ever receive that pointer.  Nothing in it needs
// clearing: k_hill_partial writes every row and block k_hill_final reads.
// All calls on one GPU share the array, so calls that overlap on one GPU
// (two streams, two threads) race.
#define FG_NOALLOC_DOUBLES (4 * FG_HILL_MAX * FG_RED_BLOCKS)
#define FG_NOALLOC_DEVICES 64
static __device__ __align__(256) double scratch[FG_NOALLOC_DOUBLES];
static std::atomic<double *> g_scratch_dev[FG_NOALLOC_DEVICES];
static std::mutex g_scratch_mu;

fg_status fg_noalloc_setup(int device) {
  if (device < 0 || device >= FG_NOALLOC_DEVICES) {
    fg_set_error("device %d beyond the %d the allocation-free entry point supports", device, FG_NOALLOC_DEVICES);
    return FG_E_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device].load()) return FG_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    fg_set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    return FG_E_HIP;
  }
  g_scratch_dev[device].store((double *)p);
  return FG_OK;
}

extern "C" fg_status fg_disk_force_hill_noalloc(fg_ctx *c, double x, double y, double rsmoothing, int dimfxy,
                                                double a, double rh, double *globalforce_dev) {
  if (!c || !globalforce_dev) {
    fg_set_error("fg_disk_force_hill_noalloc: NULL argument");
    return FG_E_INVALID;
  }
  if (dimfxy < 1 || dimfxy > FG_HILL_MAX) {
    fg_set_error("fg_disk_force_hill_noalloc: dimfxy = %d outside 1..%d (the scratch array holds %d variants)",
                 dimfxy, FG_HILL_MAX, FG_HILL_MAX);
    return FG_E_INVALID;
  }
  double *scr = (c->device >= 0 && c->device < FG_NOALLOC_DEVICES) ? g_scratch_dev[c->device].load() : nullptr;
  if (!scr) {
    fg_set_error("fg_disk_force_hill_noalloc: fg_create could not locate the scratch array on device %d",
                 c->device);
    return FG_E_HIP;
  }
  const long n = (long)c->nr * c->ns;
  const int B = red_blocks(n);
  if ((long)4 * dimfxy * B > FG_NOALLOC_DOUBLES) {
    fg_set_error("fg_disk_force_hill_noalloc: %d x %d partial sums exceed the %d-double scratch array", 4 * dimfxy,
                 B, FG_NOALLOC_DOUBLES);
    return FG_E_INVALID;
  }
  hill_force_src src{c->f[FG_DENS], c->t.Surf, c->t.Rmed, c->t.cos_cell, c->t.sin_cell, x, y, rsmoothing, a, rh,
                     c->ns, dimfxy};
  hipLaunchKernelGGL(k_hill_partial, dim3(B, dimfxy), dim3(RED_TPB), 0, c->stream, src, n, B, scr);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_hill_final, dim3((4 * dimfxy + 63) / 64), dim3(64), 0, c->stream, scr, B, dimfxy,
                     globalforce_dev);
  FG_LAUNCH_CHECK();
  return FG_OK;
}

// ---- CircumPlanetaryMass (SourceEuler.cu:549-563) --------------------
fg_status fg_circumplanetary_mass_async(fg_ctx *c, double xpl, double ypl, double hill_radius, double *out_dev) {
  cpm_src src{c->f[FG_DENS], c->t.Surf, c->t.Rmed, c->t.cos_cell, c->t.sin_cell, xpl, ypl, hill_radius, c->ns};
  return dreduce(c, src, (long)c->nr * c->ns, out_dev, 0, 0);
}

extern "C" fg_status fg_circumplanetary_mass(fg_ctx *c, double xpl, double ypl, double hill_radius, double *out) {
  if (!c || !out) {
    fg_set_error("fg_circumplanetary_mass: NULL argument");
    return FG_E_INVALID;
  }
  if (fg_status s = fg_circumplanetary_mass_async(c, xpl, ypl, hill_radius, c->red_out + 3)) return s;
  FG_HIP(hipMemcpyAsync(c->h_out + 3, c->red_out + 3, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FG_HIP(hipStreamSynchronize(c->stream));
  *out = c->h_out[3];
  return FG_OK;
}

// ---- non-reflecting boundary (SideEuler.cu:45-100) -------------------
// Ring gathers and their DeviceReduce means, then the two shifted-copy
// kernels and the mean correction.
__global__ void k_nr_copy(double *dens, double *energy, int i_angle, int ns) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  // ring 0 <- ring 1 shifted by i_angle (wrapped modulo NSEC)
  long dst = ((j + i_angle) % ns + ns) % ns;
  dens[dst] = dens[ns + j];
  energy[dst] = energy[ns + j];
}

__global__ void k_nr_copy_outer(double *dens, double *energy, int i_angle2, int nr, int ns) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  const long ii = nr - 1;
  // ring NRAD-1 <- ring NRAD-2 shifted by -i_angle2 (wrapped modulo NSEC)
  long src = (ii - 1) * ns + ((j - i_angle2) % ns + ns) % ns;
  dens[ii * ns + j] = dens[src];
  energy[ii * ns + j] = energy[src];
}

// NonReflectingBoundaryKernel2, Kernels.cu:575-589, bounded to j < NSEC
__global__ void k_nr_vrad(double *vr, const double *dens, const double *cs, double sm1, double sm2, int nr, int ns) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  long i = 1;
  double Vrad_med = -cs[i * ns + j] * (dens[i * ns + j] - sm1) / sm1;
  vr[i * ns + j] = 2.0 * Vrad_med - vr[(i + 1) * ns + j];
  i = nr - 1;
  Vrad_med = cs[i * ns + j] * (dens[(i - 1) * ns + j] - sm2) / sm2;
  vr[i * ns + j] = 2. * Vrad_med - vr[(i - 1) * ns + j];
}

// MinusMeanKernel, Kernels.cu:592-607
__global__ void k_nr_minus_mean(double *dens, double *energy, double a0, double a1, double e0, double e1, int nr,
                                int ns) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ns) return;
  dens[j] += a0;
  energy[j] += e0;
  long l = (long)(nr - 1) * ns + j;
  dens[l] += a1;
  energy[l] += e1;
}

static fg_status ring_means(fg_ctx *c, const double *f0, const double *f1, double *m0, double *m1) {
  // each DeviceReduce(...)/NSEC of SideEuler.cu:84-99 over one ring
  if (fg_status s = fg_reduce_array_internal(c, f0, c->ns, c->red_out + 4)) return s;
  if (fg_status s = fg_reduce_array_internal(c, f1, c->ns, c->red_out + 5)) return s;
  FG_HIP(hipMemcpyAsync(c->h_out + 4, c->red_out + 4, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FG_HIP(hipStreamSynchronize(c->stream));
  *m0 = c->h_out[4] / c->ns;
  *m1 = c->h_out[5] / c->ns;
  return FG_OK;
}

fg_status fg_nonreflecting_internal(fg_ctx *c) {
  const int nr = c->nr, ns = c->ns;
  double *cs = c->f[FG_SOUNDSPEED], *dens = c->f[FG_DENS], *en = c->f[FG_ENERGY], *vr = c->f[FG_VRAD];
  double cs0_r, cs1_r, csnrm1_r, csnrm2_r;
  if (fg_status s = ring_means(c, cs, cs + ns, &cs0_r, &cs1_r)) return s;
  if (fg_status s = ring_means(c, cs + (long)(nr - 1) * ns, cs + (long)(nr - 2) * ns, &csnrm1_r, &csnrm2_r)) return s;
  const double PI = FG_PI;
  const std::vector<double> &Rinf = c->h_Rinf, &Rmed = c->h_Rmed;
  int i = 1;
  double dangle = (pow(Rinf[i], -1.5) - 1.0) / (.5 * (cs0_r + cs1_r));
  dangle *= (Rmed[i] - Rmed[i - 1]);
  int i_angle = (int)(dangle / 2.0 / PI * (double)ns + .5);
  i = nr - 1;
  double dangle2 = (pow(Rinf[i - 1], -1.5) - 1.0) / (.5 * (csnrm1_r + csnrm2_r));
  dangle2 *= (Rmed[i] - Rmed[i - 1]);
  int i_angle2 = (int)(dangle2 / 2.0 / PI * (double)ns + .5);
  dim3 g(fg_blocks(ns, 256)), b(256);
  hipLaunchKernelGGL(k_nr_copy, g, b, 0, c->stream, dens, en, i_angle, ns);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_nr_copy_outer, g, b, 0, c->stream, dens, en, i_angle2, nr, ns);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_nr_vrad, g, b, 0, c->stream, vr, dens, cs, c->h_SigmaMed[1], c->h_SigmaMed[nr - 2], nr, ns);
  FG_LAUNCH_CHECK();
  double md0, md1, me0, me1;
  if (fg_status s = ring_means(c, dens, dens + (long)(nr - 1) * ns, &md0, &md1)) return s;
  if (fg_status s = ring_means(c, en, en + (long)(nr - 1) * ns, &me0, &me1)) return s;
  hipLaunchKernelGGL(k_nr_minus_mean, g, b, 0, c->stream, dens, en, c->h_SigmaMed[0] - md0,
                     c->h_SigmaMed[nr - 1] - md1, c->h_EnergyMed[0] - me0, c->h_EnergyMed[nr - 1] - me1, nr, ns);
  FG_LAUNCH_CHECK();
  return FG_OK;
}
