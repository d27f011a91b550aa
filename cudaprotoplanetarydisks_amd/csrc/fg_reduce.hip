// Sums in DeviceReduce order (Kernels.cu:253-394) and the disk-gravity
// reduction of ComputeAccel (SideEuler.cu:135, Force.cu:102).
//
// DeviceReduce launches B = min(pow2(n)/32, 1024) blocks of 32 threads.
// Thread t of block b accumulates, for k = 0, 1, ..., the elements
// e = 64*b + k*64*B + t and e + 32 (while e < n), then the 32 partials of
// a block meet in a shuffle-down tree (offsets 16, 8, 4, 2, 1) and the
// host adds the B block results in order.  Here one 256-thread workgroup
// replays one such block: its four waves evaluate the summands of CH
// consecutive k into LDS, then wave 0 (lanes 0-31 for x, 32-63 for y)
// adds them in the original order, and a width-32 shuffle tree finishes
// the block.  The final sequential sum over blocks runs in one lane.
// The result is bit-identical to the reference's reduction order while
// the summands are computed with full occupancy.
#include "fg_internal.h"

#define RED_TPB 256
#define RED_CH 16

static int red_blocks(long n) {
  long p = 1;
  while (p < n) p <<= 1;
  long b = (p + 31) / 32;
  return (int)(b > FG_RED_BLOCKS ? FG_RED_BLOCKS : b);
}

// gravity of cell e on a point mass at (x, y), Kernels.cu:703-727 with
// G = 1 and hill_cut = 1 (a = rh = 0 in ComputeAccel)
struct force_src {
  const double *dens, *Surf, *Rmed, *cosc, *sinc;
  double x, y, rs;
  int ns;
  __device__ __forceinline__ void operator()(long e, double &fx, double &fy) const {
    int i = (int)(e / ns);
    int j = (int)(e - (long)i * ns);
    double cellmass = Surf[i] * dens[e];
    double dx = Rmed[i] * cosc[j] - x;
    double dy = Rmed[i] * sinc[j] - y;
    double d2 = dx * dx + dy * dy;
    double dist2 = d2 + rs * rs;
    double distance = sqrt(dist2);
    double InvDist3 = 1.0 / dist2 / distance;
    fx = cellmass * dx * InvDist3;
    fy = cellmass * dy * InvDist3;
  }
};

struct array_src {
  const double *a, *b;  // b may be NULL
  __device__ __forceinline__ void operator()(long e, double &fx, double &fy) const {
    fx = a[e];
    fy = b ? b[e] : 0.0;
  }
};

template <class Src>
__global__ __launch_bounds__(RED_TPB) void k_dreduce_partial(Src src, long n, int B, double *partial) {
  __shared__ double vx[RED_CH][64];
  __shared__ double vy[RED_CH][64];
  const long G = 64L * B;
  const long b64 = 64L * blockIdx.x;
  double acc = 0.0;
  for (long k0 = 0; b64 + k0 * G < n; k0 += RED_CH) {
    for (int c = threadIdx.x; c < RED_CH * 64; c += RED_TPB) {
      int kk = c >> 6, lane = c & 63;
      long e = b64 + (k0 + kk) * G + lane;
      double fx = 0.0, fy = 0.0;
      if (e < n) src(e, fx, fy);
      vx[kk][lane] = fx;
      vy[kk][lane] = fy;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      const int t = threadIdx.x & 31;
      const bool isy = threadIdx.x >= 32;
      for (int kk = 0; kk < RED_CH; kk++) {
        long e = b64 + (k0 + kk) * G + t;
        if (e < n) {
          acc += isy ? vy[kk][t] : vx[kk][t];
          if (e + 32 < n) acc += isy ? vy[kk][t + 32] : vx[kk][t + 32];
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 64) {
    for (int off = 16; off > 0; off >>= 1) acc += __shfl_down(acc, off, 32);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
    if (threadIdx.x == 32) partial[B + blockIdx.x] = acc;
  }
}

// host loop of DeviceReduce (Kernels.cu:385-390).  With add_inner, the
// result is fx_inner + fx_outer of SideEuler.cu:144 where the inner sum
// is the DeviceReduce of an all-zero array (+0.0).
__global__ void k_dreduce_final(const double *partial, int B, double *out, int two, int add_inner) {
  __shared__ double p[2 * FG_RED_BLOCKS];
  const int nv = two ? 2 : 1;
  // stage the block sums in LDS with every lane, then add them in order
  for (int k = threadIdx.x; k < nv * B; k += blockDim.x) p[k] = partial[(k / B) * B + (k % B)];
  __syncthreads();
  int lane = threadIdx.x;
  if (lane >= nv) return;
  double s = 0.0;
  for (int b = 0; b < B; b++) s += p[lane * B + b];
  if (add_inner) s = 0.0 + s;
  out[lane] = s;
}

template <class Src>
static fg_status dreduce(fg_ctx *c, Src src, long n, double *out_dev, int two, int add_inner) {
  int B = red_blocks(n);
  hipLaunchKernelGGL(k_dreduce_partial<Src>, dim3(B), dim3(RED_TPB), 0, c->stream, src, n, B, c->red_partial);
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

// ---- non-reflecting boundary (SideEuler.cu:45-100) -------------------
// Ring gathers and their DeviceReduce means, then the two shifted-copy
// kernels and the mean correction.
__global__ void k_nr_copy(double *dens, double *energy, int i_angle, int i_angle2, int nr, int ns) {
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
  hipLaunchKernelGGL(k_nr_copy, g, b, 0, c->stream, dens, en, i_angle, i_angle2, nr, ns);
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
