// DeviceReduce-order sums (Kernels.cu:253-394) and the disk-gravity
// summand of ComputeAccel (SideEuler.cu:135, Force.cu:102) as device
// functions, shared by the single-context kernels (fg_reduce.hip) and the
// batched ensemble kernels (fg_fused.hip).
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
#pragma once

#include "fg_internal.h"

#define RED_TPB 256
#define RED_CH 16

// ring and sector of flattened cell e: a 32-bit division (one float
// reciprocal and corrections) below 2^31 cells, the emulated 64-bit one
// beyond; NARROW: the caller knows e < 2^31 (no branch)
template <bool NARROW = false>
__device__ __forceinline__ void cell_ij(long e, int ns, int &i, int &j) {
  if (NARROW || e < 0x80000000L) {
    const unsigned ue = (unsigned)e, q = ue / (unsigned)ns;
    i = (int)q;
    j = (int)(ue - q * (unsigned)ns);
  } else {
    i = (int)(e / ns);
    j = (int)(e - (long)i * ns);
  }
}

static inline int red_blocks(long n) {
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
  static constexpr bool kRow = true;
  template <bool NARROW = false>
  __device__ __forceinline__ void operator()(long e, double &fx, double &fy) const {
    int i, j;
    cell_ij<NARROW>(e, ns, i, j);
    summand(e, i, j, fx, fy);
  }
  // cell e0 + lane of a 64-cell row that lies in one ring (NSEC a multiple
  // of 64, e0 a multiple of 64 and the same across the wave): the ring's
  // index and values are the wave's, no per-lane division
  __device__ __forceinline__ void row_cell(long e0, int lane, double &fx, double &fy) const {
    int i, j0;
    cell_ij(e0, ns, i, j0);
    i = __builtin_amdgcn_readfirstlane(i);
    j0 = __builtin_amdgcn_readfirstlane(j0);
    summand(e0 + lane, i, j0 + lane, fx, fy);
  }
  __device__ __forceinline__ void summand(long e, int i, int j, double &fx, double &fy) const {
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

// ComputeForceKernel with the Hill cut, Kernels.cu:693-729 (G = 1): the
// summands of fxi, fxo, fyi, fyo for variant k of dimfxy.  A cell lands in
// the inner pair when Rmed[i] < a and in the outer pair otherwise; the
// other pair holds the 0 of the memset arrays (Force.cu:111-114).
struct hill_force_src {
  const double *dens, *Surf, *Rmed, *cosc, *sinc;
  double x, y, rs, a, rh;
  int ns, dimfxy;
  __device__ __forceinline__ void operator()(long e, int k, double v[4]) const {
    int i, j;
    cell_ij(e, ns, i, j);
    double cellmass = Surf[i] * dens[e];
    double dx = Rmed[i] * cosc[j] - x;
    double dy = Rmed[i] * sinc[j] - y;
    double d2 = dx * dx + dy * dy;
    double dist2 = d2 + rs * rs;
    double distance = sqrt(dist2);
    double InvDist3 = 1.0 / dist2 / distance;
    double hill_cut = 1.0;
    if (k != 0) {
      double rhk = rh * ((double)k / (double)(dimfxy - 1));
      hill_cut = 1. - exp(-d2 / (rhk * rhk));
    }
    double fx = cellmass * dx * InvDist3 * hill_cut;
    double fy = cellmass * dy * InvDist3 * hill_cut;
    const bool inner = Rmed[i] < a;
    v[0] = inner ? fx : 0.0;
    v[1] = inner ? 0.0 : fx;
    v[2] = inner ? fy : 0.0;
    v[3] = inner ? 0.0 : fy;
  }
};

// CircumPlanetaryMassKernel, Kernels.cu:236-251: the mass of cells closer
// than the Hill radius to (xpl, ypl)
struct cpm_src {
  const double *dens, *Surf, *Rmed, *cosc, *sinc;
  double xpl, ypl, hill;
  int ns;
  static constexpr bool kRow = false;
  template <bool NARROW = false>
  __device__ __forceinline__ void operator()(long e, double &m, double &zero) const {
    int i, j;
    cell_ij<NARROW>(e, ns, i, j);
    double dx = Rmed[i] * cosc[j] - xpl;
    double dy = Rmed[i] * sinc[j] - ypl;
    double dist = sqrt(dx * dx + dy * dy);
    m = dist < hill ? Surf[i] * dens[e] : 0.0;
    zero = 0.0;
  }
};

struct array_src {
  const double *a, *b;  // b may be NULL
  static constexpr bool kRow = false;
  template <bool NARROW = false>
  __device__ __forceinline__ void operator()(long e, double &fx, double &fy) const {
    fx = a[e];
    fy = b ? b[e] : 0.0;
  }
};

// block blk of B (a RED_TPB-thread workgroup); partial[blk] and
// partial[B + blk] receive the x and y block sums
// CH summand rows are staged per round (CH * 64 / RED_TPB per thread).  A
// round's summands are evaluated branch-free (indices past n evaluate
// cell n - 1 and are dropped), so all their loads issue together; wave 0
// then reads its LDS values 8 rows ahead of the dependent adds.
// ROW: every 64-cell row lies in one ring (Src::row_cell); a row is then
// wholly inside the grid or wholly past it, and a wave stages one row.
template <bool NARROW, bool ROW, class Src, int CH>
__device__ __forceinline__ void dreduce_stage(const Src &src, long n, long b64, long G, long k0, double (*vx)[64],
                                              double (*vy)[64]) {
#pragma unroll
  for (int c0 = 0; c0 < CH * 64 / RED_TPB; c0++) {
    const int c = threadIdx.x + c0 * RED_TPB;
    const int lane = c & 63;
    double fx, fy;
    if constexpr (ROW) {
      const int kk = __builtin_amdgcn_readfirstlane(c >> 6);
      const long e0 = b64 + (k0 + kk) * G;
      const bool ok = e0 < n;
      src.row_cell(ok ? e0 : n - 64, lane, fx, fy);
      vx[kk][lane] = ok ? fx : 0.0;
      vy[kk][lane] = ok ? fy : 0.0;
    } else {
      const int kk = c >> 6;
      const long e = b64 + (k0 + kk) * G + lane;
      const bool ok = e < n;
      src.template operator()<NARROW>(ok ? e : n - 1, fx, fy);
      vx[kk][lane] = ok ? fx : 0.0;
      vy[kk][lane] = ok ? fy : 0.0;
    }
  }
}

template <class Src, int CH = RED_CH>
__device__ __forceinline__ void dreduce_block(const Src &src, long n, int B, int blk, double *partial) {
  static_assert(CH % 8 == 0, "rows are added in groups of 8");
  __shared__ double vx[CH][64];
  __shared__ double vy[CH][64];
  const long G = 64L * B;
  const long b64 = 64L * blk;
  const bool narrow = n <= 0x80000000L;  // every cell index below 2^31
  bool row = false;
  if constexpr (Src::kRow) row = src.ns % 64 == 0;  // rows of 64 cells never cross a ring
  double acc = 0.0;
  for (long k0 = 0; b64 + k0 * G < n; k0 += CH) {
    if constexpr (Src::kRow) {
      if (row) {
        dreduce_stage<true, true, Src, CH>(src, n, b64, G, k0, vx, vy);
        goto staged;
      }
    }
    if (narrow)
      dreduce_stage<true, false, Src, CH>(src, n, b64, G, k0, vx, vy);
    else
      dreduce_stage<false, false, Src, CH>(src, n, b64, G, k0, vx, vy);
  staged:
    __syncthreads();
    if (threadIdx.x < 64) {
      const int t = threadIdx.x & 31;
      const bool isy = threadIdx.x >= 32;
      const double(*v)[64] = isy ? vy : vx;
      for (int g = 0; g < CH; g += 8) {
        double lo[8], hi[8];
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
          lo[kk] = v[g + kk][t];
          hi[kk] = v[g + kk][t + 32];
        }
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
          const long e = b64 + (k0 + g + kk) * G + t;
          if (e < n) {
            acc += lo[kk];
            if (e + 32 < n) acc += hi[kk];
          }
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 64) {
    for (int off = 16; off > 0; off >>= 1) acc += __shfl_down(acc, off, 32);
    if (threadIdx.x == 0) partial[blk] = acc;
    if (threadIdx.x == 32) partial[B + blk] = acc;
  }
}

// Four DeviceReduce-order sums at once, block blk of B for summand set k:
// the four outputs are staged in LDS as in dreduce_block, then waves 0 and
// 1 replay the reference block (lanes 0-31 of wave w add output 2w, lanes
// 32-63 output 2w + 1).  partial[o * B + blk] receives output o.
template <class Src>
__device__ __forceinline__ void dreduce_block4(const Src &src, int k, long n, int B, int blk, double *partial) {
  __shared__ double v[4][RED_CH][64];
  const long G = 64L * B;
  const long b64 = 64L * blk;
  const int o = threadIdx.x >> 5, t = threadIdx.x & 31;
  double acc = 0.0;
  for (long k0 = 0; b64 + k0 * G < n; k0 += RED_CH) {
    for (int c = threadIdx.x; c < RED_CH * 64; c += RED_TPB) {
      int kk = c >> 6, lane = c & 63;
      long e = b64 + (k0 + kk) * G + lane;
      double s[4] = {0.0, 0.0, 0.0, 0.0};
      if (e < n) src(e, k, s);
      for (int q = 0; q < 4; q++) v[q][kk][lane] = s[q];
    }
    __syncthreads();
    if (threadIdx.x < 128) {
      for (int kk = 0; kk < RED_CH; kk++) {
        long e = b64 + (k0 + kk) * G + t;
        if (e < n) {
          acc += v[o][kk][t];
          if (e + 32 < n) acc += v[o][kk][t + 32];
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 128) {
    for (int off = 16; off > 0; off >>= 1) acc += __shfl_down(acc, off, 32);
    if (t == 0) partial[o * B + blk] = acc;
  }
}

// Left-to-right sum m + ring[0] + ring[1] + ... + ring[n-1] in one lane,
// the order of the reference's one-thread ring loops (Kernels.cu:1276-1279).
// Values come from LDS in blocks of 16, the next block's loads issued
// before the current block's adds, so the LDS latency overlaps the
// dependent adds.
__device__ __forceinline__ double seq_sum_lds(const double *ring, int n, double m = 0.0) {
  int j = 0;
  if (n >= 32) {
    double a[16], b[16];
#pragma unroll
    for (int q = 0; q < 16; q++) a[q] = ring[q];
    for (; j + 32 <= n; j += 32) {
#pragma unroll
      for (int q = 0; q < 16; q++) b[q] = ring[j + 16 + q];
      __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the adds
#pragma unroll
      for (int q = 0; q < 16; q++) m += a[q];
      // the next a block, or (past the last full pair) a harmless re-read
      const int jn = (j + 48 <= n) ? j + 32 : j;
#pragma unroll
      for (int q = 0; q < 16; q++) a[q] = ring[jn + q];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < 16; q++) m += b[q];
    }
  }
  for (; j < n; j++) m += ring[j];
  return m;
}

// host loop of DeviceReduce (Kernels.cu:385-390).  With add_inner, the
// result is fx_inner + fx_outer of SideEuler.cu:144 where the inner sum
// is the DeviceReduce of an all-zero array (+0.0).
__device__ __forceinline__ void dreduce_final(const double *partial, int B, double *out, int two, int add_inner) {
  __shared__ double p[2 * FG_RED_BLOCKS];
  const int nv = two ? 2 : 1;
  // stage the block sums in LDS with every lane, then add them in order
  for (int k = threadIdx.x; k < nv * B; k += blockDim.x) p[k] = partial[(k / B) * B + (k % B)];
  __syncthreads();
  int lane = threadIdx.x;
  if (lane >= nv) return;
  double s = seq_sum_lds(p + lane * B, B);
  if (add_inner) s = 0.0 + s;
  out[lane] = s;
}


// One thread's part of ConditionCFLKernel2D3 (Kernels.cu:1337-1360): the
// minimum of the ring minima newdt[i] (i >= 1) and the 1D shear limits
// dt1d[i] (stored) over rings i = i0, i0 + stride, ...  The loads of four
// rings are issued before their dt1d stores (the compiler cannot move a load
// across a store to a pointer that might alias it), so a thread waits for
// memory once per four rings; a minimum does not depend on the order of its
// comparisons.
__device__ __forceinline__ double cfl_ring_min(double *dt1d, const double *vmoy, const double *newdt,
                                               const double *invRmed, double coef, int nr, int i0, int stride) {
  double best = 1e300;
  for (int ib = i0; ib < nr; ib += 4 * stride) {
    double nd[4], v0[4], v1[4], r0[4], r1[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = ib + u * stride;
      nd[u] = v0[u] = v1[u] = r0[u] = r1[u] = 0.0;
      if (i < nr) {
        nd[u] = newdt[i];
        v0[u] = vmoy[i];
        r0[u] = invRmed[i];
        if (i < nr - 1) {
          v1[u] = vmoy[i + 1];
          r1[u] = invRmed[i + 1];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = ib + u * stride;
      if (i >= nr) break;
      if (i >= 1 && nd[u] < best) best = nd[u];
      if (i < nr - 1) {
        const double d = coef / fabs(v0[u] * r0[u] - v1[u] * r1[u]);
        dt1d[i] = d;
        if (d < best) best = d;
      }
    }
  }
  return best;
}

// minimum over a 256-thread block into every thread's return value
// (shuffles within each wave, then the four wave minima)
__device__ __forceinline__ double block_min256(double best) {
  __shared__ double wmin[4];
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(best, off, 64);
    if (o < best) best = o;
  }
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = best;
  __syncthreads();
  double m = wmin[0];
  for (int w = 1; w < 4; w++)
    if (wmin[w] < m) m = wmin[w];
  return m;
}

// ConditionCFLKernel2D3 and the host side of ConditionCFL
// (SourceEuler.cu:520-546) in one 256-thread block: min of the ring minima
// and the 1D shear limits, then the sub-cycle count ceil(DeltaT / dt)
// cfl: an int, or a double slot of a result block copied to the host at once
template <class T>
__device__ __forceinline__ void cfl_count_store(T *cfl, double best, double DeltaT) {
  if (threadIdx.x == 0) {
    double newdt_ = best;
    if (DeltaT < newdt_) newdt_ = DeltaT;
    cfl[0] = (T)(int)(ceil(DeltaT / newdt_));
  }
}
template <class T>
__device__ __forceinline__ void cfl_final(T *cfl, double *dt1d, const double *vmoy, const double *newdt,
                                          const double *invRmed, double coef, double DeltaT, int nr) {
  const double best = block_min256(cfl_ring_min(dt1d, vmoy, newdt, invRmed, coef, nr, threadIdx.x, 256));
  cfl_count_store(cfl, best, DeltaT);
}
