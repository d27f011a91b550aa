// Self-gravity: compute_selfgravity (SgMain.cu:23-45) with the log-polar
// convolution of Kernels.cu:840-932 evaluated by a hand-written FFT.
//
// The 2D transform is rows (length NSEC) then columns (length 2*NRAD) of
// the zero-padded 2NRAD x NSEC complex grid, as a 2D C2C transform.  Each
// 1D transform is a mixed-radix decimation-in-time FFT run bottom-up in
// Stockham (autosort) order inside LDS, one line (rows) or a few adjacent
// lines (columns) per workgroup:
//
//   factors, top level first: 2 while even, then 3, then 5, then the
//   remaining length as one direct DFT;
//   level d has sub-length L_d = p_d * ... * p_{K-1}, S_d = N / L_d
//   sub-problems and m_d = L_d / p_d;
//   A_d[s + S_d*kk] = sum_r A_{d+1}[s + S_d*r + S_{d+1}*k] * w_{L_d}^(r*kk)
//   with k = kk mod m_d; A_K is the input, A_0 the spectrum.
//
// Every output is accumulated over r in increasing order from 0.0 with the
// twiddle cos/sin(sign*2*pi*((r*kk) mod L_d)/L_d) taken from a host table,
// i.e. the same products and sums, in the same order, as the recursive
// FFT of the host oracle.  The kernel spectra are constant and are
// transformed once (the reference redoes this every call).
#include <utility>

#include "fg_cell.h"
#include "fg_internal.h"
#include "fg_reduce.h"

extern "C" void fg_fft_twiddles(int L, int sign, double *cs);  // fg_hostmath.cpp

#define FFT_TPB 256
#define FFT_U 16  // outputs per thread: N * lines <= FFT_TPB * FFT_U
#define FFT_MAX_ELEMS (FFT_TPB * FFT_U)

struct fg_fft_plan {
  int n, nst;
  int p[FG_FFT_MAXST], L[FG_FFT_MAXST], S[FG_FFT_MAXST], off[FG_FFT_MAXST];
  int sh[FG_FFT_MAXST];  // log2 S[d] when every S[d] is a power of two
  int pow2s;             // every level above the deepest is radix 2 (k_fft_pow2 applies)
  double2 *tw[2];  // per-level twiddle tables; [0] forward (sign -1), [1] inverse (+1)
};

// Sigma of cell l after the substep's first damping pass (the density part
// of EvanescentBoundaryKernel, Kernels.cu:978-1004; the fused path applies
// that pass in its source kernel, after the self-gravity field it needs)
__device__ __forceinline__ double sigma_damped(const fg_dev_tables &t, int i, double dt, int damp, double s) {
  if (damp && t.ev_zone[i]) s = damp_toward(s, t.ev_coef[i] * dt / t.ev_period[i], t.SigmaMed[i]);
  return s;
}

// What a pass reads and writes besides its lines:
//   FFT_PLAIN  lines in, lines out (in place)
//   FFT_FILL   rows of the forward transform: row i < NRAD is formed from
//              Sigma as ComputeFFTKernel does (Kernels.cu:840-868; grid 0
//              the azimuthal, grid 1 the radial one); the zero rows of the
//              padding are never transformed
//   FFT_CONV   columns of the forward transform: rows >= zrows read as 0
//              (the padding), and the spectrum is multiplied by the
//              kernel's and stored negated into out[]
//              (ComputeConvolutionKernel, Kernels.cu:871-890)
//   FFT_HEAD   columns of the inverse transform: only rows < zrows are
//              stored (the rest of the padded grid is never read)
enum { FFT_PLAIN = 0, FFT_FILL = 1, FFT_CONV = 2, FFT_HEAD = 3 };

struct fft_args {
  int nst;
  int p[FG_FFT_MAXST], L[FG_FFT_MAXST], S[FG_FFT_MAXST], off[FG_FFT_MAXST], sh[FG_FFT_MAXST];
  const double2 *tw;
  double2 *data[2];     // up to two grids transformed by one launch (blockIdx.y)
  int n;                // transform length
  int nlines;           // number of lines
  long elem_stride;     // distance between consecutive elements of a line
  long line_step;       // distance between consecutive lines
  int lines_per_block;  // lines staged together in LDS
  // FFT_FILL: Sigma and the substep's first damping pass
  const double *dens;
  fg_dev_tables t;
  double dt;
  int damp;
  int zrows;               // FFT_CONV / FFT_HEAD: rows of the unpadded grid
  const double2 *kern[2];  // FFT_CONV: kernel spectra of grids 0 and 1
  double2 *out[2];         // FFT_CONV: products
};

// element o of line `line` of grid g (blockIdx.y), as the pass reads it
template <int MODE>
__device__ __forceinline__ double2 fft_load(const fft_args &a, int g, int line, int o) {
  if (line >= a.nlines) return make_double2(0.0, 0.0);
  if constexpr (MODE == FFT_FILL) {
    const fg_dev_tables &t = a.t;
    const double var = sigma_damped(t, line, a.dt, a.damp, a.dens[(long)line * a.n + o]) * t.sg_sqrt_ratio[line];
    return make_double2(g == 0 ? var * t.Rmed[line] / t.Rmed[0] : var, 0.0);
  } else {
    if (MODE == FFT_CONV && o >= a.zrows) return make_double2(0.0, 0.0);
    return a.data[g][line * a.line_step + o * a.elem_stride];
  }
}

template <int MODE>
__device__ __forceinline__ void fft_store(const fft_args &a, int g, int line, int o, double2 v) {
  if (line >= a.nlines) return;
  const long idx = line * a.line_step + o * a.elem_stride;
  if constexpr (MODE == FFT_CONV) {
    const double2 k = a.kern[g][idx];
    a.out[g][idx] = make_double2(-1.0 * (k.x * v.x - k.y * v.y), -1.0 * (k.x * v.y + k.y * v.x));
  } else {
    if (MODE == FFT_HEAD && o >= a.zrows) return;
    a.data[g][idx] = v;
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs.  A column pass
// reads each grid row as CB-column pieces of one 128-byte line shared with
// the neighbouring column groups, so consecutive column groups go to the
// same XCD (and L2): workgroup b takes group (b % 8) * per + b / 8 of a
// grid of 8 * per workgroups.  Rows are contiguous and keep b.
__device__ __forceinline__ int fft_block_line(const fft_args &a) {
  const int b = blockIdx.x;
  if (a.elem_stride == 1) return b;
  const int per = gridDim.x >> 3;
  return (b & 7) * per + (b >> 3);
}

template <int MODE>
__global__ __launch_bounds__(FFT_TPB) void k_fft_lines(fft_args a) {
  extern __shared__ double2 buf[];
  const int CB = a.lines_per_block, N = a.n, tid = threadIdx.x;
  const int line0 = fft_block_line(a) * CB, g = blockIdx.y;
  if (line0 >= a.nlines) return;  // (whole workgroup)
  const int E = N * CB;
  // stage: element o of line cb at buf[o*CB + cb]
  for (int e = tid; e < E; e += FFT_TPB) {
    const int cb = e % CB, o = e / CB;
    buf[e] = fft_load<MODE>(a, g, line0 + cb, o);
  }
  __syncthreads();
  for (int d = a.nst - 1; d >= 0; d--) {
    const int p = a.p[d], L = a.L[d], S = a.S[d], m = L / p;
    const double2 *w = a.tw + a.off[d];
    double2 out[FFT_U];
#pragma unroll
    for (int u = 0; u < FFT_U; u++) {
      const int e = tid + u * FFT_TPB;
      if (e < E) {
        const int cb = e % CB, o = e / CB;
        const int s = o % S, kk = o / S, k = kk % m;
        const int base = s + S * p * k;
        double re = 0.0, im = 0.0;
        int idx = 0;  // (r*kk) mod L
        for (int r = 0; r < p; r++) {
          const double2 x = buf[(base + S * r) * CB + cb];
          const double2 t = w[idx];
          re += x.x * t.x - x.y * t.y;
          im += x.x * t.y + x.y * t.x;
          idx += kk;
          if (idx >= L) idx -= L;
        }
        out[u] = make_double2(re, im);
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < FFT_U; u++) {
      const int e = tid + u * FFT_TPB;
      if (e < E) buf[e] = out[u];
    }
    __syncthreads();
  }
  for (int e = tid; e < E; e += FFT_TPB) {
    const int cb = e % CB, o = e / CB;
    fft_store<MODE>(a, g, line0 + cb, o, buf[e]);
  }
}

// k_fft_pow2 runs TPB threads with 4 outputs each (up to TPB * FFT2_U
// elements per workgroup): four complex outputs and their operands fit in
// few registers.  Rows of up to 2048 elements take 512 threads (up to four
// workgroups per CU, one's loads and stores overlapping another's levels),
// everything else 1024.
#define FFT2_U 4

// one radix-2 output, r = 0 with w0 then r = 1 with t1, accumulated from
// 0.0 as the generic p-term loop does it (same products, same order)
__device__ __forceinline__ double2 bfly2(double2 x0, double2 x1, double2 w0, double2 t1) {
  double re = 0.0, im = 0.0;
  re += x0.x * w0.x - x0.y * w0.y;
  im += x0.x * w0.y + x0.y * w0.x;
  re += x1.x * t1.x - x1.y * t1.y;
  im += x1.x * t1.y + x1.y * t1.x;
  return make_double2(re, im);
}

// The same transform for lengths N = 2^a q with at most one factor other
// than 2 (q = 1, 3, 5 or a prime done as one direct DFT): every level above
// the deepest is radix 2, so each S_d is a power of two and an output's
// indices are shifts and masks.  CB lines are staged together (element o
// of line cb at buf[o*CB + cb]); global loads and stores run along the
// line when the line is contiguous (rows) and across the CB lines
// otherwise (columns).
//
// The radix-2 levels run two at a time in registers.  With the upper level
// of a pair at sub-length L and S sub-problems (the lower one at L/2 and
// 2S), the four values A[s + j*S + 4S*k2], j = 0..3, of sub-problem s and
// k2 < L/4 are all that the lower level needs for its outputs k2 and
// k2 + L/4 of sub-problems s and s + S, and those four are all that the
// upper level needs for its outputs k2 + i*L/4, i = 0..3, of sub-problem s.
// One thread evaluates both levels of such a group: every output is the
// same radix-2 expression of the same operands as when the levels run one
// after the other through LDS, so both give the same bits (as does
// k_fft_lines), with half the LDS round trips and barriers.
template <int CB, int MODE, int TPB>
__global__ __launch_bounds__(TPB) void k_fft_pow2(fft_args a) {
  extern __shared__ double2 buf[];
  constexpr int LCB = CB == 1 ? 0 : CB == 2 ? 1 : CB == 4 ? 2 : 3;
  const int N = a.n, tid = threadIdx.x;
  const int line0 = fft_block_line(a) * CB, g = blockIdx.y;
  if (line0 >= a.nlines) return;  // (whole workgroup)
  const int E = N * CB;
  if (a.elem_stride == 1) {
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
      for (int o = tid; o < N; o += TPB) buf[o * CB + cb] = fft_load<MODE>(a, g, line0 + cb, o);
  } else {
    for (int e = tid; e < E; e += TPB) buf[e] = fft_load<MODE>(a, g, line0 + (e & (CB - 1)), e >> LCB);
  }
  __syncthreads();
  double2 out[FFT2_U];
  {
    // deepest level: L = p (any factor), S = N / p, m = 1, k = 0
    const int d = a.nst - 1;
    const int p = a.p[d], L = a.L[d], sh = a.sh[d], S = 1 << sh;
    const double2 *w = a.tw + a.off[d];
#pragma unroll
    for (int u = 0; u < FFT2_U; u++) {
      const int e = tid + u * TPB;
      out[u] = make_double2(0.0, 0.0);
      if (e < E) {
        const int cb = e & (CB - 1), o = e >> LCB;
        const int s = o & (S - 1), kk = o >> sh;
        double re = 0.0, im = 0.0;
        int idx = 0;  // (r*kk) mod L
        for (int r = 0; r < p; r++) {
          const double2 x = buf[(s + S * r) * CB + cb];
          const double2 t = w[idx];
          re += x.x * t.x - x.y * t.y;
          im += x.x * t.y + x.y * t.x;
          idx += kk;
          if (idx >= L) idx -= L;
        }
        out[u] = make_double2(re, im);
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < FFT2_U; u++) {
      const int e = tid + u * TPB;
      if (e < E) buf[e] = out[u];
    }
    __syncthreads();
  }
  int d = a.nst - 2;
  if (d >= 0 && (d & 1) == 0) {
    // an odd number of radix-2 levels: the lowest one alone (r = 0 with
    // w[0], r = 1 with w[kk])
    const int L = a.L[d], sh = a.sh[d], m = L >> 1;
    const int S = 1 << sh;
    const double2 *w = a.tw + a.off[d];
    const double2 w0 = w[0];
#pragma unroll
    for (int u = 0; u < FFT2_U; u++) {
      const int e = tid + u * TPB;
      if (e < E) {
        const int cb = e & (CB - 1), o = e >> LCB;
        const int s = o & (S - 1), kk = o >> sh;
        const int k = kk >= m ? kk - m : kk;
        const int base = s + 2 * S * k;
        out[u] = bfly2(buf[base * CB + cb], buf[(base + S) * CB + cb], w0, w[kk]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < FFT2_U; u++) {
      const int e = tid + u * TPB;
      if (e < E) buf[e] = out[u];
    }
    __syncthreads();
    d--;
  }
  // pairs of levels: lower d (L/2, 2S), upper d - 1 (L, S); one group of
  // four per thread (E / 4 <= TPB)
  for (; d >= 1; d -= 2) {
    const int L = a.L[d - 1], sh = a.sh[d - 1], q = L >> 2;  // q = L/4
    const int S = 1 << sh;
    const double2 *wl = a.tw + a.off[d], *wu = a.tw + a.off[d - 1];
    const double2 wl0 = wl[0], wu0 = wu[0];
    const bool live = tid < (E >> 2);
    const int cb = tid & (CB - 1), rest = tid >> LCB;  // group: line cb, sub-problem s, k2
    const int s = rest & (S - 1), k2 = rest >> sh;
    double2 z0 = make_double2(0.0, 0.0), z1 = z0, z2 = z0, z3 = z0;
    if (live) {
      const int b = s + 4 * S * k2;
      const double2 x0 = buf[b * CB + cb], x1 = buf[(b + S) * CB + cb];
      const double2 x2 = buf[(b + 2 * S) * CB + cb], x3 = buf[(b + 3 * S) * CB + cb];
      // lower level: sub-problem s (x0, x2) and s + S (x1, x3) at k2, k2 + L/4
      const double2 ta = wl[k2], tb = wl[k2 + q];
      const double2 y0 = bfly2(x0, x2, wl0, ta), y1 = bfly2(x0, x2, wl0, tb);
      const double2 y2 = bfly2(x1, x3, wl0, ta), y3 = bfly2(x1, x3, wl0, tb);
      // upper level: outputs k2 + i L/4 of sub-problem s
      z0 = bfly2(y0, y2, wu0, wu[k2]);
      z1 = bfly2(y1, y3, wu0, wu[k2 + q]);
      z2 = bfly2(y0, y2, wu0, wu[k2 + 2 * q]);
      z3 = bfly2(y1, y3, wu0, wu[k2 + 3 * q]);
    }
    __syncthreads();
    if (live) {
      const int o = s + S * k2;
      buf[o * CB + cb] = z0;
      buf[(o + S * q) * CB + cb] = z1;
      buf[(o + 2 * S * q) * CB + cb] = z2;
      buf[(o + 3 * S * q) * CB + cb] = z3;
    }
    __syncthreads();
  }
  if (a.elem_stride == 1) {
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
      for (int o = tid; o < N; o += TPB) fft_store<MODE>(a, g, line0 + cb, o, buf[o * CB + cb]);
  } else {
    for (int e = tid; e < E; e += TPB) fft_store<MODE>(a, g, line0 + (e & (CB - 1)), e >> LCB, buf[e]);
  }
}

__global__ void k_sg_kernel_fill(double2 *Kr, double2 *Kt, fg_dev_tables t, long n) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  Kr[l] = make_double2(t.sg_Kr[l], 0.0);
  Kt[l] = make_double2(t.sg_Kt[l], 0.0);
}

// ComputeSgAccKernel (Kernels.cu:893-912) and the ring means of g_r
// (Make1Dprofile, Kernels.cu:610-623: the sectors summed in j order) for
// one ring per workgroup.  damp: the density after the first damping pass
// of a substep of length dt.  kick_r / kick_t (the fused path; NULL on the
// per-routine path, which adds them with k_sg_kick): the increments
// Update_sgvelocityKernel (Kernels.cu:915-932) adds to VradInt /
// VthetaInt, 0 for VradInt on ring 0, which it leaves alone; the
// neighbours' accelerations are recomputed with the same expressions.
__global__ __launch_bounds__(256) void k_sg_acc_ring(double *accr, double *acct, double *kick_r, double *kick_t,
                                                     double *mean, const double *dens, const double2 *Gr,
                                                     const double2 *Gt, fg_dev_tables t, double rstep,
                                                     double tstep, double eps, double dt, int damp, int ns) {
  extern __shared__ double ring_acc[];
  const int i = blockIdx.x;
  const long base = (long)i * ns;
  auto acc_r = [&](long m, int ii) {
    double a = Gr[m].x * t.sg_norm_r[ii];
    a += 1.0 * sigma_damped(t, ii, dt, damp, dens[m]) * rstep * tstep / eps;
    return a;
  };
  for (int j = threadIdx.x; j < ns; j += 256) {
    const double ar = acc_r(base + j, i);
    accr[base + j] = ar;
    ring_acc[j] = ar;
  }
  __syncthreads();
  // lane 0 sums the ring while waves 1-3 form g_theta and the kicks
  if (threadIdx.x == 0) {
    mean[i] = seq_sum_lds(ring_acc, ns) / (double)ns;
  } else if (threadIdx.x >= 64) {
    for (int j = threadIdx.x - 64; j < ns; j += 192) {
      const long l = base + j;
      const double at = Gt[l].x * t.sg_norm_t[i];
      acct[l] = at;
      if (kick_r) {
        const double atm = Gt[base + (j == 0 ? ns - 1 : j - 1)].x * t.sg_norm_t[i];
        kick_r[l] = i > 0 ? dt * ((t.Rinf[i] - t.Rmed[i - 1]) * ring_acc[j] +
                                  (t.Rmed[i] - t.Rinf[i]) * acc_r(l - ns, i - 1)) *
                                t.invdiffRmed[i]
                          : 0.0;
        kick_t[l] = 0.5 * dt * (at + atm);
      }
    }
  }
}

// Update_sgvelocityKernel, Kernels.cu:915-932, on VradInt / VthetaInt;
// the azimuthal neighbour of j = 0 is j = NSEC-1 of the same ring
__global__ void k_sg_kick(double *vr, double *vt, const double *accr, const double *acct, fg_dev_tables t,
                          double dt, int nr, int ns) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= (long)nr * ns) return;
  const int i = (int)(l / ns), j = (int)(l - (long)i * ns);
  if (i > 0)
    vr[l] += dt * ((t.Rinf[i] - t.Rmed[i - 1]) * accr[l] + (t.Rmed[i] - t.Rinf[i]) * accr[l - ns]) *
             t.invdiffRmed[i];
  const long lm = (long)i * ns + (j == 0 ? ns - 1 : j - 1);
  vt[l] += 0.5 * dt * (acct[l] + acct[lm]);
}

// ------------------------------------------------------------ host ----
static fg_status plan_create(int n, fg_fft_plan **out) {
  fg_fft_plan *pl = new fg_fft_plan();
  pl->n = n;
  int rem = n, k = 0;
  while (rem > 1) {
    if (k >= FG_FFT_MAXST) {
      delete pl;
      fg_set_error("FFT length %d has too many factors", n);
      return FG_E_UNSUPPORTED;
    }
    int p = (rem % 2 == 0) ? 2 : (rem % 3 == 0) ? 3 : (rem % 5 == 0) ? 5 : rem;
    pl->p[k++] = p;
    rem /= p;
  }
  pl->nst = k;
  pl->pow2s = k >= 1;
  for (int d = 0; d + 1 < k; d++)
    if (pl->p[d] != 2) pl->pow2s = 0;
  int L = 1, total = 0;
  for (int d = k - 1; d >= 0; d--) {
    L *= pl->p[d];
    pl->L[d] = L;
    pl->S[d] = n / L;
    pl->sh[d] = 0;
    while ((1 << pl->sh[d]) < pl->S[d]) pl->sh[d]++;
    pl->off[d] = total;
    total += L;
  }
  for (int sg = 0; sg < 2; sg++) {
    const int sign = sg == 0 ? -1 : 1;
    std::vector<double2> tw((size_t)(total > 0 ? total : 1));
    for (int d = 0; d < k; d++) fg_fft_twiddles(pl->L[d], sign, (double *)&tw[pl->off[d]]);
    hipError_t e = hipMalloc((void **)&pl->tw[sg], tw.size() * sizeof(double2));
    if (e == hipSuccess) e = hipMemcpy(pl->tw[sg], tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      fg_set_error("FFT twiddles: %s", hipGetErrorString(e));
      if (pl->tw[0]) (void)hipFree(pl->tw[0]);
      if (pl->tw[1]) (void)hipFree(pl->tw[1]);
      delete pl;
      return FG_E_HIP;
    }
  }
  *out = pl;
  return FG_OK;
}

void fg_sg_release(fg_ctx *c) {
  for (fg_fft_plan **pp : {&c->fft_row, &c->fft_col}) {
    if (!*pp) continue;
    for (int k = 0; k < 2; k++)
      if ((*pp)->tw[k]) (void)hipFree((*pp)->tw[k]);
    delete *pp;
    *pp = nullptr;
  }
}

// one pass (rows: pass 0, columns: pass 1) of a 2D transform of one or two
// 2NRAD x NSEC grids; a mode other than FFT_PLAIN also forms the input or
// the output (fft_load / fft_store): FFT_FILL transforms the NRAD rows of
// the density, FFT_CONV and FFT_HEAD take zrows = NRAD
template <int MODE>
static void fft_launch(fg_ctx *c, const fft_args &a, const fg_fft_plan *pl, int ngrids) {
  int blocks = (a.nlines + a.lines_per_block - 1) / a.lines_per_block;
  if (a.elem_stride != 1) blocks = 8 * ((blocks + 7) / 8);  // column groups in XCD order (fft_block_line)
  const size_t shm = (size_t)a.n * a.lines_per_block * sizeof(double2);
  const dim3 grid(blocks, ngrids);
  const bool wide = a.n * a.lines_per_block > 512 * FFT2_U;  // more than 2048 elements: 1024 threads
  if (!pl->pow2s)
    hipLaunchKernelGGL(k_fft_lines<MODE>, grid, dim3(FFT_TPB), shm, c->stream, a);
  else if (wide && a.lines_per_block == 8)
    hipLaunchKernelGGL((k_fft_pow2<8, MODE, 1024>), grid, dim3(1024), shm, c->stream, a);
  else if (wide && a.lines_per_block == 4)
    hipLaunchKernelGGL((k_fft_pow2<4, MODE, 1024>), grid, dim3(1024), shm, c->stream, a);
  else if (wide && a.lines_per_block == 2)
    hipLaunchKernelGGL((k_fft_pow2<2, MODE, 1024>), grid, dim3(1024), shm, c->stream, a);
  else if (wide)
    hipLaunchKernelGGL((k_fft_pow2<1, MODE, 1024>), grid, dim3(1024), shm, c->stream, a);
  else if (a.lines_per_block == 8)
    hipLaunchKernelGGL((k_fft_pow2<8, MODE, 512>), grid, dim3(512), shm, c->stream, a);
  else if (a.lines_per_block == 4)
    hipLaunchKernelGGL((k_fft_pow2<4, MODE, 512>), grid, dim3(512), shm, c->stream, a);
  else if (a.lines_per_block == 2)
    hipLaunchKernelGGL((k_fft_pow2<2, MODE, 512>), grid, dim3(512), shm, c->stream, a);
  else
    hipLaunchKernelGGL((k_fft_pow2<1, MODE, 512>), grid, dim3(512), shm, c->stream, a);
}

static fg_status fft_pass(fg_ctx *c, int pass, int inverse, int mode, double *g0, double *g1, double dt = 0.0,
                          int damp = 0, double *out0 = nullptr, double *out1 = nullptr) {
  const int n1 = 2 * c->nr, n2 = c->ns;
  const int ngrids = g1 ? 2 : 1;
  const fg_fft_plan *pl = pass == 0 ? c->fft_row : c->fft_col;
  fft_args a;
  a.nst = pl->nst;
  for (int d = 0; d < pl->nst; d++) {
    a.p[d] = pl->p[d];
    a.L[d] = pl->L[d];
    a.S[d] = pl->S[d];
    a.off[d] = pl->off[d];
    a.sh[d] = pl->sh[d];
  }
  a.tw = pl->tw[inverse ? 1 : 0];
  a.data[0] = (double2 *)g0;
  a.data[1] = (double2 *)(g1 ? g1 : g0);
  a.n = pl->n;
  a.dens = c->f[FG_DENS];
  a.t = c->t;
  a.dt = dt;
  a.damp = damp;
  a.zrows = c->nr;
  a.kern[0] = (const double2 *)c->sg_Ktc;
  a.kern[1] = (const double2 *)c->sg_Krc;
  a.out[0] = (double2 *)out0;
  a.out[1] = (double2 *)(out1 ? out1 : out0);
  // lines staged per workgroup: k_fft_lines up to FFT_MAX_ELEMS elements;
  // k_fft_pow2 rows: up to 2 lines of at most 2048 elements together (512
  // threads), a longer line alone (1024 threads); columns: a power of two
  // up to 8 lines of at most FFT_MAX_ELEMS elements (1024 threads; four
  // columns give 64-byte pieces of each row, two only 32: measured 53 vs
  // 59 us for the forward column pass at 1024 x 1536)
  int cb = ((pl->pow2s && pass == 0) ? 512 * FFT2_U : FFT_MAX_ELEMS) / pl->n;
  cb = cb >= 8 ? 8 : cb >= 4 ? 4 : cb >= 2 ? 2 : 1;
  if (pass == 0) {  // rows: contiguous
    a.nlines = mode == FFT_FILL ? c->nr : n1;
    a.elem_stride = 1;
    a.line_step = n2;
    a.lines_per_block = pl->pow2s ? (cb > 2 ? 2 : cb) : 1;
  } else {  // columns: stride NSEC, adjacent columns staged together
    a.nlines = n2;
    a.elem_stride = n2;
    a.line_step = 1;
    a.lines_per_block = cb;
  }
  switch (mode) {
    case FFT_FILL: fft_launch<FFT_FILL>(c, a, pl, ngrids); break;
    case FFT_CONV: fft_launch<FFT_CONV>(c, a, pl, ngrids); break;
    case FFT_HEAD: fft_launch<FFT_HEAD>(c, a, pl, ngrids); break;
    default: fft_launch<FFT_PLAIN>(c, a, pl, ngrids);
  }
  FG_LAUNCH_CHECK();
  return FG_OK;
}

// one 2D transform (rows, then columns) of one or two 2NRAD x NSEC grids
static fg_status fft2d(fg_ctx *c, double *g0, double *g1, int inverse) {
  if (fg_status s = fft_pass(c, 0, inverse, FFT_PLAIN, g0, g1)) return s;
  return fft_pass(c, 1, inverse, FFT_PLAIN, g0, g1);
}

// The potential's spectra G = -K * FFT(S) and their inverse transforms of
// the density (after the first damping pass of a substep of length dt when
// damp): the steps of compute_selfgravity (SgMain.cu:23-45) with the fill
// in the forward row pass, the convolution in the forward column pass, and
// the padding rows neither transformed forward nor stored backward.  Grid
// pairs are (azimuthal, radial).
static fg_status sg_field(fg_ctx *c, double dt, int damp) {
  if (fg_status s = fft_pass(c, 0, 0, FFT_FILL, c->sg_St, c->sg_Sr, dt, damp)) return s;
  if (fg_status s = fft_pass(c, 1, 0, FFT_CONV, c->sg_St, c->sg_Sr, 0.0, 0, c->sg_Gt, c->sg_Gr)) return s;
  if (fg_status s = fft_pass(c, 0, 1, FFT_PLAIN, c->sg_Gt, c->sg_Gr)) return s;
  return fft_pass(c, 1, 1, FFT_HEAD, c->sg_Gt, c->sg_Gr);
}

static fg_status sg_prepare(fg_ctx *c) {
  if (c->sg_kernel_ready) return FG_OK;
  const int n1 = 2 * c->nr, n2 = c->ns;
  if (n1 > FFT_MAX_ELEMS || n2 > FFT_MAX_ELEMS) {
    fg_set_error("self-gravity FFT: 2*NRAD = %d and NSEC = %d must not exceed %d", n1, n2, FFT_MAX_ELEMS);
    return FG_E_UNSUPPORTED;
  }
  if (!c->fft_row)
    if (fg_status s = plan_create(n2, &c->fft_row)) return s;
  if (!c->fft_col)
    if (fg_status s = plan_create(n1, &c->fft_col)) return s;
  const long n = (long)n1 * n2;
  hipLaunchKernelGGL(k_sg_kernel_fill, dim3(fg_blocks(n, 256)), dim3(256), 0, c->stream, (double2 *)c->sg_Krc,
                     (double2 *)c->sg_Ktc, c->t, n);
  FG_LAUNCH_CHECK();
  if (fg_status s = fft2d(c, c->sg_Krc, c->sg_Ktc, 0)) return s;
  c->sg_kernel_ready = 1;
  return FG_OK;
}

extern "C" fg_status fg_selfgravity(fg_ctx *c, double dt, int update) {
  if (!c) return FG_E_INVALID;
  if (!c->p.self_gravity) {
    fg_set_error("fg_selfgravity: context created without self-gravity");
    return FG_E_INVALID;
  }
  if (fg_status s = sg_prepare(c)) return s;
  const int nr = c->nr, ns = c->ns;
  if (fg_status s = sg_field(c, 0.0, 0)) return s;
  hipLaunchKernelGGL(k_sg_acc_ring, dim3(nr), dim3(256), ns * sizeof(double), c->stream, c->f[FG_SG_ACCR],
                     c->f[FG_SG_ACCT], (double *)nullptr, (double *)nullptr, c->sgacc_mean, c->f[FG_DENS],
                     (const double2 *)c->sg_Gr, (const double2 *)c->sg_Gt, c->t, c->p.sg_rstep, c->p.sg_tstep,
                     c->p.sg_eps, 0.0, 0, ns);
  FG_LAUNCH_CHECK();
  if (update) {
    hipLaunchKernelGGL(k_sg_kick, dim3(fg_blocks((long)nr * ns, 256)), dim3(256), 0, c->stream, c->f[FG_VRADINT],
                       c->f[FG_VTHETAINT], c->f[FG_SG_ACCR], c->f[FG_SG_ACCT], c->t, dt, nr, ns);
    FG_LAUNCH_CHECK();
  }
  return FG_OK;
}

fg_status fg_selfgravity_kick(fg_ctx *c, double dt) { return fg_selfgravity(c, dt, 1); }

// The self-gravity part of a fused substep, run before its source kernel:
// the field of the density after the first damping pass, its ring means
// (the previous evaluation's move to sgacc_prev, which that damping pass
// still uses), and the kick the source kernel adds after Substep1.
fg_status fg_sg_fused(fg_ctx *c, double dt, double *kick_r, double *kick_t) {
  if (fg_status s = sg_prepare(c)) return s;
  const int nr = c->nr, ns = c->ns;
  const int damp = c->p.evanescent;
  std::swap(c->sgacc_mean, c->sgacc_prev);
  if (fg_status s = sg_field(c, dt, damp)) return s;
  hipLaunchKernelGGL(k_sg_acc_ring, dim3(nr), dim3(256), ns * sizeof(double), c->stream, c->f[FG_SG_ACCR],
                     c->f[FG_SG_ACCT], kick_r, kick_t, c->sgacc_mean, c->f[FG_DENS], (const double2 *)c->sg_Gr,
                     (const double2 *)c->sg_Gt, c->t, c->p.sg_rstep, c->p.sg_tstep, c->p.sg_eps, dt, damp, ns);
  FG_LAUNCH_CHECK();
  return FG_OK;
}
