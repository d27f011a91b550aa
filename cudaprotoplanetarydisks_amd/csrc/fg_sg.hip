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
};

__global__ __launch_bounds__(FFT_TPB) void k_fft_lines(fft_args a) {
  extern __shared__ double2 buf[];
  const int CB = a.lines_per_block, N = a.n, tid = threadIdx.x;
  const int line0 = blockIdx.x * CB;
  double2 *data = a.data[blockIdx.y];
  const int E = N * CB;
  // stage: element o of line cb at buf[o*CB + cb]
  for (int e = tid; e < E; e += FFT_TPB) {
    const int cb = e % CB, o = e / CB, line = line0 + cb;
    buf[e] = (line < a.nlines) ? data[line * a.line_step + o * a.elem_stride] : make_double2(0.0, 0.0);
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
    const int cb = e % CB, o = e / CB, line = line0 + cb;
    if (line < a.nlines) data[line * a.line_step + o * a.elem_stride] = buf[e];
  }
}

// k_fft_pow2 runs 1024 threads with 4 outputs each (FFT2_TPB * FFT2_U =
// FFT_MAX_ELEMS): four complex outputs and their operands fit in few
// registers, so several workgroups share a CU.
#define FFT2_TPB 1024
#define FFT2_U 4

// The same transform for lengths N = 2^a q with at most one factor other
// than 2 (q = 1, 3, 5 or a prime done as one direct DFT): every level above
// the deepest is radix 2, so each S_d is a power of two and an output's
// indices are shifts and masks; k = kk mod m_d is kk or kk - m_d for the
// radix-2 levels and 0 on the deepest level (m = 1).  CB lines are staged
// together (element o of line cb at buf[o*CB + cb]); global loads and
// stores run along the line when the line is contiguous (rows) and across
// the CB lines otherwise (columns).  Outputs are accumulated exactly as in
// k_fft_lines (same twiddles, same order), so both give the same bits.
template <int CB>
__global__ __launch_bounds__(FFT2_TPB) void k_fft_pow2(fft_args a) {
  extern __shared__ double2 buf[];
  constexpr int LCB = CB == 1 ? 0 : CB == 2 ? 1 : CB == 4 ? 2 : 3;
  const int N = a.n, tid = threadIdx.x;
  const int line0 = blockIdx.x * CB;
  double2 *data = a.data[blockIdx.y];
  const int E = N * CB;
  if (a.elem_stride == 1) {
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      const int line = line0 + cb;
      const double2 *src = data + (long)line * a.line_step;
      for (int o = tid; o < N; o += FFT2_TPB)
        buf[o * CB + cb] = (line < a.nlines) ? src[o] : make_double2(0.0, 0.0);
    }
  } else {
    for (int e = tid; e < E; e += FFT2_TPB) {
      const int cb = e & (CB - 1), o = e >> LCB, line = line0 + cb;
      buf[e] = (line < a.nlines) ? data[line * a.line_step + o * a.elem_stride] : make_double2(0.0, 0.0);
    }
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
      const int e = tid + u * FFT2_TPB;
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
      const int e = tid + u * FFT2_TPB;
      if (e < E) buf[e] = out[u];
    }
    __syncthreads();
  }
  // the radix-2 levels above it: r = 0 with w[0], r = 1 with w[kk] (kk < L)
  for (int d = a.nst - 2; d >= 0; d--) {
    const int L = a.L[d], sh = a.sh[d], m = L >> 1;
    const int S = 1 << sh;
    const double2 *w = a.tw + a.off[d];
    const double2 w0 = w[0];
#pragma unroll
    for (int u = 0; u < FFT2_U; u++) {
      const int e = tid + u * FFT2_TPB;
      if (e < E) {
        const int cb = e & (CB - 1), o = e >> LCB;
        const int s = o & (S - 1), kk = o >> sh;
        const int k = kk >= m ? kk - m : kk;
        const int base = s + 2 * S * k;
        const double2 x0 = buf[base * CB + cb];
        const double2 x1 = buf[(base + S) * CB + cb];
        const double2 t1 = w[kk];
        double re = 0.0, im = 0.0;
        re += x0.x * w0.x - x0.y * w0.y;
        im += x0.x * w0.y + x0.y * w0.x;
        re += x1.x * t1.x - x1.y * t1.y;
        im += x1.x * t1.y + x1.y * t1.x;
        out[u] = make_double2(re, im);
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < FFT2_U; u++) {
      const int e = tid + u * FFT2_TPB;
      if (e < E) buf[e] = out[u];
    }
    __syncthreads();
  }
  if (a.elem_stride == 1) {
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      const int line = line0 + cb;
      if (line >= a.nlines) continue;
      double2 *dst = data + (long)line * a.line_step;
      for (int o = tid; o < N; o += FFT2_TPB) dst[o] = buf[o * CB + cb];
    }
  } else {
    for (int e = tid; e < E; e += FFT2_TPB) {
      const int cb = e & (CB - 1), o = e >> LCB, line = line0 + cb;
      if (line < a.nlines) data[line * a.line_step + o * a.elem_stride] = buf[e];
    }
  }
}

// Sigma of cell l after the substep's first damping pass (the density part
// of EvanescentBoundaryKernel, Kernels.cu:978-1004; the fused path applies
// that pass in its source kernel, after the self-gravity field it needs)
__device__ __forceinline__ double sigma_damped(const fg_dev_tables &t, int i, double dt, int damp, double s) {
  if (damp && t.ev_zone[i]) s = damp_toward(s, t.ev_coef[i] * dt / t.ev_period[i], t.SigmaMed[i]);
  return s;
}

// ComputeFFTKernel, Kernels.cu:840-868 (density half of the padded grids);
// damp: Sigma as after the first damping pass of a substep of length dt
__global__ void k_sg_fill(double2 *Sr, double2 *St, const double *dens, fg_dev_tables t, double dt, int damp,
                          int nr, int ns) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= 2L * nr * ns) return;
  const int i = (int)(l / ns);
  double sr = 0.0, st = 0.0;
  if (i < nr) {
    const double var = sigma_damped(t, i, dt, damp, dens[l]) * t.sg_sqrt_ratio[i];
    sr = var;
    st = var * t.Rmed[i] / t.Rmed[0];
  }
  Sr[l] = make_double2(sr, 0.0);
  St[l] = make_double2(st, 0.0);
}

__global__ void k_sg_kernel_fill(double2 *Kr, double2 *Kt, fg_dev_tables t, long n) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  Kr[l] = make_double2(t.sg_Kr[l], 0.0);
  Kt[l] = make_double2(t.sg_Kt[l], 0.0);
}

// ComputeConvolutionKernel, Kernels.cu:871-890 (G = 1)
__global__ void k_sg_convolve(double2 *Gr, double2 *Gt, const double2 *Kr, const double2 *Kt, const double2 *Sr,
                              const double2 *St, long n) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  const double2 kt = Kt[l], st = St[l], kr = Kr[l], sr = Sr[l];
  Gt[l] = make_double2(-1.0 * (kt.x * st.x - kt.y * st.y), -1.0 * (kt.x * st.y + kt.y * st.x));
  Gr[l] = make_double2(-1.0 * (kr.x * sr.x - kr.y * sr.y), -1.0 * (kr.x * sr.y + kr.y * sr.x));
}

// ComputeSgAccKernel, Kernels.cu:893-912
__global__ void k_sg_acc(double *accr, double *acct, const double *dens, const double2 *Gr, const double2 *Gt,
                         fg_dev_tables t, double rstep, double tstep, double eps, int nr, int ns) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= (long)nr * ns) return;
  const int i = (int)(l / ns);
  acct[l] = Gt[l].x * t.sg_norm_t[i];
  double a = Gr[l].x * t.sg_norm_r[i];
  a += 1.0 * dens[l] * rstep * tstep / eps;
  accr[l] = a;
}

// The fused path's ComputeSgAccKernel: the accelerations from the density
// after the first damping pass, and the kick that Update_sgvelocityKernel
// (Kernels.cu:915-932) adds to VradInt / VthetaInt, for the source kernel
// to add (0 on ring 0, which the kick leaves alone).  The neighbours'
// accelerations are recomputed from the same expressions.
__global__ void k_sg_acc_fused(double *accr, double *acct, double *kick_r, double *kick_t, const double *dens,
                               const double2 *Gr, const double2 *Gt, fg_dev_tables t, double rstep, double tstep,
                               double eps, double dt, int damp, int nr, int ns) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= (long)nr * ns) return;
  const int i = (int)(l / ns), j = (int)(l - (long)i * ns);
  auto acc_r = [&](long m, int ii) {
    double a = Gr[m].x * t.sg_norm_r[ii];
    a += 1.0 * sigma_damped(t, ii, dt, damp, dens[m]) * rstep * tstep / eps;
    return a;
  };
  const long lm = (long)i * ns + (j == 0 ? ns - 1 : j - 1);
  const double ar = acc_r(l, i), at = Gt[l].x * t.sg_norm_t[i], atm = Gt[lm].x * t.sg_norm_t[i];
  accr[l] = ar;
  acct[l] = at;
  kick_r[l] = i > 0 ? dt * ((t.Rinf[i] - t.Rmed[i - 1]) * ar + (t.Rmed[i] - t.Rinf[i]) * acc_r(l - ns, i - 1)) *
                          t.invdiffRmed[i]
                    : 0.0;
  kick_t[l] = 0.5 * dt * (at + atm);
}

// Make1Dprofile of SG_Accr: sequential sum over the sectors of each ring
// by lane 0 of one wave, streamed through LDS
__global__ __launch_bounds__(64) void k_sg_ring_mean(const double *accr, double *mean, int ns) {
  __shared__ double chunk[64];
  const int lane = threadIdx.x;
  const double *v = accr + (long)blockIdx.x * ns;
  double sum = 0.0;
  for (int c0 = 0; c0 < ns; c0 += 64) {
    chunk[lane] = (c0 + lane < ns) ? v[c0 + lane] : 0.0;
    __syncthreads();
    if (lane == 0) {
      const int n = min(64, ns - c0);
      for (int k = 0; k < n; k++) sum += chunk[k];
    }
    __syncthreads();
  }
  if (lane == 0) mean[blockIdx.x] = sum / (double)ns;
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

// one 2D transform (rows, then columns) of one or two 2NRAD x NSEC grids
static fg_status fft2d(fg_ctx *c, double *g0, double *g1, int inverse) {
  const int n1 = 2 * c->nr, n2 = c->ns;
  const int ngrids = g1 ? 2 : 1;
  for (int pass = 0; pass < 2; pass++) {
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
    // lines staged per workgroup: up to FFT_MAX_ELEMS elements, a power of
    // two up to 8 for k_fft_pow2 (rows 2 at a time, columns 4 at NSEC 1536)
    int cb = FFT_MAX_ELEMS / pl->n;
    cb = cb >= 8 ? 8 : cb >= 4 ? 4 : cb >= 2 ? 2 : 1;
    if (pass == 0) {  // rows: contiguous
      a.nlines = n1;
      a.elem_stride = 1;
      a.line_step = n2;
      a.lines_per_block = pl->pow2s ? (cb > 2 ? 2 : cb) : 1;
    } else {  // columns: stride NSEC, adjacent columns staged together
      a.nlines = n2;
      a.elem_stride = n2;
      a.line_step = 1;
      a.lines_per_block = cb;
    }
    const int blocks = (a.nlines + a.lines_per_block - 1) / a.lines_per_block;
    const size_t shm = (size_t)a.n * a.lines_per_block * sizeof(double2);
    if (!pl->pow2s)
      hipLaunchKernelGGL(k_fft_lines, dim3(blocks, ngrids), dim3(FFT_TPB), shm, c->stream, a);
    else if (a.lines_per_block == 8)
      hipLaunchKernelGGL(k_fft_pow2<8>, dim3(blocks, ngrids), dim3(FFT2_TPB), shm, c->stream, a);
    else if (a.lines_per_block == 4)
      hipLaunchKernelGGL(k_fft_pow2<4>, dim3(blocks, ngrids), dim3(FFT2_TPB), shm, c->stream, a);
    else if (a.lines_per_block == 2)
      hipLaunchKernelGGL(k_fft_pow2<2>, dim3(blocks, ngrids), dim3(FFT2_TPB), shm, c->stream, a);
    else
      hipLaunchKernelGGL(k_fft_pow2<1>, dim3(blocks, ngrids), dim3(FFT2_TPB), shm, c->stream, a);
    FG_LAUNCH_CHECK();
  }
  return FG_OK;
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
  const long n = 2L * nr * ns, z = (long)nr * ns;
  hipLaunchKernelGGL(k_sg_fill, dim3(fg_blocks(n, 256)), dim3(256), 0, c->stream, (double2 *)c->sg_Sr,
                     (double2 *)c->sg_St, c->f[FG_DENS], c->t, 0.0, 0, nr, ns);
  FG_LAUNCH_CHECK();
  if (fg_status s = fft2d(c, c->sg_St, c->sg_Sr, 0)) return s;
  hipLaunchKernelGGL(k_sg_convolve, dim3(fg_blocks(n, 256)), dim3(256), 0, c->stream, (double2 *)c->sg_Gr,
                     (double2 *)c->sg_Gt, (const double2 *)c->sg_Krc, (const double2 *)c->sg_Ktc,
                     (const double2 *)c->sg_Sr, (const double2 *)c->sg_St, n);
  FG_LAUNCH_CHECK();
  if (fg_status s = fft2d(c, c->sg_Gt, c->sg_Gr, 1)) return s;
  hipLaunchKernelGGL(k_sg_acc, dim3(fg_blocks(z, 256)), dim3(256), 0, c->stream, c->f[FG_SG_ACCR],
                     c->f[FG_SG_ACCT], c->f[FG_DENS], (const double2 *)c->sg_Gr, (const double2 *)c->sg_Gt, c->t,
                     c->p.sg_rstep, c->p.sg_tstep, c->p.sg_eps, nr, ns);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sg_ring_mean, dim3(nr), dim3(64), 0, c->stream, c->f[FG_SG_ACCR], c->sgacc_mean, ns);
  FG_LAUNCH_CHECK();
  if (update) {
    hipLaunchKernelGGL(k_sg_kick, dim3(fg_blocks(z, 256)), dim3(256), 0, c->stream, c->f[FG_VRADINT],
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
  const long n = 2L * nr * ns, z = (long)nr * ns;
  const int damp = c->p.evanescent;
  std::swap(c->sgacc_mean, c->sgacc_prev);
  hipLaunchKernelGGL(k_sg_fill, dim3(fg_blocks(n, 256)), dim3(256), 0, c->stream, (double2 *)c->sg_Sr,
                     (double2 *)c->sg_St, c->f[FG_DENS], c->t, dt, damp, nr, ns);
  FG_LAUNCH_CHECK();
  if (fg_status s = fft2d(c, c->sg_St, c->sg_Sr, 0)) return s;
  hipLaunchKernelGGL(k_sg_convolve, dim3(fg_blocks(n, 256)), dim3(256), 0, c->stream, (double2 *)c->sg_Gr,
                     (double2 *)c->sg_Gt, (const double2 *)c->sg_Krc, (const double2 *)c->sg_Ktc,
                     (const double2 *)c->sg_Sr, (const double2 *)c->sg_St, n);
  FG_LAUNCH_CHECK();
  if (fg_status s = fft2d(c, c->sg_Gt, c->sg_Gr, 1)) return s;
  hipLaunchKernelGGL(k_sg_acc_fused, dim3(fg_blocks(z, 256)), dim3(256), 0, c->stream, c->f[FG_SG_ACCR],
                     c->f[FG_SG_ACCT], kick_r, kick_t, c->f[FG_DENS], (const double2 *)c->sg_Gr,
                     (const double2 *)c->sg_Gt, c->t, c->p.sg_rstep, c->p.sg_tstep, c->p.sg_eps, dt, damp, nr, ns);
  FG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sg_ring_mean, dim3(nr), dim3(64), 0, c->stream, c->f[FG_SG_ACCR], c->sgacc_mean, ns);
  FG_LAUNCH_CHECK();
  return FG_OK;
}
