// MLP emulator: BF16 weights and activations, FP32 accumulation, on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16), with training (MSE loss,
// backpropagation, Adam) and inference.  It maps disk parameters to
// ring-averaged profiles (cudaprotoplanetarydisks_amd/emulator.py); no
// reference counterpart.
//
// Every dense layer is one GEMM C = A B (M x K times K x N) with a fused
// epilogue:
//   forward hidden   Z = H W^T + b,  H' = bf16(GELU(Z))      (Z kept, FP32)
//   forward output   Y = H W^T + b                          (FP32)
//   backward         dZ_below = (dZ W) * GELU'(Z_below)     (FP32 + BF16)
//   weight gradient  dW = dZ^T H                            (FP32)
// Operands are addressed by (row stride, column stride), so transposes are
// free; a 256-thread workgroup computes a 64 x 64 tile as four waves of
// 2 x 2 MFMA blocks, staging 64 x 32 slices of A and B^T in LDS.
// Master weights, gradients and Adam moments are FP32; the BF16 weight
// copy the GEMMs read is refreshed by the optimizer step.
#include <cmath>
#include <cstring>
#include <vector>

#include "fg_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

#define GT 64  // output tile (rows and columns)
#define GK 32  // k per MFMA step
#define GKP (GK + 8)  // LDS row length (16-byte aligned rows, fewer bank conflicts)
#define MLP_MAX_LAYERS 8

__device__ __forceinline__ bf16_t f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

// GELU(x) = x Phi(x) and its derivative Phi(x) + x phi(x)
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

enum { EPI_STORE = 0, EPI_GELU = 1, EPI_GELU_BACK = 2 };

struct gemm_args {
  const bf16_t *A;
  long sai, sak;  // A[i][k] = A[i*sai + k*sak]
  const bf16_t *B;
  long sbk, sbj;  // B[k][j] = B[k*sbk + j*sbj]
  int M, N, K;
  int mode;
  const float *bias;   // EPI_STORE / EPI_GELU (may be NULL)
  float *C;            // FP32 result (Y, Z or dZ), row stride ldc
  bf16_t *Cb;          // BF16 copy (GELU(Z) or dZ; optional for EPI_STORE)
  const float *Zprev;  // EPI_GELU_BACK: pre-activation below, row stride ldc
  int ldc;
};

__global__ __launch_bounds__(256) void k_gemm_bf16(gemm_args g) {
  __shared__ __attribute__((aligned(16))) bf16_t As[GT][GKP];  // As[i][k]
  __shared__ __attribute__((aligned(16))) bf16_t Bs[GT][GKP];  // Bs[j][k] (B transposed)
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  const int lr = lane & 15, kq = (lane >> 4) * 8;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += GK) {
    for (int e = tid; e < GT * GK; e += 256) {
      const int r = e / GK, k = e % GK, gi = i0 + r, gk = k0 + k;
      As[r][k] = (gi < g.M && gk < g.K) ? g.A[gi * g.sai + gk * g.sak] : (bf16_t)0;
    }
    for (int e = tid; e < GT * GK; e += 256) {
      const int c = e / GK, k = e % GK, gj = j0 + c, gk = k0 + k;
      Bs[c][k] = (gj < g.N && gk < g.K) ? g.B[gk * g.sbk + gj * g.sbj] : (bf16_t)0;
    }
    __syncthreads();
    bf16x8 a[2], b[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      a[t] = *(const bf16x8 *)&As[wi + 16 * t + lr][kq];
      b[t] = *(const bf16x8 *)&Bs[wj + 16 * t + lr][kq];
    }
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
      for (int tj = 0; tj < 2; tj++) acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
    __syncthreads();
  }
  // lane holds D[4*(lane/16) + v][lane%16] of each 16 x 16 block
#pragma unroll
  for (int ti = 0; ti < 2; ti++)
#pragma unroll
    for (int tj = 0; tj < 2; tj++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = i0 + wi + 16 * ti + 4 * (lane >> 4) + v;
        const int j = j0 + wj + 16 * tj + lr;
        if (i >= g.M || j >= g.N) continue;
        const long o = (long)i * g.ldc + j;
        float x = acc[ti][tj][v];
        if (g.mode == EPI_GELU_BACK) {
          x *= gelu_d(g.Zprev[o]);
          g.C[o] = x;
          g.Cb[o] = f2bf(x);
        } else {
          if (g.bias) x += g.bias[j];
          g.C[o] = x;
          if (g.mode == EPI_GELU)
            g.Cb[o] = f2bf(gelu(x));
          else if (g.Cb)
            g.Cb[o] = f2bf(x);
        }
      }
}

__global__ void k_f32_to_bf16(const float *src, bf16_t *dst, long n) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l < n) dst[l] = f2bf(src[l]);
}

// MSE over the batch: dY = 2 (Y - T) / n (FP32 and BF16), loss into
// *loss (a fixed-order tree over one workgroup)
__global__ __launch_bounds__(1024) void k_mse(const float *Y, const float *T, long n, float *dY, bf16_t *dYb,
                                              float *loss) {
  __shared__ float red[1024];
  const float scale = 2.0f / (float)n;
  float s = 0.f;
  for (long l = threadIdx.x; l < n; l += 1024) {
    const float d = Y[l] - T[l];
    s += d * d;
    const float g = scale * d;
    dY[l] = g;
    dYb[l] = f2bf(g);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] / (float)n;
}

// bias gradient: column sums of dZ (rows in order)
__global__ void k_colsum(const float *dZ, int rows, int cols, float *db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  float s = 0.f;
  for (int i = 0; i < rows; i++) s += dZ[(long)i * cols + j];
  db[j] = s;
}

// Adam (bias-corrected), then the BF16 copy of the weights
__global__ void k_adam(float *w, const float *gr, float *m, float *v, bf16_t *wb, long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  const float g = gr[l];
  const float mm = b1 * m[l] + (1.f - b1) * g;
  const float vv = b2 * v[l] + (1.f - b2) * g * g;
  m[l] = mm;
  v[l] = vv;
  const float x = w[l] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  w[l] = x;
  wb[l] = f2bf(x);
}

// ------------------------------------------------------------ host ----
struct fg_mlp {
  int L;                        // dense layers
  int width[MLP_MAX_LAYERS + 1];  // width[0] inputs ... width[L] outputs
  int max_batch;
  long nparam;
  long offW[MLP_MAX_LAYERS], offb[MLP_MAX_LAYERS];
  hipStream_t stream;
  float *p, *g, *m, *v;  // parameters, gradients, Adam moments (FP32)
  bf16_t *pb;            // BF16 parameters
  bf16_t *H[MLP_MAX_LAYERS];     // BF16 layer inputs (H[0] = x)
  float *Z[MLP_MAX_LAYERS];      // pre-activations (Z[L-1] = output)
  float *dZ[MLP_MAX_LAYERS];     // FP32 gradients of the pre-activations
  bf16_t *dZb[MLP_MAX_LAYERS];   // and their BF16 copies
  float *X, *T;                  // FP32 batch inputs and targets
  float *loss;                   // per-step losses of the last fg_mlp_train call
  int loss_cap;
  long step;                     // Adam steps taken
};

static fg_status mlp_gemm(fg_mlp *m, const gemm_args &g) {
  dim3 grid((g.N + GT - 1) / GT, (g.M + GT - 1) / GT);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, m->stream, g);
  FG_LAUNCH_CHECK();
  return FG_OK;
}

static uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

extern "C" fg_status fg_mlp_set_params(fg_mlp *m, const float *params);

extern "C" fg_status fg_mlp_destroy(fg_mlp *m) {
  if (!m) return FG_OK;
  (void)hipStreamSynchronize(m->stream);
  void *ptrs[] = {m->p, m->g, m->m, m->v, m->pb, m->X, m->T, m->loss};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  for (int l = 0; l < MLP_MAX_LAYERS; l++) {
    void *q[] = {m->H[l], m->Z[l], m->dZ[l], m->dZb[l]};
    for (void *x : q)
      if (x) (void)hipFree(x);
  }
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return FG_OK;
}

#define MLP_TRY(call)              \
  do {                             \
    hipError_t e_ = (call);        \
    if (e_ != hipSuccess) {        \
      fg_set_error("mlp: %s", hipGetErrorString(e_)); \
      fg_mlp_destroy(m);           \
      return FG_E_HIP;             \
    }                              \
  } while (0)

// widths[0..nlayers]: inputs, hidden widths, outputs.  Weights drawn
// uniform in +-sqrt(6 / (fan_in + fan_out)) from seed, biases zero.
extern "C" fg_status fg_mlp_create(const int *widths, int nlayers, int max_batch, unsigned long long seed,
                                   int device, fg_mlp **out) {
  if (!widths || !out || nlayers < 1 || nlayers > MLP_MAX_LAYERS || max_batch < 1) {
    fg_set_error("fg_mlp_create: bad argument");
    return FG_E_INVALID;
  }
  for (int l = 0; l <= nlayers; l++)
    if (widths[l] < 1 || widths[l] > 65536) {
      fg_set_error("fg_mlp_create: width %d", widths[l]);
      return FG_E_INVALID;
    }
  FG_HIP(hipSetDevice(device));
  fg_mlp *m = new fg_mlp();
  m->L = nlayers;
  for (int l = 0; l <= nlayers; l++) m->width[l] = widths[l];
  m->max_batch = max_batch;
  long off = 0;
  for (int l = 0; l < nlayers; l++) {
    m->offW[l] = off;
    off += (long)widths[l + 1] * widths[l];
    m->offb[l] = off;
    off += widths[l + 1];
  }
  m->nparam = off;
  MLP_TRY(hipStreamCreate(&m->stream));
  const size_t P = (size_t)off;
  MLP_TRY(hipMalloc((void **)&m->p, P * 4));
  MLP_TRY(hipMalloc((void **)&m->g, P * 4));
  MLP_TRY(hipMalloc((void **)&m->m, P * 4));
  MLP_TRY(hipMalloc((void **)&m->v, P * 4));
  MLP_TRY(hipMalloc((void **)&m->pb, P * 2));
  MLP_TRY(hipMemset(m->m, 0, P * 4));
  MLP_TRY(hipMemset(m->v, 0, P * 4));
  MLP_TRY(hipMemset(m->g, 0, P * 4));
  for (int l = 0; l < nlayers; l++) {
    const size_t in = (size_t)max_batch * widths[l], outn = (size_t)max_batch * widths[l + 1];
    MLP_TRY(hipMalloc((void **)&m->H[l], in * 2));
    MLP_TRY(hipMalloc((void **)&m->Z[l], outn * 4));
    MLP_TRY(hipMalloc((void **)&m->dZ[l], outn * 4));
    MLP_TRY(hipMalloc((void **)&m->dZb[l], outn * 2));
  }
  MLP_TRY(hipMalloc((void **)&m->X, (size_t)max_batch * widths[0] * 4));
  MLP_TRY(hipMalloc((void **)&m->T, (size_t)max_batch * widths[nlayers] * 4));
  std::vector<float> host(P, 0.0f);
  uint64_t state = seed;
  for (int l = 0; l < nlayers; l++) {
    const double lim = sqrt(6.0 / (double)(widths[l] + widths[l + 1]));
    for (long q = 0; q < (long)widths[l + 1] * widths[l]; q++) {
      state = mix64(state);
      const double u = (double)(state >> 11) * (1.0 / 9007199254740992.0);
      host[m->offW[l] + q] = (float)((2.0 * u - 1.0) * lim);
    }
  }
  fg_status st = fg_mlp_set_params(m, host.data());
  if (st) {
    fg_mlp_destroy(m);
    return st;
  }
  *out = m;
  return FG_OK;
}

extern "C" fg_status fg_mlp_set_params(fg_mlp *m, const float *params) {
  if (!m || !params) return FG_E_INVALID;
  FG_HIP(hipMemcpyAsync(m->p, params, m->nparam * 4, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(fg_blocks(m->nparam, 256)), dim3(256), 0, m->stream, m->p, m->pb,
                     m->nparam);
  FG_LAUNCH_CHECK();
  FG_HIP(hipStreamSynchronize(m->stream));
  return FG_OK;
}

extern "C" long fg_mlp_nparam(const fg_mlp *m) { return m ? m->nparam : 0; }

extern "C" fg_status fg_mlp_get_params(fg_mlp *m, float *params) {
  if (!m || !params) return FG_E_INVALID;
  FG_HIP(hipMemcpyAsync(params, m->p, m->nparam * 4, hipMemcpyDeviceToHost, m->stream));
  FG_HIP(hipStreamSynchronize(m->stream));
  return FG_OK;
}

// forward pass of the batch in m->X (rows = batch); output in Z[L-1]
static fg_status mlp_forward(fg_mlp *m, int batch) {
  const int L = m->L;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(fg_blocks((long)batch * m->width[0], 256)), dim3(256), 0, m->stream, m->X,
                     m->H[0], (long)batch * m->width[0]);
  FG_LAUNCH_CHECK();
  for (int l = 0; l < L; l++) {
    const int K = m->width[l], N = m->width[l + 1];
    gemm_args g{};
    g.A = m->H[l];
    g.sai = K;
    g.sak = 1;
    g.B = m->pb + m->offW[l];  // B[k][j] = W[j][k]
    g.sbk = 1;
    g.sbj = K;
    g.M = batch;
    g.N = N;
    g.K = K;
    g.bias = m->p + m->offb[l];
    g.C = m->Z[l];
    g.ldc = N;
    if (l < L - 1) {
      g.mode = EPI_GELU;
      g.Cb = m->H[l + 1];
    } else {
      g.mode = EPI_STORE;
      g.Cb = nullptr;
    }
    if (fg_status s = mlp_gemm(m, g)) return s;
  }
  return FG_OK;
}

// gradients of the MSE of Z[L-1] against m->T into m->g; loss to *loss_dev
static fg_status mlp_backward(fg_mlp *m, int batch, float *loss_dev) {
  const int L = m->L;
  const long nout = (long)batch * m->width[L];
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(1024), 0, m->stream, m->Z[L - 1], m->T, nout, m->dZ[L - 1],
                     m->dZb[L - 1], loss_dev);
  FG_LAUNCH_CHECK();
  for (int l = L - 1; l >= 0; l--) {
    const int K = m->width[l], N = m->width[l + 1];
    // dW = dZ^T H: M = N rows, N = K columns, contraction over the batch
    gemm_args gw{};
    gw.A = m->dZb[l];
    gw.sai = 1;
    gw.sak = N;
    gw.B = m->H[l];
    gw.sbk = K;
    gw.sbj = 1;
    gw.M = N;
    gw.N = K;
    gw.K = batch;
    gw.mode = EPI_STORE;
    gw.C = m->g + m->offW[l];
    gw.ldc = K;
    if (fg_status s = mlp_gemm(m, gw)) return s;
    hipLaunchKernelGGL(k_colsum, dim3(fg_blocks(N, 256)), dim3(256), 0, m->stream, m->dZ[l], batch, N,
                       m->g + m->offb[l]);
    FG_LAUNCH_CHECK();
    if (l == 0) break;
    // dZ below = (dZ W) * GELU'(Z below)
    gemm_args gb{};
    gb.A = m->dZb[l];
    gb.sai = N;
    gb.sak = 1;
    gb.B = m->pb + m->offW[l];  // B[k][j] = W[k][j]
    gb.sbk = K;
    gb.sbj = 1;
    gb.M = batch;
    gb.N = K;
    gb.K = N;
    gb.mode = EPI_GELU_BACK;
    gb.C = m->dZ[l - 1];
    gb.Cb = m->dZb[l - 1];
    gb.Zprev = m->Z[l - 1];
    gb.ldc = K;
    if (fg_status s = mlp_gemm(m, gb)) return s;
  }
  return FG_OK;
}

static fg_status check_batch(fg_mlp *m, int batch) {
  if (!m || batch < 1 || batch > m->max_batch) {
    fg_set_error("mlp: batch %d outside 1..%d", batch, m ? m->max_batch : 0);
    return FG_E_INVALID;
  }
  return FG_OK;
}

// y[batch x outputs] = MLP(x[batch x inputs]), row-major FP32
extern "C" fg_status fg_mlp_forward(fg_mlp *m, const float *x, int batch, float *y) {
  if (fg_status s = check_batch(m, batch)) return s;
  FG_HIP(hipMemcpyAsync(m->X, x, (size_t)batch * m->width[0] * 4, hipMemcpyHostToDevice, m->stream));
  if (fg_status s = mlp_forward(m, batch)) return s;
  FG_HIP(hipMemcpyAsync(y, m->Z[m->L - 1], (size_t)batch * m->width[m->L] * 4, hipMemcpyDeviceToHost, m->stream));
  FG_HIP(hipStreamSynchronize(m->stream));
  return FG_OK;
}

// MSE loss and its gradient with respect to every parameter (flat, in the
// fg_mlp_get_params order) for one batch; no update
extern "C" fg_status fg_mlp_gradient(fg_mlp *m, const float *x, const float *t, int batch, float *grad,
                                     float *loss) {
  if (fg_status s = check_batch(m, batch)) return s;
  FG_HIP(hipMemcpyAsync(m->X, x, (size_t)batch * m->width[0] * 4, hipMemcpyHostToDevice, m->stream));
  FG_HIP(hipMemcpyAsync(m->T, t, (size_t)batch * m->width[m->L] * 4, hipMemcpyHostToDevice, m->stream));
  if (!m->loss) {
    FG_HIP(hipMalloc((void **)&m->loss, 4));
    m->loss_cap = 1;
  }
  if (fg_status s = mlp_forward(m, batch)) return s;
  if (fg_status s = mlp_backward(m, batch, m->loss)) return s;
  if (grad) FG_HIP(hipMemcpyAsync(grad, m->g, m->nparam * 4, hipMemcpyDeviceToHost, m->stream));
  if (loss) FG_HIP(hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, m->stream));
  FG_HIP(hipStreamSynchronize(m->stream));
  return FG_OK;
}

// `steps` Adam steps on the n samples (x: n x inputs, t: n x outputs),
// cycling through consecutive batches of `batch` rows (n a multiple of
// batch); losses[s] = the loss of step s before its update
extern "C" fg_status fg_mlp_train(fg_mlp *m, const float *x, const float *t, int n, int batch, int steps, float lr,
                                  float *losses) {
  if (fg_status s = check_batch(m, batch)) return s;
  if (n < batch || n % batch || steps < 1) {
    fg_set_error("fg_mlp_train: n %d must be a positive multiple of batch %d", n, batch);
    return FG_E_INVALID;
  }
  const int din = m->width[0], dout = m->width[m->L];
  float *Xall = nullptr, *Tall = nullptr;
  FG_HIP(hipMalloc((void **)&Xall, (size_t)n * din * 4));
  FG_HIP(hipMalloc((void **)&Tall, (size_t)n * dout * 4));
  fg_status st = FG_OK;
  if (m->loss_cap < steps) {
    if (m->loss) (void)hipFree(m->loss);
    m->loss = nullptr;
    if (hipMalloc((void **)&m->loss, (size_t)steps * 4) != hipSuccess) st = FG_E_NOMEM;
    m->loss_cap = steps;
  }
  if (!st && (hipMemcpyAsync(Xall, x, (size_t)n * din * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
              hipMemcpyAsync(Tall, t, (size_t)n * dout * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess))
    st = FG_E_HIP;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const int nb = n / batch;
  for (int s = 0; s < steps && !st; s++) {
    const int q = s % nb;
    if (hipMemcpyAsync(m->X, Xall + (size_t)q * batch * din, (size_t)batch * din * 4, hipMemcpyDeviceToDevice,
                       m->stream) != hipSuccess ||
        hipMemcpyAsync(m->T, Tall + (size_t)q * batch * dout, (size_t)batch * dout * 4, hipMemcpyDeviceToDevice,
                       m->stream) != hipSuccess) {
      st = FG_E_HIP;
      break;
    }
    if ((st = mlp_forward(m, batch))) break;
    if ((st = mlp_backward(m, batch, m->loss + s))) break;
    m->step++;
    const float c1 = 1.0f - powf(b1, (float)m->step), c2 = 1.0f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(k_adam, dim3(fg_blocks(m->nparam, 256)), dim3(256), 0, m->stream, m->p, m->g, m->m, m->v,
                       m->pb, m->nparam, lr, b1, b2, eps, c1, c2);
    if (hipGetLastError() != hipSuccess) st = FG_E_HIP;
  }
  if (!st && losses &&
      hipMemcpyAsync(losses, m->loss, (size_t)steps * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
    st = FG_E_HIP;
  (void)hipStreamSynchronize(m->stream);
  (void)hipFree(Xall);
  (void)hipFree(Tall);
  if (st && st != FG_E_NOMEM) fg_set_error("fg_mlp_train: HIP error");
  return st;
}

// C[M x N] = A[M x K] B[K x N] with A, B rounded to BF16 and FP32
// accumulation (the layer GEMM on its own, for tests and tools)
extern "C" fg_status fg_bf16_gemm(const float *A, const float *B, int M, int N, int K, float *C) {
  if (!A || !B || !C || M < 1 || N < 1 || K < 1) return FG_E_INVALID;
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  bf16_t *bA = nullptr, *bB = nullptr;
  hipStream_t st = nullptr;
  fg_status rc = FG_OK;
  const size_t nA = (size_t)M * K, nB = (size_t)K * N, nC = (size_t)M * N;
  if (hipStreamCreate(&st) != hipSuccess || hipMalloc((void **)&dA, nA * 4) != hipSuccess ||
      hipMalloc((void **)&dB, nB * 4) != hipSuccess || hipMalloc((void **)&dC, nC * 4) != hipSuccess ||
      hipMalloc((void **)&bA, nA * 2) != hipSuccess || hipMalloc((void **)&bB, nB * 2) != hipSuccess ||
      hipMemcpyAsync(dA, A, nA * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(dB, B, nB * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
    rc = FG_E_HIP;
  } else {
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(fg_blocks((long)nA, 256)), dim3(256), 0, st, dA, bA, (long)nA);
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(fg_blocks((long)nB, 256)), dim3(256), 0, st, dB, bB, (long)nB);
    gemm_args g{};
    g.A = bA;
    g.sai = K;
    g.sak = 1;
    g.B = bB;
    g.sbk = N;
    g.sbj = 1;
    g.M = M;
    g.N = N;
    g.K = K;
    g.mode = EPI_STORE;
    g.C = dC;
    g.ldc = N;
    hipLaunchKernelGGL(k_gemm_bf16, dim3((N + GT - 1) / GT, (M + GT - 1) / GT), dim3(256), 0, st, g);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(C, dC, nC * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = FG_E_HIP;
  }
  if (rc) fg_set_error("fg_bf16_gemm: HIP error");
  void *ptrs[] = {dA, dB, dC, bA, bB};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}
