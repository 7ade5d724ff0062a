// mlp.hip — MLP surrogate: inputs -> features (as is, or cos/sin pairs of angle
// inputs) -> hidden x [width, GELU] -> 1, trained and run on the GPU. Hidden layers are BF16 MFMA GEMMs (v_mfma_f32_16x16x16bf16_1k) with
// FP32 accumulation; the scalar output layer is an FP32 dot product. Forward and
// backward are hand-written; FP32 master weights are updated with Adam.
//
// GEMM form used everywhere: C[M x N] = A[M x K] * B[N x K]^T, A and B row-major
// BF16 (both contiguous along K). One workgroup (4 waves) computes a 64 x 64 tile,
// each wave a 32 x 32 block as 2 x 2 MFMA 16x16 tiles, K in steps of 16.
//   forward      z_l = h_{l-1} W_l^T + b_l      (A = h_{l-1}, B = W_l)
//   input grads  dh_{l-1} = dz_l W_l            (A = dz_l,   B = W_l^T)
//   weight grads dW_l = dz_l^T h_{l-1}          (A = dz_l^T, B = h_{l-1}^T)
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "kernels.h"
#include "runtime.h"

namespace yh {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 round to nearest even (finite inputs), and back.
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t b) {
  return __builtin_bit_cast(float, uint32_t(b) << 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum Epi : int { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_GELU_GRAD = 2 };

// C = A * B^T. EPI_STORE: c32 = acc. EPI_BIAS_GELU: z = acc + bias -> c32 = z,
// c16 = bf16(gelu(z)). EPI_GELU_GRAD: g = acc * gelu'(zsrc) -> c32 = g, c16 = bf16(g).
__global__ __launch_bounds__(256) void k_gemm_nt(const uint16_t* __restrict__ A, int lda,
                                                 const uint16_t* __restrict__ B, int ldb, int M, int N, int K,
                                                 int epi, const float* __restrict__ bias,
                                                 const float* __restrict__ zsrc, float* __restrict__ c32,
                                                 uint16_t* __restrict__ c16, int ldc, int Mpad, int kc,
                                                 size_t zstride) {
  // split K: workgroup z sums k in [z kc, (z + 1) kc) and stores (EPI_STORE only) to
  // c32 + z zstride; kc is a multiple of 16
  const int kb = blockIdx.z * kc, ke = min(K, kb + kc);
  c32 += blockIdx.z * zstride;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  // rows past M / N are clamped to the last valid row; their outputs are not stored
  const uint16_t* ap[2];
  const uint16_t* bp[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) ap[i] = A + size_t(min(m0 + 16 * i + r, M - 1)) * lda + kq;
#pragma unroll
  for (int j = 0; j < 2; ++j) bp[j] = B + size_t(min(n0 + 16 * j + r, N - 1)) * ldb + kq;
  for (int k0 = kb; k0 < ke; k0 += 16) {
    s4 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const s4*>(ap[i] + k0);
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const s4*>(bp[j] + k0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  // D layout: lane holds rows 4*(lane/16)+e, column lane%16 of each 16x16 tile
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = m0 + 16 * i + (lane >> 4) * 4 + e;
        const int col = n0 + 16 * j + (lane & 15);
        if (col >= N) continue;
        const size_t idx = size_t(row) * ldc + col;
        if (row >= M) {
          // gradient rows between M and the padded row count are zeroed: the weight
          // gradient GEMM sums over the padded rows
          if (epi == EPI_GELU_GRAD && row < Mpad) {
            c32[idx] = 0.f;
            c16[idx] = 0;
          }
          continue;
        }
        float v = acc[i][j][e];
        if (epi == EPI_BIAS_GELU) {
          v += bias[col];
          c32[idx] = v;
          c16[idx] = f2bf(gelu(v));
        } else if (epi == EPI_GELU_GRAD) {
          v *= gelu_d(zsrc[idx]);
          c32[idx] = v;
          c16[idx] = f2bf(v);
        } else {
          c32[idx] = v;
        }
      }
}

// out[m] = sum_k bf16 h[m][k] * w[k] + b   (one wave per row)
__global__ __launch_bounds__(256) void k_out_fwd(const uint16_t* __restrict__ h, int K, const float* __restrict__ w,
                                                 const float* __restrict__ b, int M, float* __restrict__ out) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= M) return;
  float s = 0.f;
  for (int k = lane; k < K; k += 64) s += bf2f(h[size_t(row) * K + k]) * w[k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[row] = s + b[0];
}

// dout = 2 (out - y) / M; loss partials
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ out, const float* __restrict__ y, int M,
                                              float* __restrict__ dout, double* __restrict__ loss_partial) {
  __shared__ double sh[4];
  const int m = blockIdx.x * 256 + threadIdx.x;
  double l = 0.0;
  if (m < M) {
    const float d = out[m] - y[m];
    dout[m] = 2.f * d / float(M);
    l = double(d) * double(d);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) loss_partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// output layer backward: gw[k] = sum_m h[m][k] dout[m] (one wave per column k;
// column K is the bias: gb = sum_m dout[m])
__global__ __launch_bounds__(256) void k_out_bwd_w(const uint16_t* __restrict__ h, int K, int M,
                                                   const float* __restrict__ dout, float* __restrict__ gw,
                                                   float* __restrict__ gb) {
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (k > K) return;
  float s = 0.f;
  if (k < K)
    for (int m = lane; m < M; m += 64) s += bf2f(h[size_t(m) * K + k]) * dout[m];
  else
    for (int m = lane; m < M; m += 64) s += dout[m];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) {
    if (k < K)
      gw[k] = s;
    else
      gb[0] = s;
  }
}

// dz[m][k] = dout[m] w[k] gelu'(z[m][k]) for m < M, zero for M <= m < Mpad (fp32 and bf16)
__global__ __launch_bounds__(256) void k_out_bwd_dz(const float* __restrict__ dout, const float* __restrict__ w,
                                                    const float* __restrict__ z, int M, int Mpad, int K,
                                                    float* __restrict__ dz32, uint16_t* __restrict__ dz16) {
  const size_t e = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (e >= size_t(Mpad) * K) return;
  const int m = int(e / K), k = int(e % K);
  const float g = m < M ? dout[m] * w[k] * gelu_d(z[e]) : 0.f;
  dz32[e] = g;
  dz16[e] = f2bf(g);
}

// out[i] = sum over z = 0..S-1, in that order, of part[z stride + i]
__global__ __launch_bounds__(256) void k_sum_splits(const float* __restrict__ part, int S, size_t stride,
                                                    float* __restrict__ out, size_t n) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  float s = part[i];
  for (int z = 1; z < S; ++z) s += part[z * stride + i];
  out[i] = s;
}

// column sums of a row-major fp32 [M x N] matrix (one wave per column)
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ x, int M, int N, float* __restrict__ out) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (c >= N) return;
  float s = 0.f;
  for (int m = lane; m < M; m += 64) s += x[size_t(m) * N + c];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[c] = s;
}

// bf16 transpose [R x C] -> [C x R]
__global__ __launch_bounds__(256) void k_transpose(const uint16_t* __restrict__ in, int R, int C,
                                                   uint16_t* __restrict__ out) {
  __shared__ uint16_t t[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const int i = e / 32, j = e % 32;
    if (r0 + i < R && c0 + j < C) t[i][j] = in[size_t(r0 + i) * C + c0 + j];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const int i = e / 32, j = e % 32;  // out row c0+i, col r0+j
    if (c0 + i < C && r0 + j < R) out[size_t(c0 + i) * R + r0 + j] = t[j][i];
  }
}

// input rows -> bf16 features, optionally gathering rows (output row r is source row
// idx[r]). trig = 0: feature c = x_c. trig = 1: features 2j, 2j+1 = cos x_j, sin x_j.
__global__ __launch_bounds__(256) void k_encode_rows(const float* __restrict__ src, const int32_t* __restrict__ idx,
                                                     int rows, int cols, int trig, uint16_t* __restrict__ dst) {
  const int fcols = trig ? 2 * cols : cols;
  const size_t e = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (e >= size_t(rows) * fcols) return;
  const int r = int(e / fcols), c = int(e % fcols);
  const int sr = idx ? idx[r] : r;
  float v;
  if (trig) {
    const float a = src[size_t(sr) * cols + (c >> 1)];
    v = (c & 1) ? sinf(a) : cosf(a);
  } else {
    v = src[size_t(sr) * cols + c];
  }
  dst[e] = f2bf(v);
}

__global__ __launch_bounds__(256) void k_gather_f32(const float* __restrict__ src, const int32_t* __restrict__ idx,
                                                    int n, float* __restrict__ dst) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < n) dst[e] = src[idx[e]];
}

// acc[0] += (sum of the loss partials) / M   (one wave; fixed order)
__global__ __launch_bounds__(64) void k_loss_accum(const double* __restrict__ partial, int nblk, int M,
                                                   double* __restrict__ acc) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 64) s += partial[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (threadIdx.x == 0) acc[0] += s / M;
}

// Adam (bias-corrected) over the flat parameter vector
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
  const size_t e = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (e >= n) return;
  const float ge = g[e];
  const float me = b1 * m[e] + (1.f - b1) * ge;
  const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
  m[e] = me;
  v[e] = ve;
  p[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
}

// W (fp32 [N x K]) -> bf16 W and bf16 W^T
__global__ __launch_bounds__(256) void k_weights_bf16(const float* __restrict__ w, int N, int K, uint16_t* __restrict__ wb,
                                                      uint16_t* __restrict__ wtb) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= N * K) return;
  const int n = e / K, k = e % K;
  const uint16_t b = f2bf(w[e]);
  wb[e] = b;
  wtb[size_t(k) * N + n] = b;
}

inline unsigned blocks(size_t n) { return unsigned((n + 255) / 256); }

constexpr int kMaxSplit = 16;  // K-splits of a weight-gradient GEMM

}  // namespace

}  // namespace yh

using namespace yh;

struct yh_mlp {
  int in_dim = 0, feat = 0, trig = 0, width = 0, hidden = 0, max_batch = 0;
  std::vector<size_t> woff, boff;  // per layer (hidden layers then the output layer)
  size_t nparams = 0;
  Workspace w;
  DevBuf params, grads, am, av;
  std::vector<DevBuf*> wbf, wtbf, zbuf, hbf, dz32, dz16, tmp32;
  DevBuf x32, xbf, hT, dzT, out, dout, y, lossp, eacc, idx, data_x, data_y;
  DevBuf wsplit;  // split-K partial weight gradients, kMaxSplit x width x max fan-in
  size_t data_rows = 0;
  int64_t adam_t = 0;
  ~yh_mlp() {
    for (auto* v : {&wbf, &wtbf, &zbuf, &hbf, &dz32, &dz16, &tmp32})
      for (DevBuf* b : *v) delete b;
  }
  int fan_in(int l) const { return l == 0 ? feat : width; }
  int fan_out(int l) const { return l == hidden ? 1 : width; }
};

namespace yh {

static float* fptr(DevBuf& b, size_t off = 0) { return static_cast<float*>(b.p) + off; }

// K-splits of a weight-gradient GEMM over Mp rows: chunks of at least 64 rows
static int split_count(int Mp) { return std::max(1, std::min(kMaxSplit, Mp / 64)); }
static uint16_t* hptr(DevBuf& b) { return static_cast<uint16_t*>(b.p); }

// forward over `rows` rows already in m->xbf (bf16 [rows x feat]); output in m->out
static int mlp_forward(yh_mlp* m, int rows) {
  hipStream_t st = m->w.stream;
  const int Mp = (rows + 63) / 64 * 64;
  const uint16_t* hin = hptr(m->xbf);
  for (int l = 0; l < m->hidden; ++l) {
    const int K = m->fan_in(l), N = m->width;
    dim3 grid(N / 64, Mp / 64);
    hipLaunchKernelGGL(k_gemm_nt, grid, dim3(256), 0, st, hin, K, hptr(*m->wbf[l]), K, rows, N, K, int(EPI_BIAS_GELU),
                       fptr(m->params, m->boff[l]), static_cast<const float*>(nullptr), fptr(*m->zbuf[l]),
                       hptr(*m->hbf[l]), N, Mp, K, size_t(0));
    hin = hptr(*m->hbf[l]);
  }
  const int L = m->hidden;
  hipLaunchKernelGGL(k_out_fwd, dim3((rows + 3) / 4), dim3(256), 0, st, hin, m->width, fptr(m->params, m->woff[L]),
                     fptr(m->params, m->boff[L]), rows, fptr(m->out));
  return YH_OK;
}

// backward of the last forward (rows <= max_batch, targets in m->y); grads in m->grads
static int mlp_backward(yh_mlp* m, int rows) {
  hipStream_t st = m->w.stream;
  const int Mp = (rows + 63) / 64 * 64;
  const int W = m->width, L = m->hidden;
  hipLaunchKernelGGL(k_loss, dim3(blocks(rows)), dim3(256), 0, st, fptr(m->out), fptr(m->y), rows, fptr(m->dout),
                     static_cast<double*>(m->lossp.p));
  // output layer
  hipLaunchKernelGGL(k_out_bwd_w, dim3((W + 1 + 3) / 4), dim3(256), 0, st, hptr(*m->hbf[L - 1]), W, rows, fptr(m->dout),
                     fptr(m->grads, m->woff[L]), fptr(m->grads, m->boff[L]));
  hipLaunchKernelGGL(k_out_bwd_dz, dim3(blocks(size_t(Mp) * W)), dim3(256), 0, st, fptr(m->dout),
                     fptr(m->params, m->woff[L]), fptr(*m->zbuf[L - 1]), rows, Mp, W, fptr(*m->dz32[L - 1]),
                     hptr(*m->dz16[L - 1]));
  for (int l = L - 1; l >= 0; --l) {
    const int K = m->fan_in(l), N = W;
    const uint16_t* hprev = l == 0 ? hptr(m->xbf) : hptr(*m->hbf[l - 1]);
    // db_l = column sums of dz_l
    hipLaunchKernelGGL(k_colsum, dim3((N + 3) / 4), dim3(256), 0, st, fptr(*m->dz32[l]), rows, N, fptr(m->grads, m->boff[l]));
    // dW_l [N x K] = dz_l^T [N x M] * (h_{l-1}^T [K x M])^T
    hipLaunchKernelGGL(k_transpose, dim3((N + 31) / 32, (Mp + 31) / 32), dim3(256), 0, st, hptr(*m->dz16[l]), Mp, N,
                       hptr(m->dzT));
    hipLaunchKernelGGL(k_transpose, dim3((K + 31) / 32, (Mp + 31) / 32), dim3(256), 0, st, hprev, Mp, K, hptr(m->hT));
    // the reduction runs over the Mp batch rows: split it so that the launch has
    // (K / 64) x (N / 64) x S workgroups, then sum the S partials in a fixed order
    const int S = split_count(Mp), kc = (Mp / S + 15) / 16 * 16, nz = (Mp + kc - 1) / kc;
    float* dw = nz > 1 ? fptr(m->wsplit) : fptr(m->grads, m->woff[l]);
    hipLaunchKernelGGL(k_gemm_nt, dim3((K + 63) / 64, N / 64, nz), dim3(256), 0, st, hptr(m->dzT), Mp, hptr(m->hT), Mp,
                       N, K, Mp, int(EPI_STORE), static_cast<const float*>(nullptr),
                       static_cast<const float*>(nullptr), dw, static_cast<uint16_t*>(nullptr), K, N, kc,
                       size_t(N) * K);
    if (nz > 1)
      hipLaunchKernelGGL(k_sum_splits, dim3(blocks(size_t(N) * K)), dim3(256), 0, st, dw, nz, size_t(N) * K,
                         fptr(m->grads, m->woff[l]), size_t(N) * K);
    if (l > 0) {
      // dz_{l-1} = (dz_l W_l) * gelu'(z_{l-1})
      hipLaunchKernelGGL(k_gemm_nt, dim3((K + 63) / 64, Mp / 64), dim3(256), 0, st, hptr(*m->dz16[l]), N,
                         hptr(*m->wtbf[l]), N, rows, K, N, int(EPI_GELU_GRAD), static_cast<const float*>(nullptr),
                         fptr(*m->zbuf[l - 1]), fptr(*m->dz32[l - 1]), hptr(*m->dz16[l - 1]), K, Mp, N, size_t(0));
    }
  }
  return YH_OK;
}

static int mlp_refresh_bf16(yh_mlp* m) {
  for (int l = 0; l < m->hidden; ++l) {
    const int K = m->fan_in(l), N = m->width;
    hipLaunchKernelGGL(k_weights_bf16, dim3(blocks(size_t(N) * K)), dim3(256), 0, m->w.stream,
                       fptr(m->params, m->woff[l]), N, K, hptr(*m->wbf[l]), hptr(*m->wtbf[l]));
  }
  return YH_OK;
}

static int mlp_alloc(yh_mlp* m) {
  const size_t B = size_t(m->max_batch), W = size_t(m->width);
  YH_TRY(m->w.init());
  YH_TRY(m->params.ensure(m->nparams * 4));
  YH_TRY(m->grads.ensure(m->nparams * 4));
  YH_TRY(m->am.ensure(m->nparams * 4));
  YH_TRY(m->av.ensure(m->nparams * 4));
  for (int l = 0; l < m->hidden; ++l) {
    const size_t K = size_t(m->fan_in(l));
    for (auto* v : {&m->wbf, &m->wtbf, &m->zbuf, &m->hbf, &m->dz32, &m->dz16}) v->push_back(new DevBuf());
    YH_TRY(m->wbf[l]->ensure(W * K * 2));
    YH_TRY(m->wtbf[l]->ensure(W * K * 2));
    YH_TRY(m->zbuf[l]->ensure(B * W * 4));
    YH_TRY(m->hbf[l]->ensure(B * W * 2));
    YH_TRY(m->dz32[l]->ensure(B * W * 4));
    YH_TRY(m->dz16[l]->ensure(B * W * 2));
  }
  const size_t maxk = std::max<size_t>(W, size_t(m->feat));
  YH_TRY(m->x32.ensure(B * m->in_dim * 4));
  YH_TRY(m->xbf.ensure(B * m->feat * 2));
  YH_TRY(m->hT.ensure(B * maxk * 2));
  YH_TRY(m->dzT.ensure(B * W * 2));
  YH_TRY(m->wsplit.ensure(size_t(kMaxSplit) * W * maxk * 4));
  YH_TRY(m->out.ensure(B * 4));
  YH_TRY(m->dout.ensure(B * 4));
  YH_TRY(m->y.ensure(B * 4));
  YH_TRY(m->lossp.ensure((B / 256 + 1) * 8));
  YH_TRY(m->idx.ensure(B * 4));
  YH_TRY(m->eacc.ensure(8));
  // zero the padding rows once so GEMMs over padded row counts read zeros
  YH_HIP(hipMemsetAsync(m->xbf.p, 0, B * m->feat * 2, m->w.stream));
  for (int l = 0; l < m->hidden; ++l) {
    YH_HIP(hipMemsetAsync(m->hbf[l]->p, 0, B * W * 2, m->w.stream));
    YH_HIP(hipMemsetAsync(m->dz16[l]->p, 0, B * W * 2, m->w.stream));
  }
  return m->w.sync();
}

// rows -> xbf (bf16) from a host fp32 matrix
static int mlp_load_rows(yh_mlp* m, const float* x, int rows) {
  YH_TRY(m->w.upload(x, size_t(rows) * m->in_dim * 4, m->x32));
  hipLaunchKernelGGL(k_encode_rows, dim3(blocks(size_t(rows) * m->feat)), dim3(256), 0, m->w.stream, fptr(m->x32),
                     static_cast<const int32_t*>(nullptr), rows, m->in_dim, m->trig, hptr(m->xbf));
  return YH_OK;
}

static double loss_sum(yh_mlp* m, int rows) {
  std::vector<double> p(blocks(rows));
  if (m->w.fetch(p.data(), m->lossp.p, p.size() * 8) != YH_OK) return -1.0;
  double s = 0.0;
  for (double v : p) s += v;
  return s / rows;
}

}  // namespace yh

extern "C" {

int yh_mlp_create(int in_dim, int encoding, int width, int hidden, int max_batch, uint64_t seed, yh_mlp** out) {
  if (!out) return fail(YH_ERR_INVALID_ARG, "null output pointer");
  if (encoding != YH_MLP_LINEAR && encoding != YH_MLP_TRIG) return fail(YH_ERR_INVALID_ARG, "unknown input encoding");
  const int feat = encoding == YH_MLP_TRIG ? 2 * in_dim : in_dim;
  if (in_dim < 1 || feat % 16 || width < 64 || width % 64 || hidden < 1 || hidden > 16 || max_batch < 64 ||
      max_batch % 64)
    return fail(YH_ERR_INVALID_ARG,
                "layer-0 features % 16 == 0, width % 64 == 0, 1 <= hidden <= 16, max_batch % 64 == 0");
  YH_TRY(ensure_init());
  yh_mlp* m = new yh_mlp();
  m->in_dim = in_dim;
  m->feat = feat;
  m->trig = encoding == YH_MLP_TRIG;
  m->width = width;
  m->hidden = hidden;
  m->max_batch = max_batch;
  size_t off = 0;
  for (int l = 0; l <= hidden; ++l) {
    m->woff.push_back(off);
    off += size_t(m->fan_out(l)) * m->fan_in(l);
    m->boff.push_back(off);
    off += size_t(m->fan_out(l));
  }
  m->nparams = off;
  int st = mlp_alloc(m);
  if (st != YH_OK) {
    delete m;
    return st;
  }
  // U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases, seeded
  std::mt19937_64 rng(seed);
  std::vector<float> p(off);
  for (int l = 0; l <= hidden; ++l) {
    const float bound = 1.0f / std::sqrt(float(m->fan_in(l)));
    std::uniform_real_distribution<float> U(-bound, bound);
    for (size_t k = m->woff[l]; k < m->boff[l] + size_t(m->fan_out(l)); ++k) p[k] = U(rng);
  }
  st = yh_mlp_set_params(m, p.data(), p.size());
  if (st != YH_OK) {
    delete m;
    return st;
  }
  *out = m;
  return YH_OK;
}

int yh_mlp_destroy(yh_mlp* m) {
  delete m;
  return YH_OK;
}

int yh_mlp_num_params(const yh_mlp* m, size_t* n) {
  if (!m || !n) return fail(YH_ERR_INVALID_ARG, "null argument");
  *n = m->nparams;
  return YH_OK;
}

int yh_mlp_get_params(const yh_mlp* cm, float* params, size_t n) {
  if (!cm || !params || n != cm->nparams) return fail(YH_ERR_INVALID_ARG, "parameter count mismatch");
  yh_mlp* m = const_cast<yh_mlp*>(cm);
  return m->w.fetch(params, m->params.p, n * 4);
}

int yh_mlp_set_params(yh_mlp* m, const float* params, size_t n) {
  if (!m || !params || n != m->nparams) return fail(YH_ERR_INVALID_ARG, "parameter count mismatch");
  YH_TRY(m->w.upload(params, n * 4, m->params));
  YH_HIP(hipMemsetAsync(m->am.p, 0, n * 4, m->w.stream));
  YH_HIP(hipMemsetAsync(m->av.p, 0, n * 4, m->w.stream));
  m->adam_t = 0;
  YH_TRY(mlp_refresh_bf16(m));
  return m->w.sync();
}

int yh_mlp_predict(yh_mlp* m, const float* x, size_t n, float* y) {
  if (!m || !x || !y) return fail(YH_ERR_INVALID_ARG, "null argument");
  for (size_t r0 = 0; r0 < n; r0 += size_t(m->max_batch)) {
    const int rows = int(std::min<size_t>(size_t(m->max_batch), n - r0));
    YH_TRY(mlp_load_rows(m, x + r0 * m->in_dim, rows));
    YH_TRY(mlp_forward(m, rows));
    YH_TRY(m->w.fetch(y + r0, m->out.p, size_t(rows) * 4));
  }
  return YH_OK;
}

int yh_mlp_loss_grad(yh_mlp* m, const float* x, const float* y, int n, float* loss, float* grads) {
  if (!m || !x || !y || !loss) return fail(YH_ERR_INVALID_ARG, "null argument");
  if (n < 1 || n > m->max_batch) return fail(YH_ERR_INVALID_ARG, "batch larger than max_batch");
  YH_TRY(mlp_load_rows(m, x, n));
  YH_TRY(m->w.upload(y, size_t(n) * 4, m->y));
  YH_TRY(mlp_forward(m, n));
  YH_TRY(mlp_backward(m, n));
  YH_TRY(m->w.sync());
  *loss = float(loss_sum(m, n));
  if (grads) YH_TRY(m->w.fetch(grads, m->grads.p, m->nparams * 4));
  return YH_OK;
}

int yh_mlp_train(yh_mlp* m, const float* x, const float* y, size_t n, int batch, int epochs, float lr, uint64_t seed,
                 float* epoch_loss) {
  if (!m || !x || !y) return fail(YH_ERR_INVALID_ARG, "null argument");
  if (batch < 64 || batch > m->max_batch || batch % 64 || n < size_t(batch))
    return fail(YH_ERR_INVALID_ARG, "batch must be a multiple of 64, <= max_batch and <= n");
  // the whole training set stays on the device (fp32 rows, gathered per batch)
  YH_TRY(m->w.upload(x, n * m->in_dim * 4, m->data_x));
  YH_TRY(m->w.upload(y, n * 4, m->data_y));
  std::mt19937_64 rng(seed);
  std::vector<int32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  DevBuf dperm;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const size_t nb = n / size_t(batch);
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(perm.begin(), perm.end(), rng);
    YH_TRY(m->w.upload(perm.data(), n * 4, dperm));
    YH_HIP(hipMemsetAsync(m->eacc.p, 0, 8, m->w.stream));
    for (size_t bi = 0; bi < nb; ++bi) {
      const int32_t* bidx = static_cast<const int32_t*>(dperm.p) + bi * batch;
      hipLaunchKernelGGL(k_encode_rows, dim3(blocks(size_t(batch) * m->feat)), dim3(256), 0, m->w.stream,
                         fptr(m->data_x), bidx, batch, m->in_dim, m->trig, hptr(m->xbf));
      hipLaunchKernelGGL(k_gather_f32, dim3(blocks(batch)), dim3(256), 0, m->w.stream, fptr(m->data_y), bidx, batch,
                         fptr(m->y));
      YH_TRY(mlp_forward(m, batch));
      YH_TRY(mlp_backward(m, batch));
      hipLaunchKernelGGL(k_loss_accum, dim3(1), dim3(64), 0, m->w.stream, static_cast<const double*>(m->lossp.p),
                         int(blocks(batch)), batch, static_cast<double*>(m->eacc.p));
      m->adam_t += 1;
      const float c1 = 1.f - std::pow(b1, float(m->adam_t)), c2 = 1.f - std::pow(b2, float(m->adam_t));
      hipLaunchKernelGGL(k_adam, dim3(blocks(m->nparams)), dim3(256), 0, m->w.stream, fptr(m->params), fptr(m->grads),
                         fptr(m->am), fptr(m->av), m->nparams, lr, b1, b2, eps, c1, c2);
      YH_TRY(mlp_refresh_bf16(m));
    }
    double acc = 0.0;
    YH_TRY(m->w.fetch(&acc, m->eacc.p, 8));
    acc /= double(nb);
    if (epoch_loss) epoch_loss[ep] = float(acc);
  }
  return YH_OK;
}

}  // extern "C"
