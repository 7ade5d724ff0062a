// mlp.hip — MLP surrogate: inputs -> features (as is, or cos/sin pairs of angle
// inputs) -> hidden x [width, GELU] -> 1, trained and run on the GPU. Hidden layers
// are BF16 MFMA products (v_mfma_f32_16x16x32_bf16) with FP32 accumulation; the
// scalar output layer is an FP32 dot product. Forward and backward are
// hand-written; FP32 master weights are updated with Adam.
//
// A training step is four launches, replayed from a captured hipGraph:
//   k_mlp_fwd     per block of 16 batch rows, the whole forward pass: features,
//                 every hidden layer (activations stay in LDS, weights from L2),
//                 output; stores z and the bf16 activations, transposed;
//   k_mlp_bwd     per block of 16 rows, the row-local backward chain: loss
//                 gradient, output layer, dz_{l-1} = (dz_l W_l) * gelu'(z_{l-1}) for
//                 every layer; stores the transposed bf16 dz and per-block partial
//                 sums (bias gradients, output-layer gradients, squared error);
//   k_mlp_wgrad   dW_l = dz_l^T h_{l-1} of every layer, split over the batch rows;
//   k_mlp_update  every partial sum in a fixed order (deterministic gradients), Adam,
//                 the bf16 weight copies, and the epoch's loss sum.
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <thread>
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

// GELU and its derivative at z, sharing erf(z / sqrt 2):
// gelu(z) = z Phi(z), gelu'(z) = Phi(z) + z phi(z), Phi(z) = (1 + erf(z / sqrt 2)) / 2
__device__ __forceinline__ void gelu_both(float z, float& g, float& gd) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  g = z * cdf;
  gd = cdf + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kRows = 16;       // batch rows per workgroup of the row-block kernels
constexpr int kMaxLayers = 16;  // hidden layers
constexpr int kLdsPad = 8;      // bf16 elements of padding per activation row in LDS

__device__ __forceinline__ f4 mfma32(bf8 a, bf8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// MFMA fragment order of an R x K bf16 matrix (R % 16 == 0, K % 16 == 0): the layout
// of every bf16 operand the products read from global memory (weights, transposed
// activations and gradients). Rows go in groups of 16, 16 K elements per group;
// inside a group, 32-column steps of 512 elements in which lane 16 q + row % 16
// holds columns 32 s + 8 q .. + 7 (the B operand of v_mfma_f32_16x16x32_bf16), and
// for K % 32 == 16 a last step of 256 elements in which that lane holds columns
// 32 s + 4 q .. + 3 (v_mfma_f32_16x16x16_bf16). A wave reads one step of one group as
// 1 KB of contiguous memory.
__host__ __device__ __forceinline__ size_t frag_off(int row, int col, int K) {
  const size_t base = size_t(row >> 4) * 16 * K + size_t(col >> 5) * 512;
  const int r = row & 15, kk = col & 31;
  if ((col | 31) < K) return base + size_t(((kk >> 3) * 16 + r) * 8 + (kk & 7));
  return base + size_t(((kk >> 2) * 16 + r) * 4 + (kk & 3));
}

// GELU derivatives gelu'(z) [Mp x W] in accumulator order: 16 x 16 tiles (16 batch rows x 16
// units), tile (m / 16, n / 16) at 256 (m / 16 (W / 16) + n / 16); in a tile, lane
// 16 ((m % 16) / 4) + n % 16 holds rows 4 ((m % 16) / 4) .. + 3 of unit n, as 16 bytes.
__device__ __forceinline__ size_t zoff(int m, int n, int W) {
  return (size_t(m >> 4) * (W >> 4) + (n >> 4)) * 256 + size_t((((m & 15) >> 2) * 16 + (n & 15)) * 4 + (m & 3));
}

// One wave's 16 x 64 block of A B^T: rows 0..15 of A (bf16 in LDS, row stride lda
// elements) against rows n0..n0+63 of B (bf16 [N x K] in global memory in fragment
// order; row groups past N read the last group and their results are not stored),
// K % 16 == 0.
// v_mfma_f32_16x16x32_bf16: lane l feeds row (A) or column (B) l % 16 with the 8
// consecutive k at 8 (l / 16) of each 32-k step (A and B take k in the same lane
// order, so the sum does not depend on the instruction's order of k within a step);
// D holds rows 4 (l / 16) + e, column l % 16 of each 16 x 16 tile.
// The 32-k steps run in the rotated order s = (i + rot) mod nsteps, i = 0, 1, ...:
// workgroups that read the same weights at the same time start at different steps,
// so their requests spread over the L2 instead of all asking for the same lines
// (the k-step order, and so the FP32 rounding, is fixed by rot). B is loaded in
// chunks of kChunk steps, the next chunk's loads in flight while the current chunk's
// MFMAs run; inside the loop the loads are unconditional and the last chunks are
// peeled, so the wait before each MFMA group covers only that group's loads. The
// caller starts the first chunk's loads (gemm_first) early, e.g. before the
// previous layer's epilogue, and hands them in.
constexpr int kChunk = 4;
struct BChunk {
  bf8 b[kChunk][4];
};
struct BRows {
  const uint16_t* p[4];
  int nsteps, rot;
  __device__ __forceinline__ int at(int i) const {  // step of position i in the rotated order
    const int s = i + rot;
    return s >= nsteps ? s - nsteps : s;
  }
};
__device__ __forceinline__ BRows gemm_rows(const uint16_t* __restrict__ B, int K, int N, int n0, int rot) {
  const int lane = threadIdx.x & 63;
  BRows br;
#pragma unroll
  for (int j = 0; j < 4; ++j) br.p[j] = B + size_t(min((n0 >> 4) + j, (N >> 4) - 1)) * 16 * K + lane * 8;
  br.nsteps = K >> 5;
  br.rot = br.nsteps > 0 ? rot % br.nsteps : 0;
  return br;
}
__device__ __forceinline__ void load_chunk(BChunk& o, const BRows& br, int c) {
#pragma unroll
  for (int t = 0; t < kChunk; ++t) {
    const int s = br.at(c * kChunk + t);
#pragma unroll
    for (int j = 0; j < 4; ++j) o.b[t][j] = *reinterpret_cast<const bf8*>(br.p[j] + 512 * s);
  }
}
// loads of the first chunk: positions 0 .. kChunk - 1, clamped to the last whole
// step (none for K < 32)
__device__ __forceinline__ void gemm_first(BChunk& o, const BRows& br) {
  const int last = br.nsteps - 1;
  if (last >= 0) {
#pragma unroll
    for (int t = 0; t < kChunk; ++t) {
      const int s = br.at(min(t, last));
#pragma unroll
      for (int j = 0; j < 4; ++j) o.b[t][j] = *reinterpret_cast<const bf8*>(br.p[j] + 512 * s);
    }
  }
}

__device__ __forceinline__ void block_gemm(const uint16_t* A, int lda, const BRows& br, int K, const BChunk& first,
                                           f4 (&acc)[4]) {
  const int lane = threadIdx.x & 63, kq = (lane >> 4) * 8;
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = f4{0.f, 0.f, 0.f, 0.f};
  const uint16_t* ap = A + (lane & 15) * lda + kq;
  auto step = [&](const bf8 (&bv)[4], int s) __attribute__((always_inline)) {
    const bf8 av = *reinterpret_cast<const bf8*>(ap + 32 * s);
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = mfma32(av, bv[j], acc[j]);
  };
  auto compute = [&](const BChunk& o, int c) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < kChunk; ++t) step(o.b[t], br.at(c * kChunk + t));
  };
  const int nsteps = br.nsteps, nfull = nsteps / kChunk;
  if (nfull > 0) {
    BChunk o0 = first, o1;
    int c = 0;
    for (; c + 2 < nfull; c += 2) {
      load_chunk(o1, br, c + 1);
      compute(o0, c);
      load_chunk(o0, br, c + 2);
      compute(o1, c + 1);
    }
    if (c + 1 < nfull) {  // the last two chunks
      load_chunk(o1, br, c + 1);
      compute(o0, c);
      compute(o1, c + 1);
    } else {  // the last chunk
      compute(o0, c);
    }
  } else {  // fewer than kChunk steps, all loaded by gemm_first
#pragma unroll
    for (int t = 0; t < kChunk; ++t)
      if (t < nsteps) step(first.b[t], br.at(t));
  }
  for (int i = nfull > 0 ? nfull * kChunk : nsteps; i < nsteps; ++i) {
    const int s = br.at(i);
    bf8 bv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) bv[j] = *reinterpret_cast<const bf8*>(br.p[j] + 512 * s);
    step(bv, s);
  }
  if (K & 16) {  // a final 16-k step on v_mfma_f32_16x16x16_bf16: 4 k per lane at 4 (l / 16)
    const int k0 = 32 * nsteps - kq + (lane >> 4) * 4;
    const s4 av = *reinterpret_cast<const s4*>(ap + k0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const s4 bv = *reinterpret_cast<const s4*>(br.p[j] + 512 * nsteps - 4 * lane);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[j], 0, 0, 0);
    }
  }
}

// four consecutive rows of one column, packed as the 8 bytes of a transposed row
__device__ __forceinline__ void store4t(uint16_t* dst, const uint16_t (&v)[4]) {
  uint2 pk;
  pk.x = uint32_t(v[0]) | (uint32_t(v[1]) << 16);
  pk.y = uint32_t(v[2]) | (uint32_t(v[3]) << 16);
  *reinterpret_cast<uint2*>(dst) = pk;
}

// Everything the row-block kernels read and write, by value.
struct MlpArgs {
  int in_dim, feat, trig, width, hidden, rows, Mp;
  const float* x;             // fp32 input rows
  const int32_t* idx;         // gather: row r of the batch is x row idx[off + r] (or r without idx)
  const int64_t* step;        // off = *step * rows (a replayed step's batch index), else 0
  int64_t* ctr;               // training: the backward pass advances ctr[0] (batch index) and
                              // ctr[1] (Adam's step count) once the forward pass has read them
  const float* ysrc;          // with idx: targets y[r] = ysrc[idx[off + r]]
  float* y;                   // targets of the batch
  const float* params;        // fp32 master parameters
  size_t woff[kMaxLayers + 1], boff[kMaxLayers + 1];
  const uint16_t* wbf[kMaxLayers];   // W_l bf16 [width x fan_in], fragment order
  const uint16_t* wtbf[kMaxLayers];  // W_l^T bf16 [fan_in x width], fragment order
  float* gd[kMaxLayers];             // GELU derivatives at the pre-activations [Mp x width], accumulator order
  uint16_t* hT[kMaxLayers];          // activations, transposed [width x Mp], fragment order
  uint16_t* xT;                      // input features, transposed [feat x Mp], fragment order
  uint16_t* dzT[kMaxLayers];         // pre-activation gradients, transposed [width x Mp], fragment order
  float* bpart[kMaxLayers];          // bias-gradient partials [Mp / kRows x width]
  float* out;                        // network output [Mp]
  float* gwp;                        // output-layer weight-gradient partials [Mp / kRows x width]
  float* gbp;                        // output-bias partials [Mp / kRows]
  double* lossp;                     // squared-error partials [Mp / kRows]
};

constexpr int kEncBatch = 8;  // input loads in flight per thread while encoding

// Forward pass of a block of kRows batch rows: the input features (gathered rows,
// optionally as cos / sin pairs) go to LDS and, transposed, to xT; each hidden
// layer is one LDS-to-LDS MFMA product (weights from global memory, L2-resident)
// whose epilogue adds the bias and applies GELU, storing gelu'(z) and the bf16
// activations (transposed) for the backward pass. The scalar output layer is folded into the last
// hidden layer's epilogue: out[m] = sum_c bf16(h[m][c]) w[c] + b, summed per lane
// over its columns, then over the 16 lanes of a row group (fixed butterfly), then
// over the waves in order. Rows past `rows` are zero inputs; their activations are
// stored transposed as zeros.
__global__ __launch_bounds__(256) void k_mlp_fwd(MlpArgs a) {
  extern __shared__ uint16_t smem[];
  __shared__ int64_t src_s[kRows];
  __shared__ float out_s[4][kRows];
  const int F = a.feat, W = a.width, L = a.hidden, ldF = max(F, W) + kLdsPad, ldW = W + kLdsPad;
  uint16_t* buf0 = smem;                 // kRows x ldF
  uint16_t* buf1 = smem + kRows * ldF;   // kRows x ldW
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * kRows;
  // the first weight chunk of layer 0 is in flight while the inputs are encoded
  const int rot = blockIdx.x >> 3;  // the row blocks on one XCD start their k-steps apart
  BRows br = gemm_rows(a.wbf[0], F, W, 64 * wave, rot);
  BChunk pre;
  gemm_first(pre, br);
  if (tid < kRows) {
    const int m = m0 + tid;
    int64_t sr = -1;
    if (m < a.rows) {
      const int64_t off = a.step ? *a.step * int64_t(a.rows) : 0;
      sr = a.idx ? a.idx[off + m] : m;
      if (a.ysrc) a.y[m] = a.ysrc[sr];
    }
    src_s[tid] = sr;
  }
  __syncthreads();
  // ---- input features: coalesced row reads, kEncBatch loads in flight per thread;
  // groups of 4 inputs as one 16-byte load where the rows allow it ----
  const int cols = a.in_dim, vec = (cols & 3) == 0 ? 4 : 1, cv = cols / vec, total = kRows * cv;
  for (int base = 0; base < total; base += 256 * kEncBatch) {
    f4 v[kEncBatch];
#pragma unroll
    for (int u = 0; u < kEncBatch; ++u) {
      const int e = min(base + 256 * u + tid, total - 1);
      const int r = e / cv, j = e - r * cv;
      const float* src = a.x + max(src_s[r], int64_t(0)) * cols + j * vec;
      if (vec == 4) {
        v[u] = *reinterpret_cast<const f4*>(src);
      } else {
        v[u] = f4{*src, 0.f, 0.f, 0.f};
      }
    }
#pragma unroll
    for (int u = 0; u < kEncBatch; ++u) {
      const int e = base + 256 * u + tid;
      if (e < total) {
        const int r = e / cv, j = e - r * cv;
        const bool live = src_s[r] >= 0;
        uint16_t* dst = buf0 + r * ldF + (a.trig ? 2 : 1) * j * vec;
        if (a.trig) {
          uint32_t pk[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            pk[q] = 0u;
            if (q < vec) {
              float sn, cs;
              sincosf(v[u][q], &sn, &cs);
              if (live) pk[q] = uint32_t(f2bf(cs)) | (uint32_t(f2bf(sn)) << 16);
            }
          }
          if (vec == 4) {
            *reinterpret_cast<uint4*>(dst) = uint4{pk[0], pk[1], pk[2], pk[3]};
          } else {
            *reinterpret_cast<uint32_t*>(dst) = pk[0];
          }
        } else {
          uint16_t hb[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) hb[q] = live ? f2bf(v[u][q]) : uint16_t(0);
          if (vec == 4) {
            *reinterpret_cast<uint2*>(dst) = uint2{uint32_t(hb[0]) | (uint32_t(hb[1]) << 16),
                                                   uint32_t(hb[2]) | (uint32_t(hb[3]) << 16)};
          } else {
            dst[0] = hb[0];
          }
        }
      }
    }
  }
  __syncthreads();
  // transposed features: 8 rows of one column per 16-byte store
  for (int it = tid; it < 2 * F; it += 256) {
    const int c = it >> 1, g = (it & 1) * 8;
    uint32_t w4[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      w4[q] = uint32_t(buf0[(g + 2 * q) * ldF + c]) | (uint32_t(buf0[(g + 2 * q + 1) * ldF + c]) << 16);
    *reinterpret_cast<uint4*>(a.xT + frag_off(c, m0 + g, a.Mp)) = uint4{w4[0], w4[1], w4[2], w4[3]};
  }
  // ---- hidden layers ----
  uint16_t* in = buf0;
  uint16_t* outb = buf1;
  int ldin = ldF, ldout = ldW;
  const int r0 = (lane >> 4) * 4;
  const float* wout = a.params + a.woff[L];
  float osum[4] = {0.f, 0.f, 0.f, 0.f};  // output-layer partials of rows r0 .. r0 + 3
  for (int l = 0; l < L; ++l) {
    const int K = l == 0 ? F : W;
    const bool last = l == L - 1;
    const float* bias = a.params + a.boff[l];
    for (int n0 = 64 * wave; n0 < W; n0 += 256) {
      if (n0 != 64 * wave || l > 0) {
        br = gemm_rows(a.wbf[l], K, W, n0, rot);
        gemm_first(pre, br);
      }
      float bn[4], wn[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        bn[j] = bias[n0 + 16 * j + (lane & 15)];
        wn[j] = last ? wout[n0 + 16 * j + (lane & 15)] : 0.f;
      }
      f4 acc[4];
      block_gemm(in, ldin, br, K, pre, acc);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int n = n0 + 16 * j + (lane & 15);
        f4 dv;
        uint16_t tv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float g, gdv;
          gelu_both(acc[j][e] + bn[j], g, gdv);
          dv[e] = gdv;
          const uint16_t hb = f2bf(g);
          if (last) {
            osum[e] += bf2f(hb) * wn[j];
          } else {
            outb[(r0 + e) * ldout + n] = hb;
          }
          tv[e] = m0 + r0 + e < a.rows ? hb : uint16_t(0);
        }
        *reinterpret_cast<f4*>(a.gd[l] + zoff(m0 + r0, n, W)) = dv;
        store4t(a.hT[l] + frag_off(n, m0 + r0, a.Mp), tv);
      }
    }
    if (!last) {
      __syncthreads();
      uint16_t* t = in;
      in = outb;
      outb = t;
      const int tl = ldin;
      ldin = ldout;
      ldout = tl;
    }
  }
  // ---- output layer: the 16 lanes of each row group, then the waves in order ----
#pragma unroll
  for (int e = 0; e < 4; ++e) {
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) osum[e] += __shfl_xor(osum[e], o, 64);
  }
  if ((lane & 15) == 0) {
#pragma unroll
    for (int e = 0; e < 4; ++e) out_s[wave][r0 + e] = osum[e];
  }
  __syncthreads();
  if (tid < kRows) {
    float s = out_s[0][tid];
    for (int w = 1; w < 4; ++w) s += out_s[w][tid];
    a.out[m0 + tid] = s + a.params[a.boff[L]];
  }
}

// Backward pass of a block of kRows rows (targets in a.y, outputs in a.out):
//  - dout[m] = 2 (out[m] - y[m]) / rows and the block's squared-error partial;
//  - output layer: dz_{L-1}[m][c] = dout[m] w[c] gelu'(z_{L-1}[m][c]), with the block's
//    partials of gw[c] = sum_m h_{L-1}[m][c] dout[m], gb = sum_m dout[m] and the last
//    hidden layer's bias gradient sum_m dz_{L-1}[m][c];
//  - layers l = L-1 .. 1: dz_{l-1} = (dz_l W_l) * gelu'(z_{l-1}), an LDS-to-LDS MFMA
//    product against W_l^T, with the block's column sums of dz_{l-1} (bias partials).
// Every dz_l is stored transposed (bf16) for the weight gradients; rows past `rows`
// have zero gradient. Partial sums run over the block's rows in order.
__global__ __launch_bounds__(256) void k_mlp_bwd(MlpArgs a) {
  extern __shared__ uint16_t smem[];
  const int W = a.width, L = a.hidden, ldW = W + kLdsPad;
  uint16_t* bufA = smem;
  uint16_t* bufB = smem + kRows * ldW;
  __shared__ float dout_s[kRows];
  __shared__ double err_s[kRows];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * kRows, blk = blockIdx.x;
  // the first weight chunk of the top hidden product is in flight meanwhile
  const int rot = blockIdx.x >> 3;  // the row blocks on one XCD start their k-steps apart
  BRows br = gemm_rows(a.wtbf[L - 1], W, W, 64 * wave, rot);
  BChunk pre;
  if (L > 1) gemm_first(pre, br);
  if (tid < kRows) {
    const int m = m0 + tid;
    float d = 0.f;
    if (m < a.rows) d = a.out[m] - a.y[m];
    dout_s[tid] = m < a.rows ? 2.f * d / float(a.rows) : 0.f;
    err_s[tid] = double(d) * double(d);
  }
  __syncthreads();
  if (tid == 0) {
    double e = 0.0;
    float gb = 0.f;
    for (int r = 0; r < kRows; ++r) {
      e += err_s[r];
      gb += dout_s[r];
    }
    a.lossp[blk] = e;
    a.gbp[blk] = gb;
    if (a.ctr && blk == 0) {
      a.ctr[0] += 1;
      a.ctr[1] += 1;
    }
  }
  // ---- output layer ----
  const float* wout = a.params + a.woff[L];
  for (int c = tid; c < W; c += 256) {
    const float wc = wout[c];
    f4 dd[kRows / 4];
#pragma unroll
    for (int q = 0; q < kRows / 4; ++q) dd[q] = *reinterpret_cast<const f4*>(a.gd[L - 1] + zoff(m0 + 4 * q, c, W));
    uint4 hv[kRows / 8];  // the forward's bf16 activations h_{L-1}, rows 8 h .. 8 h + 7
#pragma unroll
    for (int h = 0; h < kRows / 8; ++h) hv[h] = *reinterpret_cast<const uint4*>(a.hT[L - 1] + frag_off(c, m0 + 8 * h, a.Mp));
    float gw = 0.f, db = 0.f;
#pragma unroll
    for (int q = 0; q < kRows / 4; ++q) {
      uint16_t tv[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = 4 * q + e;
        const float d = dout_s[r];
        const float g = m0 + r < a.rows ? d * wc * dd[q][e] : 0.f;
        const uint32_t hw = (&hv[r >> 3].x)[(r & 7) >> 1];
        gw += bf2f(uint16_t((r & 1) ? hw >> 16 : hw & 0xFFFFu)) * d;
        db += g;
        tv[e] = f2bf(g);
        bufA[r * ldW + c] = tv[e];
      }
      store4t(a.dzT[L - 1] + frag_off(c, m0 + 4 * q, a.Mp), tv);
    }
    a.gwp[size_t(blk) * W + c] = gw;
    a.bpart[L - 1][size_t(blk) * W + c] = db;
  }
  __syncthreads();
  // ---- hidden layers, from the top down ----
  uint16_t* in = bufA;
  uint16_t* outb = bufB;
  const int r0 = (lane >> 4) * 4;
  for (int l = L - 1; l >= 1; --l) {
    for (int n0 = 64 * wave; n0 < W; n0 += 256) {
      if (n0 != 64 * wave || l < L - 1) {
        br = gemm_rows(a.wtbf[l], W, W, n0, rot);
        gemm_first(pre, br);
      }
      f4 dd[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        dd[j] = *reinterpret_cast<const f4*>(a.gd[l - 1] + zoff(m0 + r0, n0 + 16 * j + (lane & 15), W));
      f4 acc[4];
      block_gemm(in, ldW, br, W, pre, acc);
      float csum[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int n = n0 + 16 * j + (lane & 15);
        uint16_t tv[4];
        csum[j] = 0.f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float g = m0 + r0 + e < a.rows ? acc[j][e] * dd[j][e] : 0.f;
          csum[j] += g;
          tv[e] = f2bf(g);
          outb[(r0 + e) * ldW + n] = tv[e];
        }
        store4t(a.dzT[l - 1] + frag_off(n, m0 + r0, a.Mp), tv);
      }
      // column sums over the block's 16 rows: the lane's 4 rows, then the four lane
      // groups (fixed butterfly); each wave owns its 64 columns
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        csum[j] += __shfl_xor(csum[j], 16, 64);
        csum[j] += __shfl_xor(csum[j], 32, 64);
        if (lane < 16) a.bpart[l - 1][size_t(blk) * W + n0 + 16 * j + lane] = csum[j];
      }
    }
    __syncthreads();
    uint16_t* t = in;
    in = outb;
    outb = t;
  }
}

// Weight gradients of every hidden layer in one launch: dW_l [W x K_l] =
// dzT_l [W x Mp] * (h_{l-1}^T [K_l x Mp])^T (h_{-1}^T = xT; both operands in fragment
// order, so each operand load of a wave is 1 KB of contiguous memory). A workgroup
// computes a 128 x 128 tile of dW over one of S chunks of the Mp rows (its partial
// product goes to wpart_l + z W K_l, summed in order afterwards); each wave a 64 x 64
// quarter, so that every 16-byte operand load feeds four MFMAs. Workgroups are
// numbered so that those on one XCD (the hardware deals workgroup ids to the 8 XCDs
// in turn) take consecutive tiles of one (layer, chunk, row band), whose A rows then
// come from that XCD's L2.
constexpr int kWgTile = 128;
struct WgradJobs {
  int nl, S, kc, W, Mp, tiles_m;
  int tiles_k[kMaxLayers];      // K_l / 128 (rounded up)
  int k[kMaxLayers];            // K_l
  int first[kMaxLayers + 1];    // prefix of workgroups per layer: tiles_k tiles_m S
  const uint16_t* A[kMaxLayers];
  const uint16_t* B[kMaxLayers];
  float* part[kMaxLayers];
};
__global__ __launch_bounds__(256) void k_mlp_wgrad(WgradJobs J) {
  const int total = J.first[J.nl];
  int id = blockIdx.x;
  if ((total & 7) == 0) id = (id & 7) * (total >> 3) + (id >> 3);
  int l = 0;
  while (id >= J.first[l + 1]) ++l;
  int q = id - J.first[l];
  const int tk = J.tiles_k[l], tm = J.tiles_m;
  const int x = q % tk;
  q /= tk;
  const int y = q % tm, z = q / tm;
  const int K = J.k[l], N = J.W;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m0 = y * kWgTile + (wave >> 1) * 64;  // rows of dW (W dimension)
  const int n0 = x * kWgTile + (wave & 1) * 64;   // columns of dW (K dimension)
  if (m0 >= N || n0 >= K) return;                 // a quarter past the matrix (no barriers follow)
  const int kb = z * J.kc, ke = min(J.Mp, kb + J.kc);
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const uint16_t* ap[4];
  const uint16_t* bp[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) ap[i] = J.A[l] + size_t((m0 >> 4) + i) * 16 * J.Mp + 16 * kb + lane * 8;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    bp[j] = J.B[l] + size_t(min((n0 >> 4) + j, (K >> 4) - 1)) * 16 * J.Mp + 16 * kb + lane * 8;
  constexpr int kC = 2;  // steps per chunk
  struct Ops {
    bf8 a[kC][4], b[kC][4];
  };
  auto load = [&](Ops& o, int c) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < kC; ++t) {
      const int k0 = 512 * (c * kC + t);
#pragma unroll
      for (int i = 0; i < 4; ++i) o.a[t][i] = *reinterpret_cast<const bf8*>(ap[i] + k0);
#pragma unroll
      for (int j = 0; j < 4; ++j) o.b[t][j] = *reinterpret_cast<const bf8*>(bp[j] + k0);
    }
  };
  auto compute = [&](const Ops& o) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < kC; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mfma32(o.a[t][i], o.b[t][j], acc[i][j]);
  };
  const int nsteps = (ke - kb) >> 5, nfull = nsteps / kC;
  if (nfull > 0) {
    Ops o0, o1;
    load(o0, 0);
    int c = 0;
    for (; c + 2 < nfull; c += 2) {
      load(o1, c + 1);
      compute(o0);
      load(o0, c + 2);
      compute(o1);
    }
    if (c + 1 < nfull) {  // the last two chunks
      load(o1, c + 1);
      compute(o0);
      compute(o1);
    } else {
      compute(o0);
    }
  }
  for (int st = nfull * kC; st < nsteps; ++st) {
    const int k0 = 512 * st;
    bf8 a1[4], b1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a1[i] = *reinterpret_cast<const bf8*>(ap[i] + k0);
#pragma unroll
    for (int j = 0; j < 4; ++j) b1[j] = *reinterpret_cast<const bf8*>(bp[j] + k0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma32(a1[i], b1[j], acc[i][j]);
  }
  float* out = J.part[l] + size_t(z) * N * K;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + 16 * j + (lane & 15);
      if (col >= K) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) out[size_t(m0 + 16 * i + (lane >> 4) * 4 + e) * K + col] = acc[i][j][e];
    }
}

// The gradient and, in a training step, the parameter update, every job in one
// launch: grad[i] = the sum over p = 0..nparts-1 of src[p * stride + i], written to
// the flat gradient vector at the job's parameter offset. A job with few parts
// (the weight-gradient splits) gives each element one thread, which adds the parts
// in order (loaded in groups of eight, so the loads overlap); a job with many parts
// (the per-row-block partials, kWideParts or more) gives each element one wave: lane
// q adds parts q, q + 64, ... in order, then the lanes are added by a fixed
// butterfly, and lane 0 carries on. Either way the order is fixed. With
// Adam on, the same thread then updates that parameter (bias-corrected Adam, step
// count *t, which the backward pass has advanced) and, for a hidden layer's weight,
// writes its bf16 copies (W and W^T, fragment order). One extra wave adds the
// batch's squared-error partials, in a fixed order, to the epoch's loss sum.
constexpr int kWideParts = 64;
struct UpdateJobs {
  int nj;
  const float* src[2 * kMaxLayers + 2];
  int nparts[2 * kMaxLayers + 2];
  size_t stride[2 * kMaxLayers + 2];
  size_t poff[2 * kMaxLayers + 2];   // parameter offset of the job's first element
  int wl[2 * kMaxLayers + 2];        // hidden layer whose weights the job holds, or -1
  size_t count[2 * kMaxLayers + 2];  // elements of the job
  size_t first[2 * kMaxLayers + 3];  // first thread of each job (multiples of 64; 64 per element of a wide job)
  int n[kMaxLayers], k[kMaxLayers];  // W_l is [n x k]
  uint16_t* wb[kMaxLayers];
  uint16_t* wtb[kMaxLayers];
  float* grads;
  // Adam, when params is not null
  float* params;
  float* am;
  float* av;
  const int64_t* t;
  float lr, b1, b2, eps;
  // loss, when lossp is not null: eacc[0] += (sum of lossp[0 .. nloss), in order) / rows
  const double* lossp;
  int nloss, rows;
  double* eacc;
};
__global__ __launch_bounds__(256) void k_mlp_update(UpdateJobs J) {
  const size_t total = J.first[J.nj];
  if (size_t(blockIdx.x) * 256 >= total) {  // the extra block: the loss sum, one wave
    if (threadIdx.x < 64) {
      double s = 0.0;
      for (int i = threadIdx.x; i < J.nloss; i += 64) s += J.lossp[i];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (threadIdx.x == 0) J.eacc[0] += s / J.rows;
    }
    return;
  }
  const size_t e = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (e >= total) return;
  int j = 0;
  while (e >= J.first[j + 1]) ++j;
  const size_t st = J.stride[j];
  const int np = J.nparts[j];
  size_t i = e - J.first[j];
  float acc = 0.f;
  if (np >= kWideParts) {  // one wave per element (the wave is uniform in j)
    const int q = int(i & 63);
    i >>= 6;
    const float* s = J.src[j] + i;
    int p = q;
    for (; p + 64 * 3 < np; p += 64 * 4) {
      float v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) v[k] = s[size_t(p + 64 * k) * st];
#pragma unroll
      for (int k = 0; k < 4; ++k) acc = (p + 64 * k == q) ? v[k] : acc + v[k];
    }
    for (; p < np; p += 64) acc = (p == q) ? s[size_t(p) * st] : acc + s[size_t(p) * st];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (q != 0) return;
  } else {
    if (i >= J.count[j]) return;
    const float* s = J.src[j] + i;
    int p = 0;
    for (; p + 8 <= np; p += 8) {
      float v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = s[size_t(p + k) * st];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc = (p + k == 0) ? v[k] : acc + v[k];
    }
    for (; p < np; ++p) acc = (p == 0) ? s[0] : acc + s[size_t(p) * st];
  }
  const size_t pi = J.poff[j] + i;
  J.grads[pi] = acc;
  if (!J.params) return;
  const float tf = float(*J.t);
  const float c1 = 1.f - powf(J.b1, tf), c2 = 1.f - powf(J.b2, tf);
  const float me = J.b1 * J.am[pi] + (1.f - J.b1) * acc;
  const float ve = J.b2 * J.av[pi] + (1.f - J.b2) * acc * acc;
  J.am[pi] = me;
  J.av[pi] = ve;
  const float pn = J.params[pi] - J.lr * (me / c1) / (sqrtf(ve / c2) + J.eps);
  J.params[pi] = pn;
  const int l = J.wl[j];
  if (l >= 0) {
    const int N = J.n[l], K = J.k[l];
    const int nn = int(i / K), kk = int(i - size_t(nn) * K);
    const uint16_t b = f2bf(pn);
    J.wb[l][frag_off(nn, kk, K)] = b;
    J.wtb[l][frag_off(kk, nn, N)] = b;
  }
}

// W (fp32 [N x K]) -> bf16 W and bf16 W^T (fragment order) for every hidden layer in
// one launch (after the parameters are set from the host)
struct WLayers {
  int nl;
  int n[kMaxLayers], k[kMaxLayers];
  size_t woff[kMaxLayers];
  uint16_t* wb[kMaxLayers];
  uint16_t* wtb[kMaxLayers];
  size_t first[kMaxLayers + 1];  // prefix of n k: element e belongs to the layer l with first[l] <= e < first[l + 1]
};
__global__ __launch_bounds__(256) void k_weights_bf16(const float* __restrict__ params, WLayers L) {
  const size_t e = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (e >= L.first[L.nl]) return;
  int l = 0;
  while (e >= L.first[l + 1]) ++l;
  const size_t r = e - L.first[l];
  const int N = L.n[l], K = L.k[l];
  const int nn = int(r / K), kk = int(r % K);
  const uint16_t b = f2bf(params[L.woff[l] + r]);
  L.wb[l][frag_off(nn, kk, K)] = b;
  L.wtb[l][frag_off(kk, nn, N)] = b;
}

// start of an epoch: batch index 0, loss sum 0 (Adam's step count is kept)
__global__ void k_reset_epoch(int64_t* counters, double* eacc) {
  if (threadIdx.x == 0) {
    counters[0] = 0;
    eacc[0] = 0.0;
  }
}

inline unsigned blocks(size_t n) { return unsigned((n + 255) / 256); }

constexpr int kMaxSplit = 8;  // K-splits of a weight-gradient GEMM

}  // namespace

}  // namespace yh

using namespace yh;

struct yh_mlp {
  int in_dim = 0, feat = 0, trig = 0, width = 0, hidden = 0, max_batch = 0;
  std::vector<size_t> woff, boff;  // per layer (hidden layers then the output layer)
  size_t nparams = 0;
  Workspace w;
  DevBuf params, grads, am, av;
  // per hidden layer: bf16 W and W^T, pre-activations z (fp32), the transposed bf16
  // activations and gradients, and the bias-gradient partials per row block
  std::vector<DevBuf*> wbf, wtbf, zbuf, hT, dzT, bpart;
  DevBuf x32, xT, out, y, lossp, eacc, data_x, data_y, perm;
  DevBuf wsplit;    // split partial weight gradients of every layer
  DevBuf gwp, gbp;  // output-layer gradient partials per row block
  DevBuf ctr;       // int64 [batch index within the epoch, Adam steps taken]
  PinBuf pin_x, pin_y;  // prediction staging: FP32 input rows and outputs, page-locked
  // the training step captured as a graph, for one batch size, rate and set of buffers
  hipGraphExec_t step_graph = nullptr;
  int graph_batch = 0;
  float graph_lr = 0.f;
  const void* graph_key[3] = {nullptr, nullptr, nullptr};
  ~yh_mlp() {
    if (step_graph) (void)hipGraphExecDestroy(step_graph);
    for (auto* v : {&wbf, &wtbf, &zbuf, &hT, &dzT, &bpart})
      for (DevBuf* b : *v) delete b;
  }
  int fan_in(int l) const { return l == 0 ? feat : width; }
  int fan_out(int l) const { return l == hidden ? 1 : width; }
  size_t fwd_lds() const { return size_t(kRows) * (std::max(feat, width) + kLdsPad + width + kLdsPad) * 2; }
  size_t bwd_lds() const { return size_t(2) * kRows * (width + kLdsPad) * 2; }
};

namespace yh {

static float* fptr(DevBuf& b, size_t off = 0) { return static_cast<float*>(b.p) + off; }
static uint16_t* hptr(DevBuf& b) { return static_cast<uint16_t*>(b.p); }
static int64_t* cptr(yh_mlp* m) { return static_cast<int64_t*>(m->ctr.p); }
static int pad_rows(int rows) { return (rows + 63) / 64 * 64; }

// splits of a weight-gradient reduction over Mp rows: chunks of at least 64 rows
static int split_count(int Mp) { return std::max(1, std::min(kMaxSplit, Mp / 64)); }

static MlpArgs make_args(yh_mlp* m, int rows, const float* x, const int32_t* idx, const int64_t* step,
                         const float* ysrc) {
  MlpArgs a{};
  a.in_dim = m->in_dim;
  a.feat = m->feat;
  a.trig = m->trig;
  a.width = m->width;
  a.hidden = m->hidden;
  a.rows = rows;
  a.Mp = pad_rows(rows);
  a.x = x;
  a.idx = idx;
  a.step = step;
  a.ysrc = ysrc;
  a.y = fptr(m->y);
  a.params = fptr(m->params);
  for (int l = 0; l <= m->hidden; ++l) {
    a.woff[l] = m->woff[l];
    a.boff[l] = m->boff[l];
  }
  for (int l = 0; l < m->hidden; ++l) {
    a.wbf[l] = hptr(*m->wbf[l]);
    a.wtbf[l] = hptr(*m->wtbf[l]);
    a.gd[l] = fptr(*m->zbuf[l]);
    a.hT[l] = hptr(*m->hT[l]);
    a.dzT[l] = hptr(*m->dzT[l]);
    a.bpart[l] = fptr(*m->bpart[l]);
  }
  a.xT = hptr(m->xT);
  a.out = fptr(m->out);
  a.gwp = fptr(m->gwp);
  a.gbp = fptr(m->gbp);
  a.lossp = static_cast<double*>(m->lossp.p);
  return a;
}

// forward over `rows` rows of x (gathered through idx / *step when given); output in m->out
static int mlp_forward(yh_mlp* m, const MlpArgs& a) {
  hipLaunchKernelGGL(k_mlp_fwd, dim3(a.Mp / kRows), dim3(256), m->fwd_lds(), m->w.stream, a);
  return YH_OK;
}

// backward of the last forward (targets in m->y); every gradient in m->grads
struct AdamStep {
  float lr;
};

// backward pass and weight gradients; with adam, also the update and the loss sum
static int mlp_backward(yh_mlp* m, const MlpArgs& a, const AdamStep* adam) {
  hipStream_t st = m->w.stream;
  const int Mp = a.Mp, T = Mp / kRows, W = m->width, L = m->hidden;
  hipLaunchKernelGGL(k_mlp_bwd, dim3(T), dim3(256), m->bwd_lds(), st, a);
  // weight gradients of every layer, split over the rows
  const int S = split_count(Mp), kc = (Mp / S + 31) / 32 * 32, nz = (Mp + kc - 1) / kc;
  WgradJobs J{};
  J.nl = L;
  J.S = nz;
  J.kc = kc;
  J.W = W;
  J.Mp = Mp;
  J.tiles_m = (W + kWgTile - 1) / kWgTile;
  size_t poff = 0;
  J.first[0] = 0;
  for (int l = 0; l < L; ++l) {
    J.k[l] = m->fan_in(l);
    J.tiles_k[l] = (J.k[l] + kWgTile - 1) / kWgTile;
    J.first[l + 1] = J.first[l] + J.tiles_k[l] * J.tiles_m * nz;
    J.A[l] = hptr(*m->dzT[l]);
    J.B[l] = l == 0 ? hptr(m->xT) : hptr(*m->hT[l - 1]);
    J.part[l] = fptr(m->wsplit, poff);
    poff += size_t(nz) * W * J.k[l];
  }
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(J.first[L]), dim3(256), 0, st, J);
  // every partial sum, in part order (and, in a training step, the update)
  UpdateJobs U{};
  size_t acc = 0;
  auto job = [&](const float* src, int np, size_t stride, size_t poff, size_t n, int wl) {
    U.src[U.nj] = src;
    U.nparts[U.nj] = np;
    U.stride[U.nj] = stride;
    U.poff[U.nj] = poff;
    U.wl[U.nj] = wl;
    U.count[U.nj] = n;
    U.first[U.nj] = acc;
    acc += (np >= kWideParts ? n * 64 : (n + 63) / 64 * 64);
    ++U.nj;
  };
  for (int l = 0; l < L; ++l) {
    const size_t nk = size_t(W) * J.k[l];
    job(J.part[l], nz, nk, m->woff[l], nk, l);
    job(fptr(*m->bpart[l]), T, size_t(W), m->boff[l], size_t(W), -1);
    U.n[l] = W;
    U.k[l] = J.k[l];
    U.wb[l] = hptr(*m->wbf[l]);
    U.wtb[l] = hptr(*m->wtbf[l]);
  }
  job(fptr(m->gwp), T, size_t(W), m->woff[L], size_t(W), -1);
  job(fptr(m->gbp), T, size_t(1), m->boff[L], size_t(1), -1);
  U.first[U.nj] = acc;
  U.grads = fptr(m->grads);
  unsigned nblk = blocks(acc);
  if (adam) {
    U.params = fptr(m->params);
    U.am = fptr(m->am);
    U.av = fptr(m->av);
    U.t = cptr(m) + 1;
    U.lr = adam->lr;
    U.b1 = 0.9f;
    U.b2 = 0.999f;
    U.eps = 1e-8f;
    U.lossp = static_cast<const double*>(m->lossp.p);
    U.nloss = T;
    U.rows = a.rows;
    U.eacc = static_cast<double*>(m->eacc.p);
    ++nblk;
  }
  hipLaunchKernelGGL(k_mlp_update, dim3(nblk), dim3(256), 0, st, U);
  return YH_OK;
}

// bf16 copies of every hidden layer's weights; counters (may be null) advance by one
static int mlp_refresh_bf16(yh_mlp* m) {
  WLayers L{};
  L.nl = m->hidden;
  size_t acc = 0;
  for (int l = 0; l < m->hidden; ++l) {
    L.n[l] = m->width;
    L.k[l] = m->fan_in(l);
    L.woff[l] = m->woff[l];
    L.wb[l] = hptr(*m->wbf[l]);
    L.wtb[l] = hptr(*m->wtbf[l]);
    L.first[l] = acc;
    acc += size_t(L.n[l]) * L.k[l];
  }
  L.first[m->hidden] = acc;
  hipLaunchKernelGGL(k_weights_bf16, dim3(blocks(acc)), dim3(256), 0, m->w.stream, fptr(m->params), L);
  return YH_OK;
}

static int mlp_alloc(yh_mlp* m) {
  const size_t B = size_t(m->max_batch), W = size_t(m->width), T = B / kRows;
  YH_TRY(m->w.init());
  YH_TRY(m->params.ensure(m->nparams * 4));
  YH_TRY(m->grads.ensure(m->nparams * 4));
  YH_TRY(m->am.ensure(m->nparams * 4));
  YH_TRY(m->av.ensure(m->nparams * 4));
  size_t wk = 0;
  for (int l = 0; l < m->hidden; ++l) {
    const size_t K = size_t(m->fan_in(l));
    wk += K;
    for (auto* v : {&m->wbf, &m->wtbf, &m->zbuf, &m->hT, &m->dzT, &m->bpart}) v->push_back(new DevBuf());
    YH_TRY(m->wbf[l]->ensure(W * K * 2));
    YH_TRY(m->wtbf[l]->ensure(W * K * 2));
    YH_TRY(m->zbuf[l]->ensure(B * W * 4));
    YH_TRY(m->hT[l]->ensure(B * W * 2));
    YH_TRY(m->dzT[l]->ensure(B * W * 2));
    YH_TRY(m->bpart[l]->ensure(T * W * 4));
  }
  YH_TRY(m->x32.ensure(B * m->in_dim * 4));
  YH_TRY(m->xT.ensure(B * m->feat * 2));
  YH_TRY(m->wsplit.ensure(size_t(kMaxSplit) * W * wk * 4));
  YH_TRY(m->gwp.ensure(T * W * 4));
  YH_TRY(m->gbp.ensure(T * 4));
  YH_TRY(m->out.ensure(B * 4));
  YH_TRY(m->y.ensure(B * 4));
  YH_TRY(m->lossp.ensure(T * 8));
  YH_TRY(m->eacc.ensure(8));
  YH_TRY(m->ctr.ensure(2 * sizeof(int64_t)));
  YH_HIP(hipMemsetAsync(m->ctr.p, 0, 2 * sizeof(int64_t), m->w.stream));
  // the row-block kernels keep their activations in dynamic LDS
  YH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp_fwd), hipFuncAttributeMaxDynamicSharedMemorySize,
                             int(m->fwd_lds())));
  YH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp_bwd), hipFuncAttributeMaxDynamicSharedMemorySize,
                             int(m->bwd_lds())));
  return m->w.sync();
}

static double loss_sum(yh_mlp* m, int rows) {
  std::vector<double> p(size_t(pad_rows(rows) / kRows));
  if (m->w.fetch(p.data(), m->lossp.p, p.size() * 8) != YH_OK) return -1.0;
  double s = 0.0;
  for (double v : p) s += v;
  return s / rows;
}

// One Adam training step on a batch of `batch` rows gathered from the device-resident
// data set (rows perm[ctr[0] * batch ..]); the counters advance at its end.
static int mlp_train_step(yh_mlp* m, int batch, float lr) {
  MlpArgs a = make_args(m, batch, fptr(m->data_x), static_cast<const int32_t*>(m->perm.p), cptr(m),
                        fptr(m->data_y));
  a.ctr = cptr(m);
  YH_TRY(mlp_forward(m, a));
  const AdamStep adam{lr};
  return mlp_backward(m, a, &adam);
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
  const size_t lds = size_t(kRows) * (std::max(feat, width) + kLdsPad + width + kLdsPad) * 2;
  if (lds > 160 * 1024)
    return fail(YH_ERR_UNSUPPORTED, "layer widths too large: a block of 32 activation rows must fit in LDS");
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
  YH_HIP(hipMemsetAsync(m->ctr.p, 0, 2 * sizeof(int64_t), m->w.stream));  // Adam restarts
  YH_TRY(mlp_refresh_bf16(m));
  return m->w.sync();
}

}  // extern "C"

namespace yh {

// Predictions for n rows of input (FP32 or FP64, converted to FP32): the rows are
// copied into page-locked staging by several host threads (the copy, not the GPU,
// bounds a call: one thread moves about 13 GB/s), a part at a time, and each part's
// chunks of max_batch rows are queued (upload, forward, download of the outputs)
// while the threads fill the next part; one synchronisation at the end.
template <class T>
static int mlp_predict_rows(yh_mlp* m, const T* x, size_t n, float* y) {
  if (!m || !x || !y) return fail(YH_ERR_INVALID_ARG, "null argument");
  if (n == 0) return YH_OK;
  const size_t d = size_t(m->in_dim), mb = size_t(m->max_batch);
  YH_TRY(m->x32.ensure(mb * d * sizeof(float)));
  YH_HIP(hipStreamSynchronize(m->w.stream));  // the staging may still feed an earlier call
  YH_TRY(m->pin_x.ensure(n * d * sizeof(float)));
  YH_TRY(m->pin_y.ensure(n * sizeof(float)));
  float* px = static_cast<float*>(m->pin_x.p);
  float* py = static_cast<float*>(m->pin_y.p);
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t nth = std::max<size_t>(1, std::min<size_t>(8, hw ? hw : 1));
  const size_t part = std::max<size_t>(1, (n + 3) / 4 + mb - 1) / mb * mb;  // about a quarter, whole chunks
  for (size_t p0 = 0; p0 < n; p0 += part) {
    const size_t p1 = std::min(n, p0 + part);
    auto fill = [&](size_t t) {
      const size_t r0 = p0 + (p1 - p0) * t / nth, r1 = p0 + (p1 - p0) * (t + 1) / nth;
      for (size_t i = r0 * d; i < r1 * d; ++i) px[i] = float(x[i]);
    };
    std::vector<std::thread> th;
    for (size_t t = 1; t < nth; ++t) th.emplace_back(fill, t);
    fill(0);
    for (auto& t : th) t.join();
    for (size_t r0 = p0; r0 < p1; r0 += mb) {
      const int rows = int(std::min(mb, p1 - r0));
      YH_HIP(hipMemcpyAsync(m->x32.p, px + r0 * d, size_t(rows) * d * sizeof(float), hipMemcpyHostToDevice,
                            m->w.stream));
      YH_TRY(mlp_forward(m, make_args(m, rows, fptr(m->x32), nullptr, nullptr, nullptr)));
      YH_HIP(hipMemcpyAsync(py + r0, m->out.p, size_t(rows) * sizeof(float), hipMemcpyDeviceToHost, m->w.stream));
    }
  }
  YH_HIP(hipStreamSynchronize(m->w.stream));
  std::memcpy(y, py, n * sizeof(float));
  return YH_OK;
}

}  // namespace yh

extern "C" {

int yh_mlp_predict(yh_mlp* m, const float* x, size_t n, float* y) { return mlp_predict_rows(m, x, n, y); }

int yh_mlp_predict_f64(yh_mlp* m, const double* x, size_t n, float* y) { return mlp_predict_rows(m, x, n, y); }

int yh_mlp_loss_grad(yh_mlp* m, const float* x, const float* y, int n, float* loss, float* grads) {
  if (!m || !x || !y || !loss) return fail(YH_ERR_INVALID_ARG, "null argument");
  if (n < 1 || n > m->max_batch) return fail(YH_ERR_INVALID_ARG, "batch larger than max_batch");
  YH_TRY(m->w.upload(x, size_t(n) * m->in_dim * 4, m->x32));
  YH_TRY(m->w.upload(y, size_t(n) * 4, m->y));
  const MlpArgs a = make_args(m, n, fptr(m->x32), nullptr, nullptr, nullptr);
  YH_TRY(mlp_forward(m, a));
  YH_TRY(mlp_backward(m, a, nullptr));
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
  YH_TRY(m->perm.ensure(n * 4));
  // One training step is captured once as a graph (for this batch size and these
  // buffers) and replayed for every batch: its launches read the batch index and
  // Adam's step count from device counters, which the step's last kernel advances.
  const void* key[3] = {m->data_x.p, m->data_y.p, m->perm.p};
  if (!m->step_graph || m->graph_batch != batch || std::memcmp(key, m->graph_key, sizeof(key)) != 0 ||
      m->graph_lr != lr) {
    if (m->step_graph) (void)hipGraphExecDestroy(m->step_graph);
    m->step_graph = nullptr;
    YH_TRY(m->w.sync());
    hipGraph_t g = nullptr;
    YH_HIP(hipStreamBeginCapture(m->w.stream, hipStreamCaptureModeThreadLocal));
    const int st = mlp_train_step(m, batch, lr);
    const hipError_t ec = hipStreamEndCapture(m->w.stream, &g);
    if (st != YH_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (ec != hipSuccess) return fail(YH_ERR_HIP, std::string("graph capture failed: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&m->step_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      m->step_graph = nullptr;
      return fail(YH_ERR_HIP, std::string("graph instantiation failed: ") + hipGetErrorString(ei));
    }
    m->graph_batch = batch;
    m->graph_lr = lr;
    std::memcpy(m->graph_key, key, sizeof(key));
  }
  std::mt19937_64 rng(seed);
  std::vector<int32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  const size_t nb = n / size_t(batch);
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(perm.begin(), perm.end(), rng);
    YH_TRY(m->w.upload(perm.data(), n * 4, m->perm));
    // batch index back to 0 (Adam's count carries on); the epoch's loss sum to 0
    hipLaunchKernelGGL(k_reset_epoch, dim3(1), dim3(64), 0, m->w.stream, cptr(m), static_cast<double*>(m->eacc.p));
    for (size_t bi = 0; bi < nb; ++bi) YH_HIP(hipGraphLaunch(m->step_graph, m->w.stream));
    double acc = 0.0;
    YH_TRY(m->w.fetch(&acc, m->eacc.p, 8));
    acc /= double(nb);
    if (epoch_loss) epoch_loss[ep] = float(acc);
  }
  return YH_OK;
}

}  // extern "C"
