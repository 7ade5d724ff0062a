// dm_kernels.hip — density-matrix kernels for gfx950.
//
// rho is stored as in yao-rs (src/density_matrix.rs:9-12): rho[row * 2^n + col],
// i.e. a 2n-bit vector whose high n bits are the row index. For a gate or channel on
// K <= 2 qubits, one thread owns the 2^K x 2^K block of entries that share all bits
// except the K row bits and K column bits, and applies a short list of micro-ops to
// it in registers:
//   gate    : U on the row bits, then conj(U) on the column bits
//             (DensityMatrix::apply_gate, density_matrix.rs:161-178), with the row-
//             and column-side control predicates of the reference's two dispatches;
//   channel : sum_k K_k block K_k^dagger, accumulated in Kraus order
//             (DensityMatrix::apply_channel, density_matrix.rs:181-224).
// The reference spends two full passes per gate and (copy + 2 passes + accumulate)
// per Kraus operator; here every element list is one read and one write of rho.
#include <algorithm>

#include "kernels.h"

namespace yh {

namespace {

template <int B>
struct Blk {
  double2 a[B][B];
};

// new a[r'][c] = sum_r M[r'][r] a[r][c] over the rows spanned by block qubit mask QM
// (QM = B-1: all block qubits, M is B x B; QM = 1<<q: one qubit, M is 2 x 2).
template <int B, int QM>
__device__ __forceinline__ void row_apply(Blk<B>& x, const double2* __restrict__ M) {
  constexpr int D = (QM == B - 1) ? B : 2;
  double2 m[D * D];
#pragma unroll
  for (int e = 0; e < D * D; ++e) m[e] = M[e];
  if constexpr (QM == B - 1) {
#pragma unroll
    for (int c = 0; c < B; ++c) {
      double2 in[B];
#pragma unroll
      for (int r = 0; r < B; ++r) in[r] = x.a[r][c];
#pragma unroll
      for (int r = 0; r < B; ++r) {
        double2 acc = cmul(m[r * B], in[0]);
#pragma unroll
        for (int k = 1; k < B; ++k) acc = cadd(acc, cmul(m[r * B + k], in[k]));
        x.a[r][c] = acc;
      }
    }
  } else {
#pragma unroll
    for (int c = 0; c < B; ++c)
#pragma unroll
      for (int r = 0; r < B; ++r)
        if ((r & QM) == 0) {
          double2 w = x.a[r][c], v = x.a[r | QM][c];
          x.a[r][c] = cmul2(m[0], w, m[1], v);
          x.a[r | QM][c] = cmul2(m[2], w, m[3], v);
        }
  }
}

// new a[r][c'] = sum_c conj(M[c'][c]) a[r][c]
template <int B, int QM>
__device__ __forceinline__ void col_apply_conj(Blk<B>& x, const double2* __restrict__ M) {
  constexpr int D = (QM == B - 1) ? B : 2;
  double2 m[D * D];
#pragma unroll
  for (int e = 0; e < D * D; ++e) {
    m[e] = M[e];
    m[e].y = -m[e].y;
  }
  if constexpr (QM == B - 1) {
#pragma unroll
    for (int r = 0; r < B; ++r) {
      double2 in[B];
#pragma unroll
      for (int c = 0; c < B; ++c) in[c] = x.a[r][c];
#pragma unroll
      for (int c = 0; c < B; ++c) {
        double2 acc = cmul(m[c * B], in[0]);
#pragma unroll
        for (int k = 1; k < B; ++k) acc = cadd(acc, cmul(m[c * B + k], in[k]));
        x.a[r][c] = acc;
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int c = 0; c < B; ++c)
        if ((c & QM) == 0) {
          double2 w = x.a[r][c], v = x.a[r][c | QM];
          x.a[r][c] = cmul2(m[0], w, m[1], v);
          x.a[r][c | QM] = cmul2(m[2], w, m[3], v);
        }
  }
}

template <int B, int QM>
__device__ __forceinline__ void channel_apply(Blk<B>& x, const double2* __restrict__ mats, int nk) {
  constexpr int D = (QM == B - 1) ? B : 2;
  Blk<B> acc;
#pragma unroll
  for (int r = 0; r < B; ++r)
#pragma unroll
    for (int c = 0; c < B; ++c) acc.a[r][c] = make_double2(0.0, 0.0);
  for (int k = 0; k < nk; ++k) {
    Blk<B> br = x;
    row_apply<B, QM>(br, mats + k * D * D);
    col_apply_conj<B, QM>(br, mats + k * D * D);
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int c = 0; c < B; ++c) acc.a[r][c] = cadd(acc.a[r][c], br.a[r][c]);
  }
  x = acc;
}

template <int B>
__device__ __forceinline__ void run_micro(Blk<B>& x, const DmMicro& mo, const double2* __restrict__ mats,
                                          bool row_ok, bool col_ok) {
  const double2* M = mats + mo.moff;
  if (mo.type == 0) {
    if (mo.q < 0) {
      if (row_ok) row_apply<B, B - 1>(x, M);
      if (col_ok) col_apply_conj<B, B - 1>(x, M);
    } else if (mo.q == 0) {
      if (row_ok) row_apply<B, 1>(x, M);
      if (col_ok) col_apply_conj<B, 1>(x, M);
    } else if constexpr (B == 4) {
      if (row_ok) row_apply<B, 2>(x, M);
      if (col_ok) col_apply_conj<B, 2>(x, M);
    }
  } else {
    if (mo.q < 0) channel_apply<B, B - 1>(x, M, mo.nk);
    else if (mo.q == 0) channel_apply<B, 1>(x, M, mo.nk);
    else if constexpr (B == 4) channel_apply<B, 2>(x, M, mo.nk);
  }
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_dm_block(double2* __restrict__ rho, DmArgs a,
                                                     const double2* __restrict__ mats) {
  constexpr int B = 1 << K;
  for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < a.ngroups; g += uint64_t(gridDim.x) * kBlock) {
    const uint64_t base = insert_zeros(g, a.lk.pos, a.lk.npos);
    Blk<B> x;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int c = 0; c < B; ++c) x.a[r][c] = rho[base + a.roff[r] + a.coff[c]];
    const bool row_ok = (base & a.rc_mask) == a.rc_val;
    const bool col_ok = (base & a.cc_mask) == a.cc_val;
    for (int k = 0; k < a.nmicro; ++k) {
      // control predicates bind the first micro-op only (the gate)
      run_micro<B>(x, a.micro[k], mats, k == 0 ? row_ok : true, k == 0 ? col_ok : true);
    }
    if (a.nmicro > 1 || row_ok || col_ok) {
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) rho[base + a.roff[r] + a.coff[c]] = x.a[r][c];
    }
  }
}

__device__ __forceinline__ void bsum2(double& x, double& y, double* out) {
  __shared__ double sh[2][kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    x += __shfl_xor(x, off, 64);
    y += __shfl_xor(y, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = x;
    sh[1][threadIdx.x >> 6] = y;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      a += sh[0][w];
      b += sh[1][w];
    }
    out[0] = a;
    out[1] = b;
  }
}

__global__ __launch_bounds__(kBlock) void k_dm_trace(const double2* __restrict__ rho, uint64_t dim,
                                                     double* __restrict__ partial) {
  double x = 0, y = 0;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < dim; i += uint64_t(gridDim.x) * kBlock) {
    double2 v = rho[i * dim + i];
    x += v.x;
    y += v.y;
  }
  double r[2];
  bsum2(x, y, r);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = r[0];
    partial[2 * blockIdx.x + 1] = r[1];
  }
}

// purity = Re sum_{r,i} rho[r][i] rho[i][r]; tile pairs (R, I) with R <= I, the
// off-diagonal ones counted twice (rho[r][i]rho[i][r] == rho[i][r]rho[r][i] exactly).
constexpr int kTile = 32;
__global__ __launch_bounds__(kBlock) void k_dm_purity(const double2* __restrict__ rho, uint64_t dim,
                                                      uint64_t ts, uint64_t ntile, double* __restrict__ partial) {
  __shared__ double2 A[kTile][kTile + 1];
  __shared__ double2 Bt[kTile][kTile + 1];
  double x = 0, y = 0;
  for (uint64_t t = blockIdx.x; t < ntile * ntile; t += gridDim.x) {
    const uint64_t bx = t / ntile, by = t % ntile;
    if (bx > by) continue;
    for (uint64_t e = threadIdx.x; e < ts * ts; e += kBlock) {
      const uint64_t i = e / ts, j = e % ts;
      A[i][j] = rho[(bx * ts + i) * dim + by * ts + j];   // rho[R+i][I+j]
      Bt[i][j] = rho[(by * ts + i) * dim + bx * ts + j];  // rho[I+i][R+j]
    }
    __syncthreads();
    double acc = 0;
    for (uint64_t e = threadIdx.x; e < ts * ts; e += kBlock) {
      const uint64_t i = e / ts, j = e % ts;
      const double2 a = A[i][j], b = Bt[j][i];  // rho[R+i][I+j] * rho[I+j][R+i]
      acc += a.x * b.x - a.y * b.y;
    }
    x += (bx == by) ? acc : 2.0 * acc;
    __syncthreads();
  }
  double r[2];
  bsum2(x, y, r);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = r[0];
    partial[2 * blockIdx.x + 1] = 0.0;
  }
}

__global__ __launch_bounds__(kBlock) void k_fold2(const double* __restrict__ partial, int nb,
                                                  double* __restrict__ out) {
  double x = 0, y = 0;
  for (int b = threadIdx.x; b < nb; b += kBlock) {
    x += partial[2 * b];
    y += partial[2 * b + 1];
  }
  double r[2];
  bsum2(x, y, r);
  if (threadIdx.x == 0) {
    out[0] = r[0];
    out[1] = r[1];
  }
}

// rho[r][c] = psi[r] * conj(psi[c]) (DensityMatrix::from_reg, density_matrix.rs:23-35),
// unfused like the reference's products
__global__ __launch_bounds__(kBlock) void k_dm_outer(const double2* __restrict__ psi, double2* __restrict__ rho,
                                                     uint64_t dim) {
#pragma clang fp contract(off)
  for (uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x; e < dim * dim; e += uint64_t(gridDim.x) * kBlock) {
    const uint64_t r = e / dim, c = e % dim;
    const double2 a = psi[r], b = psi[c];
    // a * conj(b), each product rounded as num-complex computes it
    rho[e] = double2{a.x * b.x - a.y * (-b.y), a.x * (-b.y) + a.y * b.x};
  }
}

// partial_tr (density_matrix.rs:80-111): out[rr][rc] = sum over the traced bits t
// of rho[row(rr, t)][col(rc, t)]. Traced and kept positions are bit positions
// (the reference's loc), kept bit idx of the reduced index <- kept[idx]. One thread
// per output entry sums in increasing t, which is the reference's row order, so
// the sums are bit-identical to its loop.
__global__ __launch_bounds__(kBlock) void k_partial_tr(const double2* __restrict__ rho, uint64_t dim,
                                                       PtrArgs a, double2* __restrict__ out) {
  const uint64_t dr = uint64_t(1) << a.nkept;
  for (uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x; e < dr * dr; e += uint64_t(gridDim.x) * kBlock) {
    const uint64_t rr = e / dr, rc = e % dr;
    uint64_t rb = 0, cb = 0;
    for (int k = 0; k < a.nkept; ++k) {
      rb |= ((rr >> k) & 1) << a.kept[k];
      cb |= ((rc >> k) & 1) << a.kept[k];
    }
    const uint64_t nt = uint64_t(1) << a.ntr;
    double sr = 0.0, si = 0.0;
    for (uint64_t t = 0; t < nt; ++t) {
      uint64_t tb = 0;
      for (int k = 0; k < a.ntr; ++k) tb |= ((t >> k) & 1) << a.tr[k];
      const double2 v = rho[(rb | tb) * dim + (cb | tb)];
      sr += v.x;
      si += v.y;
    }
    out[e] = double2{sr, si};
  }
}

__global__ __launch_bounds__(kBlock) void k_add_inplace(double2* __restrict__ acc, const double2* __restrict__ b,
                                                        uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    const double2 x = acc[i], y = b[i];
    acc[i] = double2{x.x + y.x, x.y + y.y};
  }
}

}  // namespace

int launch_add_inplace(double2* acc, const double2* b, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_add_inplace, dim3(capped_grid(n, kBlock)), dim3(kBlock), 0, st, acc, b, n);
  return YH_OK;
}

int launch_partial_tr(const double2* rho, uint64_t dim, const PtrArgs& a, double2* out, hipStream_t st) {
  const uint64_t n = (uint64_t(1) << a.nkept) << a.nkept;
  hipLaunchKernelGGL(k_partial_tr, dim3(capped_grid(n, kBlock)), dim3(kBlock), 0, st, rho, dim, a, out);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int launch_dm_block(double2* rho, const DmArgs& a, int K, const double2* mats, hipStream_t st) {
  dim3 grid(capped_grid(a.ngroups, kBlock)), block(kBlock);
  if (K == 1) hipLaunchKernelGGL(k_dm_block<1>, grid, block, 0, st, rho, a, mats);
  else if (K == 2) hipLaunchKernelGGL(k_dm_block<2>, grid, block, 0, st, rho, a, mats);
  else return fail(YH_ERR_UNSUPPORTED, "density-matrix block kernel supports 1 or 2 qubits");
  return YH_OK;
}

int launch_dm_trace(const double2* rho, uint64_t dim, double* partial, double* out, hipStream_t st) {
  int nb = int(std::min<uint64_t>(kRedBlocks, (dim + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_dm_trace, dim3(nb), dim3(kBlock), 0, st, rho, dim, partial);
  hipLaunchKernelGGL(k_fold2, dim3(1), dim3(kBlock), 0, st, partial, nb, out);
  return YH_OK;
}

int launch_dm_purity(const double2* rho, uint64_t dim, double* partial, double* out, hipStream_t st) {
  uint64_t ts = dim < kTile ? dim : kTile;
  uint64_t ntile = dim / ts;
  int nb = int(std::min<uint64_t>(kRedBlocks, ntile * ntile));
  hipLaunchKernelGGL(k_dm_purity, dim3(nb), dim3(kBlock), 0, st, rho, dim, ts, ntile, partial);
  hipLaunchKernelGGL(k_fold2, dim3(1), dim3(kBlock), 0, st, partial, nb, out);
  return YH_OK;
}

int launch_dm_outer(const double2* psi, double2* rho, uint64_t dim, hipStream_t st) {
  uint64_t n = dim * dim;
  hipLaunchKernelGGL(k_dm_outer, dim3(capped_grid(n, kBlock)), dim3(kBlock), 0, st, psi, rho, dim);
  return YH_OK;
}

}  // namespace yh
