// sv_kernels.hip — per-gate state-vector kernels for gfx950.
//
// One launch applies one gate element the way yao-rs's dispatch_arrayreg_gate
// (src/apply.rs:7-120) routes it to an instruct (src/instruct_qubit.rs):
//   k_1q         instruct_1q            dense 2x2, no controls: amplitude-pair stream
//   k_dense<NT>  instruct_{1q,2q,nq}[_controlled]: 2^NT-amplitude groups, targets
//                and controls locked by zero-bit insertion from the thread id
//   k_diag<NT>   instruct_{1q,2q}_diag[_controlled]: element-wise phase multiply
//   k_xflip      instruct_x[_controlled]: amplitude-pair swap
//   k_swap       instruct_swap
//   k_dense_oop  instruct_nq for more than 5 targets (out of place)
// Each thread owns whole groups, so every amplitude is read and written once per
// gate: 2 * 16 B per amplitude for complex128 is the algorithmic traffic.
#include "kernels.h"

namespace yh {

// ---------------------------------------------------------------------------
// Dense 2x2 on one target bit, no controls. Each thread handles kPairs pairs
// spaced blockDim apart, so a wave touches 64 consecutive lower amplitudes and
// 64 consecutive upper amplitudes per pair slot (coalesced for bit >= 2; for
// lower bits the two loads of a wave still cover whole cache lines).
// ---------------------------------------------------------------------------
constexpr int kPairs = 2;

template <class R>
__global__ __launch_bounds__(kBlock) void k_1q(typename Cplx<R>::V* __restrict__ s, int bit,
                                                Mat2<R> m, uint64_t npairs) {
  using V = typename Cplx<R>::V;
  const uint64_t lowmask = (uint64_t(1) << bit) - 1;
  const uint64_t step = uint64_t(1) << bit;
  const uint64_t stride = uint64_t(gridDim.x) * (kBlock * kPairs);
  for (uint64_t g0 = uint64_t(blockIdx.x) * (kBlock * kPairs) + threadIdx.x; g0 < npairs; g0 += stride) {
    V w[kPairs], v[kPairs];
    uint64_t i0[kPairs];
#pragma unroll
    for (int u = 0; u < kPairs; ++u) {
      uint64_t g = g0 + uint64_t(u) * kBlock;
      i0[u] = ((g >> bit) << (bit + 1)) | (g & lowmask);
      if (g < npairs) {
        w[u] = s[i0[u]];
        v[u] = s[i0[u] + step];
      }
    }
#pragma unroll
    for (int u = 0; u < kPairs; ++u) {
      uint64_t g = g0 + uint64_t(u) * kBlock;
      if (g < npairs) {
        s[i0[u]] = cmul2(m.a, w[u], m.b, v[u]);
        s[i0[u] + step] = cmul2(m.c, w[u], m.d, v[u]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Target bits 0..5: the partner of amplitude i, i ^ (1 << bit), sits in the same
// aligned run of 64 amplitudes, so each lane loads one amplitude of a contiguous
// run and takes its partner's from lane (lane ^ (1 << bit)) by a cross-lane
// exchange inside the wave (no memory traffic). Every load and store of a wave is
// one contiguous 1 KB run (complex128). A lane with the target bit clear computes
// a w + b v, one with it set c w + d v (w the amplitude with the bit clear, v the
// one with it set): the same operations in the same order as k_1q, so the results
// are the same bit for bit. The state has at least 64 amplitudes, a multiple of 64,
// so a wave is either wholly inside it or wholly outside.
// ---------------------------------------------------------------------------
constexpr int kWaveRuns = 2;  // 64-amplitude runs per wave per iteration

template <class R>
__global__ __launch_bounds__(kBlock) void k_1q_wave(typename Cplx<R>::V* __restrict__ s, int bit, Mat2<R> m,
                                                     uint64_t n) {
  using V = typename Cplx<R>::V;
  const int lane = threadIdx.x & 63;
  const int x = 1 << bit;
  const bool hi = (lane & x) != 0;
  const V c0 = hi ? m.c : m.a, c1 = hi ? m.d : m.b;
  const uint64_t stride = uint64_t(gridDim.x) * (kBlock * kWaveRuns);
  for (uint64_t g0 = uint64_t(blockIdx.x) * (kBlock * kWaveRuns) + threadIdx.x; g0 < n; g0 += stride) {
    V a[kWaveRuns];
#pragma unroll
    for (int u = 0; u < kWaveRuns; ++u) {
      const uint64_t g = g0 + uint64_t(u) * kBlock;
      if (g < n) a[u] = s[g];
    }
#pragma unroll
    for (int u = 0; u < kWaveRuns; ++u) {
      const uint64_t g = g0 + uint64_t(u) * kBlock;
      if (g < n) {  // uniform over the wave
        V p;
        p.x = __shfl_xor(a[u].x, x, 64);
        p.y = __shfl_xor(a[u].y, x, 64);
        const V w = hi ? p : a[u], v = hi ? a[u] : p;
        s[g] = cmul2(c0, w, c1, v);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Dense 2^NT x 2^NT on NT target bits with optional controls. Locked bits (targets
// and controls) are inserted as zeros; control bits with value 1 are then set.
// The matrix is row-major with gate-index bit k <-> offs[1<<k] (yao order).
// ---------------------------------------------------------------------------
template <class R, int NT>
__global__ __launch_bounds__(kBlock) void k_dense(typename Cplx<R>::V* __restrict__ s,
                                                  GroupArgs ga,
                                                  const typename Cplx<R>::V* __restrict__ mat) {
  using V = typename Cplx<R>::V;
  constexpr int D = 1 << NT;
  for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < ga.ngroups; g += uint64_t(gridDim.x) * kBlock) {
    uint64_t base = insert_zeros(g, ga.lk.pos, ga.lk.npos) | ga.ctrl_set;
    V in[D];
#pragma unroll
    for (int j = 0; j < D; ++j) in[j] = s[base + ga.offs[j]];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      V acc = cmul(mat[r * D], in[0]);
#pragma unroll
      for (int c = 1; c < D; ++c) acc = cadd(acc, cmul(mat[r * D + c], in[c]));
      s[base + ga.offs[r]] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Diagonal gate on NT (1 or 2) target bits. Only control bits are locked, so
// consecutive threads touch consecutive amplitudes whatever the target bits.
// ---------------------------------------------------------------------------
template <class R, int NT>
__global__ __launch_bounds__(kBlock) void k_diag(typename Cplx<R>::V* __restrict__ s, GroupArgs ga,
                                                 DiagArgs<R> d) {
  for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < ga.ngroups; g += uint64_t(gridDim.x) * kBlock) {
    uint64_t i = insert_zeros(g, ga.lk.pos, ga.lk.npos) | ga.ctrl_set;
    int j = int((i >> d.tbit[0]) & 1);
    if (NT == 2) j |= int((i >> d.tbit[1]) & 1) << 1;
    s[i] = cmul(s[i], d.d[j]);
  }
}

// ---------------------------------------------------------------------------
// X on one target bit with optional controls: swap the pair (i, i | 2^bit).
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(kBlock) void k_xflip(typename Cplx<R>::V* __restrict__ s, GroupArgs ga,
                                                  uint64_t tmask) {
  using V = typename Cplx<R>::V;
  for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < ga.ngroups; g += uint64_t(gridDim.x) * kBlock) {
    uint64_t i = insert_zeros(g, ga.lk.pos, ga.lk.npos) | ga.ctrl_set;
    V a = s[i], b = s[i | tmask];
    s[i] = b;
    s[i | tmask] = a;
  }
}

// SWAP: exchange the amplitudes with (b0, b1) = (0, 1) and (1, 0).
template <class R>
__global__ __launch_bounds__(kBlock) void k_swap(typename Cplx<R>::V* __restrict__ s, GroupArgs ga,
                                                 uint64_t m0, uint64_t m1) {
  using V = typename Cplx<R>::V;
  for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < ga.ngroups; g += uint64_t(gridDim.x) * kBlock) {
    uint64_t i = insert_zeros(g, ga.lk.pos, ga.lk.npos) | ga.ctrl_set | m1;
    uint64_t j = (i ^ m1) | m0;
    V a = s[i], b = s[j];
    s[i] = b;
    s[j] = a;
  }
}

// ---------------------------------------------------------------------------
// Out-of-place dense gate for NT > 5: src holds a copy of the state; one thread
// per (group, output row).
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(kBlock) void k_dense_oop(typename Cplx<R>::V* __restrict__ dst,
                                                      const typename Cplx<R>::V* __restrict__ src,
                                                      GroupArgs ga, int nt,
                                                      const typename Cplx<R>::V* __restrict__ mat,
                                                      const uint64_t* __restrict__ offs) {
  using V = typename Cplx<R>::V;
  const uint64_t D = uint64_t(1) << nt;
  for (uint64_t t = uint64_t(blockIdx.x) * kBlock + threadIdx.x; t < ga.ngroups * D; t += uint64_t(gridDim.x) * kBlock) {
    uint64_t g = t >> nt, r = t & (D - 1);
    uint64_t base = insert_zeros(g, ga.lk.pos, ga.lk.npos) | ga.ctrl_set;
    V acc = cmk<V>(0, 0);
    for (uint64_t c = 0; c < D; ++c) acc = cadd(acc, cmul(mat[r * D + c], src[base + offs[c]]));
    dst[base + offs[r]] = acc;
  }
}

// ============================== launchers ==================================
static inline unsigned grid_for(uint64_t work, uint64_t per_block) { return capped_grid(work, per_block); }

template <class R>
int launch_1q(typename Cplx<R>::V* s, int nbits, int bit, const Mat2<R>& m, hipStream_t st) {
  if (bit < 6 && nbits >= 6) {
    const uint64_t n = uint64_t(1) << nbits;
    hipLaunchKernelGGL(k_1q_wave<R>, dim3(grid_for(n, kBlock * kWaveRuns)), dim3(kBlock), 0, st, s, bit, m, n);
    return YH_OK;
  }
  uint64_t npairs = uint64_t(1) << (nbits - 1);
  hipLaunchKernelGGL(k_1q<R>, dim3(grid_for(npairs, kBlock * kPairs)), dim3(kBlock), 0, st, s, bit, m,
                     npairs);
  return YH_OK;
}

template <class R>
int launch_dense(typename Cplx<R>::V* s, const GroupArgs& ga, int nt,
                 const typename Cplx<R>::V* mat, hipStream_t st) {
  dim3 grid(grid_for(ga.ngroups, kBlock)), block(kBlock);
  switch (nt) {
    case 1: hipLaunchKernelGGL((k_dense<R, 1>), grid, block, 0, st, s, ga, mat); break;
    case 2: hipLaunchKernelGGL((k_dense<R, 2>), grid, block, 0, st, s, ga, mat); break;
    case 3: hipLaunchKernelGGL((k_dense<R, 3>), grid, block, 0, st, s, ga, mat); break;
    case 4: hipLaunchKernelGGL((k_dense<R, 4>), grid, block, 0, st, s, ga, mat); break;
    case 5: hipLaunchKernelGGL((k_dense<R, 5>), grid, block, 0, st, s, ga, mat); break;
    default: return fail(YH_ERR_INVALID_ARG, "launch_dense: nt out of range");
  }
  return YH_OK;
}

template <class R>
int launch_diag(typename Cplx<R>::V* s, const GroupArgs& ga, int nt, const DiagArgs<R>& d,
                hipStream_t st) {
  dim3 grid(grid_for(ga.ngroups, kBlock)), block(kBlock);
  if (nt == 1) hipLaunchKernelGGL((k_diag<R, 1>), grid, block, 0, st, s, ga, d);
  else hipLaunchKernelGGL((k_diag<R, 2>), grid, block, 0, st, s, ga, d);
  return YH_OK;
}

template <class R>
int launch_xflip(typename Cplx<R>::V* s, const GroupArgs& ga, uint64_t tmask, hipStream_t st) {
  hipLaunchKernelGGL(k_xflip<R>, dim3(grid_for(ga.ngroups, kBlock)), dim3(kBlock), 0, st, s, ga, tmask);
  return YH_OK;
}

template <class R>
int launch_swap(typename Cplx<R>::V* s, const GroupArgs& ga, uint64_t m0, uint64_t m1, hipStream_t st) {
  hipLaunchKernelGGL(k_swap<R>, dim3(grid_for(ga.ngroups, kBlock)), dim3(kBlock), 0, st, s, ga, m0, m1);
  return YH_OK;
}

template <class R>
int launch_dense_oop(typename Cplx<R>::V* dst, const typename Cplx<R>::V* src, const GroupArgs& ga,
                     int nt, const typename Cplx<R>::V* mat, const uint64_t* offs, hipStream_t st) {
  uint64_t work = ga.ngroups << nt;
  hipLaunchKernelGGL(k_dense_oop<R>, dim3(grid_for(work, kBlock)), dim3(kBlock), 0, st, dst, src, ga, nt,
                     mat, offs);
  return YH_OK;
}

#define YH_INST(R)                                                                                  \
  template int launch_1q<R>(Cplx<R>::V*, int, int, const Mat2<R>&, hipStream_t);                     \
  template int launch_dense<R>(Cplx<R>::V*, const GroupArgs&, int, const Cplx<R>::V*, hipStream_t);  \
  template int launch_diag<R>(Cplx<R>::V*, const GroupArgs&, int, const DiagArgs<R>&, hipStream_t);  \
  template int launch_xflip<R>(Cplx<R>::V*, const GroupArgs&, uint64_t, hipStream_t);                \
  template int launch_swap<R>(Cplx<R>::V*, const GroupArgs&, uint64_t, uint64_t, hipStream_t);      \
  template int launch_dense_oop<R>(Cplx<R>::V*, const Cplx<R>::V*, const GroupArgs&, int,           \
                                   const Cplx<R>::V*, const uint64_t*, hipStream_t);
YH_INST(double)
YH_INST(float)
#undef YH_INST

}  // namespace yh
