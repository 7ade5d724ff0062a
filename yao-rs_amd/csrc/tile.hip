// tile.hip — multi-gate tile passes: one HBM read and write of the state applies
// every gate that acts inside a 12-qubit tile.
//
// A workgroup loads its 2^12-amplitude tile into LDS (the 5 lowest memory bits plus
// 7 high bits chosen per pass, so each load is a run of 512 B), then walks the
// pass's gate groups: for each group of <= 4 tile qubits every thread pulls one
// 16-amplitude sub-cube into registers, applies the group's gates one by one
// (2x2 / 4x4 / diagonal, with controls local to the group), and writes it back.
// The arithmetic of each gate is the per-gate kernels' arithmetic (yao-rs's
// u1rows / unrows, src/instruct_qubit.rs:23-36, 189-206); only the HBM traffic
// is shared between gates.
#include <type_traits>
#include <utility>

#include "kernels.h"
#include "tile.h"

namespace yh {

namespace {

template <class R>
__device__ __forceinline__ typename Cplx<R>::V ldm(const double (&m)[2]) {
  return cmk<typename Cplx<R>::V>(R(m[0]), R(m[1]));
}

__device__ __forceinline__ int lpad(int p) { return p + (p >> 4); }

// A 1-qubit micro gate's control form C: -1 none, otherwise the gate's control mask
// and values tested at run time. With C = -1 every amplitude is acted on and the
// test disappears; the run-time test costs a select per register (4 per complex128
// amplitude) whether or not the gate has controls.
template <int C>
__device__ __forceinline__ bool mg_on(int i, int cm, int cv) {
  if constexpr (C < 0) return true;
  else return (i & cm) == cv;
}

template <class R, int T, int C>
__device__ __forceinline__ void mg_dense1(typename Cplx<R>::V (&a)[16], const MicroGate& g) {
  using V = typename Cplx<R>::V;
  const V m0 = ldm<R>(g.m[0]), m1 = ldm<R>(g.m[1]), m2 = ldm<R>(g.m[2]), m3 = ldm<R>(g.m[3]);
  const int cm = C == 4 ? __builtin_amdgcn_readfirstlane(g.cmask) : 0;
  const int cv = C == 4 ? __builtin_amdgcn_readfirstlane(g.cval) : 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    if (i & (1 << T)) continue;
    const int j = i | (1 << T);
    if (mg_on<C>(i, cm, cv)) {
      const V w = a[i], v = a[j];
      a[i] = cmul2(m0, w, m1, v);
      a[j] = cmul2(m2, w, m3, v);
    }
  }
}

template <class R, int T, int C>
__device__ __forceinline__ void mg_perm1(typename Cplx<R>::V (&a)[16], const MicroGate& g) {
  const int cm = C == 4 ? __builtin_amdgcn_readfirstlane(g.cmask) : 0;
  const int cv = C == 4 ? __builtin_amdgcn_readfirstlane(g.cval) : 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    if (i & (1 << T)) continue;
    const int j = i | (1 << T);
    if (mg_on<C>(i, cm, cv)) {
      const typename Cplx<R>::V w = a[i];
      a[i] = a[j];
      a[j] = w;
    }
  }
}

template <class R, int T, int C>
__device__ __forceinline__ void mg_diag1(typename Cplx<R>::V (&a)[16], const MicroGate& g) {
  using V = typename Cplx<R>::V;
  const V d0 = ldm<R>(g.m[0]), d1 = ldm<R>(g.m[1]);
  const int cm = C == 4 ? __builtin_amdgcn_readfirstlane(g.cmask) : 0;
  const int cv = C == 4 ? __builtin_amdgcn_readfirstlane(g.cval) : 0;
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (mg_on<C>(i, cm, cv)) a[i] = cmul(a[i], (i & (1 << T)) ? d1 : d0);
}

// f(std::integral_constant<int, V>) for the V in [0, N) equal to the uniform v
template <int N, class F, int... Vs>
__device__ __forceinline__ void mg_pick(int v, F&& f, std::integer_sequence<int, Vs...>) {
  (
      [&] {
        if (v == Vs) f(std::integral_constant<int, Vs>{});
      }(),
      ...);
}
template <int N, class F>
__device__ __forceinline__ void mg_pick(int v, F&& f) {
  mg_pick<N>(v, f, std::make_integer_sequence<int, N>{});
}

template <class R, int T0, int T1>
__device__ __forceinline__ void mg_dense2(typename Cplx<R>::V (&a)[16], const MicroGate& g) {
  using V = typename Cplx<R>::V;
  V m[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) m[e] = ldm<R>(g.m[e]);
  const int cm = __builtin_amdgcn_readfirstlane(g.cmask), cv = __builtin_amdgcn_readfirstlane(g.cval);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    if (i & ((1 << T0) | (1 << T1))) continue;
    if ((i & cm) != cv) continue;
    const int idx[4] = {i, i | (1 << T0), i | (1 << T1), i | (1 << T0) | (1 << T1)};
    V in[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) in[c] = a[idx[c]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      V acc = cmul(m[r * 4], in[0]);
#pragma unroll
      for (int c = 1; c < 4; ++c) acc = cadd(acc, cmul(m[r * 4 + c], in[c]));
      a[idx[r]] = acc;
    }
  }
}

template <class R>
__device__ __forceinline__ void run_micro_gate(typename Cplx<R>::V (&a)[16], const MicroGate& g) {
  // the gate is the same for every lane: its header in scalar registers, so the
  // dispatch below is scalar branches
  const int kind = __builtin_amdgcn_readfirstlane(g.kind);
  if (kind != MG_DENSE2) {
    const int t0 = __builtin_amdgcn_readfirstlane(g.t0) & 3;
    const int cm = __builtin_amdgcn_readfirstlane(g.cmask);
    // Uncontrolled dense gates (the merged rotations of a qubit) skip the per-register
    // control test: on the 28-qubit brickwork that took a pass from 3.24 to 3.15 ms.
    // One-control forms for diagonal or permutation gates (control bit as a template
    // argument) measured 7-10 % faster on a 12-qubit QFT but 5-15 % slower on the
    // brickwork, whose CNOTs are controlled permutations, so every other gate keeps
    // the general form.
    const bool plain = kind == MG_DENSE1 && cm == 0;
    mg_pick<4>(t0, [&](auto tc) {
      constexpr int T = decltype(tc)::value;
      if (plain) mg_dense1<R, T, -1>(a, g);
      else if (kind == MG_DENSE1) mg_dense1<R, T, 4>(a, g);
      else if (kind == MG_DIAG1) mg_diag1<R, T, 4>(a, g);
      else mg_perm1<R, T, 4>(a, g);
    });
    return;
  }
  switch (kind) {
    default:
      switch (__builtin_amdgcn_readfirstlane(g.t0 * 4 + g.t1)) {
        case 1: mg_dense2<R, 0, 1>(a, g); break;
        case 2: mg_dense2<R, 0, 2>(a, g); break;
        case 3: mg_dense2<R, 0, 3>(a, g); break;
        case 4: mg_dense2<R, 1, 0>(a, g); break;
        case 6: mg_dense2<R, 1, 2>(a, g); break;
        case 7: mg_dense2<R, 1, 3>(a, g); break;
        case 8: mg_dense2<R, 2, 0>(a, g); break;
        case 9: mg_dense2<R, 2, 1>(a, g); break;
        case 11: mg_dense2<R, 2, 3>(a, g); break;
        case 12: mg_dense2<R, 3, 0>(a, g); break;
        case 13: mg_dense2<R, 3, 1>(a, g); break;
        default: mg_dense2<R, 3, 2>(a, g); break;
      }
      break;
  }
}

template <class R>
__global__ __launch_bounds__(kTileThreads) void k_tile(typename Cplx<R>::V* __restrict__ s, TilePass pass,
                                                       LockedBits tile_lk, const TileGroup* __restrict__ groups,
                                                       const MicroGate* __restrict__ mg) {
  using V = typename Cplx<R>::V;
  __shared__ V lds[kTileN + kTileN / 16];
  __shared__ uint64_t hoff[1 << kTileHigh];
  __shared__ MicroGate win[kTileMicroWin];  // micro gates [wfirst, wfirst + kTileMicroWin)
  __shared__ TileGroup sgr[kTileGroupsLds];
  const int tid = threadIdx.x;
  const bool glds = pass.ngroups <= kTileGroupsLds;
  if (glds)
    for (int i = tid; i < pass.ngroups; i += kTileThreads) sgr[i] = groups[pass.first_group + i];
  const uint64_t base = insert_zeros(uint64_t(blockIdx.x), tile_lk.pos, tile_lk.npos);
  if (tid < (1 << kTileHigh)) {
    uint64_t o = 0;
#pragma unroll
    for (int k = 0; k < kTileHigh; ++k)
      if ((tid >> k) & 1) o |= uint64_t(1) << pass.hbits[k];
    hoff[tid] = o;
  }
  __syncthreads();
  V v[kAmpsPerThread];
#pragma unroll
  for (int k = 0; k < kAmpsPerThread; ++k) {
    const int p = tid + k * kTileThreads;
    v[k] = s[base + (p & ((1 << kTileLow) - 1)) + hoff[p >> kTileLow]];
  }
#pragma unroll
  for (int k = 0; k < kAmpsPerThread; ++k) lds[lpad(tid + k * kTileThreads)] = v[k];
  __syncthreads();

  int wfirst = 0, wend = 0;  // the window's micro gates (none yet)
  for (int gi = 0; gi < pass.ngroups; ++gi) {
    const TileGroup G = glds ? sgr[gi] : groups[pass.first_group + gi];
    const int gfirst = __builtin_amdgcn_readfirstlane(G.first), gcount = __builtin_amdgcn_readfirstlane(G.count);
    if (gfirst < wfirst || gfirst + gcount > wend) {  // uniform: refill the window from this group on
      wfirst = gfirst;
      wend = min(gfirst + kTileMicroWin, pass.mg_end);
      const uint64_t* src = reinterpret_cast<const uint64_t*>(mg + wfirst);
      uint64_t* dst = reinterpret_cast<uint64_t*>(win);
      const int words = (wend - wfirst) * int(sizeof(MicroGate) / sizeof(uint64_t));
      for (int i = tid; i < words; i += kTileThreads) dst[i] = src[i];
      __syncthreads();
    }
    int qs[kGroupBits];
#pragma unroll
    for (int k = 0; k < kGroupBits; ++k) qs[k] = __builtin_amdgcn_readfirstlane(G.q[k]);
    // sub-cube base: insert zeros at the group's tile qubits (ascending) into tid
    int sorted[kGroupBits] = {qs[0], qs[1], qs[2], qs[3]};
#pragma unroll
    for (int x = 0; x < kGroupBits; ++x)
#pragma unroll
      for (int y = 0; y < kGroupBits - 1 - x; ++y)
        if (sorted[y] > sorted[y + 1]) {
          int t = sorted[y];
          sorted[y] = sorted[y + 1];
          sorted[y + 1] = t;
        }
    int b = tid;
#pragma unroll
    for (int k = 0; k < kGroupBits; ++k) b = ((b >> sorted[k]) << (sorted[k] + 1)) | (b & ((1 << sorted[k]) - 1));
    V a[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int p = b | ((i & 1) << qs[0]) | (((i >> 1) & 1) << qs[1]) | (((i >> 2) & 1) << qs[2]) |
                    (((i >> 3) & 1) << qs[3]);
      a[i] = lds[lpad(p)];
    }
    for (int k = 0; k < gcount; ++k) run_micro_gate<R>(a, win[gfirst - wfirst + k]);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int p = b | ((i & 1) << qs[0]) | (((i >> 1) & 1) << qs[1]) | (((i >> 2) & 1) << qs[2]) |
                    (((i >> 3) & 1) << qs[3]);
      lds[lpad(p)] = a[i];
    }
    __syncthreads();
  }

#pragma unroll
  for (int k = 0; k < kAmpsPerThread; ++k) v[k] = lds[lpad(tid + k * kTileThreads)];
#pragma unroll
  for (int k = 0; k < kAmpsPerThread; ++k) {
    const int p = tid + k * kTileThreads;
    s[base + (p & ((1 << kTileLow) - 1)) + hoff[p >> kTileLow]] = v[k];
  }
}

}  // namespace

template <class R>
int launch_tile(typename Cplx<R>::V* s, int nbits, const TilePass& pass, const TileGroup* groups,
                const MicroGate* mg, hipStream_t st) {
  LockedBits lk{};
  int pos[kTileBits];
  for (int k = 0; k < kTileLow; ++k) pos[k] = k;
  for (int k = 0; k < kTileHigh; ++k) pos[kTileLow + k] = pass.hbits[k];
  for (int a = 1; a < kTileBits; ++a)
    for (int b = a; b > 0 && pos[b - 1] > pos[b]; --b) std::swap(pos[b - 1], pos[b]);
  for (int k = 0; k < kTileBits; ++k) lk.pos[k] = uint8_t(pos[k]);
  lk.npos = kTileBits;
  const unsigned ntiles = unsigned(uint64_t(1) << (nbits - kTileBits));
  hipLaunchKernelGGL(k_tile<R>, dim3(ntiles), dim3(kTileThreads), 0, st, s, pass, lk, groups, mg);
  return YH_OK;
}

template int launch_tile<double>(double2*, int, const TilePass&, const TileGroup*, const MicroGate*, hipStream_t);
template int launch_tile<float>(float2*, int, const TilePass&, const TileGroup*, const MicroGate*, hipStream_t);

}  // namespace yh
