// block.hip — yh_apply_fused with max_fuse_qubits = 5: gates fused into 5-qubit
// blocks (32 x 32 complex matrices, multiplied in FP64 on the host), applied as
// real-split MFMA contractions: a block on its own in one HBM pass (k_block5), or
// several blocks whose bits fit one 12-bit LDS tile in one pass (k_tblocks).
//
// Planner: a block starts at the first gate not yet applied and takes, in circuit
// order, every gate whose qubits (targets and controls) keep the block within 5
// qubits and that no skipped earlier gate shares a qubit with; a skipped gate
// blocks its qubits for the rest of the scan. Gates commute past each other only
// when their qubit sets are disjoint, so the order of yao-rs's sequential
// apply_inplace (src/apply.rs:123-147) is kept. Gates on more than 5 qubits run
// through the per-gate path.
//
// Kernel (per block): the state is a 32 x 2^(n-5) complex matrix S (row r = the
// block-qubit bits of an amplitude index, column c = the other bits). out = U S is
// computed as the real 64 x 64 product [[Ur, -Ui], [Ui, Ur]] [Sr; Si] with MFMA
// 16x16x4 (FP64 or FP32). Each wave takes 16 columns per step: the amplitudes go
// from HBM straight into B-operand registers, the 64 MFMAs read U's A fragments
// from LDS, and the accumulators hold whole output amplitudes, which are stored
// directly; the next step's amplitudes are loaded while the MFMAs run. No
// barrier in the loop.
#include <algorithm>
#include <cstring>
#include <functional>
#include <mutex>
#include <vector>

#include "host.h"
#include "kernels.h"
#include "runtime.h"

namespace yh {

namespace {

constexpr int kBQ = 5;          // block qubits
constexpr int kBD = 1 << kBQ;   // 32 rows
constexpr int kCols = 16;       // columns per wave step
constexpr int kFrag = 64 * 64;  // A fragments of one block: 4 row tiles x 16 k-steps x 64 lanes

template <class R>
struct MfmaOps;

template <>
struct MfmaOps<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // D[i][j] sits in lane j + 16 (i % 4), element i / 4
  static __device__ __forceinline__ int drow(int lane, int e) { return 4 * e + (lane >> 4); }
};
template <>
struct MfmaOps<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // D[i][j] sits in lane j + 16 (i / 4), element i % 4
  static __device__ __forceinline__ int drow(int lane, int e) { return 4 * (lane >> 4) + e; }
};

struct BlockArgs {
  int spos[kBQ];   // block bits ascending (zero insertion)
  int p[kBQ];      // bit j of a row index <-> memory bit p[j]
  uint64_t ncol;   // 2^(n-5)
  uint64_t nchunk; // ceil(ncol / 16)
};

__device__ __forceinline__ uint64_t row_offset(int r, const int* p) {
  uint64_t o = 0;
#pragma unroll
  for (int j = 0; j < kBQ; ++j) o |= uint64_t((r >> j) & 1) << p[j];
  return o;
}

// MFMA 16x16x4 operand layouts (wave64): A[i][k] in lane i + 16 k, B[k][j] in
// lane j + 16 k; the row of D held by (lane, element) is MfmaOps<R>::drow.
// Each wave owns 16 columns per step: lane l reads the 8 amplitudes of column
// l % 16 on rows l / 16 + 4 m (m = 0..7), which are exactly its B operands
// (real parts for k-steps 0..7, imaginary parts for 8..15); output row tiles
// 0, 1 hold real parts and 2, 3 imaginary parts of the same rows, so each lane
// stores 8 whole amplitudes. U's A fragments, in fragment order, live in LDS.
template <class R>
__global__ __launch_bounds__(256, 3) void k_block5(typename Cplx<R>::V* __restrict__ s, BlockArgs a,
                                                const R* __restrict__ ufrag) {
  using V = typename Cplx<R>::V;
  using M = MfmaOps<R>;
  __shared__ R A[kFrag];
  for (int e = threadIdx.x; e < kFrag; e += 256) A[e] = ufrag[e];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  // row offsets split into a lane part and a wave-uniform part: input rows are
  // grp + 4 m; output rows 16 (m / 4) + drow(lane, m % 4)
  const uint64_t in_lane = row_offset(grp, a.p);
  const uint64_t out_lane = row_offset(M::drow(lane, 0), a.p);
  uint64_t in_u[8], out_u[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    in_u[m] = row_offset(4 * m, a.p);
    out_u[m] = row_offset(16 * (m >> 2) + (M::drow(lane, m & 3) - M::drow(lane, 0)), a.p);
  }
  auto base_of = [&](uint64_t c) {
    uint64_t b = c;
#pragma unroll
    for (int j = 0; j < kBQ; ++j) {
      const uint64_t lo = b & ((uint64_t(1) << a.spos[j]) - 1);
      b = ((b >> a.spos[j]) << (a.spos[j] + 1)) | lo;
    }
    return b;
  };
  const uint64_t nwaves = uint64_t(gridDim.x) * 4;
  const uint64_t q0 = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (q0 >= a.nchunk) return;  // wave-uniform
  // Columns past the end are clamped to the last column: such lanes load and
  // store exactly what the lane of the last column does (same data, same MFMA
  // column), so every load and store is unconditional.
  auto base_at = [&](uint64_t q) {
    const uint64_t c = q * kCols + col;
    return base_of(c < a.ncol ? c : a.ncol - 1);
  };
  V v[8];
  {
    const uint64_t b = base_at(q0);
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = s[b | in_lane | in_u[m]];
  }
  // Software pipeline, per step: store the previous step's results, issue the
  // next step's loads, then run this step's MFMAs. The wait for the loads at the
  // top of the next step then also covers stores issued before the MFMAs.
  typename M::acc_t prev[4];
  uint64_t bprev = 0;
  bool have_prev = false;
  for (uint64_t q = q0; q < a.nchunk; q += nwaves) {
    const uint64_t b = base_at(q);
    R bre[8], bim[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      bre[m] = v[m].x;
      bim[m] = v[m].y;
    }
    if (have_prev) {
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        V o;
        o.x = prev[m >> 2][m & 3];
        o.y = prev[2 + (m >> 2)][m & 3];
        s[bprev | out_lane | out_u[m]] = o;
      }
    }
    {
      const uint64_t qn = q + nwaves < a.nchunk ? q + nwaves : a.nchunk - 1;
      const uint64_t bn = base_at(qn);
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = s[bn | in_lane | in_u[m]];
    }
    typename M::acc_t acc[4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) acc[rt] = typename M::acc_t{0, 0, 0, 0};
    // an empty asm keeps the A-fragment reads inside the loop (hoisted, they
    // would hold 64 fragments in registers for the whole launch)
    int ai = lane;
    asm volatile("" : "+v"(ai));
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
      const R bv = ks < 8 ? bre[ks] : bim[ks - 8];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) acc[rt] = M::mfma(A[(rt * 16 + ks) * 64 + ai], bv, acc[rt]);
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) prev[rt] = acc[rt];
    bprev = b;
    have_prev = true;
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    V o;
    o.x = prev[m >> 2][m & 3];
    o.y = prev[2 + (m >> 2)][m & 3];
    s[bprev | out_lane | out_u[m]] = o;
  }
}

// ---- several blocks per HBM pass: an LDS tile of 2^12 amplitudes ----
// The tile holds memory bits 0..3 (runs of 16 consecutive amplitudes, 256 B in
// complex128) plus 8 more; every block of the pass has its 5 bits among them. One
// 512-thread workgroup per CU keeps the tile and the A values of all the pass's
// blocks in LDS (the A values once per launch). Each block is an MFMA product on
// the tile viewed as a 32 x 128 matrix (one 16-column step per wave), with a
// barrier between blocks. The complex product U S is formed from three real
// products (Gauss): Ur Sr, Ui Si and (Ur + Ui)(Sr + Si), 48 MFMAs per 16-column step
// instead of the 64 of the real-split form; the A values are Ur and Ui (32 per
// lane). Rounding differs from k_block5's at the level of a few ulps of |U||S|.
// Moving a tile in and out:
//  - DIN: the first block's operands go from HBM straight into registers, in its
//    MFMA layout (loaded during the previous tile), and its results fill LDS;
//    otherwise the tile is loaded into registers during the previous tile, in
//    thread order, and written to LDS before the first block.
//  - DOUT: the last block's results go from registers straight to HBM; otherwise
//    the tile is stored from LDS after the last block.
// The host sets DIN / DOUT for a block whose rows avoid memory bits 0..3: its 16
// columns per row are then 16 consecutive amplitudes in memory. Barriers inside the
// loop order LDS only, so the next tile's loads stay in flight while the MFMAs run.
constexpr int kTT = 12;                // tile bits
constexpr int kTN = 1 << kTT;          // 4096 amplitudes
constexpr int kTLow = 4;               // memory bits 0..3 are always tile bits
constexpr int kTHigh = kTT - kTLow;    // 8 tile bits chosen per pass
constexpr int kTThreads = 512;         // 8 waves: one 16-column step each per block
constexpr int kPassMax = 5;            // blocks per pass
constexpr int kCFrag = 32 * 64;        // compact A values of one block

// A block of the pass is a 32 x 32 matrix on 5 tile bits, or (bit b of `small`) a
// 16 x 16 matrix on 4: then row bit 4 does not exist and its q / qs entries hold
// kNoBit, a position above every tile bit (inserting a zero there, or setting it
// for row bit 4 of a row index below 16, changes nothing).
constexpr int kNoBit = 30;
// modelled time of a 16 x 16 block in a multi-block pass, in 32 x 32 block-times
constexpr double kSmallCost = 0.55;

struct TBlockArgs {
  uint8_t tbit[kTT];          // memory bit of tile position k, ascending (0..3 first)
  uint8_t q[kPassMax][kBQ];   // tile position of row bit j of block b
  uint8_t qs[kPassMax][kBQ];  // the same positions, ascending
  int32_t nblk;
  uint32_t small;             // bit b: block b is a 4-bit (16 x 16) block
  uint64_t ntiles;            // 2^(n - 12)
};

__device__ __forceinline__ int tpad(int p) { return p + (p >> 4); }

// NTL = 2: the pass's blocks are all 16 x 16 and each step of a workgroup takes two
// tiles (adjacent tile indices) in LDS at once, so each block's geometry, A values
// and barrier serve twice the MFMA work; their A values are staged kSFrag apart.
constexpr int kSFrag = 8 * 64;  // compact A values of one 16 x 16 block

template <class R, bool DIN, bool DOUT, int NTL>
__global__ __launch_bounds__(kTThreads, 1) void k_tblocks(typename Cplx<R>::V* __restrict__ s, TBlockArgs a,
                                                        const R* __restrict__ frag) {
  using V = typename Cplx<R>::V;
  using M = MfmaOps<R>;
  constexpr int kAStride = NTL == 2 ? kSFrag : kCFrag;
  __shared__ V lds[NTL][kTN + kTN / 16];
  __shared__ R afr[kPassMax * kAStride];
  __shared__ uint64_t hoff[1 << kTHigh];
  __shared__ int bq[kPassMax][2 * kBQ];  // per block: q[0..4], then qs[0..4]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, grp = lane >> 4;
  if (tid < (1 << kTHigh)) {
    uint64_t o = 0;
#pragma unroll
    for (int k = 0; k < kTHigh; ++k)
      if ((tid >> k) & 1) o |= uint64_t(1) << a.tbit[kTLow + k];
    hoff[tid] = o;
  }
  if (tid < kPassMax * kBQ) {
    const int b = tid / kBQ, j = tid % kBQ;
    bq[b][j] = a.q[b][j];
    bq[b][kBQ + j] = a.qs[b][j];
  }
  for (int e = tid; e < a.nblk * kAStride; e += kTThreads) afr[e] = frag[e];
  __syncthreads();
  constexpr int kPer = kTN / kTThreads;
  // memory offset (from the tile's base) of tile position p
  auto mem = [&](int p) { return uint64_t(p & ((1 << kTLow) - 1)) + hoff[p >> kTLow]; };
  // geometry of block b for this lane: row offsets and the column's tile position
  // a 5-bit block views the tile as 32 rows x 128 columns, one 16-column group per
  // wave (column cb); a 4-bit block as 16 x 256, two groups per wave (cb, cb1)
  struct Geo {
    int q[kBQ];
    int cb, cb1, in_lane, out_lane;
    int xb0, xb1, ob0, ob1;  // padded LDS index of the lane's column | row bits 0..1, operands / results
    bool small;
  };
  auto roff = [](const Geo& g, int r) {
    int o = 0;
#pragma unroll
    for (int j = 0; j < kBQ; ++j) o |= ((r >> j) & 1) << g.q[j];
    return o;
  };
  auto geo = [&](int b, Geo& g) {
    int qs[kBQ];
    // the block's bit positions are the same in every lane: scalar registers, so
    // the uniform row offsets below are scalar arithmetic
#pragma unroll
    for (int j = 0; j < kBQ; ++j) {
      g.q[j] = __builtin_amdgcn_readfirstlane(bq[b][j]);
      qs[j] = __builtin_amdgcn_readfirstlane(bq[b][kBQ + j]);
    }
    g.small = NTL == 2 || ((a.small >> b) & 1);
    g.in_lane = roff(g, grp);
    g.out_lane = roff(g, M::drow(lane, 0));
    // this lane's column(s), then their tile positions
    auto at = [&](int c) {
#pragma unroll
      for (int j = 0; j < kBQ; ++j) c = ((c >> qs[j]) << (qs[j] + 1)) | (c & ((1 << qs[j]) - 1));
      return c;
    };
    if (g.small) {
      g.cb = at(wave * 2 * kCols + col);
      g.cb1 = at(wave * 2 * kCols + kCols + col);
    } else {
      g.cb = at(wave * kCols + col);
      g.cb1 = 0;
    }
    g.xb0 = tpad(g.cb | g.in_lane);
    g.xb1 = tpad(g.cb1 | g.in_lane);
    g.ob0 = tpad(g.cb | g.out_lane);
    g.ob1 = tpad(g.cb1 | g.out_lane);
  };
  // tile position of operand m (B fragment: k-step m, or group m / 4 and k-step
  // m % 4) and of result m (row tile m / 4, or group m / 4; element m % 4)
  auto xpos = [&](const Geo& g, int m) {
    return g.small ? ((m >> 2 ? g.cb1 : g.cb) | g.in_lane | roff(g, 4 * (m & 3))) : (g.cb | g.in_lane | roff(g, 4 * m));
  };
  auto opos = [&](const Geo& g, int m) {
    const int d = M::drow(lane, m & 3) - M::drow(lane, 0);
    return g.small ? ((m >> 2 ? g.cb1 : g.cb) | g.out_lane | roff(g, d)) : (g.cb | g.out_lane | roff(g, 16 * (m >> 2) + d));
  };
  // their padded LDS indices: p + p / 16 is additive over disjoint bit fields, so
  // each index is the lane's base (column and row bits 0..1) plus a row offset that
  // is the same in every lane
  auto xlds = [&](const Geo& g, int m) {
    return g.small ? (m >> 2 ? g.xb1 : g.xb0) + tpad(roff(g, 4 * (m & 3))) : g.xb0 + tpad(roff(g, 4 * m));
  };
  auto olds = [&](const Geo& g, int m) {
    const int d = M::drow(lane, m & 3) - M::drow(lane, 0);
    return g.small ? (m >> 2 ? g.ob1 : g.ob0) + tpad(roff(g, d)) : g.ob0 + tpad(roff(g, 16 * (m >> 2) + d));
  };
  // a workgroup step takes NTL adjacent tiles (a "unit"); units are dealt out in
  // turn to the persistent workgroups
  const uint64_t nunits = a.ntiles / NTL;
  uint64_t u = blockIdx.x;
  if (u >= nunits) return;  // uniform
  auto unit_base = [&](uint64_t uu, int i) { return insert_zeros(uu * NTL + i, a.tbit, kTT); };
  static_assert(kPer == 8, "eight amplitudes per thread");
  Geo g0;
  geo(0, g0);
  // the next unit's amplitudes: with DIN in block 0's MFMA operand layout, else in
  // thread order (position tid + kTThreads k)
  R vx[NTL][8], vy[NTL][8];
  auto load_unit = [&](uint64_t uu) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < NTL; ++i) {
      const uint64_t bb = unit_base(uu, i);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int p = DIN ? xpos(g0, k) : tid + kTThreads * k;
        const V x = s[bb + mem(p)];
        vx[i][k] = x.x;
        vy[i][k] = x.y;
      }
    }
  };
  auto fill = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < NTL; ++i)
#pragma unroll
      for (int k = 0; k < kPer; ++k) lds[i][tpad(tid + kTThreads * k)] = cmk<V>(vx[i][k], vy[i][k]);
  };
  // block b of the pass on the unit uc; `first` (DIN) takes the operands from
  // vx/vy and then loads the unit un into them, `to_hbm` (DOUT) writes the results
  // to HBM instead of LDS. Both are constants where the block is inlined.
  auto block = [&](int b, bool first, bool to_hbm, uint64_t uc, uint64_t un) __attribute__((always_inline)) {
    // the block's geometry first (its LDS reads return in order, so they go
    // before the A values'), then per tile the operands, the MFMAs and the results
    Geo g;
    if (b == 0) g = g0;
    else geo(b, g);
    const R* f = afr + b * kAStride;
    // tile i's operands: registers (DIN, first block) or LDS
    auto operands = [&](int i, V (&x)[8]) __attribute__((always_inline)) {
      if (DIN && first) {
#pragma unroll
        for (int m = 0; m < 8; ++m) x[m] = cmk<V>(vx[i][m], vy[i][m]);
      } else {
#pragma unroll
        for (int m = 0; m < 8; ++m) x[m] = lds[i][xlds(g, m)];
      }
    };
    // three real products instead of four (Gauss): P1 = Ur Sr, P2 = Ui Si,
    // P3 = (Ur + Ui)(Sr + Si); out = (P1 - P2) + i (P3 - P1 - P2). Accumulator
    // pair index: row tile (5-bit block) or column group (4-bit block).
    typename M::acc_t p1[2], p2[2], p3[2];
    auto results = [&](int i) __attribute__((always_inline)) {
      if (DIN && first && i == NTL - 1) {
        // every tile's operands are consumed: the next unit's loads reuse their
        // registers (kept below the MFMAs, so no second register set is needed)
        __builtin_amdgcn_sched_barrier(0);
        load_unit(un);
      }
      const uint64_t tb = to_hbm ? unit_base(uc, i) : 0;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        V o;
        const R a1 = p1[m >> 2][m & 3], a2 = p2[m >> 2][m & 3];
        o.x = a1 - a2;
        o.y = p3[m >> 2][m & 3] - a1 - a2;
        if (to_hbm) s[tb + mem(opos(g, m))] = o;
        else lds[i][olds(g, m)] = o;
      }
    };
    auto zero = [&]() __attribute__((always_inline)) {
#pragma unroll
      for (int rt = 0; rt < 2; ++rt) {
        p1[rt] = typename M::acc_t{0, 0, 0, 0};
        p2[rt] = typename M::acc_t{0, 0, 0, 0};
        p3[rt] = typename M::acc_t{0, 0, 0, 0};
      }
    };
    V x[8];
    if (g.small) {
      R ur[4], ui[4];
      operands(0, x);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        ur[e] = f[e * 64 + lane];
        ui[e] = f[(4 + e) * 64 + lane];
      }
      auto mfmas = [&]() __attribute__((always_inline)) {
        zero();
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
          for (int gi = 0; gi < 2; ++gi) {
            V xv = x[gi * 4 + ks];
            // an empty asm keeps this path's operand sums from being merged with the
            // other path's and hoisted above the branch (with every wait for the
            // prefetched operands of a direct-in block)
            asm volatile("" : "+v"(xv.x));
            p1[gi] = M::mfma(ur[ks], xv.x, p1[gi]);
            p2[gi] = M::mfma(ui[ks], xv.y, p2[gi]);
            p3[gi] = M::mfma(ur[ks] + ui[ks], xv.x + xv.y, p3[gi]);
          }
      };
      mfmas();
      if (NTL == 2) {
        // the second tile's operands go into the first tile's operand registers
        // once its MFMAs are issued, and are read while those MFMAs run
        __builtin_amdgcn_sched_barrier(0);
        operands(NTL - 1, x);
        results(0);
        mfmas();
        results(NTL - 1);
      } else {
        results(0);
      }
    } else {
      R ur[16], ui[16];
      operands(0, x);
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        ur[e] = f[e * 64 + lane];
        ui[e] = f[(16 + e) * 64 + lane];
      }
      zero();
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        const R br = x[ks].x, bi = x[ks].y, bs = x[ks].x + x[ks].y;
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) {
          p1[rt] = M::mfma(ur[rt * 8 + ks], br, p1[rt]);
          p2[rt] = M::mfma(ui[rt * 8 + ks], bi, p2[rt]);
          p3[rt] = M::mfma(ur[rt * 8 + ks] + ui[rt * 8 + ks], bs, p3[rt]);
        }
      }
      results(0);  // 32 x 32 blocks run one tile per step only
    }
    lds_barrier();
  };
  // One unit: the next unit's loads are issued before this unit's stores, and every
  // wait on them follows those stores in straight-line code (the fill, or the next
  // unit's first block after the peeled first unit), so the stores stay in flight
  // while the next unit starts. A pass holds at least two blocks (the host runs a
  // block alone with k_block5), so the first and the DOUT block differ.
  auto run_unit = [&](uint64_t uc, uint64_t un) __attribute__((always_inline)) {
    if (!DIN) load_unit(un);
    const int nmid = DOUT ? a.nblk - 1 : a.nblk;
    block(0, true, false, uc, un);
    for (int b = 1; b < nmid; ++b) block(b, false, false, uc, un);
    if (DOUT) block(a.nblk - 1, false, true, uc, un);
    if (!DOUT) {
      // without DIN each thread stores, then refills, the same LDS positions, so no
      // barrier is needed between the two; with DIN the next unit's first block
      // writes other positions, so a barrier follows the store
#pragma unroll
      for (int i = 0; i < NTL; ++i) {
        const uint64_t tb = unit_base(uc, i);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
          const int p = tid + kTThreads * k;
          s[tb + mem(p)] = lds[i][tpad(p)];
        }
      }
      if (DIN) lds_barrier();
    }
  };
  // the last unit of a workgroup loads itself again (unused): unconditional loads
  // keep the operands in registers
  load_unit(u);
  if (!DIN) {
    fill();
    lds_barrier();
  }
  uint64_t un = u + gridDim.x < nunits ? u + gridDim.x : u;
  run_unit(u, un);  // peeled: the loop is entered with the same operations in flight as at its back edge
  while (u + gridDim.x < nunits) {  // uniform
    if (!DIN) {
      fill();
      lds_barrier();
    }
    u += gridDim.x;
    un = u + gridDim.x < nunits ? u + gridDim.x : u;
    run_unit(u, un);
  }
}

struct BGate {
  size_t op_index = 0;
  int sub = 0;  // target index for X gates split per target
  bool blockable = false;
  uint64_t mask = 0;
  int nt = 0;
  std::vector<int> tb;               // target bits, gate-index order
  std::vector<std::pair<int, int>> ctl;  // (bit, value)
  Mat m;
};

int popc(uint64_t x) { return __builtin_popcountll(x); }

int to_bgates(int nbits, const yh_op* ops, size_t nops, std::vector<BGate>& out) {
  auto bit_of = [nbits](int loc) { return nbits - 1 - loc; };
  for (size_t e = 0; e < nops; ++e) {
    const yh_op& op = ops[e];
    if (op.elem != YH_ELEM_GATE) continue;
    const int subs = op.kind == YH_GATE_X ? op.nt : 1;
    Mat m;
    if (op.kind == YH_GATE_X) {
      m = {cd{0, 0}, cd{1, 0}, cd{1, 0}, cd{0, 0}};
    } else {
      const int d = gate_matrix(op, m);
      if (d != (1 << op.nt)) return fail(YH_ERR_MATRIX_SIZE, "gate matrix size does not match targets");
    }
    for (int sidx = 0; sidx < subs; ++sidx) {
      BGate g;
      g.op_index = e;
      g.sub = sidx;
      g.m = m;
      if (op.kind == YH_GATE_X) {
        g.nt = 1;
        g.tb.push_back(bit_of(op.targets[sidx]));
      } else {
        g.nt = op.nt;
        for (int k = 0; k < op.nt; ++k) g.tb.push_back(bit_of(op.targets[k]));
      }
      for (int k = 0; k < op.nc; ++k) g.ctl.push_back({bit_of(op.ctrls[k]), op.ctrl_vals[k]});
      for (int b : g.tb) g.mask |= uint64_t(1) << b;
      for (auto& c : g.ctl) g.mask |= uint64_t(1) << c.first;
      g.blockable = popc(g.mask) <= kBQ;
      out.push_back(std::move(g));
    }
  }
  return YH_OK;
}

struct BStep {
  bool block = false;
  size_t gate = 0;               // single launch: index into the gate list
  int p[kBQ] = {0, 0, 0, 0, 0};  // block: row bit j <-> memory bit p[j]
  uint64_t support = 0;          // block: bits its gates act on (p[] adds identity bits)
  bool small = false;            // block: applied as a 4-bit block on p[0..3] (p[4] identity)
  std::vector<uint32_t> members; // block: gate indices, circuit order
};

// U <- G U for a gate on local rows (bit k of the gate index <- local bit lt[k]).
void left_multiply(std::vector<cd>& U, const BGate& g, const int* lt, const std::vector<std::pair<int, int>>& lc) {
  const int D = 1 << g.nt;
  int tmask = 0;
  for (int k = 0; k < g.nt; ++k) tmask |= 1 << lt[k];
  std::vector<cd> in(D), outv(D);
  std::vector<int> idx(D);
  for (int col = 0; col < kBD; ++col) {
    for (int base = 0; base < kBD; ++base) {
      if (base & tmask) continue;
      bool on = true;
      for (auto& c : lc) on = on && (((base >> c.first) & 1) == c.second);
      if (!on) continue;
      for (int j = 0; j < D; ++j) {
        int r = base;
        for (int k = 0; k < g.nt; ++k)
          if ((j >> k) & 1) r |= 1 << lt[k];
        idx[j] = r;
        in[j] = U[size_t(r) * kBD + col];
      }
      for (int j = 0; j < D; ++j) {
        cd acc{0, 0};
        for (int l = 0; l < D; ++l) acc = acc + g.m[size_t(j) * D + l] * in[l];
        outv[j] = acc;
      }
      for (int j = 0; j < D; ++j) U[size_t(idx[j]) * kBD + col] = outv[j];
    }
  }
}

// ---- planner ----
// A block is a set Q of at most 5 memory bits; it absorbs, in circuit order, every
// remaining gate inside Q that no skipped earlier gate shares a bit with (a
// skipped gate blocks its bits). Strategies for choosing Q at each step:
//   0 first-come: Q grows with the qubits of the gates met in circuit order;
//   1 windows: the best of the 5-bit windows [b, b+5), plus strategy 0's set;
//   2 local search: from each of the first ready gates, Q grows one bit at a
//     time by the bit that lets the most gates in, plus strategy 0's set;
//   3 both 1 and 2.
// "Best" is the Q that absorbs the most gates. The plan with the fewest blocks
// over the four strategies is used.
struct Planner {
  const std::vector<BGate>& bg;
  int nbits;
  uint64_t all;
  std::vector<uint32_t> rem;  // remaining gate indices, circuit order
  Planner(const std::vector<BGate>& g, int n) : bg(g), nbits(n), all(n >= 64 ? ~uint64_t(0) : (uint64_t(1) << n) - 1) {}

  void absorb(uint64_t Q, std::vector<uint32_t>& taken) const {
    taken.clear();
    uint64_t B = 0;
    for (uint32_t i : rem) {
      const BGate& g = bg[i];
      if ((g.mask & B) || !g.blockable || (g.mask & ~Q)) {
        B |= g.mask;
        if (B == all) break;
        continue;
      }
      taken.push_back(i);
    }
  }
  int lim = kBQ;  // bits of the block being chosen (4 or 5)
  uint64_t first_come() const {
    uint64_t Q = 0, B = 0;
    for (uint32_t i : rem) {
      const BGate& g = bg[i];
      if (!g.blockable || (g.mask & B) || popc(Q | g.mask) > lim)
        B |= g.mask;
      else
        Q |= g.mask;
      if (B == all) break;
    }
    return Q;
  }
  void ready(std::vector<uint32_t>& out, size_t cap) const {
    out.clear();
    uint64_t B = 0;
    for (uint32_t i : rem) {
      if (!(bg[i].mask & B) && bg[i].blockable) out.push_back(i);
      B |= bg[i].mask;
      if (out.size() >= cap || B == all) break;
    }
  }
  uint64_t grow(uint64_t Q, std::vector<uint32_t>& tmp) const {
    while (popc(Q) < lim) {
      absorb(Q, tmp);
      std::vector<char> in_blk;
      uint64_t cand = 0;
      int nc = 0;
      size_t t = 0;
      for (uint32_t i : rem) {
        if (t < tmp.size() && tmp[t] == i) {
          ++t;
          continue;
        }
        const uint64_t m = bg[i].mask;
        if ((m & Q) && popc(m | Q) <= lim && (m & ~Q & ~cand)) {
          cand |= m & ~Q;
          nc = popc(cand);
          if (nc >= 8) break;
        }
      }
      if (!cand) break;
      uint64_t best_bit = 0;
      size_t best_n = 0;
      for (int b = 0; b < 64; ++b)
        if ((cand >> b) & 1) {
          absorb(Q | (uint64_t(1) << b), tmp);
          if (!best_bit || tmp.size() > best_n) {
            best_n = tmp.size();
            best_bit = uint64_t(1) << b;
          }
        }
      Q |= best_bit;
    }
    return Q;
  }
  // the candidate block of at most `lim` bits that absorbs the most gates
  void best_block(int strategy, std::vector<uint32_t>& best, uint64_t& bestQ) {
    std::vector<uint32_t> taken, tmp, rd;
    std::vector<uint64_t> cands{first_come()};
    if (strategy == 1 || strategy == 3)
      for (int b = 0; b + lim <= nbits; ++b) cands.push_back(((uint64_t(1) << lim) - 1) << b);
    if (strategy == 2 || strategy == 3) {
      ready(rd, 16);
      for (uint32_t r : rd) cands.push_back(grow(bg[r].mask, tmp));
    }
    bestQ = 0;
    best.clear();
    for (uint64_t Q : cands) {
      absorb(Q, taken);
      if (taken.size() > best.size()) {
        best.swap(taken);
        bestQ = Q;
      }
    }
  }
  // small4: also try blocks of 4 bits (16 x 16, about half the time of a 5-bit
  // block in a multi-block pass) and keep one when it absorbs more gates per unit
  // of time
  void run(int strategy, bool small4, std::vector<BStep>& steps) {
    rem.resize(bg.size());
    for (size_t i = 0; i < bg.size(); ++i) rem[i] = uint32_t(i);
    std::vector<uint32_t> best, best4;
    std::vector<char> done(bg.size(), 0);
    while (!rem.empty()) {
      const uint32_t head = rem.front();
      if (!bg[head].blockable) {
        BStep s;
        s.gate = head;
        steps.push_back(s);
        done[head] = 1;
      } else {
        uint64_t bestQ = 0, bestQ4 = 0;
        lim = kBQ;
        best_block(strategy, best, bestQ);
        if (small4 && popc(bg[head].mask) <= kBQ - 1) {
          lim = kBQ - 1;
          best_block(strategy, best4, bestQ4);
          lim = kBQ;
          if (!best4.empty() && double(best4.size()) > kSmallCost * double(best.size())) {
            best.swap(best4);
            bestQ = bestQ4;
          }
        }
        BStep s;
        s.block = true;
        int np = 0;
        for (int b = 0; b < nbits && np < kBQ; ++b)
          if ((bestQ >> b) & 1) s.p[np++] = b;
        for (int b = 0; b < nbits && np < kBQ; ++b)
          if (!((bestQ >> b) & 1)) s.p[np++] = b;
        s.members = best;
        for (uint32_t i : best) {
          s.support |= bg[i].mask;
          done[i] = 1;
        }
        steps.push_back(std::move(s));
      }
      size_t k = 0;
      for (uint32_t i : rem)
        if (!done[i]) rem[k++] = i;
      rem.resize(k);
    }
  }
};

// structural signature of a circuit: everything but gate parameters and matrix
// values (element kind, gate or channel kind, positions, control values), which is
// all the planner reads; `salt` separates plans of different planners
std::vector<int32_t> structure_sig(int nbits, const yh_op* ops, size_t nops, int32_t salt) {
  std::vector<int32_t> sig{salt, int32_t(nbits), int32_t(nops)};
  for (size_t e = 0; e < nops; ++e) {
    const yh_op& o = ops[e];
    sig.push_back(o.elem);
    sig.push_back(o.kind);
    sig.push_back(o.nt);
    sig.push_back(o.nc);
    sig.push_back(o.elem == YH_ELEM_CHANNEL ? o.nmat : 0);
    sig.push_back(o.elem == YH_ELEM_CHANNEL ? int32_t(o.params[1]) : 0);  // depolarizing width
    for (int k = 0; k < o.nt && k < YH_MAX_TARGETS; ++k) sig.push_back(o.targets[k]);
    for (int k = 0; k < o.nc && k < YH_MAX_CTRLS; ++k) sig.push_back(o.ctrls[k] * 2 + (o.ctrl_vals[k] & 1));
  }
  return sig;
}

// 64-bit FNV-1a of a signature: a fast pre-filter for cache lookups; a hit also
// compares the whole signature
uint64_t sig_hash(const std::vector<int32_t>& sig) {
  uint64_t h = 1469598103934665603ull;
  for (int32_t v : sig) {
    h ^= uint64_t(uint32_t(v));
    h *= 1099511628211ull;
  }
  return h;
}

struct PlanCache {
  std::mutex mu;
  struct Entry {
    uint64_t key;
    std::vector<int32_t> sig;
    std::vector<BStep> steps;
  };
  std::vector<Entry> entries;  // most recent last, at most 8
};
PlanCache& plan_cache() {
  static PlanCache c;
  return c;
}

// lowest-cost plan of a gate list, cached by `sig` (the structure of the ops the
// gate list was built from)
double plan_cost(const std::vector<BStep>& steps, int nbits);

int plan_gates(const std::vector<BGate>& bg, int nbits, const std::vector<int32_t>& sig, std::vector<BStep>& steps) {
  const uint64_t key = sig_hash(sig);
  {
    PlanCache& c = plan_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    for (auto& e : c.entries)
      if (e.key == key && e.sig == sig) {
        steps = e.steps;
        return YH_OK;
      }
  }
  // every strategy, with and without 4-bit blocks; the plan with the lowest
  // modelled launch time is kept (ties: fewer blocks)
  Planner pl(bg, nbits);
  std::vector<BStep> bestp;
  double best_cost = 1e300;
  size_t best_blocks = SIZE_MAX;
  for (int small4 = 0; small4 < 2; ++small4)
    for (int strategy = 0; strategy < 4; ++strategy) {
      std::vector<BStep> cand;
      pl.run(strategy, small4 != 0, cand);
      size_t nb = 0;
      for (auto& s : cand) nb += s.block;
      const double c = plan_cost(cand, nbits);
      if (c < best_cost - 1e-9 || (c < best_cost + 1e-9 && nb < best_blocks)) {
        best_cost = c;
        best_blocks = nb;
        bestp.swap(cand);
      }
    }
  steps = bestp;
  PlanCache& c = plan_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  c.entries.push_back({key, sig, std::move(bestp)});
  if (c.entries.size() > 8) c.entries.erase(c.entries.begin());
  return YH_OK;
}

int plan_blocks(int nbits, const yh_op* ops, size_t nops, std::vector<BStep>& steps, std::vector<BGate>& bg) {
  YH_TRY(to_bgates(nbits, ops, nops, bg));
  return plan_gates(bg, nbits, structure_sig(nbits, ops, nops, 0), steps);
}

// ---- density matrices: rho as a vector of 2n bits (row bits n..2n-1 above the
// column bits 0..n-1, the row-major layout); a gate is U (x) conj(U) on the row and
// column bits of its qubits, a channel sum_k K_k (x) conj(K_k) (density_matrix.rs
// :161-224 computes the same maps as U rho U^dagger and sum_k K_k rho K_k^dagger).

// the gate's full matrix on local qubits [targets..., controls...] (local bit k <->
// qubit k): the target matrix where every control matches, identity elsewhere
Mat with_controls(const Mat& g, const yh_op& op) {
  const int nt = op.nt, k = op.nt + op.nc, D = 1 << k, dt = 1 << nt;
  Mat F(size_t(D) * D, cd{0, 0});
  int want = 0;
  for (int c = 0; c < op.nc; ++c) want |= (op.ctrl_vals[c] ? 1 : 0) << c;
  for (int i = 0; i < D; ++i) {
    if ((i >> nt) != want) {
      F[size_t(i) * D + i] = cd{1, 0};
      continue;
    }
    for (int j = 0; j < dt; ++j) F[size_t(i) * D + ((i >> nt) << nt | j)] = g[size_t(i & (dt - 1)) * dt + j];
  }
  return F;
}

// superoperator of the maps rho -> M rho M^dagger summed over `ms` (each D x D on
// the local qubits), on local bits [row bits..., column bits...]
Mat superop(const std::vector<Mat>& ms, int k) {
  const int D = 1 << k, S = D * D;
  Mat out(size_t(S) * S, cd{0, 0});
  for (const Mat& m : ms)
    for (int i = 0; i < S; ++i)
      for (int j = 0; j < S; ++j) {
        const cd a = m[size_t(i & (D - 1)) * D + (j & (D - 1))];
        const cd b = cconj(m[size_t(i >> k) * D + (j >> k)]);
        out[size_t(i) * S + j] = out[size_t(i) * S + j] + a * b;
      }
  return out;
}

int to_bgates_dm(int n, const yh_op* ops, size_t nops, std::vector<BGate>& out) {
  for (size_t e = 0; e < nops; ++e) {
    const yh_op& op = ops[e];
    std::vector<Mat> ms;
    std::vector<int> q;
    if (op.elem == YH_ELEM_GATE) {
      Mat g;
      const int d = gate_matrix(op, g);
      if (d != (1 << op.nt)) return fail(YH_ERR_MATRIX_SIZE, "gate matrix size does not match targets");
      for (int k = 0; k < op.nt; ++k) q.push_back(op.targets[k]);
      for (int k = 0; k < op.nc; ++k) q.push_back(op.ctrls[k]);
      if (2 * int(q.size()) <= kBQ) ms.push_back(with_controls(g, op));
    } else if (op.elem == YH_ELEM_CHANNEL) {
      int d = 0;
      const int st = kraus_operators(op, ms, d);
      if (st != YH_OK) return fail(st, "invalid channel parameters");
      for (int k = 0; k < op.nt; ++k) q.push_back(op.targets[k]);
      if (2 * int(q.size()) > kBQ) ms.clear();
    } else {
      continue;
    }
    BGate g;
    g.op_index = e;
    const int k = int(q.size());
    g.nt = 2 * k;
    for (int j = 0; j < k; ++j) g.tb.push_back(2 * n - 1 - q[j]);
    for (int j = 0; j < k; ++j) g.tb.push_back(n - 1 - q[j]);
    for (int b : g.tb) g.mask |= uint64_t(1) << b;
    g.blockable = !ms.empty();
    if (g.blockable) g.m = superop(ms, k);
    out.push_back(std::move(g));
  }
  return YH_OK;
}

// FP64 block matrix of one block step (product of its gates, embedded in 5 bits)
void block_matrix(const BStep& s, const std::vector<BGate>& bg, std::vector<cd>& U) {
  auto local = [&](int bit) {
    for (int j = 0; j < kBQ; ++j)
      if (s.p[j] == bit) return j;
    return -1;
  };
  U.assign(size_t(kBD) * kBD, cd{0, 0});
  for (int r = 0; r < kBD; ++r) U[size_t(r) * kBD + r] = cd{1, 0};
  for (uint32_t gi : s.members) {
    const BGate& g = bg[gi];
    int lt[kBQ];
    for (int k = 0; k < g.nt; ++k) lt[k] = local(g.tb[k]);
    std::vector<std::pair<int, int>> lc;
    for (auto& c : g.ctl) lc.push_back({local(c.first), c.second});
    left_multiply(U, g, lt, lc);
  }
}

// real-split [[Ur, -Ui], [Ui, Ur]] in MFMA A-fragment order: entry
// (rt * 16 + ks) * 64 + lane holds real row 16 rt + lane % 16, column 4 ks + lane / 16
template <class R>
void to_fragments(const std::vector<cd>& U, R* o) {
  constexpr int RD = 2 * kBD;
  auto at = [&](int r, int c) -> R {
    const bool ri = r >= kBD, ci = c >= kBD;
    const cd u = U[size_t(r % kBD) * kBD + (c % kBD)];
    if (!ri && !ci) return R(u.re);
    if (!ri && ci) return R(-u.im);
    if (ri && !ci) return R(u.im);
    return R(u.re);
  };
  for (int rt = 0; rt < 4; ++rt)
    for (int ks = 0; ks < 16; ++ks)
      for (int l = 0; l < 64; ++l) o[(rt * 16 + ks) * 64 + l] = at(16 * rt + l % 16, 4 * ks + l / 16);
  (void)RD;
}

// compact A values for k_tblocks: value e of lane l is Ur (e < 16) or Ui (e >= 16)
// at row 16 (e % 16 / 8) + l % 16, column 4 (e % 8) + l / 16
template <class R>
void to_compact(const std::vector<cd>& U, R* o) {
  for (int e = 0; e < 32; ++e)
    for (int l = 0; l < 64; ++l) {
      const int r = 16 * ((e % 16) / 8) + l % 16, c = 4 * (e % 8) + l / 16;
      const cd u = U[size_t(r) * kBD + c];
      o[e * 64 + l] = R(e < 16 ? u.re : u.im);
    }
}

// the same for a 4-bit block, from the 32 x 32 matrix built with row bit 4 an
// identity bit (so the upper-left 16 x 16 corner is the block): value e of lane l
// is Ur (e < 4) or Ui (e >= 4) at row l % 16, column 4 (e % 4) + l / 16
template <class R>
void to_compact_small(const std::vector<cd>& U, R* o) {
  for (int e = 0; e < 8; ++e)
    for (int l = 0; l < 64; ++l) {
      const int r = l % 16, c = 4 * (e % 4) + l / 16;
      const cd u = U[size_t(r) * kBD + c];
      o[e * 64 + l] = R(e < 4 ? u.re : u.im);
    }
}

// Launch order of a block plan: gates run alone stay where they are; each run of
// consecutive blocks is split into passes of k_tblocks. A pass starts at the first
// block not yet placed and takes, in order, every later block whose bits keep the
// tile (bits 0..3 plus the pass's block bits) within 12 bits and that shares no bit
// with a block skipped before it, so blocks only move past blocks on disjoint bits.
struct PassItem {
  bool single = false;
  size_t step = 0;              // single: index into the plan
  std::vector<size_t> blocks;   // pass: plan indices, in application order
  uint64_t bits = 0;            // pass: tile bits (before padding)
};

uint64_t step_mask(const BStep& st) {
  uint64_t m = 0;
  for (int j = 0; j < (st.small ? kBQ - 1 : kBQ); ++j) m |= uint64_t(1) << st.p[j];
  return m;
}

// A block acting on fewer than 5 bits is an identity on the rest of its rows, so in
// a multi-block pass those rows may be any other bits of the tile: bits above the
// low run are taken first, which lets a block that acts on none of memory bits
// 0..3 read or write HBM directly. A block acting on at most 4 bits runs as a
// 16 x 16 block on p[0..3] (half the MFMA work); p[4] is then one more identity
// bit, used only to build its matrix.
void pad_in_tile(BStep& st, uint64_t tile) {
  const uint64_t sup = st.support;
  st.small = popc(sup) <= kBQ - 1;
  int np = 0;
  for (int b = 0; b < 64 && np < kBQ; ++b)
    if ((sup >> b) & 1) st.p[np++] = b;
  for (int b = 63; b >= kTLow && np < kBQ; --b)
    if (((tile & ~sup) >> b) & 1) st.p[np++] = b;
  for (int b = 0; b < kTLow && np < kBQ; ++b)
    if (((tile & ~sup) >> b) & 1) st.p[np++] = b;
}

// Consecutive blocks are packed greedily into passes, each block by its 5 rows
// (`by_support` false) or by the bits it acts on (true: its identity rows go to
// other tile bits, and blocks that meet only on identity rows commute).
std::vector<PassItem> group_passes(const std::vector<BStep>& steps, int nbits, bool by_support) {
  auto bits_of = [&](const BStep& st) { return by_support ? st.support : step_mask(st); };
  std::vector<PassItem> items;
  size_t i = 0;
  while (i < steps.size()) {
    if (!steps[i].block) {
      PassItem it;
      it.single = true;
      it.step = i++;
      items.push_back(it);
      continue;
    }
    size_t j = i;
    while (j < steps.size() && steps[j].block) ++j;
    std::vector<size_t> rem;
    for (size_t k = i; k < j; ++k) rem.push_back(k);
    while (!rem.empty()) {
      PassItem it;
      uint64_t cur = (uint64_t(1) << kTLow) - 1, blocked = 0;
      std::vector<size_t> left;
      for (size_t k : rem) {
        const uint64_t bm = bits_of(steps[k]);
        if (nbits >= kTT && !(bm & blocked) && popc(cur | bm) <= kTT && it.blocks.size() < size_t(kPassMax)) {
          cur |= bm;
          it.blocks.push_back(k);
        } else {
          blocked |= bm;
          left.push_back(k);
        }
      }
      if (it.blocks.empty()) {  // n < 12: one block per pass (k_block5)
        it.blocks.push_back(left.front());
        cur |= bits_of(steps[left.front()]);
        left.erase(left.begin());
      }
      it.bits = cur;
      items.push_back(std::move(it));
      rem.swap(left);
    }
    i = j;
  }
  return items;
}

// One launch of a block plan, host side: a gate on its own path, a block alone in
// its HBM pass (k_block5), or a multi-block pass (k_tblocks) with its blocks in
// application order, their identity rows placed in the tile, and the direct-in /
// direct-out forms it allows.
struct PassPlan {
  int kind = 0;               // 0 gate alone, 1 block alone, 2 multi-block pass
  size_t step = 0;            // kinds 0 and 1: index into the plan
  std::vector<BStep> blocks;  // kind 2
  uint64_t tile = 0;          // kind 2: the 12 tile bits
  bool din = false, dout = false;
};

int plan_passes(const std::vector<BStep>& steps, int nbits, std::vector<PassPlan>& out) {
  // of the two packings, the one with the lower modelled time: a pass of k >= 2
  // blocks costs about 0.46 + k block-times, a block alone in its pass 1.8
  auto cost = [&](const std::vector<PassItem>& v) {
    double c = 0;
    for (const PassItem& it : v) {
      if (it.single) continue;
      if (it.blocks.size() == 1) {
        c += 1.8;
        continue;
      }
      c += 0.46;
      for (size_t k : it.blocks) c += popc(steps[k].support) <= kBQ - 1 ? kSmallCost : 1.0;
    }
    return c;
  };
  std::vector<PassItem> items = group_passes(steps, nbits, false);
  {
    std::vector<PassItem> alt = group_passes(steps, nbits, true);
    if (cost(alt) < cost(items)) items.swap(alt);
  }
  out.clear();
  for (const PassItem& it : items) {
    PassPlan pp;
    if (it.single || it.blocks.size() == 1) {
      pp.kind = it.single ? 0 : 1;
      pp.step = it.single ? it.step : it.blocks[0];
      out.push_back(std::move(pp));
      continue;
    }
    // tile bits = the pass's bits padded with the lowest free bits to exactly 12
    uint64_t tb = it.bits;
    for (int b = 0; popc(tb) < kTT && b < nbits; ++b) tb |= uint64_t(1) << b;
    if (popc(tb) != kTT || nbits < kTT) return fail(YH_ERR_UNSUPPORTED, "block pass tile does not fit");
    if (it.blocks.size() > size_t(kPassMax)) return fail(YH_ERR_UNSUPPORTED, "block pass size out of range");
    pp.kind = 2;
    pp.tile = tb;
    std::vector<BStep> ps;  // the pass's blocks with their identity rows in the tile
    for (size_t k : it.blocks) {
      ps.push_back(steps[k]);
      pad_in_tile(ps.back(), tb);
    }
    // A block whose rows avoid memory bits 0..3 can take its operands from HBM
    // (first) or put its results to HBM (last) with whole 16-amplitude runs; such a
    // block is moved to the front / the end when it acts on no bit of the blocks it
    // moves past.
    std::vector<size_t> order(ps.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = k;
    const uint64_t low = (uint64_t(1) << kTLow) - 1;
    auto rows = [&](size_t k) { return step_mask(ps[order[k]]); };
    auto mask = [&](size_t k) { return ps[order[k]].support; };
    for (size_t j = 0; j < order.size() && !pp.din; ++j) {
      bool ok = !(rows(j) & low);
      for (size_t k = 0; k < j && ok; ++k) ok = !(mask(j) & mask(k));
      if (ok) {
        std::rotate(order.begin(), order.begin() + j, order.begin() + j + 1);
        pp.din = true;
      }
    }
    for (size_t j = order.size(); j-- > (pp.din ? 1u : 0u) && !pp.dout;) {
      bool ok = !(rows(j) & low);
      for (size_t k = j + 1; k < order.size() && ok; ++k) ok = !(mask(j) & mask(k));
      if (ok) {
        std::rotate(order.begin() + j, order.begin() + j + 1, order.end());
        pp.dout = true;
      }
    }
    for (size_t k : order) pp.blocks.push_back(std::move(ps[k]));
    out.push_back(std::move(pp));
  }
  return YH_OK;
}

// modelled time of a plan's launches, in 32 x 32 block-times: a block alone in its
// pass (k_block5, HBM-bound) 1.8, a multi-block pass 0.46 plus 1 per 32 x 32 and
// kSmallCost per 16 x 16 block, a gate on its own path 1
double plan_cost(const std::vector<BStep>& steps, int nbits) {
  std::vector<PassPlan> plan;
  if (plan_passes(steps, nbits, plan) != YH_OK) return 1e300;
  double c = 0;
  for (const PassPlan& pp : plan) {
    if (pp.kind == 0) c += 1.0;
    else if (pp.kind == 1) c += 1.8;
    else {
      c += 0.46;
      for (const BStep& b : pp.blocks) c += b.small ? kSmallCost : 1.0;
    }
  }
  return c;
}

}  // namespace

// Launches a block plan on a state of `nbits` bits; steps that are not blocks go
// to `single` (the gate's own path).
template <class R>
int run_block_plan(Workspace& w, void* state, int nbits, const std::vector<BStep>& steps,
                   const std::vector<BGate>& bg, const std::function<int(const BGate&)>& single) {
  using V = typename Cplx<R>::V;
  // persistent grids: as many workgroups as the CUs hold at once
  static int resident[2][2] = {{0, 0}, {0, 0}};
  int(&res)[2] = resident[sizeof(R) == 8 ? 0 : 1];
  if (res[0] == 0) {
    int dev = 0, per_cu = 0, per_cu_t = 0;
    hipDeviceProp_t prop;
    YH_HIP(hipGetDevice(&dev));
    YH_HIP(hipGetDeviceProperties(&prop, dev));
    YH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_block5<R>, 256, 0));
    YH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_t, k_tblocks<R, false, false, 2>, kTThreads, 0));
    res[0] = prop.multiProcessorCount * std::max(1, per_cu);
    res[1] = prop.multiProcessorCount * std::max(1, per_cu_t);
  }
  std::vector<PassPlan> plan;
  YH_TRY(plan_passes(steps, nbits, plan));
  // Block matrices are built on the host while the GPU runs the previous passes:
  // each pass's operands go to their own part of a pinned staging area and are
  // copied with an async copy ordered before its launch (nothing else writes the
  // area in this call).
  size_t total = 0;
  for (const PassPlan& pp : plan) total += pp.kind == 1 ? kFrag : pp.kind == 2 ? pp.blocks.size() * kCFrag : 0;
  YH_TRY(w.sync());
  YH_TRY(w.bpin.ensure(std::max<size_t>(1, total) * sizeof(R)));
  YH_TRY(w.tiles.ensure(std::max<size_t>(1, total) * sizeof(R)));
  R* hst = static_cast<R*>(w.bpin.p);
  R* dev = static_cast<R*>(w.tiles.p);
  V* s = static_cast<V*>(state);
  std::vector<cd> U;
  size_t off = 0;
  for (const PassPlan& pp : plan) {
    if (pp.kind == 0) {
      YH_TRY(single(bg[steps[pp.step].gate]));
      continue;
    }
    if (pp.kind == 1) {
      const BStep& st = steps[pp.step];
      block_matrix(st, bg, U);
      to_fragments<R>(U, hst + off);
      YH_HIP(hipMemcpyAsync(dev + off, hst + off, kFrag * sizeof(R), hipMemcpyHostToDevice, w.stream));
      BlockArgs a{};
      int sp[kBQ];
      std::memcpy(sp, st.p, sizeof(sp));
      std::sort(sp, sp + kBQ);
      for (int j = 0; j < kBQ; ++j) {
        a.spos[j] = sp[j];
        a.p[j] = st.p[j];
      }
      a.ncol = uint64_t(1) << (nbits - kBQ);
      a.nchunk = (a.ncol + kCols - 1) / kCols;
      const uint64_t grid = std::min<uint64_t>((a.nchunk + 3) / 4, uint64_t(res[0]));
      YH_TRY(w.time_begin(YH_K_FUSED));
      hipLaunchKernelGGL(k_block5<R>, dim3(unsigned(grid)), dim3(256), 0, w.stream, s, a, dev + off);
      YH_HIP(hipGetLastError());
      YH_TRY(w.time_end());
      off += kFrag;
      continue;
    }
    TBlockArgs a{};
    int pos_of[64];
    for (int b = 0, k = 0; b < nbits; ++b)
      if ((pp.tile >> b) & 1) {
        a.tbit[k] = uint8_t(b);
        pos_of[b] = k++;
      }
    a.nblk = int(pp.blocks.size());
    a.ntiles = uint64_t(1) << (nbits - kTT);
    // a pass of 16 x 16 blocks only takes two tiles per workgroup step
    bool all_small = true;
    for (const BStep& st : pp.blocks) all_small = all_small && st.small;
    const bool pair = all_small && a.ntiles >= 2;
    const size_t stride = pair ? kSFrag : kCFrag;
    for (size_t bi = 0; bi < pp.blocks.size(); ++bi) {
      const BStep& st = pp.blocks[bi];
      const int nq = st.small ? kBQ - 1 : kBQ;
      int qs[kBQ];
      for (int j = 0; j < kBQ; ++j) {
        qs[j] = j < nq ? pos_of[st.p[j]] : kNoBit;
        a.q[bi][j] = uint8_t(qs[j]);
      }
      std::sort(qs, qs + kBQ);
      for (int j = 0; j < kBQ; ++j) a.qs[bi][j] = uint8_t(qs[j]);
      block_matrix(st, bg, U);
      if (st.small) {
        a.small |= 1u << bi;
        to_compact_small<R>(U, hst + off + bi * stride);
      } else {
        to_compact<R>(U, hst + off + bi * stride);
      }
    }
    const bool din = pp.din, dout = pp.dout;
    const size_t sz = pp.blocks.size() * stride;
    YH_HIP(hipMemcpyAsync(dev + off, hst + off, sz * sizeof(R), hipMemcpyHostToDevice, w.stream));
    const uint64_t grid = std::min<uint64_t>(a.ntiles / (pair ? 2 : 1), uint64_t(res[1]));
    YH_TRY(w.time_begin(YH_K_TBLOCK));
    auto launch = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(unsigned(grid)), dim3(kTThreads), 0, w.stream, s, a, dev + off);
    };
    if (pair) {
      if (din && dout) launch(k_tblocks<R, true, true, 2>);
      else if (din) launch(k_tblocks<R, true, false, 2>);
      else if (dout) launch(k_tblocks<R, false, true, 2>);
      else launch(k_tblocks<R, false, false, 2>);
    } else {
      if (din && dout) launch(k_tblocks<R, true, true, 1>);
      else if (din) launch(k_tblocks<R, true, false, 1>);
      else if (dout) launch(k_tblocks<R, false, true, 1>);
      else launch(k_tblocks<R, false, false, 1>);
    }
    YH_HIP(hipGetLastError());
    YH_TRY(w.time_end());
    off += sz;
  }
  return w.sync();
}

template <class R>
int apply_blocks(Workspace& w, void* state, int nbits, const yh_op* ops, size_t nops) {
  if (nbits < kBQ) return apply_gates<R>(w, state, nbits, ops, nops);
  std::vector<BStep> steps;
  std::vector<BGate> bg;
  YH_TRY(plan_blocks(nbits, ops, nops, steps, bg));
  return run_block_plan<R>(w, state, nbits, steps, bg, [&](const BGate& g) {
    yh_op one = ops[g.op_index];
    if (one.kind == YH_GATE_X && one.nt > 1) {
      one.targets[0] = one.targets[g.sub];
      one.nt = 1;
    }
    return apply_gates<R>(w, state, nbits, &one, 1);
  });
}

int apply_dm_blocks(Workspace& w, void* rho, int n, const yh_op* ops, size_t nops,
                    const std::function<int(size_t)>& single_op) {
  std::vector<BGate> bg;
  YH_TRY(to_bgates_dm(n, ops, nops, bg));
  if (bg.empty()) return YH_OK;
  if (2 * n < kBQ) {  // too few bits for a block: every element on its own path
    for (const BGate& g : bg) YH_TRY(single_op(g.op_index));
    return YH_OK;
  }
  std::vector<BStep> steps;
  YH_TRY(plan_gates(bg, 2 * n, structure_sig(2 * n, ops, nops, 1), steps));
  return run_block_plan<double>(w, rho, 2 * n, steps, bg, [&](const BGate& g) { return single_op(g.op_index); });
}

template int apply_blocks<double>(Workspace&, void*, int, const yh_op*, size_t);
template int apply_blocks<float>(Workspace&, void*, int, const yh_op*, size_t);

// every step of the block plan: 5 memory bits (row bit j <-> p[j]) per block, or
// five -1 for a gate run as its own launch
int block_plan_bits(int nbits, const yh_op* ops, size_t nops, int32_t* bits, size_t cap, size_t* nsteps) {
  std::vector<BStep> steps;
  std::vector<BGate> bg;
  YH_TRY(plan_blocks(nbits, ops, nops, steps, bg));
  *nsteps = steps.size();
  if (bits)
    for (size_t i = 0; i < steps.size() && i < cap; ++i)
      for (int j = 0; j < kBQ; ++j) bits[i * kBQ + j] = steps[i].block ? steps[i].p[j] : -1;
  return YH_OK;
}

// out[0] blocks, out[1] gates (X split per target) in blocks, out[2] single launches
int block_plan_stats(int nbits, const yh_op* ops, size_t nops, int64_t out[3]) {
  std::vector<BStep> steps;
  std::vector<BGate> bg;
  YH_TRY(plan_blocks(nbits, ops, nops, steps, bg));
  out[0] = out[1] = out[2] = 0;
  for (auto& s : steps) {
    if (s.block) {
      out[0] += 1;
      out[1] += int64_t(s.members.size());
    } else {
      out[2] += 1;
    }
  }
  return YH_OK;
}

// Host-only launch plan of the block path. dm = 0: the state-vector plan of a
// circuit on nbits qubits; dm = 1: the superoperator plan of a density matrix of
// nbits qubits. out[0] launches of the block kernels (k_block5 + k_tblocks),
// out[1] multi-block passes, out[2] blocks in them, out[3] of those the 4-bit ones,
// out[4] multi-block passes with direct-in, out[5] with direct-out, out[6] gates
// run as their own launch.
int block_pass_stats(int nbits, const yh_op* ops, size_t nops, int dm, int64_t out[7]) {
  std::vector<BStep> steps;
  std::vector<BGate> bg;
  int nb = nbits;
  for (int k = 0; k < 7; ++k) out[k] = 0;
  if (dm) {
    YH_TRY(to_bgates_dm(nbits, ops, nops, bg));
    nb = 2 * nbits;
    if (bg.empty()) return YH_OK;
    if (nb < kBQ) {
      out[6] = int64_t(bg.size());
      return YH_OK;
    }
    YH_TRY(plan_gates(bg, nb, structure_sig(nb, ops, nops, 1), steps));
  } else {
    if (nbits < kBQ) return fail(YH_ERR_UNSUPPORTED, "fewer qubits than a block");
    YH_TRY(plan_blocks(nbits, ops, nops, steps, bg));
  }
  std::vector<PassPlan> plan;
  YH_TRY(plan_passes(steps, nb, plan));
  for (const PassPlan& pp : plan) {
    if (pp.kind == 0) {
      out[6] += 1;
      continue;
    }
    out[0] += 1;
    if (pp.kind == 2) {
      out[1] += 1;
      out[2] += int64_t(pp.blocks.size());
      for (const BStep& st : pp.blocks) out[3] += st.small;
      out[4] += pp.din;
      out[5] += pp.dout;
    }
  }
  return YH_OK;
}

}  // namespace yh
