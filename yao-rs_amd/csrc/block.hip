// block.hip — yh_apply_fused with max_fuse_qubits = 5: gates fused into 5-qubit
// blocks (32 x 32 complex matrices, multiplied in FP64 on the host), each applied
// in one HBM pass as a real-split MFMA contraction.
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
  uint64_t first_come() const {
    uint64_t Q = 0, B = 0;
    for (uint32_t i : rem) {
      const BGate& g = bg[i];
      if (!g.blockable || (g.mask & B) || popc(Q | g.mask) > kBQ)
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
    while (popc(Q) < kBQ) {
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
        if ((m & Q) && popc(m | Q) <= kBQ && (m & ~Q & ~cand)) {
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
  void run(int strategy, std::vector<BStep>& steps) {
    rem.resize(bg.size());
    for (size_t i = 0; i < bg.size(); ++i) rem[i] = uint32_t(i);
    std::vector<uint32_t> taken, best, tmp, rd;
    std::vector<char> done(bg.size(), 0);
    while (!rem.empty()) {
      const uint32_t head = rem.front();
      if (!bg[head].blockable) {
        BStep s;
        s.gate = head;
        steps.push_back(s);
        done[head] = 1;
      } else {
        std::vector<uint64_t> cands{first_come()};
        if (strategy == 1 || strategy == 3)
          for (int b = 0; b + kBQ <= nbits; ++b) cands.push_back(((uint64_t(1) << kBQ) - 1) << b);
        if (strategy == 2 || strategy == 3) {
          ready(rd, 16);
          for (uint32_t r : rd) cands.push_back(grow(bg[r].mask, tmp));
        }
        uint64_t bestQ = 0;
        best.clear();
        for (uint64_t Q : cands) {
          absorb(Q, taken);
          if (taken.size() > best.size()) {
            best.swap(taken);
            bestQ = Q;
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
        for (uint32_t i : best) done[i] = 1;
        steps.push_back(std::move(s));
      }
      size_t k = 0;
      for (uint32_t i : rem)
        if (!done[i]) rem[k++] = i;
      rem.resize(k);
    }
  }
};

// structural key of a circuit: everything but gate parameters and matrix values
uint64_t structure_key(int nbits, const yh_op* ops, size_t nops) {
  uint64_t h = 1469598103934665603ull ^ uint64_t(nbits);
  auto mix = [&](uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  for (size_t e = 0; e < nops; ++e) {
    const yh_op& o = ops[e];
    mix(uint64_t(o.elem) << 40 | uint64_t(uint32_t(o.kind)) << 8 | uint64_t(o.nt) << 4 | uint64_t(o.nc));
    for (int k = 0; k < o.nt && k < YH_MAX_TARGETS; ++k) mix(uint64_t(o.targets[k]));
    for (int k = 0; k < o.nc && k < YH_MAX_CTRLS; ++k) mix(uint64_t(o.ctrls[k]) << 1 | uint64_t(o.ctrl_vals[k] & 1));
  }
  return h ^ nops;
}

struct PlanCache {
  std::mutex mu;
  struct Entry {
    uint64_t key;
    int nbits;
    size_t nops;
    std::vector<BStep> steps;
  };
  std::vector<Entry> entries;  // most recent last, at most 8
};
PlanCache& plan_cache() {
  static PlanCache c;
  return c;
}

int plan_blocks(int nbits, const yh_op* ops, size_t nops, std::vector<BStep>& steps, std::vector<BGate>& bg) {
  YH_TRY(to_bgates(nbits, ops, nops, bg));
  const uint64_t key = structure_key(nbits, ops, nops);
  {
    PlanCache& c = plan_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    for (auto& e : c.entries)
      if (e.key == key && e.nbits == nbits && e.nops == nops) {
        steps = e.steps;
        return YH_OK;
      }
  }
  Planner pl(bg, nbits);
  std::vector<BStep> bestp;
  size_t best_blocks = SIZE_MAX;
  for (int strategy = 0; strategy < 4; ++strategy) {
    std::vector<BStep> cand;
    pl.run(strategy, cand);
    size_t nb = 0;
    for (auto& s : cand) nb += s.block;
    if (nb < best_blocks) {
      best_blocks = nb;
      bestp.swap(cand);
    }
  }
  steps = bestp;
  PlanCache& c = plan_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  c.entries.push_back({key, nbits, nops, std::move(bestp)});
  if (c.entries.size() > 8) c.entries.erase(c.entries.begin());
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

}  // namespace

template <class R>
int apply_blocks(Workspace& w, void* state, int nbits, const yh_op* ops, size_t nops) {
  using V = typename Cplx<R>::V;
  if (nbits < kBQ) return apply_gates<R>(w, state, nbits, ops, nops);
  std::vector<BStep> steps;
  std::vector<BGate> bg;
  YH_TRY(plan_blocks(nbits, ops, nops, steps, bg));
  size_t nblk = 0;
  for (auto& s : steps) nblk += s.block;
  // persistent grid: as many workgroups as the CUs hold at once
  static int resident[2] = {0, 0};
  int& res = resident[sizeof(R) == 8 ? 0 : 1];
  if (res == 0) {
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    YH_HIP(hipGetDevice(&dev));
    YH_HIP(hipGetDeviceProperties(&prop, dev));
    YH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_block5<R>, 256, 0));
    res = prop.multiProcessorCount * std::max(1, per_cu);
  }
  // Block matrices are built on the host while the GPU runs the previous blocks:
  // each goes to its own slot of a pinned staging area and is copied with an
  // async copy ordered before its launch (nothing else writes the area in this call).
  const size_t slot = kFrag * sizeof(R);
  YH_TRY(w.sync());
  YH_TRY(w.bpin.ensure(std::max<size_t>(1, nblk) * slot));
  YH_TRY(w.tiles.ensure(std::max<size_t>(1, nblk) * slot));
  unsigned char* hst = static_cast<unsigned char*>(w.bpin.p);
  unsigned char* dev = static_cast<unsigned char*>(w.tiles.p);
  V* s = static_cast<V*>(state);
  std::vector<cd> U;
  size_t bi = 0;
  for (const BStep& st : steps) {
    if (st.block) {
      block_matrix(st, bg, U);
      R* frag = reinterpret_cast<R*>(hst + bi * slot);
      to_fragments<R>(U, frag);
      YH_HIP(hipMemcpyAsync(dev + bi * slot, frag, slot, hipMemcpyHostToDevice, w.stream));
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
      const uint64_t grid = std::min<uint64_t>((a.nchunk + 3) / 4, uint64_t(res));
      YH_TRY(w.time_begin(YH_K_FUSED));
      hipLaunchKernelGGL(k_block5<R>, dim3(unsigned(grid)), dim3(256), 0, w.stream, s, a,
                         reinterpret_cast<const R*>(dev + bi * slot));
      YH_HIP(hipGetLastError());
      YH_TRY(w.time_end());
      ++bi;
    } else {
      const BGate& g = bg[st.gate];
      yh_op one = ops[g.op_index];
      if (one.kind == YH_GATE_X && one.nt > 1) {
        one.targets[0] = one.targets[g.sub];
        one.nt = 1;
      }
      YH_TRY(apply_gates<R>(w, state, nbits, &one, 1));
    }
  }
  return w.sync();
}

template int apply_blocks<double>(Workspace&, void*, int, const yh_op*, size_t);
template int apply_blocks<float>(Workspace&, void*, int, const yh_op*, size_t);

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

}  // namespace yh
