// fused.hip — host planner for yh_apply_fused: gates -> tile passes.
//
// Pass selection: the tile always holds memory bits 0..kTileLow-1; its kTileHigh
// high bits are the first high bits touched, in circuit order, by gates that can
// still run. A pass then takes, in circuit order, every gate whose qubits lie in
// the tile and whose earlier gates (on any of its qubits) are done or taken; any
// other gate blocks its qubits for the rest of the scan, so the dependency order
// of yao-rs's sequential apply_inplace (src/apply.rs:123-147) is kept.
// Grouping: gates of a pass are packed into groups of <= 4 tile qubits, again by
// a dependency-respecting scan, so a group takes a whole light-cone of gates.
// Gates on more than 2 targets or more than 4 qubits run as their own launch.
#include <algorithm>
#include <cstring>
#include <vector>

#include "host.h"
#include "kernels.h"
#include "runtime.h"
#include "tile.h"

namespace yh {

template <class R>
int launch_tile(typename Cplx<R>::V* s, int nbits, const TilePass& pass, const TileGroup* groups,
                const MicroGate* mg, hipStream_t st);

namespace {

struct TGate {
  size_t op_index = 0;
  int sub = 0;          // target index for X gates split per target
  bool tileable = false;
  uint64_t mask = 0;    // memory bits touched (targets and controls)
  int nt = 0, nc = 0;
  int tb[2] = {0, 0};   // target bits, gate-index order
  int cb[4] = {0, 0, 0, 0};
  int cv[4] = {0, 0, 0, 0};
  MicroGate proto{};    // kind and matrix
};

int popc(uint64_t x) { return __builtin_popcountll(x); }

int to_tgates(int nbits, const yh_op* ops, size_t nops, std::vector<TGate>& out) {
  for (size_t e = 0; e < nops; ++e) {
    const yh_op& op = ops[e];
    if (op.elem != YH_ELEM_GATE) continue;
    auto bit_of = [nbits](int loc) { return nbits - 1 - loc; };
    const int subs = op.kind == YH_GATE_X ? op.nt : 1;
    Mat m;
    if (op.kind != YH_GATE_X) {
      int d = gate_matrix(op, m);
      if (d != (1 << op.nt)) return fail(YH_ERR_MATRIX_SIZE, "gate matrix size does not match targets");
    }
    for (int sidx = 0; sidx < subs; ++sidx) {
      TGate g;
      g.op_index = e;
      g.sub = sidx;
      g.nc = op.nc;
      for (int k = 0; k < op.nc; ++k) {
        if (k < 4) {
          g.cb[k] = bit_of(op.ctrls[k]);
          g.cv[k] = op.ctrl_vals[k];
        }
        g.mask |= uint64_t(1) << bit_of(op.ctrls[k]);
      }
      if (op.kind == YH_GATE_X) {
        g.nt = 1;
        g.tb[0] = bit_of(op.targets[sidx]);
        g.proto.kind = MG_PERM1;
        const double xm[4][2] = {{0, 0}, {1, 0}, {1, 0}, {0, 0}};
        std::memcpy(g.proto.m, xm, sizeof(xm));
      } else {
        g.nt = op.nt;
        for (int k = 0; k < op.nt && k < 2; ++k) g.tb[k] = bit_of(op.targets[k]);
        const bool diag = op.kind != YH_GATE_SWAP && gate_is_diagonal(op);
        if (op.nt == 1 && diag) {
          g.proto.kind = MG_DIAG1;
          g.proto.m[0][0] = m[0].re;
          g.proto.m[0][1] = m[0].im;
          g.proto.m[1][0] = m[3].re;
          g.proto.m[1][1] = m[3].im;
        } else if (op.nt <= 2) {
          g.proto.kind = op.nt == 1 ? MG_DENSE1 : MG_DENSE2;
          for (size_t k = 0; k < m.size(); ++k) {
            g.proto.m[k][0] = m[k].re;
            g.proto.m[k][1] = m[k].im;
          }
        }
      }
      for (int k = 0; k < g.nt; ++k) g.mask |= uint64_t(1) << g.tb[k];
      if (op.kind != YH_GATE_X)
        for (int k = 0; k < op.nt; ++k) g.mask |= uint64_t(1) << bit_of(op.targets[k]);
      g.tileable = g.nt <= 2 && popc(g.mask) <= kGroupBits && g.nc <= 4;
      out.push_back(g);
    }
  }
  return YH_OK;
}

struct Step {
  bool tile = false;
  size_t op_index = 0;  // single launch
  TilePass pass{};
};

// Pack the gates `idx` (in circuit order, all inside the tile) into groups.
void make_groups(const std::vector<TGate>& tg, const std::vector<size_t>& idx, const int tpos[64],
                 std::vector<TileGroup>& groups, std::vector<MicroGate>& micro) {
  std::vector<bool> used(idx.size(), false);
  size_t left = idx.size();
  while (left) {
    // start from the first unused gate
    size_t s0 = 0;
    while (used[s0]) ++s0;
    uint64_t q = 0, blocked = 0;
    std::vector<size_t> members;
    for (size_t k = s0; k < idx.size(); ++k) {
      if (used[k]) continue;
      const TGate& g = tg[idx[k]];
      if (!(g.mask & blocked) && popc(q | g.mask) <= kGroupBits) {
        q |= g.mask;
        used[k] = true;
        members.push_back(idx[k]);
        --left;
      } else {
        blocked |= g.mask;
      }
    }
    // local qubit order: the group's bits ascending, padded with other tile qubits
    int qb[kGroupBits], nq = 0;
    for (int b = 0; b < 64 && nq < kGroupBits; ++b)
      if ((q >> b) & 1) qb[nq++] = b;
    TileGroup G{};
    bool usedpos[kTileBits] = {};
    for (int k = 0; k < nq; ++k) {
      G.q[k] = uint8_t(tpos[qb[k]]);
      usedpos[tpos[qb[k]]] = true;
    }
    for (int p = 0; nq < kGroupBits && p < kTileBits; ++p)
      if (!usedpos[p]) {
        G.q[nq] = uint8_t(p);
        qb[nq] = -1;
        ++nq;
      }
    auto local = [&](int bit) {
      for (int k = 0; k < kGroupBits; ++k)
        if (qb[k] == bit) return k;
      return -1;
    };
    G.first = int(micro.size());
    // last micro gate touching each local qubit; an uncontrolled 1-qubit gate that
    // follows an uncontrolled 1-qubit gate on the same qubit is multiplied into it
    // (FP64, on the host): gates in between act on other qubits and commute.
    int last_touch[kGroupBits] = {-1, -1, -1, -1};
    for (size_t gi : members) {
      const TGate& g = tg[gi];
      MicroGate mg = g.proto;
      mg.t0 = uint8_t(local(g.tb[0]));
      mg.t1 = uint8_t(g.nt == 2 ? local(g.tb[1]) : 0);
      mg.cmask = mg.cval = 0;
      for (int k = 0; k < g.nc; ++k) {
        mg.cmask |= uint8_t(1 << local(g.cb[k]));
        if (g.cv[k]) mg.cval |= uint8_t(1 << local(g.cb[k]));
      }
      const bool single = mg.kind != MG_DENSE2 && mg.cmask == 0;
      const int lt = single ? last_touch[mg.t0] : -1;
      if (lt >= 0 && micro[lt].kind != MG_DENSE2 && micro[lt].cmask == 0 && micro[lt].t0 == mg.t0) {
        MicroGate& p = micro[lt];
        cd P[4], Gm[4], M[4];
        auto dense = [](const MicroGate& x, cd out[4]) {
          if (x.kind == MG_DIAG1) {
            out[0] = cd{x.m[0][0], x.m[0][1]};
            out[1] = out[2] = cd{0, 0};
            out[3] = cd{x.m[1][0], x.m[1][1]};
          } else {
            for (int e = 0; e < 4; ++e) out[e] = cd{x.m[e][0], x.m[e][1]};
          }
        };
        dense(p, P);
        dense(mg, Gm);
        for (int r = 0; r < 2; ++r)
          for (int c = 0; c < 2; ++c) M[r * 2 + c] = Gm[r * 2] * P[c] + Gm[r * 2 + 1] * P[2 + c];
        const bool diag = M[1].re == 0 && M[1].im == 0 && M[2].re == 0 && M[2].im == 0;
        if (diag) {
          p.kind = MG_DIAG1;
          p.m[0][0] = M[0].re;
          p.m[0][1] = M[0].im;
          p.m[1][0] = M[3].re;
          p.m[1][1] = M[3].im;
        } else {
          p.kind = MG_DENSE1;
          for (int e = 0; e < 4; ++e) {
            p.m[e][0] = M[e].re;
            p.m[e][1] = M[e].im;
          }
        }
        continue;
      }
      micro.push_back(mg);
      const int idx = int(micro.size()) - 1;
      last_touch[mg.t0] = idx;
      if (mg.kind == MG_DENSE2) last_touch[mg.t1] = idx;
      for (int k = 0; k < kGroupBits; ++k)
        if ((mg.cmask >> k) & 1) last_touch[k] = idx;
    }
    // a group longer than the kernel's LDS window of micro gates runs as consecutive
    // groups on the same qubits (the gates keep their order)
    const int end = int(micro.size());
    for (int f = G.first; f < end; f += kTileMicroWin) {
      TileGroup part = G;
      part.first = f;
      part.count = std::min(kTileMicroWin, end - f);
      groups.push_back(part);
    }
  }
}

}  // namespace

// Plans the passes for `ops`; the tile steps' groups/micro gates are appended.
int plan_tiles(int nbits, const yh_op* ops, size_t nops, std::vector<Step>& steps,
               std::vector<TileGroup>& groups, std::vector<MicroGate>& micro, std::vector<TGate>& tg) {
  YH_TRY(to_tgates(nbits, ops, nops, tg));
  const size_t G = tg.size();
  const uint64_t lowmask = (uint64_t(1) << kTileLow) - 1;
  std::vector<bool> done(G, false);
  size_t head = 0;
  while (head < G) {
    if (done[head]) {
      ++head;
      continue;
    }
    if (!tg[head].tileable) {
      Step s;
      s.op_index = head;
      steps.push_back(s);
      done[head] = true;
      continue;
    }
    // choose the high bits
    uint64_t hset = 0, blocked = 0;
    for (size_t i = head; i < G; ++i) {
      if (done[i]) continue;
      const TGate& g = tg[i];
      if (!g.tileable || (g.mask & blocked)) {
        blocked |= g.mask;
        continue;
      }
      uint64_t u = hset | (g.mask & ~lowmask);
      if (popc(u) <= kTileHigh) hset = u;
      else blocked |= g.mask;
      if (popc(hset) == kTileHigh && ((blocked | ~(lowmask | hset)) == ~uint64_t(0))) break;
    }
    for (int b = kTileLow; b < nbits && popc(hset) < kTileHigh; ++b)
      if (!((hset >> b) & 1)) hset |= uint64_t(1) << b;
    const uint64_t T = lowmask | hset;
    // tile position of each memory bit
    int tpos[64];
    for (int b = 0; b < 64; ++b) tpos[b] = -1;
    for (int b = 0; b < kTileLow; ++b) tpos[b] = b;
    Step s;
    s.tile = true;
    int k = 0;
    for (int b = kTileLow; b < nbits; ++b)
      if ((hset >> b) & 1) {
        s.pass.hbits[k] = uint8_t(b);
        tpos[b] = kTileLow + k;
        ++k;
      }
    // take the gates
    std::vector<size_t> idx;
    blocked = 0;
    for (size_t i = head; i < G; ++i) {
      if (done[i]) continue;
      const TGate& g = tg[i];
      if (g.tileable && !(g.mask & ~T) && !(g.mask & blocked)) {
        idx.push_back(i);
        done[i] = true;
      } else {
        blocked |= g.mask;
        if ((blocked & T) == T) break;
      }
    }
    s.pass.first_group = int(groups.size());
    make_groups(tg, idx, tpos, groups, micro);
    s.pass.ngroups = int(groups.size()) - s.pass.first_group;
    s.pass.mg_end = int(micro.size());
    steps.push_back(s);
  }
  return YH_OK;
}

template <class R>
int apply_tiled(Workspace& w, void* state, int nbits, const yh_op* ops, size_t nops) {
  using V = typename Cplx<R>::V;
  std::vector<Step> steps;
  std::vector<TileGroup> groups;
  std::vector<MicroGate> micro;
  std::vector<TGate> tg;
  YH_TRY(plan_tiles(nbits, ops, nops, steps, groups, micro, tg));
  // one upload: groups then micro gates
  const size_t gbytes = (groups.size() * sizeof(TileGroup) + 255) & ~size_t(255);
  std::vector<unsigned char> blob(gbytes + micro.size() * sizeof(MicroGate));
  if (!groups.empty()) std::memcpy(blob.data(), groups.data(), groups.size() * sizeof(TileGroup));
  if (!micro.empty()) std::memcpy(blob.data() + gbytes, micro.data(), micro.size() * sizeof(MicroGate));
  YH_TRY(w.upload(blob.data(), blob.size(), w.tiles));
  const TileGroup* dg = static_cast<const TileGroup*>(w.tiles.p);
  const MicroGate* dm = reinterpret_cast<const MicroGate*>(static_cast<const unsigned char*>(w.tiles.p) + gbytes);
  V* s = static_cast<V*>(state);
  for (const Step& st : steps) {
    if (st.tile) {
      YH_TRY(w.time_begin(YH_K_TILE));
      YH_TRY(launch_tile<R>(s, nbits, st.pass, dg, dm, w.stream));
      YH_TRY(w.time_end());
    } else {
      // a gate outside the tile form: the per-gate path (one launch per gate)
      const TGate& g = tg[st.op_index];
      yh_op one = ops[g.op_index];
      if (one.kind == YH_GATE_X && one.nt > 1) {
        one.targets[0] = one.targets[g.sub];
        one.nt = 1;
      }
      YH_TRY(w.sync());
      YH_TRY(apply_gates<R>(w, state, nbits, &one, 1));
    }
  }
  return w.sync();
}

template int apply_tiled<double>(Workspace&, void*, int, const yh_op*, size_t);
template int apply_tiled<float>(Workspace&, void*, int, const yh_op*, size_t);

// Plan statistics for tests and tuning: number of tile passes, groups, gates
// placed in tiles, and single launches.
int plan_stats(int nbits, const yh_op* ops, size_t nops, int64_t out[4]) {
  std::vector<Step> steps;
  std::vector<TileGroup> groups;
  std::vector<MicroGate> micro;
  std::vector<TGate> tg;
  YH_TRY(plan_tiles(nbits, ops, nops, steps, groups, micro, tg));
  out[0] = out[3] = 0;
  for (auto& s : steps) (s.tile ? out[0] : out[3]) += 1;
  out[1] = int64_t(groups.size());
  out[2] = int64_t(tg.size()) - out[3];  // gates (X split per target) placed in tiles
  return YH_OK;
}

}  // namespace yh
