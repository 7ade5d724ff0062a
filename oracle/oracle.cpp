// oracle.cpp — CPU restatement of the yao-rs dense simulation path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP backend: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
// product library never links or calls it.
//
// Each function follows the yao-rs source it cites (paths relative to the yao-rs
// repository). Arithmetic is kept in the reference's order; the file is compiled
// with -ffp-contract=off because Rust does not fuse multiply-adds, so results are
// bit-identical to the reference wherever libm agrees (glibc on both sides).
//
// Pinning (what checks each part, tests/test_oracle.py):
//   instructs / dispatch  : tests/golden/instruct.json (53 Yao.jl cases) and
//                           apply.json (68 qubit cases), plus a brute-force dense
//                           embedding on <= 8 qubits;
//   gate matrices         : tests/golden/gates.json (49 entries);
//   kraus / superop       : noise.json "kraus"; DM evolution: noise.json "dm_cases";
//   expect_dm             : noise.json "expectations" and the F5 bit-reversal identity;
//   probs                 : measure.json (11 cases);
//   itercontrol           : bitbasis/src/iterate_control.rs:166-219 index lists;
//   sampling              : cumsum boundary cases (r exactly at a partial sum);
//   collapse/reset/remove : closed forms on GHZ / product states.
//
// Threads: orc_set_threads(k) lets the loops over independent amplitude groups of
// the instructs (and the density matrix's branch accumulation and outer product)
// run on k OpenMP threads. Each group is updated by the same arithmetic in the same
// order, so results do not depend on k; the default is one thread (the
// reference is single-threaded), and the CPU baseline is timed that way.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>

#include "../include/yaohip.h"

namespace {

struct C {
  double re, im;
};
inline C cx(double r, double i) { return C{r, i}; }
// num-complex 0.4 Mul: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re)
inline C operator*(C a, C b) { return C{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C operator+(C a, C b) { return C{a.re + b.re, a.im + b.im}; }
inline C operator-(C a, C b) { return C{a.re - b.re, a.im - b.im}; }
inline C& operator+=(C& a, C b) { a = a + b; return a; }
inline C& operator*=(C& a, C b) { a = a * b; return a; }
inline C conj(C a) { return C{a.re, -a.im}; }
inline double norm_sqr(C a) { return a.re * a.re + a.im * a.im; }
inline C from_polar(double r, double t) { return C{r * std::cos(t), r * std::sin(t)}; }
inline C divr(C a, double d) { return C{a.re / d, a.im / d}; }

using Vec = std::vector<C>;
using usize = uint64_t;

int g_threads = 1;                        // orc_set_threads
constexpr usize kParMin = usize(1) << 14;  // smaller loops stay on one thread
#define ORC_PAR_REGION(n) if (g_threads > 1 && (n) >= kParMin) num_threads(g_threads)
#define ORC_PAR(n) ORC_PAR_REGION(n) schedule(static)

// ---------------- bitbasis (bitbasis/src/bit_ops.rs, iterate_control.rs) ----------------
inline usize indicator(usize k) { return usize(1) << k; }
inline usize low_mask(usize w) { return w == 0 ? 0 : (w >= 64 ? ~usize(0) : (usize(1) << w) - 1); }
inline usize bmask_range(usize start, usize stop) {  // bit_ops.rs:15-23
  if (stop <= start) return 0;
  return low_mask(stop - start) << start;
}

struct IterControl {  // iterate_control.rs:12-56
  usize n = 0, base = 0;
  std::vector<usize> masks, factors;
  usize get(usize k) const {
    usize out = base;
    for (size_t i = 0; i < masks.size(); ++i) out += (k & masks[i]) * factors[i];
    return out;
  }
};

// group_shift (iterate_control.rs:108-142)
void group_shift(usize nbits, std::vector<usize> positions, std::vector<usize>& masks,
                 std::vector<usize>& factors) {
  std::sort(positions.begin(), positions.end());
  usize previous = 0, free_index = 0;
  for (usize position : positions) {
    position += 1;
    if (position != previous + 1) {
      factors.push_back(usize(1) << (previous - free_index));
      usize gap = position - previous - 1;
      masks.push_back(bmask_range(free_index, free_index + gap));
      free_index += gap;
    }
    previous = position;
  }
  usize nfree = nbits - positions.size();
  if (free_index != nfree) {
    factors.push_back(usize(1) << (previous - free_index));
    masks.push_back(bmask_range(free_index, nfree));
  }
}

// itercontrol (iterate_control.rs:82-106)
IterControl itercontrol(usize nbits, const std::vector<usize>& positions,
                        const std::vector<usize>& configs) {
  IterControl ic;
  for (size_t i = 0; i < positions.size(); ++i)
    if (configs[i] == 1) ic.base |= indicator(positions[i]);
  group_shift(nbits, positions, ic.masks, ic.factors);
  ic.n = usize(1) << (nbits - positions.size());
  return ic;
}

// controller (iterate_control.rs:144-159)
struct Controller {
  usize mask = 0, target = 0;
  bool operator()(usize b) const { return (b & mask) == target; }
};
Controller controller(const std::vector<usize>& cbits, const std::vector<usize>& cvals) {
  Controller c;
  for (size_t i = 0; i < cbits.size(); ++i) {
    c.mask |= indicator(cbits[i]);
    if (cvals[i] == 1) c.target |= indicator(cbits[i]);
  }
  return c;
}

// ---------------- instructs (src/instruct_qubit.rs) ----------------
inline usize loc_to_bit(usize nbits, usize loc) { return nbits - 1 - loc; }  // :15-17

inline void u1rows(C* s, usize i, usize j, C a, C b, C c, C d) {  // :23-36
  C w = s[i], v = s[j];
  s[i] = a * w + b * v;
  s[j] = c * w + d * v;
}

template <class F>
void for_each_single_control_base(usize total, usize target_bit, usize ctrl_bit, usize ctrl_val,
                                  F f) {  // :39-79
  if (target_bit < ctrl_bit) {
    usize low_span = usize(1) << target_bit;
    usize middle_iters = usize(1) << (ctrl_bit - target_bit - 1);
    usize outer_step = usize(1) << (ctrl_bit + 1);
    usize ctrl_offset = ctrl_val << ctrl_bit;
#pragma omp parallel for collapse(3) ORC_PAR(total)
    for (usize outer = 0; outer < total; outer += outer_step) {
      for (usize m = 0; m < middle_iters; ++m) {
        for (usize low = 0; low < low_span; ++low) f(outer + ctrl_offset + (m << (target_bit + 1)) + low);
      }
    }
  } else {
    usize low_span = usize(1) << ctrl_bit;
    usize middle_iters = usize(1) << (target_bit - ctrl_bit - 1);
    usize outer_step = usize(1) << (target_bit + 1);
    usize ctrl_offset = ctrl_val << ctrl_bit;
#pragma omp parallel for collapse(3) ORC_PAR(total)
    for (usize outer = 0; outer < total; outer += outer_step) {
      for (usize m = 0; m < middle_iters; ++m) {
        for (usize low = 0; low < low_span; ++low) f(outer + (m << (ctrl_bit + 1)) + ctrl_offset + low);
      }
    }
  }
}

void instruct_x(C* s, usize nbits, usize loc) {  // :82-92
  usize mask = indicator(loc_to_bit(nbits, loc));
  usize dim = usize(1) << nbits;
#pragma omp parallel for ORC_PAR(dim)
  for (usize b = 0; b < dim; ++b)
    if ((b & mask) == 0) std::swap(s[b], s[b ^ mask]);
}

std::vector<usize> to_bits(usize nbits, const std::vector<usize>& locs) {
  std::vector<usize> r;
  for (usize l : locs) r.push_back(loc_to_bit(nbits, l));
  return r;
}

void instruct_x_controlled(C* s, usize nbits, usize loc, const std::vector<usize>& cl,
                           const std::vector<usize>& cb) {  // :95-126
  usize bit = loc_to_bit(nbits, loc);
  usize mask = indicator(bit);
  usize dim = usize(1) << nbits;
  if (cl.size() == 1) {
    for_each_single_control_base(dim, bit, loc_to_bit(nbits, cl[0]), cb[0],
                                 [&](usize base) { std::swap(s[base], s[base + mask]); });
    return;
  }
  Controller ctrl = controller(to_bits(nbits, cl), cb);
#pragma omp parallel for ORC_PAR(dim)
  for (usize b = 0; b < dim; ++b)
    if ((b & mask) == 0 && ctrl(b)) std::swap(s[b], s[b ^ mask]);
}

void instruct_1q(C* s, usize len, usize loc, C a, C b, C c, C d) {  // :138-158
  usize nbits = __builtin_ctzll(len);
  usize bit = loc_to_bit(nbits, loc);
  usize step1 = usize(1) << bit, step2 = usize(1) << (bit + 1);
#pragma omp parallel for collapse(2) ORC_PAR(len)
  for (usize j = 0; j < len; j += step2)
    for (usize t = 0; t < step1; ++t) u1rows(s, j + t, j + t + step1, a, b, c, d);
}

void instruct_1q_diag(C* s, usize len, usize loc, C d0, C d1) {  // :170-184
  usize nbits = __builtin_ctzll(len);
  usize bit = loc_to_bit(nbits, loc);
  usize step1 = usize(1) << bit, step2 = usize(1) << (bit + 1);
#pragma omp parallel for collapse(2) ORC_PAR(len)
  for (usize j = 0; j < len; j += step2)
    for (usize t = 0; t < step1; ++t) {
      s[j + t] *= d0;
      s[j + t + step1] *= d1;
    }
}

void unrows(C* s, const usize* inds, usize d, const C* gate) {  // :189-206
  C old[4];
  for (usize k = 0; k < d; ++k) old[k] = s[inds[k]];
  for (usize i = 0; i < d; ++i) {
    C acc = cx(0, 0);
    for (usize j = 0; j < d; ++j) acc += gate[i * d + j] * old[j];
    s[inds[i]] = acc;
  }
}

// sort_unitary_2q (:215-231)
void sort_unitary_2q(const C* gate, usize b0, usize b1, C* out, usize sorted[2]) {
  if (b0 < b1) {
    std::memcpy(out, gate, 16 * sizeof(C));
    sorted[0] = b0;
    sorted[1] = b1;
    return;
  }
  sorted[0] = std::min(b0, b1);
  sorted[1] = std::max(b0, b1);
  const usize perm[4] = {0, 2, 1, 3};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[perm[i] * 4 + perm[j]] = gate[i * 4 + j];
}

std::vector<usize> locs_raw_2(usize nbits, const usize sorted[2]) {
  std::vector<usize> nt;
  for (usize b = 0; b < nbits; ++b)
    if (b != sorted[0] && b != sorted[1]) nt.push_back(b);
  IterControl ic = itercontrol(nbits, nt, std::vector<usize>(nt.size(), 0));
  std::vector<usize> r;
  for (usize k = 0; k < ic.n; ++k) r.push_back(ic.get(k));
  return r;
}

void instruct_2q(C* s, usize nbits, const std::vector<usize>& locs, const C* gate) {  // :238-265
  C g[16];
  usize sb[2];
  sort_unitary_2q(gate, loc_to_bit(nbits, locs[0]), loc_to_bit(nbits, locs[1]), g, sb);
  std::vector<usize> raw = locs_raw_2(nbits, sb);
  IterControl ic = itercontrol(nbits, {sb[0], sb[1]}, {0, 0});
#pragma omp parallel for ORC_PAR(ic.n)
  for (usize k = 0; k < ic.n; ++k) {
    usize base = ic.get(k);
    usize inds[4] = {raw[0] + base, raw[1] + base, raw[2] + base, raw[3] + base};
    unrows(s, inds, 4, g);
  }
}

void instruct_2q_diag(C* s, usize nbits, const std::vector<usize>& locs, const C diag_in[4]) {
  // :270-302
  usize b0 = loc_to_bit(nbits, locs[0]), b1 = loc_to_bit(nbits, locs[1]);
  usize sb[2] = {b0, b1};
  C diag[4] = {diag_in[0], diag_in[1], diag_in[2], diag_in[3]};
  if (sb[0] > sb[1]) {
    std::swap(sb[0], sb[1]);
    diag[1] = diag_in[2];
    diag[2] = diag_in[1];
  }
  std::vector<usize> raw = locs_raw_2(nbits, sb);
  IterControl ic = itercontrol(nbits, {sb[0], sb[1]}, {0, 0});
#pragma omp parallel for ORC_PAR(ic.n)
  for (usize k = 0; k < ic.n; ++k) {
    usize base = ic.get(k);
    for (int q = 0; q < 4; ++q) s[raw[q] + base] *= diag[q];
  }
}

void instruct_swap(C* s, usize nbits, const std::vector<usize>& locs) {  // :305-320
  usize m0 = indicator(loc_to_bit(nbits, locs[0])), m1 = indicator(loc_to_bit(nbits, locs[1]));
  usize sm = m0 | m1, dim = usize(1) << nbits;
#pragma omp parallel for ORC_PAR(dim)
  for (usize b = 0; b < dim; ++b)
    if ((b & m0) == 0 && (b & m1) == m1) std::swap(s[b], s[b ^ sm]);
}

void instruct_nq(C* s, usize nbits, const std::vector<usize>& locs, const C* gate,
                 const std::vector<usize>& cl, const std::vector<usize>& cb) {  // :323-372
  usize nl = locs.size(), gd = usize(1) << nl;
  std::vector<usize> tb = to_bits(nbits, locs);
  std::vector<usize> locked = to_bits(nbits, cl);
  locked.insert(locked.end(), tb.begin(), tb.end());
  std::vector<usize> vals = cb;
  vals.insert(vals.end(), nl, 0);
  std::vector<usize> raw(gd);
  for (usize cfg = 0; cfg < gd; ++cfg) {
    usize acc = 0;
    for (usize idx = 0; idx < nl; ++idx) acc |= ((cfg >> idx) & 1) << tb[idx];
    raw[cfg] = acc;
  }
  IterControl ic = itercontrol(nbits, locked, vals);
#pragma omp parallel ORC_PAR_REGION(ic.n)
  {
    std::vector<C> temp(gd);
#pragma omp for schedule(static)
    for (usize k = 0; k < ic.n; ++k) {
      usize base = ic.get(k);
      for (usize row = 0; row < gd; ++row) {
        temp[row] = cx(0, 0);
        for (usize col = 0; col < gd; ++col) temp[row] += gate[row * gd + col] * s[base + raw[col]];
      }
      for (usize r = 0; r < gd; ++r) s[base + raw[r]] = temp[r];
    }
  }
}

void instruct_1q_controlled(C* s, usize nbits, usize loc, C a, C b, C c, C d,
                            const std::vector<usize>& cl, const std::vector<usize>& cb) {
  // :385-423
  usize tbit = loc_to_bit(nbits, loc), step = usize(1) << tbit, total = usize(1) << nbits;
  if (cl.size() == 1) {
    for_each_single_control_base(total, tbit, loc_to_bit(nbits, cl[0]), cb[0],
                                 [&](usize base) { u1rows(s, base, base + step, a, b, c, d); });
    return;
  }
  std::vector<usize> locked = to_bits(nbits, cl);
  locked.push_back(tbit);
  std::vector<usize> vals = cb;
  vals.push_back(0);
  IterControl ic = itercontrol(nbits, locked, vals);
#pragma omp parallel for ORC_PAR(ic.n)
  for (usize k = 0; k < ic.n; ++k) {
    usize base = ic.get(k);
    u1rows(s, base, base + step, a, b, c, d);
  }
}

void instruct_2q_controlled(C* s, usize nbits, const std::vector<usize>& locs, const C* gate,
                            const std::vector<usize>& cl, const std::vector<usize>& cb) {
  // :428-470
  C g[16];
  usize sb[2];
  sort_unitary_2q(gate, loc_to_bit(nbits, locs[0]), loc_to_bit(nbits, locs[1]), g, sb);
  std::vector<usize> raw = locs_raw_2(nbits, sb);
  std::vector<usize> locked = to_bits(nbits, cl);
  locked.push_back(sb[0]);
  locked.push_back(sb[1]);
  std::vector<usize> vals = cb;
  vals.push_back(0);
  vals.push_back(0);
  IterControl ic = itercontrol(nbits, locked, vals);
#pragma omp parallel for ORC_PAR(ic.n)
  for (usize k = 0; k < ic.n; ++k) {
    usize base = ic.get(k);
    usize inds[4] = {raw[0] + base, raw[1] + base, raw[2] + base, raw[3] + base};
    unrows(s, inds, 4, g);
  }
}

void instruct_1q_diag_controlled(C* s, usize nbits, usize loc, C d0, C d1,
                                 const std::vector<usize>& cl, const std::vector<usize>& cb) {
  // :475-511
  usize tbit = loc_to_bit(nbits, loc), step = usize(1) << tbit, total = usize(1) << nbits;
  if (cl.size() == 1) {
    for_each_single_control_base(total, tbit, loc_to_bit(nbits, cl[0]), cb[0], [&](usize base) {
      s[base] *= d0;
      s[base + step] *= d1;
    });
    return;
  }
  std::vector<usize> locked = to_bits(nbits, cl);
  locked.push_back(tbit);
  std::vector<usize> vals = cb;
  vals.push_back(0);
  IterControl ic = itercontrol(nbits, locked, vals);
#pragma omp parallel for ORC_PAR(ic.n)
  for (usize k = 0; k < ic.n; ++k) {
    usize base = ic.get(k);
    s[base] *= d0;
    s[base + step] *= d1;
  }
}

void instruct_2q_diag_controlled(C* s, usize nbits, const std::vector<usize>& locs,
                                 const C diag_in[4], const std::vector<usize>& cl,
                                 const std::vector<usize>& cb) {  // :516-555
  usize sb[2] = {loc_to_bit(nbits, locs[0]), loc_to_bit(nbits, locs[1])};
  C diag[4] = {diag_in[0], diag_in[1], diag_in[2], diag_in[3]};
  if (sb[0] > sb[1]) {
    std::swap(sb[0], sb[1]);
    diag[1] = diag_in[2];
    diag[2] = diag_in[1];
  }
  std::vector<usize> raw = locs_raw_2(nbits, sb);
  std::vector<usize> locked = to_bits(nbits, cl);
  locked.push_back(sb[0]);
  locked.push_back(sb[1]);
  std::vector<usize> vals = cb;
  vals.push_back(0);
  vals.push_back(0);
  IterControl ic = itercontrol(nbits, locked, vals);
#pragma omp parallel for ORC_PAR(ic.n)
  for (usize k = 0; k < ic.n; ++k) {
    usize base = ic.get(k);
    for (int q = 0; q < 4; ++q) s[raw[q] + base] *= diag[q];
  }
}

// ---------------- gate matrices (src/gate.rs:284-393) ----------------
const double FRAC_1_SQRT_2 = 0.70710678118654752440;
const double FRAC_PI_4 = 0.78539816339744830962;

// Returns dim (2 or 4) and writes dim*dim entries; Custom gates copy their matrix.
int gate_matrix(const yh_op& op, std::vector<C>& m) {
  C zero = cx(0, 0), one = cx(1, 0), neg_one = cx(-1, 0), i = cx(0, 1), neg_i = cx(0, -1);
  const double* p = op.params;
  switch (op.kind) {
    case YH_GATE_X: m = {zero, one, one, zero}; return 2;
    case YH_GATE_Y: m = {zero, neg_i, i, zero}; return 2;
    case YH_GATE_Z: m = {one, zero, zero, neg_one}; return 2;
    case YH_GATE_H: {
      C s = cx(FRAC_1_SQRT_2, 0), ns = cx(-FRAC_1_SQRT_2, 0);
      m = {s, s, s, ns};
      return 2;
    }
    case YH_GATE_S: m = {one, zero, zero, i}; return 2;
    case YH_GATE_T: m = {one, zero, zero, from_polar(1.0, FRAC_PI_4)}; return 2;
    case YH_GATE_SWAP:
      m.assign(16, zero);
      m[0] = one; m[1 * 4 + 2] = one; m[2 * 4 + 1] = one; m[15] = one;
      return 4;
    case YH_GATE_PHASE: m = {one, zero, zero, from_polar(1.0, p[0])}; return 2;
    case YH_GATE_RX: {
      C c = cx(std::cos(p[0] / 2.0), 0), nis = cx(0, -std::sin(p[0] / 2.0));
      m = {c, nis, nis, c};
      return 2;
    }
    case YH_GATE_RY: {
      C c = cx(std::cos(p[0] / 2.0), 0), s = cx(std::sin(p[0] / 2.0), 0),
        ns = cx(-std::sin(p[0] / 2.0), 0);
      m = {c, ns, s, c};
      return 2;
    }
    case YH_GATE_RZ: m = {from_polar(1.0, -p[0] / 2.0), zero, zero, from_polar(1.0, p[0] / 2.0)};
      return 2;
    case YH_GATE_SQRTX: {
      C f = cx(0.5, 0.5);
      m = {f * one, f * neg_i, f * neg_i, f * one};
      return 2;
    }
    case YH_GATE_SQRTY: {
      C f = cx(0.5, 0.5);
      m = {f * one, f * neg_one, f * one, f * one};
      return 2;
    }
    case YH_GATE_SQRTW: {
      C cv = cx(FRAC_1_SQRT_2, 0), nis = cx(0, -FRAC_1_SQRT_2);
      C g01 = cx(FRAC_1_SQRT_2, -FRAC_1_SQRT_2), g10 = cx(FRAC_1_SQRT_2, FRAC_1_SQRT_2);
      m = {cv, nis * g01, nis * g10, cv};
      return 2;
    }
    case YH_GATE_ISWAP:
      m.assign(16, zero);
      m[0] = one; m[1 * 4 + 2] = i; m[2 * 4 + 1] = i; m[15] = one;
      return 4;
    case YH_GATE_FSIM: {
      m.assign(16, zero);
      C ct = cx(std::cos(p[0]), 0), nist = cx(0, -std::sin(p[0]));
      m[0] = one; m[5] = ct; m[6] = nist; m[9] = nist; m[10] = ct;
      m[15] = from_polar(1.0, -p[1]);
      return 4;
    }
    case YH_GATE_CUSTOM: {
      int d = op.dim;
      m.resize(size_t(d) * d);
      for (int k = 0; k < d * d; ++k) m[k] = cx(op.matrix[2 * k], op.matrix[2 * k + 1]);
      return d;
    }
  }
  return -1;
}

bool gate_is_diagonal(const yh_op& op) {  // gate.rs:98-104
  switch (op.kind) {
    case YH_GATE_Z: case YH_GATE_S: case YH_GATE_T: case YH_GATE_PHASE: case YH_GATE_RZ:
      return true;
    case YH_GATE_CUSTOM: return op.is_diag != 0;
    default: return false;
  }
}

// ---------------- dispatch (src/apply.rs:7-120) ----------------
void dispatch(usize nbits, C* s, const yh_op& op) {
  bool has_c = op.nc > 0;
  std::vector<usize> cl, cb, tl;
  for (int k = 0; k < op.nc; ++k) {
    cl.push_back(op.ctrls[k]);
    cb.push_back(op.ctrl_vals[k] ? 1 : 0);
  }
  for (int k = 0; k < op.nt; ++k) tl.push_back(op.targets[k]);
  usize len = usize(1) << nbits;
  std::vector<C> m;
  if (op.kind == YH_GATE_X) {
    for (usize l : tl) {
      if (!has_c) instruct_x(s, nbits, l);
      else instruct_x_controlled(s, nbits, l, cl, cb);
    }
    return;
  }
  if (op.kind == YH_GATE_SWAP && !has_c) {
    instruct_swap(s, nbits, tl);
    return;
  }
  int d = gate_matrix(op, m);
  if (op.kind == YH_GATE_SWAP) {
    instruct_2q_controlled(s, nbits, tl, m.data(), cl, cb);
    return;
  }
  bool diag = gate_is_diagonal(op);
  if (diag && tl.size() == 1) {
    C d0 = m[0], d1 = m[d + 1];
    if (has_c) instruct_1q_diag_controlled(s, nbits, tl[0], d0, d1, cl, cb);
    else instruct_1q_diag(s, len, tl[0], d0, d1);
    return;
  }
  if (diag && tl.size() == 2) {
    C dg[4] = {m[0], m[5], m[10], m[15]};
    if (has_c) instruct_2q_diag_controlled(s, nbits, tl, dg, cl, cb);
    else instruct_2q_diag(s, nbits, tl, dg);
    return;
  }
  if (tl.size() == 1) {
    if (has_c) instruct_1q_controlled(s, nbits, tl[0], m[0], m[1], m[2], m[3], cl, cb);
    else instruct_1q(s, len, tl[0], m[0], m[1], m[2], m[3]);
    return;
  }
  if (tl.size() == 2) {
    if (has_c) instruct_2q_controlled(s, nbits, tl, m.data(), cl, cb);
    else instruct_2q(s, nbits, tl, m.data());
    return;
  }
  instruct_nq(s, nbits, tl, m.data(), cl, cb);
}

// ---------------- noise (src/noise.rs) ----------------
using Mat = std::vector<C>;  // row-major 2x2 (or d x d)

bool in01(double p) { return p >= 0.0 && p <= 1.0; }

int validate_channel(const yh_op& op) {  // noise.rs:90-176
  const double* p = op.params;
  switch (op.kind) {
    case YH_CH_BITFLIP: case YH_CH_PHASEFLIP: case YH_CH_DEPOLARIZING:
      return in01(p[0]) ? 0 : YH_ERR_CHANNEL_PARAM;
    case YH_CH_PAULI:
      return (p[0] >= 0 && p[1] >= 0 && p[2] >= 0 && p[0] + p[1] + p[2] <= 1.0) ? 0
                                                                              : YH_ERR_CHANNEL_PARAM;
    case YH_CH_RESET:
      return (p[0] >= 0 && p[1] >= 0 && p[0] + p[1] <= 1.0) ? 0 : YH_ERR_CHANNEL_PARAM;
    case YH_CH_PHASE_AMPLITUDE_DAMPING:
      if (!(p[0] >= 0 && p[1] >= 0 && p[0] + p[1] <= 1.0)) return YH_ERR_CHANNEL_PARAM;
      return in01(p[2]) ? 0 : YH_ERR_CHANNEL_PARAM;
    case YH_CH_AMPLITUDE_DAMPING:
      return (in01(p[0]) && in01(p[1])) ? 0 : YH_ERR_CHANNEL_PARAM;
    case YH_CH_PHASE_DAMPING: return in01(p[0]) ? 0 : YH_ERR_CHANNEL_PARAM;
    case YH_CH_THERMAL_RELAXATION:
      if (!(p[0] > 0 && p[1] > 0 && p[2] >= 0 && p[1] <= 2.0 * p[0])) return YH_ERR_CHANNEL_PARAM;
      return in01(p[3]) ? 0 : YH_ERR_CHANNEL_PARAM;
    case YH_CH_CUSTOM: return op.nmat >= 1 ? 0 : YH_ERR_CHANNEL_PARAM;
    case YH_CH_COHERENT: return 0;
  }
  return YH_ERR_INVALID_ARG;
}

Mat m2(C a, C b, C c, C d) { return Mat{a, b, c, d}; }

// phase_amplitude_damping_kraus (noise.rs:383-458)
std::vector<Mat> pad_kraus(double a, double b, double p1) {
  const double EPS = 2.220446049250313e-16;
  std::vector<Mat> ops;
  double rest = std::sqrt(1.0 - a - b);
  C z = cx(0, 0);
  if (std::fabs(p1 - 1.0) > EPS) {
    double s = std::sqrt(1.0 - p1);
    ops.push_back(m2(cx(s, 0), z, z, cx(s * rest, 0)));
    if (a > 0.0) ops.push_back(m2(z, cx(std::sqrt(a) * s, 0), z, z));
    if (b > 0.0) ops.push_back(m2(z, z, z, cx(std::sqrt(b) * s, 0)));
  }
  if (std::fabs(p1) > EPS) {
    double s = std::sqrt(p1);
    ops.push_back(m2(cx(s * rest, 0), z, z, cx(s, 0)));
    if (a > 0.0) ops.push_back(m2(z, z, cx(std::sqrt(a) * s, 0), z));
    if (b > 0.0) ops.push_back(m2(cx(std::sqrt(b) * s, 0), z, z, z));
  }
  return ops;
}

std::vector<Mat> pauli_kraus(double px, double py, double pz) {  // noise.rs:250-279
  C z = cx(0, 0);
  double s0 = std::sqrt(1.0 - px - py - pz), sx = std::sqrt(px), sy = std::sqrt(py),
         sz = std::sqrt(pz);
  return {m2(cx(s0, 0), z, z, cx(s0, 0)), m2(z, cx(sx, 0), cx(sx, 0), z),
          m2(z, cx(0, -sy), cx(0, sy), z), m2(cx(sz, 0), z, z, cx(-sz, 0))};
}

Mat kron(const Mat& a, int ad, const Mat& b, int bd) {  // noise.rs:512-526
  Mat r(size_t(ad * bd) * (ad * bd));
  for (int i = 0; i < ad; ++i)
    for (int j = 0; j < ad; ++j)
      for (int k = 0; k < bd; ++k)
        for (int l = 0; l < bd; ++l) r[(i * bd + k) * (ad * bd) + j * bd + l] = a[i * ad + j] * b[k * bd + l];
  return r;
}

std::vector<Mat> depol_multi(int n, double p) {  // noise.rs:462-508
  C z = cx(0, 0), o = cx(1, 0);
  Mat pa[4] = {m2(o, z, z, o), m2(z, o, o, z), m2(z, cx(0, -1), cx(0, 1), z), m2(o, z, z, cx(-1, 0))};
  usize dim = usize(1) << (2 * n);
  std::vector<Mat> ops;
  for (usize idx = 0; idx < dim; ++idx) {
    Mat mat{o};
    int md = 1;
    usize tmp = idx;
    for (int q = 0; q < n; ++q) {
      mat = kron(mat, md, pa[tmp % 4], 2);
      md *= 2;
      tmp /= 4;
    }
    double w = idx == 0 ? 1.0 - p + p / double(dim) : p / double(dim);
    double s = std::sqrt(w);
    for (auto& v : mat) v = v * cx(s, 0);
    ops.push_back(mat);
  }
  return ops;
}

int channel_num_qubits(const yh_op& op) {  // noise.rs:64-86
  if (op.kind == YH_CH_DEPOLARIZING) return int(op.params[1] > 0 ? op.params[1] : op.nt);
  if (op.kind == YH_CH_CUSTOM || op.kind == YH_CH_COHERENT) return __builtin_ctz(op.dim);
  return 1;
}

// kraus_operators (noise.rs:183-352); returns status, fills ops and dim
int kraus_operators(const yh_op& op, std::vector<Mat>& ops, int& dim) {
  int st = validate_channel(op);
  if (st) return st;
  const double* p = op.params;
  C z = cx(0, 0);
  dim = 2;
  switch (op.kind) {
    case YH_CH_PHASE_AMPLITUDE_DAMPING: ops = pad_kraus(p[0], p[1], p[2]); break;
    case YH_CH_AMPLITUDE_DAMPING: ops = pad_kraus(p[0], 0.0, p[1]); break;
    case YH_CH_PHASE_DAMPING: ops = pad_kraus(0.0, p[0], 0.0); break;
    case YH_CH_THERMAL_RELAXATION: {
      double t1 = p[0], t2 = p[1], time = p[2];
      double t_phi = (t1 * t2) / (2.0 * t1 - t2);
      double a = 1.0 - std::exp(-time / t1);
      double b = 1.0 - std::exp(-time / t_phi);
      ops = pad_kraus(a, b, p[3]);
      break;
    }
    case YH_CH_BITFLIP: {
      double s0 = std::sqrt(1.0 - p[0]), s1 = std::sqrt(p[0]);
      ops = {m2(cx(s0, 0), z, z, cx(s0, 0)), m2(z, cx(s1, 0), cx(s1, 0), z)};
      break;
    }
    case YH_CH_PHASEFLIP: {
      double s0 = std::sqrt(1.0 - p[0]), s1 = std::sqrt(p[0]);
      ops = {m2(cx(s0, 0), z, z, cx(s0, 0)), m2(cx(s1, 0), z, z, cx(-s1, 0))};
      break;
    }
    case YH_CH_PAULI: ops = pauli_kraus(p[0], p[1], p[2]); break;
    case YH_CH_DEPOLARIZING: {
      int n = channel_num_qubits(op);
      if (n == 1) {
        double q = p[0] / 4.0;
        ops = pauli_kraus(q, q, q);
      } else {
        ops = depol_multi(n, p[0]);
        dim = 1 << n;
      }
      break;
    }
    case YH_CH_RESET: {
      double s = std::sqrt(1.0 - p[0] - p[1]);
      ops = {m2(cx(s, 0), z, z, cx(s, 0))};
      if (p[0] > 0.0) {
        double sp = std::sqrt(p[0]);
        ops.push_back(m2(cx(sp, 0), z, z, z));
        ops.push_back(m2(z, cx(sp, 0), z, z));
      }
      if (p[1] > 0.0) {
        double sp = std::sqrt(p[1]);
        ops.push_back(m2(z, z, z, cx(sp, 0)));
        ops.push_back(m2(z, z, cx(sp, 0), z));
      }
      break;
    }
    case YH_CH_COHERENT:
    case YH_CH_CUSTOM: {
      dim = op.dim;
      int nm = op.kind == YH_CH_COHERENT ? 1 : op.nmat;
      for (int k = 0; k < nm; ++k) {
        Mat m(size_t(dim) * dim);
        const double* src = op.matrix + size_t(2) * dim * dim * k;
        for (int e = 0; e < dim * dim; ++e) m[e] = cx(src[2 * e], src[2 * e + 1]);
        ops.push_back(m);
      }
      break;
    }
    default: return YH_ERR_INVALID_ARG;
  }
  return 0;
}


// ---------------- Circuit::new validation (src/circuit.rs:150-272), qubit-only ----------------
int validate(int nbits, const yh_op* ops, size_t nops) {
  for (size_t e = 0; e < nops; ++e) {
    const yh_op& op = ops[e];
    if (op.elem == YH_ELEM_GATE) {
      if (op.nconf != op.nc) return YH_ERR_CONTROL_CONFIG_LEN;
      if (op.nt < 0 || op.nt > YH_MAX_TARGETS || op.nc < 0 || op.nc > YH_MAX_CTRLS)
        return YH_ERR_INVALID_ARG;
      for (int k = 0; k < op.nt; ++k)
        if (op.targets[k] < 0 || op.targets[k] >= nbits) return YH_ERR_LOC_OUT_OF_RANGE;
      for (int k = 0; k < op.nc; ++k)
        if (op.ctrls[k] < 0 || op.ctrls[k] >= nbits) return YH_ERR_LOC_OUT_OF_RANGE;
      for (int a = 0; a < op.nt; ++a)
        for (int b = 0; b < op.nc; ++b)
          if (op.targets[a] == op.ctrls[b]) return YH_ERR_OVERLAPPING_LOCS;
      int md;
      if (op.kind == YH_GATE_CUSTOM) md = op.dim;
      else if (op.kind == YH_GATE_SWAP || op.kind == YH_GATE_ISWAP || op.kind == YH_GATE_FSIM) md = 4;
      else md = 2;
      if (md != (1 << op.nt)) return YH_ERR_MATRIX_SIZE;
    } else if (op.elem == YH_ELEM_CHANNEL) {
      for (int k = 0; k < op.nt; ++k)
        if (op.targets[k] < 0 || op.targets[k] >= nbits) return YH_ERR_LOC_OUT_OF_RANGE;
      if (op.nt != channel_num_qubits(op)) return YH_ERR_MATRIX_SIZE;
      for (int a = 0; a < op.nt; ++a)
        for (int b = 0; b < a; ++b)
          if (op.targets[a] == op.targets[b]) return YH_ERR_OVERLAPPING_LOCS;
    } else if (op.elem == YH_ELEM_ANNOTATION) {
      if (op.nt >= 1 && (op.targets[0] < 0 || op.targets[0] >= nbits)) return YH_ERR_LOC_OUT_OF_RANGE;
    } else {
      return YH_ERR_INVALID_ARG;
    }
  }
  return 0;
}

// ---------------- operators (src/operator.rs) ----------------
Mat op_matrix(int op) {  // operator.rs:53-65
  C z = cx(0, 0), o = cx(1, 0);
  switch (op) {
    case YH_OP_I: return m2(o, z, z, o);
    case YH_OP_X: return m2(z, o, o, z);
    case YH_OP_Y: return m2(z, cx(0, -1), cx(0, 1), z);
    case YH_OP_Z: return m2(o, z, z, cx(-1, 0));
    case YH_OP_P0: return m2(o, z, z, z);
    case YH_OP_P1: return m2(z, z, z, o);
    case YH_OP_PU: return m2(z, o, z, z);
    case YH_OP_PD: return m2(z, z, o, z);
  }
  return m2(z, z, z, z);
}

struct Term {
  C coeff;
  std::vector<std::pair<usize, int>> ops;  // OperatorString::new: sorted by site, I removed
};

std::vector<Term> read_terms(size_t nterms, const double* coeffs, const int32_t* term_len,
                             const int32_t* sites, const int32_t* ops) {
  std::vector<Term> t(nterms);
  size_t off = 0;
  for (size_t k = 0; k < nterms; ++k) {
    t[k].coeff = cx(coeffs[2 * k], coeffs[2 * k + 1]);
    std::vector<std::pair<usize, int>> v;
    for (int j = 0; j < term_len[k]; ++j) v.push_back({usize(sites[off + j]), ops[off + j]});
    off += term_len[k];
    std::stable_sort(v.begin(), v.end(), [](auto& a, auto& b) { return a.first < b.first; });
    for (auto& x : v)
      if (x.second != YH_OP_I) t[k].ops.push_back(x);
  }
  return t;
}

// expect_opstring_pure (expect.rs:13-36)
C expect_opstring_pure(const Vec& st, const Term& t) {
  Vec s = st;
  for (auto& so : t.ops) {
    Mat m = op_matrix(so.second);
    instruct_1q(s.data(), s.size(), so.first, m[0], m[1], m[2], m[3]);
  }
  C acc = cx(0, 0);
  for (size_t i = 0; i < st.size(); ++i) acc += conj(st[i]) * s[i];
  return acc;
}

// expect_opstring_dm (expect.rs:44-72), sites are LSB-first bit positions
C expect_opstring_dm(usize nbits, const C* rho, const Term& t) {
  usize dim = usize(1) << nbits;
  std::vector<Mat> mats;
  for (auto& so : t.ops) mats.push_back(op_matrix(so.second));
  C tr = cx(0, 0);
  for (usize row = 0; row < dim; ++row)
    for (usize col = 0; col < dim; ++col) {
      C e = cx(1, 0);
      for (size_t q = 0; q < t.ops.size(); ++q) {
        usize loc = t.ops[q].first;
        usize rb = (row >> loc) & 1, cb = (col >> loc) & 1;
        e *= mats[q][rb * 2 + cb];
      }
      bool match = true;
      for (usize loc = 0; loc < nbits && match; ++loc) {
        bool touched = false;
        for (auto& so : t.ops)
          if (so.first == loc) touched = true;
        if (!touched && (((row >> loc) & 1) != ((col >> loc) & 1))) match = false;
      }
      if (match) tr += e * rho[col * dim + row];
    }
  return tr;
}

// ---------------- density matrix (src/density_matrix.rs) ----------------
bool is_diagonal_matrix(const Mat& m, int d) {  // :125-129
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c)
      if (r != c && std::hypot(m[r * d + c].re, m[r * d + c].im) >= 1e-15) return false;
  return true;
}

yh_op custom_gate(const Mat& m, int d, bool diag, const yh_op& pos, std::vector<double>& store) {
  yh_op g = pos;
  g.elem = YH_ELEM_GATE;
  g.kind = YH_GATE_CUSTOM;
  g.dim = d;
  g.is_diag = diag ? 1 : 0;
  store.resize(size_t(2) * d * d);
  for (int e = 0; e < d * d; ++e) {
    store[2 * e] = m[e].re;
    store[2 * e + 1] = m[e].im;
  }
  g.matrix = store.data();
  return g;
}

// ---------------- adjoint differentiation (src/ad.rs) ----------------
int gate_num_params(const yh_op& op) {  // gate.rs:107-113
  if (op.elem != YH_ELEM_GATE) return 0;
  if (op.kind == YH_GATE_RX || op.kind == YH_GATE_RY || op.kind == YH_GATE_RZ || op.kind == YH_GATE_PHASE) return 1;
  return op.kind == YH_GATE_FSIM ? 2 : 0;
}

Mat generator_matrix(const yh_op& op, int i) {  // gate.rs:153-193
  C z = cx(0, 0);
  switch (op.kind) {
    case YH_GATE_RX: return m2(z, cx(0, -0.5), cx(0, -0.5), z);
    case YH_GATE_RY: return m2(z, cx(-0.5, 0), cx(0.5, 0), z);
    case YH_GATE_RZ: return m2(cx(0, -0.5), z, z, cx(0, 0.5));
    case YH_GATE_PHASE: return m2(z, z, z, cx(0, 1));
    default: {
      Mat m(16, z);
      if (i == 0) {
        m[1 * 4 + 2] = cx(0, -1);
        m[2 * 4 + 1] = cx(0, -1);
      } else {
        m[3 * 4 + 3] = cx(0, -1);
      }
      return m;
    }
  }
}

Mat conjugate_transpose(const Mat& m, int d) {
  Mat o(m.size());
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c) o[r * d + c] = conj(m[c * d + r]);
  return o;
}

// Gate::dagger (gate.rs:198-281), with the same element position
yh_op dagger_gate(const yh_op& op, std::vector<double>& store) {
  yh_op d = op;
  switch (op.kind) {
    case YH_GATE_H: case YH_GATE_X: case YH_GATE_Y: case YH_GATE_Z: case YH_GATE_SWAP:
      return d;
    case YH_GATE_S:
      d.kind = YH_GATE_PHASE;
      d.params[0] = -M_PI / 2;
      return d;
    case YH_GATE_T:
      d.kind = YH_GATE_PHASE;
      d.params[0] = -M_PI / 4;
      return d;
    case YH_GATE_RX: case YH_GATE_RY: case YH_GATE_RZ: case YH_GATE_PHASE:
      d.params[0] = -op.params[0];
      return d;
    case YH_GATE_FSIM:
      d.params[0] = -op.params[0];
      d.params[1] = -op.params[1];
      return d;
    default: {  // SqrtX, SqrtY, SqrtW, ISWAP, Custom: conjugate transpose as Custom
      std::vector<C> m;
      int dim = gate_matrix(op, m);
      bool diag = op.kind == YH_GATE_CUSTOM ? op.is_diag != 0 : false;
      return custom_gate(conjugate_transpose(m, dim), dim, diag, op, store);
    }
  }
}

bool controls_match(usize basis, usize nbits, const yh_op& op) {  // ad.rs:19-27
  for (int k = 0; k < op.nc; ++k) {
    usize bit = nbits - 1 - usize(op.ctrls[k]);
    if (((basis >> bit) & 1) != usize(op.ctrl_vals[k])) return false;
  }
  return true;
}

// apply_generator (ad.rs:29-64): the generator as a non-diagonal Custom gate at
// the same position; amplitudes on inactive control branches become zero
void apply_generator(Vec& state, usize nbits, const yh_op& op, const Mat& g) {
  std::vector<double> store;
  int d = 1 << op.nt;
  yh_op temp = custom_gate(g, d, false, op, store);
  if (op.nc == 0) {
    dispatch(nbits, state.data(), temp);
    return;
  }
  Vec transformed = state;
  dispatch(nbits, transformed.data(), temp);
  for (usize b = 0; b < state.size(); ++b) state[b] = controls_match(b, nbits, op) ? transformed[b] : cx(0, 0);
}

void dm_apply_gate(usize n, C* rho, const yh_op& op) {  // :161-178
  dispatch(2 * n, rho, op);
  std::vector<C> m;
  int d = gate_matrix(op, m);
  for (auto& v : m) v = conj(v);
  std::vector<double> store;
  yh_op c = custom_gate(m, d, gate_is_diagonal(op), op, store);
  for (int k = 0; k < c.nt; ++k) c.targets[k] += int(n);
  for (int k = 0; k < c.nc; ++k) c.ctrls[k] += int(n);
  dispatch(2 * n, rho, c);
}

int dm_apply_channel(usize n, Vec& rho, const yh_op& op) {  // :181-224
  std::vector<Mat> kr;
  int d;
  int st = kraus_operators(op, kr, d);
  if (st) return st;
  Vec original = rho;
  Vec acc(rho.size(), cx(0, 0));
  Vec branch(rho.size(), cx(0, 0));
  yh_op pos{};
  pos.nt = op.nt;
  for (int k = 0; k < op.nt; ++k) pos.targets[k] = op.targets[k];
  for (auto& K : kr) {
    branch = original;
    bool dg = is_diagonal_matrix(K, d);
    std::vector<double> s1, s2;
    yh_op g = custom_gate(K, d, dg, pos, s1);
    dispatch(2 * n, branch.data(), g);
    Mat Kc = K;
    for (auto& v : Kc) v = conj(v);
    yh_op gc = custom_gate(Kc, d, dg, pos, s2);
    for (int k = 0; k < gc.nt; ++k) gc.targets[k] += int(n);
    dispatch(2 * n, branch.data(), gc);
    const usize len = acc.size();
#pragma omp parallel for ORC_PAR(len)
    for (usize i = 0; i < len; ++i) acc[i] += branch[i];
  }
  rho = acc;
  return 0;
}

// ---------------- measure (src/measure.rs) ----------------
std::vector<double> marginal_probs_qubits(const C* s, usize nbits, const std::vector<usize>& locs) {
  // :89-104
  std::vector<double> pv(usize(1) << locs.size(), 0.0);
  usize len = usize(1) << nbits;
  for (usize b = 0; b < len; ++b) {
    usize mi = 0;
    for (size_t idx = 0; idx < locs.size(); ++idx) {
      usize bit = nbits - 1 - locs[idx];
      mi |= ((b >> bit) & 1) << (locs.size() - 1 - idx);
    }
    pv[mi] += norm_sqr(s[b]);
  }
  return pv;
}

int validate_measure_locs(usize nbits, const std::vector<usize>& locs) {  // :124-133
  for (size_t i = 0; i < locs.size(); ++i) {
    if (locs[i] >= nbits) return YH_ERR_MEASURE_LOC;
    for (size_t j = 0; j < i; ++j)
      if (locs[j] == locs[i]) return YH_ERR_MEASURE_LOC;
  }
  return 0;
}

usize sample_from_probs(const std::vector<double>& p, double r) {  // :140-150
  double cs = 0.0;
  for (size_t i = 0; i < p.size(); ++i) {
    cs += p[i];
    if (r < cs) return i;
  }
  return p.size() - 1;
}

void collapse_qubits_to(C* s, usize nbits, const std::vector<usize>& locs,
                        const std::vector<usize>& vals) {  // :158-185
  usize len = usize(1) << nbits;
  double ns = 0.0;
  for (usize b = 0; b < len; ++b) {
    bool match = true;
    for (size_t k = 0; k < locs.size(); ++k)
      if (((b >> (nbits - 1 - locs[k])) & 1) != vals[k]) match = false;
    if (match) ns += norm_sqr(s[b]);
    else s[b] = cx(0, 0);
  }
  double nrm = std::sqrt(ns);
  if (nrm > 1e-15)
    for (usize b = 0; b < len; ++b) s[b] = divr(s[b], nrm);
}

std::vector<usize> decode_outcome_bits(usize idx, usize nb) {  // :152-156
  std::vector<usize> r;
  for (usize k = 0; k < nb; ++k) r.push_back((idx >> (nb - 1 - k)) & 1);
  return r;
}

void reset_qubits_to(C* s, usize nbits, const std::vector<usize>& locs,
                     const std::vector<usize>& from, usize reset_val) {  // :187-214
  std::vector<usize> to = decode_outcome_bits(reset_val, locs.size());
  if (from == to) return;
  usize sm = 0;
  for (size_t k = 0; k < locs.size(); ++k)
    if (from[k] != to[k]) sm |= usize(1) << (nbits - 1 - locs[k]);
  usize len = usize(1) << nbits;
  for (usize b = 0; b < len; ++b) {
    bool mf = true;
    for (size_t k = 0; k < locs.size(); ++k)
      if (((b >> (nbits - 1 - locs[k])) & 1) != from[k]) mf = false;
    if (!mf) continue;
    C a = s[b];
    s[b ^ sm] = a;
    s[b] = cx(0, 0);
  }
}

Vec remove_measured_qubits(const C* s, usize nbits, const std::vector<usize>& locs) {
  // :216-240
  std::vector<usize> kept;
  for (usize l = 0; l < nbits; ++l)
    if (std::find(locs.begin(), locs.end(), l) == locs.end()) kept.push_back(l);
  Vec ns(usize(1) << kept.size(), cx(0, 0));
  usize len = usize(1) << nbits;
  for (usize b = 0; b < len; ++b) {
    if (norm_sqr(s[b]) < 1e-30) continue;
    usize nb = 0;
    for (size_t idx = 0; idx < kept.size(); ++idx) {
      usize bit = nbits - 1 - kept[idx];
      nb |= ((b >> bit) & 1) << (kept.size() - 1 - idx);
    }
    ns[nb] += s[b];
  }
  // ArrayReg::normalize (register.rs:106-113)
  double acc = 0.0;
  for (auto& a : ns) acc += norm_sqr(a);
  double nrm = std::sqrt(acc);
  if (nrm > 0.0)
    for (auto& a : ns) a = divr(a, nrm);
  return ns;
}

C* asC(double* p) { return reinterpret_cast<C*>(p); }
const C* asC(const double* p) { return reinterpret_cast<const C*>(p); }

}  // namespace

// ======================= exported C API (test infrastructure) =======================
extern "C" {

// itercontrol index list (bitbasis iterate_control.rs:82-106); out gets 2^(n-m) values
int orc_itercontrol(int nbits, const int32_t* positions, const int32_t* configs, int m,
                    uint64_t* out) {
  std::vector<usize> p, c;
  for (int k = 0; k < m; ++k) {
    p.push_back(positions[k]);
    c.push_back(configs[k]);
  }
  IterControl ic = itercontrol(nbits, p, c);
  for (usize k = 0; k < ic.n; ++k) out[k] = ic.get(k);
  return 0;
}

int orc_gate_matrix(const yh_op* op, double* out, int* dim) {
  std::vector<C> m;
  int d = gate_matrix(*op, m);
  if (d < 0) return YH_ERR_INVALID_ARG;
  for (int e = 0; e < d * d; ++e) {
    out[2 * e] = m[e].re;
    out[2 * e + 1] = m[e].im;
  }
  *dim = d;
  return 0;
}

int orc_validate(int nbits, const yh_op* ops, size_t nops) { return validate(nbits, ops, nops); }

// Direct instruct calls (instruct_qubit.rs), as the reference unit tests make them.
// fn: 0 instruct_1q, 1 instruct_1q_diag, 2 instruct_2q, 3 instruct_2q_diag, 4 instruct_nq,
//     5 instruct_1q_controlled, 6 instruct_2q_controlled, 7 instruct_1q_diag_controlled,
//     8 instruct_2q_diag_controlled, 9 instruct_x, 10 instruct_x_controlled, 11 instruct_swap
// gate: row-major interleaved matrix (diag variants take the diagonal: 2 or 4 entries).
int orc_instruct(int fn, int nbits, double* state, const int32_t* locs, int nlocs,
                 const double* gate, const int32_t* ctrl_locs, const int32_t* ctrl_bits, int nc) {
  C* s = asC(state);
  const C* g = gate ? asC(gate) : nullptr;
  std::vector<usize> l, cl, cb;
  for (int k = 0; k < nlocs; ++k) l.push_back(locs[k]);
  for (int k = 0; k < nc; ++k) {
    cl.push_back(ctrl_locs[k]);
    cb.push_back(ctrl_bits[k]);
  }
  usize len = usize(1) << nbits;
  switch (fn) {
    case 0: instruct_1q(s, len, l[0], g[0], g[1], g[2], g[3]); break;
    case 1: instruct_1q_diag(s, len, l[0], g[0], g[1]); break;
    case 2: instruct_2q(s, nbits, l, g); break;
    case 3: instruct_2q_diag(s, nbits, l, g); break;
    case 4: instruct_nq(s, nbits, l, g, cl, cb); break;
    case 5: instruct_1q_controlled(s, nbits, l[0], g[0], g[1], g[2], g[3], cl, cb); break;
    case 6: instruct_2q_controlled(s, nbits, l, g, cl, cb); break;
    case 7: instruct_1q_diag_controlled(s, nbits, l[0], g[0], g[1], cl, cb); break;
    case 8: instruct_2q_diag_controlled(s, nbits, l, g, cl, cb); break;
    case 9: instruct_x(s, nbits, l[0]); break;
    case 10: instruct_x_controlled(s, nbits, l[0], cl, cb); break;
    case 11: instruct_swap(s, nbits, l); break;
    default: return YH_ERR_INVALID_ARG;
  }
  return 0;
}

// apply_inplace (apply.rs:123-147) after Circuit::new validation.
int orc_apply(int nbits, int circuit_nbits, double* state, const yh_op* ops, size_t nops) {
  if (nbits != circuit_nbits) return YH_ERR_NBITS_MISMATCH;
  int st = validate(circuit_nbits, ops, nops);
  if (st) return st;
  for (size_t e = 0; e < nops; ++e)
    if (ops[e].elem == YH_ELEM_GATE) dispatch(nbits, asC(state), ops[e]);
  return 0;
}

// Threads for the parallel loops (k >= 1); returns the previous value.
int orc_set_threads(int k) {
  const int prev = g_threads;
  g_threads = k < 1 ? 1 : k;
  return prev;
}

int orc_dispatch(int nbits, double* state, const yh_op* op) {
  dispatch(nbits, asC(state), *op);
  return 0;
}

int orc_expect(int nbits, const double* state, size_t nterms, const double* coeffs,
               const int32_t* term_len, const int32_t* sites, const int32_t* ops, double* out) {
  usize len = usize(1) << nbits;
  Vec st(asC(state), asC(state) + len);
  std::vector<Term> t = read_terms(nterms, coeffs, term_len, sites, ops);
  C acc = cx(0, 0);  // expect_arrayreg (expect.rs:7-11): sum of coeff * <opstring>
  for (auto& term : t) acc += term.coeff * expect_opstring_pure(st, term);
  out[0] = acc.re;
  out[1] = acc.im;
  return 0;
}

// expect_grad (ad.rs:75-156). Returns 1 for a circuit with a channel (the
// reference panics), 2 for a wrong gradient length.
int orc_expect_grad(int nbits, const double* psi0, const void* opsv, size_t nops, size_t nterms, const double* coeffs,
                    const int32_t* term_len, const int32_t* sites, const int32_t* opcodes, double* value,
                    double* grad, size_t ngrad) {
  const yh_op* ops = static_cast<const yh_op*>(opsv);
  for (size_t e = 0; e < nops; ++e)
    if (ops[e].elem == YH_ELEM_CHANNEL) return 1;
  usize len = usize(1) << nbits;
  Vec psi(asC(psi0), asC(psi0) + len);
  for (size_t e = 0; e < nops; ++e)  // apply_inplace
    if (ops[e].elem == YH_ELEM_GATE) dispatch(nbits, psi.data(), ops[e]);
  std::vector<Term> terms = read_terms(nterms, coeffs, term_len, sites, opcodes);
  C v = cx(0, 0);
  for (auto& t : terms) v += t.coeff * expect_opstring_pure(psi, t);
  *value = v.re;
  size_t n_params = 0;
  for (size_t e = 0; e < nops; ++e) n_params += size_t(gate_num_params(ops[e]));
  if (ngrad != n_params) return 2;
  if (n_params == 0) return 0;
  Vec lambda(len, cx(0, 0));
  for (auto& t : terms) {
    Vec tmp = psi;
    for (auto& so : t.ops) {
      Mat m = op_matrix(so.second);
      instruct_1q(tmp.data(), len, so.first, m[0], m[1], m[2], m[3]);
    }
    for (usize i = 0; i < len; ++i) lambda[i] += t.coeff * tmp[i];
  }
  size_t slot = n_params;
  Vec phi(len);
  std::vector<double> store;
  for (size_t e = nops; e-- > 0;) {
    const yh_op& op = ops[e];
    if (op.elem != YH_ELEM_GATE) continue;
    int np = gate_num_params(op);
    for (int i = np - 1; i >= 0; --i) {
      --slot;
      phi = psi;
      apply_generator(phi, nbits, op, generator_matrix(op, i));
      C acc = cx(0, 0);
      for (usize k = 0; k < len; ++k) acc += conj(lambda[k]) * phi[k];
      grad[slot] = 2.0 * acc.re;
    }
    yh_op d = dagger_gate(op, store);
    dispatch(nbits, psi.data(), d);
    dispatch(nbits, lambda.data(), d);
  }
  return 0;
}

int orc_probs(int nbits, const double* state, const int32_t* locs, int nlocs, double* out) {
  usize len = usize(1) << nbits;
  const C* s = asC(state);
  if (!locs) {  // full_probs (measure.rs:38-40)
    for (usize b = 0; b < len; ++b) out[b] = norm_sqr(s[b]);
    return 0;
  }
  std::vector<usize> l(locs, locs + nlocs);
  int st = validate_measure_locs(nbits, l);
  if (st) return st;
  std::vector<double> p = marginal_probs_qubits(s, nbits, l);
  std::copy(p.begin(), p.end(), out);
  return 0;
}

uint64_t orc_sample_from_probs(const double* probs, size_t n, double r) {
  return sample_from_probs(std::vector<double>(probs, probs + n), r);
}

// Shot loop: each shot draws one uniform and samples (measure_with_postprocess with
// NoPostProcess, as called per shot in the reference CLI run.rs:32-42).
int orc_sample(int nbits, const double* state, const int32_t* locs, int nlocs,
               const double* uniforms, size_t nshots, uint64_t* out) {
  usize len = usize(1) << nbits;
  const C* s = asC(state);
  std::vector<double> p;
  if (!locs) {
    p.resize(len);
    for (usize b = 0; b < len; ++b) p[b] = norm_sqr(s[b]);
  } else {
    std::vector<usize> l(locs, locs + nlocs);
    int st = validate_measure_locs(nbits, l);
    if (st) return st;
    p = marginal_probs_qubits(s, nbits, l);
  }
  for (size_t k = 0; k < nshots; ++k) out[k] = sample_from_probs(p, uniforms[k]);
  return 0;
}

// measure_with_postprocess (measure.rs:243-274). For RemoveMeasured the new state
// (2^(n-nlocs) amps) is written to new_state.
int orc_measure(int nbits, double* state, const int32_t* locs, int nlocs, int post,
                uint64_t reset_val, double r, int32_t* outcome_bits, double* new_state) {
  std::vector<usize> l(locs, locs + nlocs);
  int st = validate_measure_locs(nbits, l);
  if (st) return st;
  C* s = asC(state);
  std::vector<double> p = marginal_probs_qubits(s, nbits, l);
  usize idx = sample_from_probs(p, r);
  std::vector<usize> oc = decode_outcome_bits(idx, l.size());
  for (size_t k = 0; k < oc.size(); ++k) outcome_bits[k] = int32_t(oc[k]);
  if (post == 1) {
    if (reset_val >= (usize(1) << l.size())) return YH_ERR_RESET_VALUE;
    collapse_qubits_to(s, nbits, l, oc);
    reset_qubits_to(s, nbits, l, oc, reset_val);
  } else if (post == 2) {
    collapse_qubits_to(s, nbits, l, oc);
    Vec ns = remove_measured_qubits(s, nbits, l);
    std::memcpy(new_state, ns.data(), ns.size() * sizeof(C));
  }
  return 0;
}

double orc_norm(int nbits, const double* state) {  // register.rs:98-104
  usize len = usize(1) << nbits;
  double acc = 0.0;
  for (usize b = 0; b < len; ++b) acc += norm_sqr(asC(state)[b]);
  return std::sqrt(acc);
}

double orc_fidelity(int nbits, const double* a, const double* b) {  // register.rs:115-124
  usize len = usize(1) << nbits;
  C acc = cx(0, 0);
  for (usize k = 0; k < len; ++k) acc += conj(asC(a)[k]) * asC(b)[k];
  return norm_sqr(acc);
}

int orc_kraus(const yh_op* ch, double* out, int max_ops, int* nops, int* dim) {
  std::vector<Mat> kr;
  int d;
  int st = kraus_operators(*ch, kr, d);
  if (st) return st;
  if (int(kr.size()) > max_ops) return YH_ERR_INVALID_ARG;
  size_t off = 0;
  for (auto& K : kr)
    for (auto& v : K) {
      out[off++] = v.re;
      out[off++] = v.im;
    }
  *nops = int(kr.size());
  *dim = d;
  return 0;
}

// superop (noise.rs:359-379): S = sum kron(conj(K), K); out is d^2 x d^2.
int orc_superop(const yh_op* ch, double* out, int* d2out) {
  std::vector<Mat> kr;
  int d;
  int st = kraus_operators(*ch, kr, d);
  if (st) return st;
  int d2 = d * d;
  std::vector<C> S(size_t(d2) * d2, cx(0, 0));
  for (auto& K : kr)
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j)
        for (int kr_ = 0; kr_ < d; ++kr_)
          for (int kc = 0; kc < d; ++kc)
            S[size_t(i * d + kr_) * d2 + (j * d + kc)] += conj(K[i * d + j]) * K[kr_ * d + kc];
  for (size_t e = 0; e < S.size(); ++e) {
    out[2 * e] = S[e].re;
    out[2 * e + 1] = S[e].im;
  }
  *d2out = d2;
  return 0;
}

// DensityMatrix::from_reg (density_matrix.rs:23-35)
// deterministic_state (register.rs:71-84): (cos 0.1k, sin 0.2k), divided by the
// square root of the sequential sum of norm_sqr
int orc_deterministic_state(int nbits, double* out) {
  const usize dim = usize(1) << nbits;
  C* s = asC(out);
  for (usize k = 0; k < dim; ++k) {
    const double kf = double(k);
    s[k] = cx(std::cos(0.1 * kf), std::sin(0.2 * kf));
  }
  double acc = 0.0;
  for (usize k = 0; k < dim; ++k) acc += norm_sqr(s[k]);
  const double nrm = std::sqrt(acc);
  for (usize k = 0; k < dim; ++k) s[k] = divr(s[k], nrm);
  return 0;
}

// rand_state (register.rs:35-45) with the uniforms of the caller's stream in the
// reference's draw order: re then im per amplitude
int orc_rand_state(int nbits, const double* u, double* out) {
  const usize dim = usize(1) << nbits;
  C* s = asC(out);
  for (usize k = 0; k < dim; ++k) s[k] = cx(u[2 * k] - 0.5, u[2 * k + 1] - 0.5);
  double acc = 0.0;
  for (usize k = 0; k < dim; ++k) acc += norm_sqr(s[k]);
  const double nrm = std::sqrt(acc);
  for (usize k = 0; k < dim; ++k) s[k] = divr(s[k], nrm);
  return 0;
}

// mixed (density_matrix.rs:37-61): states holds nregs registers back to back
int orc_dm_mixed(int nbits, int nregs, const double* weights, const double* states, double* rho) {
  const usize dim = usize(1) << nbits;
  C* r = asC(rho);
  for (usize e = 0; e < dim * dim; ++e) r[e] = cx(0, 0);
  for (int j = 0; j < nregs; ++j) {
    const C* psi = asC(states) + usize(j) * dim;
    for (usize row = 0; row < dim; ++row)
      for (usize col = 0; col < dim; ++col) r[row * dim + col] += cx(weights[j], 0.0) * psi[row] * conj(psi[col]);
  }
  return 0;
}

int orc_dm_from_reg(int nbits, const double* state, double* rho) {
  usize dim = usize(1) << nbits;
  const C* s = asC(state);
  C* r = asC(rho);
#pragma omp parallel for ORC_PAR(dim * dim)
  for (usize row = 0; row < dim; ++row)
    for (usize col = 0; col < dim; ++col) r[row * dim + col] = s[row] * conj(s[col]);
  return 0;
}

// Register::apply for DensityMatrix (density_matrix.rs:244-259)
int orc_dm_apply(int nbits, int circuit_nbits, double* rho, const yh_op* ops, size_t nops) {
  if (nbits != circuit_nbits) return YH_ERR_NBITS_MISMATCH;
  int st = validate(circuit_nbits, ops, nops);
  if (st) return st;
  usize len = usize(1) << (2 * nbits);
  Vec r(asC(rho), asC(rho) + len);
  for (size_t e = 0; e < nops; ++e) {
    if (ops[e].elem == YH_ELEM_GATE) dm_apply_gate(nbits, r.data(), ops[e]);
    else if (ops[e].elem == YH_ELEM_CHANNEL) {
      st = dm_apply_channel(nbits, r, ops[e]);
      if (st) return st;
    }
  }
  std::memcpy(rho, r.data(), len * sizeof(C));
  return 0;
}

int orc_dm_trace(int nbits, const double* rho, double* out) {  // :63-67
  usize dim = usize(1) << nbits;
  C acc = cx(0, 0);
  for (usize i = 0; i < dim; ++i) acc += asC(rho)[i * dim + i];
  out[0] = acc.re;
  out[1] = acc.im;
  return 0;
}

double orc_dm_purity(int nbits, const double* rho) {  // :69-78
  usize dim = usize(1) << nbits;
  const C* r = asC(rho);
  C acc = cx(0, 0);
  for (usize row = 0; row < dim; ++row)
    for (usize in = 0; in < dim; ++in) acc += r[row * dim + in] * r[in * dim + row];
  return acc.re;
}

int orc_dm_expect(int nbits, const double* rho, size_t nterms, const double* coeffs,
                  const int32_t* term_len, const int32_t* sites, const int32_t* ops, double* out) {
  std::vector<Term> t = read_terms(nterms, coeffs, term_len, sites, ops);
  C acc = cx(0, 0);  // expect_dm (expect.rs:38-42)
  for (auto& term : t) acc += term.coeff * expect_opstring_dm(nbits, asC(rho), term);
  out[0] = acc.re;
  out[1] = acc.im;
  return 0;
}

int orc_dm_probs(int nbits, const double* rho, const int32_t* locs, int nlocs, double* out) {
  usize dim = usize(1) << nbits;
  const C* r = asC(rho);
  if (!locs) {  // measure.rs:50-56
    for (usize b = 0; b < dim; ++b) out[b] = std::max(r[b * dim + b].re, 0.0);
    return 0;
  }
  std::vector<usize> l(locs, locs + nlocs);
  int st = validate_measure_locs(nbits, l);
  if (st) return st;
  std::fill(out, out + (usize(1) << nlocs), 0.0);  // measure.rs:106-122
  for (usize b = 0; b < dim; ++b) {
    usize mi = 0;
    for (int idx = 0; idx < nlocs; ++idx) {
      usize bit = nbits - 1 - l[idx];
      mi |= ((b >> bit) & 1) << (nlocs - 1 - idx);
    }
    out[mi] += std::max(r[b * dim + b].re, 0.0);
  }
  return 0;
}

// partial_tr (density_matrix.rs:80-112), traced locs are LSB-first bit positions.
// von_neumann_entropy (density_matrix.rs:113-122). Eigenvalues by cyclic Jacobi
// rotations on the real symmetric embedding [[A, -B], [B, A]] of rho = A + iB
// (each eigenvalue appears twice), then the reference's clamp and filter.
int orc_dm_entropy(int nbits, const double* rho, double* out) {
  const usize d = usize(1) << nbits, n = 2 * d;
  const C* r = asC(rho);
  std::vector<double> a(n * n);
  for (usize i = 0; i < d; ++i)
    for (usize j = 0; j < d; ++j) {
      const C h = r[i * d + j];
      a[i * n + j] = h.re;
      a[(i + d) * n + j + d] = h.re;
      a[i * n + j + d] = -h.im;
      a[(i + d) * n + j] = h.im;
    }
  // a sweep rotates every pair whose off-diagonal entry is not negligible against
  // its two diagonal entries; the iteration ends after a sweep with no rotation
  for (int sweep = 0; sweep < 60; ++sweep) {
    int rotated = 0;
    for (usize p = 0; p + 1 < n; ++p)
      for (usize q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        const double scale = std::fabs(a[p * n + p]) + std::fabs(a[q * n + q]);
        if (std::fabs(apq) <= 1e-17 * scale || std::fabs(apq) < 1e-300) continue;
        ++rotated;
        const double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (usize k = 0; k < n; ++k) {  // columns p, q
          const double kp = a[k * n + p], kq = a[k * n + q];
          a[k * n + p] = c * kp - s * kq;
          a[k * n + q] = s * kp + c * kq;
        }
        for (usize k = 0; k < n; ++k) {  // rows p, q
          const double pk = a[p * n + k], qk = a[q * n + k];
          a[p * n + k] = c * pk - s * qk;
          a[q * n + k] = s * pk + c * qk;
        }
      }
    if (rotated == 0) break;
  }
  std::vector<double> ev(n);
  for (usize i = 0; i < n; ++i) ev[i] = a[i * n + i];
  std::sort(ev.begin(), ev.end());
  double acc = 0.0;
  for (usize k = 0; k < d; ++k) {
    double v = std::max(ev[2 * k], 0.0);
    if (v > 1e-15) acc += v * std::log(v);
  }
  *out = -acc;
  return 0;
}

int orc_dm_partial_tr(int nbits, const double* rho, const int32_t* traced, int ntr, double* out,
                      int* kept_nbits) {
  std::vector<usize> tl(traced, traced + ntr), kept;
  for (usize l = 0; l < usize(nbits); ++l)
    if (std::find(tl.begin(), tl.end(), l) == tl.end()) kept.push_back(l);
  usize dim = usize(1) << nbits, dr = usize(1) << kept.size();
  std::vector<C> red(dr * dr, cx(0, 0));
  const C* r = asC(rho);
  for (usize row = 0; row < dim; ++row)
    for (usize col = 0; col < dim; ++col) {
      bool all = true;
      for (usize l : tl)
        if (((row >> l) & 1) != ((col >> l) & 1)) all = false;
      if (!all) continue;
      usize rr = 0, rc = 0;
      for (size_t idx = 0; idx < kept.size(); ++idx) {
        rr |= ((row >> kept[idx]) & 1) << idx;
        rc |= ((col >> kept[idx]) & 1) << idx;
      }
      red[rr * dr + rc] += r[row * dim + col];
    }
  std::memcpy(out, red.data(), red.size() * sizeof(C));
  *kept_nbits = int(kept.size());
  return 0;
}

}  // extern "C"
