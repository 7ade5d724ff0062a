// host.cpp — host-side gate, channel and operator data for the gfx950 backend.
// Built with -ffp-contract=off so matrices are rounded exactly as yao-rs builds them.
#include "host.h"

#include <algorithm>
#include <cmath>

namespace yh {

namespace {
const double kFrac1Sqrt2 = 0.70710678118654752440;
const double kFracPi4 = 0.78539816339744830962;
inline cd C(double r, double i) { return cd{r, i}; }
inline cd polar1(double t) { return cd{1.0 * std::cos(t), 1.0 * std::sin(t)}; }
}  // namespace

int gate_matrix(const yh_op& op, Mat& m) {
  const cd z = C(0, 0), one = C(1, 0), neg_one = C(-1, 0), i = C(0, 1), neg_i = C(0, -1);
  const double* p = op.params;
  switch (op.kind) {
    case YH_GATE_X: m = {z, one, one, z}; return 2;
    case YH_GATE_Y: m = {z, neg_i, i, z}; return 2;
    case YH_GATE_Z: m = {one, z, z, neg_one}; return 2;
    case YH_GATE_H: m = {C(kFrac1Sqrt2, 0), C(kFrac1Sqrt2, 0), C(kFrac1Sqrt2, 0), C(-kFrac1Sqrt2, 0)};
      return 2;
    case YH_GATE_S: m = {one, z, z, i}; return 2;
    case YH_GATE_T: m = {one, z, z, polar1(kFracPi4)}; return 2;
    case YH_GATE_SWAP:
      m.assign(16, z);
      m[0] = one; m[6] = one; m[9] = one; m[15] = one;
      return 4;
    case YH_GATE_PHASE: m = {one, z, z, polar1(p[0])}; return 2;
    case YH_GATE_RX: {
      cd c = C(std::cos(p[0] / 2.0), 0), s = C(0, -std::sin(p[0] / 2.0));
      m = {c, s, s, c};
      return 2;
    }
    case YH_GATE_RY: {
      double c = std::cos(p[0] / 2.0), s = std::sin(p[0] / 2.0);
      m = {C(c, 0), C(-s, 0), C(s, 0), C(c, 0)};
      return 2;
    }
    case YH_GATE_RZ: m = {polar1(-p[0] / 2.0), z, z, polar1(p[0] / 2.0)}; return 2;
    case YH_GATE_SQRTX: {
      cd f = C(0.5, 0.5);
      m = {f * one, f * neg_i, f * neg_i, f * one};
      return 2;
    }
    case YH_GATE_SQRTY: {
      cd f = C(0.5, 0.5);
      m = {f * one, f * neg_one, f * one, f * one};
      return 2;
    }
    case YH_GATE_SQRTW: {
      cd c = C(kFrac1Sqrt2, 0), nis = C(0, -kFrac1Sqrt2);
      m = {c, nis * C(kFrac1Sqrt2, -kFrac1Sqrt2), nis * C(kFrac1Sqrt2, kFrac1Sqrt2), c};
      return 2;
    }
    case YH_GATE_ISWAP:
      m.assign(16, z);
      m[0] = one; m[6] = i; m[9] = i; m[15] = one;
      return 4;
    case YH_GATE_FSIM: {
      m.assign(16, z);
      cd ct = C(std::cos(p[0]), 0), ns = C(0, -std::sin(p[0]));
      m[0] = one; m[5] = ct; m[6] = ns; m[9] = ns; m[10] = ct; m[15] = polar1(-p[1]);
      return 4;
    }
    case YH_GATE_CUSTOM: {
      int d = op.dim;
      if (d <= 0 || !op.matrix) return -1;
      m.resize(size_t(d) * d);
      for (int k = 0; k < d * d; ++k) m[k] = C(op.matrix[2 * k], op.matrix[2 * k + 1]);
      return d;
    }
  }
  return -1;
}

bool gate_is_diagonal(const yh_op& op) {
  switch (op.kind) {
    case YH_GATE_Z: case YH_GATE_S: case YH_GATE_T: case YH_GATE_PHASE: case YH_GATE_RZ:
      return true;
    case YH_GATE_CUSTOM: return op.is_diag != 0;
    default: return false;
  }
}

int channel_num_qubits(const yh_op& op) {
  switch (op.kind) {
    case YH_CH_DEPOLARIZING: return op.params[1] > 0 ? int(op.params[1]) : op.nt;
    case YH_CH_COHERENT:
    case YH_CH_CUSTOM: {
      int d = op.dim, n = 0;
      while ((1 << n) < d) ++n;
      return n;
    }
    default: return 1;
  }
}

int validate_circuit(int nbits, const yh_op* ops, size_t nops) {
  if (nops && !ops) return YH_ERR_INVALID_ARG;
  for (size_t e = 0; e < nops; ++e) {
    const yh_op& op = ops[e];
    if (op.elem == YH_ELEM_GATE) {
      if (op.nt < 0 || op.nt > YH_MAX_TARGETS || op.nc < 0 || op.nc > YH_MAX_CTRLS)
        return YH_ERR_INVALID_ARG;
      if (op.kind < YH_GATE_X || op.kind > YH_GATE_CUSTOM) return YH_ERR_INVALID_ARG;
      if (op.nconf != op.nc) return YH_ERR_CONTROL_CONFIG_LEN;                // rule 1
      for (int k = 0; k < op.nt; ++k)                                          // rule 2
        if (op.targets[k] < 0 || op.targets[k] >= nbits) return YH_ERR_LOC_OUT_OF_RANGE;
      for (int k = 0; k < op.nc; ++k)
        if (op.ctrls[k] < 0 || op.ctrls[k] >= nbits) return YH_ERR_LOC_OUT_OF_RANGE;
      for (int a = 0; a < op.nt; ++a)                                          // rule 3
        for (int b = 0; b < op.nc; ++b)
          if (op.targets[a] == op.ctrls[b]) return YH_ERR_OVERLAPPING_LOCS;
      for (int k = 0; k < op.nc; ++k)
        if (op.ctrl_vals[k] != 0 && op.ctrl_vals[k] != 1) return YH_ERR_INVALID_ARG;
      int md;                                                                  // rule 6
      if (op.kind == YH_GATE_CUSTOM) {
        if (!op.matrix || op.dim <= 0) return YH_ERR_INVALID_ARG;
        md = op.dim;
      } else if (op.kind == YH_GATE_SWAP || op.kind == YH_GATE_ISWAP || op.kind == YH_GATE_FSIM) {
        md = 4;
      } else {
        md = 2;
      }
      if (op.nt >= 31 || md != (1 << op.nt)) return YH_ERR_MATRIX_SIZE;
      // duplicate target locs are not rejected by the reference's Circuit::new; the
      // instructs would then read and write one amplitude twice. They are refused here.
      for (int a = 0; a < op.nt; ++a)
        for (int b = 0; b < a; ++b)
          if (op.targets[a] == op.targets[b]) return YH_ERR_OVERLAPPING_LOCS;
      for (int a = 0; a < op.nc; ++a)
        for (int b = 0; b < a; ++b)
          if (op.ctrls[a] == op.ctrls[b]) return YH_ERR_OVERLAPPING_LOCS;
    } else if (op.elem == YH_ELEM_CHANNEL) {
      if (op.nt < 0 || op.nt > YH_MAX_TARGETS) return YH_ERR_INVALID_ARG;
      if (op.kind < YH_CH_BITFLIP || op.kind > YH_CH_CUSTOM) return YH_ERR_INVALID_ARG;
      if ((op.kind == YH_CH_COHERENT || op.kind == YH_CH_CUSTOM) && (!op.matrix || op.dim <= 0))
        return YH_ERR_INVALID_ARG;
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
  return YH_OK;
}

namespace {
bool in01(double p) { return p >= 0.0 && p <= 1.0; }

int validate_channel(const yh_op& op) {
  const double* p = op.params;
  switch (op.kind) {
    case YH_CH_BITFLIP: case YH_CH_PHASEFLIP: case YH_CH_DEPOLARIZING:
      return in01(p[0]) ? YH_OK : YH_ERR_CHANNEL_PARAM;
    case YH_CH_PAULI:
      return (p[0] >= 0 && p[1] >= 0 && p[2] >= 0 && p[0] + p[1] + p[2] <= 1.0) ? YH_OK
                                                                               : YH_ERR_CHANNEL_PARAM;
    case YH_CH_RESET:
      return (p[0] >= 0 && p[1] >= 0 && p[0] + p[1] <= 1.0) ? YH_OK : YH_ERR_CHANNEL_PARAM;
    case YH_CH_PHASE_AMPLITUDE_DAMPING:
      return (p[0] >= 0 && p[1] >= 0 && p[0] + p[1] <= 1.0 && in01(p[2])) ? YH_OK
                                                                         : YH_ERR_CHANNEL_PARAM;
    case YH_CH_AMPLITUDE_DAMPING: return (in01(p[0]) && in01(p[1])) ? YH_OK : YH_ERR_CHANNEL_PARAM;
    case YH_CH_PHASE_DAMPING: return in01(p[0]) ? YH_OK : YH_ERR_CHANNEL_PARAM;
    case YH_CH_THERMAL_RELAXATION:
      return (p[0] > 0 && p[1] > 0 && p[2] >= 0 && p[1] <= 2.0 * p[0] && in01(p[3]))
                 ? YH_OK : YH_ERR_CHANNEL_PARAM;
    case YH_CH_CUSTOM: return op.nmat >= 1 ? YH_OK : YH_ERR_CHANNEL_PARAM;
    case YH_CH_COHERENT: return YH_OK;
  }
  return YH_ERR_INVALID_ARG;
}

Mat M2(cd a, cd b, cd c, cd d) { return Mat{a, b, c, d}; }

// phase_amplitude_damping_kraus (noise.rs:383-458)
std::vector<Mat> pad_kraus(double a, double b, double p1) {
  const double eps = 2.220446049250313e-16;
  const cd z = C(0, 0);
  std::vector<Mat> ops;
  double rest = std::sqrt(1.0 - a - b);
  if (std::fabs(p1 - 1.0) > eps) {
    double s = std::sqrt(1.0 - p1);
    ops.push_back(M2(C(s, 0), z, z, C(s * rest, 0)));
    if (a > 0.0) ops.push_back(M2(z, C(std::sqrt(a) * s, 0), z, z));
    if (b > 0.0) ops.push_back(M2(z, z, z, C(std::sqrt(b) * s, 0)));
  }
  if (std::fabs(p1) > eps) {
    double s = std::sqrt(p1);
    ops.push_back(M2(C(s * rest, 0), z, z, C(s, 0)));
    if (a > 0.0) ops.push_back(M2(z, z, C(std::sqrt(a) * s, 0), z));
    if (b > 0.0) ops.push_back(M2(C(std::sqrt(b) * s, 0), z, z, z));
  }
  return ops;
}

std::vector<Mat> pauli_kraus(double px, double py, double pz) {
  const cd z = C(0, 0);
  double s0 = std::sqrt(1.0 - px - py - pz), sx = std::sqrt(px), sy = std::sqrt(py), sz = std::sqrt(pz);
  return {M2(C(s0, 0), z, z, C(s0, 0)), M2(z, C(sx, 0), C(sx, 0), z), M2(z, C(0, -sy), C(0, sy), z),
          M2(C(sz, 0), z, z, C(-sz, 0))};
}

// all n-qubit Pauli products, weights 1-p+p/4^n (identity) and p/4^n (noise.rs:462-508)
std::vector<Mat> depolarizing_multi(int n, double p) {
  const cd z = C(0, 0), o = C(1, 0);
  const Mat pauli[4] = {M2(o, z, z, o), M2(z, o, o, z), M2(z, C(0, -1), C(0, 1), z), M2(o, z, z, C(-1, 0))};
  uint64_t dim = uint64_t(1) << (2 * n);
  std::vector<Mat> ops;
  for (uint64_t idx = 0; idx < dim; ++idx) {
    Mat acc{o};
    int ad = 1;
    uint64_t t = idx;
    for (int q = 0; q < n; ++q) {
      const Mat& b = pauli[t % 4];
      t /= 4;
      Mat r(size_t(ad * 2) * (ad * 2));
      for (int i = 0; i < ad; ++i)
        for (int j = 0; j < ad; ++j)
          for (int k = 0; k < 2; ++k)
            for (int l = 0; l < 2; ++l) r[(i * 2 + k) * (ad * 2) + j * 2 + l] = acc[i * ad + j] * b[k * 2 + l];
      acc = r;
      ad *= 2;
    }
    double w = idx == 0 ? 1.0 - p + p / double(dim) : p / double(dim);
    double s = std::sqrt(w);
    for (auto& v : acc) v = v * C(s, 0);
    ops.push_back(acc);
  }
  return ops;
}
}  // namespace

int kraus_operators(const yh_op& op, std::vector<Mat>& ops, int& dim) {
  int st = validate_channel(op);
  if (st != YH_OK) return st;
  const double* p = op.params;
  const cd z = C(0, 0);
  dim = 2;
  ops.clear();
  switch (op.kind) {
    case YH_CH_PHASE_AMPLITUDE_DAMPING: ops = pad_kraus(p[0], p[1], p[2]); break;
    case YH_CH_AMPLITUDE_DAMPING: ops = pad_kraus(p[0], 0.0, p[1]); break;
    case YH_CH_PHASE_DAMPING: ops = pad_kraus(0.0, p[0], 0.0); break;
    case YH_CH_THERMAL_RELAXATION: {
      double t_phi = (p[0] * p[1]) / (2.0 * p[0] - p[1]);
      double a = 1.0 - std::exp(-p[2] / p[0]);
      double b = 1.0 - std::exp(-p[2] / t_phi);
      ops = pad_kraus(a, b, p[3]);
      break;
    }
    case YH_CH_BITFLIP: {
      double s0 = std::sqrt(1.0 - p[0]), s1 = std::sqrt(p[0]);
      ops = {M2(C(s0, 0), z, z, C(s0, 0)), M2(z, C(s1, 0), C(s1, 0), z)};
      break;
    }
    case YH_CH_PHASEFLIP: {
      double s0 = std::sqrt(1.0 - p[0]), s1 = std::sqrt(p[0]);
      ops = {M2(C(s0, 0), z, z, C(s0, 0)), M2(C(s1, 0), z, z, C(-s1, 0))};
      break;
    }
    case YH_CH_PAULI: ops = pauli_kraus(p[0], p[1], p[2]); break;
    case YH_CH_DEPOLARIZING: {
      int n = channel_num_qubits(op);
      if (n == 1) {
        double q = p[0] / 4.0;
        ops = pauli_kraus(q, q, q);
      } else {
        if (n > 4) return YH_ERR_UNSUPPORTED;
        ops = depolarizing_multi(n, p[0]);
        dim = 1 << n;
      }
      break;
    }
    case YH_CH_RESET: {
      double s = std::sqrt(1.0 - p[0] - p[1]);
      ops = {M2(C(s, 0), z, z, C(s, 0))};
      if (p[0] > 0.0) {
        double sp = std::sqrt(p[0]);
        ops.push_back(M2(C(sp, 0), z, z, z));
        ops.push_back(M2(z, C(sp, 0), z, z));
      }
      if (p[1] > 0.0) {
        double sp = std::sqrt(p[1]);
        ops.push_back(M2(z, z, z, C(sp, 0)));
        ops.push_back(M2(z, z, C(sp, 0), z));
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
        for (int e = 0; e < dim * dim; ++e) m[e] = C(src[2 * e], src[2 * e + 1]);
        ops.push_back(m);
      }
      break;
    }
    default: return YH_ERR_INVALID_ARG;
  }
  return YH_OK;
}

bool is_diagonal_matrix(const Mat& m, int d) {
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c)
      if (r != c && std::hypot(m[r * d + c].re, m[r * d + c].im) >= 1e-15) return false;
  return true;
}

std::vector<Term> read_terms(size_t nterms, const double* coeffs, const int32_t* term_len,
                             const int32_t* sites, const int32_t* ops) {
  std::vector<Term> t(nterms);
  size_t off = 0;
  for (size_t k = 0; k < nterms; ++k) {
    t[k].coeff = C(coeffs[2 * k], coeffs[2 * k + 1]);
    std::vector<std::pair<int, int>> v;
    for (int j = 0; j < term_len[k]; ++j) v.push_back({sites[off + j], ops[off + j]});
    off += size_t(term_len[k]);
    std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (auto& x : v)
      if (x.second != YH_OP_I) t[k].ops.push_back(x);
  }
  return t;
}

void op_row_table(int op, int src[2], cd fac[2]) {
  const cd z = C(0, 0), o = C(1, 0);
  switch (op) {
    case YH_OP_X: src[0] = 1; src[1] = 0; fac[0] = o; fac[1] = o; return;
    case YH_OP_Y: src[0] = 1; src[1] = 0; fac[0] = C(0, -1); fac[1] = C(0, 1); return;
    case YH_OP_Z: src[0] = 0; src[1] = 1; fac[0] = o; fac[1] = C(-1, 0); return;
    case YH_OP_P0: src[0] = 0; src[1] = 1; fac[0] = o; fac[1] = z; return;
    case YH_OP_P1: src[0] = 0; src[1] = 1; fac[0] = z; fac[1] = o; return;
    case YH_OP_PU: src[0] = 1; src[1] = 1; fac[0] = o; fac[1] = z; return;
    case YH_OP_PD: src[0] = 0; src[1] = 0; fac[0] = z; fac[1] = o; return;
    default: src[0] = 0; src[1] = 1; fac[0] = o; fac[1] = o; return;  // I
  }
}

// Hermitian eigenvalues through the real symmetric embedding S = [[A, -B], [B, A]]
// of H = A + iB (every eigenvalue of H appears twice in S): Householder reduction
// to tridiagonal form, then implicit-shift QL iterations (the textbook tred2/tqli
// pair, eigenvalues only), then every second sorted value.
std::vector<double> hermitian_eigenvalues(const Mat& m, int d) {
  const int n = 2 * d;
  std::vector<double> a(size_t(n) * n);
  auto A = [&](int r, int c) -> double& { return a[size_t(r) * n + c]; };
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c) {
      const cd h = m[size_t(r) * d + c];
      A(r, c) = h.re;
      A(r + d, c + d) = h.re;
      A(r, c + d) = -h.im;
      A(r + d, c) = h.im;
    }
  std::vector<double> diag(n), off(n, 0.0);
  // Householder tridiagonalisation (rows n-1 .. 1)
  for (int i = n - 1; i > 0; --i) {
    const int l = i - 1;
    double h = 0.0, scale = 0.0;
    if (l > 0) {
      for (int k = 0; k <= l; ++k) scale += std::fabs(A(i, k));
      if (scale == 0.0) {
        off[i] = A(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          A(i, k) /= scale;
          h += A(i, k) * A(i, k);
        }
        double f = A(i, l);
        const double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
        off[i] = scale * g;
        h -= f * g;
        A(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          double gg = 0.0;
          for (int k = 0; k <= j; ++k) gg += A(j, k) * A(i, k);
          for (int k = j + 1; k <= l; ++k) gg += A(k, j) * A(i, k);
          off[j] = gg / h;
          f += off[j] * A(i, j);
        }
        const double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          const double fj = A(i, j);
          const double gj = off[j] - hh * fj;
          off[j] = gj;
          for (int k = 0; k <= j; ++k) A(j, k) -= fj * off[k] + gj * A(i, k);
        }
      }
    } else {
      off[i] = A(i, l);
    }
    diag[i] = h;
  }
  for (int i = 0; i < n; ++i) diag[i] = A(i, i);
  // implicit QL on the tridiagonal matrix (diag, off[1..n-1])
  for (int i = 1; i < n; ++i) off[i - 1] = off[i];
  off[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0, mm;
    do {
      for (mm = l; mm < n - 1; ++mm) {
        const double dd = std::fabs(diag[mm]) + std::fabs(diag[mm + 1]);
        if (std::fabs(off[mm]) <= 1e-15 * dd) break;
      }
      if (mm != l) {
        if (++iter > 60) break;
        double g = (diag[l + 1] - diag[l]) / (2.0 * off[l]);
        double r = std::hypot(g, 1.0);
        g = diag[mm] - diag[l] + off[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = mm - 1; i >= l; --i) {
          double f = s * off[i];
          const double b = c * off[i];
          r = std::hypot(f, g);
          off[i + 1] = r;
          if (r == 0.0) {
            diag[i + 1] -= p;
            off[mm] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = diag[i + 1] - p;
          r = (diag[i] - g) * s + 2.0 * c * b;
          p = s * r;
          diag[i + 1] = g + p;
          g = c * r - b;
        }
        if (r == 0.0 && i >= l) continue;
        diag[l] -= p;
        off[l] = g;
        off[mm] = 0.0;
      }
    } while (mm != l);
  }
  std::sort(diag.begin(), diag.end());
  std::vector<double> ev(d);
  for (int k = 0; k < d; ++k) ev[k] = diag[2 * k];
  return ev;
}

}  // namespace yh
