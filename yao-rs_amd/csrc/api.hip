// api.hip — the C ABI (include/yaohip.h): device handles, the per-gate dispatch
// planner and the measurement / density-matrix entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include <rocsolver/rocsolver.h>

#include <type_traits>

#include "host.h"
#include "kernels.h"
#include "runtime.h"

namespace yh {

static thread_local std::string g_last_error;
void set_error(const std::string& m) { g_last_error = m; }
int fail(int status, const std::string& m) {
  g_last_error = m;
  return status;
}

// ------------------------------ device buffers ------------------------------
DevBuf::~DevBuf() { release(); }
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}
int DevBuf::ensure(size_t b) {
  if (b <= bytes) return YH_OK;
  release();
  if (b == 0) return YH_OK;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? YH_ERR_OUT_OF_MEMORY : YH_ERR_HIP,
                std::string("hipMalloc(") + std::to_string(b) + ") failed: " + hipGetErrorString(e));
  }
  bytes = b;
  return YH_OK;
}
PinBuf::~PinBuf() {
  if (p) (void)hipHostFree(p);
}
int PinBuf::ensure(size_t b) {
  if (b <= bytes) return YH_OK;
  if (p) (void)hipHostFree(p);
  p = nullptr;
  bytes = 0;
  hipError_t e = hipHostMalloc(&p, b, hipHostMallocDefault);
  if (e != hipSuccess) {
    p = nullptr;
    return fail(YH_ERR_OUT_OF_MEMORY, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
  }
  bytes = b;
  return YH_OK;
}

Workspace::~Workspace() {
  if (stream) (void)hipStreamDestroy(stream);
}
int Workspace::init() {
  YH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  YH_TRY(red.ensure(sizeof(double) * (size_t(kRedBlocks) * 2 * kMaxTermsPerPass + 256)));
  return YH_OK;
}
double* Workspace::partial() { return static_cast<double*>(red.p); }
double* Workspace::result() { return static_cast<double*>(red.p) + size_t(kRedBlocks) * 2 * kMaxTermsPerPass; }
int Workspace::sync() {
  YH_HIP(hipGetLastError());
  YH_HIP(hipStreamSynchronize(stream));
  return time_collect();
}
int Workspace::time_begin(int kclass) {
  if (!timing) return YH_OK;
  TimedLaunch t{kclass, nullptr, nullptr};
  YH_HIP(hipEventCreate(&t.a));
  YH_HIP(hipEventCreate(&t.b));
  YH_HIP(hipEventRecord(t.a, stream));
  pending.push_back(t);
  return YH_OK;
}
int Workspace::time_end() {
  if (!timing || pending.empty()) return YH_OK;
  YH_HIP(hipEventRecord(pending.back().b, stream));
  return YH_OK;
}
int Workspace::time_collect() {
  for (auto& t : pending) {
    float ms = 0.f;
    YH_HIP(hipEventSynchronize(t.b));
    YH_HIP(hipEventElapsedTime(&ms, t.a, t.b));
    t_count[t.kclass] += 1;
    t_ms[t.kclass] += double(ms);
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  pending.clear();
  return YH_OK;
}
int Workspace::upload(const void* host, size_t bytes, DevBuf& dst) {
  YH_TRY(dst.ensure(bytes));
  if (bytes == 0) return YH_OK;
  YH_TRY(pin.ensure(bytes));
  // the staging buffer may still feed an earlier copy: drain the stream first
  YH_HIP(hipStreamSynchronize(stream));
  std::memcpy(pin.p, host, bytes);
  YH_HIP(hipMemcpyAsync(dst.p, pin.p, bytes, hipMemcpyHostToDevice, stream));
  return YH_OK;
}
int Workspace::fetch(void* host, const void* dev, size_t bytes) {
  YH_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream));
  YH_HIP(hipStreamSynchronize(stream));
  return YH_OK;
}

static bool g_inited = false;
int ensure_init() {
  if (g_inited) return YH_OK;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) return fail(YH_ERR_NO_DEVICE, "no HIP device visible");
  YH_HIP(hipSetDevice(0));
  hipDeviceProp_t prop;
  YH_HIP(hipGetDeviceProperties(&prop, 0));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(YH_ERR_NO_DEVICE, std::string("device 0 is ") + prop.gcnArchName + ", this build targets gfx950");
  g_inited = true;
  return YH_OK;
}

// ------------------------------ gate planning -------------------------------
// Arena of kernel operands (matrices, offset tables) uploaded once per call.
template <class R>
struct Arena {
  using V = typename Cplx<R>::V;
  std::vector<unsigned char> bytes;
  size_t push_mat(const Mat& m) {
    size_t off = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(off + m.size() * sizeof(V));
    V* dst = reinterpret_cast<V*>(bytes.data() + off);
    for (size_t k = 0; k < m.size(); ++k) dst[k] = cmk<V>(R(m[k].re), R(m[k].im));
    return off;
  }
  size_t push_u64(const std::vector<uint64_t>& v) {
    size_t off = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(off + v.size() * sizeof(uint64_t));
    std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(uint64_t));
    return off;
  }
};

enum LaunchType { L_1Q, L_DENSE, L_DIAG, L_X, L_SWAP, L_OOP };

template <class R>
struct Launch {
  LaunchType type;
  int nt = 0;
  int bit = 0;
  GroupArgs ga{};
  uint64_t m0 = 0, m1 = 0;
  size_t moff = 0, ooff = 0;
  Mat2<R> m2{};
  DiagArgs<R> dg{};
};

static void set_locked(GroupArgs& ga, int nbits, std::vector<int> pos) {
  std::sort(pos.begin(), pos.end());
  ga.lk.npos = int(pos.size());
  for (size_t k = 0; k < pos.size(); ++k) ga.lk.pos[k] = uint8_t(pos[k]);
  ga.ngroups = uint64_t(1) << (nbits - int(pos.size()));
}

// dispatch_arrayreg_gate (apply.rs:7-120): classify one gate and record its launches.
template <class R>
static int plan_gate(int nbits, const yh_op& op, std::vector<Launch<R>>& L, Arena<R>& arena) {
  using V = typename Cplx<R>::V;
  auto bit_of = [nbits](int loc) { return nbits - 1 - loc; };
  std::vector<int> cbits;
  uint64_t ctrl_set = 0;
  for (int k = 0; k < op.nc; ++k) {
    cbits.push_back(bit_of(op.ctrls[k]));
    if (op.ctrl_vals[k]) ctrl_set |= uint64_t(1) << bit_of(op.ctrls[k]);
  }
  const bool has_c = op.nc > 0;
  if (op.kind == YH_GATE_X) {  // instruct_x / instruct_x_controlled per target
    for (int k = 0; k < op.nt; ++k) {
      Launch<R> l;
      l.type = L_X;
      std::vector<int> pos = cbits;
      pos.push_back(bit_of(op.targets[k]));
      set_locked(l.ga, nbits, pos);
      l.ga.ctrl_set = ctrl_set;
      l.m0 = uint64_t(1) << bit_of(op.targets[k]);
      L.push_back(l);
    }
    return YH_OK;
  }
  if (op.kind == YH_GATE_SWAP && !has_c) {  // instruct_swap
    Launch<R> l;
    l.type = L_SWAP;
    set_locked(l.ga, nbits, {bit_of(op.targets[0]), bit_of(op.targets[1])});
    l.m0 = uint64_t(1) << bit_of(op.targets[0]);
    l.m1 = uint64_t(1) << bit_of(op.targets[1]);
    L.push_back(l);
    return YH_OK;
  }
  Mat m;
  int d = gate_matrix(op, m);
  if (d != (1 << op.nt)) return fail(YH_ERR_MATRIX_SIZE, "gate matrix size does not match targets");
  const bool diag = op.kind != YH_GATE_SWAP && gate_is_diagonal(op);
  if (diag && op.nt <= 2) {  // instruct_1q_diag / instruct_2q_diag [_controlled]
    Launch<R> l;
    l.type = L_DIAG;
    l.nt = op.nt;
    set_locked(l.ga, nbits, cbits);
    l.ga.ctrl_set = ctrl_set;
    for (int j = 0; j < d; ++j) l.dg.d[j] = cmk<V>(R(m[j * d + j].re), R(m[j * d + j].im));
    for (int k = 0; k < op.nt; ++k) l.dg.tbit[k] = bit_of(op.targets[k]);
    L.push_back(l);
    return YH_OK;
  }
  if (op.nt == 1 && !has_c) {  // instruct_1q
    Launch<R> l;
    l.type = L_1Q;
    l.bit = bit_of(op.targets[0]);
    l.m2.a = cmk<V>(R(m[0].re), R(m[0].im));
    l.m2.b = cmk<V>(R(m[1].re), R(m[1].im));
    l.m2.c = cmk<V>(R(m[2].re), R(m[2].im));
    l.m2.d = cmk<V>(R(m[3].re), R(m[3].im));
    L.push_back(l);
    return YH_OK;
  }
  // instruct_1q_controlled, instruct_2q[_controlled], instruct_nq
  Launch<R> l;
  l.nt = op.nt;
  std::vector<int> pos = cbits;
  std::vector<uint64_t> offs(size_t(1) << op.nt, 0);
  for (int k = 0; k < op.nt; ++k) pos.push_back(bit_of(op.targets[k]));
  for (size_t j = 0; j < offs.size(); ++j)
    for (int k = 0; k < op.nt; ++k)
      if ((j >> k) & 1) offs[j] |= uint64_t(1) << bit_of(op.targets[k]);
  set_locked(l.ga, nbits, pos);
  l.ga.ctrl_set = ctrl_set;
  l.moff = arena.push_mat(m);
  if (op.nt <= 5) {
    l.type = L_DENSE;
    for (size_t j = 0; j < offs.size(); ++j) l.ga.offs[j] = offs[j];
  } else {
    l.type = L_OOP;
    l.ooff = arena.push_u64(offs);
  }
  L.push_back(l);
  return YH_OK;
}

template <class R>
static int run_launches(Workspace& w, typename Cplx<R>::V* s, int nbits, std::vector<Launch<R>>& L,
                        Arena<R>& arena) {
  using V = typename Cplx<R>::V;
  YH_TRY(w.upload(arena.bytes.data(), arena.bytes.size(), w.mats));
  const unsigned char* base = static_cast<const unsigned char*>(w.mats.p);
  static const int kclass_of[] = {YH_K_1Q, YH_K_DENSE, YH_K_DIAG, YH_K_X, YH_K_SWAP, YH_K_OOP};
  for (auto& l : L) {
    if (l.type == L_OOP) {
      size_t bytes = (size_t(1) << nbits) * sizeof(V);
      YH_TRY(w.scratch.ensure(bytes));
      YH_HIP(hipMemcpyAsync(w.scratch.p, s, bytes, hipMemcpyDeviceToDevice, w.stream));
    }
    YH_TRY(w.time_begin(kclass_of[l.type]));
    switch (l.type) {
      case L_1Q: YH_TRY(launch_1q<R>(s, nbits, l.bit, l.m2, w.stream)); break;
      case L_DENSE:
        YH_TRY(launch_dense<R>(s, l.ga, l.nt, reinterpret_cast<const V*>(base + l.moff), w.stream));
        break;
      case L_DIAG: YH_TRY(launch_diag<R>(s, l.ga, l.nt, l.dg, w.stream)); break;
      case L_X: YH_TRY(launch_xflip<R>(s, l.ga, l.m0, w.stream)); break;
      case L_SWAP: YH_TRY(launch_swap<R>(s, l.ga, l.m0, l.m1, w.stream)); break;
      case L_OOP:
        YH_TRY(launch_dense_oop<R>(s, static_cast<const V*>(w.scratch.p), l.ga, l.nt,
                                   reinterpret_cast<const V*>(base + l.moff),
                                   reinterpret_cast<const uint64_t*>(base + l.ooff), w.stream));
        break;
    }
    YH_TRY(w.time_end());
  }
  return w.sync();
}

template <class R>
int apply_gates(Workspace& w, void* state, int nbits, const yh_op* ops, size_t nops) {
  std::vector<Launch<R>> L;
  Arena<R> arena;
  for (size_t e = 0; e < nops; ++e)
    if (ops[e].elem == YH_ELEM_GATE) YH_TRY(plan_gate<R>(nbits, ops[e], L, arena));
  if (L.empty()) return YH_OK;
  return run_launches<R>(w, static_cast<typename Cplx<R>::V*>(state), nbits, L, arena);
}
template int apply_gates<double>(Workspace&, void*, int, const yh_op*, size_t);
template int apply_gates<float>(Workspace&, void*, int, const yh_op*, size_t);

// ------------------------------ terms ---------------------------------------
// Lower terms to device form. lsb_sites: sites are bit positions (expect_dm);
// otherwise loc l is bit nbits-1-l. Sites beyond the register contribute the
// operator's [0][0] entry (expect_dm reads bit 0 there) and are dropped.
int lower_terms(int nbits, const std::vector<Term>& terms, bool lsb_sites, std::vector<TermDev>& out,
                       std::vector<cd>& coeffs) {
  out.clear();
  coeffs.clear();
  for (const Term& t : terms) {
    TermDev d{};
    cd c = t.coeff;
    int ns = 0;
    for (auto& so : t.ops) {
      if (so.second < YH_OP_I || so.second > YH_OP_PD) return fail(YH_ERR_INVALID_ARG, "bad operator code");
      if (so.first < 0) return fail(YH_ERR_LOC_OUT_OF_RANGE, "negative operator site");
      if (so.first >= nbits) {
        if (!lsb_sites) return fail(YH_ERR_LOC_OUT_OF_RANGE, "operator site beyond the register");
        int src[2];
        cd fac[2];
        op_row_table(so.second, src, fac);
        cd m00 = src[0] == 0 ? fac[0] : cd{0, 0};
        c = c * m00;
        continue;
      }
      if (ns >= kMaxSites) return fail(YH_ERR_UNSUPPORTED, "too many sites in one operator string");
      int src[2];
      cd fac[2];
      op_row_table(so.second, src, fac);
      d.bit[ns] = lsb_sites ? so.first : nbits - 1 - so.first;
      for (int v = 0; v < 2; ++v) {
        d.src[ns][v] = uint8_t(src[v]);
        d.fac[ns][v][0] = fac[v].re;
        d.fac[ns][v][1] = fac[v].im;
      }
      ++ns;
    }
    d.nsites = ns;
    out.push_back(d);
    coeffs.push_back(c);
  }
  return YH_OK;
}

}  // namespace yh

using namespace yh;

// ====================================== C ABI ======================================
#define YH_GUARD(cond, st, msg) \
  do {                          \
    if (!(cond)) return fail(st, msg); \
  } while (0)

extern "C" {

const char* yh_last_error(void) { return g_last_error.c_str(); }

int yh_init(void) { return ensure_init(); }

int yh_device_name(char* buf, size_t len) {
  YH_TRY(ensure_init());
  hipDeviceProp_t prop;
  YH_HIP(hipGetDeviceProperties(&prop, 0));
  std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return YH_OK;
}

int yh_copy_bandwidth(size_t bytes, int reps, double* gbps) {
  YH_GUARD(gbps && bytes > 0 && reps >= 1, YH_ERR_INVALID_ARG, "bytes > 0, reps >= 1 and an output pointer");
  YH_TRY(ensure_init());
  DevBuf a, b;
  YH_TRY(a.ensure(bytes));
  YH_TRY(b.ensure(bytes));
  hipStream_t st;
  YH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  YH_HIP(hipEventCreate(&e0));
  YH_HIP(hipEventCreate(&e1));
  float best = 0.0f;
  hipError_t err = hipMemsetAsync(a.p, 1, bytes, st);
  for (int r = 0; err == hipSuccess && r <= reps; ++r) {  // r = 0 warms up
    err = hipEventRecord(e0, st);
    if (err == hipSuccess) err = hipMemcpyAsync(b.p, a.p, bytes, hipMemcpyDeviceToDevice, st);
    if (err == hipSuccess) err = hipEventRecord(e1, st);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (err == hipSuccess && r > 0 && (best == 0.0f || ms < best)) best = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  if (err != hipSuccess) return fail(YH_ERR_HIP, std::string("copy timing failed: ") + hipGetErrorString(err));
  *gbps = 2.0 * double(bytes) / (double(best) * 1e-3) / 1e9;
  return YH_OK;
}

int yh_reg_create(int nbits, int dtype, yh_reg** out) {
  YH_GUARD(out, YH_ERR_INVALID_ARG, "null output pointer");
  YH_GUARD(nbits >= 1 && nbits <= 35, YH_ERR_INVALID_ARG, "nbits must be in [1, 35]");
  YH_GUARD(dtype == YH_C128 || dtype == YH_C64, YH_ERR_DTYPE, "unknown dtype");
  YH_TRY(ensure_init());
  yh_reg* r = new yh_reg();
  r->nbits = nbits;
  r->dtype = dtype;
  int st = r->w.init();
  if (st == YH_OK) st = r->state.ensure(r->bytes());
  if (st != YH_OK) {
    delete r;
    return st;
  }
  // zero_state (register.rs:21-26)
  YH_HIP(hipMemsetAsync(r->state.p, 0, r->bytes(), r->w.stream));
  if (dtype == YH_C128) {
    const double one[2] = {1.0, 0.0};
    YH_HIP(hipMemcpyAsync(r->state.p, one, sizeof(one), hipMemcpyHostToDevice, r->w.stream));
  } else {
    const float one[2] = {1.0f, 0.0f};
    YH_HIP(hipMemcpyAsync(r->state.p, one, sizeof(one), hipMemcpyHostToDevice, r->w.stream));
  }
  st = r->w.sync();
  if (st != YH_OK) {
    delete r;
    return st;
  }
  *out = r;
  return YH_OK;
}

static int reg_norm2(yh_reg* reg, double* sumsq_dev);
static int reg_scale(yh_reg* reg, const double* sumsq_dev, double thresh);

// amplitude idx = (re, 0), written with one host-to-device copy on the register's stream
static int reg_set_amp(yh_reg* r, uint64_t idx, double re) {
  if (r->dtype == YH_C128) {
    const double v[2] = {re, 0.0};
    YH_HIP(hipMemcpyAsync(static_cast<double2*>(r->state.p) + idx, v, sizeof(v), hipMemcpyHostToDevice, r->w.stream));
  } else {
    const float v[2] = {float(re), 0.0f};
    YH_HIP(hipMemcpyAsync(static_cast<float2*>(r->state.p) + idx, v, sizeof(v), hipMemcpyHostToDevice, r->w.stream));
  }
  return r->w.sync();
}

// a register of nbits qubits whose amplitudes `fill` writes on its stream
static int reg_make(int nbits, int dtype, yh_reg** out, const std::function<int(yh_reg*)>& fill) {
  YH_GUARD(out, YH_ERR_INVALID_ARG, "null output pointer");
  YH_GUARD(nbits >= 1 && nbits <= 35, YH_ERR_INVALID_ARG, "nbits must be in [1, 35]");
  YH_GUARD(dtype == YH_C128 || dtype == YH_C64, YH_ERR_DTYPE, "unknown dtype");
  YH_TRY(ensure_init());
  std::unique_ptr<yh_reg> r(new yh_reg());
  r->nbits = nbits;
  r->dtype = dtype;
  YH_TRY(r->w.init());
  YH_TRY(r->state.ensure(r->bytes()));
  YH_TRY(fill(r.get()));
  YH_TRY(r->w.sync());
  *out = r.release();
  return YH_OK;
}

// amplitude `idx` = (re, 0), the rest zero
static int reg_set_onehot(yh_reg* r, uint64_t idx, double re) {
  YH_HIP(hipMemsetAsync(r->state.p, 0, r->bytes(), r->w.stream));
  return reg_set_amp(r, idx, re);
}

int yh_reg_product_state(int nbits, int dtype, uint64_t index, yh_reg** out) {
  YH_GUARD(nbits >= 1 && nbits < 64 && index < (uint64_t(1) << nbits), YH_ERR_INVALID_ARG,
           "product_state: the bit string does not fit in nbits");
  return reg_make(nbits, dtype, out, [&](yh_reg* r) { return reg_set_onehot(r, index, 1.0); });
}

int yh_reg_uniform_state(int nbits, int dtype, yh_reg** out) {
  return reg_make(nbits, dtype, out, [&](yh_reg* r) {
    const double amp = 1.0 / std::sqrt(double(uint64_t(1) << nbits));
    const uint64_t n = uint64_t(1) << nbits;
    if (dtype == YH_C128) return launch_fill<double>(static_cast<double2*>(r->state.p), n, amp, r->w.stream);
    return launch_fill<float>(static_cast<float2*>(r->state.p), n, amp, r->w.stream);
  });
}

int yh_reg_ghz_state(int nbits, int dtype, yh_reg** out) {
  return reg_make(nbits, dtype, out, [&](yh_reg* r) {
    const double amp = 1.0 / std::sqrt(2.0);
    YH_TRY(reg_set_onehot(r, 0, amp));
    return reg_set_amp(r, (uint64_t(1) << nbits) - 1, amp);
  });
}

int yh_reg_deterministic_state(int nbits, int dtype, yh_reg** out) {
  return reg_make(nbits, dtype, out, [&](yh_reg* r) {
    const uint64_t n = uint64_t(1) << nbits;
    if (dtype == YH_C128) YH_TRY(launch_deterministic<double>(static_cast<double2*>(r->state.p), n, r->w.stream));
    else YH_TRY(launch_deterministic<float>(static_cast<float2*>(r->state.p), n, r->w.stream));
    YH_TRY(reg_norm2(r, r->w.result()));
    return reg_scale(r, r->w.result(), 0.0);
  });
}

int yh_reg_rand_state(int nbits, int dtype, const double* uniforms, size_t n_uniforms, yh_reg** out) {
  YH_GUARD(uniforms, YH_ERR_INVALID_ARG, "null uniforms");
  YH_GUARD(nbits >= 1 && nbits <= 35 && n_uniforms == (size_t(2) << nbits), YH_ERR_INVALID_ARG,
           "rand_state: 2 * 2^nbits uniforms (re, im per amplitude)");
  return reg_make(nbits, dtype, out, [&](yh_reg* r) {
    const uint64_t n = uint64_t(1) << nbits;
    YH_TRY(r->w.upload(uniforms, n_uniforms * sizeof(double), r->w.scratch));
    const double* u = static_cast<const double*>(r->w.scratch.p);
    if (dtype == YH_C128) YH_TRY(launch_from_uniforms<double>(static_cast<double2*>(r->state.p), u, n, r->w.stream));
    else YH_TRY(launch_from_uniforms<float>(static_cast<float2*>(r->state.p), u, n, r->w.stream));
    YH_TRY(reg_norm2(r, r->w.result()));
    return reg_scale(r, r->w.result(), 0.0);
  });
}

int yh_reg_destroy(yh_reg* reg) {
  delete reg;
  return YH_OK;
}

int yh_reg_nbits(const yh_reg* reg, int* nbits) {
  YH_GUARD(reg && nbits, YH_ERR_INVALID_ARG, "null argument");
  *nbits = reg->nbits;
  return YH_OK;
}

int yh_reg_dtype(const yh_reg* reg, int* dtype) {
  YH_GUARD(reg && dtype, YH_ERR_INVALID_ARG, "null argument");
  *dtype = reg->dtype;
  return YH_OK;
}

int yh_reg_device_ptr(const yh_reg* reg, void** ptr) {
  YH_GUARD(reg && ptr, YH_ERR_INVALID_ARG, "null argument");
  *ptr = reg->state.p;
  return YH_OK;
}

int yh_reg_upload(yh_reg* reg, const double* amps, size_t n_amps) {
  YH_GUARD(reg && amps, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(n_amps == (size_t(1) << reg->nbits), YH_ERR_INVALID_ARG, "from_vec: length must be 2^nbits");
  if (reg->dtype == YH_C128) {
    YH_HIP(hipMemcpy(reg->state.p, amps, reg->bytes(), hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(2 * n_amps);
    for (size_t k = 0; k < 2 * n_amps; ++k) f[k] = float(amps[k]);
    YH_HIP(hipMemcpy(reg->state.p, f.data(), reg->bytes(), hipMemcpyHostToDevice));
  }
  return YH_OK;
}

int yh_reg_download(const yh_reg* reg, double* amps, size_t n_amps) {
  YH_GUARD(reg && amps, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(n_amps == (size_t(1) << reg->nbits), YH_ERR_INVALID_ARG, "length must be 2^nbits");
  YH_HIP(hipStreamSynchronize(reg->w.stream));
  if (reg->dtype == YH_C128) {
    YH_HIP(hipMemcpy(amps, reg->state.p, reg->bytes(), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> f(2 * n_amps);
    YH_HIP(hipMemcpy(f.data(), reg->state.p, reg->bytes(), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < 2 * n_amps; ++k) amps[k] = double(f[k]);
  }
  return YH_OK;
}

int yh_reg_clone(const yh_reg* src, yh_reg** out) {
  YH_GUARD(src && out, YH_ERR_INVALID_ARG, "null argument");
  yh_reg* r = nullptr;
  YH_TRY(yh_reg_create(src->nbits, src->dtype, &r));
  int st = yh_reg_copy(r, src);
  if (st != YH_OK) {
    delete r;
    return st;
  }
  *out = r;
  return YH_OK;
}

int yh_reg_copy(yh_reg* dst, const yh_reg* src) {
  YH_GUARD(dst && src, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(dst->nbits == src->nbits && dst->dtype == src->dtype, YH_ERR_INVALID_ARG, "shape mismatch");
  YH_HIP(hipStreamSynchronize(src->w.stream));
  YH_HIP(hipMemcpyAsync(dst->state.p, src->state.p, src->bytes(), hipMemcpyDeviceToDevice, dst->w.stream));
  return dst->w.sync();
}

static int reg_norm2(yh_reg* reg, double* sumsq_dev) {
  uint64_t n = uint64_t(1) << reg->nbits;
  if (reg->dtype == YH_C128)
    return launch_norm2<double>(static_cast<const double2*>(reg->state.p), n, reg->w.partial(), sumsq_dev,
                                reg->w.stream);
  return launch_norm2<float>(static_cast<const float2*>(reg->state.p), n, reg->w.partial(), sumsq_dev,
                             reg->w.stream);
}

static int reg_scale(yh_reg* reg, const double* sumsq_dev, double thresh) {
  uint64_t n = uint64_t(1) << reg->nbits;
  if (reg->dtype == YH_C128)
    return launch_scale<double>(static_cast<double2*>(reg->state.p), n, sumsq_dev, thresh, reg->w.stream);
  return launch_scale<float>(static_cast<float2*>(reg->state.p), n, sumsq_dev, thresh, reg->w.stream);
}

int yh_reg_norm(const yh_reg* creg, double* out) {
  YH_GUARD(creg && out, YH_ERR_INVALID_ARG, "null argument");
  yh_reg* reg = const_cast<yh_reg*>(creg);
  YH_TRY(reg_norm2(reg, reg->w.result()));
  double s = 0;
  YH_TRY(reg->w.fetch(&s, reg->w.result(), sizeof(double)));
  *out = std::sqrt(s);
  return YH_OK;
}

int yh_reg_normalize(yh_reg* reg) {
  YH_GUARD(reg, YH_ERR_INVALID_ARG, "null argument");
  YH_TRY(reg_norm2(reg, reg->w.result()));
  YH_TRY(reg_scale(reg, reg->w.result(), 0.0));
  return reg->w.sync();
}

int yh_reg_fidelity(const yh_reg* a, const yh_reg* b, double* out) {
  YH_GUARD(a && b && out, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(a->nbits == b->nbits && a->dtype == b->dtype, YH_ERR_INVALID_ARG, "fidelity: register mismatch");
  yh_reg* r = const_cast<yh_reg*>(a);
  YH_HIP(hipStreamSynchronize(b->w.stream));
  uint64_t n = uint64_t(1) << a->nbits;
  if (a->dtype == YH_C128)
    YH_TRY(launch_inner<double>(static_cast<const double2*>(a->state.p), static_cast<const double2*>(b->state.p),
                                n, r->w.partial(), r->w.result(), r->w.stream));
  else
    YH_TRY(launch_inner<float>(static_cast<const float2*>(a->state.p), static_cast<const float2*>(b->state.p), n,
                               r->w.partial(), r->w.result(), r->w.stream));
  double v[2];
  YH_TRY(r->w.fetch(v, r->w.result(), sizeof(v)));
  *out = v[0] * v[0] + v[1] * v[1];
  return YH_OK;
}

int yh_validate(int nbits, const yh_op* ops, size_t nops) {
  int st = validate_circuit(nbits, ops, nops);
  if (st != YH_OK) {
    const char* what[] = {"ok", "invalid argument", "control_configs length does not match control_locs",
                          "location out of range", "overlapping locations", "matrix size mismatch"};
    return fail(st, st <= 5 ? what[st] : "invalid circuit");
  }
  return YH_OK;
}

int yh_gate_matrix(const yh_op* op, double* out, int* dim) {
  YH_GUARD(op && out && dim, YH_ERR_INVALID_ARG, "null argument");
  Mat m;
  int d = gate_matrix(*op, m);
  YH_GUARD(d > 0, YH_ERR_INVALID_ARG, "unknown gate");
  for (size_t k = 0; k < m.size(); ++k) {
    out[2 * k] = m[k].re;
    out[2 * k + 1] = m[k].im;
  }
  *dim = d;
  return YH_OK;
}

int yh_apply(yh_reg* reg, int circuit_nbits, const yh_op* ops, size_t nops) {
  YH_GUARD(reg, YH_ERR_INVALID_ARG, "null register");
  YH_GUARD(circuit_nbits == reg->nbits, YH_ERR_NBITS_MISMATCH, "Register and circuit qubit count mismatch");
  YH_TRY(yh_validate(circuit_nbits, ops, nops));
  if (reg->dtype == YH_C128) return apply_gates<double>(reg->w, reg->state.p, reg->nbits, ops, nops);
  return apply_gates<float>(reg->w, reg->state.p, reg->nbits, ops, nops);
}

int yh_apply_fused(yh_reg* reg, int circuit_nbits, const yh_op* ops, size_t nops, int max_fuse_qubits) {
  YH_GUARD(reg, YH_ERR_INVALID_ARG, "null register");
  YH_GUARD(max_fuse_qubits >= 0 && max_fuse_qubits <= 5, YH_ERR_INVALID_ARG, "max_fuse_qubits must be in [0, 5]");
  YH_GUARD(circuit_nbits == reg->nbits, YH_ERR_NBITS_MISMATCH, "Register and circuit qubit count mismatch");
  YH_TRY(yh_validate(circuit_nbits, ops, nops));
  if (max_fuse_qubits == 5 || max_fuse_qubits == 0) {  // 5-qubit blocks, one MFMA pass each (default)
    if (reg->dtype == YH_C128) return apply_blocks<double>(reg->w, reg->state.p, reg->nbits, ops, nops);
    return apply_blocks<float>(reg->w, reg->state.p, reg->nbits, ops, nops);
  }
  if (reg->nbits < 12) {  // smaller than one tile: the per-gate path
    if (reg->dtype == YH_C128) return apply_gates<double>(reg->w, reg->state.p, reg->nbits, ops, nops);
    return apply_gates<float>(reg->w, reg->state.p, reg->nbits, ops, nops);
  }
  if (reg->dtype == YH_C128) return apply_tiled<double>(reg->w, reg->state.p, reg->nbits, ops, nops);
  return apply_tiled<float>(reg->w, reg->state.p, reg->nbits, ops, nops);
}

int yh_plan_stats(int nbits, const yh_op* ops, size_t nops, int64_t out[4]) {
  YH_GUARD(out, YH_ERR_INVALID_ARG, "null argument");
  YH_TRY(yh_validate(nbits, ops, nops));
  return plan_stats(nbits, ops, nops, out);
}

int yh_block_plan(int nbits, const yh_op* ops, size_t nops, int32_t* bits, size_t cap, size_t* nsteps) {
  YH_GUARD(nsteps, YH_ERR_INVALID_ARG, "null argument");
  YH_TRY(yh_validate(nbits, ops, nops));
  return block_plan_bits(nbits, ops, nops, bits, cap, nsteps);
}

int yh_block_plan_stats(int nbits, const yh_op* ops, size_t nops, int64_t out[3]) {
  YH_GUARD(out, YH_ERR_INVALID_ARG, "null argument");
  YH_TRY(yh_validate(nbits, ops, nops));
  return block_plan_stats(nbits, ops, nops, out);
}

int yh_block_pass_stats(int nbits, const yh_op* ops, size_t nops, int dm, int64_t out[7]) {
  YH_GUARD(out, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(dm == 0 || dm == 1, YH_ERR_INVALID_ARG, "dm must be 0 or 1");
  YH_TRY(yh_validate(nbits, ops, nops));
  return block_pass_stats(nbits, ops, nops, dm, out);
}

int yh_instruct(yh_reg* reg, const yh_op* op) {
  YH_GUARD(reg && op, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(op->elem == YH_ELEM_GATE, YH_ERR_INVALID_ARG, "yh_instruct takes a gate element");
  return yh_apply(reg, reg->nbits, op, 1);
}

int yh_channel_kraus(const yh_op* ch, double* out, int max_ops, int* nops, int* dim) {
  YH_GUARD(ch && out && nops && dim, YH_ERR_INVALID_ARG, "null argument");
  std::vector<Mat> kr;
  int d = 0;
  int st = kraus_operators(*ch, kr, d);
  if (st != YH_OK) return fail(st, "invalid channel parameters");
  YH_GUARD(int(kr.size()) <= max_ops, YH_ERR_INVALID_ARG, "output too small for the Kraus operators");
  size_t off = 0;
  for (auto& K : kr)
    for (auto& v : K) {
      out[off++] = v.re;
      out[off++] = v.im;
    }
  *nops = int(kr.size());
  *dim = d;
  return YH_OK;
}

static int expect_impl(Workspace& w, int nbits, const void* data, int dtype, bool dm, size_t nterms,
                       const double* coeffs, const int32_t* term_len, const int32_t* sites, const int32_t* ops,
                       double out[2]) {
  YH_GUARD(out && (nterms == 0 || (coeffs && term_len)), YH_ERR_INVALID_ARG, "null argument");
  size_t total = 0;
  for (size_t k = 0; k < nterms; ++k) {
    YH_GUARD(term_len[k] >= 0, YH_ERR_INVALID_ARG, "negative term length");
    total += size_t(term_len[k]);
  }
  YH_GUARD(total == 0 || (sites && ops), YH_ERR_INVALID_ARG, "null sites/ops");
  std::vector<Term> terms = read_terms(nterms, coeffs, term_len, sites, ops);
  std::vector<TermDev> td;
  std::vector<cd> cf;
  YH_TRY(lower_terms(nbits, terms, dm, td, cf));
  if (!td.empty()) YH_TRY(w.upload(td.data(), td.size() * sizeof(TermDev), w.terms));
  cd acc{0, 0};
  const uint64_t n = uint64_t(1) << nbits;
  for (size_t t0 = 0; t0 < td.size(); t0 += kMaxTermsPerPass) {
    int nt = int(std::min<size_t>(kMaxTermsPerPass, td.size() - t0));
    const TermDev* tp = static_cast<const TermDev*>(w.terms.p) + t0;
    if (dm)
      YH_TRY(launch_dm_expect_terms(static_cast<const double2*>(data), n, tp, nt, w.partial(), w.result(), w.stream));
    else if (dtype == YH_C128)
      YH_TRY(launch_expect_terms<double>(static_cast<const double2*>(data), n, tp, nt, w.partial(), w.result(),
                                         w.stream));
    else
      YH_TRY(launch_expect_terms<float>(static_cast<const float2*>(data), n, tp, nt, w.partial(), w.result(),
                                        w.stream));
    double v[2 * kMaxTermsPerPass];
    YH_TRY(w.fetch(v, w.result(), sizeof(v)));
    for (int k = 0; k < nt; ++k) acc = acc + cf[t0 + k] * cd{v[2 * k], v[2 * k + 1]};
  }
  out[0] = acc.re;
  out[1] = acc.im;
  return YH_OK;
}

int yh_expect(const yh_reg* reg, size_t nterms, const double* coeffs, const int32_t* term_len,
              const int32_t* sites, const int32_t* ops, double out[2]) {
  YH_GUARD(reg, YH_ERR_INVALID_ARG, "null register");
  yh_reg* r = const_cast<yh_reg*>(reg);
  return expect_impl(r->w, r->nbits, r->state.p, r->dtype, false, nterms, coeffs, term_len, sites, ops, out);
}

extern "C++" {
// ------------------------- adjoint differentiation --------------------------
// Gate::num_params (gate.rs:107-113).
static int gate_num_params(const yh_op& op) {
  if (op.elem != YH_ELEM_GATE) return 0;
  switch (op.kind) {
    case YH_GATE_RX:
    case YH_GATE_RY:
    case YH_GATE_RZ:
    case YH_GATE_PHASE:
      return 1;
    case YH_GATE_FSIM:
      return 2;
    default:
      return 0;
  }
}

// Gate::generator_matrix (gate.rs:153-193): G = (dU/dtheta_i) U^-1, angle independent.
static Mat generator_matrix(const yh_op& op, int i) {
  const cd z{0, 0};
  switch (op.kind) {
    case YH_GATE_RX: return {z, cd{0, -0.5}, cd{0, -0.5}, z};
    case YH_GATE_RY: return {z, cd{-0.5, 0}, cd{0.5, 0}, z};
    case YH_GATE_RZ: return {cd{0, -0.5}, z, z, cd{0, 0.5}};
    case YH_GATE_PHASE: return {z, z, z, cd{0, 1}};
    default: {  // FSim
      Mat m(16, z);
      if (i == 0) {
        m[1 * 4 + 2] = cd{0, -1};
        m[2 * 4 + 1] = cd{0, -1};
      } else {
        m[3 * 4 + 3] = cd{0, -1};
      }
      return m;
    }
  }
}

// Gate::dagger (gate.rs:198-281) as an element; `store` keeps a Custom matrix alive.
static int dagger_op(const yh_op& op, yh_op& d, std::vector<double>& store) {
  d = op;
  auto custom_from = [&](const Mat& m, int dim, int is_diag) {
    store.assign(size_t(dim) * dim * 2, 0.0);
    for (int r = 0; r < dim; ++r)
      for (int c = 0; c < dim; ++c) {  // (m^T)* : out[r][c] = conj(m[c][r])
        store[2 * (size_t(r) * dim + c)] = m[size_t(c) * dim + r].re;
        store[2 * (size_t(r) * dim + c) + 1] = -m[size_t(c) * dim + r].im;
      }
    d.kind = YH_GATE_CUSTOM;
    d.matrix = store.data();
    d.dim = dim;
    d.is_diag = is_diag;
  };
  switch (op.kind) {
    case YH_GATE_H:
    case YH_GATE_X:
    case YH_GATE_Y:
    case YH_GATE_Z:
    case YH_GATE_SWAP:
      return YH_OK;
    case YH_GATE_S:
      d.kind = YH_GATE_PHASE;
      d.params[0] = -1.5707963267948966;  // -pi/2
      return YH_OK;
    case YH_GATE_T:
      d.kind = YH_GATE_PHASE;
      d.params[0] = -0.7853981633974483;  // -pi/4
      return YH_OK;
    case YH_GATE_RX:
    case YH_GATE_RY:
    case YH_GATE_RZ:
    case YH_GATE_PHASE:
      d.params[0] = -op.params[0];
      return YH_OK;
    case YH_GATE_FSIM:
      d.params[0] = -op.params[0];
      d.params[1] = -op.params[1];
      return YH_OK;
    case YH_GATE_CUSTOM: {
      Mat m;
      const int dim = gate_matrix(op, m);
      custom_from(m, dim, op.is_diag);
      return YH_OK;
    }
    default: {  // SqrtX, SqrtY, SqrtW, ISWAP: conjugate transpose as a Custom gate
      Mat m;
      const int dim = gate_matrix(op, m);
      custom_from(m, dim, 0);
      return YH_OK;
    }
  }
}

struct TermArgs {
  size_t nterms;
  const double* coeffs;
  const int32_t* term_len;
  const int32_t* sites;
  const int32_t* opcodes;
};

template <class R>
static int expect_grad_impl(yh_reg* reg, int nbits, const yh_op* ops, size_t nops, const TermArgs& ta,
                            const std::vector<TermDev>& td, const std::vector<cd>& cf, size_t nparams, double* value,
                            double* grad) {
  using V = typename Cplx<R>::V;
  Workspace& w = reg->w;
  const uint64_t n = uint64_t(1) << nbits;
  DevBuf psi, lam, gout, gens, coeffs;
  YH_TRY(psi.ensure(reg->bytes()));
  YH_TRY(lam.ensure(reg->bytes()));
  YH_HIP(hipMemcpyAsync(psi.p, reg->state.p, reg->bytes(), hipMemcpyDeviceToDevice, w.stream));
  YH_TRY(apply_gates<R>(w, psi.p, nbits, ops, nops));
  // value = Re expect_arrayreg(psi, H) (ad.rs:92)
  double ev[2];
  YH_TRY(expect_impl(w, nbits, psi.p, reg->dtype, false, ta.nterms, ta.coeffs, ta.term_len, ta.sites, ta.opcodes, ev));
  *value = ev[0];
  if (nparams == 0) return YH_OK;
  V* P = static_cast<V*>(psi.p);
  V* Lm = static_cast<V*>(lam.p);
  // lambda = sum_t c_t P_t psi (ad.rs:101-116)
  if (td.empty()) {
    YH_HIP(hipMemsetAsync(lam.p, 0, reg->bytes(), w.stream));
  } else {
    std::vector<double> c2(2 * cf.size());
    for (size_t k = 0; k < cf.size(); ++k) {
      c2[2 * k] = cf[k].re;
      c2[2 * k + 1] = cf[k].im;
    }
    YH_TRY(w.upload(td.data(), td.size() * sizeof(TermDev), w.terms));
    YH_TRY(w.upload(c2.data(), c2.size() * sizeof(double), coeffs));
    for (size_t t0 = 0; t0 < td.size(); t0 += kMaxTermsPerPass) {
      const int nt = int(std::min<size_t>(kMaxTermsPerPass, td.size() - t0));
      YH_TRY(launch_pauli_accum<R>(P, Lm, n, static_cast<const TermDev*>(w.terms.p) + t0,
                                   static_cast<const double*>(coeffs.p) + 2 * t0, nt, t0 == 0, w.stream));
    }
  }
  // all generator matrices, in parameter order, uploaded once
  std::vector<double2> gm;
  std::vector<size_t> goff(nparams);
  {
    size_t slot = 0;
    for (size_t e = 0; e < nops; ++e)
      for (int i = 0; i < gate_num_params(ops[e]); ++i) {
        const Mat g = generator_matrix(ops[e], i);
        goff[slot++] = gm.size();
        for (const cd& v : g) gm.push_back(double2{v.re, v.im});
      }
  }
  YH_TRY(w.upload(gm.data(), gm.size() * sizeof(double2), gens));
  YH_TRY(gout.ensure(2 * nparams * sizeof(double)));
  // reverse sweep (ad.rs:118-152)
  size_t slot = nparams;
  std::vector<double> store;
  for (size_t e = nops; e-- > 0;) {
    const yh_op& op = ops[e];
    if (op.elem != YH_ELEM_GATE) continue;
    const int np = gate_num_params(op);
    if (np > 0) {
      auto bit_of = [nbits](int loc) { return nbits - 1 - loc; };
      GroupArgs ga{};
      std::vector<int> pos;
      for (int k = 0; k < op.nt; ++k) pos.push_back(bit_of(op.targets[k]));
      for (int k = 0; k < op.nc; ++k) {
        pos.push_back(bit_of(op.ctrls[k]));
        if (op.ctrl_vals[k]) ga.ctrl_set |= uint64_t(1) << bit_of(op.ctrls[k]);
      }
      for (int j = 0; j < (1 << op.nt); ++j) {
        uint64_t o = 0;
        for (int k = 0; k < op.nt; ++k)
          if ((j >> k) & 1) o |= uint64_t(1) << bit_of(op.targets[k]);
        ga.offs[j] = o;
      }
      set_locked(ga, nbits, pos);
      for (int i = np - 1; i >= 0; --i) {
        --slot;
        YH_TRY(launch_gen_inner<R>(P, Lm, ga, op.nt, static_cast<const double2*>(gens.p) + goff[slot], w.partial(),
                                   static_cast<double*>(gout.p) + 2 * slot, w.stream));
      }
    }
    yh_op d;
    YH_TRY(dagger_op(op, d, store));
    YH_TRY(apply_gates<R>(w, psi.p, nbits, &d, 1));
    YH_TRY(apply_gates<R>(w, lam.p, nbits, &d, 1));
  }
  std::vector<double> g2(2 * nparams);
  if (nparams) YH_TRY(w.fetch(g2.data(), gout.p, g2.size() * sizeof(double)));
  YH_TRY(w.sync());
  for (size_t k = 0; k < nparams; ++k) grad[k] = 2.0 * g2[2 * k];
  return YH_OK;
}

}  // extern "C++"

int yh_circuit_num_params(const yh_op* ops, size_t nops, size_t* n) {
  YH_GUARD(n && (nops == 0 || ops), YH_ERR_INVALID_ARG, "null argument");
  size_t c = 0;
  for (size_t e = 0; e < nops; ++e) c += size_t(gate_num_params(ops[e]));
  *n = c;
  return YH_OK;
}

int yh_expect_grad(yh_reg* reg, int circuit_nbits, const yh_op* ops, size_t nops, size_t nterms,
                   const double* coeffs, const int32_t* term_len, const int32_t* sites, const int32_t* opcodes,
                   double* value, double* grad, size_t ngrad) {
  YH_GUARD(reg && value && (nops == 0 || ops), YH_ERR_INVALID_ARG, "null argument");
  for (size_t e = 0; e < nops; ++e)
    YH_GUARD(ops[e].elem != YH_ELEM_CHANNEL, YH_ERR_INVALID_ARG,
             "expect_grad: circuit contains a noise channel; AD is unsupported for channels");
  YH_GUARD(circuit_nbits == reg->nbits, YH_ERR_NBITS_MISMATCH, "Register and circuit qubit count mismatch");
  YH_TRY(yh_validate(circuit_nbits, ops, nops));
  size_t np = 0;
  YH_TRY(yh_circuit_num_params(ops, nops, &np));
  YH_GUARD(ngrad == np, YH_ERR_INVALID_ARG, "gradient length does not match the circuit's parameter count");
  YH_GUARD(np == 0 || grad, YH_ERR_INVALID_ARG, "null gradient buffer");
  size_t total = 0;
  for (size_t k = 0; k < nterms; ++k) {
    YH_GUARD(term_len && term_len[k] >= 0, YH_ERR_INVALID_ARG, "bad term length");
    total += size_t(term_len[k]);
  }
  YH_GUARD(nterms == 0 || coeffs, YH_ERR_INVALID_ARG, "null coefficients");
  YH_GUARD(total == 0 || (sites && opcodes), YH_ERR_INVALID_ARG, "null sites/ops");
  std::vector<Term> terms = read_terms(nterms, coeffs, term_len, sites, opcodes);
  std::vector<TermDev> td;
  std::vector<cd> cf;
  YH_TRY(lower_terms(reg->nbits, terms, false, td, cf));
  const TermArgs ta{nterms, coeffs, term_len, sites, opcodes};
  if (reg->dtype == YH_C128) return expect_grad_impl<double>(reg, reg->nbits, ops, nops, ta, td, cf, np, value, grad);
  return expect_grad_impl<float>(reg, reg->nbits, ops, nops, ta, td, cf, np, value, grad);
}

static int measure_locs(int nbits, const int32_t* locs, int nlocs, std::vector<int32_t>& locbits) {
  YH_GUARD(nlocs >= 0 && nlocs <= nbits && nlocs <= 30, YH_ERR_MEASURE_LOC, "too many measured locations");
  YH_GUARD(nlocs == 0 || locs, YH_ERR_INVALID_ARG, "null locs");
  locbits.clear();
  for (int i = 0; i < nlocs; ++i) {
    if (locs[i] < 0 || locs[i] >= nbits)
      return fail(YH_ERR_MEASURE_LOC, "measurement location " + std::to_string(locs[i]) + " is out of range for " +
                                           std::to_string(nbits) + " qubits");
    for (int j = 0; j < i; ++j)
      if (locs[j] == locs[i]) return fail(YH_ERR_MEASURE_LOC, "duplicate measurement location " + std::to_string(locs[i]));
    locbits.push_back(nbits - 1 - locs[i]);
  }
  return YH_OK;
}

static uint64_t choose_chunks(int nbits, int nlocs) {
  uint64_t members = uint64_t(1) << (nbits - nlocs);
  const uint64_t exact_limit = uint64_t(1) << 18, chunk = uint64_t(1) << 14;
  return members <= exact_limit ? 1 : members / chunk;
}

// probabilities into device buffer dst (2^nlocs or 2^nbits doubles)
static int probs_to_dev(yh_reg* reg, const int32_t* locs, int nlocs, double* dst) {
  Workspace& w = reg->w;
  const uint64_t n = uint64_t(1) << reg->nbits;
  if (!locs) {
    if (reg->dtype == YH_C128) return launch_probs_full<double>(static_cast<const double2*>(reg->state.p), n, dst, w.stream);
    return launch_probs_full<float>(static_cast<const float2*>(reg->state.p), n, dst, w.stream);
  }
  std::vector<int32_t> lb;
  YH_TRY(measure_locs(reg->nbits, locs, nlocs, lb));
  uint64_t nch = choose_chunks(reg->nbits, nlocs);
  YH_TRY(w.ws.ensure(sizeof(double) * (size_t(nch) << nlocs)));
  YH_TRY(w.upload(lb.data(), lb.size() * sizeof(int32_t), w.locs));
  return probs_marginal_dev(reg->state.p, reg->dtype, reg->nbits, static_cast<const int32_t*>(w.locs.p), lb.data(),
                            nlocs, dst, static_cast<double*>(w.ws.p), nch, w.stream);
}

int yh_probs(const yh_reg* creg, const int32_t* locs, int nlocs, double* out) {
  YH_GUARD(creg && out, YH_ERR_INVALID_ARG, "null argument");
  yh_reg* reg = const_cast<yh_reg*>(creg);
  uint64_t m = locs ? (uint64_t(1) << nlocs) : (uint64_t(1) << reg->nbits);
  if (locs) {
    std::vector<int32_t> lb;
    YH_TRY(measure_locs(reg->nbits, locs, nlocs, lb));
  }
  YH_TRY(reg->w.probs.ensure(m * sizeof(double)));
  YH_TRY(probs_to_dev(reg, locs, nlocs, static_cast<double*>(reg->w.probs.p)));
  return reg->w.fetch(out, reg->w.probs.p, m * sizeof(double));
}

// distribution -> sequential CDF -> per-shot search, all on the device
static int sample_dev(yh_reg* reg, const int32_t* locs, int nlocs, const double* uniforms, size_t nshots,
                      uint64_t* outcomes) {
  Workspace& w = reg->w;
  uint64_t m = locs ? (uint64_t(1) << nlocs) : (uint64_t(1) << reg->nbits);
  YH_TRY(w.probs.ensure(2 * m * sizeof(double)));
  double* p = static_cast<double*>(w.probs.p);
  YH_TRY(probs_to_dev(reg, locs, nlocs, p));
  const size_t csb = cdf_scratch_bytes(m);
  if (csb) YH_TRY(w.cdfws.ensure(csb));
  YH_TRY(launch_cdf(p, m, p + m, csb ? w.cdfws.p : nullptr, w.stream));
  YH_TRY(w.upload(uniforms, nshots * sizeof(double), w.shots));
  YH_TRY(w.shots_out.ensure(nshots * sizeof(uint64_t)));
  YH_TRY(launch_search(p + m, m, static_cast<const double*>(w.shots.p), nshots,
                       static_cast<uint64_t*>(w.shots_out.p), w.stream));
  return w.fetch(outcomes, w.shots_out.p, nshots * sizeof(uint64_t));
}

int yh_sample(const yh_reg* creg, const int32_t* locs, int nlocs, const double* uniforms, size_t nshots,
              uint64_t* outcomes) {
  YH_GUARD(creg && uniforms && outcomes && nshots > 0, YH_ERR_INVALID_ARG, "null argument");
  yh_reg* reg = const_cast<yh_reg*>(creg);
  if (locs) {
    std::vector<int32_t> lb;
    YH_TRY(measure_locs(reg->nbits, locs, nlocs, lb));
  }
  return sample_dev(reg, locs, nlocs, uniforms, nshots, outcomes);
}

int yh_measure(yh_reg* reg, const int32_t* locs, int nlocs, int post, uint64_t reset_val, double uniform,
               int32_t* outcome_bits) {
  YH_GUARD(reg && locs && outcome_bits, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(post >= 0 && post <= 2, YH_ERR_INVALID_ARG, "unknown post-processing");
  std::vector<int32_t> lb;
  YH_TRY(measure_locs(reg->nbits, locs, nlocs, lb));
  uint64_t idx = 0;
  YH_TRY(sample_dev(reg, locs, nlocs, &uniform, 1, &idx));
  // decode_outcome_bits (measure.rs:152-156)
  uint64_t mask = 0, val = 0;
  for (int k = 0; k < nlocs; ++k) {
    outcome_bits[k] = int32_t((idx >> (nlocs - 1 - k)) & 1);
    mask |= uint64_t(1) << lb[k];
    if (outcome_bits[k]) val |= uint64_t(1) << lb[k];
  }
  if (post == 0) return YH_OK;
  if (post == 1 && reset_val >= (uint64_t(1) << nlocs))
    return fail(YH_ERR_RESET_VALUE, "reset value " + std::to_string(reset_val) + " does not fit in " +
                                         std::to_string(nlocs) + " measured qubits");
  Workspace& w = reg->w;
  const uint64_t n = uint64_t(1) << reg->nbits;
  // collapse_qubits_to (measure.rs:158-185)
  if (reg->dtype == YH_C128) YH_TRY(launch_collapse_mask<double>(static_cast<double2*>(reg->state.p), n, mask, val, w.stream));
  else YH_TRY(launch_collapse_mask<float>(static_cast<float2*>(reg->state.p), n, mask, val, w.stream));
  YH_TRY(reg_norm2(reg, w.result()));
  YH_TRY(reg_scale(reg, w.result(), 1e-15));
  if (post == 1) {  // reset_qubits_to (measure.rs:187-214)
    uint64_t sm = 0;
    for (int k = 0; k < nlocs; ++k) {
      int to = int((reset_val >> (nlocs - 1 - k)) & 1);
      if (to != outcome_bits[k]) sm |= uint64_t(1) << lb[k];
    }
    if (sm) {
      std::vector<int32_t> sorted = lb;
      std::sort(sorted.begin(), sorted.end());
      LockedBits lk{};
      for (int k = 0; k < nlocs; ++k) lk.pos[k] = uint8_t(sorted[k]);
      lk.npos = nlocs;
      YH_TRY(launch_reset_dev(reg->state.p, reg->dtype, lk, n >> nlocs, val, sm, w.stream));
    }
    return w.sync();
  }
  // RemoveMeasured: gather the kept qubits, then normalize (measure.rs:216-240)
  YH_GUARD(nlocs < reg->nbits, YH_ERR_UNSUPPORTED, "RemoveMeasured of every qubit leaves an empty register");
  std::vector<int32_t> sorted = lb;
  std::sort(sorted.begin(), sorted.end());
  LockedBits lk{};
  for (int k = 0; k < nlocs; ++k) lk.pos[k] = uint8_t(sorted[k]);
  lk.npos = nlocs;
  const int nnew_bits = reg->nbits - nlocs;
  const uint64_t nnew = uint64_t(1) << nnew_bits;
  DevBuf fresh;
  YH_TRY(fresh.ensure(nnew * reg->amp_bytes()));
  if (reg->dtype == YH_C128)
    YH_TRY(launch_remove_gather<double>(static_cast<const double2*>(reg->state.p), static_cast<double2*>(fresh.p), nnew,
                                        lk, val, w.stream));
  else
    YH_TRY(launch_remove_gather<float>(static_cast<const float2*>(reg->state.p), static_cast<float2*>(fresh.p), nnew,
                                       lk, val, w.stream));
  YH_TRY(w.sync());
  std::swap(reg->state.p, fresh.p);
  std::swap(reg->state.bytes, fresh.bytes);
  reg->nbits = nnew_bits;
  YH_TRY(reg_norm2(reg, w.result()));
  YH_TRY(reg_scale(reg, w.result(), 0.0));
  return w.sync();
}

// ---------------------------------- DensityMatrix ----------------------------------
int yh_dm_create(int nbits, yh_dm** out) {
  YH_GUARD(out, YH_ERR_INVALID_ARG, "null output pointer");
  YH_GUARD(nbits >= 1 && nbits <= 17, YH_ERR_INVALID_ARG, "density matrix nbits must be in [1, 17]");
  YH_TRY(ensure_init());
  yh_dm* d = new yh_dm();
  d->nbits = nbits;
  int st = d->w.init();
  if (st == YH_OK) st = d->rho.ensure(d->bytes());
  if (st != YH_OK) {
    delete d;
    return st;
  }
  YH_HIP(hipMemsetAsync(d->rho.p, 0, d->bytes(), d->w.stream));
  const double one[2] = {1.0, 0.0};
  YH_HIP(hipMemcpyAsync(d->rho.p, one, sizeof(one), hipMemcpyHostToDevice, d->w.stream));
  st = d->w.sync();
  if (st != YH_OK) {
    delete d;
    return st;
  }
  *out = d;
  return YH_OK;
}

int yh_dm_from_reg(const yh_reg* reg, yh_dm** out) {
  YH_GUARD(reg && out, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(reg->dtype == YH_C128, YH_ERR_DTYPE, "density matrices are complex128");
  yh_dm* d = nullptr;
  YH_TRY(yh_dm_create(reg->nbits, &d));
  YH_HIP(hipStreamSynchronize(reg->w.stream));
  int st = launch_dm_outer(static_cast<const double2*>(reg->state.p), static_cast<double2*>(d->rho.p), d->dim(),
                           d->w.stream);
  if (st == YH_OK) st = d->w.sync();
  if (st != YH_OK) {
    delete d;
    return st;
  }
  *out = d;
  return YH_OK;
}

int yh_dm_mixed(const double* weights, const yh_reg* const* regs, size_t nregs, yh_dm** out) {
  YH_GUARD(out && weights && regs, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(nregs >= 1, YH_ERR_INVALID_ARG, "mixed: at least one register");
  const int n = regs[0] ? regs[0]->nbits : 0;
  std::vector<const double2*> ptrs(nregs);
  for (size_t j = 0; j < nregs; ++j) {
    YH_GUARD(regs[j], YH_ERR_INVALID_ARG, "null register");
    YH_GUARD(regs[j]->nbits == n, YH_ERR_NBITS_MISMATCH, "mixed: registers of different sizes");
    YH_GUARD(regs[j]->dtype == YH_C128, YH_ERR_DTYPE, "density matrices are complex128");
    YH_HIP(hipStreamSynchronize(regs[j]->w.stream));
    ptrs[j] = static_cast<const double2*>(regs[j]->state.p);
  }
  yh_dm* d = nullptr;
  YH_TRY(yh_dm_create(n, &d));
  std::unique_ptr<yh_dm> hold(d);
  DevBuf dp, dw;
  YH_TRY(d->w.upload(ptrs.data(), nregs * sizeof(void*), dp));
  YH_TRY(d->w.upload(weights, nregs * sizeof(double), dw));
  YH_TRY(launch_mixed(static_cast<double2*>(d->rho.p), d->dim(), static_cast<const double2* const*>(dp.p),
                      static_cast<const double*>(dw.p), int(nregs), d->w.stream));
  YH_TRY(d->w.sync());
  *out = hold.release();
  return YH_OK;
}

int yh_dm_from_reg_locs(const yh_reg* reg, const int32_t* locs, int nlocs, yh_dm** out) {
  YH_GUARD(reg && out && nlocs >= 0 && (nlocs == 0 || locs), YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(reg->dtype == YH_C128, YH_ERR_DTYPE, "density matrices are complex128");
  // kept locs are bit positions, as partial_tr reads them (the traced ones are
  // 0..nbits without locs, density_matrix.rs:266-272)
  PtrArgs a{};
  for (int l = 0; l < reg->nbits; ++l) {
    bool keep = false;
    for (int k = 0; k < nlocs; ++k) keep = keep || locs[k] == l;
    if (keep) a.kept[a.nkept++] = l;
    else a.tr[a.ntr++] = l;
  }
  YH_GUARD(a.nkept <= 17, YH_ERR_INVALID_ARG, "the reduced density matrix has more than 17 qubits");
  std::unique_ptr<yh_dm> d(new yh_dm());
  d->nbits = a.nkept;
  YH_TRY(d->w.init());
  YH_TRY(d->rho.ensure(d->bytes()));
  YH_HIP(hipStreamSynchronize(reg->w.stream));
  YH_TRY(launch_reduced<double>(static_cast<const double2*>(reg->state.p), a, static_cast<double2*>(d->rho.p),
                                d->w.stream));
  YH_TRY(d->w.sync());
  *out = d.release();
  return YH_OK;
}

int yh_dm_destroy(yh_dm* dm) {
  delete dm;
  return YH_OK;
}

int yh_dm_nbits(const yh_dm* dm, int* nbits) {
  YH_GUARD(dm && nbits, YH_ERR_INVALID_ARG, "null argument");
  *nbits = dm->nbits;
  return YH_OK;
}

int yh_dm_upload(yh_dm* dm, const double* rho, size_t n) {
  YH_GUARD(dm && rho, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(n == size_t(dm->dim()) * dm->dim(), YH_ERR_INVALID_ARG, "length must be 4^nbits");
  YH_HIP(hipMemcpy(dm->rho.p, rho, dm->bytes(), hipMemcpyHostToDevice));
  return YH_OK;
}

}  // extern "C"

// Eigenvalues (ascending) of a density matrix on the device, for matrices too large
// for the host path: rocSOLVER's divide-and-conquer Hermitian eigensolver (zheevd,
// eigenvalues only) on a copy of rho, which it overwrites. The row-major rho read
// as column-major is its transpose, conj(rho), which has the same eigenvalues.
static int dm_eigenvalues_device(yh_dm* dm, std::vector<double>& ev) {
  Workspace& w = dm->w;
  const int d = int(dm->dim());
  YH_TRY(w.acc.ensure(dm->bytes()));
  DevBuf dv, ev_off, info;
  YH_TRY(dv.ensure(sizeof(double) * size_t(d)));
  YH_TRY(ev_off.ensure(sizeof(double) * size_t(d)));
  YH_TRY(info.ensure(sizeof(rocblas_int)));
  YH_HIP(hipMemcpyAsync(w.acc.p, dm->rho.p, dm->bytes(), hipMemcpyDeviceToDevice, w.stream));
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return fail(YH_ERR_HIP, "rocblas_create_handle failed");
  std::unique_ptr<std::remove_pointer<rocblas_handle>::type, decltype(&rocblas_destroy_handle)> hold(
      h, &rocblas_destroy_handle);
  if (rocblas_set_stream(h, w.stream) != rocblas_status_success) return fail(YH_ERR_HIP, "rocblas_set_stream failed");
  const rocblas_status st =
      rocsolver_zheevd(h, rocblas_evect_none, rocblas_fill_lower, d, static_cast<rocblas_double_complex*>(w.acc.p), d,
                       static_cast<double*>(dv.p), static_cast<double*>(ev_off.p), static_cast<rocblas_int*>(info.p));
  if (st != rocblas_status_success)
    return fail(st == rocblas_status_memory_error ? YH_ERR_OUT_OF_MEMORY : YH_ERR_HIP,
                std::string("rocsolver_zheevd failed: ") + rocblas_status_to_string(st));
  rocblas_int inf = 0;
  ev.assign(size_t(d), 0.0);
  YH_TRY(w.fetch(ev.data(), dv.p, sizeof(double) * size_t(d)));
  YH_TRY(w.fetch(&inf, info.p, sizeof(inf)));
  if (inf != 0) return fail(YH_ERR_HIP, "eigendecomposition failed (zheevd did not converge)");
  return YH_OK;
}

extern "C" {

int yh_dm_partial_tr(const yh_dm* cdm, const int32_t* traced, int ntraced, yh_dm** out) {
  YH_GUARD(cdm && out && (ntraced == 0 || traced) && ntraced >= 0, YH_ERR_INVALID_ARG, "null argument");
  yh_dm* dm = const_cast<yh_dm*>(cdm);
  // traced locs are bit positions; locs at or beyond nbits trace nothing
  // (density_matrix.rs:80-84 filters 0..nbits)
  PtrArgs a{};
  for (int l = 0; l < dm->nbits; ++l) {
    bool t = false;
    for (int k = 0; k < ntraced; ++k) {
      YH_GUARD(traced[k] >= 0 && traced[k] < 64, YH_ERR_LOC_OUT_OF_RANGE, "traced location out of range");
      if (traced[k] == l) t = true;
    }
    if (t)
      a.tr[a.ntr++] = l;
    else
      a.kept[a.nkept++] = l;
  }
  std::unique_ptr<yh_dm> d(new yh_dm());
  d->nbits = a.nkept;
  YH_TRY(d->w.init());
  YH_TRY(d->rho.ensure(d->bytes()));
  YH_HIP(hipStreamSynchronize(dm->w.stream));
  YH_TRY(launch_partial_tr(static_cast<const double2*>(dm->rho.p), dm->dim(), a, static_cast<double2*>(d->rho.p),
                           d->w.stream));
  YH_TRY(d->w.sync());
  *out = d.release();
  return YH_OK;
}

int yh_dm_entropy(const yh_dm* dm, double* out) {
  YH_GUARD(dm && out, YH_ERR_INVALID_ARG, "null argument");
  const int d = int(dm->dim());
  std::vector<double> ev;
  if (dm->nbits <= 10) {
    // small matrices: downloaded, eigenvalues on the host
    std::vector<double> buf(size_t(2) * d * d);
    YH_TRY(yh_dm_download(dm, buf.data(), size_t(d) * d));
    Mat m(size_t(d) * d);
    for (size_t k = 0; k < m.size(); ++k) m[k] = cd{buf[2 * k], buf[2 * k + 1]};
    ev = hermitian_eigenvalues(m, d);
  } else {
    YH_TRY(dm_eigenvalues_device(const_cast<yh_dm*>(dm), ev));
  }
  // von_neumann_entropy (density_matrix.rs:113-122): eigenvalues clamped at 0
  // (hermitian_eigenvalues :227-242), those above 1e-15 contribute -v ln v
  double s = 0.0;
  for (double v : ev) {
    v = std::max(v, 0.0);
    if (v > 1e-15) s += v * std::log(v);
  }
  *out = -s;
  return YH_OK;
}

int yh_dm_download(const yh_dm* dm, double* rho, size_t n) {
  YH_GUARD(dm && rho, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(n == size_t(dm->dim()) * dm->dim(), YH_ERR_INVALID_ARG, "length must be 4^nbits");
  YH_HIP(hipStreamSynchronize(dm->w.stream));
  YH_HIP(hipMemcpy(rho, dm->rho.p, dm->bytes(), hipMemcpyDeviceToHost));
  return YH_OK;
}

}  // extern "C"

namespace yh {

struct DmLaunch {
  // 0: one k_dm_block pass; 1: the passes in `sv` on the 2n-bit vector (a gate on
  // more than 2 qubits: U on the row bits, conj(U) on the column bits); 2: a channel
  // as a sum of Kraus branches, branch k being the passes in branches[k]
  int mode = 0;
  int K = 1;
  DmArgs a{};
  std::vector<Launch<double>> sv;
  std::vector<std::vector<Launch<double>>> branches;
};

static void dm_block_geometry(int n, const std::vector<int>& qubits, DmArgs& a) {
  const int K = int(qubits.size());
  std::vector<int> pos;
  for (int q : qubits) {
    pos.push_back(2 * n - 1 - q);  // row bit of loc q
    pos.push_back(n - 1 - q);      // column bit
  }
  std::sort(pos.begin(), pos.end());
  a.lk.npos = int(pos.size());
  for (size_t k = 0; k < pos.size(); ++k) a.lk.pos[k] = uint8_t(pos[k]);
  a.ngroups = uint64_t(1) << (2 * n - 2 * K);
  for (int j = 0; j < (1 << K); ++j) {
    a.roff[j] = a.coff[j] = 0;
    for (int k = 0; k < K; ++k)
      if ((j >> k) & 1) {
        a.roff[j] |= uint64_t(1) << (2 * n - 1 - qubits[k]);
        a.coff[j] |= uint64_t(1) << (n - 1 - qubits[k]);
      }
  }
}

// Controlled matrix over block qubits [targets..., ctrls...] (block index bit k <-> qubit k).
static Mat controlled_block_matrix(const Mat& u, int nt, const yh_op& op) {
  const int K = nt + op.nc, B = 1 << K, T = 1 << nt;
  Mat m(size_t(B) * B, cd{0, 0});
  int want = 0;
  for (int k = 0; k < op.nc; ++k)
    if (op.ctrl_vals[k]) want |= 1 << k;
  for (int r = 0; r < B; ++r)
    for (int c = 0; c < B; ++c) {
      const int rc = r >> nt, cc = c >> nt;
      if (rc != cc) continue;
      if (rc == want) m[size_t(r) * B + c] = u[size_t(r & (T - 1)) * T + (c & (T - 1))];
      else if (r == c) m[size_t(r) * B + c] = cd{1, 0};
    }
  return m;
}

// A channel on more than 2 qubits, or with more Kraus operators than the block
// kernel folds in registers, as apply_channel (density_matrix.rs:181-224) does it:
// per Kraus operator K (in order) a branch copy of rho, K on the row bits and
// conj(K) on the column bits (dispatch_arrayreg_gate of a Custom gate, diagonal when
// K is), then the branch added into the accumulated sum.
static int dm_plan_kraus_branches(int n, const yh_op& op, const std::vector<Mat>& kr, int d, std::vector<DmLaunch>& L,
                                  Arena<double>& arena) {
  DmLaunch l;
  l.mode = 2;
  for (const Mat& K : kr) {
    const int dg = is_diagonal_matrix(K, d) ? 1 : 0;
    std::vector<double> mk(2 * K.size()), mc(2 * K.size());
    for (size_t k = 0; k < K.size(); ++k) {
      mk[2 * k] = mc[2 * k] = K[k].re;
      mk[2 * k + 1] = K[k].im;
      mc[2 * k + 1] = -K[k].im;
    }
    yh_op row{}, col{};
    row.elem = col.elem = YH_ELEM_GATE;
    row.kind = col.kind = YH_GATE_CUSTOM;
    row.nt = col.nt = op.nt;
    for (int k = 0; k < op.nt; ++k) {
      row.targets[k] = op.targets[k];
      col.targets[k] = op.targets[k] + n;
    }
    row.dim = col.dim = d;
    row.is_diag = col.is_diag = dg;
    row.matrix = mk.data();
    col.matrix = mc.data();
    std::vector<Launch<double>> br;
    YH_TRY(plan_gate<double>(2 * n, row, br, arena));
    YH_TRY(plan_gate<double>(2 * n, col, br, arena));
    l.branches.push_back(std::move(br));
  }
  L.push_back(std::move(l));
  return YH_OK;
}

static int dm_plan_channel(int n, const yh_op& op, std::vector<DmLaunch>& L, Arena<double>& arena) {
  std::vector<Mat> kr;
  int d = 0;
  int st = kraus_operators(op, kr, d);
  if (st != YH_OK) return fail(st, "invalid channel parameters");
  if (op.nt > 2 || kr.size() > 64) return dm_plan_kraus_branches(n, op, kr, d, L, arena);
  DmLaunch l;
  l.K = op.nt;
  std::vector<int> q(op.targets, op.targets + op.nt);
  dm_block_geometry(n, q, l.a);
  Mat all;
  for (auto& K : kr) all.insert(all.end(), K.begin(), K.end());
  l.a.nmicro = 1;
  l.a.micro[0] = DmMicro{1, -1, int(kr.size()), int(arena.push_mat(all) / sizeof(double2))};
  L.push_back(l);
  return YH_OK;
}

static int dm_plan_gate(int n, const yh_op& op, std::vector<DmLaunch>& L, Arena<double>& arena) {
  Mat u;
  int d = gate_matrix(op, u);
  if (d != (1 << op.nt)) return fail(YH_ERR_MATRIX_SIZE, "gate matrix size does not match targets");
  if (op.nt <= 2) {
    DmLaunch l;
    l.K = op.nt;
    std::vector<int> q(op.targets, op.targets + op.nt);
    dm_block_geometry(n, q, l.a);
    for (int k = 0; k < op.nc; ++k) {
      l.a.rc_mask |= uint64_t(1) << (2 * n - 1 - op.ctrls[k]);
      l.a.cc_mask |= uint64_t(1) << (n - 1 - op.ctrls[k]);
      if (op.ctrl_vals[k]) {
        l.a.rc_val |= uint64_t(1) << (2 * n - 1 - op.ctrls[k]);
        l.a.cc_val |= uint64_t(1) << (n - 1 - op.ctrls[k]);
      }
    }
    l.a.nmicro = 1;
    l.a.micro[0] = DmMicro{0, -1, 1, int(arena.push_mat(u) / sizeof(double2))};
    L.push_back(l);
    return YH_OK;
  }
  // nt > 2: U on the row bits, conj(U) on the column bits, as two dense passes
  DmLaunch l;
  l.mode = 1;
  yh_op row = op, col = op;
  row.kind = col.kind = YH_GATE_CUSTOM;
  std::vector<double> mu(2 * u.size()), mc(2 * u.size());
  for (size_t k = 0; k < u.size(); ++k) {
    mu[2 * k] = u[k].re;
    mu[2 * k + 1] = u[k].im;
    mc[2 * k] = u[k].re;
    mc[2 * k + 1] = -u[k].im;
  }
  row.matrix = mu.data();
  col.matrix = mc.data();
  row.dim = col.dim = d;
  row.is_diag = col.is_diag = 0;
  for (int k = 0; k < col.nt; ++k) col.targets[k] += n;
  for (int k = 0; k < col.nc; ++k) col.ctrls[k] += n;
  YH_TRY(plan_gate<double>(2 * n, row, l.sv, arena));
  YH_TRY(plan_gate<double>(2 * n, col, l.sv, arena));
  L.push_back(l);
  return YH_OK;
}

// Fusion: a gate on <= 2 qubits (targets plus controls) absorbs the 1-qubit
// channels that follow it on those qubits.
static int dm_plan_fused(int n, const yh_op* ops, size_t nops, std::vector<DmLaunch>& L, Arena<double>& arena) {
  size_t e = 0;
  while (e < nops) {
    const yh_op& op = ops[e];
    if (op.elem == YH_ELEM_ANNOTATION) {
      ++e;
      continue;
    }
    if (op.elem == YH_ELEM_CHANNEL) {
      YH_TRY(dm_plan_channel(n, op, L, arena));
      ++e;
      continue;
    }
    if (op.nt + op.nc > 2) {
      YH_TRY(dm_plan_gate(n, op, L, arena));
      ++e;
      continue;
    }
    Mat u;
    int d = gate_matrix(op, u);
    if (d != (1 << op.nt)) return fail(YH_ERR_MATRIX_SIZE, "gate matrix size does not match targets");
    std::vector<int> q(op.targets, op.targets + op.nt);
    for (int k = 0; k < op.nc; ++k) q.push_back(op.ctrls[k]);
    DmLaunch l;
    l.K = int(q.size());
    dm_block_geometry(n, q, l.a);
    Mat m = controlled_block_matrix(u, op.nt, op);
    l.a.nmicro = 1;
    l.a.micro[0] = DmMicro{0, -1, 1, int(arena.push_mat(m) / sizeof(double2))};
    ++e;
    while (e < nops && l.a.nmicro < kMaxMicro) {
      const yh_op& ch = ops[e];
      if (ch.elem == YH_ELEM_ANNOTATION) {
        ++e;
        continue;
      }
      if (ch.elem != YH_ELEM_CHANNEL || ch.nt != 1) break;
      int qi = -1;
      for (int k = 0; k < l.K; ++k)
        if (q[k] == ch.targets[0]) qi = k;
      if (qi < 0) break;
      std::vector<Mat> kr;
      int dk = 0;
      int st = kraus_operators(ch, kr, dk);
      if (st != YH_OK) return fail(st, "invalid channel parameters");
      Mat all;
      for (auto& K : kr) all.insert(all.end(), K.begin(), K.end());
      l.a.micro[l.a.nmicro++] = DmMicro{1, l.K == 1 ? -1 : qi, int(kr.size()),
                                        int(arena.push_mat(all) / sizeof(double2))};
      ++e;
    }
    L.push_back(l);
  }
  return YH_OK;
}

// per-gate passes on a 2n-bit vector `buf` (a copy for an out-of-place pass goes to w.scratch)
static int dm_run_sv(Workspace& w, double2* buf, int nbits2, size_t bytes, const std::vector<Launch<double>>& sv,
                     const unsigned char* base) {
  for (auto& s : sv) {
    switch (s.type) {
      case L_1Q: YH_TRY(launch_1q<double>(buf, nbits2, s.bit, s.m2, w.stream)); break;
      case L_DENSE:
        YH_TRY(launch_dense<double>(buf, s.ga, s.nt, reinterpret_cast<const double2*>(base + s.moff), w.stream));
        break;
      case L_DIAG: YH_TRY(launch_diag<double>(buf, s.ga, s.nt, s.dg, w.stream)); break;
      case L_X: YH_TRY(launch_xflip<double>(buf, s.ga, s.m0, w.stream)); break;
      case L_SWAP: YH_TRY(launch_swap<double>(buf, s.ga, s.m0, s.m1, w.stream)); break;
      case L_OOP: {
        YH_TRY(w.scratch.ensure(bytes));
        YH_HIP(hipMemcpyAsync(w.scratch.p, buf, bytes, hipMemcpyDeviceToDevice, w.stream));
        YH_TRY(launch_dense_oop<double>(buf, static_cast<const double2*>(w.scratch.p), s.ga, s.nt,
                                        reinterpret_cast<const double2*>(base + s.moff),
                                        reinterpret_cast<const uint64_t*>(base + s.ooff), w.stream));
        break;
      }
    }
  }
  return YH_OK;
}

static int dm_run(yh_dm* dm, std::vector<DmLaunch>& L, Arena<double>& arena) {
  Workspace& w = dm->w;
  YH_TRY(w.upload(arena.bytes.data(), arena.bytes.size(), w.mats));
  const unsigned char* base = static_cast<const unsigned char*>(w.mats.p);
  const size_t bytes = dm->bytes();
  for (auto& l : L) {
    double2* rho = static_cast<double2*>(dm->rho.p);
    if (l.mode == 0) {
      YH_TRY(w.time_begin(YH_K_DM));
      YH_TRY(launch_dm_block(rho, l.a, l.K, reinterpret_cast<const double2*>(base), w.stream));
      YH_TRY(w.time_end());
    } else if (l.mode == 1) {
      YH_TRY(dm_run_sv(w, rho, 2 * dm->nbits, bytes, l.sv, base));
    } else {
      // the sum of the branches builds up in w.acc (the first branch is computed
      // there directly), each later branch in w.branch; rho stays the original
      // until the sum is copied over it (callers such as the block path hold the
      // rho pointer across elements, so it never changes)
      YH_TRY(w.acc.ensure(bytes));
      for (size_t k = 0; k < l.branches.size(); ++k) {
        DevBuf& dst = k == 0 ? w.acc : w.branch;
        YH_TRY(dst.ensure(bytes));
        YH_HIP(hipMemcpyAsync(dst.p, rho, bytes, hipMemcpyDeviceToDevice, w.stream));
        YH_TRY(dm_run_sv(w, static_cast<double2*>(dst.p), 2 * dm->nbits, bytes, l.branches[k], base));
        if (k > 0)
          YH_TRY(launch_add_inplace(static_cast<double2*>(w.acc.p), static_cast<const double2*>(w.branch.p),
                                    bytes / sizeof(double2), w.stream));
      }
      YH_HIP(hipMemcpyAsync(rho, w.acc.p, bytes, hipMemcpyDeviceToDevice, w.stream));
    }
  }
  return w.sync();
}

}  // namespace yh

extern "C" {

static int timing_read(const Workspace& w, int kclass, uint64_t* launches, double* total_ms) {
  YH_GUARD(launches && total_ms, YH_ERR_INVALID_ARG, "null argument");
  YH_GUARD(kclass >= 0 && kclass < YH_K_NCLASS, YH_ERR_INVALID_ARG, "unknown kernel class");
  *launches = w.t_count[kclass];
  *total_ms = w.t_ms[kclass];
  return YH_OK;
}

int yh_reg_timing_enable(yh_reg* reg, int on) {
  YH_GUARD(reg, YH_ERR_INVALID_ARG, "null register");
  reg->w.timing = on != 0;
  for (int k = 0; k < YH_K_NCLASS; ++k) {
    reg->w.t_count[k] = 0;
    reg->w.t_ms[k] = 0;
  }
  return YH_OK;
}
int yh_reg_timing_read(const yh_reg* reg, int kclass, uint64_t* launches, double* total_ms) {
  YH_GUARD(reg, YH_ERR_INVALID_ARG, "null register");
  return timing_read(reg->w, kclass, launches, total_ms);
}
int yh_dm_timing_enable(yh_dm* dm, int on) {
  YH_GUARD(dm, YH_ERR_INVALID_ARG, "null density matrix");
  dm->w.timing = on != 0;
  for (int k = 0; k < YH_K_NCLASS; ++k) {
    dm->w.t_count[k] = 0;
    dm->w.t_ms[k] = 0;
  }
  return YH_OK;
}
int yh_dm_timing_read(const yh_dm* dm, int kclass, uint64_t* launches, double* total_ms) {
  YH_GUARD(dm, YH_ERR_INVALID_ARG, "null density matrix");
  return timing_read(dm->w, kclass, launches, total_ms);
}

int yh_dm_apply(yh_dm* dm, int circuit_nbits, const yh_op* ops, size_t nops) {
  YH_GUARD(dm, YH_ERR_INVALID_ARG, "null density matrix");
  YH_GUARD(circuit_nbits == dm->nbits, YH_ERR_NBITS_MISMATCH, "Register and circuit qubit count mismatch");
  YH_TRY(yh_validate(circuit_nbits, ops, nops));
  std::vector<DmLaunch> L;
  Arena<double> arena;
  for (size_t e = 0; e < nops; ++e) {
    if (ops[e].elem == YH_ELEM_GATE) YH_TRY(dm_plan_gate(dm->nbits, ops[e], L, arena));
    else if (ops[e].elem == YH_ELEM_CHANNEL) YH_TRY(dm_plan_channel(dm->nbits, ops[e], L, arena));
  }
  if (L.empty()) return YH_OK;
  return dm_run(dm, L, arena);
}

int yh_dm_apply_fused(yh_dm* dm, int circuit_nbits, const yh_op* ops, size_t nops) {
  YH_GUARD(dm, YH_ERR_INVALID_ARG, "null density matrix");
  YH_GUARD(circuit_nbits == dm->nbits, YH_ERR_NBITS_MISMATCH, "Register and circuit qubit count mismatch");
  YH_TRY(yh_validate(circuit_nbits, ops, nops));
  std::vector<DmLaunch> L;
  Arena<double> arena;
  YH_TRY(dm_plan_fused(dm->nbits, ops, nops, L, arena));
  if (L.empty()) return YH_OK;
  return dm_run(dm, L, arena);
}

int yh_dm_apply_blocks(yh_dm* dm, int circuit_nbits, const yh_op* ops, size_t nops) {
  YH_GUARD(dm, YH_ERR_INVALID_ARG, "null density matrix");
  YH_GUARD(circuit_nbits == dm->nbits, YH_ERR_NBITS_MISMATCH, "Register and circuit qubit count mismatch");
  YH_TRY(yh_validate(circuit_nbits, ops, nops));
  for (size_t e = 0; e < nops; ++e)  // the same parameter checks as the other paths, before any launch
    if (ops[e].elem == YH_ELEM_CHANNEL) {
      std::vector<Mat> kr;
      int d = 0;
      if (kraus_operators(ops[e], kr, d) != YH_OK) return fail(YH_ERR_CHANNEL_PARAM, "invalid channel parameters");
    }
  return apply_dm_blocks(dm->w, dm->rho.p, dm->nbits, ops, nops, [&](size_t e) {
    std::vector<DmLaunch> L;
    Arena<double> arena;
    if (ops[e].elem == YH_ELEM_GATE) YH_TRY(dm_plan_gate(dm->nbits, ops[e], L, arena));
    else YH_TRY(dm_plan_channel(dm->nbits, ops[e], L, arena));
    return L.empty() ? YH_OK : dm_run(dm, L, arena);
  });
}

int yh_dm_trace(const yh_dm* cdm, double out[2]) {
  YH_GUARD(cdm && out, YH_ERR_INVALID_ARG, "null argument");
  yh_dm* dm = const_cast<yh_dm*>(cdm);
  YH_TRY(launch_dm_trace(static_cast<const double2*>(dm->rho.p), dm->dim(), dm->w.partial(), dm->w.result(),
                         dm->w.stream));
  return dm->w.fetch(out, dm->w.result(), 2 * sizeof(double));
}

int yh_dm_purity(const yh_dm* cdm, double* out) {
  YH_GUARD(cdm && out, YH_ERR_INVALID_ARG, "null argument");
  yh_dm* dm = const_cast<yh_dm*>(cdm);
  YH_TRY(launch_dm_purity(static_cast<const double2*>(dm->rho.p), dm->dim(), dm->w.partial(), dm->w.result(),
                          dm->w.stream));
  double v[2];
  YH_TRY(dm->w.fetch(v, dm->w.result(), sizeof(v)));
  *out = v[0];
  return YH_OK;
}

int yh_dm_expect(const yh_dm* cdm, size_t nterms, const double* coeffs, const int32_t* term_len,
                 const int32_t* sites, const int32_t* ops, double out[2]) {
  YH_GUARD(cdm, YH_ERR_INVALID_ARG, "null density matrix");
  yh_dm* dm = const_cast<yh_dm*>(cdm);
  return expect_impl(dm->w, dm->nbits, dm->rho.p, YH_C128, true, nterms, coeffs, term_len, sites, ops, out);
}

int yh_dm_probs(const yh_dm* cdm, const int32_t* locs, int nlocs, double* out) {
  YH_GUARD(cdm && out, YH_ERR_INVALID_ARG, "null argument");
  yh_dm* dm = const_cast<yh_dm*>(cdm);
  Workspace& w = dm->w;
  uint64_t m = locs ? (uint64_t(1) << nlocs) : dm->dim();
  YH_TRY(w.probs.ensure(m * sizeof(double)));
  double* p = static_cast<double*>(w.probs.p);
  const double2* rho = static_cast<const double2*>(dm->rho.p);
  if (!locs) {
    YH_TRY(dm_probs_dev(rho, dm->nbits, nullptr, nullptr, 0, p, nullptr, 1, w.stream));
  } else {
    std::vector<int32_t> lb;
    YH_TRY(measure_locs(dm->nbits, locs, nlocs, lb));
    uint64_t nch = choose_chunks(dm->nbits, nlocs);
    YH_TRY(w.ws.ensure(sizeof(double) * (size_t(nch) << nlocs)));
    YH_TRY(w.upload(lb.data(), lb.size() * sizeof(int32_t), w.locs));
    YH_TRY(dm_probs_dev(rho, dm->nbits, static_cast<const int32_t*>(w.locs.p), lb.data(), nlocs, p,
                        static_cast<double*>(w.ws.p), nch, w.stream));
  }
  return w.fetch(out, p, m * sizeof(double));
}

}  // extern "C"
