// common.h — shared definitions of the gfx950 backend (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/yaohip.h"

namespace yh {

// ---- error plumbing: every entry point returns a status and sets the message ----
void set_error(const std::string& msg);
int fail(int status, const std::string& msg);

#define YH_HIP(expr)                                                                   \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return ::yh::fail(YH_ERR_HIP, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
  } while (0)

#define YH_TRY(expr)          \
  do {                        \
    int _s = (expr);          \
    if (_s != YH_OK) return _s; \
  } while (0)

// ---- complex helpers on device (V = double2 for complex128, float2 for complex64) ----
template <class R> struct Cplx;
template <> struct Cplx<double> { using V = double2; };
template <> struct Cplx<float> { using V = float2; };

template <class V>
__host__ __device__ __forceinline__ V cmk(decltype(V::x) re, decltype(V::x) im) {
  V v;
  v.x = re;
  v.y = im;
  return v;
}
template <class V>
__device__ __forceinline__ V cmul(V a, V b) {
  return cmk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class V>
__device__ __forceinline__ V cadd(V a, V b) {
  return cmk<V>(a.x + b.x, a.y + b.y);
}
// a * w + b * v   (the u1rows update, src/instruct_qubit.rs:23-36)
template <class V>
__device__ __forceinline__ V cmul2(V a, V w, V b, V v) {
  return cadd(cmul(a, w), cmul(b, v));
}

// |z|^2 = re*re + im*im rounded after each operation, as num-complex's norm_sqr in
// Rust (which never fuses); contraction into an FMA is switched off for this scope.
__device__ __forceinline__ double norm_sqr_exact(double re, double im) {
#pragma clang fp contract(off)
  const double a = re * re;
  const double b = im * im;
  return a + b;
}

// Insert a zero bit at each of the npos sorted (ascending) positions of `pos`.
// Maps a free index k to the basis index with those bits locked to 0
// (bitbasis itercontrol / group_shift, iterate_control.rs:82-142).
__device__ __forceinline__ uint64_t insert_zeros(uint64_t k, const uint8_t* pos, int npos) {
  for (int j = 0; j < npos; ++j) {
    uint32_t p = pos[j];
    uint64_t lo = k & ((uint64_t(1) << p) - 1);
    k = ((k >> p) << (p + 1)) | lo;
  }
  return k;
}

// Workgroup barrier that orders LDS accesses only. __syncthreads() also waits for
// every outstanding global load and store of the calling wave, which would make a
// register prefetch issued before the barrier complete at the barrier; here global
// accesses are waited for only where their registers are used.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Device limits used by the launchers.
constexpr int kMaxLocked = 32;
constexpr int kBlock = 256;

// Kernels that give each thread one item (or a fixed number of items) walk their
// items with a grid stride, over at most kMaxGridBlocks workgroups: HIP refuses a
// launch of 2^32 or more work-items in one dimension, which a 33-qubit state with one
// thread per amplitude would need. Grids of states up to 30 qubits are unchanged.
constexpr uint64_t kMaxGridBlocks = uint64_t(1) << 20;
inline unsigned capped_grid(uint64_t work, uint64_t per_block) {
  uint64_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  return unsigned(b > kMaxGridBlocks ? kMaxGridBlocks : b);
}

struct LockedBits {
  uint8_t pos[kMaxLocked];  // ascending bit positions locked to zero
  int32_t npos;
};

}  // namespace yh
