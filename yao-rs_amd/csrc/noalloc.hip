// noalloc.hip — yh_expect_noalloc: expect_arrayreg (src/expect.rs:7-11) on a
// caller-owned device state, for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// yh_expect keeps its reduction partials, results and the device form of the
// Pauli terms in buffers its register handle allocates. This entry point takes the
// same memory from one file-scope device array, `scratch`, of kScratchBytes: every
// call on a GPU shares that GPU's copy, which keeps the last call's contents, so
// calls that overlap on one GPU race. Nothing in it has to start at zero (each
// reduction block writes its partial, each fold writes its results). The array's
// device address is taken once per device with hipGetSymbolAddress (the array's
// name in host code is only a host-side stand-in) and passed to the kernels.
#include <cstring>
#include <mutex>
#include <vector>

#include "runtime.h"

namespace yh {

namespace {

constexpr int kNoAllocMaxTerms = 256;
constexpr int kNoAllocMaxPasses = (kNoAllocMaxTerms + kMaxTermsPerPass - 1) / kMaxTermsPerPass;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }
// layout: reduction partials of one pass | results of every pass | term table
constexpr size_t kPartialBytes = sizeof(double) * size_t(kRedBlocks) * 2 * kMaxTermsPerPass;
constexpr size_t kResultOff = align_up(kPartialBytes);
constexpr size_t kResultBytes = sizeof(double) * 2 * kMaxTermsPerPass * kNoAllocMaxPasses;
constexpr size_t kTermOff = align_up(kResultOff + kResultBytes);
constexpr size_t kScratchBytes = align_up(kTermOff + sizeof(TermDev) * kNoAllocMaxTerms);

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, resolved on the first call
// made with that device current.
int scratch_ptr(unsigned char** out) {
  int dev = 0;
  YH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(YH_ERR_NO_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    YH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(YH_ERR_HIP, "hipGetSymbolAddress returned a null address");
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return YH_OK;
}

}  // namespace

}  // namespace yh

using namespace yh;

extern "C" int yh_expect_noalloc(const void* amps, int nbits, int dtype, size_t nterms, const double* coeffs,
                                 const int32_t* term_len, const int32_t* sites, const int32_t* ops, void* stream,
                                 double out[2]) {
  if (!amps || !out || (nterms > 0 && (!coeffs || !term_len)))
    return fail(YH_ERR_INVALID_ARG, "null argument");
  if (nbits < 1 || nbits > 35) return fail(YH_ERR_INVALID_ARG, "nbits must be in [1, 35]");
  if (dtype != YH_C128 && dtype != YH_C64) return fail(YH_ERR_DTYPE, "unknown dtype");
  if (nterms > size_t(kNoAllocMaxTerms))
    return fail(YH_ERR_INVALID_ARG, "yh_expect_noalloc takes at most 256 terms (its static scratch holds 256)");
  size_t total = 0;
  for (size_t k = 0; k < nterms; ++k) {
    if (term_len[k] < 0) return fail(YH_ERR_INVALID_ARG, "negative term length");
    total += size_t(term_len[k]);
  }
  if (total > 0 && (!sites || !ops)) return fail(YH_ERR_INVALID_ARG, "null sites/ops");
  YH_TRY(ensure_init());
  unsigned char* base = nullptr;
  YH_TRY(scratch_ptr(&base));
  std::vector<Term> terms = read_terms(nterms, coeffs, term_len, sites, ops);
  std::vector<TermDev> td;
  std::vector<cd> cf;
  YH_TRY(lower_terms(nbits, terms, false, td, cf));
  const hipStream_t st = static_cast<hipStream_t>(stream);
  double* partial = reinterpret_cast<double*>(base);
  double* results = reinterpret_cast<double*>(base + kResultOff);
  TermDev* tdev = reinterpret_cast<TermDev*>(base + kTermOff);
  if (!td.empty()) YH_HIP(hipMemcpyAsync(tdev, td.data(), td.size() * sizeof(TermDev), hipMemcpyHostToDevice, st));
  const uint64_t n = uint64_t(1) << nbits;
  const size_t npass = (td.size() + kMaxTermsPerPass - 1) / kMaxTermsPerPass;
  for (size_t p = 0; p < npass; ++p) {
    const size_t t0 = p * kMaxTermsPerPass;
    const int nt = int(std::min<size_t>(kMaxTermsPerPass, td.size() - t0));
    double* res = results + p * 2 * kMaxTermsPerPass;
    if (dtype == YH_C128)
      YH_TRY(launch_expect_terms<double>(static_cast<const double2*>(amps), n, tdev + t0, nt, partial, res, st));
    else
      YH_TRY(launch_expect_terms<float>(static_cast<const float2*>(amps), n, tdev + t0, nt, partial, res, st));
    YH_HIP(hipGetLastError());
  }
  std::vector<double> v(npass * 2 * kMaxTermsPerPass);
  if (npass) YH_HIP(hipMemcpyAsync(v.data(), results, v.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  YH_HIP(hipStreamSynchronize(st));
  // the same order of accumulation as yh_expect
  cd acc{0, 0};
  for (size_t t = 0; t < td.size(); ++t) acc = acc + cf[t] * cd{v[2 * t], v[2 * t + 1]};
  out[0] = acc.re;
  out[1] = acc.im;
  return YH_OK;
}
