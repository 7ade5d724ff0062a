// runtime.h — device buffers and the per-handle workspace (one HIP stream each).
#pragma once

#include <functional>
#include <vector>

#include "common.h"
#include "host.h"
#include "kernels.h"

namespace yh {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  int ensure(size_t b);  // grows (contents are not preserved)
  void release();
};

struct PinBuf {  // page-locked host staging (hipHostMalloc)
  void* p = nullptr;
  size_t bytes = 0;
  PinBuf() = default;
  PinBuf(const PinBuf&) = delete;
  PinBuf& operator=(const PinBuf&) = delete;
  ~PinBuf();
  int ensure(size_t b);
};

struct Workspace {
  hipStream_t stream = nullptr;
  DevBuf red;        // reduction partials followed by results
  DevBuf mats;       // per-call operand arena (matrices, offset tables)
  DevBuf scratch;    // out-of-place copies
  DevBuf terms;      // Pauli terms in device form
  DevBuf locs;       // measured loc bits
  DevBuf ws;         // chunked probability sums
  DevBuf probs;      // probabilities and their CDF
  DevBuf shots;      // caller-supplied uniforms
  DevBuf shots_out;  // sampled outcomes
  DevBuf cdfws;      // cumulative-sum records of large distributions
  DevBuf tiles;      // tile-pass groups and micro gates
  DevBuf acc;        // density matrix: sum of Kraus branches (copied into rho)
  DevBuf branch;     // density matrix: one Kraus branch
  PinBuf pin;
  PinBuf bpin;       // block matrices staged for async copies (yh_apply_fused)
  // launch timing (measurement hooks): events around each launch while enabled
  struct TimedLaunch {
    int kclass;
    hipEvent_t a, b;
  };
  bool timing = false;
  std::vector<TimedLaunch> pending;
  uint64_t t_count[YH_K_NCLASS] = {};
  double t_ms[YH_K_NCLASS] = {};
  int time_begin(int kclass);  // call before a launch
  int time_end();              // call after it
  int time_collect();          // after the stream has drained
  Workspace() = default;
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
  ~Workspace();
  int init();
  double* partial();
  double* result();
  int sync();
  int upload(const void* host, size_t bytes, DevBuf& dst);
  int fetch(void* host, const void* dev, size_t bytes);
};

int ensure_init();

// Terms of an operator polynomial in device form. lsb_sites: sites are bit
// positions (expect_dm); otherwise loc l is bit nbits-1-l.
int lower_terms(int nbits, const std::vector<Term>& terms, bool lsb_sites, std::vector<TermDev>& out,
                std::vector<cd>& coeffs);

template <class R>
int apply_gates(Workspace& w, void* state, int nbits, const yh_op* ops, size_t nops);
template <class R>
int apply_tiled(Workspace& w, void* state, int nbits, const yh_op* ops, size_t nops);
int plan_stats(int nbits, const yh_op* ops, size_t nops, int64_t out[4]);
template <class R>
int apply_blocks(Workspace& w, void* state, int nbits, const yh_op* ops, size_t nops);
int block_plan_stats(int nbits, const yh_op* ops, size_t nops, int64_t out[3]);
int block_plan_bits(int nbits, const yh_op* ops, size_t nops, int32_t* bits, size_t cap, size_t* nsteps);
int block_pass_stats(int nbits, const yh_op* ops, size_t nops, int dm, int64_t out[7]);
// Density matrix (2n-bit vector, row bits above column bits) through the block
// path: gates and channels as superoperators; an element on more than 5 bits goes
// to single_op(op index).
int apply_dm_blocks(Workspace& w, void* rho, int n, const yh_op* ops, size_t nops,
                    const std::function<int(size_t)>& single_op);

}  // namespace yh

// Register and density-matrix handles (opaque in the C ABI).
struct yh_reg {
  int nbits = 0;
  int dtype = YH_C128;
  yh::DevBuf state;
  yh::Workspace w;
  size_t amp_bytes() const { return dtype == YH_C128 ? 16 : 8; }
  size_t bytes() const { return (size_t(1) << nbits) * amp_bytes(); }
};

struct yh_dm {
  int nbits = 0;
  yh::DevBuf rho;
  yh::Workspace w;
  uint64_t dim() const { return uint64_t(1) << nbits; }
  size_t bytes() const { return size_t(dim()) * dim() * 16; }
};
