// obs_kernels.hip — observables on gfx950: FP64 reductions, Pauli expectations,
// probabilities, exact-order CDF sampling and measurement post-processing.
//
// Reductions are two-stage and deterministic: a fixed grid writes one partial per
// block (wave64 shuffles, then LDS across waves) and one block folds the partials
// in a fixed order. All sums accumulate in FP64, also for complex64 states.
#include "kernels.h"

namespace yh {

namespace {

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* out) {
  __shared__ double sh[NV][kBlock / 64];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[k] = x;
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double acc = 0.0;
      for (int j = 0; j < kBlock / 64; ++j) acc += sh[k][j];
      out[k] = acc;
    }
  }
}

// Stage 2: fold nblocks partials of NV values each into out[NV].
template <int NV>
__global__ __launch_bounds__(kBlock) void k_fold(const double* __restrict__ partial, int nblocks,
                                                 double* __restrict__ out) {
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kBlock) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += partial[size_t(b) * NV + k];
  }
  double r[NV];
  block_sum<NV>(v, r);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = r[k];
  }
}

inline int red_blocks(uint64_t n) {
  uint64_t b = (n + uint64_t(kBlock) * 4 - 1) / (uint64_t(kBlock) * 4);
  return int(b < 1 ? 1 : (b > uint64_t(kRedBlocks) ? kRedBlocks : b));
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_norm2(const typename Cplx<R>::V* __restrict__ s, uint64_t n,
                                                  double* __restrict__ partial) {
  double v[1] = {0.0};
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    double re = s[i].x, im = s[i].y;
    v[0] += re * re + im * im;
  }
  double r[1];
  block_sum<1>(v, r);
  if (threadIdx.x == 0) partial[blockIdx.x] = r[0];
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_inner(const typename Cplx<R>::V* __restrict__ a,
                                                  const typename Cplx<R>::V* __restrict__ b, uint64_t n,
                                                  double* __restrict__ partial) {
  double v[2] = {0.0, 0.0};
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
    // conj(a) * b
    v[0] += ar * br + ai * bi;
    v[1] += ar * bi - ai * br;
  }
  double r[2];
  block_sum<2>(v, r);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = r[0];
    partial[2 * blockIdx.x + 1] = r[1];
  }
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_scale(typename Cplx<R>::V* __restrict__ s, uint64_t n,
                                                  const double* __restrict__ sumsq, double thresh) {
  const double nrm = sqrt(*sumsq);
  if (!(nrm > thresh)) return;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    typename Cplx<R>::V v = s[i];
    v.x = R(double(v.x) / nrm);
    v.y = R(double(v.y) / nrm);
    s[i] = v;
  }
}

// (P psi)_i = factor(i) * psi_j: each site replaces bit `bit` of i by src[v] and
// multiplies by fac[v], v being bit `bit` of i (rows of I/X/Y/Z/P0/P1/Pu/Pd have
// a single non-zero entry).
__device__ __forceinline__ void term_source(const TermDev& t, uint64_t i, uint64_t& j, double& fr,
                                            double& fi) {
  j = i;
  fr = 1.0;
  fi = 0.0;
  for (int k = 0; k < t.nsites; ++k) {
    const int b = t.bit[k];
    const int v = int((i >> b) & 1);
    j = (j & ~(uint64_t(1) << b)) | (uint64_t(t.src[k][v]) << b);
    const double gr = t.fac[k][v][0], gi = t.fac[k][v][1];
    const double nr = fr * gr - fi * gi, ni = fr * gi + fi * gr;
    fr = nr;
    fi = ni;
  }
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_expect(const typename Cplx<R>::V* __restrict__ s, uint64_t n,
                                                   const TermDev* __restrict__ terms, int nterms,
                                                   double* __restrict__ partial) {
  double v[2 * kMaxTermsPerPass];
#pragma unroll
  for (int k = 0; k < 2 * kMaxTermsPerPass; ++k) v[k] = 0.0;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    const double pr = s[i].x, pi = s[i].y;
#pragma unroll
    for (int t = 0; t < kMaxTermsPerPass; ++t) {
      if (t < nterms) {
        uint64_t j;
        double fr, fi;
        term_source(terms[t], i, j, fr, fi);
        const double qr = s[j].x, qi = s[j].y;
        const double yr = fr * qr - fi * qi, yi = fr * qi + fi * qr;  // (P psi)_i
        v[2 * t] += pr * yr + pi * yi;                                // conj(psi_i) * (P psi)_i
        v[2 * t + 1] += pr * yi - pi * yr;
      }
    }
  }
  double r[2 * kMaxTermsPerPass];
  block_sum<2 * kMaxTermsPerPass>(v, r);
  if (threadIdx.x == 0)
    for (int k = 0; k < 2 * kMaxTermsPerPass; ++k) partial[size_t(blockIdx.x) * 2 * kMaxTermsPerPass + k] = r[k];
}

// tr(P rho) with LSB-first sites: sum_r factor(r) * rho[c(r) * dim + r]
__global__ __launch_bounds__(kBlock) void k_dm_expect(const double2* __restrict__ rho, uint64_t dim,
                                                      const TermDev* __restrict__ terms, int nterms,
                                                      double* __restrict__ partial) {
  double v[2 * kMaxTermsPerPass];
#pragma unroll
  for (int k = 0; k < 2 * kMaxTermsPerPass; ++k) v[k] = 0.0;
  for (uint64_t r = uint64_t(blockIdx.x) * kBlock + threadIdx.x; r < dim; r += uint64_t(gridDim.x) * kBlock) {
#pragma unroll
    for (int t = 0; t < kMaxTermsPerPass; ++t) {
      if (t < nterms) {
        uint64_t c;
        double fr, fi;
        term_source(terms[t], r, c, fr, fi);
        const double2 q = rho[c * dim + r];
        v[2 * t] += fr * q.x - fi * q.y;
        v[2 * t + 1] += fr * q.y + fi * q.x;
      }
    }
  }
  double rr[2 * kMaxTermsPerPass];
  block_sum<2 * kMaxTermsPerPass>(v, rr);
  if (threadIdx.x == 0)
    for (int k = 0; k < 2 * kMaxTermsPerPass; ++k) partial[size_t(blockIdx.x) * 2 * kMaxTermsPerPass + k] = rr[k];
}

// Full probabilities: norm_sqr = re*re + im*im, rounded as the reference (no FMA).
template <class R>
__global__ __launch_bounds__(kBlock) void k_probs_full(const typename Cplx<R>::V* __restrict__ s, uint64_t n,
                                                       double* __restrict__ out) {
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    const double re = s[i].x, im = s[i].y;
    out[i] = norm_sqr_exact(re, im);
  }
}

__device__ __forceinline__ uint64_t bucket_base(uint64_t m, const int32_t* locbits, int nlocs) {
  uint64_t b = 0;
  for (int idx = 0; idx < nlocs; ++idx) b |= ((m >> (nlocs - 1 - idx)) & 1) << locbits[idx];
  return b;
}

// One thread per (bucket, chunk); members of a bucket are visited in increasing
// basis order, so with one chunk the sum is the reference's sequential fold.
template <class R, bool DM>
__global__ __launch_bounds__(kBlock) void k_probs_marginal(const typename Cplx<R>::V* __restrict__ s,
                                                           uint64_t dim, LockedBits lk, const int32_t* locbits,
                                                           int nlocs, uint64_t nbuckets, uint64_t nchunks,
                                                           uint64_t chunk_len, double* __restrict__ ws) {
  for (uint64_t t = uint64_t(blockIdx.x) * kBlock + threadIdx.x; t < nbuckets * nchunks;
       t += uint64_t(gridDim.x) * kBlock) {
    const uint64_t m = t / nchunks, c = t % nchunks;
    const uint64_t base = bucket_base(m, locbits, nlocs);
    double acc = 0.0;
    const uint64_t k0 = c * chunk_len, k1 = k0 + chunk_len;
    for (uint64_t k = k0; k < k1; ++k) {
      const uint64_t b = insert_zeros(k, lk.pos, lk.npos) | base;
      double p;
      if (DM) {
        p = fmax(double(s[b * dim + b].x), 0.0);
      } else {
        const double re = s[b].x, im = s[b].y;
        p = norm_sqr_exact(re, im);
      }
      acc = __dadd_rn(acc, p);
    }
    ws[t] = acc;
  }
}

__global__ __launch_bounds__(kBlock) void k_chunk_fold(const double* __restrict__ ws, uint64_t nbuckets,
                                                       uint64_t nchunks, double* __restrict__ out) {
  for (uint64_t m = uint64_t(blockIdx.x) * kBlock + threadIdx.x; m < nbuckets; m += uint64_t(gridDim.x) * kBlock) {
    double acc = 0.0;
    for (uint64_t c = 0; c < nchunks; ++c) acc = __dadd_rn(acc, ws[m * nchunks + c]);
    out[m] = acc;
  }
}

__global__ __launch_bounds__(kBlock) void k_dm_diag_probs(const double2* __restrict__ rho, uint64_t dim,
                                                          double* __restrict__ out) {
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < dim; i += uint64_t(gridDim.x) * kBlock)
    out[i] = fmax(rho[i * dim + i].x, 0.0);
}

// Sequential cumulative sum in the reference's order (measure.rs:140-150):
// cdf[i] = (((p[0] + p[1]) + p[2]) + ... + p[i]), one FP64 rounding per addition.
// The additions form one dependent chain, folded by one lane (thread 0) out of LDS;
// waves 1..3 stream the distribution through two LDS chunks around it (store the
// previous chunk's sums to cdf, load the next chunk's probabilities), so the chain
// never waits for HBM. Thread 0 reads its chunk eight values ahead of its writes.
constexpr int kCdfChunk = 4096;
constexpr int kCdfThreads = 256;
__global__ __launch_bounds__(kCdfThreads) void k_cdf(const double* __restrict__ p, uint64_t n,
                                                     double* __restrict__ cdf) {
  __shared__ double buf[2][kCdfChunk];
  const int tid = threadIdx.x;
  const int mover = tid - 64, nmov = kCdfThreads - 64;
  const uint64_t nch = (n + kCdfChunk - 1) / kCdfChunk;
  auto load = [&](uint64_t c) {
    const uint64_t base = c * kCdfChunk;
    for (int j = mover; j < kCdfChunk; j += nmov) buf[c & 1][j] = base + j < n ? p[base + j] : 0.0;
  };
  auto store = [&](uint64_t c) {
    const uint64_t base = c * kCdfChunk;
    for (int j = mover; j < kCdfChunk; j += nmov)
      if (base + j < n) cdf[base + j] = buf[c & 1][j];
  };
  if (mover >= 0) load(0);
  __syncthreads();
  double cs = 0.0;
  for (uint64_t c = 0; c < nch; ++c) {
    if (mover >= 0) {
      // the same thread stores, then refills, the same positions of the other chunk
      if (c >= 1) store(c - 1);
      if (c + 1 < nch) load(c + 1);
    } else if (tid == 0) {
      double* b = buf[c & 1];
      const int m = c + 1 < nch ? kCdfChunk : int(n - c * kCdfChunk);
      int j = 0;
      double x[8];
      if (m >= 8) {
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = b[k];
      }
      for (; j + 8 <= m; j += 8) {
        double y[8];
        const bool more = j + 16 <= m;
        if (more) {
#pragma unroll
          for (int k = 0; k < 8; ++k) y[k] = b[j + 8 + k];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          cs = __dadd_rn(cs, x[k]);
          b[j + k] = cs;
        }
        if (more) {
#pragma unroll
          for (int k = 0; k < 8; ++k) x[k] = y[k];
        }
      }
      for (; j < m; ++j) {
        cs = __dadd_rn(cs, b[j]);
        b[j] = cs;
      }
    }
    __syncthreads();
  }
  if (mover >= 0) store(nch - 1);
}

// First i with r < cdf[i] (cdf is non-decreasing), else n-1.
__global__ __launch_bounds__(kBlock) void k_search(const double* __restrict__ cdf, uint64_t n,
                                                   const double* __restrict__ r, uint64_t nshots,
                                                   uint64_t* __restrict__ out) {
  for (uint64_t k = uint64_t(blockIdx.x) * kBlock + threadIdx.x; k < nshots; k += uint64_t(gridDim.x) * kBlock) {
    const double x = r[k];
    uint64_t lo = 0, hi = n;  // invariant: answer in [lo, hi]; hi == n means "none"
    while (lo < hi) {
      uint64_t mid = lo + (hi - lo) / 2;
      if (x < cdf[mid]) hi = mid;
      else lo = mid + 1;
    }
    out[k] = lo < n ? lo : n - 1;
  }
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_collapse(typename Cplx<R>::V* __restrict__ s, uint64_t n,
                                                     uint64_t mask, uint64_t val) {
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock)
    if ((i & mask) != val) s[i] = cmk<typename Cplx<R>::V>(0, 0);
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_reset(typename Cplx<R>::V* __restrict__ s, LockedBits lk,
                                                  uint64_t ngroups, uint64_t from, uint64_t sm) {
  for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < ngroups; g += uint64_t(gridDim.x) * kBlock) {
    const uint64_t b = insert_zeros(g, lk.pos, lk.npos) | from;
    typename Cplx<R>::V a = s[b];
    s[b ^ sm] = a;
    s[b] = cmk<typename Cplx<R>::V>(0, 0);
  }
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_remove(const typename Cplx<R>::V* __restrict__ s,
                                                   typename Cplx<R>::V* __restrict__ dst, uint64_t nnew,
                                                   LockedBits lk, uint64_t fixed) {
  for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < nnew; g += uint64_t(gridDim.x) * kBlock) {
    typename Cplx<R>::V a = s[insert_zeros(g, lk.pos, lk.npos) | fixed];
    const double ns = double(a.x) * double(a.x) + double(a.y) * double(a.y);
    dst[g] = ns < 1e-30 ? cmk<typename Cplx<R>::V>(0, 0) : a;
  }
}

inline unsigned nblk(uint64_t n) { return capped_grid(n, kBlock); }

}  // namespace

template <class R>
int launch_norm2(const typename Cplx<R>::V* s, uint64_t n, double* partial, double* out, hipStream_t st) {
  int nb = red_blocks(n);
  hipLaunchKernelGGL(k_norm2<R>, dim3(nb), dim3(kBlock), 0, st, s, n, partial);
  hipLaunchKernelGGL(k_fold<1>, dim3(1), dim3(kBlock), 0, st, partial, nb, out);
  return YH_OK;
}

template <class R>
int launch_inner(const typename Cplx<R>::V* a, const typename Cplx<R>::V* b, uint64_t n, double* partial,
                 double* out, hipStream_t st) {
  int nb = red_blocks(n);
  hipLaunchKernelGGL(k_inner<R>, dim3(nb), dim3(kBlock), 0, st, a, b, n, partial);
  hipLaunchKernelGGL(k_fold<2>, dim3(1), dim3(kBlock), 0, st, partial, nb, out);
  return YH_OK;
}

template <class R>
int launch_scale(typename Cplx<R>::V* s, uint64_t n, const double* sumsq, double thresh, hipStream_t st) {
  int nb = red_blocks(n);
  hipLaunchKernelGGL(k_scale<R>, dim3(nb), dim3(kBlock), 0, st, s, n, sumsq, thresh);
  return YH_OK;
}

template <class R>
int launch_expect_terms(const typename Cplx<R>::V* s, uint64_t n, const TermDev* terms, int nterms,
                        double* partial, double* out, hipStream_t st) {
  int nb = red_blocks(n);
  hipLaunchKernelGGL(k_expect<R>, dim3(nb), dim3(kBlock), 0, st, s, n, terms, nterms, partial);
  hipLaunchKernelGGL(k_fold<2 * kMaxTermsPerPass>, dim3(1), dim3(kBlock), 0, st, partial, nb, out);
  return YH_OK;
}

int launch_dm_expect_terms(const double2* rho, uint64_t dim, const TermDev* terms, int nterms,
                           double* partial, double* out, hipStream_t st) {
  int nb = red_blocks(dim);
  hipLaunchKernelGGL(k_dm_expect, dim3(nb), dim3(kBlock), 0, st, rho, dim, terms, nterms, partial);
  hipLaunchKernelGGL(k_fold<2 * kMaxTermsPerPass>, dim3(1), dim3(kBlock), 0, st, partial, nb, out);
  return YH_OK;
}

template <class R>
int launch_probs_full(const typename Cplx<R>::V* s, uint64_t n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_probs_full<R>, dim3(nblk(n)), dim3(kBlock), 0, st, s, n, out);
  return YH_OK;
}

static LockedBits sorted_bits(const int32_t* locbits, int nlocs) {
  LockedBits lk{};
  int tmp[kMaxLocked];
  for (int k = 0; k < nlocs; ++k) tmp[k] = locbits[k];
  for (int a = 1; a < nlocs; ++a)
    for (int b = a; b > 0 && tmp[b - 1] > tmp[b]; --b) std::swap(tmp[b - 1], tmp[b]);
  for (int k = 0; k < nlocs; ++k) lk.pos[k] = uint8_t(tmp[k]);
  lk.npos = nlocs;
  return lk;
}

// locbits_dev: device copy of the loc bit positions (read by the kernel);
// locbits: host copy (to sort for the free-bit enumeration).
template <class R, bool DM>
static int probs_marginal_impl(const typename Cplx<R>::V* s, int nbits, const int32_t* locbits_dev,
                               const int32_t* locbits, int nlocs, double* out, double* ws,
                               uint64_t nchunks, hipStream_t st) {
  LockedBits lk = sorted_bits(locbits, nlocs);
  const uint64_t nb = uint64_t(1) << nlocs;
  const uint64_t members = uint64_t(1) << (nbits - nlocs);
  if (nchunks < 1 || members % nchunks) return fail(YH_ERR_INVALID_ARG, "probs: bad chunking");
  const uint64_t dim = uint64_t(1) << nbits;
  hipLaunchKernelGGL((k_probs_marginal<R, DM>), dim3(nblk(nb * nchunks)), dim3(kBlock), 0, st, s, dim, lk,
                     locbits_dev, nlocs, nb, nchunks, members / nchunks, ws);
  hipLaunchKernelGGL(k_chunk_fold, dim3(nblk(nb)), dim3(kBlock), 0, st, ws, nb, nchunks, out);
  return YH_OK;
}

int probs_marginal_dev(const void* s, int dtype, int nbits, const int32_t* locbits_dev,
                       const int32_t* locbits, int nlocs, double* out, double* ws, uint64_t nchunks,
                       hipStream_t st) {
  if (dtype == YH_C128)
    return probs_marginal_impl<double, false>(static_cast<const double2*>(s), nbits, locbits_dev, locbits,
                                              nlocs, out, ws, nchunks, st);
  return probs_marginal_impl<float, false>(static_cast<const float2*>(s), nbits, locbits_dev, locbits, nlocs,
                                           out, ws, nchunks, st);
}

int dm_probs_dev(const double2* rho, int nbits, const int32_t* locbits_dev, const int32_t* locbits,
                 int nlocs, double* out, double* ws, uint64_t nchunks, hipStream_t st) {
  if (nlocs == 0 && locbits == nullptr) {
    uint64_t dim = uint64_t(1) << nbits;
    hipLaunchKernelGGL(k_dm_diag_probs, dim3(nblk(dim)), dim3(kBlock), 0, st, rho, dim, out);
    return YH_OK;
  }
  return probs_marginal_impl<double, true>(rho, nbits, locbits_dev, locbits, nlocs, out, ws, nchunks, st);
}

int launch_cdf(const double* p, uint64_t n, double* cdf, hipStream_t st) {
  hipLaunchKernelGGL(k_cdf, dim3(1), dim3(kCdfThreads), 0, st, p, n, cdf);
  return YH_OK;
}

int launch_search(const double* cdf, uint64_t n, const double* r, uint64_t nshots, uint64_t* out,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_search, dim3(nblk(nshots)), dim3(kBlock), 0, st, cdf, n, r, nshots, out);
  return YH_OK;
}

template <class R>
int launch_collapse_mask(typename Cplx<R>::V* s, uint64_t n, uint64_t mask, uint64_t val, hipStream_t st) {
  hipLaunchKernelGGL(k_collapse<R>, dim3(nblk(n)), dim3(kBlock), 0, st, s, n, mask, val);
  return YH_OK;
}

int launch_reset_dev(void* s, int dtype, const LockedBits& lk, uint64_t ngroups, uint64_t from, uint64_t sm,
                     hipStream_t st) {
  if (dtype == YH_C128)
    hipLaunchKernelGGL(k_reset<double>, dim3(nblk(ngroups)), dim3(kBlock), 0, st, static_cast<double2*>(s),
                       lk, ngroups, from, sm);
  else
    hipLaunchKernelGGL(k_reset<float>, dim3(nblk(ngroups)), dim3(kBlock), 0, st, static_cast<float2*>(s), lk,
                       ngroups, from, sm);
  return YH_OK;
}

template <class R>
int launch_remove_gather(const typename Cplx<R>::V* s, typename Cplx<R>::V* dst, uint64_t nnew,
                         const LockedBits& lk, uint64_t fixed, hipStream_t st) {
  hipLaunchKernelGGL(k_remove<R>, dim3(nblk(nnew)), dim3(kBlock), 0, st, s, dst, nnew, lk, fixed);
  return YH_OK;
}

#define YH_INST(R)                                                                                      \
  template int launch_norm2<R>(const Cplx<R>::V*, uint64_t, double*, double*, hipStream_t);             \
  template int launch_inner<R>(const Cplx<R>::V*, const Cplx<R>::V*, uint64_t, double*, double*,        \
                               hipStream_t);                                                             \
  template int launch_scale<R>(Cplx<R>::V*, uint64_t, const double*, double, hipStream_t);              \
  template int launch_expect_terms<R>(const Cplx<R>::V*, uint64_t, const TermDev*, int, double*,        \
                                      double*, hipStream_t);                                             \
  template int launch_probs_full<R>(const Cplx<R>::V*, uint64_t, double*, hipStream_t);                 \
  template int launch_collapse_mask<R>(Cplx<R>::V*, uint64_t, uint64_t, uint64_t, hipStream_t);         \
  template int launch_remove_gather<R>(const Cplx<R>::V*, Cplx<R>::V*, uint64_t, const LockedBits&,     \
                                       uint64_t, hipStream_t);
YH_INST(double)
YH_INST(float)
#undef YH_INST

}  // namespace yh
