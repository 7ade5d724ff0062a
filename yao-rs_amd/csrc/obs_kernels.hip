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

// Cumulative sum of sample_from_probs (measure.rs:140-150) with the bits of its
// sequential left fold, cdf[i] = fl(cdf[i-1] + p[i]) from cdf[-1] = 0, computed in
// parallel.
//
// While the running sum c stays inside one binade [2^e, 2^(e+1)), every addition
// rounds to the grid u = 2^(e-52): fl(c + p) = c + u * rint(p / u), unless p / u is
// an exact half-integer (a tie, whose rounding depends on c's last bit). So a run of
// additions inside one binade is c0 + u * (integer prefix sum of R = rint(p / u)),
// and each of its partial sums is exact in FP64. The distribution is cut into
// chunks of 4096 buckets, each taken by one workgroup of 512 lanes:
//  1. an FP64 prefix scan (any order) gives approximate sums s; a bucket whose s,
//     widened by a bound on the fold's rounding error, sits inside one binade is a
//     run bucket of that binade; ties, buckets near a change of binade and the first
//     non-zero bucket are single buckets, folded one at a time; a zero bucket leaves
//     c unchanged and joins the segment before it;
//  2. segments (runs and single buckets) are numbered by a scan of their starts, and
//     R is prefix-summed in uint64 (a run's share is exact modulo 2^64);
//  3. one wave walks the segments in order, 64 at a time held in its lanes: it runs
//     the chain of additions (p for a single bucket, u * T for a run), then checks in
//     parallel that each run's starting c lies in the run's binade and c + u * T stays
//     within it, which makes every run's addition exact; if one fails, the group is
//     walked again segment by segment and the failing run folded bucket by bucket, so
//     the result never depends on the error bound being right;
//  4. every lane writes its buckets: run start + u * prefix, or the folded value.
// Up to kCdfSmallMax buckets (one chunk) one workgroup does all four steps
// (k_cdf_small). Above that, steps 1-2 and 4 run over all chunks at once, around a
// one-workgroup walk over every chunk's segment records (k_cdf_tot, k_cdf_base,
// k_cdf_plan, k_cdf_roff, k_cdf_pack, k_cdf_walk, k_cdf_fill); a chunk with more than
// kCdfMaxSeg segments is folded bucket by bucket in the walk.
constexpr int kCdfT = 512;             // lanes
constexpr int kCdfE = 8;               // consecutive buckets per lane
constexpr int kCdfCh = kCdfT * kCdfE;  // buckets per chunk
constexpr int kCdfW = kCdfT / 64;
constexpr int kCdfMaxSeg = 64;         // segment records per chunk (one per lane of the walking wave)
constexpr uint64_t kCdfSmallMax = kCdfCh;
// bucket kinds
constexpr int kBZero = 0;  // every bucket so far is zero: cdf = 0
constexpr int kBOne = 1;   // folded on its own
constexpr int kBRun = 2;   // member of a binade run
constexpr int kBJoin = 3;  // zero bucket (or past the end) joining the segment before it
// segment kinds, as stored in sinfo bits 12..13
constexpr int kSOne = 1, kSRun = 2, kSFolded = 3;

__device__ __forceinline__ int seg_first(int inf) { return inf & (kCdfCh - 1); }
__device__ __forceinline__ int seg_kind(int inf) { return (inf >> 12) & 3; }
__device__ __forceinline__ int seg_exp(int inf) { return (inf >> 14) - 1024; }

// DPP lane moves (CTRL: row_shr:n = 0x110 + n, row_bcast:15 = 0x142, row_bcast:31 =
// 0x143, wave_shr:1 = 0x138); lanes without a source read 0
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_mov(uint64_t v) {
  const uint32_t lo = uint32_t(dpp_mov<CTRL>(int(uint32_t(v))));
  const uint32_t hi = uint32_t(dpp_mov<CTRL>(int(uint32_t(v >> 32))));
  return (uint64_t(hi) << 32) | lo;
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  return __longlong_as_double((long long)dpp_mov<CTRL>(uint64_t(__double_as_longlong(v))));
}

__device__ __forceinline__ double readlane_f64(double v, int j) {
  const uint64_t b = uint64_t(__double_as_longlong(v));
  const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b)), j));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b >> 32)), j));
  return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}

// inclusive prefix over the 64 lanes of a wave (every lane active)
template <class T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const int lane = threadIdx.x & 63, rl = lane & 15;
  T t = dpp_mov<0x111>(x);
  if (rl >= 1) x += t;
  t = dpp_mov<0x112>(x);
  if (rl >= 2) x += t;
  t = dpp_mov<0x114>(x);
  if (rl >= 4) x += t;
  t = dpp_mov<0x118>(x);
  if (rl >= 8) x += t;
  t = dpp_mov<0x142>(x);  // lane 15 of each row into the next row
  if (lane & 16) x += t;
  t = dpp_mov<0x143>(x);  // lane 31 into rows 2 and 3
  if (lane >= 32) x += t;
  return x;
}

// exclusive prefix of x over the workgroup's kCdfT lanes, and the workgroup total (the
// same value, summed in the same order, in every lane); sh holds kCdfW values
template <class T>
__device__ __forceinline__ T block_excl_scan(T x, T* sh, T& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const T inc = wave_incl_scan(x);
  T ex = dpp_mov<0x138>(inc);
  if (lane == 0) ex = T(0);
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  T base = T(0), tot = T(0);
#pragma unroll
  for (int j = 0; j < kCdfW; ++j) {
    const T v = sh[j];
    if (j < w) base += v;
    tot += v;
  }
  total = tot;
  return base + ex;
}

// Kind of bucket k (value p) given the approximate sum s_prev of the buckets before it
// (kd = k as a double). The fold's own sum differs from the exact prefix by at most
// (k + 1) 2^-53 of itself, and the any-order scan by far less; the margin is
// (k + 257) 2^-51 of s_prev. For a run bucket, r = rint(p / u) with u = 2^(e-52).
__device__ __forceinline__ int cdf_classify(double p, double s_prev, double kd, int& e, uint64_t& r) {
  e = 0;
  r = 0;
  if (s_prev == 0.0) return p == 0.0 ? kBZero : kBOne;
  const uint64_t sbits = uint64_t(__double_as_longlong(s_prev));
  const int bexp = int(sbits >> 52);  // s_prev > 0: the biased exponent
  const double lo = __longlong_as_double((long long)(sbits & 0xfff0000000000000ull));  // 2^(bexp - 1023)
  const double margin = (kd + 257.0) * 0x1p-51 * s_prev;
  if (bexp < 3 || bexp > 2043 || !(s_prev - margin >= lo) || !(s_prev + margin + p * (1.0 + 0x1p-50) < 2.0 * lo))
    return kBOne;
  const int be = bexp - 1023;
  const double x = ldexp(p, 52 - be);  // exact, below 2^53
  // x < 2^52: x + 2^52 rounds x to an integer (ties to even) in the low mantissa bits;
  // from 2^52 on x is an integer already
  const bool small = x < 0x1p52;
  const double y = small ? x + 0x1p52 : x;
  if (small && fabs((y - 0x1p52) - x) == 0.5) return kBOne;  // a tie: rounding depends on c
  e = be;
  r = (uint64_t(__double_as_longlong(y)) & ((uint64_t(1) << 52) - 1)) | (small ? 0 : (uint64_t(1) << 52));
  return kBRun;
}

struct CdfLds {
  int sinfo[kCdfCh];           // per segment: first bucket | kind << 12 | (e + 1024) << 14
  uint64_t sbase[kCdfCh + 1];  // per segment: R prefix before its first bucket; [nseg] = the chunk's total
  int lastke[kCdfT];           // per lane: kind | (e + 1024) << 2 of its last bucket
  double shd[kCdfW];
  int shn[kCdfW];
  uint64_t shu[kCdfW];
};

struct CdfLane {
  uint64_t uk[kCdfE];  // chunk-local R prefix through each bucket (modulo 2^64), unless kept in LDS
  int kinds;           // bucket kinds, 2 bits each
  int starts;          // bit j: bucket j starts a segment
  int seg0;            // segments started before this lane's buckets
  __device__ int kind(int j) const { return (kinds >> (2 * j)) & 3; }
  __device__ int seg(int j) const { return seg0 + __popc(starts & ((2 << j) - 1)) - 1; }
};

// Segments of the chunk at bucket b whose buckets before it sum to about sb (every
// lane of the workgroup calls it, x = this lane's kCdfE buckets): records in
// L.sinfo / L.sbase, per-lane state in ln, the R prefixes in ukl[bucket] when ukl is
// given (else in ln.uk); returns the number of segments, and the chunk's approximate
// sum in tot. Ends with a barrier.
__device__ int cdf_segments(const double (&x)[kCdfE], uint64_t b, uint64_t n, double sb, CdfLds& L, CdfLane& ln,
                            double& tot, uint64_t* ukl = nullptr) {
  const int tid = threadIdx.x, pos0 = tid * kCdfE;
  const uint64_t i0 = b + uint64_t(pos0);
  double a[kCdfE];
  a[0] = x[0];
#pragma unroll
  for (int j = 1; j < kCdfE; ++j) a[j] = a[j - 1] + x[j];
  const double ex = sb + block_excl_scan(a[kCdfE - 1], L.shd, tot);
  const double kd0 = double(i0);
  int kinds = 0, ee[kCdfE];
  uint64_t r[kCdfE];
#pragma unroll
  for (int j = 0; j < kCdfE; ++j) {
    const double s_prev = j == 0 ? ex : ex + a[j - 1];
    int k = cdf_classify(x[j], s_prev, kd0 + double(j), ee[j], r[j]);
    if (i0 + j >= n || (k == kBOne && x[j] == 0.0 && pos0 + j > 0)) {
      k = kBJoin;  // pos0 + j > 0 here: bucket 0 of a chunk is always below n
      r[j] = 0;
    }
    kinds |= k << (2 * j);
  }
  L.lastke[tid] = ((kinds >> (2 * (kCdfE - 1))) & 3) | ((ee[kCdfE - 1] + 1024) << 2);
  __syncthreads();
  const int pke = tid > 0 ? L.lastke[tid - 1] : 0;
  int starts = 0, nst = 0;
  uint64_t rs = 0;
#pragma unroll
  for (int j = 0; j < kCdfE; ++j) {
    const int k = (kinds >> (2 * j)) & 3;
    const int pk = j == 0 ? (pke & 3) : ((kinds >> (2 * (j - 1))) & 3);
    const int pe = j == 0 ? (pke >> 2) - 1024 : ee[j - 1];
    if (k == kBOne || (k == kBRun && (pos0 + j == 0 || pk != kBRun || pe != ee[j]))) {
      starts |= 1 << j;
      ++nst;
    }
    rs += r[j];
  }
  int nseg;
  uint64_t utot;
  int cnt = block_excl_scan(nst, L.shn, nseg);
  uint64_t u = block_excl_scan(rs, L.shu, utot);
  ln.seg0 = cnt;
#pragma unroll
  for (int j = 0; j < kCdfE; ++j) {
    if ((starts >> j) & 1) {
      const int k = (kinds >> (2 * j)) & 3;
      L.sinfo[cnt] = (pos0 + j) | ((k == kBRun ? kSRun : kSOne) << 12) | ((ee[j] + 1024) << 14);
      L.sbase[cnt] = u;
      ++cnt;
    }
    u += r[j];
    if (ukl) ukl[pos0 + j] = u;
    else ln.uk[j] = u;
  }
  if (tid == 0) L.sbase[nseg] = utot;
  ln.kinds = kinds;
  ln.starts = starts;
  __syncthreads();
  return nseg;
}

// A run's addend u * T, or -1 when T is too large for one binade (forcing its fold).
__device__ __forceinline__ double run_addend(const CdfLds& L, int s, int e) {
  const uint64_t t = L.sbase[s + 1] - L.sbase[s];
  return t < (uint64_t(1) << 53) ? ldexp(double(t), e - 52) : -1.0;
}

// c after adding a run's addend, if the run stays in binade e from c; else false
__device__ __forceinline__ bool run_fits(double c, int e, double add) {
  const double lo = ldexp(1.0, e), hi = 2.0 * lo;
  return c >= lo && c < hi && add >= 0.0 && add <= hi - c;  // hi - c is exact for c in [lo, hi)
}

// One wave folds v[first..last) into c bucket by bucket (every lane holds the same c),
// writing each partial sum to out[first..last).
__device__ double cdf_fold(const double* v, double* out, uint64_t first, uint64_t last, double c) {
  const int lane = threadIdx.x & 63;
  for (uint64_t k0 = first; k0 < last; k0 += 64) {
    const uint64_t k = k0 + lane;
    const double x = k < last ? v[k] : 0.0;
    const int m = last - k0 < 64 ? int(last - k0) : 64;
    double mine = 0.0;
    for (int j = 0; j < m; ++j) {
      c = __dadd_rn(c, readlane_f64(x, j));
      if (lane == j) mine = c;
    }
    if (k < last) out[k] = mine;
  }
  return c;
}

// The chain of additions over the first m records held one per lane (add; lanes from
// m on must hold 0): c after them; before and after receive each lane's own record's
// sums. Exact for every run that fits its binade from its `before`, so the chain
// holds if all of them do.
__device__ __forceinline__ double walk_chain(double add, int m, double c, double& before, double& after) {
  const int lane = threadIdx.x & 63;
  for (int j0 = 0; j0 < m; j0 += 8) {
#pragma unroll
    for (int j = j0; j < j0 + 8; ++j) {
      const double a = readlane_f64(add, j);
      if (lane == j) before = c;
      c = __dadd_rn(c, a);
      if (lane == j) after = c;
    }
  }
  return c;
}

// The chain of walk_chain over records of the flat list (run: the record is a binade
// run of exponent e; lanes from m on must hold add 0). Consecutive chunks inside one
// binade leave consecutive runs of the same exponent: each such group is summed first
// (its shares are multiples of that binade's grid, so the sum is exact while the runs
// fit) and costs one step of the serial chain instead of one per record. Every lane
// then forms its record's before and after from its group's start and checks them
// against its neighbour's (before[k] == after[k-1], before[0] == c): by induction they
// all pass exactly when every value equals the one-by-one chain's. ok is false
// otherwise, and the caller walks the group record by record.
__device__ __forceinline__ double walk_merged(double add, bool run, int e, int m, double c, double& before,
                                              double& after, bool& ok) {
  const int lane = threadIdx.x & 63, rl = lane & 15;
  const bool live = lane < m;
  const int eprev = dpp_mov<0x138>(e);  // lane - 1's (wave_shr:1)
  const int rprev = dpp_mov<0x138>(run ? 1 : 0);
  const bool head = !live || lane == 0 || !run || rprev == 0 || eprev != e;
  const uint64_t heads = __ballot(head);
  const uint64_t upto = lane == 63 ? ~0ull : (2ull << lane) - 1;
  const int hk = 63 - __clzll((long long)(heads & upto));  // the lane's group head (lane 0 is one)
  // sum over the group up to this lane: the steps of wave_incl_scan, each taken only
  // from a lane of the same group
  double incl = add, t;
  t = dpp_mov<0x111>(incl);
  if (rl >= 1 && lane - 1 >= hk) incl = __dadd_rn(incl, t);
  t = dpp_mov<0x112>(incl);
  if (rl >= 2 && lane - 2 >= hk) incl = __dadd_rn(incl, t);
  t = dpp_mov<0x114>(incl);
  if (rl >= 4 && lane - 4 >= hk) incl = __dadd_rn(incl, t);
  t = dpp_mov<0x118>(incl);
  if (rl >= 8 && lane - 8 >= hk) incl = __dadd_rn(incl, t);
  t = dpp_mov<0x142>(incl);  // lane 15 of each row into the next row
  if ((lane & 16) && hk < (lane & ~15)) incl = __dadd_rn(incl, t);
  t = dpp_mov<0x143>(incl);  // lane 31 into rows 2 and 3
  if (lane >= 32 && hk < 32) incl = __dadd_rn(incl, t);
  const double c0 = c;
  const uint64_t livem = m >= 64 ? ~0ull : (1ull << m) - 1;
  uint64_t hs = heads & livem;
  double hb = 0.0;  // c before the lane's group
  while (hs) {
    const int j = __builtin_ctzll(hs);
    hs &= hs - 1;
    const uint64_t later = j == 63 ? 0ull : heads & ~((2ull << j) - 1);
    const int end = later ? __builtin_ctzll(later) - 1 : 63;  // the group's last lane
    const double gsum = readlane_f64(incl, end);
    if (lane >= j && lane <= end) hb = c;
    c = __dadd_rn(c, gsum);
  }
  const double pin = dpp_mov<0x138>(incl);
  before = lane == hk ? hb : __dadd_rn(hb, pin);
  after = __dadd_rn(before, add);
  const double aprev = dpp_mov<0x138>(after);
  const bool good = !live || before == (lane == 0 ? c0 : aprev);
  ok = __ballot(!good) == 0;
  return readlane_f64(after, m - 1);
}

// One wave walks a chunk's nseg segments (records in L) from c: per segment, the value
// of a single bucket or c before a run goes to sc. A group of 64 runs the addition
// chain first and keeps it when every run fits; otherwise it is walked segment by
// segment, a run that does not fit folded from pb in place over its buckets and marked
// folded in L.sinfo. Returns c after the last segment.
__device__ double cdf_walk(CdfLds& L, int nseg, double* sc, double* pb, double c) {
  const int lane = threadIdx.x & 63;
  nseg = __builtin_amdgcn_readfirstlane(nseg);  // the same in every lane
  for (int g = 0; g < nseg; g += 64) {
    const int s = g + lane;
    int kd = 0, e = 0;
    double add = 0.0;
    if (s < nseg) {
      const int inf = L.sinfo[s];
      kd = seg_kind(inf);
      e = seg_exp(inf);
      add = kd == kSOne ? pb[seg_first(inf)] : run_addend(L, s, e);
    }
    const int m = nseg - g < 64 ? nseg - g : 64;
    double before = 0.0, after = 0.0;
    const double cf = walk_chain(add, m, c, before, after);
    if (__ballot(s < nseg && kd == kSRun && !run_fits(before, e, add)) == 0) {
      if (s < nseg) sc[s] = kd == kSOne ? after : before;
      c = cf;
      continue;
    }
    double mine = 0.0;
    for (int j = 0; j < m; ++j) {
      const int kdj = __builtin_amdgcn_readlane(kd, j);
      const double addj = readlane_f64(add, j);
      if (kdj == kSOne) {
        c = __dadd_rn(c, addj);
        if (lane == j) mine = c;
        continue;
      }
      if (run_fits(c, __builtin_amdgcn_readlane(e, j), addj)) {
        if (lane == j) mine = c;
        c = __dadd_rn(c, addj);  // exact
        continue;
      }
      const int inf = L.sinfo[g + j];
      const int last = g + j + 1 < nseg ? seg_first(L.sinfo[g + j + 1]) : kCdfCh;
      c = cdf_fold(pb, pb, seg_first(inf), last, c);
      if (lane == 0) L.sinfo[g + j] = (inf & ~(3 << 12)) | (kSFolded << 12);
    }
    if (s < nseg) sc[s] = mine;
  }
  return c;
}

// The value of bucket j of a lane whose R prefix is uk (not past the end): 0 before
// the first non-zero bucket, a single bucket's value, a run's start plus u times the
// run's prefix, or (folded) the value in fv.
__device__ __forceinline__ double cdf_value(const CdfLane& ln, int j, uint64_t uk, const CdfLds& L,
                                            const double* sc, double fv) {
  if (ln.kind(j) == kBZero) return 0.0;
  const int s = ln.seg(j);
  const int inf = L.sinfo[s];
  const double c0 = sc[s];
  const uint64_t b0 = L.sbase[s];
  const int kd = seg_kind(inf);
  if (kd == kSFolded) return fv;
  if (kd == kSOne) return c0;
  return c0 + ldexp(double(uk - b0), seg_exp(inf) - 52);
}

__device__ __forceinline__ void cdf_load(const double* p, uint64_t n, uint64_t i0, double (&x)[kCdfE]) {
#pragma unroll
  for (int j = 0; j < kCdfE; ++j) x[j] = i0 + j < n ? p[i0 + j] : 0.0;
}

// Up to kCdfSmallMax buckets: one workgroup, chunk after chunk.
__global__ __launch_bounds__(kCdfT) void k_cdf_small(const double* __restrict__ p, uint64_t n,
                                                     double* __restrict__ cdf) {
  __shared__ CdfLds L;
  __shared__ double pb[kCdfCh];    // the chunk's probabilities; a folded run's values
  __shared__ uint64_t uk[kCdfCh];  // per bucket: chunk-local R prefix
  __shared__ double sc[kCdfCh];    // per segment: c before the run, or the single bucket's value
  const int tid = threadIdx.x, pos0 = tid * kCdfE;
  double sb = 0.0;  // approximate sum of the chunks before (the same in every lane)
  double c = 0.0;   // the fold's running sum (wave 0)
  for (uint64_t b = 0; b < n; b += kCdfCh) {
    const uint64_t i0 = b + uint64_t(pos0);
    double x[kCdfE];
    cdf_load(p, n, i0, x);
#pragma unroll
    for (int j = 0; j < kCdfE; ++j) pb[pos0 + j] = x[j];
    CdfLane ln;
    double tot;
    const int nseg = cdf_segments(x, b, n, sb, L, ln, tot, uk);
    if (tid < 64) c = cdf_walk(L, nseg, sc, pb, c);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kCdfE; ++j)
      if (i0 + j < n) cdf[i0 + j] = cdf_value(ln, j, uk[pos0 + j], L, sc, pb[pos0 + j]);
    sb = sb + tot;
    __syncthreads();  // the next chunk reuses LDS
  }
}

// ---- more than kCdfSmallMax buckets: every chunk at once around one walk ----

// approximate sum of each chunk
__global__ __launch_bounds__(kCdfT) void k_cdf_tot(const double* __restrict__ p, uint64_t n,
                                                   double* __restrict__ tot) {
  __shared__ double sh[kCdfW];
  const uint64_t b = uint64_t(blockIdx.x) * kCdfCh;
  double x[kCdfE];
  cdf_load(p, n, b + uint64_t(threadIdx.x) * kCdfE, x);
  double a = x[0];
#pragma unroll
  for (int j = 1; j < kCdfE; ++j) a += x[j];
  double t;
  (void)block_excl_scan(a, sh, t);
  if (threadIdx.x == 0) tot[blockIdx.x] = t;
}

// base[ch] = approximate sum of the chunks before ch (one workgroup)
__global__ __launch_bounds__(kCdfT) void k_cdf_base(const double* __restrict__ tot, uint64_t nch,
                                                    double* __restrict__ base) {
  __shared__ double sh[kCdfW];
  double carry = 0.0;
  for (uint64_t b = 0; b < nch; b += kCdfCh) {
    const uint64_t i0 = b + uint64_t(threadIdx.x) * kCdfE;
    double x[kCdfE];
    cdf_load(tot, nch, i0, x);
    double a = x[0];
#pragma unroll
    for (int j = 1; j < kCdfE; ++j) a += x[j];
    double t;
    double ex = carry + block_excl_scan(a, sh, t);
#pragma unroll
    for (int j = 0; j < kCdfE; ++j) {
      if (i0 + j < nch) base[i0 + j] = ex;
      ex += x[j];
    }
    carry += t;
    __syncthreads();
  }
}

// segment records of each chunk: nsegs[ch] (-1: more than kCdfMaxSeg, fold the
// chunk), and per segment its info word and addend
__global__ __launch_bounds__(kCdfT) void k_cdf_plan(const double* __restrict__ p, uint64_t n,
                                                    const double* __restrict__ base, int* __restrict__ nsegs,
                                                    int* __restrict__ rinfo, double* __restrict__ radd) {
  __shared__ CdfLds L;
  const uint64_t ch = blockIdx.x, b = ch * kCdfCh;
  double x[kCdfE];
  cdf_load(p, n, b + uint64_t(threadIdx.x) * kCdfE, x);
  CdfLane ln;
  double tot;
  const int nseg = cdf_segments(x, b, n, base[ch], L, ln, tot);
  if (threadIdx.x == 0) nsegs[ch] = nseg <= kCdfMaxSeg ? nseg : -1;
  if (nseg <= kCdfMaxSeg && int(threadIdx.x) < nseg) {
    const int s = threadIdx.x, inf = L.sinfo[s];
    rinfo[ch * kCdfMaxSeg + s] = inf;
    radd[ch * kCdfMaxSeg + s] = seg_kind(inf) == kSOne ? p[b + seg_first(inf)] : run_addend(L, s, seg_exp(inf));
  }
}

// roff[ch]: records before chunk ch in the flat list (a chunk folded whole is one
// record); roff[nch] = all records (one workgroup)
__global__ __launch_bounds__(kCdfT) void k_cdf_roff(const int* __restrict__ nsegs, uint64_t nch,
                                                    int* __restrict__ roff) {
  __shared__ int sh[kCdfW];
  int carry = 0;
  for (uint64_t b = 0; b < nch; b += kCdfCh) {
    const uint64_t i0 = b + uint64_t(threadIdx.x) * kCdfE;
    int cnt[kCdfE], t = 0;
#pragma unroll
    for (int j = 0; j < kCdfE; ++j) {
      const int ns = i0 + j < nch ? nsegs[i0 + j] : 0;
      cnt[j] = ns < 0 ? 1 : ns;
      t += cnt[j];
    }
    int tot;
    int ex = carry + block_excl_scan(t, sh, tot);
#pragma unroll
    for (int j = 0; j < kCdfE; ++j) {
      if (i0 + j < nch) roff[i0 + j] = ex;
      ex += cnt[j];
    }
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) roff[nch] = carry;
}

// the flat record list in walking order: info word (-1: fold chunk dst whole), addend,
// and the record's slot dst = chunk * kCdfMaxSeg + segment (one wave per chunk)
__global__ __launch_bounds__(64) void k_cdf_pack(const int* __restrict__ nsegs, const int* __restrict__ roff,
                                                 const int* __restrict__ rinfo, const double* __restrict__ radd,
                                                 int* __restrict__ finfo, double* __restrict__ fadd,
                                                 int* __restrict__ fdst) {
  const uint64_t ch = blockIdx.x;
  const int ns = nsegs[ch], s = threadIdx.x, at = roff[ch] + s;
  if (ns < 0) {
    if (s == 0) {
      finfo[at] = -1;
      fadd[at] = 0.0;
      fdst[at] = int(ch);
    }
  } else if (s < ns) {
    finfo[at] = rinfo[ch * kCdfMaxSeg + s];
    fadd[at] = radd[ch * kCdfMaxSeg + s];
    fdst[at] = int(ch * kCdfMaxSeg + s);
  }
}

// The walk over all records in order (one workgroup): waves 1.. stage the next
// kCdfCh records in LDS while wave 0 walks the current ones, 64 at a time: the chain
// of additions when every run fits, else record by record, folding a chunk marked
// whole or a run that does not fit straight into cdf (and marking the run folded in
// rinfo). rc[slot] receives each segment's start value.
__global__ __launch_bounds__(kCdfT) void k_cdf_walk(const double* __restrict__ p, uint64_t n,
                                                    const int* __restrict__ roff, uint64_t nch,
                                                    const int* __restrict__ finfo, const double* __restrict__ fadd,
                                                    const int* __restrict__ fdst, int* __restrict__ rinfo,
                                                    double* __restrict__ rc, double* __restrict__ cdf) {
  __shared__ int si[2][kCdfCh];
  __shared__ int sd[2][kCdfCh];
  __shared__ double sa[2][kCdfCh];
  const int tid = threadIdx.x, lane = tid & 63;
  const int nrec = roff[nch];
  const int nblk = (nrec + kCdfCh - 1) / kCdfCh;
  auto stage = [&](int blk, int buf, int from, int stride) {
    for (int q = from; q < kCdfCh; q += stride) {
      const int r = blk * kCdfCh + q;
      if (r >= nrec) break;
      si[buf][q] = finfo[r];
      sd[buf][q] = fdst[r];
      sa[buf][q] = fadd[r];
    }
  };
  if (nblk > 0) stage(0, 0, tid, kCdfT);
  __syncthreads();
  double c = 0.0;
  for (int blk = 0; blk < nblk; ++blk) {
    const int buf = blk & 1;
    if (tid >= 64) {
      if (blk + 1 < nblk) stage(blk + 1, buf ^ 1, tid - 64, kCdfT - 64);
    } else {
      const int cnt = nrec - blk * kCdfCh < kCdfCh ? nrec - blk * kCdfCh : kCdfCh;
      for (int g = 0; g < cnt; g += 64) {
        const int q = g + lane;
        int inf = -1, dst = 0;
        double add = 0.0;
        if (q < cnt) {
          inf = si[buf][q];
          dst = sd[buf][q];
          add = sa[buf][q];
        }
        const int m = cnt - g < 64 ? cnt - g : 64;
        double before = 0.0, after = 0.0;
        const bool run = inf >= 0 && seg_kind(inf) == kSRun;
        bool ok;
        const double cf = walk_merged(inf < 0 ? 0.0 : add, run, seg_exp(inf), m, c, before, after, ok);
        const bool bad = q < cnt && (inf < 0 || (run && !run_fits(before, seg_exp(inf), add)));
        if (ok && __ballot(bad) == 0) {
          if (q < cnt) rc[dst] = seg_kind(inf) == kSOne ? after : before;
          c = cf;
          continue;
        }
        for (int j = 0; j < m; ++j) {
          const int infj = __builtin_amdgcn_readlane(inf, j);
          const int dstj = __builtin_amdgcn_readlane(dst, j);
          const double addj = readlane_f64(add, j);
          if (infj < 0) {  // a chunk folded whole
            const uint64_t b = uint64_t(dstj) * kCdfCh;
            c = cdf_fold(p, cdf, b, b + kCdfCh < n ? b + kCdfCh : n, c);
            continue;
          }
          if (seg_kind(infj) == kSOne) {
            c = __dadd_rn(c, addj);
            if (lane == 0) rc[dstj] = c;
            continue;
          }
          if (run_fits(c, seg_exp(infj), addj)) {
            if (lane == 0) rc[dstj] = c;
            c = __dadd_rn(c, addj);  // exact
            continue;
          }
          // fold the run's buckets up to the chunk's next segment, else to its end
          const uint64_t ch = uint64_t(dstj) / kCdfMaxSeg, b = ch * kCdfCh;
          uint64_t last = b + kCdfCh < n ? b + kCdfCh : n;
          if (g + j + 1 < cnt) {
            const int in2 = si[buf][g + j + 1], d2 = sd[buf][g + j + 1];
            if (in2 >= 0 && uint64_t(d2) / kCdfMaxSeg == ch) last = b + seg_first(in2);
          } else if (blk * kCdfCh + g + j + 1 < nrec) {
            const int r2 = blk * kCdfCh + g + j + 1;
            const int in2 = finfo[r2], d2 = fdst[r2];
            if (in2 >= 0 && uint64_t(d2) / kCdfMaxSeg == ch) last = b + seg_first(in2);
          }
          c = cdf_fold(p, cdf, b + seg_first(infj), last, c);
          if (lane == 0) rinfo[dstj] = (infj & ~(3 << 12)) | (kSFolded << 12);
        }
      }
    }
    __syncthreads();
  }
}

// every chunk's buckets from its segments' start values (folded chunks and runs are
// already written)
__global__ __launch_bounds__(kCdfT) void k_cdf_fill(const double* __restrict__ p, uint64_t n,
                                                    const double* __restrict__ base, const int* __restrict__ nsegs,
                                                    const int* __restrict__ rinfo, const double* __restrict__ rc,
                                                    double* __restrict__ cdf) {
  __shared__ CdfLds L;
  __shared__ double sc[kCdfMaxSeg];
  const uint64_t ch = blockIdx.x, b = ch * kCdfCh;
  const int ns = nsegs[ch];
  if (ns < 0) return;  // the same for the whole workgroup
  const uint64_t i0 = b + uint64_t(threadIdx.x) * kCdfE;
  double x[kCdfE];
  cdf_load(p, n, i0, x);
  CdfLane ln;
  double tot;
  (void)cdf_segments(x, b, n, base[ch], L, ln, tot);  // the same segments as k_cdf_plan
  if (int(threadIdx.x) < ns) {
    sc[threadIdx.x] = rc[ch * kCdfMaxSeg + threadIdx.x];
    L.sinfo[threadIdx.x] = rinfo[ch * kCdfMaxSeg + threadIdx.x];
  }
  __syncthreads();
  // when the lane's buckets start 16-byte aligned (always for yh_sample: cdf sits n
  // doubles after p, n a power of two above one chunk) and none of them is past the
  // end or already folded, four 16-byte stores write them
  double v[kCdfE];
  bool keep[kCdfE], all = (reinterpret_cast<uintptr_t>(cdf + i0) & 15) == 0;
#pragma unroll
  for (int j = 0; j < kCdfE; ++j) {
    keep[j] = i0 + j < n && !(ln.kind(j) != kBZero && seg_kind(L.sinfo[ln.seg(j)]) == kSFolded);
    all = all && keep[j];
    v[j] = keep[j] ? cdf_value(ln, j, ln.uk[j], L, sc, 0.0) : 0.0;
  }
  if (all) {
    double2* d = reinterpret_cast<double2*>(cdf + i0);
#pragma unroll
    for (int h = 0; h < kCdfE / 2; ++h) d[h] = make_double2(v[2 * h], v[2 * h + 1]);
  } else {
#pragma unroll
    for (int j = 0; j < kCdfE; ++j)
      if (keep[j]) cdf[i0 + j] = v[j];
  }
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

size_t cdf_scratch_bytes(uint64_t n) {
  if (n <= kCdfSmallMax) return 0;
  const uint64_t nch = (n + kCdfCh - 1) / kCdfCh;
  // tot, base; radd, rc, fadd per slot (doubles); nsegs, roff; rinfo, finfo, fdst per slot (ints)
  return nch * (2 * sizeof(double) + 3 * kCdfMaxSeg * sizeof(double)) +
         (2 * nch + 1 + 3 * nch * kCdfMaxSeg) * sizeof(int) + 256;
}

int launch_cdf(const double* p, uint64_t n, double* cdf, void* scratch, hipStream_t st) {
  if (n <= kCdfSmallMax) {
    hipLaunchKernelGGL(k_cdf_small, dim3(1), dim3(kCdfT), 0, st, p, n, cdf);
    return YH_OK;
  }
  const uint64_t nch = (n + kCdfCh - 1) / kCdfCh;
  if (nch * kCdfMaxSeg > 0x7fffffffull)  // record slots are int
    return fail(YH_ERR_INVALID_ARG, "distribution too large for the device cumulative sum");
  const uint64_t ns = nch * kCdfMaxSeg;
  double* tot = static_cast<double*>(scratch);
  double* base = tot + nch;
  double* radd = base + nch;
  double* rc = radd + ns;
  double* fadd = rc + ns;
  int* nsegs = reinterpret_cast<int*>(fadd + ns);
  int* roff = nsegs + nch;
  int* rinfo = roff + nch + 1;
  int* finfo = rinfo + ns;
  int* fdst = finfo + ns;
  const dim3 g(static_cast<unsigned>(nch));
  hipLaunchKernelGGL(k_cdf_tot, g, dim3(kCdfT), 0, st, p, n, tot);
  hipLaunchKernelGGL(k_cdf_base, dim3(1), dim3(kCdfT), 0, st, tot, nch, base);
  hipLaunchKernelGGL(k_cdf_plan, g, dim3(kCdfT), 0, st, p, n, base, nsegs, rinfo, radd);
  hipLaunchKernelGGL(k_cdf_roff, dim3(1), dim3(kCdfT), 0, st, nsegs, nch, roff);
  hipLaunchKernelGGL(k_cdf_pack, g, dim3(64), 0, st, nsegs, roff, rinfo, radd, finfo, fadd, fdst);
  hipLaunchKernelGGL(k_cdf_walk, dim3(1), dim3(kCdfT), 0, st, p, n, roff, nch, finfo, fadd, fdst, rinfo, rc, cdf);
  hipLaunchKernelGGL(k_cdf_fill, g, dim3(kCdfT), 0, st, p, n, base, nsegs, rinfo, rc, cdf);
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
