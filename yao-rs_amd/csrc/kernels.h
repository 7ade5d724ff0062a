// kernels.h — kernel argument structs and launcher declarations.
#pragma once

#include "common.h"

namespace yh {

// Group enumeration: thread id -> basis index with locked bits zero, then ctrl_set.
struct GroupArgs {
  uint64_t ngroups;
  uint64_t ctrl_set;   // control bits whose required value is 1
  uint64_t offs[32];   // offsets of the 2^nt group members (gate-index order)
  LockedBits lk;
};

template <class R>
struct Mat2 {
  typename Cplx<R>::V a, b, c, d;
};

template <class R>
struct DiagArgs {
  typename Cplx<R>::V d[4];
  int32_t tbit[2];
};

// ---- sv_kernels.hip ----
template <class R>
int launch_1q(typename Cplx<R>::V* s, int nbits, int bit, const Mat2<R>& m, hipStream_t st);
template <class R>
int launch_dense(typename Cplx<R>::V* s, const GroupArgs& ga, int nt,
                 const typename Cplx<R>::V* mat, hipStream_t st);
template <class R>
int launch_diag(typename Cplx<R>::V* s, const GroupArgs& ga, int nt, const DiagArgs<R>& d,
                hipStream_t st);
template <class R>
int launch_xflip(typename Cplx<R>::V* s, const GroupArgs& ga, uint64_t tmask, hipStream_t st);
template <class R>
int launch_swap(typename Cplx<R>::V* s, const GroupArgs& ga, uint64_t m0, uint64_t m1,
                hipStream_t st);
template <class R>
int launch_dense_oop(typename Cplx<R>::V* dst, const typename Cplx<R>::V* src, const GroupArgs& ga,
                     int nt, const typename Cplx<R>::V* mat, const uint64_t* offs, hipStream_t st);

// ---- obs_kernels.hip ----
// Partial-sum workspace: kRedBlocks blocks x up to kMaxTermsPerPass complex terms.
constexpr int kRedBlocks = 1024;
constexpr int kMaxTermsPerPass = 8;
constexpr int kMaxSites = 64;

// One Pauli-product term in device form: for each site, the bit, the source-bit
// table and factor table of its operator (row v of the 2x2 reads input bit src[v]).
struct TermDev {
  int32_t nsites;
  int32_t bit[kMaxSites];
  uint8_t src[kMaxSites][2];
  double fac[kMaxSites][2][2];  // [site][v] = (re, im)
};

// ---- ad_kernels.hip ----
template <class R>
int launch_pauli_accum(const typename Cplx<R>::V* psi, typename Cplx<R>::V* lam, uint64_t n, const TermDev* terms,
                       const double* coeffs, int nt, bool first, hipStream_t st);
template <class R>
int launch_gen_inner(const typename Cplx<R>::V* psi, const typename Cplx<R>::V* lam, const GroupArgs& ga, int nt,
                     const double2* G, double* partial, double* out, hipStream_t st);

template <class R>
int launch_norm2(const typename Cplx<R>::V* s, uint64_t n, double* partial, double* out,
                 hipStream_t st);
template <class R>
int launch_inner(const typename Cplx<R>::V* a, const typename Cplx<R>::V* b, uint64_t n,
                 double* partial, double* out, hipStream_t st);
// s /= sqrt(*sumsq) when sqrt(*sumsq) > thresh
template <class R>
int launch_scale(typename Cplx<R>::V* s, uint64_t n, const double* sumsq, double thresh,
                 hipStream_t st);
// expectation of up to kMaxTermsPerPass terms; out receives 2*nterms doubles
// (sum_i conj(psi_i) (P_t psi)_i for each term)
template <class R>
int launch_expect_terms(const typename Cplx<R>::V* s, uint64_t n, const TermDev* terms_dev,
                        int nterms, double* partial, double* out, hipStream_t st);
// density-matrix flavour: tr(P_t rho) with LSB-first sites, rho of dimension dim
int launch_dm_expect_terms(const double2* rho, uint64_t dim, const TermDev* terms_dev, int nterms,
                           double* partial, double* out, hipStream_t st);
template <class R>
int launch_probs_full(const typename Cplx<R>::V* s, uint64_t n, double* out, hipStream_t st);
// marginal probabilities: bucket m has output bit nlocs-1-idx <- basis bit locbits[idx];
// each bucket is summed in increasing basis order over `nchunks` chunks, then the
// chunks are summed in order (nchunks == 1 is the reference's sequential fold).
// locbits_dev is the device copy of the host array locbits.
int probs_marginal_dev(const void* s, int dtype, int nbits, const int32_t* locbits_dev,
                       const int32_t* locbits, int nlocs, double* out, double* ws,
                       uint64_t nchunks, hipStream_t st);
// density matrix: full diagonal when locbits == nullptr, else marginal as above
int dm_probs_dev(const double2* rho, int nbits, const int32_t* locbits_dev,
                 const int32_t* locbits, int nlocs, double* out, double* ws, uint64_t nchunks,
                 hipStream_t st);
// cumulative sum with the bits of the sequential left fold (scratch: cdf_scratch_bytes(n)
// bytes of device memory, none up to 4096 buckets) and per-shot search
size_t cdf_scratch_bytes(uint64_t n);
int launch_cdf(const double* p, uint64_t n, double* cdf, void* scratch, hipStream_t st);
int launch_search(const double* cdf, uint64_t n, const double* r, uint64_t nshots, uint64_t* out,
                  hipStream_t st);
// collapse to outcome (zero the rest), reset permutation, removed-qubit gather
template <class R>
int launch_collapse_mask(typename Cplx<R>::V* s, uint64_t n, uint64_t mask, uint64_t val,
                         hipStream_t st);
// for every basis b with measured bits == from: s[b ^ sm] = s[b]; s[b] = 0
int launch_reset_dev(void* s, int dtype, const LockedBits& lk, uint64_t ngroups, uint64_t from,
                     uint64_t sm, hipStream_t st);
template <class R>
int launch_remove_gather(const typename Cplx<R>::V* s, typename Cplx<R>::V* dst, uint64_t nnew,
                         const LockedBits& lk, uint64_t fixed, hipStream_t st);

// ---- dm_kernels.hip ----
constexpr int kMaxMicro = 8;
struct DmMicro {
  int32_t type;   // 0 gate, 1 channel
  int32_t q;      // block qubit index (0..K-1) or -1 for all K block qubits
  int32_t nk;     // Kraus operator count (channel) or 1 (gate)
  int32_t moff;   // offset (in complex entries) into the matrix table
};
struct DmArgs {
  uint64_t ngroups;
  LockedBits lk;          // row and column target bits (sorted)
  uint64_t roff[4];       // row offsets of block index r (gate-index order)
  uint64_t coff[4];       // column offsets
  uint64_t rc_mask, rc_val;  // row-side control predicate (gate micro-op 0 only)
  uint64_t cc_mask, cc_val;  // column-side control predicate
  int32_t nmicro;
  DmMicro micro[kMaxMicro];
};
int launch_dm_block(double2* rho, const DmArgs& a, int K, const double2* mats, hipStream_t st);
int launch_dm_trace(const double2* rho, uint64_t dim, double* partial, double* out, hipStream_t st);
int launch_dm_purity(const double2* rho, uint64_t dim, double* partial, double* out, hipStream_t st);
// partial trace: kept and traced bit positions, ascending
struct PtrArgs {
  int32_t nkept, ntr;
  int32_t kept[32], tr[32];
};
int launch_partial_tr(const double2* rho, uint64_t dim, const PtrArgs& a, double2* out, hipStream_t st);
int launch_dm_outer(const double2* psi, double2* rho, uint64_t dim, hipStream_t st);
// ---- init_kernels.hip ----
template <class R>
int launch_deterministic(typename Cplx<R>::V* s, uint64_t n, hipStream_t st);
template <class R>
int launch_from_uniforms(typename Cplx<R>::V* s, const double* u, uint64_t n, hipStream_t st);
template <class R>
int launch_fill(typename Cplx<R>::V* s, uint64_t n, double re, hipStream_t st);
int launch_mixed(double2* rho, uint64_t dim, const double2* const* psi, const double* w, int nreg, hipStream_t st);
template <class R>
int launch_reduced(const typename Cplx<R>::V* psi, const PtrArgs& a, double2* out, hipStream_t st);
// acc[i] += b[i] for n complex128 entries (the branch accumulation of apply_channel)
int launch_add_inplace(double2* acc, const double2* b, uint64_t n, hipStream_t st);

}  // namespace yh
