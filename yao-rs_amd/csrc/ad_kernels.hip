// ad_kernels.hip — kernels of adjoint-mode differentiation (expect_grad,
// src/ad.rs:75-156): the co-state lambda = H psi for a Pauli polynomial H, and the
// generator overlap <lambda| P_ctrl G |psi> of one gate parameter.
#include "kernels.h"

namespace yh {

namespace {

// lam_i (+)= sum_t c_t (P_t psi)_i over nt terms. (P psi)_i = f(i) psi_j, where each
// site replaces bit b of i by the operator row's source bit and multiplies by the
// row's single entry (operator.rs:53-65); the entries are 0, +-1, +-i, so f psi_j is
// exact, and c_t (f psi_j) is the plain complex product, accumulated in term order
// like the reference's lambda_state[i] += coeff * tmp[i] (ad.rs:101-116).
template <class R>
__global__ __launch_bounds__(kBlock) void k_pauli_accum(const typename Cplx<R>::V* __restrict__ psi,
                                                        typename Cplx<R>::V* __restrict__ lam, uint64_t n,
                                                        const TermDev* __restrict__ terms,
                                                        const double* __restrict__ coeffs, int nt, int first) {
#pragma clang fp contract(off)
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    double ar = 0.0, ai = 0.0;
    if (!first) {
      ar = lam[i].x;
      ai = lam[i].y;
    }
    for (int t = 0; t < nt; ++t) {
      const TermDev& T = terms[t];
      uint64_t j = i;
      double fr = 1.0, fi = 0.0;
      for (int k = 0; k < T.nsites; ++k) {
        const int b = T.bit[k];
        const int v = int((i >> b) & 1);
        j = (j & ~(uint64_t(1) << b)) | (uint64_t(T.src[k][v]) << b);
        const double gr = T.fac[k][v][0], gi = T.fac[k][v][1];
        const double nr = fr * gr - fi * gi, ni = fr * gi + fi * gr;
        fr = nr;
        fi = ni;
      }
      const double pr = psi[j].x, pi = psi[j].y;
      const double tr = fr * pr - fi * pi, ti = fr * pi + fi * pr;
      const double cr = coeffs[2 * t], ci = coeffs[2 * t + 1];
      ar += cr * tr - ci * ti;
      ai += cr * ti + ci * tr;
    }
    lam[i].x = R(ar);
    lam[i].y = R(ai);
  }
}

// Per block: sum over groups g of sum_j conj(lam[idx_j]) sum_k G[j][k] psi[idx_k],
// idx_j = base(g) + offs[j]; groups whose controls do not match contribute zero
// (the generator maps inactive control branches to zero, ad.rs:29-64).
template <class R, int NT>
__global__ __launch_bounds__(kBlock) void k_gen_inner(const typename Cplx<R>::V* __restrict__ psi,
                                                      const typename Cplx<R>::V* __restrict__ lam, GroupArgs ga,
                                                      const double2* __restrict__ G, double* __restrict__ partial) {
  constexpr int D = 1 << NT;
  double2 g[D * D];
#pragma unroll
  for (int e = 0; e < D * D; ++e) g[e] = G[e];
  double sr = 0.0, si = 0.0;
  for (uint64_t q = uint64_t(blockIdx.x) * kBlock + threadIdx.x; q < ga.ngroups;
       q += uint64_t(gridDim.x) * kBlock) {
    const uint64_t base = insert_zeros(q, ga.lk.pos, ga.lk.npos) | ga.ctrl_set;
    double xr[D], xi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      xr[k] = psi[base + ga.offs[k]].x;
      xi[k] = psi[base + ga.offs[k]].y;
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      double yr = 0.0, yi = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        yr += g[j * D + k].x * xr[k] - g[j * D + k].y * xi[k];
        yi += g[j * D + k].x * xi[k] + g[j * D + k].y * xr[k];
      }
      const double lr = lam[base + ga.offs[j]].x, li = lam[base + ga.offs[j]].y;
      sr += lr * yr + li * yi;  // conj(l) * y
      si += lr * yi - li * yr;
    }
  }
  __shared__ double sh[2][kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    si += __shfl_xor(si, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = sr;
    sh[1][threadIdx.x >> 6] = si;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) {
      a += sh[0][w];
      b += sh[1][w];
    }
    partial[2 * blockIdx.x] = a;
    partial[2 * blockIdx.x + 1] = b;
  }
}

// out[0..1] = fold of nblocks partial pairs in a fixed order
__global__ __launch_bounds__(64) void k_fold_pair(const double* __restrict__ partial, int nblocks,
                                                  double* __restrict__ out) {
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < nblocks; k += 64) {
    a += partial[2 * k];
    b += partial[2 * k + 1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

unsigned grid_for(uint64_t work) {
  uint64_t b = (work + kBlock - 1) / kBlock;
  return unsigned(b < 1 ? 1 : (b > uint64_t(kRedBlocks) ? kRedBlocks : b));
}

}  // namespace

template <class R>
int launch_pauli_accum(const typename Cplx<R>::V* psi, typename Cplx<R>::V* lam, uint64_t n, const TermDev* terms,
                       const double* coeffs, int nt, bool first, hipStream_t st) {
  hipLaunchKernelGGL(k_pauli_accum<R>, dim3(grid_for(n)), dim3(kBlock), 0, st, psi, lam, n, terms, coeffs, nt,
                     first ? 1 : 0);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

template <class R>
int launch_gen_inner(const typename Cplx<R>::V* psi, const typename Cplx<R>::V* lam, const GroupArgs& ga, int nt,
                     const double2* G, double* partial, double* out, hipStream_t st) {
  const unsigned grid = grid_for(ga.ngroups);
  if (nt == 1)
    hipLaunchKernelGGL((k_gen_inner<R, 1>), dim3(grid), dim3(kBlock), 0, st, psi, lam, ga, G, partial);
  else if (nt == 2)
    hipLaunchKernelGGL((k_gen_inner<R, 2>), dim3(grid), dim3(kBlock), 0, st, psi, lam, ga, G, partial);
  else
    return fail(YH_ERR_UNSUPPORTED, "generators act on 1 or 2 targets");
  YH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_fold_pair, dim3(1), dim3(64), 0, st, static_cast<const double*>(partial), int(grid), out);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

template int launch_pauli_accum<double>(const double2*, double2*, uint64_t, const TermDev*, const double*, int, bool,
                                        hipStream_t);
template int launch_pauli_accum<float>(const float2*, float2*, uint64_t, const TermDev*, const double*, int, bool,
                                       hipStream_t);
template int launch_gen_inner<double>(const double2*, const double2*, const GroupArgs&, int, const double2*, double*,
                                      double*, hipStream_t);
template int launch_gen_inner<float>(const float2*, const float2*, const GroupArgs&, int, const double2*, double*,
                                     double*, hipStream_t);

}  // namespace yh
