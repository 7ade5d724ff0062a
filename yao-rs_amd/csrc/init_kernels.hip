// init_kernels.hip — register and density-matrix constructors on the device:
// deterministic_state, rand_state (from caller-supplied uniforms), uniform_state,
// DensityMatrix::mixed and density_matrix_from_reg. Products follow the
// reference's unfused complex arithmetic (contraction into FMA is switched off
// where a result is compared bit for bit).
#include "kernels.h"

namespace yh {

namespace {

// deterministic_state (register.rs:71-84) before normalisation:
// amplitude k = (cos(0.1 k), sin(0.2 k))
template <class R>
__global__ __launch_bounds__(kBlock) void k_deterministic(typename Cplx<R>::V* __restrict__ s, uint64_t n) {
  for (uint64_t k = uint64_t(blockIdx.x) * kBlock + threadIdx.x; k < n; k += uint64_t(gridDim.x) * kBlock) {
    const double kf = double(k);
    s[k] = cmk<typename Cplx<R>::V>(R(cos(0.1 * kf)), R(sin(0.2 * kf)));
  }
}

// rand_state (register.rs:35-45) before normalisation: amplitude k = (u[2k] - 0.5,
// u[2k+1] - 0.5) from the caller's stream of f64 uniforms
template <class R>
__global__ __launch_bounds__(kBlock) void k_from_uniforms(typename Cplx<R>::V* __restrict__ s,
                                                          const double* __restrict__ u, uint64_t n) {
  for (uint64_t k = uint64_t(blockIdx.x) * kBlock + threadIdx.x; k < n; k += uint64_t(gridDim.x) * kBlock)
    s[k] = cmk<typename Cplx<R>::V>(R(u[2 * k] - 0.5), R(u[2 * k + 1] - 0.5));
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_fill(typename Cplx<R>::V* __restrict__ s, uint64_t n, double re) {
  for (uint64_t k = uint64_t(blockIdx.x) * kBlock + threadIdx.x; k < n; k += uint64_t(gridDim.x) * kBlock)
    s[k] = cmk<typename Cplx<R>::V>(R(re), R(0));
}

// DensityMatrix::mixed (density_matrix.rs:37-61): rho[r][c] = sum over registers j,
// in order, of (w_j psi_j[r]) conj(psi_j[c]), each product rounded as written
__global__ __launch_bounds__(kBlock) void k_mixed(double2* __restrict__ rho, uint64_t dim,
                                                  const double2* const* __restrict__ psi,
                                                  const double* __restrict__ w, int nreg) {
#pragma clang fp contract(off)
  for (uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x; e < dim * dim; e += uint64_t(gridDim.x) * kBlock) {
    const uint64_t r = e / dim, c = e % dim;
    double ar = 0.0, ai = 0.0;
    for (int j = 0; j < nreg; ++j) {
      const double2 x = psi[j][r], y = psi[j][c];
      // Complex64::new(w, 0) * x: (w x.re - 0 x.im, w x.im + 0 x.re)
      const double wr = w[j] * x.x - 0.0 * x.y, wi = w[j] * x.y + 0.0 * x.x;
      // times conj(y)
      const double pr = wr * y.x - wi * (-y.y), pi = wr * (-y.y) + wi * y.x;
      ar += pr;
      ai += pi;
    }
    rho[e] = double2{ar, ai};
  }
}

// density_matrix_from_reg (density_matrix.rs:266-272) without the full matrix:
// out[rr][rc] = sum over the traced bit values t (increasing) of
// psi[row] conj(psi[col]), the entries from_reg (:23-35) would hold and partial_tr
// (:80-111) would add in increasing row order
template <class R>
__global__ __launch_bounds__(kBlock) void k_reduced(const typename Cplx<R>::V* __restrict__ psi, PtrArgs a,
                                                    double2* __restrict__ out) {
#pragma clang fp contract(off)
  const uint64_t dr = uint64_t(1) << a.nkept;
  for (uint64_t e = uint64_t(blockIdx.x) * kBlock + threadIdx.x; e < dr * dr; e += uint64_t(gridDim.x) * kBlock) {
    const uint64_t rr = e / dr, rc = e % dr;
    uint64_t rb = 0, cb = 0;
    for (int k = 0; k < a.nkept; ++k) {
      rb |= ((rr >> k) & 1) << a.kept[k];
      cb |= ((rc >> k) & 1) << a.kept[k];
    }
    const uint64_t nt = uint64_t(1) << a.ntr;
    double sr = 0.0, si = 0.0;
    for (uint64_t t = 0; t < nt; ++t) {
      uint64_t tb = 0;
      for (int k = 0; k < a.ntr; ++k) tb |= ((t >> k) & 1) << a.tr[k];
      const double xr = double(psi[rb | tb].x), xi = double(psi[rb | tb].y);
      const double yr = double(psi[cb | tb].x), yi = -double(psi[cb | tb].y);
      sr += xr * yr - xi * yi;
      si += xr * yi + xi * yr;
    }
    out[e] = double2{sr, si};
  }
}

}  // namespace

template <class R>
int launch_deterministic(typename Cplx<R>::V* s, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_deterministic<R>, dim3(capped_grid(n, kBlock)), dim3(kBlock), 0, st, s, n);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

template <class R>
int launch_from_uniforms(typename Cplx<R>::V* s, const double* u, uint64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_from_uniforms<R>, dim3(capped_grid(n, kBlock)), dim3(kBlock), 0, st, s, u, n);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

template <class R>
int launch_fill(typename Cplx<R>::V* s, uint64_t n, double re, hipStream_t st) {
  hipLaunchKernelGGL(k_fill<R>, dim3(capped_grid(n, kBlock)), dim3(kBlock), 0, st, s, n, re);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

int launch_mixed(double2* rho, uint64_t dim, const double2* const* psi, const double* w, int nreg, hipStream_t st) {
  hipLaunchKernelGGL(k_mixed, dim3(capped_grid(dim * dim, kBlock)), dim3(kBlock), 0, st, rho, dim, psi, w, nreg);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

template <class R>
int launch_reduced(const typename Cplx<R>::V* psi, const PtrArgs& a, double2* out, hipStream_t st) {
  const uint64_t n = (uint64_t(1) << a.nkept) << a.nkept;
  hipLaunchKernelGGL(k_reduced<R>, dim3(capped_grid(n, kBlock)), dim3(kBlock), 0, st, psi, a, out);
  YH_HIP(hipGetLastError());
  return YH_OK;
}

#define YH_INST(R)                                                                        \
  template int launch_deterministic<R>(Cplx<R>::V*, uint64_t, hipStream_t);               \
  template int launch_from_uniforms<R>(Cplx<R>::V*, const double*, uint64_t, hipStream_t); \
  template int launch_fill<R>(Cplx<R>::V*, uint64_t, double, hipStream_t);                \
  template int launch_reduced<R>(const Cplx<R>::V*, const PtrArgs&, double2*, hipStream_t);
YH_INST(double)
YH_INST(float)
#undef YH_INST

}  // namespace yh
