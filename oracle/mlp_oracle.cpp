// mlp_oracle.cpp — CPU restatement of the MLP surrogate's forward pass, MSE loss
// and backward pass, in FP64. TEST INFRASTRUCTURE ONLY (the checker of the HIP
// training kernels); nothing in the product library links it.
//
// The surrogate has no counterpart in the reference: it is the learned energy model
// that screens ansatz parameters before the batched simulator re-evaluates them.
// Network: in_dim inputs -> features -> hidden x [width, exact GELU] -> 1. With
// trig = 0 the features are the inputs; with trig = 1 features 2j, 2j+1 are
// cos x_j, sin x_j. Parameters are flat, layer by layer: W_l (fan_out x fan_in,
// row-major) then b_l.
//
// quant = 1 rounds to bfloat16 (round to nearest even) at the points where the GPU
// stores bfloat16 operands: the input rows, the hidden-layer weights, the hidden
// activations and the hidden pre-activation gradients that feed the weight and
// input-gradient products. quant = 0 is the plain FP64 network (pinned against
// finite differences and an independent autograd in the tests).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

double bf16_round(double v) {
  float f = float(v);
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  std::memcpy(&f, &u, 4);
  return double(f);
}

double gelu(double z) { return 0.5 * z * (1.0 + std::erf(z / std::sqrt(2.0))); }
double gelu_d(double z) {
  const double phi = std::exp(-0.5 * z * z) / std::sqrt(2.0 * M_PI);
  return 0.5 * (1.0 + std::erf(z / std::sqrt(2.0))) + z * phi;
}

}  // namespace

extern "C" {

// Forward + backward over n rows. out[n] = predictions, *loss = mean squared error,
// grads[nparams] = d loss / d params (may be null).
int orc_mlp_loss_grad(int in_dim, int trig, int width, int hidden, const double* params, const float* x,
                      const float* y, int n, int quant, double* out, double* loss, double* grads) {
  if (in_dim < 1 || width < 1 || hidden < 1 || n < 1) return 1;
  const int nraw = in_dim;
  if (trig) in_dim *= 2;
  auto q = [&](double v) { return quant ? bf16_round(v) : v; };
  const int L = hidden;
  std::vector<size_t> woff(L + 1), boff(L + 1);
  size_t off = 0;
  for (int l = 0; l <= L; ++l) {
    const int fi = l == 0 ? in_dim : width, fo = l == L ? 1 : width;
    woff[l] = off;
    off += size_t(fo) * fi;
    boff[l] = off;
    off += fo;
  }
  // rounded hidden weights
  std::vector<std::vector<double>> Wq(L);
  for (int l = 0; l < L; ++l) {
    const int fi = l == 0 ? in_dim : width;
    Wq[l].resize(size_t(width) * fi);
    for (size_t k = 0; k < Wq[l].size(); ++k) Wq[l][k] = q(params[woff[l] + k]);
  }
  // forward: h[0] = rounded input, z[l], h[l+1] = rounded gelu(z[l])
  std::vector<std::vector<double>> h(L + 1), z(L);
  h[0].resize(size_t(n) * in_dim);
  for (int m = 0; m < n; ++m)
    for (int c = 0; c < in_dim; ++c) {
      double v;
      if (trig) {
        const double a = double(x[size_t(m) * nraw + c / 2]);
        v = (c & 1) ? std::sin(a) : std::cos(a);
      } else {
        v = double(x[size_t(m) * nraw + c]);
      }
      h[0][size_t(m) * in_dim + c] = q(v);
    }
  for (int l = 0; l < L; ++l) {
    const int fi = l == 0 ? in_dim : width;
    z[l].assign(size_t(n) * width, 0.0);
    h[l + 1].assign(size_t(n) * width, 0.0);
    for (int m = 0; m < n; ++m)
      for (int o = 0; o < width; ++o) {
        double s = params[boff[l] + o];
        for (int i = 0; i < fi; ++i) s += h[l][size_t(m) * fi + i] * Wq[l][size_t(o) * fi + i];
        z[l][size_t(m) * width + o] = s;
        h[l + 1][size_t(m) * width + o] = q(gelu(s));
      }
  }
  const double* wout = params + woff[L];
  const double bout = params[boff[L]];
  double lsum = 0.0;
  std::vector<double> dout(n);
  for (int m = 0; m < n; ++m) {
    double s = bout;
    for (int i = 0; i < width; ++i) s += h[L][size_t(m) * width + i] * wout[i];
    out[m] = s;
    const double d = s - double(y[m]);
    lsum += d * d;
    dout[m] = 2.0 * d / n;
  }
  *loss = lsum / n;
  if (!grads) return 0;
  std::memset(grads, 0, off * sizeof(double));
  // output layer
  for (int m = 0; m < n; ++m) {
    for (int i = 0; i < width; ++i) grads[woff[L] + i] += h[L][size_t(m) * width + i] * dout[m];
    grads[boff[L]] += dout[m];
  }
  // dz for the last hidden layer
  std::vector<double> dz(size_t(n) * width), dzq(size_t(n) * width);
  for (int m = 0; m < n; ++m)
    for (int i = 0; i < width; ++i) {
      const size_t e = size_t(m) * width + i;
      dz[e] = dout[m] * wout[i] * gelu_d(z[L - 1][e]);
    }
  for (int l = L - 1; l >= 0; --l) {
    const int fi = l == 0 ? in_dim : width;
    for (size_t e = 0; e < dz.size(); ++e) dzq[e] = q(dz[e]);
    for (int m = 0; m < n; ++m)
      for (int o = 0; o < width; ++o) {
        const double g = dzq[size_t(m) * width + o];
        grads[boff[l] + o] += dz[size_t(m) * width + o];
        for (int i = 0; i < fi; ++i) grads[woff[l] + size_t(o) * fi + i] += g * h[l][size_t(m) * fi + i];
      }
    if (l > 0) {
      std::vector<double> nd(size_t(n) * width, 0.0);
      for (int m = 0; m < n; ++m)
        for (int i = 0; i < fi; ++i) {
          double s = 0.0;
          for (int o = 0; o < width; ++o) s += dzq[size_t(m) * width + o] * Wq[l][size_t(o) * fi + i];
          nd[size_t(m) * fi + i] = s * gelu_d(z[l - 1][size_t(m) * fi + i]);
        }
      dz.swap(nd);
    }
  }
  return 0;
}

}  // extern "C"
