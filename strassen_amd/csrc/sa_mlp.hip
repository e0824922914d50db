// sa_mlp.hip -- the learned path selector's network: a small FP64 MLP with
// hand-written inference, backward and Adam kernels.
//
// y = W3 tanh(W2 tanh(W1 x' + b1) + b2) + b3,  x' = (x - mu) / sd
// with F = 10 node features, two hidden layers of 32 and O = 2 outputs: the
// predicted log2 cost ratios of the Strassen-Winograd and dense paths of a
// recursion node against its screened-sparse cost.  Parameters are one flat
// array [W1 (32xF, row-major) | b1 | W2 (32x32) | b2 | W3 (Ox32) | b3].
//
// Training is full-batch: per-sample kernels compute the output error and the
// back-propagated hidden-layer errors; per-parameter kernels sum the products
// over samples in sample order (deterministic, no atomics); Adam updates the
// parameters in place.  Loss: L = 1/(2N) sum_n sum_o m_no (y_no - t_no)^2.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "sa_internal.hpp"

namespace sa {

constexpr int kF = kMlpFeatures, kH = 32, kO = kMlpOutputs;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kH * kF, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kO * kH, kNumParams = kOffB3 + kO;
static_assert(kNumParams == kMlpParams, "parameter count");

// one thread per sample; also stores the activations for the backward pass
__global__ void k_mlp_fwd(const double* __restrict__ X, int64_t n, const double* __restrict__ prm,
                          const double* __restrict__ mu, const double* __restrict__ sd, double* __restrict__ Y,
                          double* __restrict__ XN, double* __restrict__ H1, double* __restrict__ H2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[kF], h1[kH], h2[kH];
#pragma unroll
  for (int f = 0; f < kF; ++f) x[f] = (X[i * kF + f] - mu[f]) / sd[f];
  for (int h = 0; h < kH; ++h) {
    double z = prm[kOffB1 + h];
#pragma unroll
    for (int f = 0; f < kF; ++f) z = fma(prm[kOffW1 + h * kF + f], x[f], z);
    h1[h] = tanh(z);
  }
  for (int h = 0; h < kH; ++h) {
    double z = prm[kOffB2 + h];
    for (int g = 0; g < kH; ++g) z = fma(prm[kOffW2 + h * kH + g], h1[g], z);
    h2[h] = tanh(z);
  }
#pragma unroll
  for (int o = 0; o < kO; ++o) {
    double z = prm[kOffB3 + o];
    for (int g = 0; g < kH; ++g) z = fma(prm[kOffW3 + o * kH + g], h2[g], z);
    Y[i * kO + o] = z;
  }
  if (XN) {
#pragma unroll
    for (int f = 0; f < kF; ++f) XN[i * kF + f] = x[f];
    for (int h = 0; h < kH; ++h) {
      H1[i * kH + h] = h1[h];
      H2[i * kH + h] = h2[h];
    }
  }
}

// per-sample errors: DY = m (y - t) / N, D2 = (W3^T DY) (1 - h2^2), D1 = (W2^T D2) (1 - h1^2)
__global__ void k_mlp_bwd_samples(int64_t n, const double* __restrict__ prm, const double* __restrict__ Y,
                                  const double* __restrict__ T, const double* __restrict__ M,
                                  const double* __restrict__ H1, const double* __restrict__ H2,
                                  double* __restrict__ DY, double* __restrict__ D1, double* __restrict__ D2,
                                  double* __restrict__ LS) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double dy[kO], d2[kH];
  double l = 0.0;
#pragma unroll
  for (int o = 0; o < kO; ++o) {
    const double e = Y[i * kO + o] - T[i * kO + o];
    const double m = M[i * kO + o];
    l += m * e * e;
    dy[o] = m * e / (double)n;
    DY[i * kO + o] = dy[o];
  }
  LS[i] = l;
  for (int h = 0; h < kH; ++h) {
    double s = 0.0;
#pragma unroll
    for (int o = 0; o < kO; ++o) s = fma(prm[kOffW3 + o * kH + h], dy[o], s);
    const double a = H2[i * kH + h];
    d2[h] = s * (1.0 - a * a);
    D2[i * kH + h] = d2[h];
  }
  for (int g = 0; g < kH; ++g) {
    double s = 0.0;
    for (int h = 0; h < kH; ++h) s = fma(prm[kOffW2 + h * kH + g], d2[h], s);
    const double a = H1[i * kH + g];
    D1[i * kH + g] = s * (1.0 - a * a);
  }
}

// one thread per parameter: sum over samples in order
__global__ void k_mlp_bwd_params(int64_t n, const double* __restrict__ XN, const double* __restrict__ H1,
                                 const double* __restrict__ H2, const double* __restrict__ DY,
                                 const double* __restrict__ D1, const double* __restrict__ D2,
                                 double* __restrict__ G) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNumParams) return;
  const double *u = nullptr, *v = nullptr;  // grad = sum_i u[i*su + a] * (v ? v[i*sv + b] : 1)
  int su = 0, sv = 0, a = 0, b = 0;
  if (p < kOffB1) {
    u = D1; su = kH; a = (p - kOffW1) / kF; v = XN; sv = kF; b = (p - kOffW1) % kF;
  } else if (p < kOffW2) {
    u = D1; su = kH; a = p - kOffB1;
  } else if (p < kOffB2) {
    u = D2; su = kH; a = (p - kOffW2) / kH; v = H1; sv = kH; b = (p - kOffW2) % kH;
  } else if (p < kOffW3) {
    u = D2; su = kH; a = p - kOffB2;
  } else if (p < kOffB3) {
    u = DY; su = kO; a = (p - kOffW3) / kH; v = H2; sv = kH; b = (p - kOffW3) % kH;
  } else {
    u = DY; su = kO; a = p - kOffB3;
  }
  double g = 0.0;
  if (v) {
    for (int64_t i = 0; i < n; ++i) g = fma(u[i * su + a], v[i * sv + b], g);
  } else {
    for (int64_t i = 0; i < n; ++i) g += u[i * su + a];
  }
  G[p] = g;
}

__global__ void k_sum_ordered(const double* __restrict__ x, int64_t n, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += x[i];
  *out = s;
}

// Adam (Kingma & Ba 2015), step t >= 1
__global__ void k_adam(double* __restrict__ prm, const double* __restrict__ g, double* __restrict__ m1,
                       double* __restrict__ m2, int t, double lr, double beta1, double beta2, double eps) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNumParams) return;
  m1[p] = beta1 * m1[p] + (1.0 - beta1) * g[p];
  m2[p] = beta2 * m2[p] + (1.0 - beta2) * g[p] * g[p];
  const double mh = m1[p] / (1.0 - pow(beta1, (double)t));
  const double vh = m2[p] / (1.0 - pow(beta2, (double)t));
  prm[p] -= lr * mh / (sqrt(vh) + eps);
}

// ----------------------------------------------------------- model state --
// context buffers: "mlp_prm" = params | mu (F) | sd (F)
bool mlp_loaded(sa_context* c) { return c->mlp_loaded; }

static double* prm_of(sa_context* c) { return c->buf("mlp_prm").as<double>(); }

void mlp_forward_device(sa_context* c, const double* dX, int64_t n, double* dY) {
  SA_REQUIRE(c->mlp_loaded, "mlp: no model loaded");
  if (n == 0) return;
  const double* prm = prm_of(c);
  hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)ceil_div(n, 128)), dim3(128), 0, c->stream, dX, n, prm,
                     prm + kNumParams, prm + kNumParams + kF, dY, (double*)nullptr, (double*)nullptr,
                     (double*)nullptr);
  SA_CHECK_HIP(hipGetLastError());
}

static void set_model(sa_context* c, const double* prm, const double* mu, const double* sd) {
  DevBuf& b = c->buf("mlp_prm");
  b.reserve(sizeof(double) * (kNumParams + 2 * kF));
  std::vector<double> h(prm, prm + kNumParams);
  for (int f = 0; f < kF; ++f) h.push_back(mu ? mu[f] : 0.0);
  for (int f = 0; f < kF; ++f) {
    const double s = sd ? sd[f] : 1.0;
    SA_REQUIRE(s > 0.0, "mlp: feature scale %d must be positive", f);
    h.push_back(s);
  }
  SA_CHECK_HIP(hipMemcpyAsync(b.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  c->mlp_loaded = true;
  c->mlp_step = 0;
}

// forward + backward on device arrays; returns the loss (sync); grads in "mlp_grad"
static double grad_device(sa_context* c, const double* dX, const double* dT, const double* dM, int64_t n) {
  const double* prm = prm_of(c);
  DevBuf &bxn = c->buf("mlp_xn"), &bh1 = c->buf("mlp_h1"), &bh2 = c->buf("mlp_h2"), &by = c->buf("mlp_y");
  DevBuf &bdy = c->buf("mlp_dy"), &bd1 = c->buf("mlp_d1"), &bd2 = c->buf("mlp_d2"), &bls = c->buf("mlp_ls");
  DevBuf& bg = c->buf("mlp_grad");
  bxn.reserve(sizeof(double) * (size_t)(n * kF + 1));
  bh1.reserve(sizeof(double) * (size_t)(n * kH + 1));
  bh2.reserve(sizeof(double) * (size_t)(n * kH + 1));
  by.reserve(sizeof(double) * (size_t)(n * kO + 1));
  bdy.reserve(sizeof(double) * (size_t)(n * kO + 1));
  bd1.reserve(sizeof(double) * (size_t)(n * kH + 1));
  bd2.reserve(sizeof(double) * (size_t)(n * kH + 1));
  bls.reserve(sizeof(double) * (size_t)(n + 2));
  bg.reserve(sizeof(double) * kNumParams);
  const unsigned gs = (unsigned)ceil_div(n, 128);
  hipLaunchKernelGGL(k_mlp_fwd, dim3(gs), dim3(128), 0, c->stream, dX, n, prm, prm + kNumParams,
                     prm + kNumParams + kF, by.as<double>(), bxn.as<double>(), bh1.as<double>(), bh2.as<double>());
  hipLaunchKernelGGL(k_mlp_bwd_samples, dim3(gs), dim3(128), 0, c->stream, n, prm, (const double*)by.p, dT, dM,
                     (const double*)bh1.p, (const double*)bh2.p, bdy.as<double>(), bd1.as<double>(), bd2.as<double>(),
                     bls.as<double>());
  hipLaunchKernelGGL(k_mlp_bwd_params, dim3((unsigned)ceil_div(kNumParams, 64)), dim3(64), 0, c->stream, n,
                     (const double*)bxn.p, (const double*)bh1.p, (const double*)bh2.p, (const double*)bdy.p,
                     (const double*)bd1.p, (const double*)bd2.p, bg.as<double>());
  hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(1), 0, c->stream, (const double*)bls.p, n, bls.as<double>() + n);
  double loss = 0.0;
  SA_CHECK_HIP(hipMemcpyAsync(&loss, bls.as<double>() + n, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  SA_CHECK_HIP(hipGetLastError());
  return loss / (2.0 * (double)n);
}

struct HostArrays {
  double *x, *t, *m;
};

static HostArrays upload_xtm(sa_context* c, const double* X, const double* T, const double* M, int64_t n) {
  DevBuf &bx = c->buf("mlp_x"), &bt = c->buf("mlp_t"), &bm = c->buf("mlp_m");
  bx.reserve(sizeof(double) * (size_t)(n * kF + 1));
  bt.reserve(sizeof(double) * (size_t)(n * kO + 1));
  bm.reserve(sizeof(double) * (size_t)(n * kO + 1));
  SA_CHECK_HIP(hipMemcpyAsync(bx.p, X, sizeof(double) * (size_t)(n * kF), hipMemcpyHostToDevice, c->stream));
  SA_CHECK_HIP(hipMemcpyAsync(bt.p, T, sizeof(double) * (size_t)(n * kO), hipMemcpyHostToDevice, c->stream));
  if (M) {
    SA_CHECK_HIP(hipMemcpyAsync(bm.p, M, sizeof(double) * (size_t)(n * kO), hipMemcpyHostToDevice, c->stream));
  } else {
    fill_f64(c, bm.as<double>(), n * kO, 1.0);
  }
  return HostArrays{bx.as<double>(), bt.as<double>(), bm.as<double>()};
}

}  // namespace sa

using namespace sa;

extern "C" {

int sa_mlp_info(int32_t* n_features, int32_t* n_outputs, int64_t* n_params) {
  return guard([&] {
    if (n_features) *n_features = kF;
    if (n_outputs) *n_outputs = kO;
    if (n_params) *n_params = kNumParams;
  });
}

int sa_mlp_set(sa_ctx_t c, const double* params, int64_t n_params, const double* feat_mu, const double* feat_sd) {
  return guard([&] {
    SA_REQUIRE(c && params, "null argument");
    SA_REQUIRE(n_params == kNumParams, "mlp: %lld parameters given, the network has %d", (long long)n_params,
               kNumParams);
    SA_CHECK_HIP(hipSetDevice(c->device));
    set_model(c, params, feat_mu, feat_sd);
  });
}

int sa_mlp_get(sa_ctx_t c, double* params, double* feat_mu, double* feat_sd) {
  return guard([&] {
    SA_REQUIRE(c && params, "null argument");
    SA_REQUIRE(c->mlp_loaded, "mlp: no model loaded");
    std::vector<double> h(kNumParams + 2 * kF);
    SA_CHECK_HIP(hipMemcpyAsync(h.data(), c->buf("mlp_prm").p, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                                c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(params, h.data(), sizeof(double) * kNumParams);
    if (feat_mu) std::memcpy(feat_mu, h.data() + kNumParams, sizeof(double) * kF);
    if (feat_sd) std::memcpy(feat_sd, h.data() + kNumParams + kF, sizeof(double) * kF);
  });
}

// Text model file: "path_mlp <F> <H> <O> <P>" then P parameters, F means, F scales.
int sa_mlp_load(sa_ctx_t c, const char* path) {
  return guard([&] {
    SA_REQUIRE(c && path, "null argument");
    std::ifstream in(path);
    SA_REQUIRE(in.good(), "mlp: cannot open %s", path);
    std::string tag;
    int f = 0, h = 0, o = 0;
    int64_t p = 0;
    in >> tag >> f >> h >> o >> p;
    SA_REQUIRE(tag == "path_mlp" && f == kF && h == kH && o == kO && p == kNumParams,
               "mlp: %s is not a %dx%dx%dx%d path model", path, kF, kH, kH, kO);
    std::vector<double> v((size_t)(kNumParams + 2 * kF));
    for (auto& x : v) {
      in >> x;
      SA_REQUIRE(!in.fail(), "mlp: %s is truncated", path);
    }
    SA_CHECK_HIP(hipSetDevice(c->device));
    set_model(c, v.data(), v.data() + kNumParams, v.data() + kNumParams + kF);
  });
}

int sa_mlp_forward(sa_ctx_t c, const double* X, int64_t n, double* Y) {
  return guard([&] {
    SA_REQUIRE(c && (n == 0 || (X && Y)), "null argument");
    SA_CHECK_HIP(hipSetDevice(c->device));
    if (n == 0) return;
    DevBuf &bx = c->buf("mlp_x"), &by = c->buf("mlp_y");
    bx.reserve(sizeof(double) * (size_t)(n * kF + 1));
    by.reserve(sizeof(double) * (size_t)(n * kO + 1));
    SA_CHECK_HIP(hipMemcpyAsync(bx.p, X, sizeof(double) * (size_t)(n * kF), hipMemcpyHostToDevice, c->stream));
    mlp_forward_device(c, bx.as<double>(), n, by.as<double>());
    SA_CHECK_HIP(hipMemcpyAsync(Y, by.p, sizeof(double) * (size_t)(n * kO), hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int sa_mlp_grad(sa_ctx_t c, const double* X, const double* T, const double* M, int64_t n, double* grad,
                double* loss) {
  return guard([&] {
    SA_REQUIRE(c && X && T && n > 0, "null argument or empty batch");
    SA_REQUIRE(c->mlp_loaded, "mlp: no model loaded");
    SA_CHECK_HIP(hipSetDevice(c->device));
    const HostArrays d = upload_xtm(c, X, T, M, n);
    const double l = grad_device(c, d.x, d.t, d.m, n);
    if (loss) *loss = l;
    if (grad) {
      SA_CHECK_HIP(hipMemcpyAsync(grad, c->buf("mlp_grad").p, sizeof(double) * kNumParams, hipMemcpyDeviceToHost,
                                  c->stream));
      SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    }
  });
}

int sa_mlp_train(sa_ctx_t c, const double* X, const double* T, const double* M, int64_t n, int32_t epochs, double lr,
                 double* loss_out) {
  return guard([&] {
    SA_REQUIRE(c && X && T && n > 0 && epochs >= 0, "bad argument");
    SA_REQUIRE(c->mlp_loaded, "mlp: no model loaded (initialise with sa_mlp_set)");
    SA_CHECK_HIP(hipSetDevice(c->device));
    const HostArrays d = upload_xtm(c, X, T, M, n);
    DevBuf &b1 = c->buf("mlp_adam1"), &b2 = c->buf("mlp_adam2");
    if (c->mlp_step == 0) {
      b1.reserve(sizeof(double) * kNumParams);
      b2.reserve(sizeof(double) * kNumParams);
      fill_f64(c, b1.as<double>(), kNumParams, 0.0);
      fill_f64(c, b2.as<double>(), kNumParams, 0.0);
    }
    double l = 0.0;
    for (int32_t e = 0; e < epochs; ++e) {
      l = grad_device(c, d.x, d.t, d.m, n);
      ++c->mlp_step;
      hipLaunchKernelGGL(k_adam, dim3((unsigned)ceil_div(kNumParams, 128)), dim3(128), 0, c->stream, prm_of(c),
                         (const double*)c->buf("mlp_grad").p, b1.as<double>(), b2.as<double>(), c->mlp_step, lr, 0.9,
                         0.999, 1e-8);
    }
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    SA_CHECK_HIP(hipGetLastError());
    if (loss_out) *loss_out = epochs > 0 ? l : grad_device(c, d.x, d.t, d.m, n);
  });
}

}  // extern "C"
