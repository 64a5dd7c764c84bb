// Host side of the far-field model: creation from a template, training on the GPU with the
// FP32 kernels of surrogate.hip (forward, backward, Adam), parameter transfer, and the
// standalone gradient and inference entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "sf_far.hpp"

using namespace sf;

#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));                       \
      return SF_E_HIP;                                                                 \
    }                                                                                  \
  } while (0)

struct sf_far_model {
  FarSpec spec{};
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  float* params = nullptr;   // [kFParams] FP32 master weights
  float* grad = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  FarSpec* d_spec = nullptr;
  unsigned short* wpack = nullptr;  // bf16 inference weights
  float* bpack = nullptr;
  // activations of one batch
  int cap = 0;
  float *x = nullptr, *y = nullptr, *z[4] = {}, *a[3] = {}, *dy = nullptr, *da = nullptr, *db = nullptr;
  double* loss_part = nullptr;

  template <typename T>
  int alloc(T** p, size_t n) {
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    allocs.push_back(q);
    *p = (T*)q;
    return SF_OK;
  }
  void release(void* p) {
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it != allocs.end()) { allocs.erase(it); (void)hipFree(p); }
  }
  ~sf_far_model() {
    for (void* p : allocs) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }

  int upload_spec() {
    HIPCHK(hipMemcpyAsync(d_spec, &spec, sizeof(FarSpec), hipMemcpyHostToDevice, st));
    return SF_OK;
  }
  int pack() { return launch_pack_weights(params, wpack, bpack, st); }

  // batch buffers for up to n rows
  int ensure(int n) {
    if (n <= cap) return SF_OK;
    for (float* p : {x, y, dy, da, db}) release(p);
    for (auto* p : z) release(p);
    for (auto* p : a) release(p);
    release(loss_part);
    if (alloc(&x, (size_t)n * 16) || alloc(&y, (size_t)n * kFOut) || alloc(&dy, (size_t)n * kFOut) ||
        alloc(&da, (size_t)n * kFHid) || alloc(&db, (size_t)n * kFHid) || alloc(&loss_part, (size_t)(n * kFOut + 255) / 256))
      return SF_E_HIP;
    for (int l = 0; l < 4; l++)
      if (alloc(&z[l], (size_t)n * (l == 3 ? kFOut : kFHid))) return SF_E_HIP;
    for (int l = 0; l < 3; l++)
      if (alloc(&a[l], (size_t)n * kFHid)) return SF_E_HIP;
    cap = n;
    return SF_OK;
  }

  // FP32 forward of n rows in x ([n][16], 13 used): z[3] = network output [n][6]
  int forward(int n) {
    const float* in = x;
    int ld = 16, kin = kFIn;
    const int wofs[4] = {kFW1, kFW2, kFW3, kFW4}, bofs[4] = {kFB1, kFB2, kFB3, kFB4};
    for (int l = 0; l < 4; l++) {
      const int nout = l == 3 ? kFOut : kFHid;
      if (int rc = launch_gemm(0, 1, n, nout, kin, in, ld, params + wofs[l], kin, z[l], nout, st)) return rc;
      float* act = l == 3 ? z[3] : a[l];
      if (int rc = launch_bias_act(z[l], params + bofs[l], act, n, nout, l < 3 ? 1 : 0, st)) return rc;
      in = act; ld = nout; kin = nout;
    }
    return SF_OK;
  }

  // MSE against y ([n][6]) and gradients of every parameter into grad; partial loss sums in
  // loss_part.
  int backward(int n) {
    if (int rc = launch_mse(z[3], y, n, kFOut, dy, loss_part, st)) return rc;
    const int wofs[4] = {kFW1, kFW2, kFW3, kFW4}, bofs[4] = {kFB1, kFB2, kFB3, kFB4};
    float* dz = dy;  // gradient of the current layer's pre-activation
    for (int l = 3; l >= 0; l--) {
      const int nout = l == 3 ? kFOut : kFHid, nin = l == 0 ? kFIn : kFHid;
      const float* in = l == 0 ? x : a[l - 1];
      const int ldin = l == 0 ? 16 : kFHid;
      // dW = dz^T in, db = column sums of dz
      if (int rc = launch_gemm(1, 0, nout, nin, n, dz, nout, in, ldin, grad + wofs[l], nin, st)) return rc;
      if (int rc = launch_colsum(dz, n, nout, grad + bofs[l], st)) return rc;
      if (l == 0) break;
      // d(in) = dz W, then through the SiLU of the previous layer
      float* dnext = (dz == da) ? db : da;
      if (int rc = launch_gemm(0, 0, n, nin, nout, dz, nout, params + wofs[l], nin, dnext, nin, st)) return rc;
      if (int rc = launch_silu_back(dnext, z[l - 1], n * nin, st)) return rc;
      dz = dnext;
    }
    return SF_OK;
  }

  int loss_sum(int n, double* out) {
    const int nb = (n * kFOut + 255) / 256;
    std::vector<double> h(nb);
    HIPCHK(hipMemcpyAsync(h.data(), loss_part, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double s = 0;
    for (double v : h) s += v;
    *out = s;
    return SF_OK;
  }
};

namespace sf {
void far_model_device(const sf_far_model* m, FarSpec* spec, const unsigned short** wpack, const float** bias) {
  *spec = m->spec;
  *wpack = m->wpack;
  *bias = m->bpack;
}
}  // namespace sf

namespace {

int upload_rows(sf_far_model* m, const float* x, int n, int in_cols, float* dst, int dst_cols) {
  std::vector<float> h((size_t)n * dst_cols, 0.0f);
  for (int r = 0; r < n; r++)
    for (int c = 0; c < in_cols; c++) h[(size_t)r * dst_cols + c] = x[(size_t)r * in_cols + c];
  HIPCHK(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, m->st));
  HIPCHK(hipStreamSynchronize(m->st));
  return SF_OK;
}

}  // namespace

extern "C" {

sf_far_model* sf_far_model_create(const sf_template* t, double far_rf, double c_lo_mM, double c_hi_mM, uint64_t seed) {
  if (!t) { set_error("null template"); return nullptr; }
  const int nq = t->nq();
  if (nq < 1 || nq > kFMaxQ) { set_error("the far-field model needs 1..8 charged beads per template"); return nullptr; }
  if (!t->frame_ok) { set_error("template beads are collinear: no body frame"); return nullptr; }
  if (!(c_lo_mM > 0) || !(c_hi_mM > c_lo_mM)) { set_error("need 0 < c_lo < c_hi"); return nullptr; }
  sf_far_model* m = new sf_far_model();
  FarSpec& f = m->spec;
  f.nq = nq;
  for (int k = 0; k < nq; k++) {
    const int l = t->qlocal[k];
    for (int a = 0; a < 3; a++) f.qbody[k][a] = t->qbody[3 * k + a];
    f.q[k] = t->q[l];
    f.sig[k] = t->sigma[l];
  }
  const double sig = t->si_sigma;
  f.lb = 0.701e-9 / sig;  // src/md.cpp:165
  f.rq = t->r_q > 0 ? t->r_q : 1.0;
  f.rf = far_rf > 0 ? far_rf : 2.5 * t->diam;
  f.kcoef = std::sqrt(8 * 3.14159 * 6.022e23 * sig * sig * sig * f.lb);  // src/md.cpp:166-169
  f.kmin = f.kcoef * std::sqrt(c_lo_mM);
  f.kmax = f.kcoef * std::sqrt(c_hi_mM);
  for (int k = 0; k < kFOut; k++) { f.ymu[k] = 0.0f; f.ysd[k] = 1.0f; }
  // rigid torque correction of the per-bead spread: M = sum_a (|r|^2 I - r r^T) about the
  // charged centroid, inverted once
  for (int a = 0; a < 3; a++) {
    f.qcen[a] = 0;
    for (int k = 0; k < nq; k++) f.qcen[a] += f.qbody[k][a] / nq;
  }
  double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < nq; k++) {
    const double r[3] = {f.qbody[k][0] - f.qcen[0], f.qbody[k][1] - f.qcen[1], f.qbody[k][2] - f.qcen[2]};
    const double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) M[3 * i + j] += (i == j ? rr : 0.0) - r[i] * r[j];
  }
  const double c00 = M[4] * M[8] - M[5] * M[7], c01 = M[5] * M[6] - M[3] * M[8], c02 = M[3] * M[7] - M[4] * M[6];
  const double det = M[0] * c00 + M[1] * c01 + M[2] * c02;
  const double tr = M[0] + M[4] + M[8];
  if (!(std::fabs(det) > 1e-12 * tr * tr * tr)) { set_error("charged beads are collinear: no torque spread"); delete m; return nullptr; }
  const double adj[9] = {c00, M[2] * M[7] - M[1] * M[8], M[1] * M[5] - M[2] * M[4],
                         c01, M[0] * M[8] - M[2] * M[6], M[2] * M[3] - M[0] * M[5],
                         c02, M[1] * M[6] - M[0] * M[7], M[0] * M[4] - M[1] * M[3]};
  for (int k = 0; k < 9; k++) f.minv[k] = adj[k] / det;
  auto fail = [&](const char* w) -> sf_far_model* { if (w) set_error(w); delete m; return nullptr; };
  if (hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess) return fail("hipStreamCreate failed (no GPU?)");
  if (m->alloc(&m->params, kFParams) || m->alloc(&m->grad, kFParams) || m->alloc(&m->adam_m, kFParams) ||
      m->alloc(&m->adam_v, kFParams) || m->alloc(&m->d_spec, 1) || m->alloc(&m->wpack, kFPackN) ||
      m->alloc(&m->bpack, 3 * kFHid + kFOutPad))
    return fail(nullptr);
  // Glorot-uniform weights, zero biases
  std::vector<float> p(kFParams, 0.0f);
  std::mt19937_64 rng(seed);
  auto init = [&](int ofs, int nout, int nin) {
    const double a = std::sqrt(6.0 / (nin + nout));
    std::uniform_real_distribution<double> u(-a, a);
    for (int i = 0; i < nout * nin; i++) p[ofs + i] = (float)u(rng);
  };
  init(kFW1, kFHid, kFIn);
  init(kFW2, kFHid, kFHid);
  init(kFW3, kFHid, kFHid);
  init(kFW4, kFOut, kFHid);
  if (hipMemcpy(m->params, p.data(), kFParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->adam_m, 0, kFParams * sizeof(float)) != hipSuccess ||
      hipMemset(m->adam_v, 0, kFParams * sizeof(float)) != hipSuccess)
    return fail("hipMemcpy failed");
  if (m->upload_spec() || m->pack() || hipStreamSynchronize(m->st) != hipSuccess) return fail(nullptr);
  return m;
}

void sf_far_model_destroy(sf_far_model* m) { delete m; }

int sf_far_model_info(const sf_far_model* m, double* far_rf, double* kappa_min, double* kappa_max, int* n_params) {
  if (!m) { set_error("null model"); return SF_E_ARG; }
  if (far_rf) *far_rf = m->spec.rf;
  if (kappa_min) *kappa_min = m->spec.kmin;
  if (kappa_max) *kappa_max = m->spec.kmax;
  if (n_params) *n_params = kFParams;
  return SF_OK;
}

int sf_far_model_params(sf_far_model* m, float* get, const float* set, float* norm_get, const float* norm_set) {
  if (!m) { set_error("null model"); return SF_E_ARG; }
  if (get) {
    HIPCHK(hipMemcpyAsync(get, m->params, kFParams * sizeof(float), hipMemcpyDeviceToHost, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
  }
  if (norm_get)
    for (int k = 0; k < kFOut; k++) { norm_get[k] = m->spec.ymu[k]; norm_get[kFOut + k] = m->spec.ysd[k]; }
  if (set) HIPCHK(hipMemcpyAsync(m->params, set, kFParams * sizeof(float), hipMemcpyHostToDevice, m->st));
  if (norm_set) {
    for (int k = 0; k < kFOut; k++) { m->spec.ymu[k] = norm_set[k]; m->spec.ysd[k] = norm_set[kFOut + k]; }
    if (int rc = m->upload_spec()) return rc;
  }
  if (set || norm_set) {
    if (int rc = m->pack()) return rc;
    HIPCHK(hipStreamSynchronize(m->st));
  }
  return SF_OK;
}

int sf_far_model_samples(sf_far_model* m, int n, uint64_t seed, double c_lo_mM, double c_hi_mM, float* x, float* y,
                         double* pose) {
  if (!m || n <= 0 || !(c_lo_mM > 0) || !(c_hi_mM >= c_lo_mM)) { set_error("bad arguments"); return SF_E_ARG; }
  float *dx, *dy;
  double* dp = nullptr;
  if (m->alloc(&dx, (size_t)n * 16) || m->alloc(&dy, (size_t)n * 8) || (pose && m->alloc(&dp, (size_t)n * 16)))
    return SF_E_HIP;
  int rc = launch_far_samples(m->d_spec, n, seed, c_lo_mM, std::max(c_hi_mM, c_lo_mM * (1 + 1e-12)), dx, dy, dp, m->st);
  std::vector<float> hx((size_t)n * 16), hy((size_t)n * 8);
  std::vector<double> hp(pose ? (size_t)n * 16 : 0);
  if (!rc) {
    if (hipMemcpyAsync(hx.data(), dx, hx.size() * 4, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
        hipMemcpyAsync(hy.data(), dy, hy.size() * 4, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
        (pose && hipMemcpyAsync(hp.data(), dp, hp.size() * 8, hipMemcpyDeviceToHost, m->st) != hipSuccess) ||
        hipStreamSynchronize(m->st) != hipSuccess) {
      set_error("sample download failed");
      rc = SF_E_HIP;
    }
  }
  m->release(dx); m->release(dy); if (dp) m->release(dp);
  if (rc) return rc;
  for (int r = 0; r < n; r++) {
    if (x) for (int k = 0; k < kFIn; k++) x[(size_t)r * kFIn + k] = hx[(size_t)r * 16 + k];
    if (y) for (int k = 0; k < kFOut; k++) y[(size_t)r * kFOut + k] = hy[(size_t)r * 8 + k];
    if (pose) for (int k = 0; k < 16; k++) pose[(size_t)r * 16 + k] = hp[(size_t)r * 16 + k];
  }
  return SF_OK;
}

int sf_far_model_grad(sf_far_model* m, int n, const float* x, const float* y, float* grad, double* loss) {
  if (!m || n <= 0 || !x || !y) { set_error("bad arguments"); return SF_E_ARG; }
  if (int rc = m->ensure(n)) return rc;
  if (int rc = upload_rows(m, x, n, kFIn, m->x, 16)) return rc;
  if (int rc = upload_rows(m, y, n, kFOut, m->y, kFOut)) return rc;
  if (int rc = m->forward(n)) return rc;
  if (int rc = m->backward(n)) return rc;
  double s = 0;
  if (int rc = m->loss_sum(n, &s)) return rc;
  if (loss) *loss = s / ((double)n * kFOut);
  if (grad) {
    HIPCHK(hipMemcpyAsync(grad, m->grad, kFParams * sizeof(float), hipMemcpyDeviceToHost, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
  }
  return SF_OK;
}

int sf_far_model_infer(sf_far_model* m, int n, const float* x, float* y, int bf16_mfma) {
  if (!m || n <= 0 || !x || !y) { set_error("bad arguments"); return SF_E_ARG; }
  if (bf16_mfma) {
    float *dx, *dy;
    if (m->alloc(&dx, (size_t)n * kFIn) || m->alloc(&dy, (size_t)n * kFOut)) return SF_E_HIP;
    int rc = SF_OK;
    if (hipMemcpyAsync(dx, x, (size_t)n * kFIn * 4, hipMemcpyHostToDevice, m->st) != hipSuccess) { set_error("copy failed"); rc = SF_E_HIP; }
    if (!rc) rc = launch_mlp_rows(m->wpack, m->bpack, n, dx, dy, m->st);
    if (!rc && (hipMemcpyAsync(y, dy, (size_t)n * kFOut * 4, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
                hipStreamSynchronize(m->st) != hipSuccess)) { set_error("copy failed"); rc = SF_E_HIP; }
    m->release(dx); m->release(dy);
    return rc;
  }
  if (int rc = m->ensure(n)) return rc;
  if (int rc = upload_rows(m, x, n, kFIn, m->x, 16)) return rc;
  if (int rc = m->forward(n)) return rc;
  HIPCHK(hipMemcpyAsync(y, m->z[3], (size_t)n * kFOut * 4, hipMemcpyDeviceToHost, m->st));
  HIPCHK(hipStreamSynchronize(m->st));
  return SF_OK;
}

int sf_far_model_train(sf_far_model* m, const sf_train_opts* o, sf_train_report* rep) {
  if (!m) { set_error("null model"); return SF_E_ARG; }
  sf_train_opts opt = {1000000, 20, 4096, 2e-3, 0.05, 12345ull, 0.0, 0.0};
  if (o) {
    if (o->n_samples > 0) opt.n_samples = o->n_samples;
    if (o->epochs > 0) opt.epochs = o->epochs;
    if (o->batch > 0) opt.batch = o->batch;
    if (o->lr > 0) opt.lr = o->lr;
    if (o->val_frac > 0) opt.val_frac = o->val_frac;
    opt.seed = o->seed;
    opt.c_lo_mM = o->c_lo_mM;
    opt.c_hi_mM = o->c_hi_mM;
  }
  const double c_lo = opt.c_lo_mM > 0 ? opt.c_lo_mM : (m->spec.kmin / m->spec.kcoef) * (m->spec.kmin / m->spec.kcoef);
  const double c_hi = opt.c_hi_mM > 0 ? opt.c_hi_mM : (m->spec.kmax / m->spec.kcoef) * (m->spec.kmax / m->spec.kcoef);
  const int n = opt.n_samples, B = std::min(opt.batch, n);
  const int n_val = std::max(1, (int)(opt.val_frac * n));
  const int n_tr = n - n_val;
  if (n_tr < B) { set_error("training set smaller than one batch"); return SF_E_ARG; }
  auto t0 = std::chrono::steady_clock::now();
  float *X, *Yraw, *Ys;
  int* perm;
  if (m->alloc(&X, (size_t)n * 16) || m->alloc(&Yraw, (size_t)n * 8) || m->alloc(&Ys, (size_t)n * kFOut) ||
      m->alloc(&perm, (size_t)std::max(n_tr, n_val)))
    return SF_E_HIP;
  auto cleanup = [&]() { m->release(X); m->release(Yraw); m->release(Ys); m->release(perm); };
  int rc = launch_far_samples(m->d_spec, n, opt.seed, c_lo, c_hi, X, Yraw, nullptr, m->st);
  // standardisation from the training rows
  std::vector<float> hy((size_t)n * 8);
  if (!rc && (hipMemcpyAsync(hy.data(), Yraw, hy.size() * 4, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
              hipStreamSynchronize(m->st) != hipSuccess)) { set_error("copy failed"); rc = SF_E_HIP; }
  if (rc) { cleanup(); return rc; }
  double mu[kFOut] = {0}, var[kFOut] = {0};
  for (int r = 0; r < n_tr; r++)
    for (int k = 0; k < kFOut; k++) mu[k] += hy[(size_t)r * 8 + k];
  for (int k = 0; k < kFOut; k++) mu[k] /= n_tr;
  for (int r = 0; r < n_tr; r++)
    for (int k = 0; k < kFOut; k++) { const double e = hy[(size_t)r * 8 + k] - mu[k]; var[k] += e * e; }
  for (int k = 0; k < kFOut; k++) {
    m->spec.ymu[k] = (float)mu[k];
    m->spec.ysd[k] = (float)std::max(std::sqrt(var[k] / n_tr), 1e-12);
  }
  std::vector<float> ys((size_t)n * kFOut);
  for (int r = 0; r < n; r++)
    for (int k = 0; k < kFOut; k++) ys[(size_t)r * kFOut + k] = (hy[(size_t)r * 8 + k] - m->spec.ymu[k]) / m->spec.ysd[k];
  if (hipMemcpyAsync(Ys, ys.data(), ys.size() * 4, hipMemcpyHostToDevice, m->st) != hipSuccess) { cleanup(); set_error("copy failed"); return SF_E_HIP; }
  if ((rc = m->upload_spec()) || (rc = m->ensure(B))) { cleanup(); return rc; }
  (void)hipMemsetAsync(m->adam_m, 0, kFParams * sizeof(float), m->st);
  (void)hipMemsetAsync(m->adam_v, 0, kFParams * sizeof(float), m->st);
  const int steps_per_epoch = n_tr / B;
  const long total = (long)steps_per_epoch * opt.epochs;
  long step = 0;
  std::vector<int> idx(n_tr);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < opt.epochs && !rc; ep++) {
    std::iota(idx.begin(), idx.end(), 0);
    std::mt19937_64 g(opt.seed * 1000003ull + (uint64_t)ep);
    std::shuffle(idx.begin(), idx.end(), g);
    if (hipMemcpyAsync(perm, idx.data(), (size_t)n_tr * 4, hipMemcpyHostToDevice, m->st) != hipSuccess ||
        hipStreamSynchronize(m->st) != hipSuccess) { set_error("copy failed"); rc = SF_E_HIP; break; }
    for (int s = 0; s < steps_per_epoch && !rc; s++, step++) {
      const int* pi = perm + (size_t)s * B;
      if ((rc = launch_gather_rows(X, pi, m->x, B, 16, m->st)) || (rc = launch_gather_rows(Ys, pi, m->y, B, kFOut, m->st)) ||
          (rc = m->forward(B)) || (rc = m->backward(B)))
        break;
      // cosine schedule from lr down to lr/50
      const double fr = (double)step / std::max<long>(total, 1);
      const float lr = (float)(opt.lr * (0.02 + 0.98 * 0.5 * (1 + std::cos(3.141592653589793 * fr))));
      const double t = (double)(step + 1);
      rc = launch_adam(m->params, m->grad, m->adam_m, m->adam_v, kFParams, lr, b1, b2, eps,
                       (float)(1.0 / (1.0 - std::pow(b1, t))), (float)(1.0 / (1.0 - std::pow(b2, t))), m->st);
    }
  }
  // losses and relative errors on held-out and training rows (FP32 forward)
  auto evaluate = [&](int r0, int cnt, double* loss, double* ef, double* et) -> int {
    double se = 0, fe = 0, fn = 0, te = 0, tn = 0;
    std::vector<int> ids(std::min(cnt, B));
    std::vector<float> yh((size_t)B * kFOut);
    for (int c0 = 0; c0 < cnt; c0 += B) {
      const int k = std::min(B, cnt - c0);
      for (int i = 0; i < k; i++) ids[i] = r0 + c0 + i;
      HIPCHK(hipMemcpyAsync(perm, ids.data(), k * 4, hipMemcpyHostToDevice, m->st));
      if (int r = launch_gather_rows(X, perm, m->x, k, 16, m->st)) return r;
      if (int r = m->forward(k)) return r;
      HIPCHK(hipMemcpyAsync(yh.data(), m->z[3], (size_t)k * kFOut * 4, hipMemcpyDeviceToHost, m->st));
      HIPCHK(hipStreamSynchronize(m->st));
      for (int i = 0; i < k; i++)
        for (int c = 0; c < kFOut; c++) {
          const double yt = ys[(size_t)(r0 + c0 + i) * kFOut + c], yp = yh[(size_t)i * kFOut + c];
          se += (yp - yt) * (yp - yt);
          const double pt = m->spec.ymu[c] + m->spec.ysd[c] * yt, pp = m->spec.ymu[c] + m->spec.ysd[c] * yp;
          if (c < 3) { fe += (pp - pt) * (pp - pt); fn += pt * pt; } else { te += (pp - pt) * (pp - pt); tn += pt * pt; }
        }
    }
    *loss = se / ((double)cnt * kFOut);
    *ef = std::sqrt(fe / std::max(fn, 1e-300));
    *et = std::sqrt(te / std::max(tn, 1e-300));
    return SF_OK;
  };
  double vl = 0, vf = 0, vt = 0, tl = 0, tf = 0, tt = 0;
  if (!rc) rc = evaluate(n_tr, n_val, &vl, &vf, &vt);
  if (!rc) rc = evaluate(0, std::min(n_tr, 50000), &tl, &tf, &tt);
  if (!rc) rc = m->pack();
  if (!rc && hipStreamSynchronize(m->st) != hipSuccess) { set_error("sync failed"); rc = SF_E_HIP; }
  cleanup();
  if (rc) return rc;
  if (rep) {
    rep->train_loss = tl; rep->val_loss = vl;
    rep->val_rel_err_force = vf; rep->val_rel_err_torque = vt;
    rep->train_rel_err_force = tf;
    rep->steps = (int)step;
    rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return SF_OK;
}

}  // extern "C"
