// Far-field model of the uncut Yukawa between two subunits: shared types of the host code and
// the HIP kernels (surrogate.hip, surrogate.cpp, capi.cpp).
#pragma once

#include "sf_internal.hpp"

namespace sf {

// Network: 13 inputs -> 128 -> 128 -> 128 (SiLU) -> 6 outputs (linear).
constexpr int kFIn = 13;     // 12 pose features + kappa
constexpr int kFInPad = 32;  // one MFMA k-step
constexpr int kFHid = 128;
constexpr int kFOut = 6;     // force (3) and torque (3) on subunit A, in A's body frame
constexpr int kFOutPad = 16;
constexpr int kFMaxQ = 8;    // charged beads per template

// Packed FP32 parameters, in this order: W1 [128][13], b1 [128], W2 [128][128], b2 [128],
// W3 [128][128], b3 [128], W4 [6][128], b4 [6].
constexpr int kFW1 = 0, kFB1 = kFW1 + kFHid * kFIn, kFW2 = kFB1 + kFHid, kFB2 = kFW2 + kFHid * kFHid,
              kFW3 = kFB2 + kFHid, kFB3 = kFW3 + kFHid * kFHid, kFW4 = kFB3 + kFHid,
              kFB4 = kFW4 + kFOut * kFHid, kFParams = kFB4 + kFOut;

// bf16 weights for inference, row-major [out][in] with padded rows (LDS bank spread):
// W1 [128][kFInPad + 8], W2 and W3 [128][136], W4 [16][136] (rows 6..15 zero).
constexpr int kFLd1 = kFInPad + 8, kFLdH = kFHid + 8;
constexpr int kFPackW1 = 0, kFPackW2 = kFPackW1 + kFHid * kFLd1, kFPackW3 = kFPackW2 + kFHid * kFLdH,
              kFPackW4 = kFPackW3 + kFHid * kFLdH, kFPackN = kFPackW4 + kFOutPad * kFLdH;

// Template constants and output scaling of a far-field model.
struct FarSpec {
  int nq;
  double qbody[kFMaxQ][3];   // charged beads in the body frame (relative to the centre bead)
  double q[kFMaxQ];          // charges
  double sig[kFMaxQ];        // diameters
  double lb;                 // Bjerrum length
  double rq;                 // largest charged-bead distance from the centre bead
  double rf;                 // split radius (centre-bead distance) the model covers beyond
  double kmin, kmax;         // kappa range of the training set
  double kcoef;              // kappa = kcoef sqrt(c_mM)
  float ymu[kFOut], ysd[kFOut];  // standardisation of the scaled targets
  double qcen[3];            // centroid of qbody
  double minv[9];            // inverse of sum_a (|r_a|^2 I - r_a r_a^T), r_a = qbody_a - qcen
};

// Output scale of a pair: the magnitude of the pair force lb (kappa + 1/r) exp(-kappa r) / r
// in the leading (parallel-projection) form of sum_ab exp(-kappa r_ab),
//   S = lb (kappa + 1/d) / d exp(-kappa d) sum_a exp(kappa p_a) sum_b exp(kappa p'_b),
// p_a = qbody_a . n (A's charges towards B), p'_b = -(Q qbody_b) . n (B's towards A), n = d/|d|
// in A's body frame. Targets are (F / S, tau / (S rq)). The sums are taken as shifted
// log-sum-exps so that S underflows only where the pair force itself does.
__host__ __device__ inline double far_scale(const FarSpec& f, const double* dv, const double* Q, double kappa) {
  const double d = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
  const double n[3] = {dv[0] / d, dv[1] / d, dv[2] / d};
  double pa[kFMaxQ], pb[kFMaxQ], ma = -1e300, mb = -1e300;
  for (int a = 0; a < f.nq; a++) {
    const double* u = f.qbody[a];
    pa[a] = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
    pb[a] = -((Q[0] * u[0] + Q[1] * u[1] + Q[2] * u[2]) * n[0] + (Q[3] * u[0] + Q[4] * u[1] + Q[5] * u[2]) * n[1] +
              (Q[6] * u[0] + Q[7] * u[1] + Q[8] * u[2]) * n[2]);
    ma = pa[a] > ma ? pa[a] : ma;
    mb = pb[a] > mb ? pb[a] : mb;
  }
  double sa = 0, sb = 0;
  for (int a = 0; a < f.nq; a++) { sa += exp(kappa * (pa[a] - ma)); sb += exp(kappa * (pb[a] - mb)); }
  return f.lb * (kappa + 1.0 / d) / d * exp(-kappa * (d - ma - mb) + log(sa) + log(sb));
}

// Device view of a model attached to an engine handle.
struct FarDev {
  const FarSpec* spec;           // device copy
  const unsigned short* wpack;   // bf16 weights [kFPackN]
  const float* bias;             // b1 b2 b3 (128 each), b4 (6, padded to 16)
  float* frames;                 // [Ns][12]: e1 e2 e3 (lab) of each subunit, FP32
};

// device pointers of a model's spec-independent state (surrogate.cpp)
void far_model_device(const sf_far_model* m, FarSpec* spec, const unsigned short** wpack, const float** bias);

// launchers (surrogate.hip)
int launch_far_samples(const FarSpec* spec_dev, int n, uint64_t seed, double c_lo, double c_hi, float* x,
                       float* y, double* pose, hipStream_t st);
int launch_gemm(int ta, int tb, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
                int ldc, hipStream_t st);
int launch_bias_act(float* Z, const float* b, float* A, int rows, int cols, int act, hipStream_t st);
int launch_silu_back(float* dA, const float* Z, int n, hipStream_t st);
int launch_colsum(const float* M, int rows, int cols, float* out, hipStream_t st);
int launch_mse(const float* yhat, const float* y, int rows, int cols, float* dy, double* loss_part, hipStream_t st);
int launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                float c1, float c2, hipStream_t st);
int launch_gather_rows(const float* src, const int* idx, float* dst, int rows, int cols, hipStream_t st);
int launch_pack_weights(const float* params, unsigned short* wpack, float* bias, hipStream_t st);
int launch_mlp_rows(const unsigned short* wpack, const float* bias, int n, const float* x, float* y, hipStream_t st);
int launch_frames(const Dev& d, const FarDev& f, hipStream_t st);
int launch_far_mlp(const Dev& d, const FarDev& f, hipStream_t st);
// per replica r: out2[3 r] = sum |F_far,model - F_far,exact|^2, out2[3 r + 1] = sum |F_far,exact|^2,
// out2[3 r + 2] = sum over its beads of |F|^2 with the exact far field
int launch_far_audit(const Dev& d, const double* qfar_exact, double* out2, hipStream_t st);

}  // namespace sf
