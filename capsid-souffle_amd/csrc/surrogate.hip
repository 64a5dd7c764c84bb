// Far-field model of the uncut Yukawa between two subunits: training-set generation scored by
// the exact pair sum, FP32 training kernels (tiled GEMM, SiLU, MSE, Adam), and BF16 MFMA
// inference (gfx950 v_mfma_f32_16x16x32_bf16, FP32 accumulation), standalone and inside the
// engine's step.
//
// Pose of subunit B seen from subunit A: d = R_A^T (c_B - c_A) (A's body frame, centre beads,
// minimum image), Q = R_A^T R_B. Features (13): d/|d| (3), rf/|d|, (rf/|d|)^2, (rf/|d|)^3,
// the first two columns of Q (6), kappa. Outputs (6): the Yukawa force on A's charged beads
// and its torque about A's centre bead, in A's body frame, divided by the pair scale S of
// far_scale (torque also by rq) and standardised.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sf_far.hpp"

namespace sf {

namespace {

int check_launch_far() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("kernel launch failed: ") + hipGetErrorString(e));
    return SF_E_HIP;
  }
  return SF_OK;
}

__device__ __forceinline__ double min_img_f(double x, double box, double hbox) {
  return x - (x > hbox ? box : (x < -hbox ? -box : 0.0));  // single shift as a select
}

// splitmix64 stream: one independent stream per sample
__device__ __forceinline__ uint64_t smix(uint64_t& s) {
  s += 0x9e3779b97f4a7c15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unif(uint64_t& s) { return (double)(smix(s) >> 11) * 0x1.0p-53; }

// Features of a pose (d in A's body frame, Q = R_A^T R_B row-major).
__device__ __forceinline__ void pose_features(const double* dv, const double* Q, double rf, double kappa, float* x) {
  const double d = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
  const double w = rf / d;
  x[0] = (float)(dv[0] / d); x[1] = (float)(dv[1] / d); x[2] = (float)(dv[2] / d);
  x[3] = (float)w; x[4] = (float)(w * w); x[5] = (float)(w * w * w);
  x[6] = (float)Q[0]; x[7] = (float)Q[3]; x[8] = (float)Q[6];
  x[9] = (float)Q[1]; x[10] = (float)Q[4]; x[11] = (float)Q[7];
  x[12] = (float)kappa;
}

// Exact Yukawa force and torque on rigid subunit A (charged beads at qbody) from rigid subunit
// B (charged beads at d + Q qbody), no periodic images: the pair terms of src/forces.cpp:104-107.
__device__ void pair_exact(const FarSpec& f, const double* dv, const double* Q, double kappa, double* F, double* T) {
  F[0] = F[1] = F[2] = T[0] = T[1] = T[2] = 0.0;
  for (int a = 0; a < f.nq; a++) {
    const double* xa = f.qbody[a];
    double fa[3] = {0, 0, 0};
    for (int b = 0; b < f.nq; b++) {
      const double* u = f.qbody[b];
      const double xb[3] = {dv[0] + Q[0] * u[0] + Q[1] * u[1] + Q[2] * u[2],
                            dv[1] + Q[3] * u[0] + Q[4] * u[1] + Q[5] * u[2],
                            dv[2] + Q[6] * u[0] + Q[7] * u[1] + Q[8] * u[2]};
      const double r[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
      const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      const double rr = sqrt(r2);
      const double shc = (f.sig[a] + f.sig[b]) / 2;
      const double Y = exp(kappa * shc) / ((1 + kappa * f.sig[a] / 2) * (1 + kappa * f.sig[b] / 2));
      const double c = (f.q[a] * f.q[b] * f.lb * exp(-kappa * rr) / r2) * (kappa + 1 / rr) * Y;
      fa[0] += r[0] * c; fa[1] += r[1] * c; fa[2] += r[2] * c;
    }
    F[0] += fa[0]; F[1] += fa[1]; F[2] += fa[2];
    T[0] += xa[1] * fa[2] - xa[2] * fa[1];
    T[1] += xa[2] * fa[0] - xa[0] * fa[2];
    T[2] += xa[0] * fa[1] - xa[1] * fa[0];
  }
}

// Training poses: centre distance log-uniform in [rf, 1.5 rf] (50%; the pairs that carry most
// of the far field) and in [1.5 rf, 3 rf] (30%), and 20% with rf/d uniform in
// (max(0.02, rf / (2 rq + 600/kappa)), 1/3); direction uniform on the sphere, orientation uniform (Shoemake quaternion),
// salt log-uniform in [c_lo, c_hi] mM.
__global__ void k_far_samples(const FarSpec* fs, int n, uint64_t seed, double c_lo, double c_hi, float* x, float* y,
                              double* pose) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FarSpec& f = *fs;
  uint64_t s = seed ^ (0xd1b54a32d192ed03ull * (uint64_t)(i + 1));
  smix(s);
  const double c = c_lo * exp(unif(s) * log(c_hi / c_lo));
  const double kappa = f.kcoef * sqrt(c);
  // the far tail stops where exp(-kappa (d - 2 rq)) = e^-600 (beyond it the pair force is
  // below double range; the engine skips such pairs)
  const double wmin = fmax(0.02, f.rf / (2.0 * f.rq + 600.0 / kappa));
  double d;
  const double u0 = unif(s);
  if (u0 < 0.5) d = f.rf * exp(unif(s) * log(1.5));
  else if (u0 < 0.8) d = 1.5 * f.rf * exp(unif(s) * log(2.0));
  else d = f.rf / (wmin + unif(s) * (1.0 / 3.0 - wmin));
  const double z = 2 * unif(s) - 1, ph = 2 * 3.141592653589793 * unif(s);
  const double sz = sqrt(fmax(0.0, 1 - z * z));
  const double dv[3] = {d * sz * cos(ph), d * sz * sin(ph), d * z};
  const double u1 = unif(s), u2 = unif(s), u3 = unif(s);
  const double qa = sqrt(1 - u1) * sin(2 * 3.141592653589793 * u2), qb = sqrt(1 - u1) * cos(2 * 3.141592653589793 * u2),
               qc = sqrt(u1) * sin(2 * 3.141592653589793 * u3), qd = sqrt(u1) * cos(2 * 3.141592653589793 * u3);
  // rotation of the unit quaternion (qd; qa, qb, qc)
  const double Q[9] = {1 - 2 * (qb * qb + qc * qc), 2 * (qa * qb - qc * qd), 2 * (qa * qc + qb * qd),
                       2 * (qa * qb + qc * qd), 1 - 2 * (qa * qa + qc * qc), 2 * (qb * qc - qa * qd),
                       2 * (qa * qc - qb * qd), 2 * (qb * qc + qa * qd), 1 - 2 * (qa * qa + qb * qb)};
  float fx[kFIn];
  pose_features(dv, Q, f.rf, kappa, fx);
  for (int k = 0; k < 16; k++) x[(size_t)i * 16 + k] = k < kFIn ? fx[k] : 0.0f;
  double F[3], T[3];
  pair_exact(f, dv, Q, kappa, F, T);
  const double S = far_scale(f, dv, Q, kappa);
  for (int k = 0; k < 3; k++) {
    y[(size_t)i * 8 + k] = (float)(F[k] / S);
    y[(size_t)i * 8 + 3 + k] = (float)(T[k] / (S * f.rq));
  }
  y[(size_t)i * 8 + 6] = 0.0f; y[(size_t)i * 8 + 7] = 0.0f;
  if (pose) {
    double* p = pose + (size_t)i * 16;
    p[0] = dv[0]; p[1] = dv[1]; p[2] = dv[2];
    for (int k = 0; k < 9; k++) p[3 + k] = Q[k];
    p[12] = kappa; p[13] = c; p[14] = S; p[15] = 0.0;
  }
}

// ---- FP32 training kernels -----------------------------------------------------------------

// C[M][N] = sum_k A(m,k) B(k,n), A(m,k) = ta ? A[k lda + m] : A[m lda + k],
// B(k,n) = tb ? B[n ldb + k] : B[k ldb + n]. 64x64 tiles, k-steps of 16, k ascending.
template <int TA, int TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float* __restrict__ A, int lda,
                                              const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc) {
  __shared__ float As[16][68];
  __shared__ float Bs[16][68];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0f;
  for (int k0 = 0; k0 < K; k0 += 16) {
    for (int e = threadIdx.x; e < 1024; e += 256) {
      int mm, kk;
      if (TA) { mm = e & 63; kk = e >> 6; } else { kk = e & 15; mm = e >> 4; }
      const int m = m0 + mm, k = k0 + kk;
      float v = 0.0f;
      if (m < M && k < K) v = TA ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k];
      As[kk][mm] = v;
    }
    for (int e = threadIdx.x; e < 1024; e += 256) {
      int nn, kk;
      if (TB) { kk = e & 15; nn = e >> 4; } else { nn = e & 63; kk = e >> 6; }
      const int nc = n0 + nn, k = k0 + kk;
      float v = 0.0f;
      if (nc < N && k < K) v = TB ? B[(size_t)nc * ldb + k] : B[(size_t)k * ldb + nc];
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int m = m0 + ty * 4 + i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int nc = n0 + tx * 4 + j;
      if (nc < N) C[(size_t)m * ldc + nc] = acc[i][j];
    }
  }
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }

// Z += bias (per column); A = act ? SiLU(Z) : Z.
__global__ void k_bias_act(float* Z, const float* b, float* A, int rows, int cols, int act) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)rows * cols) return;
  const float z = Z[i] + b[i % cols];
  Z[i] = z;
  A[i] = act ? silu(z) : z;
}

// dA <- dA * SiLU'(Z), SiLU'(z) = s (1 + z (1 - s)), s = 1/(1 + e^-z).
__global__ void k_silu_back(float* dA, const float* Z, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float z = Z[i];
  const float s = 1.0f / (1.0f + expf(-z));
  dA[i] = dA[i] * (s * (1.0f + z * (1.0f - s)));
}

// out[c] = sum over rows of M[r][c]: one block per column, fixed-order tree.
__global__ __launch_bounds__(256) void k_colsum(const float* M, int rows, int cols, float* out) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float s = 0.0f;
  for (int r = threadIdx.x; r < rows; r += 256) s += M[(size_t)r * cols + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// Mean squared error over rows x cols: dy = 2 (yhat - y) / (rows cols); per-block partial sums
// of the squared errors (fixed order).
__global__ __launch_bounds__(256) void k_mse(const float* yhat, const float* y, int n, float inv, float* dy,
                                             double* part) {
  __shared__ double red[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double e2 = 0.0;
  if (i < n) {
    const float e = yhat[i] - y[i];
    dy[i] = 2.0f * e * inv;
    e2 = (double)e * (double)e;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam (Kingma and Ba): c1 = 1/(1 - b1^t), c2 = 1/(1 - b2^t).
__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void k_gather_rows(const float* src, const int* idx, float* dst, int rows, int cols) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)rows * cols) return;
  const size_t r = i / cols, c = i - r * cols;
  dst[i] = src[(size_t)idx[r] * cols + c];
}

// ---- BF16 MFMA inference -----------------------------------------------------------------

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// SiLU for inference on log2(e)-scaled pre-activations: u = log2(e) z gives
// u / (1 + 2^-u) = log2(e) silu(z) with the bare hardware exp2 and reciprocal. The packed
// weights carry the factors (layer-1 rows and hidden biases times log2(e), output rows times
// ln 2; the hidden-to-hidden rows need none, since ln 2 log2(e) = 1).
__device__ __forceinline__ float silu_l2(float u) {
  return u * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-u));
}

// Two floats to packed bf16 (round to nearest even, v_cvt_pk_bf16_f32).
__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

// FP32 -> bf16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ unsigned short bf16_bits(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// Hidden-layer input order of the packed W2, W3, W4: the inference keeps each layer's outputs
// in the MFMA accumulator layout, where lane group g (lanes 16 g .. 16 g + 15) holds outputs
// 16 t + 4 g .. 16 t + 4 g + 3 of every 16-row tile t. The next layer's K-block kb (32 inputs)
// takes tiles 2 kb and 2 kb + 1 from the same lanes, so packed input position 32 kb + 8 g + j
// is hidden unit 32 kb + 4 g + j (j < 4) or 32 kb + 16 + 4 g + j - 4 (j >= 4).
__device__ __forceinline__ int hidden_of_position(int k) {
  const int kb = k >> 5, g = (k >> 3) & 3, j = k & 7;
  return 32 * kb + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4);
}

__global__ void k_pack_weights(const float* p, unsigned short* w, float* bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const float log2e = 1.4426950408889634f, ln2 = 0.6931471805599453f;  // exp2-based SiLU (silu_l2)
  if (i < kFHid * kFLd1) {  // W1 [128][40], inputs 13..39 zero
    const int o = i / kFLd1, k = i - o * kFLd1;
    w[kFPackW1 + i] = bf16_bits(k < kFIn ? log2e * p[kFW1 + o * kFIn + k] : 0.0f);
  }
  if (i < kFHid * kFLdH) {  // W2, W3 [128][136], inputs in the accumulator order
    const int o = i / kFLdH, k = i - o * kFLdH;
    const int h = k < kFHid ? hidden_of_position(k) : 0;
    w[kFPackW2 + i] = bf16_bits(k < kFHid ? p[kFW2 + o * kFHid + h] : 0.0f);
    w[kFPackW3 + i] = bf16_bits(k < kFHid ? p[kFW3 + o * kFHid + h] : 0.0f);
  }
  if (i < kFOutPad * kFLdH) {  // W4 [16][136], rows 6..15 zero
    const int o = i / kFLdH, k = i - o * kFLdH;
    const int h = k < kFHid ? hidden_of_position(k) : 0;
    w[kFPackW4 + i] = bf16_bits(o < kFOut && k < kFHid ? ln2 * p[kFW4 + o * kFHid + h] : 0.0f);
  }
  if (i < 3 * kFHid) bias[i] = log2e * p[(i < kFHid ? kFB1 : i < 2 * kFHid ? kFB2 - kFHid : kFB3 - 2 * kFHid) + i];
  if (i < kFOutPad) bias[3 * kFHid + i] = i < kFOut ? p[kFB4 + i] : 0.0f;
}

constexpr int kMlpWaves = 8;
constexpr int kMlpThreads = 64 * kMlpWaves;
constexpr int kRowsPerWave = 64;           // pairs (rows) a wave carries through the network
constexpr int kTileLd = kFLd1;             // feature tile row stride (bf16): 32 inputs + 8 pad
constexpr int kTileElems = kRowsPerWave * kTileLd;
constexpr size_t kMlpLds = (size_t)kFPackN * 2 + (size_t)kMlpWaves * kTileElems * 2 + (3 * kFHid + kFOutPad) * 4;

// Orders a wave's LDS writes before its following LDS reads (lanes read other lanes' rows).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wave runs the network on pairs 16 pt0 .. 16 (pt0 + NPT) - 1 of its LDS feature tile h
// (row = pair, 13 features in columns 0..12 of 32, bf16). Each layer is computed transposed,
// D = W H^T, with the weights (LDS) as the A operand and the activations as the B operand of
// v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)..+7] (a weight row), B[8(l/16)..+7][l%16]
// (one pair's inputs) and D[4(l/16)+i][l%16] (four outputs of one pair). Returns
// out[p][i] = output 4(l/16)+i of pair 16 (pt0 + p) + l%16 (outputs 0..5 used).
//
// A hidden layer runs block by block: the 16 outputs of row block mt are complete after its
// K-blocks, go through SiLU while the MFMAs of block mt + 2 run, and become half of bf16 K-block
// mt/2 of the next layer's B operand in the same lanes (block 2k gives inputs 8g..8g+3 of lane
// group g, block 2k+1 inputs 8g+4..8g+7; the next layer's weights are packed in that input
// order), so activations never leave the registers and the vector ALU overlaps the matrix core.
template <int NPT>
__device__ __forceinline__ void silu_block(const f32x4 (&u)[NPT], unsigned (&w)[NPT][2]) {
#pragma unroll
  for (int pt = 0; pt < NPT; pt++) {
    w[pt][0] = pack_bf16(silu_l2(u[pt][0]), silu_l2(u[pt][1]));
    w[pt][1] = pack_bf16(silu_l2(u[pt][2]), silu_l2(u[pt][3]));
  }
}

// One hidden layer (128 outputs) from KB input K-blocks into bout[4] (bf16 SiLU outputs).
// Iteration mt requests block mt+1's weight fragments, issues block mt's MFMAs and the SiLU of
// block mt-2 (independent, so the compiler interleaves them); a scheduling fence between
// iterations bounds the registers in flight.
template <int NPT, int KB>
__device__ __forceinline__ void hidden_layer(const unsigned short* Wl, int ldw, const float* bl,
                                             const bf16x8 (&bin)[KB][NPT], bf16x8 (&bout)[4][NPT], int r16, int kq) {
  const unsigned short* wrow = Wl + r16 * ldw + 8 * kq;
  bf16x8 a[KB], an[KB];
#pragma unroll
  for (int ks = 0; ks < KB; ks++) a[ks] = *reinterpret_cast<const bf16x8*>(wrow + 32 * ks);
  f32x4 pend[2][NPT];
  unsigned lo[NPT][2], hi[NPT][2];
#pragma unroll
  for (int mt = 0; mt < 10; mt++) {
    if (mt < 7) {
#pragma unroll
      for (int ks = 0; ks < KB; ks++) an[ks] = *reinterpret_cast<const bf16x8*>(wrow + (mt + 1) * 16 * ldw + 32 * ks);
    }
    f32x4 acc[NPT];
    if (mt < 8) {
      const float4 bv = *reinterpret_cast<const float4*>(bl + mt * 16 + 4 * kq);
#pragma unroll
      for (int pt = 0; pt < NPT; pt++) acc[pt] = f32x4{bv.x, bv.y, bv.z, bv.w};
#pragma unroll
      for (int ks = 0; ks < KB; ks++)
#pragma unroll
        for (int pt = 0; pt < NPT; pt++) acc[pt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], bin[ks][pt], acc[pt], 0, 0, 0);
    }
    if (mt >= 2) {  // SiLU of block mt - 2
      if (mt & 1) {
        silu_block<NPT>(pend[mt & 1], hi);
#pragma unroll
        for (int pt = 0; pt < NPT; pt++)
          bout[(mt - 2) >> 1][pt] = __builtin_bit_cast(bf16x8, make_uint4(lo[pt][0], lo[pt][1], hi[pt][0], hi[pt][1]));
      } else {
        silu_block<NPT>(pend[mt & 1], lo);
      }
    }
    if (mt < 8) {
#pragma unroll
      for (int pt = 0; pt < NPT; pt++) pend[mt & 1][pt] = acc[pt];
#pragma unroll
      for (int ks = 0; ks < KB; ks++) a[ks] = an[ks];
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int NPT>
__device__ __forceinline__ void mlp_tile(const unsigned short* W, const float* bias, const unsigned short* h, int pt0,
                                         int lane, f32x4 out[NPT]) {
  const int r16 = lane & 15, kq = lane >> 4;
  bf16x8 x[1][NPT], ba[4][NPT], bb[4][NPT];
#pragma unroll
  for (int pt = 0; pt < NPT; pt++) x[0][pt] = *reinterpret_cast<const bf16x8*>(h + ((pt0 + pt) * 16 + r16) * kTileLd + 8 * kq);
  hidden_layer<NPT, 1>(W + kFPackW1, kFLd1, bias, x, ba, r16, kq);
  hidden_layer<NPT, 4>(W + kFPackW2, kFLdH, bias + kFHid, ba, bb, r16, kq);
  hidden_layer<NPT, 4>(W + kFPackW3, kFLdH, bias + 2 * kFHid, bb, ba, r16, kq);
  // output layer: one 16-row block (rows 6..15 zero)
  const float4 bv = *reinterpret_cast<const float4*>(bias + 3 * kFHid + 4 * kq);
#pragma unroll
  for (int pt = 0; pt < NPT; pt++) out[pt] = f32x4{bv.x, bv.y, bv.z, bv.w};
#pragma unroll
  for (int ks = 0; ks < 4; ks++) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(W + kFPackW4 + r16 * kFLdH + 32 * ks + 8 * kq);
#pragma unroll
    for (int pt = 0; pt < NPT; pt++) out[pt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, ba[ks][pt], out[pt], 0, 0, 0);
  }
}

// Loads the packed weights and biases into LDS.
__device__ __forceinline__ void load_net(const unsigned short* wsrc, const float* bsrc, unsigned short* W, float* bias) {
  const uint4* s = reinterpret_cast<const uint4*>(wsrc);
  uint4* d = reinterpret_cast<uint4*>(W);
  for (int i = threadIdx.x; i < kFPackN / 8; i += blockDim.x) d[i] = s[i];
  for (int i = threadIdx.x; i < 3 * kFHid + kFOutPad; i += blockDim.x) bias[i] = bsrc[i];
}

// Writes one tile row: 13 features as bf16, zero to column 31.
__device__ __forceinline__ void put_features(unsigned short* row, const float* x, bool valid) {
  unsigned w[16];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const float a = (valid && 2 * k < kFIn) ? x[2 * k] : 0.0f;
    const float b = (valid && 2 * k + 1 < kFIn) ? x[2 * k + 1] : 0.0f;
    w[k] = pack_bf16(a, b);
  }
  uint4* dst = reinterpret_cast<uint4*>(row);
#pragma unroll
  for (int k = 0; k < 4; k++) dst[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
}

// Standalone inference on n rows of 13 features: y [n][6] raw network outputs.
__global__ __launch_bounds__(kMlpThreads) void k_mlp_rows(const unsigned short* wpack, const float* bpack, int n,
                                                          const float* x, float* y) {
  extern __shared__ unsigned char lds[];
  unsigned short* W = reinterpret_cast<unsigned short*>(lds);
  unsigned short* tiles = W + kFPackN;
  float* bias = reinterpret_cast<float*>(tiles + kMlpWaves * kTileElems);
  load_net(wpack, bpack, W, bias);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  unsigned short* h = tiles + wave * kTileElems;
  const int row0 = (blockIdx.x * kMlpWaves + wave) * kRowsPerWave;
  {
    const int r = row0 + lane;
    float xr[kFIn];
    for (int k = 0; k < kFIn; k++) xr[k] = r < n ? x[(size_t)r * kFIn + k] : 0.0f;
    put_features(h + lane * kTileLd, xr, r < n);
  }
  __syncthreads();
  f32x4 out[4];
  mlp_tile<4>(W, bias, h, 0, lane, out);
  const int r16 = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int pt = 0; pt < 4; pt++) {
    const int r = row0 + pt * 16 + r16;
    if (r >= n) continue;
#pragma unroll
    for (int i = 0; i < 4; i++)
      if (4 * kq + i < kFOut) y[(size_t)r * kFOut + 4 * kq + i] = out[pt][i];
  }
}

// Body frame of every subunit from its centre bead and the two frame beads (FP32 copy).
__global__ void k_frames(Dev d, FarDev f, int fp1, int fp2) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= d.Ns) return;
  const RepParams& P = d.rp[d.sub_rep[s]];
  const double4 c = d.centres[s];
  const double4 p1 = d.pos[(size_t)s * d.topo.np + fp1], p2 = d.pos[(size_t)s * d.topo.np + fp2];
  const double u1[3] = {min_img_f(p1.x - c.x, P.box, P.hbox), min_img_f(p1.y - c.y, P.box, P.hbox),
                        min_img_f(p1.z - c.z, P.box, P.hbox)};
  const double u2[3] = {min_img_f(p2.x - c.x, P.box, P.hbox), min_img_f(p2.y - c.y, P.box, P.hbox),
                        min_img_f(p2.z - c.z, P.box, P.hbox)};
  double R[9];
  body_frame(u1, u2, R);
  for (int k = 0; k < 9; k++) f.frames[(size_t)s * 12 + k] = (float)R[k];
}

// Far field of every subunit A by the network: all subunits B of A's replica whose centre
// bead lies beyond rf, 256 pairs per round (64 per wave, one per lane). Features, the pair
// scale and the spread weights are computed in FP32 (the centre-bead separation in FP64).
// Each pair's force is spread over A's charged beads with weights exp(kappa qbody_a . n) / sum
// (n: unit vector to B in A's body frame; nearer charges feel the screened field more) and the
// pair torques are summed; after all pairs a rigid correction w x (qbody_a - qcen),
// w = M^-1 (tau - tau_spread), restores the network's total torque about the centre bead. Sums
// run in a fixed order; the per-bead forces go to the lab frame with A's body frame.
// Persistent: workgroups stride over the subunits.
__global__ __launch_bounds__(kMlpThreads) void k_far_mlp(Dev d, FarDev f) {
  extern __shared__ unsigned char lds[];
  unsigned short* W = reinterpret_cast<unsigned short*>(lds);
  unsigned short* tiles = W + kFPackN;
  float* bias = reinterpret_cast<float*>(tiles + kMlpWaves * kTileElems);
  __shared__ double red[kMlpWaves][3 * kFMaxQ + 3];
  __shared__ double fbs[kFMaxQ][3];
  __shared__ FarSpec spec;
  if (threadIdx.x == 0) spec = *f.spec;
  load_net(f.wpack, f.bias, W, bias);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r16 = lane & 15, kq = lane >> 4;
  unsigned short* h = tiles + wave * kTileElems;
  const int nq = d.topo.nq;
  // charged beads (FP32) and the output standardisation in LDS; registers go to the network
  __shared__ float qb[kFMaxQ][3];
  __shared__ float ymu[kFOutPad], ysd[kFOutPad];
  if (threadIdx.x < kFMaxQ * 3) {
    const int a = threadIdx.x / 3, c = threadIdx.x % 3;
    qb[a][c] = a < nq ? (float)spec.qbody[a][c] : 0.0f;
  }
  if (threadIdx.x < kFOutPad) {
    ymu[threadIdx.x] = threadIdx.x < kFOut ? spec.ymu[threadIdx.x] : 0.0f;
    ysd[threadIdx.x] = threadIdx.x < kFOut ? spec.ysd[threadIdx.x] : 0.0f;
  }
  __syncthreads();
  const float rf = (float)spec.rf, lbf = (float)spec.lb;
  for (int A = blockIdx.x; A < d.Ns; A += gridDim.x) {
    const RepParams& P = d.rp[d.sub_rep[A]];
    const double box = P.box, hbox = P.hbox;
    const float kappa = (float)P.kappa;
    const double4 ca = d.centres[A];
    float RA[9];
#pragma unroll
    for (int k = 0; k < 9; k++) RA[k] = f.frames[(size_t)A * 12 + k];
    // kq == 0 lanes: spread force (3 components x charged beads) and torque x; kq == 1: torque y, z.
    // FP32 per lane (a fixed order over the lane's pairs), FP64 across lanes and waves.
    float accq[3][kFMaxQ], acct[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int a = 0; a < kFMaxQ; a++) accq[c][a] = 0.0f;
    for (int t0 = 0; t0 < P.ns; t0 += kRowsPerWave * kMlpWaves) {
      // this lane's pair: features, scale and spread weights
      const int t = t0 + wave * kRowsPerWave + lane;
      float x[kFIn], w[kFMaxQ];
      float sc = 0.0f;
      if (t < P.ns && P.s0 + t != A) {
        const int B = P.s0 + t;
        const double4 cb = d.centres[B];
        const double cx = min_img_f(ca.x - cb.x, box, hbox), cy = min_img_f(ca.y - cb.y, box, hbox),
                     cz = min_img_f(ca.z - cb.z, box, hbox);
        const double d2 = cx * cx + cy * cy + cz * cz;
        if (d2 > P.far_rf2) {
          const float fx = (float)cx, fy = (float)cy, fz = (float)cz;
          const float dv[3] = {-(RA[0] * fx + RA[3] * fy + RA[6] * fz), -(RA[1] * fx + RA[4] * fy + RA[7] * fz),
                               -(RA[2] * fx + RA[5] * fy + RA[8] * fz)};
          const float4* rb4 = reinterpret_cast<const float4*>(f.frames + (size_t)B * 12);  // 48-byte rows
          const float4 rb0 = rb4[0], rb1 = rb4[1], rb2 = rb4[2];
          const float RB[9] = {rb0.x, rb0.y, rb0.z, rb0.w, rb1.x, rb1.y, rb1.z, rb1.w, rb2.x};
          float Q[9];
#pragma unroll
          for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) Q[3 * i + j] = RA[i] * RB[j] + RA[3 + i] * RB[3 + j] + RA[6 + i] * RB[6 + j];
          const float inv = __builtin_amdgcn_rsqf((float)d2), dd = (float)d2 * inv;
          const float n[3] = {dv[0] * inv, dv[1] * inv, dv[2] * inv};
          const float wr = rf * inv;
          x[0] = n[0]; x[1] = n[1]; x[2] = n[2];
          x[3] = wr; x[4] = wr * wr; x[5] = wr * wr * wr;
          x[6] = Q[0]; x[7] = Q[3]; x[8] = Q[6]; x[9] = Q[1]; x[10] = Q[4]; x[11] = Q[7];
          x[12] = kappa;
          // pair scale (far_scale) and spread weights; B's charges project on n as -(Q qb).n = -qb.(Q^T n)
          const float m[3] = {Q[0] * n[0] + Q[3] * n[1] + Q[6] * n[2], Q[1] * n[0] + Q[4] * n[1] + Q[7] * n[2],
                              Q[2] * n[0] + Q[5] * n[1] + Q[8] * n[2]};
          float pa[kFMaxQ], ma = -1e30f, mb = -1e30f;
#pragma unroll
          for (int a = 0; a < kFMaxQ; a++) {
            if (a >= nq) break;
            pa[a] = qb[a][0] * n[0] + qb[a][1] * n[1] + qb[a][2] * n[2];
            const float pb = -(qb[a][0] * m[0] + qb[a][1] * m[1] + qb[a][2] * m[2]);
            w[a] = pb;  // reuse as scratch for p'_b
            ma = fmaxf(ma, pa[a]);
            mb = fmaxf(mb, pb);
          }
          float sa = 0.0f, sb = 0.0f;
#pragma unroll
          for (int a = 0; a < kFMaxQ; a++) {
            if (a >= nq) break;
            sb += __expf(kappa * (w[a] - mb));
            w[a] = __expf(kappa * (pa[a] - ma));
            sa += w[a];
          }
          const float isa = __builtin_amdgcn_rcpf(sa);
#pragma unroll
          for (int a = 0; a < kFMaxQ; a++) w[a] = a < nq ? w[a] * isa : 0.0f;
          sc = lbf * (kappa + inv) * inv * __expf(-kappa * (dd - ma - mb)) * sa * sb;
        }
      }
      if (!(sc > 0.0f)) {
        sc = 0.0f;
#pragma unroll
        for (int a = 0; a < kFMaxQ; a++) w[a] = 0.0f;
      }
      put_features(h + lane * kTileLd, x, sc > 0.0f);
      wave_lds_sync();
      // two halves of 32 pairs, so the network's registers fit beside the spread sums
#pragma unroll 1
      for (int half = 0; half < 2; half++) {
        f32x4 out[2];
        mlp_tile<2>(W, bias, h, 2 * half, lane, out);
#pragma unroll
        for (int pp = 0; pp < 2; pp++) {
          const int pt = 2 * half + pp;
          const int src = pt * 16 + r16;  // the lane that built this pair
          const float scp = __shfl(sc, src, 64);
          float wp[kFMaxQ];
#pragma unroll
          for (int a = 0; a < kFMaxQ; a++) wp[a] = __shfl(w[a], src, 64);
          if (scp == 0.0f) continue;
          const float v0 = (ymu[4 * kq] + ysd[4 * kq] * out[pp][0]) * scp,
                      v1 = (ymu[4 * kq + 1] + ysd[4 * kq + 1] * out[pp][1]) * scp,
                      v2 = (ymu[4 * kq + 2] + ysd[4 * kq + 2] * out[pp][2]) * scp,
                      v3 = (ymu[4 * kq + 3] + ysd[4 * kq + 3] * out[pp][3]) * scp;
          if (kq == 0) {
#pragma unroll
            for (int a = 0; a < kFMaxQ; a++) {
              accq[0][a] = fmaf(v0, wp[a], accq[0][a]);
              accq[1][a] = fmaf(v1, wp[a], accq[1][a]);
              accq[2][a] = fmaf(v2, wp[a], accq[2][a]);
            }
            acct[0] += v3;
          } else if (kq == 1) {
            acct[1] += v0;
            acct[2] += v1;
          }
        }
      }
      wave_lds_sync();  // the tile's reads are done before the next round writes it
    }
    // fixed-order FP64 sums over the 16 lanes of a quarter-wave, then the waves
    double dq[3][kFMaxQ], dt3[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
#pragma unroll
      for (int a = 0; a < kFMaxQ; a++) dq[c][a] = (double)accq[c][a];
      dt3[c] = (double)acct[c];
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int a = 0; a < kFMaxQ; a++) dq[c][a] += __shfl_xor(dq[c][a], o, 64);
#pragma unroll
      for (int c = 0; c < 3; c++) dt3[c] += __shfl_xor(dt3[c], o, 64);
    }
    if (lane == 0) {
#pragma unroll
      for (int a = 0; a < kFMaxQ; a++)
#pragma unroll
        for (int c = 0; c < 3; c++)
          if (a < nq) red[wave][3 * a + c] = dq[c][a];
      red[wave][3 * kFMaxQ] = dt3[0];
    } else if (lane == 16) {
      red[wave][3 * kFMaxQ + 1] = dt3[1];
      red[wave][3 * kFMaxQ + 2] = dt3[2];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double (*fb)[3] = fbs;  // LDS: indexed by the runtime charge count
      double T[3], Ts[3] = {0, 0, 0};
      for (int a = 0; a < nq; a++)
        for (int c = 0; c < 3; c++) {
          double s = 0.0;
          for (int w2 = 0; w2 < kMlpWaves; w2++) s += red[w2][3 * a + c];
          fb[a][c] = s;
        }
      for (int c = 0; c < 3; c++) {
        double s = 0.0;
        for (int w2 = 0; w2 < kMlpWaves; w2++) s += red[w2][3 * kFMaxQ + c];
        T[c] = s * spec.rq;
      }
      for (int a = 0; a < nq; a++) {
        const double* q = spec.qbody[a];
        Ts[0] += q[1] * fb[a][2] - q[2] * fb[a][1];
        Ts[1] += q[2] * fb[a][0] - q[0] * fb[a][2];
        Ts[2] += q[0] * fb[a][1] - q[1] * fb[a][0];
      }
      const double dt[3] = {T[0] - Ts[0], T[1] - Ts[1], T[2] - Ts[2]};
      double wv[3];
      for (int i = 0; i < 3; i++) wv[i] = spec.minv[3 * i] * dt[0] + spec.minv[3 * i + 1] * dt[1] + spec.minv[3 * i + 2] * dt[2];
      for (int a = 0; a < nq; a++) {
        const double r[3] = {spec.qbody[a][0] - spec.qcen[0], spec.qbody[a][1] - spec.qcen[1], spec.qbody[a][2] - spec.qcen[2]};
        const double g[3] = {fb[a][0] + (wv[1] * r[2] - wv[2] * r[1]), fb[a][1] + (wv[2] * r[0] - wv[0] * r[2]),
                             fb[a][2] + (wv[0] * r[1] - wv[1] * r[0])};
        double* o = d.qfar + 3 * ((size_t)A * nq + a);
        for (int i = 0; i < 3; i++) o[i] = (double)RA[3 * i] * g[0] + (double)RA[3 * i + 1] * g[1] + (double)RA[3 * i + 2] * g[2];
      }
    }
    __syncthreads();
  }
}

// Audit sums over all beads: out[0] = sum over charged beads |qfar_model - qfar_exact|^2,
// out[1] = sum over beads |F|^2 (the integrator force after the step's force kernel).
// Audit sums of one replica per workgroup, in a fixed order: out[3 r] = sum over its charged beads
// of |F_far,model - F_far,exact|^2, out[3 r + 1] = sum of |F_far,exact|^2 (the far field the model
// replaces), out[3 r + 2] = sum over all its beads of |F|^2 with the exact far field, F being the
// integrator force of the last step (MD: the conservative force) with the model's far field
// swapped for the exact one.
__global__ __launch_bounds__(1024) void k_far_audit(Dev d, const double* qfar_exact, double* out) {
  __shared__ double red[3][16];
  const RepParams& P = d.rp[blockIdx.x];
  const int np = d.topo.np, nq = d.topo.nq;
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = 3 * P.q0 + threadIdx.x; i < 3 * (P.q0 + P.nqr); i += 1024) {
    const double x = qfar_exact[i], e = d.qfar[i] - x;
    a += e * e;
    b += x * x;
  }
  for (int i = P.b0 + threadIdx.x; i < P.b0 + P.nb; i += 1024) {
    double fx = d.fx[i], fy = d.fy[i], fz = d.fz[i];
    const int l = i % np, qs = d.topo.qslot[l];
    if (qs >= 0) {
      const size_t k = 3 * ((size_t)(i / np) * nq + qs);
      fx += qfar_exact[k] - d.qfar[k];
      fy += qfar_exact[k + 1] - d.qfar[k + 1];
      fz += qfar_exact[k + 2] - d.qfar[k + 2];
    }
    c += fx * fx + fy * fy + fz * fz;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, o, 64);
    b += __shfl_down(b, o, 64);
    c += __shfl_down(c, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a; red[1][threadIdx.x >> 6] = b; red[2][threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0, sb = 0, sc = 0;
    for (int w = 0; w < 16; w++) { sa += red[0][w]; sb += red[1][w]; sc += red[2][w]; }
    out[3 * blockIdx.x] = sa; out[3 * blockIdx.x + 1] = sb; out[3 * blockIdx.x + 2] = sc;
  }
}

}  // namespace

int launch_far_samples(const FarSpec* spec_dev, int n, uint64_t seed, double c_lo, double c_hi, float* x, float* y,
                       double* pose, hipStream_t st) {
  hipLaunchKernelGGL(k_far_samples, dim3((n + 127) / 128), dim3(128), 0, st, spec_dev, n, seed, c_lo, c_hi, x, y, pose);
  return check_launch_far();
}

int launch_gemm(int ta, int tb, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                hipStream_t st) {
  dim3 g((N + 63) / 64, (M + 63) / 64), b(256);
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm<0, 0>), g, b, 0, st, M, N, K, A, lda, B, ldb, C, ldc);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm<0, 1>), g, b, 0, st, M, N, K, A, lda, B, ldb, C, ldc);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm<1, 0>), g, b, 0, st, M, N, K, A, lda, B, ldb, C, ldc);
  else hipLaunchKernelGGL((k_gemm<1, 1>), g, b, 0, st, M, N, K, A, lda, B, ldb, C, ldc);
  return check_launch_far();
}

int launch_bias_act(float* Z, const float* b, float* A, int rows, int cols, int act, hipStream_t st) {
  const size_t n = (size_t)rows * cols;
  hipLaunchKernelGGL(k_bias_act, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, Z, b, A, rows, cols, act);
  return check_launch_far();
}

int launch_silu_back(float* dA, const float* Z, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_silu_back, dim3((n + 255) / 256), dim3(256), 0, st, dA, Z, n);
  return check_launch_far();
}

int launch_colsum(const float* M, int rows, int cols, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_colsum, dim3(cols), dim3(256), 0, st, M, rows, cols, out);
  return check_launch_far();
}

int launch_mse(const float* yhat, const float* y, int rows, int cols, float* dy, double* part, hipStream_t st) {
  const int n = rows * cols;
  hipLaunchKernelGGL(k_mse, dim3((n + 255) / 256), dim3(256), 0, st, yhat, y, n, 1.0f / (float)n, dy, part);
  return check_launch_far();
}

int launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps, float c1,
                float c2, hipStream_t st) {
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
  return check_launch_far();
}

int launch_gather_rows(const float* src, const int* idx, float* dst, int rows, int cols, hipStream_t st) {
  const size_t n = (size_t)rows * cols;
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, idx, dst, rows, cols);
  return check_launch_far();
}

int launch_pack_weights(const float* params, unsigned short* wpack, float* bias, hipStream_t st) {
  const int n = kFHid * kFLdH;
  hipLaunchKernelGGL(k_pack_weights, dim3((n + 255) / 256), dim3(256), 0, st, params, wpack, bias);
  return check_launch_far();
}

int launch_mlp_rows(const unsigned short* wpack, const float* bias, int n, const float* x, float* y, hipStream_t st) {
  const int rows_per_block = kRowsPerWave * kMlpWaves;
  static bool lds_set = false;
  if (!lds_set) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mlp_rows), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kMlpLds) != hipSuccess) {
      set_error("cannot raise the LDS limit of the inference kernel");
      return SF_E_HIP;
    }
    lds_set = true;
  }
  hipLaunchKernelGGL(k_mlp_rows, dim3((n + rows_per_block - 1) / rows_per_block), dim3(kMlpThreads), kMlpLds, st, wpack,
                     bias, n, x, y);
  return check_launch_far();
}

int launch_frames(const Dev& d, const FarDev& f, hipStream_t st) {
  hipLaunchKernelGGL(k_frames, dim3((d.Ns + 127) / 128), dim3(128), 0, st, d, f, d.topo.fp1, d.topo.fp2);
  return check_launch_far();
}

int launch_far_mlp(const Dev& d, const FarDev& f, hipStream_t st) {
  static bool lds_set = false;
  if (!lds_set) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_far_mlp), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kMlpLds) != hipSuccess) {
      set_error("cannot raise the LDS limit of the far-field kernel");
      return SF_E_HIP;
    }
    lds_set = true;
  }
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const int grid = std::min(d.Ns, ncu);
  hipLaunchKernelGGL(k_far_mlp, dim3(grid), dim3(kMlpThreads), kMlpLds, st, d, f);
  return check_launch_far();
}

int launch_far_audit(const Dev& d, const double* qfar_exact, double* out2, hipStream_t st) {
  hipLaunchKernelGGL(k_far_audit, dim3(d.R), dim3(1024), 0, st, d, qfar_exact, out2);
  return check_launch_far();
}

size_t mlp_lds_bytes() { return kMlpLds; }

}  // namespace sf
