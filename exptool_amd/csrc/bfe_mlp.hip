// MLP neural field x, y, z -> (Phi, F_R, F_phi, F_z): 3 -> 128 -> 128 -> 128 -> 128 -> 4,
// SiLU activations.  Training and inference kernels for gfx950.
//
// Precision: the 3 -> 128 input layer and the 128 -> 4 output layer run in
// float32 on the vector ALUs (positions are never rounded to bf16); the three
// 128 x 128 layers run on v_mfma_f32_16x16x32_bf16 with bf16 weights and
// activations and float32 accumulation.  Master weights and Adam moments are
// float32.
//
// MFMA operand layout (16x16x32, wave64): lane l, g = l / 16, c = l % 16
//   A (16 x 32):  A[c][8g + j], j < 8        B (32 x 16): B[8g + j][c]
//   D (16 x 16):  D[4g + i][c], i < 4
// Activations are kept "neuron x point": M = output neurons, N = points,
// K = input neurons.  Lane l of an output tile (mt, nt) holds neurons
// 16 mt + 4 g + i of point 16 nt + c; the next layer's B operand for K-step
// kk takes the 8 values of tiles 2kk and 2kk+1, i.e. logical k = 32kk + 8g + j
// stands for neuron pi(k) = 32kk + 4g + j (j < 4) or 32kk + 16 + 4g + j - 4.
// The weights are stored permuted along K by the same pi, so no shuffle is
// needed between layers.  The same holds for the backward chain with the
// transposed weights.
#include <math.h>

#include <vector>

#include "bfe_field.hpp"

namespace bfe {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kH = 128;
constexpr int kNIn = 3;
constexpr int kNOut = 4;
// parameter offsets (float32 master vector)
constexpr int kOffW0 = 0;                       // [128][3]
constexpr int kOffB0 = kOffW0 + kH * kNIn;       // [128]
constexpr int kOffW1 = kOffB0 + kH;              // [128][128]
constexpr int kOffB1 = kOffW1 + kH * kH;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH * kH;
constexpr int kOffW4 = kOffB3 + kH;              // [4][128]
constexpr int kOffB4 = kOffW4 + kNOut * kH;      // [4]
constexpr int kNParams = kOffB4 + kNOut;         // 50564
static_assert(kNParams == BFE_MLP_NPARAMS, "parameter count");
constexpr int kHidOff[3] = {kOffW1, kOffW2, kOffW3};
constexpr int kSmallFloats = kH * kNIn + 4 * kH + kNOut * kH + kNOut;  // W0, b0..b3, W4, b4
#ifndef BFE_MLP_NT
#define BFE_MLP_NT 2
#endif
#ifndef BFE_MLP_BLOCK
#define BFE_MLP_BLOCK 512
#endif
constexpr int kNT = BFE_MLP_NT;                  // 16-point N-tiles per wave
constexpr int kWaveTile = 16 * kNT;              // points per wave
// Threads per block.  The inference and integrator kernels run 8 waves on
// one copy of the LDS weights (two waves per SIMD hide the SiLU latency);
// the training kernels keep 4 waves, whose stores of the intermediates need
// the registers.
constexpr int kBlockInf = BFE_MLP_BLOCK;
constexpr int kBlockTrain = 256;
// The conservative-field kernels carry three tangent operand sets beside the
// activations: four waves per block, one per SIMD, keep them in registers.
constexpr int kBlockGrad = 256;
template <bool TRAIN>
constexpr int block_threads() { return TRAIN ? kBlockTrain : kBlockInf; }
template <bool TRAIN>
constexpr int block_points() { return kWaveTile * block_threads<TRAIN>() / 64; }
// LDS copy of the three hidden weight matrices: rows of kH bf16 padded to
// kLdsRow.  A lane reads 16 bytes of row 16 mt + c; unpadded (256-byte)
// rows put the rows that one read cycle serves on the same four banks, with
// 16 bytes of padding they cover every bank once.
constexpr int kLdsRow = kH + 8;
constexpr int kLdsLayer = kH * kLdsRow;
constexpr size_t kLdsWBytes = 3 * kLdsLayer * sizeof(__bf16);
constexpr size_t kLdsBytes = kLdsWBytes + kSmallFloats * sizeof(float);

__device__ __forceinline__ floatx4 mfma16(bf16x8 a, bf16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__host__ __device__ __forceinline__ int pi_neuron(int kk, int g, int j) {
  return 32 * kk + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}
// inverse of pi within a 32-block: neuron n -> logical k
__host__ __device__ __forceinline__ int pi_inverse(int n) {
  const int kk = n >> 5, r = n & 31;
  const int hi = r >> 4, q = r & 15;
  return 32 * kk + 8 * (q >> 2) + 4 * hi + (q & 3);
}

// SiLU and its derivative with the hardware reciprocal (about 1 ulp).
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigmoid(z);
  return s * (1.f + z * (1.f - s));
}

// Analytic base of the field: the network predicts the field minus this
// model (all zero: no base).  Up to kMaxMN Miyamoto-Nagai discs
// Phi = -M / sqrt(R^2 + (a + sqrt(z^2 + b^2))^2) and one Hernquist sphere
// Phi = -Mh / (r + ah), plus a constant; evaluated in float64.
struct MLPBase {
  int nmn;
  double M[kMaxMN], a[kMaxMN], b[kMaxMN];
  double mh, ah, phi0;
};

struct MLPCfg {
  float sxy;        // feature 0,1 = x/sxy, y/sxy
  float hz, wz;     // feature 2 = asinh(z/hz)/wz
  float os[4];      // physical output k = os[k] * y_k + ob[k] + base_k
  float ob[4];
  MLPBase base;
  int conservative; // forces are -grad of the potential output (mlp_tile_grad)
  float lw[4];      // training loss weight of output k
};

// (Phi, F_R, F_phi = 0, F_z) of the analytic base at (R, z).
__device__ __forceinline__ void mlp_base_eval(const MLPBase& B, float Rf, float zf, float out[4]) {
  const double R = Rf, z = zf, R2 = R * R, z2 = z * z;
  double pot = B.phi0, fR = 0.0, fz = 0.0;
  for (int k = 0; k < B.nmn; ++k) {
    const double bz = sqrt(z2 + B.b[k] * B.b[k]);
    const double s = B.a[k] + bz;
    const double id = 1.0 / sqrt(R2 + s * s);
    const double m3 = B.M[k] * id * id * id;
    pot -= B.M[k] * id;
    fR -= m3 * R;
    fz -= m3 * z * s / bz;
  }
  if (B.mh != 0.0) {
    const double r = sqrt(R2 + z2);
    const double ir = 1.0 / (r + B.ah);
    pot -= B.mh * ir;
    // -Mh / (r + a)^2 along r-hat; r = 0 gives no force
    const double g = r > 0.0 ? B.mh * ir * ir / r : 0.0;
    fR -= g * R;
    fz -= g * z;
  }
  out[0] = (float)pot;
  out[1] = (float)fR;
  out[2] = 0.f;
  out[3] = (float)fz;
}

__device__ __forceinline__ void mlp_features(const MLPCfg& c, float x, float y, float z, float f[3]) {
  f[0] = x / c.sxy;
  f[1] = y / c.sxy;
  f[2] = asinhf(z / c.hz) / c.wz;
}

struct MLPKArgs {
  const float* x;
  const float* y;
  const float* z;
  int64_t n;
  const float* target;   // [4][n] normalised targets (training)
  const float* weight;   // [n] per-point loss weight or NULL
  float* out;            // [4][n] physical outputs (inference)
  float* feat;           // [3][bcap]
  float* Z;              // [4][128][bcap] pre-activations of layers 1..4
  float* gZ;             // [4][128][bcap]
  float* gY;             // [4][bcap]
  int64_t bcap;
  const float* params;
  const __bf16* Wp;      // [3][128][128] forward A operands (K-permuted)
  const __bf16* Wt;      // [3][128][128] backward A operands (transposed, K-permuted)
  double* loss_part;     // [gridDim.x]
  float inv_n;
  MLPCfg cfg;
};

__device__ __forceinline__ void load_small(const float* __restrict__ params, float* s_small) {
  // W0, b0, b1, b2, b3, W4, b4 into LDS
  for (int t = threadIdx.x; t < kSmallFloats; t += blockDim.x) {
    int src;
    if (t < kH * kNIn + kH) src = t;                                       // W0, b0
    else if (t < kH * kNIn + 2 * kH) src = kOffB1 + (t - kH * kNIn - kH);
    else if (t < kH * kNIn + 3 * kH) src = kOffB2 + (t - kH * kNIn - 2 * kH);
    else if (t < kH * kNIn + 4 * kH) src = kOffB3 + (t - kH * kNIn - 3 * kH);
    else src = kOffW4 + (t - kH * kNIn - 4 * kH);                          // W4, b4
    s_small[t] = params[src];
  }
}

__device__ __forceinline__ void load_weights(const __bf16* __restrict__ W, __bf16* s_w) {
  constexpr int kRow16 = kH / 8, kLds16 = kLdsRow / 8;   // 16-byte chunks per row
  const uint4* src = reinterpret_cast<const uint4*>(W);
  uint4* dst = reinterpret_cast<uint4*>(s_w);
  for (int t = threadIdx.x; t < 3 * kH * kRow16; t += blockDim.x)
    dst[(t / kRow16) * kLds16 + t % kRow16] = src[t];
}

// Hidden layer: acc[mt][nt] = bias + W (K-permuted, LDS) x B[kk][nt].
template <int NT>
__device__ __forceinline__ void hidden_mfma(const __bf16* __restrict__ sW, const float* bias,
                                            const bf16x8 (&B)[4][NT], floatx4 (&acc)[8][NT],
                                            int g, int c) {
#pragma unroll
  for (int mt = 0; mt < 8; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[mt][nt][i] = bias ? bias[16 * mt + 4 * g + i] : 0.f;
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    // keep the operand loads of one K-step in flight at a time (register budget)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(sW + (16 * mt + c) * kLdsRow + 32 * kk + 8 * g);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma16(a, B[kk][nt], acc[mt][nt]);
    }
  }
}

// SiLU of hidden layer `layer`'s pre-activations acc: into the next layer's
// bf16 operands B, or (LAST) in place as float32 for the output layer.
// With TRAIN the pre-activations are stored to Z for points with valid[nt].
template <bool TRAIN, bool LAST>
__device__ __forceinline__ void forward_activate(floatx4 (&acc)[8][kNT], bf16x8 (&B)[4][kNT],
                                                 float* Z, int64_t bcap,
                                                 const int64_t (&pid)[kNT],
                                                 const bool (&valid)[kNT], int layer, int g) {
#pragma unroll
  for (int mt = 0; mt < 8; ++mt)
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float zz = acc[mt][nt][i];
        if (TRAIN && valid[nt])
          Z[((size_t)(layer + 1) * kH + 16 * mt + 4 * g + i) * bcap + pid[nt]] = zz;
        const float hh = silu(zz);
        if (LAST) acc[mt][nt][i] = hh;   // H4 in place (Z4 is reloaded in backward)
        else B[mt >> 1][nt][(mt & 1) * 4 + i] = (__bf16)hh;
      }
}

// Forward pass of one wave tile (kNT x 16 points, column c of lane group g):
// features f -> normalised outputs yk (all four lane groups of a column end
// with the same yk).  With TRAIN the pre-activations are stored to Z
// ([4][128][bcap]) for points with valid[nt].
template <bool TRAIN>
__device__ __forceinline__ void mlp_tile_forward(const __bf16* __restrict__ sW,
                                                 const float* __restrict__ sS,
                                                 const float (&f)[kNT][3], float (&yk)[kNT][4],
                                                 int g, int c, float* Z, int64_t bcap,
                                                 const int64_t (&pid)[kNT], const bool (&valid)[kNT]) {
  const float* sW0 = sS;
  const float* sB0 = sS + kH * kNIn;
  const float* sBh = sS + kH * kNIn + kH;           // b1, b2, b3
  const float* sW4 = sS + kH * kNIn + 4 * kH;
  const float* sB4 = sW4 + kNOut * kH;
  // layer 0 (float32): each lane evaluates the 32 neurons it feeds to layer 1
  bf16x8 B[4][kNT];
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int nr = pi_neuron(kk, g, j);
      const float w0 = sW0[nr * 3], w1 = sW0[nr * 3 + 1], w2 = sW0[nr * 3 + 2], b = sB0[nr];
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt) {
        const float zz = fmaf(w0, f[nt][0], fmaf(w1, f[nt][1], fmaf(w2, f[nt][2], b)));
        if (TRAIN && valid[nt]) Z[(size_t)nr * bcap + pid[nt]] = zz;
        B[kk][nt][j] = (__bf16)silu(zz);
      }
    }
  floatx4 acc[8][kNT];
  // layers 1 and 2 feed the next layer's operands; the last layer is peeled
  // so that no activation carries a branch on the layer
#pragma unroll 1
  for (int layer = 0; layer < 2; ++layer) {
    hidden_mfma(sW + layer * kLdsLayer, sBh + layer * kH, B, acc, g, c);
    forward_activate<TRAIN, false>(acc, B, Z, bcap, pid, valid, layer, g);
  }
  hidden_mfma(sW + 2 * kLdsLayer, sBh + 2 * kH, B, acc, g, c);
  forward_activate<TRAIN, true>(acc, B, Z, bcap, pid, valid, 2, g);
  // output layer (float32): partial sums over this lane's neurons, then
  // reduce over the four lane groups that hold the same point column
#pragma unroll
  for (int nt = 0; nt < kNT; ++nt)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float s = 0.f;
#pragma unroll
      for (int mt = 0; mt < 8; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) s = fmaf(sW4[k * kH + 16 * mt + 4 * g + i], acc[mt][nt][i], s);
      s += __shfl_xor(s, 16);
      s += __shfl_xor(s, 32);
      yk[nt][k] = s + sB4[k];
    }
}

// Forward pass of one wave tile for the potential output y0 and its
// derivatives dy0/df_k along the three features, by forward-mode tangents:
// t_1^k = silu'(z_1) W0[:, k], t_{l+1}^k = silu'(z_{l+1}) (W_l t_l^k), each
// W_l t_l^k one more bf16 MFMA pass over the layer's LDS weights (float32
// accumulation, tangents rounded to bf16 as the activations are), and
// dy0/df_k = W4[0] . t_4^k.  Layout as mlp_tile_forward; NT 16-point tiles.
//
// One hidden layer W of it: the value pass (activations into B, or LAST: the
// potential row of the output layer into ys) and the three tangent passes
// (into T, or LAST: into ds).
template <int NT, bool LAST>
__device__ __forceinline__ void grad_layer(const __bf16* __restrict__ W, const float* bias,
                                           const float* __restrict__ sW4, bf16x8 (&B)[4][NT],
                                           bf16x8 (&T)[3][4][NT], float (&ys)[NT],
                                           float (&ds)[NT][3], int g, int c) {
  // One sweep over W: each A fragment is read from LDS once and multiplied
  // into the activations (acc[0], from the bias) and the three tangents
  // (acc[1..3], from zero), 4 NT independent products per fragment.
  floatx4 acc[4][8][NT];
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    // one K-step's operand loads in flight at a time (register budget)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(W + (16 * mt + c) * kLdsRow + 32 * kk + 8 * g);
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        if (kk == 0) {
          floatx4 b0;
#pragma unroll
          for (int i = 0; i < 4; ++i) b0[i] = bias[16 * mt + 4 * g + i];
          acc[0][mt][nt] = mfma16(a, B[0][nt], b0);
#pragma unroll
          for (int k = 0; k < 3; ++k) acc[1 + k][mt][nt] = mfma16(a, T[k][0][nt], zero);
        } else {
          acc[0][mt][nt] = mfma16(a, B[kk][nt], acc[0][mt][nt]);
#pragma unroll
          for (int k = 0; k < 3; ++k) acc[1 + k][mt][nt] = mfma16(a, T[k][kk][nt], acc[1 + k][mt][nt]);
        }
      }
    }
  }
#pragma unroll
  for (int mt = 0; mt < 8; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float zz = acc[0][mt][nt][i];
        const float sg = sigmoid(zz);
        const float hh = zz * sg;
        const float dsl = sg * (1.f + zz * (1.f - sg));
        if (LAST) ys[nt] = fmaf(sW4[16 * mt + 4 * g + i], hh, ys[nt]);
        else B[mt >> 1][nt][(mt & 1) * 4 + i] = (__bf16)hh;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const float tt = dsl * acc[1 + k][mt][nt][i];
          if (LAST) ds[nt][k] = fmaf(sW4[16 * mt + 4 * g + i], tt, ds[nt][k]);
          else T[k][mt >> 1][nt][(mt & 1) * 4 + i] = (__bf16)tt;
        }
      }
}

template <int NT>
__device__ __forceinline__ void mlp_tile_grad(const __bf16* __restrict__ sW,
                                              const float* __restrict__ sS,
                                              const float (&f)[NT][3], float (&y0)[NT],
                                              float (&dy)[NT][3], int g, int c) {
  const float* sW0 = sS;
  const float* sB0 = sS + kH * kNIn;
  const float* sBh = sS + kH * kNIn + kH;
  const float* sW4 = sS + kH * kNIn + 4 * kH;
  const float* sB4 = sW4 + kNOut * kH;
  bf16x8 B[4][NT];
  bf16x8 T[3][4][NT];
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int nr = pi_neuron(kk, g, j);
      const float w0 = sW0[nr * 3], w1 = sW0[nr * 3 + 1], w2 = sW0[nr * 3 + 2], b = sB0[nr];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const float zz = fmaf(w0, f[nt][0], fmaf(w1, f[nt][1], fmaf(w2, f[nt][2], b)));
        const float sg = sigmoid(zz);
        const float d = sg * (1.f + zz * (1.f - sg));
        B[kk][nt][j] = (__bf16)(zz * sg);
        T[0][kk][nt][j] = (__bf16)(d * w0);
        T[1][kk][nt][j] = (__bf16)(d * w1);
        T[2][kk][nt][j] = (__bf16)(d * w2);
      }
    }
  float ys[NT], ds[NT][3];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    ys[nt] = 0.f;
    ds[nt][0] = ds[nt][1] = ds[nt][2] = 0.f;
  }
  // layers 1 and 2 feed the next layer's operands; the last layer is peeled
  // so that no activation carries a branch on the layer
#pragma unroll 1
  for (int layer = 0; layer < 2; ++layer)
    grad_layer<NT, false>(sW + layer * kLdsLayer, sBh + layer * kH, sW4, B, T, ys, ds, g, c);
  grad_layer<NT, true>(sW + 2 * kLdsLayer, sBh + 2 * kH, sW4, B, T, ys, ds, g, c);
  // sums over the four lane groups that hold the same point column
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    float v = ys[nt];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    y0[nt] = v + sB4[0];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float w = ds[nt][k];
      w += __shfl_xor(w, 16);
      w += __shfl_xor(w, 32);
      dy[nt][k] = w;
    }
  }
}

// Potential and Cartesian force of the conservative field at (x, y, z): the
// network's potential os0 y0 + ob0 plus the base, F = -grad of it (chain rule
// through the features f = (x/sxy, y/sxy, asinh(z/hz)/wz)).
__device__ __forceinline__ void conservative_field(const MLPCfg& cf, float x, float y, float z,
                                                   float y0, const float (&dy)[3], float& fx,
                                                   float& fy, float& fz, float& pot) {
  const float R = sqrtf(x * x + y * y);
  float bv[4];
  mlp_base_eval(cf.base, R, z, bv);
  const float cp = R > 0.f ? x / R : 1.f, sp = R > 0.f ? y / R : 0.f;
  const float zh = z / cf.hz;
  const float dfz = 1.f / (cf.wz * cf.hz * sqrtf(1.f + zh * zh));
  fx = -cf.os[0] * dy[0] / cf.sxy + bv[1] * cp;
  fy = -cf.os[0] * dy[1] / cf.sxy + bv[1] * sp;
  fz = -cf.os[0] * dy[2] * dfz + bv[3];
  pot = cf.os[0] * y0 + cf.ob[0] + bv[0];
}

// Forward pass over 32-point tiles per wave.  TRAIN also stores the
// features and pre-activations, the loss and dL/dy.
template <bool TRAIN>
__global__ __launch_bounds__(block_threads<TRAIN>()) void mlp_fwd_kernel(MLPKArgs a) {
  constexpr int kBlock = block_threads<TRAIN>(), kBlockPts = block_points<TRAIN>();
  extern __shared__ __align__(16) unsigned char smem[];
  __bf16* sW = reinterpret_cast<__bf16*>(smem);
  float* sS = reinterpret_cast<float*>(smem + kLdsWBytes);
  __shared__ double s_loss[kBlock / 64];

  load_small(a.params, sS);
  load_weights(a.Wp, sW);
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  double loss = 0.0;
  const int64_t ntiles = (a.n + kBlockPts - 1) / kBlockPts;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t base = tile * kBlockPts + wave * kWaveTile;
    int64_t pid[kNT];
    bool valid[kNT];
    float f[kNT][3];
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt) {
      pid[nt] = base + 16 * nt + c;
      valid[nt] = pid[nt] < a.n;
      const int64_t q = valid[nt] ? pid[nt] : 0;
      mlp_features(a.cfg, a.x[q], a.y[q], a.z[q], f[nt]);
      if (TRAIN && valid[nt] && g == 0)
        for (int d = 0; d < 3; ++d) a.feat[d * a.bcap + pid[nt]] = f[nt][d];
    }
    float yk[kNT][4];
    mlp_tile_forward<TRAIN>(sW, sS, f, yk, g, c, a.Z, a.bcap, pid, valid);
    if (!TRAIN) {
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt)
        if (valid[nt] && g == 0) {
          const float px = a.x[pid[nt]], py = a.y[pid[nt]];
          float bv[4];
          mlp_base_eval(a.cfg.base, sqrtf(px * px + py * py), a.z[pid[nt]], bv);
#pragma unroll
          for (int k = 0; k < 4; ++k)
            a.out[(size_t)k * a.n + pid[nt]] = a.cfg.os[k] * yk[nt][k] + a.cfg.ob[k] + bv[k];
        }
      continue;
    }
    // ---- loss and backward ------------------------------------------------
    float gy[kNT][4];
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt) {
      const float w = valid[nt] ? (a.weight ? a.weight[pid[nt]] : 1.f) : 0.f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float t = valid[nt] ? a.target[(size_t)k * a.n + pid[nt]] : 0.f;
        const float d = valid[nt] ? yk[nt][k] - t : 0.f;
        gy[nt][k] = 2.f * w * a.cfg.lw[k] * d * a.inv_n;
        if (g == 0) loss += (double)(w * a.cfg.lw[k] * d * d);
        if (valid[nt] && g == 0) a.gY[(size_t)k * a.bcap + pid[nt]] = gy[nt][k];
      }
    }
  }
  if (TRAIN) {
    // fixed-order block sum of the loss
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (lane == 0) s_loss[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) s += s_loss[w];
      a.loss_part[blockIdx.x] = s;
    }
  }
}

// Inference of the conservative field: out = Phi, F_R, F_phi, F_z with
// F = -grad Phi (mlp_tile_grad, 16-point tiles per wave).
__global__ __launch_bounds__(kBlockGrad) void mlp_fwd_grad_kernel(MLPKArgs a) {
  constexpr int NT = 1, kBlockPts = 16 * NT * (kBlockGrad / 64);
  extern __shared__ __align__(16) unsigned char smem[];
  __bf16* sW = reinterpret_cast<__bf16*>(smem);
  float* sS = reinterpret_cast<float*>(smem + kLdsWBytes);
  load_small(a.params, sS);
  load_weights(a.Wp, sW);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int64_t ntiles = (a.n + kBlockPts - 1) / kBlockPts;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t p = tile * kBlockPts + wave * 16 * NT + c;
    const bool valid = p < a.n;
    const int64_t q = valid ? p : 0;
    const float x = a.x[q], y = a.y[q], z = a.z[q];
    float f[NT][3];
    mlp_features(a.cfg, x, y, z, f[0]);
    float y0[NT], dy[NT][3];
    mlp_tile_grad<NT>(sW, sS, f, y0, dy, g, c);
    if (valid && g == 0) {
      float fx, fy, fz, pot;
      conservative_field(a.cfg, x, y, z, y0[0], dy[0], fx, fy, fz, pot);
      const float R = sqrtf(x * x + y * y);
      const float cp = R > 0.f ? x / R : 1.f, sp = R > 0.f ? y / R : 0.f;
      a.out[p] = pot;
      a.out[a.n + p] = fx * cp + fy * sp;
      a.out[2 * a.n + p] = fy * cp - fx * sp;
      a.out[3 * a.n + p] = fz;
    }
  }
}

// Backward chain: gZ4 = (W4^T dL/dy) silu'(Z4), then gZ_l = (W_l^T gZ_{l+1})
// silu'(Z_l) for l = 3, 2, 1, with the transposed K-permuted weights in LDS.
__global__ __launch_bounds__(kBlockTrain) void mlp_bwd_kernel(MLPKArgs a) {
  constexpr int kBlockPts = block_points<true>();
  extern __shared__ __align__(16) unsigned char smem[];
  __bf16* sW = reinterpret_cast<__bf16*>(smem);
  float* sS = reinterpret_cast<float*>(smem + kLdsWBytes);
  const float* sW4 = sS + kH * kNIn + 4 * kH;
  load_small(a.params, sS);
  load_weights(a.Wt, sW);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const int64_t ntiles = (a.n + kBlockPts - 1) / kBlockPts;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t base = tile * kBlockPts + wave * kWaveTile;
    int64_t pid[kNT];
    bool valid[kNT];
    float gy[kNT][4];
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt) {
      pid[nt] = base + 16 * nt + c;
      valid[nt] = pid[nt] < a.n;
#pragma unroll
      for (int k = 0; k < 4; ++k) gy[nt][k] = valid[nt] ? a.gY[(size_t)k * a.bcap + pid[nt]] : 0.f;
    }
    bf16x8 Bg[4][kNT];
#pragma unroll
    for (int mt = 0; mt < 8; ++mt)
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int nr = 16 * mt + 4 * g + i;
          float gh = 0.f;
#pragma unroll
          for (int k = 0; k < 4; ++k) gh = fmaf(sW4[k * kH + nr], gy[nt][k], gh);
          float gz = 0.f;
          if (valid[nt]) {
            const float z4 = a.Z[((size_t)3 * kH + nr) * a.bcap + pid[nt]];
            gz = gh * dsilu(z4);
            a.gZ[((size_t)3 * kH + nr) * a.bcap + pid[nt]] = gz;
          }
          Bg[mt >> 1][nt][(mt & 1) * 4 + i] = (__bf16)gz;
        }
    floatx4 acc[8][kNT];
#pragma unroll 1
    for (int layer = 2; layer >= 0; --layer) {
      hidden_mfma(sW + layer * kLdsLayer, nullptr, Bg, acc, g, c);
#pragma unroll
      for (int mt = 0; mt < 8; ++mt)
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int nr = 16 * mt + 4 * g + i;
            float gz = 0.f;
            if (valid[nt]) {
              const float zl = a.Z[((size_t)layer * kH + nr) * a.bcap + pid[nt]];
              gz = acc[mt][nt][i] * dsilu(zl);
              a.gZ[((size_t)layer * kH + nr) * a.bcap + pid[nt]] = gz;
            }
            Bg[mt >> 1][nt][(mt & 1) * 4 + i] = (__bf16)gz;
          }
    }
  }
}

// dW_l[o][i] = sum_p gZ_{l+1}[o][p] * silu(Z_l)[i][p] for hidden layers l = 1..3,
// one block per (chunk of kChunk points, layer); 4 waves x (2 M-tiles x 8 N-tiles).
constexpr int kChunk = 256;
__global__ __launch_bounds__(256) void mlp_dw_kernel(const float* __restrict__ Z,
                                                      const float* __restrict__ gZ, int64_t n,
                                                      int64_t bcap, float* __restrict__ part) {
  const int chunk = blockIdx.x, layer = blockIdx.y;   // layer 0..2 -> W1..W3
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const float* G = gZ + (size_t)(layer + 1) * kH * bcap;   // gZ_{l+1}
  const float* H = Z + (size_t)layer * kH * bcap;          // Z_l (silu applied on load)
  floatx4 acc[2][8];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int64_t p0 = (int64_t)chunk * kChunk;
  for (int ks = 0; ks < kChunk / 32; ++ks) {
    __builtin_amdgcn_sched_barrier(0);
    const int64_t pk = p0 + 32 * ks + 8 * g;
    bf16x8 A[2], Bv[8];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const float* src = G + (size_t)(16 * (2 * wave + a) + c) * bcap + pk;
#pragma unroll
      for (int j = 0; j < 8; ++j) A[a][j] = (__bf16)((pk + j < n) ? src[j] : 0.f);
    }
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const float* src = H + (size_t)(16 * b + c) * bcap + pk;
#pragma unroll
      for (int j = 0; j < 8; ++j) Bv[b][j] = (__bf16)((pk + j < n) ? silu(src[j]) : 0.f);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 8; ++b) acc[a][b] = mfma16(A[a], Bv[b], acc[a][b]);
  }
  float* out = part + ((size_t)layer * gridDim.x + chunk) * kH * kH;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        out[(size_t)(16 * (2 * wave + a) + 4 * g + i) * kH + 16 * b + c] = acc[a][b][i];
}

// grads[W_l] = sum over chunks (fixed order) of the partials.
__global__ void mlp_dw_reduce_kernel(const float* __restrict__ part, int nchunk,
                                     float* __restrict__ grads) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int layer = blockIdx.y;
  if (e >= kH * kH) return;
  const float* p = part + (size_t)layer * nchunk * kH * kH + e;
  double s = 0.0;
  for (int k = 0; k < nchunk; ++k) s += p[(size_t)k * kH * kH];
  grads[kHidOff[layer] + e] = (float)s;
}

// Gradients of W0, b0..b3, W4, b4: one block per parameter, fixed-order sums.
__global__ __launch_bounds__(256) void mlp_small_grads_kernel(const float* __restrict__ feat,
                                                              const float* __restrict__ Z,
                                                              const float* __restrict__ gZ,
                                                              const float* __restrict__ gY,
                                                              int64_t n, int64_t bcap,
                                                              float* __restrict__ grads) {
  __shared__ double sh[256];
  int q = blockIdx.x;  // 0..(384 + 512 + 512 + 4)
  const float* u = nullptr;
  const float* v = nullptr;
  bool silu_v = false;
  int dst;
  if (q < kH * kNIn) {                      // dW0[nr][d]
    const int nr = q / 3, d = q % 3;
    u = gZ + (size_t)nr * bcap;
    v = feat + (size_t)d * bcap;
    dst = kOffW0 + q;
  } else if ((q -= kH * kNIn) < 4 * kH) {   // db0..db3
    const int l = q / kH, nr = q % kH;
    u = gZ + ((size_t)l * kH + nr) * bcap;
    dst = (l == 0 ? kOffB0 : l == 1 ? kOffB1 : l == 2 ? kOffB2 : kOffB3) + nr;
  } else if ((q -= 4 * kH) < kNOut * kH) {  // dW4[k][nr]
    const int k = q / kH, nr = q % kH;
    u = gY + (size_t)k * bcap;
    v = Z + ((size_t)3 * kH + nr) * bcap;
    silu_v = true;
    dst = kOffW4 + q;
  } else {                                  // db4[k]
    q -= kNOut * kH;
    u = gY + (size_t)q * bcap;
    dst = kOffB4 + q;
  }
  double s = 0.0;
  for (int64_t p = threadIdx.x; p < n; p += 256) {
    double t = u[p];
    if (v) t *= silu_v ? (double)silu(v[p]) : (double)v[p];
    s += t;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) grads[dst] = (float)sh[0];
}

__global__ void loss_reduce_kernel(const double* __restrict__ part, int nb, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[b];
    out[0] = s;
  }
}

// bf16 operand copies of W1..W3: Wp[l][o][k] = W[o][pi(k)], Wt[l][i][k] = W[pi(k)][i].
__global__ void mlp_sync_kernel(const float* __restrict__ params, __bf16* __restrict__ Wp,
                                __bf16* __restrict__ Wt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int layer = blockIdx.y;
  if (e >= kH * kH) return;
  const int o = e / kH, i = e % kH;
  const __bf16 w = (__bf16)params[kHidOff[layer] + e];
  Wp[(size_t)layer * kH * kH + o * kH + pi_inverse(i)] = w;
  Wt[(size_t)layer * kH * kH + i * kH + pi_inverse(o)] = w;
}

struct AdamArgs {
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
};

__global__ void mlp_adam_kernel(float* __restrict__ params, const float* __restrict__ grads,
                                float* __restrict__ m, float* __restrict__ v, AdamArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kNParams) return;
  const float gr = grads[e];
  const float mm = a.b1 * m[e] + (1.f - a.b1) * gr;
  const float vv = a.b2 * v[e] + (1.f - a.b2) * gr * gr;
  m[e] = mm;
  v[e] = vv;
  params[e] -= a.lr * (mm / a.c1) / (sqrtf(vv / a.c2) + a.eps);
}


// ---------------------------------------------------------------------------
// Orbit integration with the surrogate (integrate.py:129-143 update rule).
// ---------------------------------------------------------------------------
struct MLPOrbitAux {
  float ax, ay, az, maxerr;   // current acceleration, largest checked force error
  double e0, drift;           // initial energy (BFE potential), max |E - E0| / |E0|
  int nfallback, nchecked;    // BFE fallback evaluations, checked evaluations
};

struct MLPLeapArgs {
  float *x, *y, *z, *vx, *vy, *vz;
  int64_t n;
  double dt, t0, omega;
  int64_t k0, k1;             // steps (k0, k1] are taken
  int init;                   // 1: compute a0 and E0 at step k0 only
  int energy_check;           // 1: compare the energy at k1 with E0
  int check_period;           // every step 1 wave in check_period is checked against the BFE
  float rmax2, zmax;          // training domain R^2 < rmax2, |z| < zmax
  MLPOrbitAux* aux;
  double* partE;              // [checks][nblocks] block energy sums, or NULL
  int check_index, nblocks;
  const float* params;
  const __bf16* Wp;
  MLPCfg cfg;
  FieldTables T;
};

// The plain-BFE evaluation on the surrogate's rare paths (fallback, checks,
// energies) is an out-of-line call so that its float64 register demand does
// not add to the MLP loop's.
#ifndef BFE_MLP_INLINE_BFE
#define BFE_MLP_INLINE_BFE 0
#endif
#if BFE_MLP_INLINE_BFE
__device__ __forceinline__
#else
__device__ __noinline__
#endif
void bfe_point_call(const FieldTables& T, double x, double y, double z, double* cv) {
  field_eval_point(T, x, y, z, cv);
}

// Cartesian acceleration and potential of orbit (x, y, z) from the
// surrogate outputs, or from the BFE when the point is outside the domain.
// GRAD: yv holds (y0, dy0/df_0..2) of mlp_tile_grad, the forces are -grad.
template <bool GRAD>
__device__ __forceinline__ void surrogate_accel(const MLPLeapArgs& a, FieldTables& T, float x,
                                                float y, float z, const float (&yv)[4], float& fx,
                                                float& fy, float& fz, float& pot, bool& fb) {
  const float R2 = x * x + y * y;
  fb = !(R2 < a.rmax2 && fabsf(z) < a.zmax);
  if (fb) {
    double cv[8];
    bfe_point_call(T, x, y, z, cv);
    fx = (float)(cv[0] + cv[1]);
    fy = (float)(cv[2] + cv[3]);
    fz = (float)(cv[4] + cv[5]);
    pot = (float)(cv[6] + cv[7]);
    return;
  }
  if constexpr (GRAD) {
    const float dy[3] = {yv[1], yv[2], yv[3]};
    conservative_field(a.cfg, x, y, z, yv[0], dy, fx, fy, fz, pot);
    return;
  }
  const float R = sqrtf(R2);
  float bv[4];
  mlp_base_eval(a.cfg.base, R, z, bv);
  const float FR = a.cfg.os[1] * yv[1] + a.cfg.ob[1] + bv[1];
  const float Fp = a.cfg.os[2] * yv[2] + a.cfg.ob[2];
  const float cp = R > 0.f ? x / R : 1.f, sp = R > 0.f ? y / R : 0.f;
  fx = FR * cp - Fp * sp;
  fy = FR * sp + Fp * cp;
  fz = a.cfg.os[3] * yv[3] + a.cfg.ob[3] + bv[3];
  pot = a.cfg.os[0] * yv[0] + a.cfg.ob[0] + bv[0];
}

template <bool GRAD>
__global__ __launch_bounds__(GRAD ? kBlockGrad : kBlockInf) void mlp_leapfrog_kernel(MLPLeapArgs a) {
  // GRAD: conservative forces from mlp_tile_grad, 16-point tiles
  constexpr int NT = GRAD ? 1 : kNT;
  constexpr int kBlock = GRAD ? kBlockGrad : kBlockInf, kBlockPts = 16 * NT * (kBlock / 64);
  extern __shared__ __align__(16) unsigned char smem[];
  __bf16* sW = reinterpret_cast<__bf16*>(smem);
  float* sS = reinterpret_cast<float*>(smem + kLdsWBytes);
  __shared__ double s_e[kBlock / 64];
  load_small(a.params, sS);
  load_weights(a.Wp, sW);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  const bool owner = g < NT;                       // lane (g, c) owns orbit 16 g + c of the wave
  const int64_t ntiles = (a.n + kBlockPts - 1) / kBlockPts;
  double esum = 0.0;
  FieldTables T = a.T;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t base = tile * kBlockPts + wave * 16 * NT;
    const int64_t me = base + 16 * (owner ? g : 0) + c;
    const bool live = owner && me < a.n;
    float px = 0, py = 0, pz = 0, ux = 0, uy = 0, uz = 0;
    MLPOrbitAux o{};
    if (live) {
      px = a.x[me]; py = a.y[me]; pz = a.z[me];
      ux = a.vx[me]; uy = a.vy[me]; uz = a.vz[me];
      if (!a.init) o = a.aux[me];
    }
    const float dt = (float)a.dt, hdt = 0.5f * dt, hdt2 = 0.5f * dt * dt;
    const int64_t gwave = tile * (kBlock / 64) + wave;
    const int64_t kstart = a.init ? a.k0 : a.k0 + 1;
    for (int64_t k = kstart; k <= a.k1; ++k) {
      if (!a.init && live) {
        px = px + ux * dt + o.ax * hdt2;
        py = py + uy * dt + o.ay * hdt2;
        pz = pz + uz * dt + o.az * hdt2;
      }
      // surrogate forward for the wave's points (column c, tiles nt)
      float f[NT][3];
      int64_t pid[NT];
      bool valid[NT];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const float qx = __shfl(px, 16 * nt + c), qy = __shfl(py, 16 * nt + c), qz = __shfl(pz, 16 * nt + c);
        mlp_features(a.cfg, qx, qy, qz, f[nt]);
        pid[nt] = 0;
        valid[nt] = false;
      }
      float yk[NT][4];
      if constexpr (GRAD) {
        float y0[NT], dy[NT][3];
        mlp_tile_grad<NT>(sW, sS, f, y0, dy, g, c);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          yk[nt][0] = y0[nt];
          yk[nt][1] = dy[nt][0];
          yk[nt][2] = dy[nt][1];
          yk[nt][3] = dy[nt][2];
        }
      } else {
        mlp_tile_forward<false>(sW, sS, f, yk, g, c, nullptr, 0, pid, valid);
      }
      if (live) {
        float yv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          yv[q] = yk[0][q];
#pragma unroll
          for (int nt = 1; nt < NT; ++nt)
            if (g == nt) yv[q] = yk[nt][q];
        }
        if (a.omega != 0.0) {
          const double ang = a.omega * (a.t0 + (double)k * a.dt);
          T.crot = cos(ang);
          T.srot = sin(ang);
        }
        float nx, ny, nz, pot;
        bool fb;
        surrogate_accel<GRAD>(a, T, px, py, pz, yv, nx, ny, nz, pot, fb);
        if (fb) o.nfallback += 1;
        if (!fb && a.check_period > 0 && ((gwave + k) % a.check_period) == 0) {
          double cv[8];
          bfe_point_call(T, px, py, pz, cv);
          const double bx = cv[0] + cv[1], by = cv[2] + cv[3], bz = cv[4] + cv[5];
          const double err = sqrt((nx - bx) * (nx - bx) + (ny - by) * (ny - by) + (nz - bz) * (nz - bz)) /
                             sqrt(bx * bx + by * by + bz * bz);
          if (err > o.maxerr || !(err == err)) o.maxerr = (float)err;
          o.nchecked += 1;
        }
        if (a.init) {
          o.ax = nx; o.ay = ny; o.az = nz;
        } else {
          ux = ux + (o.ax + nx) * hdt;
          uy = uy + (o.ay + ny) * hdt;
          uz = uz + (o.az + nz) * hdt;
          o.ax = nx; o.ay = ny; o.az = nz;
        }
      }
    }
    if (live) {
      // energy with the BFE potential (the reference for measuring drift)
      if (a.init || a.energy_check) {
        if (a.omega != 0.0) {
          const double ang = a.omega * (a.t0 + (double)a.k1 * a.dt);
          T.crot = cos(ang);
          T.srot = sin(ang);
        }
        double cv[8];
        bfe_point_call(T, px, py, pz, cv);
        const double e = 0.5 * ((double)ux * ux + (double)uy * uy + (double)uz * uz) + (cv[6] + cv[7]);
        if (a.init) {
          o.e0 = e;
          o.drift = 0.0;
        } else {
          const double rel = fabs(e - o.e0) / fabs(o.e0);
          if (rel > o.drift || !(rel == rel)) o.drift = rel;
        }
        esum += e;
      }
      a.x[me] = px; a.y[me] = py; a.z[me] = pz;
      a.vx[me] = ux; a.vy[me] = uy; a.vz[me] = uz;
      a.aux[me] = o;
    }
  }
  if (a.partE) {
    for (int off = 32; off > 0; off >>= 1) esum += __shfl_xor(esum, off);
    if (lane == 0) s_e[wave] = esum;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sum = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) sum += s_e[w];
      a.partE[(size_t)a.check_index * a.nblocks + blockIdx.x] = sum;
    }
  }
}

__global__ void mlp_orbit_stats_kernel(const MLPOrbitAux* __restrict__ aux, int64_t n,
                                       double* __restrict__ drift, float* __restrict__ maxerr,
                                       int* __restrict__ counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (drift) drift[i] = aux[i].drift;
  if (maxerr) maxerr[i] = aux[i].maxerr;
  if (counts) {
    counts[2 * i] = aux[i].nfallback;
    counts[2 * i + 1] = aux[i].nchecked;
  }
}

__global__ void mlp_energy_reduce_kernel(const double* __restrict__ partE, int nchecks, int nblocks,
                                         double* __restrict__ log_out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchecks) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += partE[(size_t)c * nblocks + b];
  log_out[c] = s;
}

int field_tables(Context* c, double rotpos, FieldTables& T);

}  // namespace bfe

using namespace bfe;

namespace {

MLPCfg make_cfg(const MLPState& M) {
  const double* cfg = M.cfg;
  MLPCfg c;
  c.sxy = (float)cfg[0];
  c.hz = (float)cfg[1];
  c.wz = (float)cfg[2];
  for (int k = 0; k < 4; ++k) {
    c.os[k] = (float)cfg[3 + k];
    c.ob[k] = (float)cfg[7 + k];
  }
  // base layout: nmn, (M, a, b) x nmn, Mh, ah, phi0
  const double* b = M.base;
  c.base.nmn = (int)b[0];
  for (int k = 0; k < kMaxMN; ++k) {
    const bool on = k < c.base.nmn;
    c.base.M[k] = on ? b[1 + 3 * k] : 0.0;
    c.base.a[k] = on ? b[2 + 3 * k] : 0.0;
    c.base.b[k] = on ? b[3 + 3 * k] : 1.0;
  }
  const int o = 1 + 3 * c.base.nmn;
  c.base.mh = b[o];
  c.base.ah = b[o + 1];
  c.base.phi0 = b[o + 2];
  c.conservative = M.conservative;
  for (int k = 0; k < 4; ++k) c.lw[k] = (float)M.lw[k];
  return c;
}

int mlp_ready(Context* c) {
  if (!c->mlp.params) return set_error(c, BFE_ENOTABLE, "mlp: parameters not set (bfe_mlp_set_params)");
  return 0;
}

int mlp_batch_buffers(Context* c, int64_t n, hipStream_t s) {
  MLPState& M = c->mlp;
  if (n <= M.bcap) return 0;
  BFE_HIP(c, hipStreamSynchronize(s));
  for (float** p : {&M.feat, &M.Z, &M.gZ, &M.gY, &M.dwpart}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  if (M.loss_part) (void)hipFree(M.loss_part);
  M.loss_part = nullptr;
  const int64_t cap = ((n + kChunk - 1) / kChunk) * kChunk;
  const int64_t nchunk = cap / kChunk;
  BFE_HIP(c, hipMalloc(&M.feat, sizeof(float) * 3 * cap));
  BFE_HIP(c, hipMalloc(&M.Z, sizeof(float) * 4 * kH * cap));
  BFE_HIP(c, hipMalloc(&M.gZ, sizeof(float) * 4 * kH * cap));
  BFE_HIP(c, hipMalloc(&M.gY, sizeof(float) * 4 * cap));
  BFE_HIP(c, hipMalloc(&M.dwpart, sizeof(float) * 3 * nchunk * kH * kH));
  BFE_HIP(c, hipMalloc(&M.loss_part, sizeof(double) * ((cap + block_points<true>() - 1) / block_points<true>())));
  M.bcap = cap;
  return 0;
}

MLPKArgs base_args(Context* c) {
  MLPKArgs a{};
  a.params = c->mlp.params;
  a.Wp = c->mlp.Wp;
  a.Wt = c->mlp.Wt;
  a.cfg = make_cfg(c->mlp);
  return a;
}

// Forward + backward + gradient assembly for n points (n <= bcap).
int mlp_gradient(Context* c, const float* x, const float* y, const float* z, const float* target,
                 const float* weight, int64_t n, float* grads, double* loss_out, hipStream_t s) {
  int rc = mlp_batch_buffers(c, n, s);
  if (rc) return rc;
  MLPState& M = c->mlp;
  MLPKArgs a = base_args(c);
  a.x = x; a.y = y; a.z = z; a.n = n;
  a.target = target;
  a.weight = weight;
  a.feat = M.feat; a.Z = M.Z; a.gZ = M.gZ; a.gY = M.gY; a.bcap = M.bcap;
  a.loss_part = M.loss_part;
  a.inv_n = 1.f / (float)n;
  const int nb = (int)((n + block_points<true>() - 1) / block_points<true>());
  hipLaunchKernelGGL(mlp_fwd_kernel<true>, dim3(nb), dim3(kBlockTrain), kLdsBytes, s, a);
  hipLaunchKernelGGL(mlp_bwd_kernel, dim3(nb), dim3(kBlockTrain), kLdsBytes, s, a);
  BFE_HIP(c, hipGetLastError());
  const int nchunk = (int)((n + kChunk - 1) / kChunk);
  hipLaunchKernelGGL(mlp_dw_kernel, dim3(nchunk, 3), dim3(256), 0, s, M.Z, M.gZ, n, M.bcap, M.dwpart);
  hipLaunchKernelGGL(mlp_dw_reduce_kernel, dim3((kH * kH + 255) / 256, 3), dim3(256), 0, s, M.dwpart,
                     nchunk, grads);
  hipLaunchKernelGGL(mlp_small_grads_kernel, dim3(kH * kNIn + 4 * kH + kNOut * kH + kNOut), dim3(256),
                     0, s, M.feat, M.Z, M.gZ, M.gY, n, M.bcap, grads);
  if (loss_out) hipLaunchKernelGGL(loss_reduce_kernel, dim3(1), dim3(64), 0, s, M.loss_part, nb, loss_out);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

int mlp_sync(Context* c, hipStream_t s) {
  hipLaunchKernelGGL(mlp_sync_kernel, dim3((kH * kH + 255) / 256, 3), dim3(256), 0, s,
                     c->mlp.params, c->mlp.Wp, c->mlp.Wt);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

}  // namespace

namespace bfe {
void mlp_free(Context* c) {
  MLPState& M = c->mlp;
  for (float** p : {&M.params, &M.grads, &M.m, &M.v, &M.feat, &M.Z, &M.gZ, &M.gY, &M.dwpart}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  if (M.Wp) (void)hipFree(M.Wp);
  if (M.Wt) (void)hipFree(M.Wt);
  if (M.loss_part) (void)hipFree(M.loss_part);
  M = MLPState();
}
}  // namespace bfe

extern "C" int bfe_mlp_set_params(bfe_ctx* h, const float* params, int64_t count,
                                  const double* cfg) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (count != kNParams) return set_error(c, BFE_EINVAL, "mlp: %lld params given, %d expected", (long long)count, kNParams);
  if (!params || !cfg) return set_error(c, BFE_EINVAL, "mlp: NULL params or cfg");
  if (!(cfg[0] > 0) || !(cfg[1] > 0) || !(cfg[2] > 0)) return set_error(c, BFE_EINVAL, "mlp: feature scales must be > 0");
  BFE_HIP(c, hipSetDevice(c->device));
  MLPState& M = c->mlp;
  if (!M.params) {
    BFE_HIP(c, hipMalloc(&M.params, sizeof(float) * kNParams));
    BFE_HIP(c, hipMalloc(&M.grads, sizeof(float) * kNParams));
    BFE_HIP(c, hipMalloc(&M.m, sizeof(float) * kNParams));
    BFE_HIP(c, hipMalloc(&M.v, sizeof(float) * kNParams));
    BFE_HIP(c, hipMalloc(&M.Wp, sizeof(__bf16) * 3 * kH * kH));
    BFE_HIP(c, hipMalloc(&M.Wt, sizeof(__bf16) * 3 * kH * kH));
    BFE_HIP(c, hipFuncSetAttribute((const void*)mlp_fwd_kernel<true>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes));
    BFE_HIP(c, hipFuncSetAttribute((const void*)mlp_fwd_kernel<false>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes));
    BFE_HIP(c, hipFuncSetAttribute((const void*)mlp_bwd_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes));
    BFE_HIP(c, hipFuncSetAttribute((const void*)mlp_leapfrog_kernel<false>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes));
    BFE_HIP(c, hipFuncSetAttribute((const void*)mlp_leapfrog_kernel<true>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes));
    BFE_HIP(c, hipFuncSetAttribute((const void*)mlp_fwd_grad_kernel,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBytes));
  }
  BFE_HIP(c, hipMemcpy(M.params, params, sizeof(float) * kNParams, hipMemcpyHostToDevice));
  BFE_HIP(c, hipMemset(M.m, 0, sizeof(float) * kNParams));
  BFE_HIP(c, hipMemset(M.v, 0, sizeof(float) * kNParams));
  for (int k = 0; k < 11; ++k) M.cfg[k] = cfg[k];
  M.step = 0;
  int rc = mlp_sync(c, nullptr);
  if (rc) return rc;
  BFE_HIP(c, hipDeviceSynchronize());
  return 0;
}

extern "C" int bfe_mlp_set_base(bfe_ctx* h, const double* base, int count) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (!base || count < 1) return set_error(c, BFE_EINVAL, "mlp_set_base: NULL or empty base");
  const int nmn = (int)base[0];
  if (base[0] != (double)nmn || nmn < 0 || nmn > kMaxMN)
    return set_error(c, BFE_EINVAL, "mlp_set_base: %g Miyamoto-Nagai components (0..%d)", base[0], kMaxMN);
  if (count != 4 + 3 * nmn)
    return set_error(c, BFE_EINVAL, "mlp_set_base: %d values given, %d expected", count, 4 + 3 * nmn);
  for (int k = 0; k < count; ++k)
    if (!std::isfinite(base[k])) return set_error(c, BFE_EINVAL, "mlp_set_base: non-finite value");
  for (int k = 0; k < nmn; ++k)
    if (!(base[3 + 3 * k] > 0.0)) return set_error(c, BFE_EINVAL, "mlp_set_base: b must be > 0");
  if (base[1 + 3 * nmn] != 0.0 && !(base[2 + 3 * nmn] > 0.0))
    return set_error(c, BFE_EINVAL, "mlp_set_base: Hernquist scale must be > 0");
  MLPState& M = c->mlp;
  for (int k = 0; k < kMLPBaseMax; ++k) M.base[k] = k < count ? base[k] : 0.0;
  return 0;
}

extern "C" int bfe_mlp_set_mode(bfe_ctx* h, int conservative, const double* loss_weights) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (conservative != 0 && conservative != 1)
    return set_error(c, BFE_EINVAL, "mlp_set_mode: conservative must be 0 or 1");
  MLPState& M = c->mlp;
  for (int k = 0; k < 4; ++k) {
    const double w = loss_weights ? loss_weights[k] : 1.0;
    if (!(w >= 0.0) || !std::isfinite(w)) return set_error(c, BFE_EINVAL, "mlp_set_mode: loss weights must be >= 0");
    M.lw[k] = w;
  }
  M.conservative = conservative;
  return 0;
}

extern "C" int bfe_mlp_get_params(bfe_ctx* h, float* params, int64_t count) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (int rc = mlp_ready(c)) return rc;
  if (count != kNParams || !params) return set_error(c, BFE_EINVAL, "mlp_get_params: bad buffer");
  BFE_HIP(c, hipSetDevice(c->device));
  BFE_HIP(c, hipDeviceSynchronize());
  BFE_HIP(c, hipMemcpy(params, c->mlp.params, sizeof(float) * kNParams, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int bfe_mlp_forward(bfe_ctx* h, const float* x, const float* y, const float* z,
                               int64_t n, float* out, void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (int rc = mlp_ready(c)) return rc;
  if (n < 0) return set_error(c, BFE_EINVAL, "mlp_forward: n < 0");
  if (n == 0) return 0;
  if (!x || !y || !z || !out) return set_error(c, BFE_EINVAL, "mlp_forward: NULL pointer");
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  MLPKArgs a = base_args(c);
  a.x = x; a.y = y; a.z = z; a.n = n; a.out = out;
  if (c->mlp.conservative) {
    const int64_t ntiles = (n + 16 * (kBlockGrad / 64) - 1) / (16 * (kBlockGrad / 64));
    const int nb = (int)(ntiles < 2048 ? ntiles : 2048);
    hipLaunchKernelGGL(mlp_fwd_grad_kernel, dim3(nb), dim3(kBlockGrad), kLdsBytes, s, a);
    BFE_HIP(c, hipGetLastError());
    return 0;
  }
  const int64_t ntiles = (n + block_points<false>() - 1) / block_points<false>();
  const int nb = (int)(ntiles < 1024 ? ntiles : 1024);
  hipLaunchKernelGGL(mlp_fwd_kernel<false>, dim3(nb), dim3(kBlockInf), kLdsBytes, s, a);
  BFE_HIP(c, hipGetLastError());
  return 0;
}

extern "C" int bfe_mlp_gradient(bfe_ctx* h, const float* x, const float* y, const float* z,
                                const float* target, const float* weight, int64_t n,
                                float* grad_out, double* loss_out, void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (int rc = mlp_ready(c)) return rc;
  if (n <= 0 || !x || !y || !z || !target || !grad_out)
    return set_error(c, BFE_EINVAL, "mlp_gradient: bad arguments");
  BFE_HIP(c, hipSetDevice(c->device));
  return mlp_gradient(c, x, y, z, target, weight, n, grad_out, loss_out, (hipStream_t)stream);
}

extern "C" int bfe_mlp_train_step(bfe_ctx* h, const float* x, const float* y, const float* z,
                                  const float* target, const float* weight, int64_t n, float lr,
                                  float beta1, float beta2, float eps, double* loss_out,
                                  void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (int rc = mlp_ready(c)) return rc;
  if (n <= 0 || !x || !y || !z || !target) return set_error(c, BFE_EINVAL, "mlp_train_step: bad arguments");
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  MLPState& M = c->mlp;
  int rc = mlp_gradient(c, x, y, z, target, weight, n, M.grads, loss_out, s);
  if (rc) return rc;
  M.step += 1;
  AdamArgs a;
  a.lr = lr; a.b1 = beta1; a.b2 = beta2; a.eps = eps;
  a.c1 = (float)(1.0 - pow((double)beta1, (double)M.step));
  a.c2 = (float)(1.0 - pow((double)beta2, (double)M.step));
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kNParams + 255) / 256), dim3(256), 0, s, M.params, M.grads, M.m, M.v, a);
  BFE_HIP(c, hipGetLastError());
  return mlp_sync(c, s);
}

extern "C" int bfe_mlp_leapfrog(bfe_ctx* h, float* x, float* y, float* z, float* vx, float* vy,
                                float* vz, int64_t n, double dt, int64_t nsteps, double t0,
                                double rotfreq, int energy_every, float rmax, float zmax,
                                int check_period, double* drift_out, float* maxerr_out,
                                int* counts_out, double* energy_log, void* stream) {
  if (!h) return BFE_EINVAL;
  Context* c = &h->c;
  if (int rc = mlp_ready(c)) return rc;
  if (n < 0 || nsteps < 0 || energy_every < 0 || check_period < 0)
    return set_error(c, BFE_EINVAL, "mlp_leapfrog: negative size");
  if (n > 0 && (!x || !y || !z || !vx || !vy || !vz)) return set_error(c, BFE_EINVAL, "mlp_leapfrog: NULL state");
  if (energy_log && energy_every == 0) return set_error(c, BFE_EINVAL, "mlp_leapfrog: energy_log needs energy_every > 0");
  FieldTables T;
  int rc = field_tables(c, 0.0, T);
  if (rc) return rc;
  if (n == 0) return 0;
  BFE_HIP(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  const int64_t ntiles = (n + block_points<false>() - 1) / block_points<false>();
  const int nb = (int)(ntiles < 2048 ? ntiles : 2048);
  const int chunk = energy_every > 0 ? energy_every : (int)(nsteps < 100 ? (nsteps > 0 ? nsteps : 1) : 100);
  const int nchecks = energy_every > 0 ? (int)(nsteps / energy_every) + 1 : 1;
  const size_t aux_bytes = (size_t)n * sizeof(MLPOrbitAux);
  const size_t part_bytes = (size_t)nchecks * nb * sizeof(double);
  rc = ensure_scratch(c, ((aux_bytes + 255) & ~(size_t)255) + part_bytes);
  if (rc) return rc;
  MLPOrbitAux* aux = (MLPOrbitAux*)c->scratch;
  double* partE = (double*)((char*)c->scratch + ((aux_bytes + 255) & ~(size_t)255));
  MLPLeapArgs a{};
  a.x = x; a.y = y; a.z = z; a.vx = vx; a.vy = vy; a.vz = vz;
  a.n = n;
  a.dt = dt; a.t0 = t0; a.omega = 2.0 * kPi * rotfreq;
  a.check_period = check_period;
  a.rmax2 = rmax * rmax;
  a.zmax = zmax;
  a.aux = aux;
  a.nblocks = nb;
  a.params = c->mlp.params;
  a.Wp = c->mlp.Wp;
  a.cfg = make_cfg(c->mlp);
  a.T = T;
  // step 0: accelerations and initial energies
  a.init = 1;
  a.k0 = 0;
  a.k1 = 0;
  a.energy_check = 1;
  a.partE = energy_log ? partE : nullptr;
  a.check_index = 0;
  if (a.cfg.conservative)
    hipLaunchKernelGGL(mlp_leapfrog_kernel<true>, dim3(nb), dim3(kBlockGrad), kLdsBytes, s, a);
  else
    hipLaunchKernelGGL(mlp_leapfrog_kernel<false>, dim3(nb), dim3(kBlockInf), kLdsBytes, s, a);
  BFE_HIP(c, hipGetLastError());
  a.init = 0;
  for (int64_t k0 = 0, ci = 0; k0 < nsteps; k0 += chunk, ++ci) {
    a.k0 = k0;
    a.k1 = k0 + chunk < nsteps ? k0 + chunk : nsteps;
    a.energy_check = (energy_every > 0 && a.k1 - a.k0 == energy_every) ? 1 : 0;
    a.partE = (energy_log && a.energy_check) ? partE : nullptr;
    a.check_index = (int)ci + 1;
    if (a.cfg.conservative)
    hipLaunchKernelGGL(mlp_leapfrog_kernel<true>, dim3(nb), dim3(kBlockGrad), kLdsBytes, s, a);
  else
    hipLaunchKernelGGL(mlp_leapfrog_kernel<false>, dim3(nb), dim3(kBlockInf), kLdsBytes, s, a);
    BFE_HIP(c, hipGetLastError());
  }
  if (drift_out || maxerr_out || counts_out)
    hipLaunchKernelGGL(mlp_orbit_stats_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       aux, n, drift_out, maxerr_out, counts_out);
  if (energy_log)
    hipLaunchKernelGGL(mlp_energy_reduce_kernel, dim3((nchecks + 63) / 64), dim3(64), 0, s, partE,
                       nchecks, nb, energy_log);
  BFE_HIP(c, hipGetLastError());
  return 0;
}
