// mlp_device.h -- the learned pressure guess of the con2prim path: a
// 3 -> 64 -> 64 -> 1 ReLU network evaluated one wavefront (64 points) at a
// time.  Layer 1 (K = 3) runs on the FP32 vector units, one point per lane;
// layer 2 is a 64x64 bf16 contraction on the matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulate) computed as
// Z2^T[f][pt] = W2[f][:] . H1[pt][:], so that each lane ends with whole
// columns (points) and layer 3 needs a single cross-lane add.  No reference
// counterpart: this is the learned initial guess added by this library.
//
// MFMA lane layout (v_mfma_f32_32x32x16_bf16, one wave64):
//   A[i][k]: i = lane % 32, k = 8*(lane/32) + e   (e = 0..7 of the bf16x8)
//   B[k][j]: j = lane % 32, k = 8*(lane/32) + e
//   C[i][j]: j = lane % 32, i = 8*(v/4) + 4*(lane/32) + v%4   (v = 0..15)
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/bamhip.h"

namespace bamhip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kMlpH = 64;   // hidden width
// parameter vector: W1[64][3], b1[64], W2[64][64] ([out][in]), b2[64], W3[64], b3
constexpr int kOffW1 = 0, kOffB1 = 192, kOffW2 = 256, kOffB2 = 4352, kOffW3 = 4416, kOffB3 = 4480;
constexpr int kMlpNParam = 4481;
constexpr int kHStr = 72;   // row stride (bf16) of the per-wave activation tiles in LDS

struct MlpView {
  const float *p;       // FP32 master parameters (biases are read from here)
  const float *wq;      // W1 (192) and W3 (64) rounded to bf16, stored as float
  const __bf16 *w2;     // W2 [f][k] in bf16
  const __bf16 *w2t;    // W2^T [k][f] in bf16
  float xs[3], xk[3];   // input normalisation: (x - xs) * xk
  float ys, yk;         // output: log(p / p_hi) = ys + yk * y_net
};

// device view of a model (mlp.hip)
MlpView mlp_view(const bamhip_mlp *m);

__device__ __forceinline__ int mfma_row(int v, int lane) { return 8 * (v / 4) + 4 * (lane / 32) + (v % 4); }

// Thermal content of the conserved state, Delta = tau + D - sqrt(S^2 + D^2):
// zero for a cold flow (p = eps = 0 gives tau + D = W D and S = W D v), so
// Delta / (tau + D) carries, without cancellation, the information that
// fixes p.  Evaluated in FP64; the network sees only well-conditioned
// logarithms of it, and predicts log(p / ((Gamma-1) Delta)), which lies in
// [0, ~1.1] for the ideal gas.
__device__ __forceinline__ double mlp_thermal(double D, double S2, double T) { return T + D - sqrt(S2 + D * D); }

// raw features log D, |S| / (tau + D), log(Delta / (tau + D)): the ratios
// are formed in FP64, the logarithms in FP32 (the network's precision).
// Shared by the training-data and the inference paths so that both see the
// same values; false if the state has no thermal content or a feature is
// not finite.
__device__ __forceinline__ bool mlp_raw_features(double D, double S2, double T, double th, float r[3]) {
  r[0] = __logf((float)D);
  r[1] = (float)(sqrt(S2) / (T + D));
  r[2] = __logf((float)(th / (T + D)));
  return D > 0. && th > 0. && isfinite(r[0]) && isfinite(r[1]) && isfinite(r[2]);
}

// normalised features of a state
__device__ __forceinline__ bool mlp_features(const MlpView &M, double D, double S2, double T, double th, float x[3]) {
  float r[3];
  const bool ok = mlp_raw_features(D, S2, T, th, r);
  for (int j = 0; j < 3; j++) x[j] = (r[j] - M.xs[j]) * M.xk[j];
  return ok && isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]);
}

// layer-1 pre-activation of hidden unit k (fixed FMA order, shared by the
// forward pass and the backward pass's ReLU mask)
__device__ __forceinline__ float mlp_z1(const MlpView &M, int k, const float x[3]) {
  return fmaf(M.wq[3 * k + 2], x[2], fmaf(M.wq[3 * k + 1], x[1], fmaf(M.wq[3 * k], x[0], M.p[kOffB1 + k])));
}

// layer 1 of the lane's point into its bf16 row of the wave tile
__device__ __forceinline__ void mlp_layer1(const MlpView &M, const float x[3], __bf16 *row) {
#pragma unroll
  for (int k0 = 0; k0 < kMlpH; k0 += 8) {
    bf16x8 h;
#pragma unroll
    for (int e = 0; e < 8; e++) h[e] = (__bf16)fmaxf(mlp_z1(M, k0 + e, x), 0.f);
    *(bf16x8 *)(row + k0) = h;
  }
}

// the same with W1 (bf16-rounded, [k][3]) and b1 read from w1b1 = {W1, b1}
// (256 floats, an LDS copy in the inference kernel): the pre-activation of
// unit k is formed in mlp_z1's order, so the values are those of mlp_layer1
__device__ __forceinline__ void mlp_layer1_from(const float *w1b1, const float x[3], __bf16 *row) {
#pragma unroll
  for (int k0 = 0; k0 < kMlpH; k0 += 8) {
    bf16x8 h;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int k = k0 + e;
      const float z = fmaf(w1b1[3 * k + 2], x[2], fmaf(w1b1[3 * k + 1], x[1], fmaf(w1b1[3 * k], x[0], w1b1[192 + k])));
      h[e] = (__bf16)fmaxf(z, 0.f);
    }
    *(bf16x8 *)(row + k0) = h;
  }
}

// layers 2 and 3 for the 64 points whose bf16 rows are in h1s (row stride
// kHStr).  c receives relu(Z2^T) tiles: c[m][n] holds features 32m.. and
// points 32n..; the return value is the network output for point lane.
// b2, w3: the layer-2 bias and the bf16-rounded W3 (64 floats each, 16-byte
// aligned; LDS copies in the inference kernel).  Each lane reads them in
// groups of four consecutive features (C rows 8q + 4 (lane/32) + 0..3).
__device__ __forceinline__ float mlp_layer23(const MlpView &M, const __bf16 *h1s, f32x16 c[2][2], const float *b2,
                                             const float *w3) {
  const int l = threadIdx.x & 63, lr = l % 32, lh = l / 32;
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++)
#pragma unroll
      for (int v = 0; v < 16; v++) c[m][n][v] = 0.f;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    bf16x8 b[2], a[2];
#pragma unroll
    for (int n = 0; n < 2; n++) b[n] = *(const bf16x8 *)(h1s + (32 * n + lr) * kHStr + 16 * s + 8 * lh);
#pragma unroll
    for (int m = 0; m < 2; m++) a[m] = *(const bf16x8 *)(M.w2 + (32 * m + lr) * kMlpH + 16 * s + 8 * lh);
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int n = 0; n < 2; n++) c[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[m], b[n], c[m][n], 0, 0, 0);
  }
  float part[2] = {0.f, 0.f};
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const int f0 = 32 * m + mfma_row(4 * g, l);  // rows of v = 4g .. 4g + 3 are f0 .. f0 + 3
      const float4 bb = *(const float4 *)(b2 + f0), ww = *(const float4 *)(w3 + f0);
      const float bv[4] = {bb.x, bb.y, bb.z, bb.w}, wv[4] = {ww.x, ww.y, ww.z, ww.w};
#pragma unroll
      for (int n = 0; n < 2; n++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int v = 4 * g + e;
          const float h = fmaxf(c[m][n][v] + bv[e], 0.f);
          c[m][n][v] = h;
          part[n] = fmaf(h, wv[e], part[n]);
        }
    }
#pragma unroll
  for (int n = 0; n < 2; n++) part[n] += __shfl_xor(part[n], 32);
  // lanes 0..31 own points 0..31 (tile n = 0), lanes 32..63 points 32..63
  return (lh == 0 ? part[0] : part[1]) + M.p[kOffB3];
}

// Inference-only layers 2 and 3, one half of the points (n) at a time, so
// that only two accumulator tiles are live (fewer registers, more waves per
// SIMD).  The accumulators start at the layer-2 bias instead of adding it
// after the contraction (the FP32 sums round differently from mlp_layer23 in
// the last bits).
__device__ __forceinline__ float mlp_infer23(const MlpView &M, const __bf16 *h1s, const float *b2, const float *w3) {
  const int l = threadIdx.x & 63, lr = l % 32, lh = l / 32;
  float part0 = 0.f, part1 = 0.f;
#pragma unroll 1
  for (int n = 0; n < 2; n++) {
    float acc = 0.f;
    f32x16 c[2];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const float4 bb = *(const float4 *)(b2 + 32 * m + mfma_row(4 * g, l));
        c[m][4 * g] = bb.x;
        c[m][4 * g + 1] = bb.y;
        c[m][4 * g + 2] = bb.z;
        c[m][4 * g + 3] = bb.w;
      }
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const bf16x8 b = *(const bf16x8 *)(h1s + (32 * n + lr) * kHStr + 16 * s + 8 * lh);
#pragma unroll
      for (int m = 0; m < 2; m++) {
        const bf16x8 a = *(const bf16x8 *)(M.w2 + (32 * m + lr) * kMlpH + 16 * s + 8 * lh);
        c[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c[m], 0, 0, 0);
      }
    }
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const float4 ww = *(const float4 *)(w3 + 32 * m + mfma_row(4 * g, l));
        const float wv[4] = {ww.x, ww.y, ww.z, ww.w};
#pragma unroll
        for (int e = 0; e < 4; e++) acc = fmaf(fmaxf(c[m][4 * g + e], 0.f), wv[e], acc);
      }
    if (n == 0)
      part0 = acc;
    else
      part1 = acc;
  }
  part0 += __shfl_xor(part0, 32);
  part1 += __shfl_xor(part1, 32);
  return (lh == 0 ? part0 : part1) + M.p[kOffB3];
}

}  // namespace bamhip
