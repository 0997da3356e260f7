// warmstart.hpp -- the learned SHAKE warm start: a small network that
// predicts each H-bond cluster's converged SHAKE multipliers from its
// geometry, run on the matrix cores.  No counterpart in the reference: it
// attaches where `shake_cluster` starts (src/gpu/kernels/shake.wgsl:130-146),
// before the oracle's multiplier accumulation (src/reference/constraints.ts
// :147-184).  Plain SHAKE still iterates to the same tolerance afterwards, so
// the constraints end where they would have; the network only saves passes.
//
// Network (BF16 weights and activations as matrix operands, FP32
// accumulation and FP32 biases):
//     H1 = tanh(X W1 + b1),  X: [n, 32], W1: [32, 64]
//     H2 = tanh(H1 W2 + b2), W2: [64, 64]
//     Y  = H2 W3 + b3,       W3: [64, 16], outputs 0..3 used
// Tiles of 16 clusters go through v_mfma_f32_16x16x16_bf16: lane l holds
// A[l%16][k0 + 4(l/16) + j] and B[k0 + 4(l/16) + j][l%16], j = 0..3, and
// receives D[4(l/16) + j][l%16].  Weights sit in global memory already in
// that B-operand order, one 8-byte load per lane per tile.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define MD_WS_F 32      // features per cluster
#define MD_WS_H 64      // hidden width
#define MD_WS_O 16      // output tile width (4 used)
#define MD_WS_MAXC 4    // constraints per cluster the network covers
#define MD_WS_SCALE 100.0f  // features and targets are scaled by this
// parameters: W1 | b1 | W2 | b2 | W3 | b3
#define MD_WS_OFF_W1 0
#define MD_WS_OFF_B1 (MD_WS_OFF_W1 + MD_WS_F * MD_WS_H)
#define MD_WS_OFF_W2 (MD_WS_OFF_B1 + MD_WS_H)
#define MD_WS_OFF_B2 (MD_WS_OFF_W2 + MD_WS_H * MD_WS_H)
#define MD_WS_OFF_W3 (MD_WS_OFF_B2 + MD_WS_H)
#define MD_WS_OFF_B3 (MD_WS_OFF_W3 + MD_WS_H * MD_WS_O)
#define MD_WS_NPARAMS (MD_WS_OFF_B3 + MD_WS_O)
// per-wave LDS of one 16-row tile: X [16][32], H1 [16][64], H2 [16][64]
#define MD_WS_TILE_LDS (16 * MD_WS_F + 2 * 16 * MD_WS_H)

typedef short md_s4 __attribute__((ext_vector_type(4)));
typedef float md_f4 __attribute__((ext_vector_type(4)));

// the network as the kernels read it
struct md_ws_net {
  const md_s4 *W1s;   // B tiles of W1: [kt 2][nt 4][lane 64]
  const md_s4 *W2s;   // B tiles of W2: [kt 4][nt 4][lane 64]
  const md_s4 *W3s;   // B tiles of W3: [kt 4][lane 64]
  const md_s4 *W2ts;  // B tiles of W2^T: [kt 4][nt 4][lane 64]
  const md_s4 *W3ts;  // B tiles of W3^T: [nt 4][lane 64]
  const float *b1, *b2, *b3;
};

__device__ __forceinline__ short md_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);  // nearest, ties to even
  return (short)(u >> 16);
}

__device__ __forceinline__ md_f4 md_mfma(md_s4 a, md_s4 b, md_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A operand of columns k0..k0+15 of a row-major [16][ld] LDS matrix
__device__ __forceinline__ md_s4 md_a_rows(const float *M, int ld, int k0, int l) {
  const float *p = M + (l & 15) * ld + k0 + 4 * (l >> 4);
  md_s4 a;
  a[0] = md_bf16(p[0]);
  a[1] = md_bf16(p[1]);
  a[2] = md_bf16(p[2]);
  a[3] = md_bf16(p[3]);
  return a;
}

// B operand -- or the A operand of the transpose -- of columns c0..c0+15 of
// a row-major [16][ld] LDS matrix: lane l takes M[4(l/16) + j][c0 + l%16]
__device__ __forceinline__ md_s4 md_b_rows(const float *M, int ld, int c0, int l) {
  const float *p = M + 4 * (l >> 4) * ld + c0 + (l & 15);
  md_s4 b;
  b[0] = md_bf16(p[0]);
  b[1] = md_bf16(p[ld]);
  b[2] = md_bf16(p[2 * ld]);
  b[3] = md_bf16(p[3 * ld]);
  return b;
}

__device__ __forceinline__ float md_tanh(float x) {
  float e = __expf(-2.0f * fabsf(x));
  float t = (1.0f - e) / (1.0f + e);
  return copysignf(t, x);
}

// Forward pass of one 16-row tile: X (LDS [16][32]) -> H1, H2 (LDS
// [16][64] each); returns Y in the D layout, Y[4(l/16)+j][l%16].  Every lane
// of the wave takes part; LDS is fenced with workgroup barriers, so every
// wave of the workgroup must call it the same number of times.
__device__ __forceinline__ md_f4 md_ws_forward(const md_ws_net &N, const float *X, float *H1, float *H2, int l) {
  const int row0 = 4 * (l >> 4), col = l & 15;
  md_f4 acc[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) acc[nt] = md_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int kt = 0; kt < 2; ++kt) {
    md_s4 a = md_a_rows(X, MD_WS_F, 16 * kt, l);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[nt] = md_mfma(a, N.W1s[(kt * 4 + nt) * 64 + l], acc[nt]);
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int c = 16 * nt + col;
      H1[(row0 + j) * MD_WS_H + c] = md_tanh(acc[nt][j] + N.b1[c]);
    }
  __syncthreads();
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) acc[nt] = md_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int kt = 0; kt < 4; ++kt) {
    md_s4 a = md_a_rows(H1, MD_WS_H, 16 * kt, l);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[nt] = md_mfma(a, N.W2s[(kt * 4 + nt) * 64 + l], acc[nt]);
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int c = 16 * nt + col;
      H2[(row0 + j) * MD_WS_H + c] = md_tanh(acc[nt][j] + N.b2[c]);
    }
  __syncthreads();
  md_f4 y = md_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int kt = 0; kt < 4; ++kt) y = md_mfma(md_a_rows(H2, MD_WS_H, 16 * kt, l), N.W3s[kt * 64 + l], y);
#pragma unroll
  for (int j = 0; j < 4; ++j) y[j] += N.b3[col];
  return y;
}

// Inference for the 64 clusters of a wave (one per lane): `f` is this
// lane's feature vector, `out` receives its 4 outputs.  `lds` is this wave's
// MD_WS_TILE_LDS floats.  All lanes of every wave in the workgroup call it.
__device__ __forceinline__ void md_ws_infer(const md_ws_net &N, const float *f, float *out, float *lds) {
  const int l = threadIdx.x & 63;
  float *X = lds, *H1 = lds + 16 * MD_WS_F, *H2 = H1 + 16 * MD_WS_H;
  for (int t = 0; t < 4; ++t) {
    if ((l >> 4) == t)
#pragma unroll
      for (int k = 0; k < MD_WS_F; ++k) X[(l & 15) * MD_WS_F + k] = f[k];
    __syncthreads();
    md_f4 y = md_ws_forward(N, X, H1, H2, l);
    __syncthreads();
    // outputs 0..3 of the tile's 16 rows, through X (no longer needed)
    if ((l & 15) < MD_WS_MAXC)
#pragma unroll
      for (int j = 0; j < 4; ++j) X[(4 * (l >> 4) + j) * MD_WS_F + (l & 15)] = y[j];
    __syncthreads();
    if ((l >> 4) == t)
#pragma unroll
      for (int c = 0; c < MD_WS_MAXC; ++c) out[c] = X[(l & 15) * MD_WS_F + c];
    __syncthreads();
  }
}
