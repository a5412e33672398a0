// imc_mlp.h -- the thick-cell surrogate network 6 -> 64 -> 64 -> 4 on gfx950 matrix cores.
//
// All matrix products use v_mfma_f32_16x16x16bf16_1k (BF16 operands, FP32 accumulation).
// Operand layouts of that instruction (16x16 tiles):
//   A[i][k]: lane i + 16*(k/4), element k%4      B[k][j]: lane j + 16*(k/4), element k%4
//   D[i][j]: lane j + 16*(i/4), element i%4
// A wave evaluates the network for up to 16 samples per pass (one 16-column tile), so
// a wave only pays for the 16-lane groups that hold samples.
//
// Parameter vector (FP32 master copy, 4868 values):
//   W1[6][64] (in x out), b1[64], W2[64][64], b2[64], W3[64][4], b3[4]
// Inference copy (BF16, transposed, zero padded): W1t[64][16], W2t[64][64], W3t[16][64]
// plus FP32 biases b1[64], b2[64], b3[4].
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "imc_internal.h"

namespace imc {

constexpr int MLP_IN = 6, MLP_H = 64, MLP_OUT = 4;
constexpr int MLP_NPARAM = MLP_IN * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H * MLP_OUT + MLP_OUT;  // 4868
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_IN * MLP_H, OFF_W2 = OFF_B1 + MLP_H,
              OFF_B2 = OFF_W2 + MLP_H * MLP_H, OFF_W3 = OFF_B2 + MLP_H, OFF_B3 = OFF_W3 + MLP_H * MLP_OUT;
// bf16 inference block (uint16 units): W1t 64x16, W2t 64x64, W3t 16x64
constexpr int BF_W1T = 0, BF_W2T = BF_W1T + 64 * 16, BF_W3T = BF_W2T + 64 * 64, BF_TOTAL = BF_W3T + 16 * 64;
constexpr int BIAS_TOTAL = 64 + 64 + 4;

typedef short mshort4 __attribute__((ext_vector_type(4)));
typedef float mfloat4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (gfx950's v_cvt_pk_bf16_f32; inputs are finite).
__device__ __forceinline__ uint16_t f2bf(float f) { return __bfloat16_as_ushort(__float2bfloat16(f)); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ mfloat4 mfma_bf16(mshort4 a, mshort4 b, mfloat4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Fragment of a 16x16 operand tile: element j of lane l is
// base[(i0 + l%16) * si + (k0 + 4*(l/16) + j) * sk]  (A: i = row; B: i = column).
__device__ __forceinline__ mshort4 frag(const uint16_t* base, int si, int sk, int i0, int k0) {
  const int l = (int)(threadIdx.x & 63u);
  const int i = i0 + (l & 15), k = k0 + 4 * (l >> 4);
  mshort4 v;
  v[0] = (short)base[i * si + (k + 0) * sk];
  v[1] = (short)base[i * si + (k + 1) * sk];
  v[2] = (short)base[i * si + (k + 2) * sk];
  v[3] = (short)base[i * si + (k + 3) * sk];
  return v;
}

// Contiguous fragment (sk == 1, row stride a multiple of 4): one 8-byte load.
__device__ __forceinline__ mshort4 frag_c(const uint16_t* base, int si, int i0, int k0) {
  const int l = (int)(threadIdx.x & 63u);
  return *(const mshort4*)(base + (i0 + (l & 15)) * si + k0 + 4 * (l >> 4));
}

// Orders this wave's LDS traffic: writes by one lane before reads by another lane.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + __expf(-z)); }

// Per-wave LDS scratch of one 16-sample pass: the input tile xs[16 samples][16] (bf16)
// and the outputs y[16 samples][4] (fp32).
struct MlpScratch {
  uint16_t xs[16 * 16];
  float y[16 * 4];
};

// Evaluates the network for the lanes with `active` set (all 64 lanes must call).
// feat[6] are the lane's normalised inputs; out[4] receives the sigmoid outputs.
//
// Samples are the columns: H1^T = relu(W1^T X^T + b1), H2^T = relu(W2^T H1^T + b2),
// Y^T = W3^T H2^T + b3.  With the 16x16x16 layouts above, the D tile of one layer
// (lane holds rows 4*(l/16)..+3 of column l%16) is exactly the B fragment that the next
// layer needs for the same 16 rows as its k range, so hidden activations never leave the
// registers; the weights are the A operands (8-byte LDS loads).
__device__ __forceinline__ void mlp_forward_wave(const uint16_t* __restrict__ wbf,   // BF_TOTAL
                                                 const float* __restrict__ bias,      // BIAS_TOTAL
                                                 MlpScratch* sc, bool active, const float (&feat)[MLP_IN],
                                                 float (&out)[MLP_OUT]) {
  const int l = (int)(threadIdx.x & 63u);
  const int grp = l >> 4;
  const uint64_t mask = __ballot(active);
  const uint16_t* W1t = wbf + BF_W1T;
  const uint16_t* W2t = wbf + BF_W2T;
  const uint16_t* W3t = wbf + BF_W3T;
  const float* b1 = bias;
  const float* b2 = bias + 64;
  const float* b3 = bias + 128;
  // Active lanes are packed by rank into ceil(n/16) passes of 16 samples each.
  const int nact = __popcll(mask);
  const uint64_t lt = l == 0 ? 0ull : (~0ull >> (64 - l));
  const int rank = __popcll(mask & lt);
  for (int c = 0; 16 * c < nact; ++c) {                       // wave-uniform trip count
    // The weight and bias addresses are made opaque per pass, so the compiler cannot hoist
    // the 24 weight fragments and 36 biases out of the pass loop: held across the caller's
    // event loop, they took ~84 VGPRs and halved the transport kernel's occupancy.
    asm volatile("" : "+s"(W1t), "+s"(W2t), "+s"(W3t), "+s"(b1), "+s"(b2), "+s"(b3));
    const bool mine = active && (rank >> 4) == c;
    if (mine) {
      const int sidx = rank & 15;
#pragma unroll
      for (int j = 0; j < 16; ++j) sc->xs[sidx * 16 + j] = j < MLP_IN ? f2bf(feat[j]) : (uint16_t)0;
    }
    wave_lds_sync();
    const mshort4 bx = frag_c(sc->xs, 16, 0, 0);             // B[k][s] = X[s][k]
    mshort4 h1[4], h2[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {                          // layer 1
      mfloat4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma_bf16(frag_c(W1t, 16, 16 * mt, 0), bx, acc);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        h1[mt][r] = (short)f2bf(fmaxf(acc[r] + b1[16 * mt + 4 * grp + r], 0.f));
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {                          // layer 2
      mfloat4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) acc = mfma_bf16(frag_c(W2t, 64, 16 * mt, 16 * kt), h1[kt], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        h2[mt][r] = (short)f2bf(fmaxf(acc[r] + b2[16 * mt + 4 * grp + r], 0.f));
    }
    mfloat4 acc = {0.f, 0.f, 0.f, 0.f};                       // layer 3: rows 0..3 are outputs
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) acc = mfma_bf16(frag_c(W3t, 64, 0, 16 * kt), h2[kt], acc);
    if (grp == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) sc->y[(l & 15) * 4 + r] = acc[r] + b3[r];
    }
    wave_lds_sync();
    if (mine) {
      const int sidx = rank & 15;
#pragma unroll
      for (int o = 0; o < MLP_OUT; ++o) out[o] = sigmoidf_(sc->y[sidx * 4 + o]);
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------------
// Feature map of the single-cell problem.  All lengths in units of the collision mean
// free path 1/sigma_c (sigma_c = sigma_a(1-f) + sigma_s):
//   tau_cell = sigma_c * L, xi = x/L, mu, tau_B = sigma_c * min(census path, cut-off path),
//   kappa = f sigma_a / sigma_c (weight decay per mean free path).
// Logarithmic features are mapped to [-1, 1] over the training box: f = a ln(v) + b, with
// (a, b) formed once per launch (SurrMap) instead of per sample.
// ---------------------------------------------------------------------------------
struct SurrMap {
  float a[4], b[4];                    // tau_cell, tau_B, kappa, tau_B / tau_cell^2
};

// Formed on the host and passed by value, so the coefficients live in scalar registers.
inline void map_coef(float lo, float hi, float& a, float& b) {
  const float l0 = logf(lo), l1 = logf(hi);
  a = 2.0f / (l1 - l0);
  b = -1.0f - a * l0;
}

inline SurrMap surr_map(const SurrBox& bx) {
  SurrMap m;
  map_coef(bx.tau_cell_min, bx.tau_cell_max, m.a[0], m.b[0]);
  map_coef(bx.tau_b_min, bx.tau_b_max, m.a[1], m.b[1]);
  map_coef(bx.kappa_min, bx.kappa_max, m.a[2], m.b[2]);
  // diffusion-time ratio tau_B / tau_cell^2 over its range in the box
  map_coef(bx.tau_b_min / (bx.tau_cell_max * bx.tau_cell_max), bx.tau_b_max / (bx.tau_cell_min * bx.tau_cell_min),
           m.a[3], m.b[3]);
  return m;
}

__device__ __forceinline__ void surr_features(const SurrMap& m, float tau_cell, float xi, float mu, float tau_b,
                                              float kappa, float (&f)[MLP_IN]) {
  const float lc = __logf(tau_cell), lb = __logf(tau_b);
  f[0] = m.a[0] * lc + m.b[0];
  f[1] = 2.0f * xi - 1.0f;
  f[2] = mu;
  f[3] = m.a[1] * lb + m.b[1];
  f[4] = m.a[2] * __logf(kappa) + m.b[2];
  f[5] = m.a[3] * (lb - 2.0f * lc) + m.b[3];
}

__device__ __forceinline__ bool surr_in_box(const SurrBox& bx, float tau_cell, float tau_b, float kappa) {
  return tau_cell >= bx.tau_cell_min && tau_cell <= bx.tau_cell_max && tau_b >= bx.tau_b_min &&
         tau_b <= bx.tau_b_max && kappa >= bx.kappa_min && kappa <= bx.kappa_max;
}

}  // namespace imc
