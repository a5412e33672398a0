// imc_surrogate.hip -- the learned thick-cell surrogate: training-data generation from
// plain single-cell histories, BF16-MFMA training (forward + backward, FP32 accumulation),
// Adam, and standalone forward / gradient evaluation.
//
// The single-cell problem (lengths in collision mean free paths): a particle at x in
// [depth_min, tau_cell - depth_min] with direction mu random-walks with unit collision cross section and
// isotropic effective scattering (imc_transport.jl:72-194 with sigma_c = 1) until it
// crosses a face (exit) or its path budget tau_B is spent (stay: census or weight
// cut-off).  Its weight after path s is exp(-kappa s).  Targets per history:
//   stay (0/1), right (exit through x = tau_cell), s/tau_B at exit, 1 - exp(-kappa s).
// The network learns their conditional means: P(stay), P(right | exit), E[s/tau_B | exit],
// E[deposited fraction | exit].
#include <hip/hip_runtime.h>

#include "imc_device.h"
#include "imc_internal.h"
#include "imc_mlp.h"

namespace imc {

constexpr uint32_t TAG_GEN = 4u;
constexpr uint32_t GEN_KEY1 = 0x53555252u;
constexpr int GEN_MAX_EVENTS = 1 << 20;

// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_surr_generate(SurrBox bx, SurrMap smap, int64_t n, uint32_t seed,
                                                       float* __restrict__ feat, float* __restrict__ targ,
                                                       unsigned long long* __restrict__ events) {
  unsigned long long evcount = 0;

  const float lc0 = __logf(bx.tau_cell_min), lc1 = __logf(bx.tau_cell_max);
  const float lb0 = __logf(bx.tau_b_min), lb1 = __logf(bx.tau_b_max);
  const float lk0 = __logf(bx.kappa_min), lk1 = __logf(bx.kappa_max);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t lo = (uint32_t)i, hi = (uint32_t)(i >> 32);
    const u4 q0 = philox(lo, hi, TAG_GEN, 0u, seed, GEN_KEY1);
    const float tau_cell = __expf(lc0 + u24(q0.r[0]) * (lc1 - lc0));
    const float xlo = bx.depth_min / tau_cell;                 // the surrogate is applied only
    const float xi = xlo + u24(q0.r[1]) * (1.0f - 2.0f * xlo);  // depth_min paths from both faces
    float mu = 1.0f - 2.0f * u24(q0.r[2]);
    if (mu == 0.0f) mu = 0x1p-24f;
    const float tau_b = __expf(lb0 + u24(q0.r[3]) * (lb1 - lb0));
    const u4 q1 = philox(lo, hi, TAG_GEN, 1u, seed, GEN_KEY1);
    const float kappa = __expf(lk0 + u24(q1.r[0]) * (lk1 - lk0));
    const float L = tau_cell;
    float x = xi * L, m = mu, s = 0.0f;
    int outcome = 0;  // 0 stay, 1 exit left, 2 exit right
    for (int ev = 0; ev < GEN_MAX_EVENTS; ++ev) {
      const u4 q = philox(lo, hi, TAG_GEN, 2u + (uint32_t)ev, seed, GEN_KEY1);
      ++evcount;
      const float db = m > 0.0f ? (L - x) / m : x / (-m);
      const float dc = -logf((float)((q.r[0] >> 8) + 1) * 0x1p-24f);
      const float r = tau_b - s;
      const float d = fminf(fminf(db, dc), r);
      s += d;
      x += m * d;
      if (d == r) { outcome = 0; break; }
      if (d == db) { outcome = m > 0.0f ? 2 : 1; break; }
      m = 1.0f - 2.0f * u24(q.r[1]);
      if (m == 0.0f) m = 0x1p-24f;
    }
    float f[MLP_IN];
    surr_features(smap, tau_cell, xi, mu, tau_b, kappa, f);
#pragma unroll
    for (int j = 0; j < MLP_IN; ++j) feat[i * MLP_IN + j] = f[j];
    targ[i * 4 + 0] = outcome == 0 ? 1.0f : 0.0f;
    targ[i * 4 + 1] = outcome == 2 ? 1.0f : 0.0f;
    targ[i * 4 + 2] = outcome == 0 ? 1.0f : fminf(s / tau_b, 1.0f);
    targ[i * 4 + 3] = 1.0f - __expf(-kappa * s);
  }
  // one atomic per wave
  unsigned long long w = evcount;
#pragma unroll
  for (int m2 = 32; m2 >= 1; m2 >>= 1) w += __shfl_xor(w, m2, 64);
  if ((threadIdx.x & 63u) == 0) atomicAdd(events, w);
}

// ---------------------------------------------------------------------------------
// Master FP32 parameters -> BF16 transposed inference copy + FP32 biases.
// ---------------------------------------------------------------------------------
__global__ void k_make_inference_weights(const float* __restrict__ p, uint16_t* __restrict__ wbf,
                                         float* __restrict__ bias) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 64 * 16) {                       // W1t[o][i] = W1[i][o]
    const int o = t / 16, i = t % 16;
    wbf[BF_W1T + t] = i < MLP_IN ? f2bf(p[OFF_W1 + i * 64 + o]) : (uint16_t)0;
  }
  if (t < 64 * 64) {                       // W2t[o][i] = W2[i][o]
    const int o = t / 64, i = t % 64;
    wbf[BF_W2T + t] = f2bf(p[OFF_W2 + i * 64 + o]);
  }
  if (t < 16 * 64) {                       // W3t[o][i] = W3[i][o]
    const int o = t / 64, i = t % 64;
    wbf[BF_W3T + t] = o < MLP_OUT ? f2bf(p[OFF_W3 + i * 4 + o]) : (uint16_t)0;
  }
  if (t < 64) { bias[t] = p[OFF_B1 + t]; bias[64 + t] = p[OFF_B2 + t]; }
  if (t < MLP_OUT) bias[128 + t] = p[OFF_B3 + t];
}

// ---------------------------------------------------------------------------------
// Training step on 64 samples per block (one wave): forward, BCE loss on the four
// sigmoid outputs (outputs 1-3 masked to exiting histories), backward, and FP32 atomic
// accumulation of the gradient.  dz = (p - y) * mask * inv_batch.
// ---------------------------------------------------------------------------------
struct TrainLds {
  uint16_t w[BF_TOTAL];
  float bias[BIAS_TOTAL];
  uint16_t X[64 * 16];
  uint16_t H1[64 * 64];
  uint16_t H2[64 * 64];
  uint16_t dY[64 * 16];
  uint16_t dH2[64 * 64];
  uint16_t dH1[64 * 64];
  float Y[64 * 4];
};

__global__ __launch_bounds__(64) void k_surr_grad(const uint16_t* __restrict__ wbf, const float* __restrict__ bias,
                                                   const float* __restrict__ feat, const float* __restrict__ targ,
                                                   int64_t n0, int64_t n, float inv_batch,
                                                   float* __restrict__ grad, float* __restrict__ loss) {
  __shared__ TrainLds S;
  const int l = (int)threadIdx.x;
  const int64_t s0 = n0 + (int64_t)blockIdx.x * 64;
  for (int t = l; t < BF_TOTAL; t += 64) S.w[t] = wbf[t];
  for (int t = l; t < BIAS_TOTAL; t += 64) S.bias[t] = bias[t];
  const int64_t my = s0 + l;
  const bool valid = my < n0 + n;
#pragma unroll
  for (int j = 0; j < 16; ++j) S.X[l * 16 + j] = (valid && j < MLP_IN) ? f2bf(feat[my * MLP_IN + j]) : (uint16_t)0;
  __syncthreads();
  const uint16_t* W1t = S.w + BF_W1T;
  const uint16_t* W2t = S.w + BF_W2T;
  const uint16_t* W3t = S.w + BF_W3T;
  // ---- forward ----
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt) {
      mfloat4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma_bf16(frag(S.X, 16, 1, 16 * mt, 0), frag(W1t, 16, 1, 16 * nt, 0), acc);
      const int col = 16 * nt + (l & 15);
      for (int r = 0; r < 4; ++r)
        S.H1[(16 * mt + 4 * (l >> 4) + r) * 64 + col] = f2bf(fmaxf(acc[r] + S.bias[col], 0.f));
    }
  __syncthreads();
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt) {
      mfloat4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int kt = 0; kt < 4; ++kt)
        acc = mfma_bf16(frag(S.H1, 64, 1, 16 * mt, 16 * kt), frag(W2t, 64, 1, 16 * nt, 16 * kt), acc);
      const int col = 16 * nt + (l & 15);
      for (int r = 0; r < 4; ++r)
        S.H2[(16 * mt + 4 * (l >> 4) + r) * 64 + col] = f2bf(fmaxf(acc[r] + S.bias[64 + col], 0.f));
    }
  __syncthreads();
  for (int mt = 0; mt < 4; ++mt) {
    mfloat4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < 4; ++kt)
      acc = mfma_bf16(frag(S.H2, 64, 1, 16 * mt, 16 * kt), frag(W3t, 64, 1, 0, 16 * kt), acc);
    const int col = l & 15;
    if (col < MLP_OUT)
      for (int r = 0; r < 4; ++r) S.Y[(16 * mt + 4 * (l >> 4) + r) * 4 + col] = acc[r] + S.bias[128 + col];
  }
  __syncthreads();
  // ---- loss and output deltas (lane = sample) ----
  float dz[MLP_OUT] = {0.f, 0.f, 0.f, 0.f};
  float lsum = 0.f;
  if (valid) {
    const float stay = targ[my * 4 + 0];
    for (int o = 0; o < MLP_OUT; ++o) {
      const float z = S.Y[l * 4 + o];
      const float p = 1.0f / (1.0f + expf(-z));
      const float y = targ[my * 4 + o];
      const float mask = o == 0 ? 1.0f : (1.0f - stay);
      dz[o] = (p - y) * mask * inv_batch;
      // BCE with logits, numerically stable
      const float bce = fmaxf(z, 0.f) - z * y + log1pf(expf(-fabsf(z)));
      lsum += bce * mask;
    }
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) S.dY[l * 16 + j] = j < MLP_OUT ? f2bf(dz[j]) : (uint16_t)0;
  // db3 from the bf16 deltas (the values the MFMA products use)
  for (int o = 0; o < MLP_OUT; ++o) {
    float v = bf2f(f2bf(dz[o]));
    for (int m2 = 32; m2 >= 1; m2 >>= 1) v += __shfl_xor(v, m2, 64);
    if (l == 0) atomicAdd(&grad[OFF_B3 + o], v);
  }
  {
    float v = lsum;
    for (int m2 = 32; m2 >= 1; m2 >>= 1) v += __shfl_xor(v, m2, 64);
    if (l == 0) atomicAdd(loss, v);
  }
  __syncthreads();
  // ---- dW3 = H2^T dY ----
  for (int mt = 0; mt < 4; ++mt) {
    mfloat4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < 4; ++kt)
      acc = mfma_bf16(frag(S.H2, 1, 64, 16 * mt, 16 * kt), frag(S.dY, 1, 16, 0, 16 * kt), acc);
    const int o = l & 15;
    if (o < MLP_OUT)
      for (int r = 0; r < 4; ++r) atomicAdd(&grad[OFF_W3 + (16 * mt + 4 * (l >> 4) + r) * 4 + o], acc[r]);
  }
  // ---- dH2 = (dY W3^T) * relu'(H2) ----
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt) {
      mfloat4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma_bf16(frag(S.dY, 16, 1, 16 * mt, 0), frag(W3t, 1, 64, 16 * nt, 0), acc);
      const int col = 16 * nt + (l & 15);
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * mt + 4 * (l >> 4) + r;
        const bool on = bf2f(S.H2[row * 64 + col]) > 0.f;
        S.dH2[row * 64 + col] = on ? f2bf(acc[r]) : (uint16_t)0;
      }
    }
  __syncthreads();
  {   // db2[h] = sum_s dH2[s][h]
    float v = 0.f;
    for (int s = 0; s < 64; ++s) v += bf2f(S.dH2[s * 64 + l]);
    atomicAdd(&grad[OFF_B2 + l], v);
  }
  // ---- dW2 = H1^T dH2 ----
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt) {
      mfloat4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int kt = 0; kt < 4; ++kt)
        acc = mfma_bf16(frag(S.H1, 1, 64, 16 * mt, 16 * kt), frag(S.dH2, 1, 64, 16 * nt, 16 * kt), acc);
      const int col = 16 * nt + (l & 15);
      for (int r = 0; r < 4; ++r) atomicAdd(&grad[OFF_W2 + (16 * mt + 4 * (l >> 4) + r) * 64 + col], acc[r]);
    }
  // ---- dH1 = (dH2 W2^T) * relu'(H1) ----
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt) {
      mfloat4 acc = {0.f, 0.f, 0.f, 0.f};
      for (int kt = 0; kt < 4; ++kt)
        acc = mfma_bf16(frag(S.dH2, 64, 1, 16 * mt, 16 * kt), frag(W2t, 1, 64, 16 * nt, 16 * kt), acc);
      const int col = 16 * nt + (l & 15);
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * mt + 4 * (l >> 4) + r;
        const bool on = bf2f(S.H1[row * 64 + col]) > 0.f;
        S.dH1[row * 64 + col] = on ? f2bf(acc[r]) : (uint16_t)0;
      }
    }
  __syncthreads();
  {   // db1
    float v = 0.f;
    for (int s = 0; s < 64; ++s) v += bf2f(S.dH1[s * 64 + l]);
    atomicAdd(&grad[OFF_B1 + l], v);
  }
  // ---- dW1 = X^T dH1 (rows 0..5 of the 16-row tile are inputs) ----
  for (int nt = 0; nt < 4; ++nt) {
    mfloat4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < 4; ++kt)
      acc = mfma_bf16(frag(S.X, 1, 16, 0, 16 * kt), frag(S.dH1, 1, 64, 16 * nt, 16 * kt), acc);
    const int col = 16 * nt + (l & 15);
    for (int r = 0; r < 4; ++r) {
      const int in = 4 * (l >> 4) + r;
      if (in < MLP_IN) atomicAdd(&grad[OFF_W1 + in * 64 + col], acc[r]);
    }
  }
}

__global__ void k_adam(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NPARAM) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  g[i] = 0.f;
}

// Standalone forward (one wave per 64 samples); outputs sigmoid values.
__global__ __launch_bounds__(64) void k_surr_forward(const uint16_t* __restrict__ wbf, const float* __restrict__ bias,
                                                      const float* __restrict__ feat, int64_t n,
                                                      float* __restrict__ out) {
  __shared__ uint16_t w[BF_TOTAL];
  __shared__ float b[BIAS_TOTAL];
  __shared__ MlpScratch sc;
  for (int t = threadIdx.x; t < BF_TOTAL; t += 64) w[t] = wbf[t];
  for (int t = threadIdx.x; t < BIAS_TOTAL; t += 64) b[t] = bias[t];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const bool active = i < n;
  float f[MLP_IN], o[MLP_OUT] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < MLP_IN; ++j) f[j] = active ? feat[i * MLP_IN + j] : 0.f;
  mlp_forward_wave(w, b, &sc, active, f, o);
  if (active)
#pragma unroll
    for (int j = 0; j < MLP_OUT; ++j) out[i * MLP_OUT + j] = o[j];
}

// ---------------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------------
hipError_t surr_generate(Ctx& c, int64_t n, uint32_t seed, unsigned long long* d_events) {
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 8L * c.num_cu * 4);
  k_surr_generate<<<grid, 256, 0, c.stream>>>(c.surr_box, surr_map(c.surr_box), n, seed, c.d_surr_feat, c.d_surr_targ,
                                              d_events);
  return hipGetLastError();
}

hipError_t surr_make_weights(Ctx& c) {
  k_make_inference_weights<<<16, 256, 0, c.stream>>>(c.d_surr_params, c.d_surr_wbf, c.d_surr_bias);
  return hipGetLastError();
}

hipError_t surr_grad(Ctx& c, const float* feat, const float* targ, int64_t n0, int64_t n, float inv_batch,
                     float* grad, float* loss) {
  const unsigned grid = (unsigned)((n + 63) / 64);
  k_surr_grad<<<grid, 64, 0, c.stream>>>(c.d_surr_wbf, c.d_surr_bias, feat, targ, n0, n, inv_batch, grad, loss);
  return hipGetLastError();
}

hipError_t surr_adam(Ctx& c, float lr, int step) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
  k_adam<<<(MLP_NPARAM + 255) / 256, 256, 0, c.stream>>>(c.d_surr_params, c.d_surr_grad, c.d_surr_m, c.d_surr_v,
                                                         lr, b1, b2, 1e-8f, bc1, bc2);
  return hipGetLastError();
}

hipError_t surr_forward(Ctx& c, const float* feat, int64_t n, float* out) {
  const unsigned grid = (unsigned)((n + 63) / 64);
  k_surr_forward<<<grid, 64, 0, c.stream>>>(c.d_surr_wbf, c.d_surr_bias, feat, n, out);
  return hipGetLastError();
}

}  // namespace imc
