// imc_planck.hip -- Planck frequency sampling, Sourcing.sample_planck (imc_sourcing.jl:372-399).
//
// The normalised Planck spectrum in x = h nu / k T is (15 / pi^4) x^3 / (e^x - 1)
//   = sum_{n >= 1} (90 / (pi^4 n^4)) * [n^4 x^3 e^{-n x} / 6],
// a mixture of Gamma(4, 1/n) laws with weights 90 / (pi^4 n^4) (the method the reference names
// "from Fleck and Cummings").  One sample:
//   1. n by inversion of the weights, the cumulative formed in FP64 term by term for
//      n = 1 .. NTERMS;
//   2. beyond NTERMS (probability 1.2e-6) n by rejection: X = NTERMS / cbrt(1 - v) has density
//      3 NTERMS^3 X^-4 on [NTERMS, inf), n = ceil(X) is accepted with probability
//      3 n^-4 / ((n-1)^-3 - n^-3) <= 1, which leaves P(n) proportional to n^-4 exactly;
//   3. x = -(log xi1 + log xi2 + log xi3 + log xi4) / n, the xi in (0, 1];
// and nu = T x.  Unlike the reference, which keeps the running sum of n^-4 in the tracking
// precision and loops while its first uniform is above 90 sum / pi^4 (a sum that stops growing
// below zeta(4) in Float16 and Float32, so a draw close enough to 1 never ends), every draw ends:
// the rejection loop is bounded by MAX_TRIES (acceptance >= 0.97 per try; the bound is never met
// with probability above 1e-90).
//
// Sample i uses Philox blocks (i, TAG_PLANCK, j) under the key (seed, stream):
//   j = 0: u for n (words 0-1), xi1 (words 2-3);  j = 1: xi2, xi3;  j = 2: xi4, v;
//   j = 3 + t: the t-th tail acceptance test (words 0-1) and the next proposal v (words 2-3).
// FP64 throughout: the sampler is not on the transport path.
#include <hip/hip_runtime.h>

#include "imc_device.h"
#include "imc_internal.h"

namespace imc {

constexpr uint32_t TAG_PLANCK = 6u;
constexpr int PLANCK_NTERMS = 64;
constexpr int PLANCK_MAX_TRIES = 64;

// (k53 + 1) 2^-53: uniform on (0, 1], so its log is finite
__device__ __forceinline__ double u53_open0(uint32_t hi, uint32_t lo) {
  return (double)((((uint64_t)hi) << 21) | (lo >> 11)) * 0x1p-53 + 0x1p-53;
}

__device__ double planck_x(uint64_t i, uint32_t seed, uint32_t stream) {
  const uint32_t i0 = (uint32_t)i, i1 = (uint32_t)(i >> 32);
  const u4 q0 = philox(i0, i1, TAG_PLANCK, 0u, seed, stream);
  const u4 q1 = philox(i0, i1, TAG_PLANCK, 1u, seed, stream);
  const u4 q2 = philox(i0, i1, TAG_PLANCK, 2u, seed, stream);
  const double u = u53(q0.r[0], q0.r[1]);
  const double w = 90.0 / (PI_D * PI_D * PI_D * PI_D);
  double cum = 0.0, n = 0.0;
  for (int m = 1; m <= PLANCK_NTERMS; ++m) {
    const double dm = (double)m;
    cum += w / (dm * dm * dm * dm);
    if (u < cum) { n = dm; break; }
  }
  if (n == 0.0) {                                   // tail: n > NTERMS
    double v = u53(q2.r[2], q2.r[3]);
    for (int t = 0; t < PLANCK_MAX_TRIES; ++t) {
      n = ceil((double)PLANCK_NTERMS / cbrt(1.0 - v));
      if (n <= (double)PLANCK_NTERMS) n = (double)PLANCK_NTERMS + 1.0;
      const double a = 1.0 / (n - 1.0), b = 1.0 / n;
      const double accept = 3.0 * (b * b * b * b) / (a * a * a - b * b * b);
      const u4 qt = philox(i0, i1, TAG_PLANCK, 3u + (uint32_t)t, seed, stream);
      if (u53(qt.r[0], qt.r[1]) < accept) break;
      v = u53(qt.r[2], qt.r[3]);
    }
  }
  const double s = log(u53_open0(q0.r[2], q0.r[3])) + log(u53_open0(q1.r[0], q1.r[1])) +
                   log(u53_open0(q1.r[2], q1.r[3])) + log(u53_open0(q2.r[0], q2.r[1]));
  return -s / n;
}

__global__ __launch_bounds__(256) void k_planck(int64_t n, const double* __restrict__ temp,
                                                double* __restrict__ freq, uint32_t seed, uint32_t stream) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  freq[i] = temp[i] * planck_x((uint64_t)i, seed, stream);
}

hipError_t launch_planck(Ctx& c, int64_t n, const double* d_temp, double* d_freq, uint32_t stream) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + 255) / 256;
  k_planck<<<(unsigned)blocks, 256, 0, c.stream>>>(n, d_temp, d_freq, c.cfg.seed, stream);
  return hipGetLastError();
}

}  // namespace imc
