// imc_device.h -- device-side building blocks shared by the gfx950 kernels:
// Philox4x32-10, precision traits, uniform/exponential variates, the reference's
// sorter, and the wave64 tally reduction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace imc {

constexpr uint32_t TAG_BIRTH = 1u;   // Philox counter word 2 for births
constexpr uint32_t TAG_EVENT = 2u;   // ... for transport events
constexpr uint32_t DEAD = 0xFFFFFFFFu;
constexpr int KSHIFT = 28;           // cellk = cell | (scale index << 28)
constexpr uint32_t CELLMASK = (1u << KSHIFT) - 1u;
constexpr double PI_D = 3.141592653589793;
constexpr double TWO_PI_D = 6.283185307179586;

struct u4 { uint32_t r[4]; };

// Philox4x32-10 (Salmon et al., SC'11): multipliers 0xD2511F53/0xCD9E8D57, key bumps
// 0x9E3779B9/0xBB67AE85, key bumped before every round but the first.
__device__ __forceinline__ u4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                     uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    if (round) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    // one 32x32->64 product per word (v_mad_u64_u32) instead of separate hi and lo halves
    const uint64_t p0 = (uint64_t)c0 * 0xD2511F53ull, p1 = (uint64_t)c2 * 0xCD9E8D57ull;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
  }
  u4 o; o.r[0] = c0; o.r[1] = c1; o.r[2] = c2; o.r[3] = c3;
  return o;
}

__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
  return (double)((((uint64_t)hi) << 21) | (lo >> 11)) * 0x1p-53;
}
__device__ __forceinline__ float u24(uint32_t a) { return (float)(a >> 8) * 0x1p-24f; }
__device__ __forceinline__ float u11(uint32_t a) { return (float)(a >> 21) * 0x1p-11f; }

// ---------------------------------------------------------------------------------
// Precision traits.  P = 64: T=W=double.  P = 32: T=W=float.  P = 16: T=_Float16 for
// positions/directions/time/distances, W=float for weights.  Acc is the tally type.
// ---------------------------------------------------------------------------------
template <int P> struct Prec;
template <> struct Prec<64> { using T = double; using W = double; using Cc = double; using Acc = double; };
template <> struct Prec<32> { using T = float; using W = float; using Cc = float; using Acc = float; };
template <> struct Prec<16> { using T = _Float16; using W = float; using Cc = float; using Acc = float; };

template <class T> struct Tm;
template <> struct Tm<double> {
  static __device__ __forceinline__ double maxval() { return 1.7976931348623157e308; }
  static __device__ __forceinline__ double from(double x) { return x; }
  static __device__ __forceinline__ double d(double x) { return x; }
  static __device__ __forceinline__ double exp_(double x) { return exp(x); }
  static __device__ __forceinline__ double expm1_(double x) { return expm1(x); }
  static __device__ __forceinline__ double cos_(double x) { return cos(x); }
  static __device__ __forceinline__ double sin_(double x) { return sin(x); }
  static __device__ __forceinline__ double atan2_(double y, double x) { return atan2(y, x); }
  static __device__ __forceinline__ double abs_(double x) { return fabs(x); }
  static __device__ __forceinline__ bool nan_(double x) { return isnan(x); }
  // exponential variate -log(u), u = (k53 + 1) 2^-53 in (0, 1]
  static __device__ __forceinline__ double expvar(const u4& q) {
    return -log((double)(((((uint64_t)q.r[0]) << 21) | (q.r[1] >> 11)) + 1) * 0x1p-53);
  }
  static __device__ __forceinline__ double uscat(const u4& q, bool first) {
    return first ? u53(q.r[2], q.r[3]) : u53(q.r[0], q.r[1]);
  }
};
template <> struct Tm<float> {
  static __device__ __forceinline__ double maxval() { return 3.4028234663852886e38; }
  static __device__ __forceinline__ float from(double x) { return (float)x; }
  static __device__ __forceinline__ double d(float x) { return (double)x; }
  static __device__ __forceinline__ float exp_(float x) { return expf(x); }
  static __device__ __forceinline__ float expm1_(float x) { return expm1f(x); }
  static __device__ __forceinline__ float cos_(float x) { return cosf(x); }
  static __device__ __forceinline__ float sin_(float x) { return sinf(x); }
  static __device__ __forceinline__ float atan2_(float y, float x) { return atan2f(y, x); }
  static __device__ __forceinline__ float abs_(float x) { return fabsf(x); }
  static __device__ __forceinline__ bool nan_(float x) { return isnan(x); }
  static __device__ __forceinline__ float expvar(const u4& q) {
    return -logf((float)((q.r[0] >> 8) + 1) * 0x1p-24f);
  }
  static __device__ __forceinline__ float uscat(const u4& q, bool first) {
    return first ? u24(q.r[1]) : u24(q.r[0]);
  }
};
// float -> binary16, round to nearest even, done on the bit pattern.  An explicit
// conversion keeps the compiler from fusing fptrunc(fmul) into v_fma_mix, which would
// round the exact product once instead of rounding to FP32 first.
__device__ __forceinline__ _Float16 f2h(float f) {
  const uint32_t x = __float_as_uint(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7FFFFFFFu;
  uint32_t h;
  if (ax >= 0x7F800000u) {
    h = ax > 0x7F800000u ? 0x7E00u : 0x7C00u;             // NaN / Inf
  } else if (ax >= 0x477FF000u) {
    h = 0x7C00u;                                          // >= 65520 rounds to Inf
  } else if (ax >= 0x38800000u) {                         // normal: >= 2^-14
    h = (ax + 0xFFFu + ((ax >> 13) & 1u) - 0x38000000u) >> 13;
  } else if (ax >= 0x33000000u) {                         // subnormal: 2^-25 <= |f| < 2^-14
    const uint32_t e = ax >> 23;
    const uint32_t m = (ax & 0x7FFFFFu) | 0x800000u;
    const uint32_t s = 126u - e;                          // 14..24
    uint32_t q = m >> s;
    const uint32_t r = m & ((1u << s) - 1u), half = 1u << (s - 1u);
    if (r > half || (r == half && (q & 1u))) ++q;
    h = q;
  } else {
    h = 0u;
  }
  return __builtin_bit_cast(_Float16, (unsigned short)(sign | h));
}

template <> struct Tm<_Float16> {
  using h = _Float16;
  static __device__ __forceinline__ double maxval() { return 65504.0; }
  static __device__ __forceinline__ h from(double x) { return f2h((float)x); }
  static __device__ __forceinline__ double d(h x) { return (double)(float)x; }
  static __device__ __forceinline__ h exp_(h x) { return f2h(expf((float)x)); }
  static __device__ __forceinline__ h expm1_(h x) { return f2h(expm1f((float)x)); }
  static __device__ __forceinline__ h cos_(h x) { return f2h(cosf((float)x)); }
  static __device__ __forceinline__ h sin_(h x) { return f2h(sinf((float)x)); }
  static __device__ __forceinline__ h atan2_(h y, h x) { return f2h(atan2f((float)y, (float)x)); }
  static __device__ __forceinline__ h abs_(h x) { return (h)fabsf((float)x); }
  static __device__ __forceinline__ bool nan_(h x) { return isnan((float)x); }
  static __device__ __forceinline__ h expvar(const u4& q) {
    return f2h(-logf((float)((q.r[0] >> 8) + 1) * 0x1p-24f));
  }
  static __device__ __forceinline__ h uscat(const u4& q, bool first) {
    return f2h(first ? u11(q.r[1]) : u11(q.r[0]));
  }
};

// Julia's min propagates NaN.
template <class T> __device__ __forceinline__ T jmin(T a, T b) {
  const T m = (b < a) ? b : a;
  return Tm<T>::nan_(a) ? a : (Tm<T>::nan_(b) ? b : m);      // selects, no branches
}

// Birth uniforms: double -> 53-bit pairs (two per Philox block), float -> 24-bit words.
template <class Cc> struct BirthU;
template <> struct BirthU<double> {
  static __device__ __forceinline__ double u(uint64_t b, int j, uint32_t seed, uint32_t step) {
    const u4 q = philox((uint32_t)b, (uint32_t)(b >> 32), TAG_BIRTH, (uint32_t)(j >> 1), seed, step);
    return (j & 1) ? u53(q.r[2], q.r[3]) : u53(q.r[0], q.r[1]);
  }
  static __device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
};
template <> struct BirthU<float> {
  static __device__ __forceinline__ float u(uint64_t b, int j, uint32_t seed, uint32_t step) {
    const u4 q = philox((uint32_t)b, (uint32_t)(b >> 32), TAG_BIRTH, (uint32_t)(j >> 2), seed, step);
    return u24(q.r[j & 3]);
  }
  static __device__ __forceinline__ float sqrt_(float x) { return sqrtf(x); }
};

// ---------------------------------------------------------------------------------
// Utilities.sorter (imc_utilities.jl:23-54) for a compile-time number of values.
// Values are already rounded to W; the scale is appended, the M values sorted by an
// odd-even transposition network in registers, and smallest*largest pairs multiplied
// (pair formed in FP64 when promote64) with each pair rounded to W.
// ---------------------------------------------------------------------------------
template <class W, int N>
__device__ __forceinline__ int sorter_dev(const double (&vals)[N], bool promote64,
                                          const double* scales, int K, W& prod, double& esc) {
  constexpr int M = N + 1;
  for (int j = 0; j < K; ++j) {
    double b[M];
#pragma unroll
    for (int i = 0; i < N; ++i) b[i] = vals[i];
    b[N] = scales[j];
#pragma unroll
    for (int r = 0; r < M; ++r) {
#pragma unroll
      for (int i = (r & 1); i + 1 < M; i += 2) {
        const double lo = b[i] < b[i + 1] ? b[i] : b[i + 1];
        const double hi = b[i] < b[i + 1] ? b[i + 1] : b[i];
        b[i] = lo; b[i + 1] = hi;
      }
    }
    W p = (W)1;
#pragma unroll
    for (int i = 0; i < M / 2; ++i) {
      const W pair = promote64 ? (W)(b[i] * b[M - 1 - i]) : ((W)b[i] * (W)b[M - 1 - i]);
      p = p * pair;
    }
    if (M & 1) p = p * (W)b[M / 2];
    if (isfinite((double)p)) { prod = p; esc = scales[j]; return j + 1; }
  }
  prod = (W)0; esc = 0.0;
  return 0;
}

// ---------------------------------------------------------------------------------
// Wave64 helpers
// ---------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

template <class A> __device__ __forceinline__ A wave_sum(A v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Neumaier-compensated accumulation of v into (s, c).
__device__ __forceinline__ void neumaier(float& s, float& c, float v) {
  const float t = s + v;
  if (fabsf(s) >= fabsf(v)) c += (s - t) + v;
  else c += (v - t) + s;
  s = t;
}
__device__ __forceinline__ void neumaier(double& s, double& c, double v) {
  const double t = s + v;
  if (fabs(s) >= fabs(v)) c += (s - t) + v;
  else c += (v - t) + s;
  s = t;
}

// Tally storage modes (all keep per-block partial sums that a deterministic compensated
// reduction combines across blocks, except GLOBAL):
//  BLOCKROW: one LDS row per block; every lane adds its deposit with an LDS atomic.
//            Each block bin sums only that block's deposits (a few hundred terms), so FP32
//            rounding stays near 1e-6 relative before the compensated cross-block sum.
//  WAVEROW : every wave owns a private LDS row of (sum, comp) pairs; lanes hitting the
//            same bin are first summed by a tree inside the wave, then the leader lane
//            applies a Neumaier update -- no atomics, compensated, deterministic.
//  GLOBAL  : lanes add straight into the global FP64 bins with atomics (meshes whose
//            bins do not fit in LDS).
enum TallyMode { TM_WAVEROW = 0, TM_BLOCKROW = 1, TM_GLOBAL = 2, TM_HASH = 3 };

// TM_HASH: an open LDS table of 2^hbits (bin, value) slots per block for meshes whose bin
// row does not fit in LDS: values [2^hbits] then bin tags [2^hbits] (-1 = free).  A bin
// claims its slot (multiplicative hash) on first touch; deposits to claimed bins are LDS
// atomics, a bin whose slot holds another bin goes to the global copy, and the block adds
// every claimed slot to the global copy once at the end.
constexpr int HASH_BITS_MIN = 8, HASH_BITS_MAX = 13;
__device__ __forceinline__ uint32_t hash_slot(int bin, int hbits) {
  return ((uint32_t)bin * 2654435761u) >> (32 - hbits);
}

template <class Acc, int TM>
struct Tally {
  Acc* lds;          // WAVEROW: [nwaves][NB][2]; BLOCKROW: [NB]
  double* gbins;     // GLOBAL, HASH: [NB], this block's copy
  int NB;
  int hbits;         // HASH: log2 of the table's slots

  // All 64 lanes must call this together.  bin < 0 marks "no deposit".
  __device__ __forceinline__ void add(int bin, Acc val) {
    if (TM == TM_BLOCKROW) {
      if (bin >= 0) atomicAdd(&lds[bin], val);
      return;
    }
    if (TM == TM_GLOBAL) {
      if (bin >= 0) atomicAdd(&gbins[bin], (double)val);
      return;
    }
    if (TM == TM_HASH) {
      if (bin < 0) return;
      int* tags = (int*)(lds + (1 << hbits));
      const uint32_t s = hash_slot(bin, hbits);
      int t = tags[s];
      if (t == -1) {
        t = atomicCAS(&tags[s], -1, bin);
        if (t == -1) t = bin;
      }
      if (t == bin) atomicAdd(&lds[s], val);
      else atomicAdd(&gbins[bin], (double)val);
      return;
    }
    const int lane = lane_id();
    uint64_t pend = __ballot(bin >= 0);
    while (pend) {
      const int leader = __ffsll((unsigned long long)pend) - 1;
      const int b = __shfl(bin, leader, 64);
      const bool mine = (bin == b);
      Acc v = wave_sum(mine ? val : (Acc)0);
      if (lane == leader) {
        Acc* row = lds + (size_t)(threadIdx.x >> 6) * NB * 2;
        Acc s = row[2 * b], c = row[2 * b + 1];
        neumaier(s, c, v);
        row[2 * b] = s; row[2 * b + 1] = c;
      }
      pend &= ~__ballot(mine);
    }
  }
};

}  // namespace imc
