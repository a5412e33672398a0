// imc_oracle.cpp -- CPU restatement of MixedPrecisionIMC.jl's per-step particle engine.
//
// TEST INFRASTRUCTURE ONLY (the checker of libimc_hip; never linked by the product).
// Each function cites the reference file:line it restates.  Random numbers follow the
// Philox keying described in imc_oracle.h instead of Julia's global RNG stream.
//
// Precision model (mirrors the product's specification):
//   prec 64: tracking T=double, weights W=double, birth arithmetic in double
//   prec 32: T=float,  W=float,  birth arithmetic in float
//   prec 16: T=binary16 (emulated: every op computed in float and rounded RNE to
//            binary16, exact by the 2p+2 double-rounding theorem), W=float,
//            birth arithmetic in float
// Per-cell coefficients (sigma_a(1-f)+sigma_s, f*sigma_a, dx*s) are formed in FP64 and
// rounded once to the tracking/weight type (double -> float -> half for prec 16).
// Tallies are accumulated in double.
#include "imc_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t TAG_BIRTH = 1u;
constexpr uint32_t TAG_EVENT = 2u;
// A history still running after this many events (only a non-finite state can) ends like
// the 1% cut-off, as in the engine; the reference would loop forever.
constexpr uint32_t ORC_EV_MAX = 1u << 24;
constexpr double PI_D = 3.141592653589793;
constexpr double TWO_PI_D = 6.283185307179586;

// ---------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers: as easy
// as 1, 2, 3").  Multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key bumps 0x9E3779B9 /
// 0xBB67AE85, ten rounds, key bumped before every round but the first.
// ---------------------------------------------------------------------------------
struct U4 { uint32_t r[4]; };

inline U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
  for (int round = 0; round < 10; ++round) {
    if (round) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  return U4{{c0, c1, c2, c3}};
}

inline double u53(uint32_t hi, uint32_t lo) {            // [0,1), 53 bits
  return (double)((((uint64_t)hi) << 21) | (lo >> 11)) * 0x1p-53;
}
inline float u24(uint32_t a) { return (float)(a >> 8) * 0x1p-24f; }   // [0,1), 24 bits
inline float u11(uint32_t a) { return (float)(a >> 21) * 0x1p-11f; }  // [0,1), 11 bits

// ---------------------------------------------------------------------------------
// binary16 emulation
// ---------------------------------------------------------------------------------
double round_half_d(double x) {
  if (std::isnan(x) || std::isinf(x)) return x;
  const double a = std::fabs(x);
  if (a >= 65520.0) return std::copysign(INFINITY, x);   // 65504 + half ulp ties to even -> inf
  double q;
  if (a < 0x1p-14) {
    q = 0x1p-24;                                          // subnormal quantum
  } else {
    int e;
    std::frexp(a, &e);                                    // a in [2^(e-1), 2^e)
    q = std::ldexp(1.0, e - 11);                          // 11 significant bits
  }
  return std::copysign(std::nearbyint(a / q) * q, x);     // a/q exact; nearbyint is RNE
}

struct h16 {
  float v;
  h16() : v(0.f) {}
  explicit h16(float f) : v((float)round_half_d((double)f)) {}
  explicit h16(int i) : v((float)round_half_d((double)i)) {}
};
inline h16 operator+(h16 a, h16 b) { return h16(a.v + b.v); }
inline h16 operator-(h16 a, h16 b) { return h16(a.v - b.v); }
inline h16 operator*(h16 a, h16 b) { return h16(a.v * b.v); }
inline h16 operator/(h16 a, h16 b) { return h16(a.v / b.v); }
inline h16 operator-(h16 a) { h16 r; r.v = -a.v; return r; }
inline bool operator<(h16 a, h16 b) { return a.v < b.v; }
inline bool operator>(h16 a, h16 b) { return a.v > b.v; }
inline bool operator<=(h16 a, h16 b) { return a.v <= b.v; }
inline bool operator==(h16 a, h16 b) { return a.v == b.v; }
inline bool operator!=(h16 a, h16 b) { return a.v != b.v; }

template <class T> struct Tr;
template <> struct Tr<double> {
  static double maxval() { return 1.7976931348623157e308; }   // largest finite value
  static double from(double x) { return x; }
  static double d(double x) { return x; }
  static double exp(double x) { return std::exp(x); }
  static double expm1(double x) { return std::expm1(x); }
  static double cos(double x) { return std::cos(x); }
  static double sin(double x) { return std::sin(x); }
  static double atan2(double y, double x) { return std::atan2(y, x); }
  static double abs(double x) { return std::fabs(x); }
  static bool isnan(double x) { return std::isnan(x); }
  static double zero() { return 0.0; }
  static double one() { return 1.0; }
  // transport draws: exponential variate from (r0,r1); scatter uniform from (r2,r3) or,
  // on a redraw, (r0,r1) of the redraw block.
  static double expvar(const U4& q) {
    return -std::log((double)(((((uint64_t)q.r[0]) << 21) | (q.r[1] >> 11)) + 1) * 0x1p-53);
  }
  static double uscat(const U4& q, bool first) { return first ? u53(q.r[2], q.r[3]) : u53(q.r[0], q.r[1]); }
};
template <> struct Tr<float> {
  static double maxval() { return 3.4028234663852886e38; }
  static float from(double x) { return (float)x; }
  static double d(float x) { return (double)x; }
  static float exp(float x) { return expf(x); }
  static float expm1(float x) { return expm1f(x); }
  static float cos(float x) { return cosf(x); }
  static float sin(float x) { return sinf(x); }
  static float atan2(float y, float x) { return atan2f(y, x); }
  static float abs(float x) { return fabsf(x); }
  static bool isnan(float x) { return std::isnan(x); }
  static float zero() { return 0.f; }
  static float one() { return 1.f; }
  static float expvar(const U4& q) { return -logf((float)((q.r[0] >> 8) + 1) * 0x1p-24f); }
  static float uscat(const U4& q, bool first) { return first ? u24(q.r[1]) : u24(q.r[0]); }
};
template <> struct Tr<h16> {
  static double maxval() { return 65504.0; }
  static h16 from(double x) { return h16((float)x); }
  static double d(h16 x) { return (double)x.v; }
  static h16 exp(h16 x) { return h16(expf(x.v)); }
  static h16 expm1(h16 x) { return h16(expm1f(x.v)); }
  static h16 cos(h16 x) { return h16(cosf(x.v)); }
  static h16 sin(h16 x) { return h16(sinf(x.v)); }
  static h16 atan2(h16 y, h16 x) { return h16(atan2f(y.v, x.v)); }
  static h16 abs(h16 x) { h16 r; r.v = fabsf(x.v); return r; }
  static bool isnan(h16 x) { return std::isnan(x.v); }
  static h16 zero() { return h16(0.f); }
  static h16 one() { return h16(1.f); }
  static h16 expvar(const U4& q) { return h16(-logf((float)((q.r[0] >> 8) + 1) * 0x1p-24f)); }
  static h16 uscat(const U4& q, bool first) { return h16(first ? u11(q.r[1]) : u11(q.r[0])); }
};

// Julia's min for floats propagates NaN.
template <class T> inline T jmin(T a, T b) {
  if (Tr<T>::isnan(a)) return a;
  if (Tr<T>::isnan(b)) return b;
  return (b < a) ? b : a;
}

// Birth arithmetic type: double for prec 64, float otherwise.
template <class Cc> struct Birth;
template <> struct Birth<double> {
  static double u(uint64_t b, int j, uint32_t seed, uint32_t step) {
    const U4 q = philox((uint32_t)b, (uint32_t)(b >> 32), TAG_BIRTH, (uint32_t)(j >> 1), seed, step);
    const int i = (j & 1) * 2;
    return u53(q.r[i], q.r[i + 1]);
  }
  static double sqrt(double x) { return std::sqrt(x); }
};
template <> struct Birth<float> {
  static float u(uint64_t b, int j, uint32_t seed, uint32_t step) {
    const U4 q = philox((uint32_t)b, (uint32_t)(b >> 32), TAG_BIRTH, (uint32_t)(j >> 2), seed, step);
    return u24(q.r[j & 3]);
  }
  static float sqrt(float x) { return sqrtf(x); }
};

int scale_index(const orc_problem* pb, double scale) {
  for (int k = 0; k < pb->K; ++k)
    if (pb->scales[k] == scale) return k;          // findfirst(isequal(scale), energyscales)
  return 0;
}

// ---------------------------------------------------------------------------------
// Utilities.sorter (imc_utilities.jl:23-54).  For each scale (given in descending
// order) sort [vals; scale], multiply smallest*largest pairs each rounded to W (the
// pair product is formed in Float64 when the value vector was promoted to Float64 by
// a Float64 literal, e.g. 0.25 at imc_sourcing.jl:60), times the middle element when
// the length is odd; return the first finite product with its scale and 1-based index.
// ---------------------------------------------------------------------------------
template <class W>
int sorter_w(const double* vals, int n, bool promote64, const double* scales, int K, W* prod,
             double* scale) {
  double buf[40];
  for (int j = 0; j < K; ++j) {
    for (int i = 0; i < n; ++i) buf[i] = vals[i];
    buf[n] = scales[j];
    const int m = n + 1;
    std::sort(buf, buf + m);
    W product = (W)1;
    for (int i = 0; i < m / 2; ++i) {
      if (promote64) product = product * (W)(buf[i] * buf[m - 1 - i]);
      else product = product * ((W)buf[i] * (W)buf[m - 1 - i]);
    }
    if (m & 1) product = product * (W)buf[m / 2];
    if (std::isfinite((double)product)) {
      *prod = product; *scale = scales[j];
      return j + 1;
    }
  }
  *prod = (W)0; *scale = 0.0;
  return 0;
}

// Deterministic FP64 sum: blocks of 256 reduced by a halving tree (s[j] += s[j+h],
// h = 128..1), repeated over the block sums.  The GPU reduction uses the same tree.
double tree256(std::vector<double> v) {
  if (v.empty()) return 0.0;
  while (v.size() > 1) {
    const size_t nb = (v.size() + 255) / 256;
    std::vector<double> out(nb);
    for (size_t b = 0; b < nb; ++b) {
      double s[256];
      for (int j = 0; j < 256; ++j) {
        const size_t i = b * 256 + j;
        s[j] = i < v.size() ? v[i] : 0.0;
      }
      for (int h = 128; h >= 1; h >>= 1)
        for (int j = 0; j < h; ++j) s[j] += s[j + h];
      out[b] = s[0];
    }
    v.swap(out);
  }
  return v[0];
}

inline int nthreads_of(int req) {
#ifdef _OPENMP
  return req > 0 ? req : omp_get_max_threads();
#else
  (void)req;
  return 1;
#endif
}
inline int thread_id() {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

// ---------------------------------------------------------------------------------
// Sourcing.sourcing (imc_sourcing.jl:12-370)
// ---------------------------------------------------------------------------------
enum EntryKind { E_FACE_L1 = 0, E_FACE_R1, E_FACE_B, E_FACE_T, E_FACE_L, E_FACE_R, E_BODY, E_SRC };
struct Entry { int kind; int cell; double e; double esc; int k; int64_t n; double nrg_div; };

template <class T, class W, class Cc>
int64_t source_t(const orc_problem* pb, const orc_step* st, uint32_t step, int64_t n_census,
                 int64_t n_input, int64_t n_max, double cellmin, double* rec, int64_t cap,
                 double* emitted, double* etot) {
  const int nx = pb->nx, ny = pb->geometry == 2 ? pb->ny : 1, C = nx * ny, K = pb->K;
  const bool is2d = pb->geometry == 2;
  const double a = (double)(W)pb->phys_a, c = (double)(W)pb->phys_c;
  const double dt = (double)(W)st->dt, s = (double)(W)pb->distancescale;
  std::vector<Entry> ents;
  auto face = [&](int kind, int cell, double ts, double width, bool has_width) {
    const double Tw = (double)(W)ts;
    double v[9]; int n = 0;
    v[n++] = a; v[n++] = c; v[n++] = Tw; v[n++] = Tw; v[n++] = Tw; v[n++] = Tw;
    if (has_width) v[n++] = (double)(W)width;
    v[n++] = dt; v[n++] = 0.25;
    W e; double esc;
    const int k = sorter_w<W>(v, n, true, pb->scales, K, &e, &esc);
    ents.push_back(Entry{kind, cell, (double)e, esc, k > 0 ? k - 1 : 0, 0, 0.0});
  };
  if (!is2d) {
    face(E_FACE_L1, 0, st->tsurf[0][0], 0.0, false);             // imc_sourcing.jl:60
    face(E_FACE_R1, C - 1, st->tsurf[1][0], 0.0, false);         // imc_sourcing.jl:61
  } else {
    for (int i = 0; i < nx; ++i) face(E_FACE_B, i, st->tsurf[0][i], st->dx[i], true);   // :87
    for (int i = 0; i < nx; ++i) face(E_FACE_T, i, st->tsurf[1][i], st->dx[i], true);   // :88
    for (int j = 0; j < ny; ++j) face(E_FACE_L, j, st->tsurf[2][j], st->dy[j], true);   // :91
    for (int j = 0; j < ny; ++j) face(E_FACE_R, j, st->tsurf[3][j], st->dy[j], true);   // :92
  }
  std::vector<Entry> body(C), src(C);
  for (int k = 0; k < C * K; ++k) emitted[k] = 0.0;
  for (int cidx = 0; cidx < C; ++cidx) {
    const int ix = cidx % nx, iy = cidx / nx;
    const double f = (double)(W)st->fleck[cidx], sa = (double)(W)st->sigma_a[cidx];
    const double Tm = (double)(W)st->temp[cidx], dxw = (double)(W)st->dx[ix];
    double v[16]; int n = 0;
    // body: [f, sa, a, c, T, T, T, T, dx, (dy), dt, s]  (imc_sourcing.jl:67 / :102)
    v[n++] = f; v[n++] = sa; v[n++] = a; v[n++] = c; v[n++] = Tm; v[n++] = Tm; v[n++] = Tm; v[n++] = Tm;
    v[n++] = dxw;
    if (is2d) v[n++] = (double)(W)st->dy[iy];
    v[n++] = dt; v[n++] = s;
    W e; double esc; int k = sorter_w<W>(v, n, false, pb->scales, K, &e, &esc);
    body[cidx] = Entry{E_BODY, cidx, (double)e, esc, k > 0 ? k - 1 : 0, 0, 0.0};
    // radiation source: [S, dx, (dy), dt]  (:68 / :103)
    n = 0;
    v[n++] = (double)(W)st->radsource[cidx]; v[n++] = dxw;
    if (is2d) v[n++] = (double)(W)st->dy[iy];
    v[n++] = dt;
    k = sorter_w<W>(v, n, false, pb->scales, K, &e, &esc);
    src[cidx] = Entry{E_SRC, cidx, (double)e, esc, k > 0 ? k - 1 : 0, 0, 0.0};
    // emitted energy per volume: [f, sa, a, c, T, T, T, T, dt, s]  (:69-70 / :104-105)
    n = 0;
    v[n++] = f; v[n++] = sa; v[n++] = a; v[n++] = c; v[n++] = Tm; v[n++] = Tm; v[n++] = Tm; v[n++] = Tm;
    v[n++] = dt; v[n++] = s;
    k = sorter_w<W>(v, n, false, pb->scales, K, &e, &esc);
    if (k > 0) emitted[(k - 1) * C + cidx] = (double)e;
  }
  ents.insert(ents.end(), body.begin(), body.end());
  ents.insert(ents.end(), src.begin(), src.end());

  // Total energy (imc_sourcing.jl:121) -- FP64 tree over e/escale formed in W.
  std::vector<double> terms(ents.size());
  for (size_t i = 0; i < ents.size(); ++i) terms[i] = (double)((W)ents[i].e / (W)ents[i].esc);
  const double E_tot = tree256(terms);
  *etot = E_tot;

  // Particle budget (imc_sourcing.jl:132-136).
  double n_source = (double)n_input;
  const int ndim = is2d ? 2 : 1;
  if (n_input + n_census > n_max) n_source = std::max(cellmin, (double)(n_max - n_census - ndim - 1));

  // Counts (imc_sourcing.jl:138-146, 1D faces :150-157, 2D faces :240-263).
  std::vector<int64_t> nbody(C);
  for (size_t i = 0; i < ents.size(); ++i) {
    Entry& en = ents[i];
    const double x = terms[i] * n_source / E_tot;
    double cnt = 0.0;
    switch (en.kind) {
      case E_FACE_L1: case E_FACE_R1: cnt = en.e > 0 ? std::nearbyint(x) : 0.0; break;
      case E_FACE_B: case E_FACE_T: case E_FACE_L: case E_FACE_R:
        cnt = en.e > 0 ? std::max(std::nearbyint(x), cellmin) : 0.0; break;
      case E_BODY: cnt = std::max(std::nearbyint(x), cellmin); break;
      case E_SRC: cnt = en.e > 0 ? std::max(std::nearbyint(x), cellmin) : 0.0; break;
    }
    en.n = cnt > 0 ? (int64_t)cnt : 0;
    en.nrg_div = (double)en.n;
    if (en.kind == E_BODY) nbody[en.cell] = en.n;
  }
  if (is2d && (pb->quirks & ORC_QUIRK_Q1_RADSRC_NBODY)) {   // imc_sourcing.jl:357 loops n_body
    for (auto& en : ents)
      if (en.kind == E_SRC && en.n > 0) en.n = nbody[en.cell];
  }

  // Births, appended after the census in entry order (1D :161-236, 2D :265-366).
  int64_t total = 0;
  std::vector<int64_t> off(ents.size());
  for (size_t i = 0; i < ents.size(); ++i) { off[i] = total; total += ents[i].n; }
  if (n_census + total > cap) return -1;
  const int RW = is2d ? 10 : 9;
  const uint32_t seed = pb->seed;
  const double sD = pb->distancescale;
  for (size_t ei = 0; ei < ents.size(); ++ei) {
    const Entry& en = ents[ei];
    if (en.n <= 0) continue;
    const W nrg = (W)en.e / (W)en.nrg_div;
    for (int64_t j = 0; j < en.n; ++j) {
      const uint64_t b = (uint64_t)(off[ei] + j);
      double* r = rec + (n_census + off[ei] + j) * RW;
      auto U = [&](int jj) { return Birth<Cc>::u(b, jj, seed, step); };
      const Cc dtc = (Cc)st->dt, sc = (Cc)pb->distancescale;
      if (!is2d) {
        double origin, x, mu, t;
        int cell = en.cell;
        if (en.kind == E_FACE_L1 || en.kind == E_FACE_R1) {
          const bool left = en.kind == E_FACE_L1;
          origin = left ? 1.0 : (double)C;
          const double dxe = (double)(W)st->dx[left ? 0 : C - 1];
          x = Tr<T>::d(Tr<T>::from((left ? 0.01 : 0.99) * dxe * (double)(W)sD));
          Cc m = Birth<Cc>::sqrt(U(1));
          for (int at = 0; m == (Cc)0; ++at) m = Birth<Cc>::sqrt(U(8 + at));
          mu = Tr<T>::d(Tr<T>::from(left ? (double)m : -(double)m));
          t = Tr<T>::d(Tr<T>::from((double)(dtc * U(2))));
        } else {
          origin = (double)(cell + 1);
          x = Tr<T>::d(Tr<T>::from((double)((Cc)st->dx[cell] * U(0) * sc)));
          Cc m = (Cc)1 - (Cc)2 * U(1);
          for (int at = 0; m == (Cc)0; ++at) m = (Cc)1 - (Cc)2 * U(8 + at);
          mu = Tr<T>::d(Tr<T>::from((double)m));
          t = Tr<T>::d(Tr<T>::from((double)(dtc * U(2))));
        }
        r[0] = origin; r[1] = t; r[2] = (double)(cell + 1); r[3] = x; r[4] = mu; r[5] = 1.0;
        r[6] = (double)nrg; r[7] = (double)nrg; r[8] = en.esc;
      } else {
        int ix, iy; double x, y, th, t;
        switch (en.kind) {
          case E_FACE_B: case E_FACE_T: {
            const bool bot = en.kind == E_FACE_B;
            t = Tr<T>::d(Tr<T>::from((double)(dtc * U(0))));
            ix = en.cell; iy = bot ? 0 : ny - 1;
            x = Tr<T>::d(Tr<T>::from((double)((Cc)st->dx[ix] * U(1) * sc)));
            y = Tr<T>::d(Tr<T>::from((bot ? 0.001 : 0.999) * (double)(W)st->dy[iy] * (double)(W)sD));
            th = Tr<T>::d(Tr<T>::from((bot ? PI_D : -PI_D) * (double)U(2)));
            break;
          }
          case E_FACE_L: case E_FACE_R: {
            const bool lf = en.kind == E_FACE_L;
            t = Tr<T>::d(Tr<T>::from((double)(dtc * U(0))));
            ix = lf ? 0 : nx - 1; iy = en.cell;
            x = Tr<T>::d(Tr<T>::from((lf ? 0.001 : 0.999) * (double)(W)st->dx[ix] * (double)(W)sD));
            const double wy = (pb->quirks & ORC_QUIRK_Q2_LR_FACE_DX) ? st->dx[iy] : st->dy[iy];
            y = Tr<T>::d(Tr<T>::from((double)((Cc)wy * U(1) * sc)));
            th = Tr<T>::d(Tr<T>::from(PI_D * (lf ? (0.5 - (double)U(2)) : (0.5 + (double)U(2)))));
            break;
          }
          default: {
            ix = en.cell % nx; iy = en.cell / nx;
            x = Tr<T>::d(Tr<T>::from((double)((Cc)st->dx[ix] * U(0) * sc)));
            y = Tr<T>::d(Tr<T>::from((double)((Cc)st->dy[iy] * U(1) * sc)));
            th = Tr<T>::d(Tr<T>::from(TWO_PI_D * (double)U(2)));
            t = Tr<T>::d(Tr<T>::from((double)(dtc * U(3))));
            break;
          }
        }
        r[0] = t; r[1] = ix + 1; r[2] = iy + 1; r[3] = x; r[4] = y; r[5] = th; r[6] = 1.0;
        r[7] = (double)nrg; r[8] = (double)nrg; r[9] = en.esc;
      }
    }
  }
  return total;
}

// ---------------------------------------------------------------------------------
// Transport.MC (imc_transport.jl:13-210), event semantics of :72-194.
// ---------------------------------------------------------------------------------
template <class T, class W>
void transport1d_t(const orc_problem* pb, const orc_step* st, uint32_t step, double* rec, int64_t n,
                   double* dep, double* lost_out, int64_t* segs_out, int nth) {
  const int C = pb->nx, K = pb->K;
  std::vector<T> dxs(C), isig(C);
  std::vector<W> sigdep(C), idxw(C);
  for (int i = 0; i < C; ++i) {
    const double sa = st->sigma_a[i], f = st->fleck[i], ss = st->sigma_s[i];
    dxs[i] = Tr<T>::from(st->dx[i] * pb->distancescale);
    isig[i] = Tr<T>::from(std::min(1.0 / (sa * (1.0 - f) + ss), Tr<T>::maxval()));    // :87 1/(effective scattering)
    sigdep[i] = (W)(sa * f);                                // :95 / :120 sigma_a*f
    idxw[i] = (W)(1.0 / st->dx[i]);                         // :104 / :120 1/dx
  }
  const T cT = Tr<T>::from(pb->phys_c), sT = Tr<T>::from(pb->distancescale), dtT = Tr<T>::from(st->dt);
  const T icsT = Tr<T>::from(1.0 / (pb->phys_c * pb->distancescale));   // :126 (d/s)/c as d*(1/(c s))
  const T zero = Tr<T>::zero(), one = Tr<T>::one(), two = Tr<T>::from(2.0);
  nth = nthreads_of(nth);
  std::vector<std::vector<double>> tdep(nth, std::vector<double>((size_t)C * K, 0.0));
  std::vector<double> tlost(nth, 0.0);
  std::vector<int64_t> tsegs(nth, 0);
  const uint32_t seed = pb->seed;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nth)
  for (int64_t p = 0; p < n; ++p) {
    const int tid = thread_id();
    double* D = tdep[tid].data();
    double* r = rec + 9 * p;
    T t = Tr<T>::from(r[1]);
    int i = (int)r[2] - 1;
    T x = Tr<T>::from(r[3]), mu = Tr<T>::from(r[4]);
    W E = (W)r[6];
    const W E0 = (W)r[7];
    const double scale = r[8];
    const int k = scale_index(pb, scale);
    const W Emin = (W)(0.01 * (double)E0);                  // :61
    int64_t segs = 0;
    for (uint32_t ev = 0;; ++ev) {
      ++segs;                                               // :73
      const U4 q = philox((uint32_t)p, ev, TAG_EVENT, 0u, seed, step);
      const T db = (mu > zero) ? (dxs[i] - x) / mu : Tr<T>::abs(x / mu);   // :77-83
      const T dc = Tr<T>::expvar(q) * isig[i];              // :87
      const T dcen = cT * (dtT - t) * sT;                   // :89
      const T d = jmin(jmin(db, dc), dcen);                 // :92
      const W tau = sigdep[i] * (W)Tr<T>::d(d);
      const W Enew = E * (W)Tr<W>::exp(-tau);               // :95
      if (Enew <= Emin || ev >= ORC_EV_MAX) {              // :97-106; event cap as the engine
        D[k * C + i] += (double)(E * idxw[i]);
        r[7] = -1.0;
        break;
      }
      D[k * C + i] += (double)(-(E * idxw[i]) * (W)Tr<W>::expm1(-tau));   // :120
      x = x + mu * d;                                       // :124-126
      t = t + d * icsT;
      E = Enew;
      if (d == db) {                                        // :130-170
        bool dead = false;
        if (mu > zero) {
          if (i == C - 1) {
            if (pb->bc[1] == ORC_REFLECT) mu = -mu;
            else { tlost[tid] += (double)(E / (W)scale); r[7] = -1.0; dead = true; }
          }
          if (!dead) { i += 1; x = zero; }
        }
        if (!dead && mu < zero) {
          if (i == 0) {
            if (pb->bc[0] == ORC_REFLECT) mu = -mu;
            else { tlost[tid] += (double)(E / (W)scale); r[7] = -1.0; dead = true; }
          } else {
            i -= 1; x = dxs[i];
          }
        }
        if (dead) break;
      }
      if (d == dc) {                                        // :174-183
        bool first = true;
        U4 qq = q;
        for (uint32_t at = 1;; ++at) {
          mu = one - two * Tr<T>::from((double)Tr<T>::d(Tr<T>::uscat(qq, first)));
          if (mu != zero) break;
          first = false;
          qq = philox((uint32_t)p, ev, TAG_EVENT, at, seed, step);
        }
      }
      if (d == dcen) {                                      // :185-193
        r[1] = 0.0; r[2] = (double)(i + 1); r[3] = Tr<T>::d(x); r[4] = Tr<T>::d(mu); r[6] = (double)E;
        break;
      }
    }
    tsegs[tid] += segs;
  }
  for (int th = 0; th < nth; ++th) {
    for (int j = 0; j < C * K; ++j) dep[j] += tdep[th][j];
    *lost_out += tlost[th];
    *segs_out += tsegs[th];
  }
}

// ---------------------------------------------------------------------------------
// Transport.MC2D (imc_transport.jl:483-732), event semantics of :533-718.
// ---------------------------------------------------------------------------------
template <class T, class W>
void transport2d_t(const orc_problem* pb, const orc_step* st, uint32_t step, double* rec, int64_t n,
                   double* dep, double* lost_out, int64_t* segs_out, int nth) {
  const int nx = pb->nx, ny = pb->ny, C = nx * ny, K = pb->K;
  std::vector<T> dxs(nx), dys(ny), isig(C);
  std::vector<W> sigdep(C), idxw(nx), idyw(ny);
  for (int i = 0; i < nx; ++i) { dxs[i] = Tr<T>::from(st->dx[i] * pb->distancescale); idxw[i] = (W)(1.0 / st->dx[i]); }
  for (int j = 0; j < ny; ++j) { dys[j] = Tr<T>::from(st->dy[j] * pb->distancescale); idyw[j] = (W)(1.0 / st->dy[j]); }
  for (int c = 0; c < C; ++c) {
    const double sa = st->sigma_a[c], f = st->fleck[c], ss = st->sigma_s[c];
    isig[c] = Tr<T>::from(std::min(1.0 / (sa * (1.0 - f) + ss), Tr<T>::maxval()));    // :561
    sigdep[c] = (W)(sa * f);                                // :580 / :607
  }
  const T cT = Tr<T>::from(pb->phys_c), sT = Tr<T>::from(pb->distancescale), dtT = Tr<T>::from(st->dt);
  const T icsT = Tr<T>::from(1.0 / (pb->phys_c * pb->distancescale));   // :618 (d/s)/c as d*(1/(c s))
  const T zero = Tr<T>::zero();
  nth = nthreads_of(nth);
  std::vector<std::vector<double>> tdep(nth, std::vector<double>((size_t)C * K, 0.0));
  std::vector<double> tlost(nth, 0.0);
  std::vector<int64_t> tsegs(nth, 0);
  const uint32_t seed = pb->seed;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nth)
  for (int64_t p = 0; p < n; ++p) {
    const int tid = thread_id();
    double* D = tdep[tid].data();
    double* r = rec + 10 * p;
    T t = Tr<T>::from(r[0]);
    int ix = (int)r[1] - 1, iy = (int)r[2] - 1;
    T x = Tr<T>::from(r[3]), y = Tr<T>::from(r[4]), th = Tr<T>::from(r[5]);
    W E = (W)r[7];
    const W E0 = (W)r[8];
    const double scale = r[9];
    const int k = scale_index(pb, scale);
    const W Emin = (W)(0.01 * (double)E0);                  // :531
    int64_t segs = 0;
    for (uint32_t ev = 0;; ++ev) {
      ++segs;
      const U4 q = philox((uint32_t)p, ev, TAG_EVENT, 0u, seed, step);
      const int c = ix + nx * iy;
      const T ox = Tr<T>::cos(th), oy = Tr<T>::sin(th);     // :534
      const T dbx = (ox > zero) ? Tr<T>::abs((dxs[ix] - x) / ox) : Tr<T>::abs(x / ox);   // :538-542
      const T dby = (oy > zero) ? Tr<T>::abs((dys[iy] - y) / oy) : Tr<T>::abs(y / oy);   // :544-548
      T db;
      if (Tr<T>::isnan(dbx)) db = dby;                      // :551-557
      else if (Tr<T>::isnan(dby)) db = dbx;
      else db = jmin(dbx, dby);
      const T dc = Tr<T>::expvar(q) * isig[c];              // :561
      const T dcen = cT * (dtT - t) * sT;                   // :569
      const T d = jmin(jmin(db, dc), dcen);                 // :571
      const W tau = sigdep[c] * (W)Tr<T>::d(d);
      const W Enew = E * (W)Tr<W>::exp(-tau);               // :580
      if (Enew <= Emin || ev >= ORC_EV_MAX) {              // :586-595; event cap as the engine
        D[k * C + c] += (double)((E * idxw[ix]) * idyw[iy]);
        r[7] = -1.0;
        break;
      }
      D[k * C + c] += (double)(-((E * idxw[ix]) * idyw[iy]) * (W)Tr<W>::expm1(-tau));   // :607
      x = x + d * ox;                                       // :615-618
      y = y + d * oy;
      t = t + d * icsT;
      E = Enew;
      if (d == dbx || d == dby) {                           // :621-704
        bool dead = false;
        if (dbx < dby) {
          if (ox > zero) {
            if (ix == nx - 1) {
              if (pb->bc[1] == ORC_REFLECT) th = Tr<T>::atan2(oy, -ox);
              else dead = true;
            } else { ix += 1; x = zero; }
          } else {
            if (ix == 0) {
              if (pb->bc[0] == ORC_REFLECT) th = Tr<T>::atan2(oy, -ox);
              else dead = true;
            } else { ix -= 1; x = dxs[ix]; }
          }
        } else {
          if (oy > zero) {
            if (iy == ny - 1) {
              if (pb->bc[2] == ORC_REFLECT) th = Tr<T>::atan2(-oy, ox);
              else dead = true;
            } else { iy += 1; y = zero; }
          } else {
            if (iy == 0) {
              if (pb->bc[3] == ORC_REFLECT) th = Tr<T>::atan2(-oy, ox);
              else dead = true;
            } else { iy -= 1; y = dys[iy]; }
          }
        }
        if (dead) { tlost[tid] += (double)(E / (W)scale); r[7] = -1.0; break; }
        continue;                                           // :703
      }
      if (d == dc)                                          // :706-710
        th = Tr<T>::from(TWO_PI_D * Tr<T>::d(Tr<T>::uscat(q, true)));
      if (d == dcen) {                                      // :712-717
        r[0] = 0.0; r[1] = ix + 1; r[2] = iy + 1; r[3] = Tr<T>::d(x); r[4] = Tr<T>::d(y);
        r[5] = Tr<T>::d(th); r[7] = (double)E;
        break;
      }
    }
    tsegs[tid] += segs;
  }
  for (int thr = 0; thr < nth; ++thr) {
    for (int j = 0; j < C * K; ++j) dep[j] += tdep[thr][j];
    *lost_out += tlost[thr];
    *segs_out += tsegs[thr];
  }
}

template <class W>
void census_t(const orc_problem* pb, const orc_step* st, const double* rec, int64_t n, double* rad) {
  const int nx = pb->nx, ny = pb->geometry == 2 ? pb->ny : 1;
  for (int c = 0; c < nx * ny; ++c) rad[c] = 0.0;
  if (pb->geometry == 1) {                                  // imc_tally.jl:163-173
    for (int64_t p = 0; p < n; ++p) {
      const double* r = rec + 9 * p;
      const int i = (int)r[2] - 1;
      rad[i] += (double)((W)r[6] / ((W)st->dx[i] * (W)r[8]));
    }
  } else {                                                  // imc_tally.jl:174-190
    for (int64_t p = 0; p < n; ++p) {
      const double* r = rec + 10 * p;
      const int ix = (int)r[1] - 1, iy = (int)r[2] - 1;
      rad[ix + nx * iy] += (double)((W)r[7] / (((W)st->dx[ix] * (W)st->dy[iy]) * (W)r[9]));
    }
  }
}

// Host-side FP64 sorter (imc_update.jl:39 / :52, imc_tally.jl:108 with scale 1).
double sorter64(std::vector<double> v, double scale) {
  double prod; double sc;
  sorter_w<double>(v.data(), (int)v.size(), false, &scale, 1, &prod, &sc);
  return prod;
}

}  // namespace

// ===================================================================================
extern "C" {

void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  const U4 q = philox(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = q.r[i];
}

double orc_round_half(double x) { return round_half_d((double)(float)x); }

int orc_sorter(int prec, const double* vals, int n, int promote64, const double* scales, int K,
               double* product, double* scale) {
  if (prec == 64) {
    double p; const int k = sorter_w<double>(vals, n, promote64 != 0, scales, K, &p, scale);
    *product = p; return k;
  }
  float p; const int k = sorter_w<float>(vals, n, promote64 != 0, scales, K, &p, scale);
  *product = (double)p; return k;
}

// Birth order of the engine: a step's births are stored in a stable order by start-time
// bucket min(15, floor(16 t/dt)), so histories with the most time left are transported
// first.  The reference appends births in entry order (imc_sourcing.jl:173,276); the order
// only decides which slot, and so which event stream, a birth gets.
static void order_births(const orc_problem* pb, const orc_step* st, double* rec, int64_t n_census, int64_t nb) {
  if (nb <= 1) return;
  const int RW = pb->geometry == 2 ? 10 : 9, TI = pb->geometry == 2 ? 0 : 1;
  double* b0 = rec + (size_t)RW * n_census;
  std::vector<int64_t> cnt(17, 0);
  std::vector<int> key((size_t)nb);
  for (int64_t i = 0; i < nb; ++i) {
    int q = (int)(b0[(size_t)RW * i + TI] / st->dt * 16.0);
    q = q < 0 ? 0 : (q > 15 ? 15 : q);
    key[(size_t)i] = q;
    ++cnt[(size_t)q + 1];
  }
  for (int q = 0; q < 16; ++q) cnt[(size_t)q + 1] += cnt[(size_t)q];
  std::vector<double> out((size_t)RW * nb);
  for (int64_t i = 0; i < nb; ++i) {
    const int64_t d = cnt[(size_t)key[(size_t)i]]++;
    std::memcpy(&out[(size_t)RW * d], b0 + (size_t)RW * i, sizeof(double) * RW);
  }
  std::memcpy(b0, out.data(), sizeof(double) * out.size());
}

int64_t orc_source(const orc_problem* pb, const orc_step* st, uint32_t step, int64_t n_census,
                   int64_t n_input, int64_t n_max, double cellmin, double* rec, int64_t cap,
                   double* emitted, double* etot) {
  int64_t nb = -2;
  switch (pb->prec) {
    case 64: nb = source_t<double, double, double>(pb, st, step, n_census, n_input, n_max, cellmin, rec, cap, emitted, etot); break;
    case 32: nb = source_t<float, float, float>(pb, st, step, n_census, n_input, n_max, cellmin, rec, cap, emitted, etot); break;
    case 16: nb = source_t<h16, float, float>(pb, st, step, n_census, n_input, n_max, cellmin, rec, cap, emitted, etot); break;
  }
  if (nb > 0) order_births(pb, st, rec, n_census, nb);
  return nb;
}

int orc_transport(const orc_problem* pb, const orc_step* st, uint32_t step, double* rec, int64_t n,
                  double* energydep, double* lost, int64_t* segments, int nthreads) {
  const int C = pb->nx * (pb->geometry == 2 ? pb->ny : 1);
  for (int j = 0; j < C * pb->K; ++j) energydep[j] = 0.0;   // imc_transport.jl:45 / :499
  *segments = 0;
  if (pb->geometry == 1) {
    switch (pb->prec) {
      case 64: transport1d_t<double, double>(pb, st, step, rec, n, energydep, lost, segments, nthreads); return 0;
      case 32: transport1d_t<float, float>(pb, st, step, rec, n, energydep, lost, segments, nthreads); return 0;
      case 16: transport1d_t<h16, float>(pb, st, step, rec, n, energydep, lost, segments, nthreads); return 0;
    }
  } else {
    switch (pb->prec) {
      case 64: transport2d_t<double, double>(pb, st, step, rec, n, energydep, lost, segments, nthreads); return 0;
      case 32: transport2d_t<float, float>(pb, st, step, rec, n, energydep, lost, segments, nthreads); return 0;
      case 16: transport2d_t<h16, float>(pb, st, step, rec, n, energydep, lost, segments, nthreads); return 0;
    }
  }
  return -2;
}

// Clean.clean (imc_clean.jl:6-19): drop records whose slot 8 is -1.0, keeping order.
int64_t orc_clean(const orc_problem* pb, double* rec, int64_t n) {
  const int RW = pb->geometry == 2 ? 10 : 9;
  int64_t w = 0;
  for (int64_t p = 0; p < n; ++p) {
    if (rec[p * RW + 7] == -1.0) continue;
    if (w != p) std::memmove(rec + w * RW, rec + p * RW, sizeof(double) * RW);
    ++w;
  }
  return w;
}

void orc_census_tally(const orc_problem* pb, const orc_step* st, const double* rec, int64_t n,
                      double* radenergydens) {
  if (pb->prec == 64) census_t<double>(pb, st, rec, n, radenergydens);
  else census_t<float>(pb, st, rec, n, radenergydens);
}

// Full run: MixedPrecisionIMC.main loop (MixedPrecisionIMC.jl:136-176) with the host
// layers restated in FP64: Update.update (imc_update.jl:12-70), Tally.tally
// (imc_tally.jl:11-143), EnergyCheck.energychecker (imc_energycheck.jl:10-38) and
// timestep (MixedPrecisionIMC.jl:181-222).
int orc_run(const orc_run_cfg* cfg, orc_run_out* out) {
  const orc_problem* pb = &cfg->pb;
  const bool is2d = pb->geometry == 2;
  const int nx = pb->nx, ny = is2d ? pb->ny : 1, C = nx * ny, K = pb->K;
  const int RW = is2d ? 10 : 9;
  const double a = pb->phys_a, c = pb->phys_c, s = pb->distancescale, alpha = cfg->alpha;
  std::vector<double> temp(C, cfg->t_init), bee(cfg->bee, cfg->bee + C), beta(C, 1.0), fleck(C, 0.0),
      sa(C), ss(C), mat(C, 0.0), rad(C, 0.0), dep((size_t)C * K), emit((size_t)C * K), ones(ny, 1.0);
  std::vector<double> rec;
  int64_t n = 0;
  double t = 0.0, dt = cfg->timestepping == 0 ? cfg->dt : cfg->dt0, radold = 0.0, lost = 0.0;
  out->steps_done = 0; out->histories = 0; out->segments = 0; out->t_hot = 0.0;
  using clk = std::chrono::steady_clock;
  while (t <= cfg->t_end && (cfg->max_steps <= 0 || out->steps_done < cfg->max_steps)) {
    // Update.update
    for (int i = 0; i < C; ++i) {
      const double T = temp[i];
      if (cfg->linearized) { bee[i] = 4.0 * a * (T * T * T); beta[i] = 1.0; }
      else beta[i] = (4.0 * a * (T * T * T)) / bee[i];
      sa[i] = cfg->marshak ? cfg->sa0[i] / T / T / T : cfg->sa0[i] * std::pow(T, cfg->sa_pow[i]);
      ss[i] = cfg->ss0[i] * std::pow(T, cfg->ss_pow[i]);
      fleck[i] = 1.0 / (1.0 + sorter64({s, alpha, beta[i], c, dt, sa[i]}, 1.0));
    }
    orc_step st;
    st.dt = dt; st.dx = cfg->dx; st.dy = is2d ? cfg->dy : ones.data(); st.temp = temp.data();
    st.fleck = fleck.data(); st.sigma_a = sa.data(); st.sigma_s = ss.data(); st.radsource = cfg->radsource;
    for (int q = 0; q < 4; ++q) st.tsurf[q] = cfg->tsurf[q];
    const uint32_t stepidx = (uint32_t)out->steps_done;
    const auto t0 = clk::now();
    // Sourcing: grow the record buffer as needed.
    double etot = 0.0;
    int64_t nb;
    for (;;) {
      const int64_t cap = (int64_t)(rec.size() / RW);
      nb = orc_source(pb, &st, stepidx, n, cfg->n_input, cfg->n_max, cfg->cellmin, rec.data(), cap, emit.data(), &etot);
      if (nb >= 0) break;
      rec.resize((size_t)std::max<int64_t>(2 * cap, n + cfg->n_input + 4 * C * (int64_t)(cfg->cellmin + 1) + 1024) * RW);
    }
    n += nb;
    int64_t segs = 0;
    orc_transport(pb, &st, stepidx, rec.data(), n, dep.data(), &lost, &segs, cfg->nthreads);
    const bool counted = out->steps_done >= cfg->time_from;
    if (counted) {
      out->histories += n;
      out->segments += segs;
    }
    n = orc_clean(pb, rec.data(), n);
    orc_census_tally(pb, &st, rec.data(), n, rad.data());
    if (counted) out->t_hot += std::chrono::duration<double>(clk::now() - t0).count();
    // Tally.tally (host part)
    if (t == 0.0)
      for (int i = 0; i < C; ++i)
        mat[i] = sorter64({fleck[i], sa[i], a, c, temp[i], temp[i], temp[i], temp[i], dt, s}, 1.0);
    double totaldep = 0.0;
    for (int i = 0; i < C; ++i) {
      double inc = 0.0;
      for (int k = 0; k < K; ++k) {
        inc += (dep[(size_t)k * C + i] - emit[(size_t)k * C + i]) / pb->scales[k];
        const double vol = is2d ? cfg->dx[i % nx] * cfg->dy[i / nx] : cfg->dx[i];
        totaldep += dep[(size_t)k * C + i] * vol / pb->scales[k];
      }
      mat[i] += inc;
      if (cfg->linearized) temp[i] = std::pow(mat[i], 0.25);
      else temp[i] = temp[i] + inc / bee[i];
    }
    // EnergyCheck.energychecker
    double erad = 0.0;
    for (int i = 0; i < C; ++i)
      erad += is2d ? rad[i] * cfg->dx[i % nx] * cfg->dy[i / nx] : rad[i] * cfg->dx[i];
    const double err = (etot - totaldep - (erad - radold) - lost) / etot;
    if (out->energy_error && out->steps_done < out->err_cap) out->energy_error[out->steps_done] = err;
    radold = erad;
    lost = 0.0;
    out->steps_done += 1;
    // timestep
    if (t == cfg->t_end) { t = cfg->t_end + dt; continue; }
    if (cfg->timestepping == 1 && dt < cfg->dtmax) {
      dt = dt * cfg->k;
      if (dt > cfg->dtmax) dt = cfg->dtmax;
    }
    if (t + dt > cfg->t_end) { dt = cfg->t_end - t; t = cfg->t_end; }
    else t = t + dt;
  }
  for (int i = 0; i < C; ++i) {
    if (out->temp) out->temp[i] = temp[i];
    if (out->radenergydens) out->radenergydens[i] = rad[i];
    if (out->matenergydens) out->matenergydens[i] = mat[i];
  }
  out->t_final = t;
  out->n_alive = n;
  return 0;
}

// Planck frequency sampling (imc_sourcing.jl:372-399, the Fleck-Cummings mixture of
// Gamma(4, 1/n) laws with weights 90 / (pi^4 n^4)), restated with the engine's Philox layout:
// sample i uses blocks (i, TAG_PLANCK = 6, j) keyed (seed, stream); n by FP64 inversion for
// n <= 64, beyond that by rejection from the continuous x^-4 law (n = ceil(64 / cbrt(1 - v)),
// accepted with probability 3 n^-4 / ((n-1)^-3 - n^-3)); x = -(sum of four logs) / n.
// u_force >= 0 replaces the term uniform (to exercise the tail in tests).
static double planck_term(int64_t i, uint32_t seed, uint32_t stream, double u_force, U4* q0o, U4* q1o, U4* q2o) {
  constexpr uint32_t TAG_PLANCK = 6u;
  const uint32_t i0 = (uint32_t)i, i1 = (uint32_t)((uint64_t)i >> 32);
  const U4 q0 = philox(i0, i1, TAG_PLANCK, 0u, seed, stream);
  const U4 q1 = philox(i0, i1, TAG_PLANCK, 1u, seed, stream);
  const U4 q2 = philox(i0, i1, TAG_PLANCK, 2u, seed, stream);
  *q0o = q0; *q1o = q1; *q2o = q2;
  const double u = u_force >= 0.0 ? u_force : u53(q0.r[0], q0.r[1]);
  const double w = 90.0 / (PI_D * PI_D * PI_D * PI_D);
  double cum = 0.0, m = 0.0;
  for (int k = 1; k <= 64; ++k) {
    const double dk = (double)k;
    cum += w / (dk * dk * dk * dk);
    if (u < cum) { m = dk; break; }
  }
  if (m == 0.0) {
    double v = u53(q2.r[2], q2.r[3]);
    for (int t = 0; t < 64; ++t) {
      m = std::ceil(64.0 / std::cbrt(1.0 - v));
      if (m <= 64.0) m = 65.0;
      const double a = 1.0 / (m - 1.0), b = 1.0 / m;
      const double accept = 3.0 * (b * b * b * b) / (a * a * a - b * b * b);
      const U4 qt = philox(i0, i1, TAG_PLANCK, 3u + (uint32_t)t, seed, stream);
      if (u53(qt.r[0], qt.r[1]) < accept) break;
      v = u53(qt.r[2], qt.r[3]);
    }
  }
  return m;
}

void orc_planck(int64_t n, const double* temp, double* freq, uint32_t seed, uint32_t stream) {
  const auto open0 = [](uint32_t hi, uint32_t lo) { return u53(hi, lo) + 0x1p-53; };
  for (int64_t i = 0; i < n; ++i) {
    U4 q0, q1, q2;
    const double m = planck_term(i, seed, stream, -1.0, &q0, &q1, &q2);
    const double s = std::log(open0(q0.r[2], q0.r[3])) + std::log(open0(q1.r[0], q1.r[1])) +
                     std::log(open0(q1.r[2], q1.r[3])) + std::log(open0(q2.r[0], q2.r[1]));
    freq[i] = temp[i] * (-s / m);
  }
}

// The mixture term n of samples 0 .. n-1 (u_force < 0: the sampler's own uniform).
void orc_planck_terms(int64_t n, uint32_t seed, uint32_t stream, double u_force, int64_t* terms) {
  for (int64_t i = 0; i < n; ++i) {
    U4 q0, q1, q2;
    terms[i] = (int64_t)planck_term(i, seed, stream, u_force, &q0, &q1, &q2);
  }
}

}  // extern "C"
