// imc_kernels.hip -- gfx950 kernels of the IMC particle engine.
//
//   k_prep            per-cell tracking coefficients (FP64 -> tracking precision)
//   k_source_entries  emission / radiation-source / surface energies via the sorter
//   k_tree256         deterministic FP64 tree sum (mesh.totalenergy)
//   k_counts          particle budget and per-entry birth counts
//   k_scan_tile_sums, k_scan_tiles, k_counts_total: multi-block exclusive scan of birth counts
//   k_births          one thread per birth: position, direction, time, weight
//   k_transport       persistent event loop, one lane per history with lane refill,
//                     wave-reduced tallies
//   k_reduce_partials per-bin reduction of the per-block tally partials
//   k_clean_*         order-preserving compaction fused with the census tally
#include <hip/hip_runtime.h>

#include "imc_device.h"
#include "imc_internal.h"
#include "imc_mlp.h"

namespace imc {

constexpr int TPB = 256;           // threads per block of the persistent kernels
constexpr int NWAVES = TPB / 64;
// Slots a wave claims from the work queue at a time.  Plain transport: 128 (many short
// histories; 64 and 32 were 6% and 30% slower on the Marshak wave from queue contention).  With
// the surrogate: 32 (fewer, longer histories, where a wave's unstarted backlog lengthens the
// tail when the queue runs dry; 7% faster than 128 on the thick/thin blocks).
constexpr uint32_t CHUNK = 128, CHUNK_SURR = 32;
constexpr int CLEAN_ITEMS = 8;     // slots per thread per clean tile
constexpr int CLEAN_TILE = TPB * CLEAN_ITEMS;

// ---------------------------------------------------------------------------------
__global__ void k_fill_u8(unsigned char* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void k_fill_u64(unsigned long long* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// out[j] = sum over copies of copies[k][j], in copy order.
__global__ void k_sum_copies(const double* __restrict__ copies, int ncopies, int64_t n, double* __restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    double sum = 0.0;
    for (int k = 0; k < ncopies; ++k) sum += copies[(size_t)k * n + j];
    out[j] = sum;
  }
}

hipError_t launch_sum_copies(Ctx& c, const double* copies, int ncopies, int64_t n, double* out) {
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
  k_sum_copies<<<grid, 256, 0, c.stream>>>(copies, ncopies, n, out);
  return hipGetLastError();
}

hipError_t launch_zero(Ctx& c, void* p, size_t bytes) {
  if (bytes == 0) return hipSuccess;
  if ((bytes & 7) == 0 && (((uintptr_t)p) & 7) == 0) {
    const size_t n = bytes / 8;
    const int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    k_fill_u64<<<grid, 256, 0, c.stream>>>((unsigned long long*)p, n);
  } else {
    k_fill_u8<<<(unsigned)((bytes + 255) / 256), 256, 0, c.stream>>>((unsigned char*)p, bytes);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Per-cell coefficients: FP64 formulas rounded once to the working types.
// ---------------------------------------------------------------------------------
template <int P>
__global__ void k_prep(int nx, int ny, int C, double s, const double* __restrict__ dx,
                       const double* __restrict__ dy, const double* __restrict__ sa,
                       const double* __restrict__ ss, const double* __restrict__ f,
                       typename Prec<P>::T* dxs, typename Prec<P>::T* dys,
                       typename Prec<P>::W* dxw, typename Prec<P>::W* dyw,
                       typename Prec<P>::T* sigcol, typename Prec<P>::W* sigdep,
                       typename Prec<P>::T* isig, typename Prec<P>::W* idxw, typename Prec<P>::W* idyw,
                       Scalars* sc_out) {
#pragma clang fp contract(off)
  using T = typename Prec<P>::T;
  using W = typename Prec<P>::W;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < C) {
    const double sc = sa[i] * (1.0 - f[i]) + ss[i];
    sigcol[i] = Tm<T>::from(sc);                              // imc_transport.jl:87
    // :87 as a product.  The reciprocal is clamped to the largest finite T: an infinite one
    // (sigma_c = 0, or 1/sigma_c beyond binary16's 65504) would make 0 * inf = NaN when the
    // exponential variate is 0, where the quotient gives 0.
    const T is = Tm<T>::from(fmin(1.0 / sc, Tm<T>::maxval()));
    isig[i] = is;
    // a mean free path that rounds to 0 would make every collision distance 0
    if (sc > 0.0 && Tm<T>::d(is) == 0.0) atomicAdd(&sc_out->prep_underflow, 1u);
    sigdep[i] = (W)(sa[i] * f[i]);                            // imc_transport.jl:95
  }
  if (i < nx) { dxs[i] = Tm<T>::from(dx[i] * s); dxw[i] = (W)dx[i]; idxw[i] = (W)(1.0 / dx[i]); }
  if (i < ny) { dys[i] = Tm<T>::from(dy[i] * s); dyw[i] = (W)dy[i]; idyw[i] = (W)(1.0 / dy[i]); }
}

hipError_t launch_prep(Ctx& c) {
  const int n = std::max(c.C, std::max(c.nx, c.ny));
  const int grid = (n + 255) / 256;
  const double s = c.cfg.distancescale;
#define PREP(P)                                                                                  \
  k_prep<P><<<grid, 256, 0, c.stream>>>(                                                        \
      c.nx, c.ny, c.C, s, c.d_dx, c.d_dy, c.d_sa, c.d_ss, c.d_fleck,                             \
      (Prec<P>::T*)c.d_dxs, (Prec<P>::T*)c.d_dys, (Prec<P>::W*)c.d_dxw, (Prec<P>::W*)c.d_dyw,     \
      (Prec<P>::T*)c.d_sigcol, (Prec<P>::W*)c.d_sigdep, (Prec<P>::T*)c.d_isig, (Prec<P>::W*)c.d_idxw,  \
      (Prec<P>::W*)c.d_idyw, c.d_sc)
  if (c.prec == 64) PREP(64);
  else if (c.prec == 32) PREP(32);
  else PREP(16);
#undef PREP
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Sourcing energies (imc_sourcing.jl:56-111).  Entry order is the birth order of the
// reference: 1D [left face, right face, body cells, source cells]; 2D [bottom faces,
// top faces, left faces, right faces, body cells, source cells].
// ---------------------------------------------------------------------------------
struct SrcArgs {
  int nx, ny, C, K;
  int64_t NE;
  double a, c, dt, s;
  const double* temp; const double* fleck; const double* sa; const double* radsrc;
  const double* tsurf; const double* dx; const double* dy;
  double* ent_e; double* ent_esc; double* ent_term; int32_t* ent_k; double* emitted;
  double scales[16];
};

template <int P, bool TWO_D>
__global__ void k_source_entries(SrcArgs A) {
#pragma clang fp contract(off)
  using W = typename Prec<P>::W;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= A.NE) return;
  const double a = (double)(W)A.a, c = (double)(W)A.c, dt = (double)(W)A.dt, s = (double)(W)A.s;
  const int nf = TWO_D ? 2 * A.nx + 2 * A.ny : 2;
  W en = (W)0; double esc = 0.0; int k = 0;
  if (e < nf) {
    const double Tw = (double)(W)A.tsurf[e];
    if (!TWO_D) {                                            // imc_sourcing.jl:60-61
      const double v[8] = {a, c, Tw, Tw, Tw, Tw, dt, 0.25};
      k = sorter_dev<W, 8>(v, true, A.scales, A.K, en, esc);
    } else {                                                 // imc_sourcing.jl:86-93
      const int ei = (int)e;
      const double width = ei < 2 * A.nx ? A.dx[ei % A.nx] : A.dy[(ei - 2 * A.nx) % A.ny];
      const double v[9] = {a, c, Tw, Tw, Tw, Tw, (double)(W)width, dt, 0.25};
      k = sorter_dev<W, 9>(v, true, A.scales, A.K, en, esc);
    }
  } else if (e < nf + A.C) {                                 // body (+ emitted energy)
    const int cell = (int)(e - nf);
    const int ix = cell % A.nx, iy = cell / A.nx;
    const double f = (double)(W)A.fleck[cell], sa = (double)(W)A.sa[cell], Tm_ = (double)(W)A.temp[cell];
    const double dxw = (double)(W)A.dx[ix];
    if (!TWO_D) {                                            // imc_sourcing.jl:67
      const double v[11] = {f, sa, a, c, Tm_, Tm_, Tm_, Tm_, dxw, dt, s};
      k = sorter_dev<W, 11>(v, false, A.scales, A.K, en, esc);
    } else {                                                 // imc_sourcing.jl:102
      const double v[12] = {f, sa, a, c, Tm_, Tm_, Tm_, Tm_, dxw, (double)(W)A.dy[iy], dt, s};
      k = sorter_dev<W, 12>(v, false, A.scales, A.K, en, esc);
    }
    // emitted energy per volume (imc_sourcing.jl:69-70 / :104-105)
    const double v2[10] = {f, sa, a, c, Tm_, Tm_, Tm_, Tm_, dt, s};
    W em; double esc2;
    const int k2 = sorter_dev<W, 10>(v2, false, A.scales, A.K, em, esc2);
    for (int kk = 0; kk < A.K; ++kk) A.emitted[(size_t)kk * A.C + cell] = 0.0;
    if (k2 > 0) A.emitted[(size_t)(k2 - 1) * A.C + cell] = (double)em;
  } else {                                                   // radiation source
    const int cell = (int)(e - nf - A.C);
    const int ix = cell % A.nx, iy = cell / A.nx;
    const double S = (double)(W)A.radsrc[cell], dxw = (double)(W)A.dx[ix];
    if (!TWO_D) {                                            // imc_sourcing.jl:68
      const double v[3] = {S, dxw, dt};
      k = sorter_dev<W, 3>(v, false, A.scales, A.K, en, esc);
    } else {                                                 // imc_sourcing.jl:103
      const double v[4] = {S, dxw, (double)(W)A.dy[iy], dt};
      k = sorter_dev<W, 4>(v, false, A.scales, A.K, en, esc);
    }
  }
  A.ent_e[e] = (double)en;
  A.ent_esc[e] = esc;
  A.ent_k[e] = k > 0 ? k - 1 : 0;
  A.ent_term[e] = (double)(en / (W)esc);
}

hipError_t launch_source_entries(Ctx& c) {
  SrcArgs A;
  A.nx = c.nx; A.ny = c.ny; A.C = c.C; A.K = c.K; A.NE = c.NE;
  A.a = c.cfg.phys_a; A.c = c.cfg.phys_c; A.dt = c.dt; A.s = c.cfg.distancescale;
  A.temp = c.d_temp; A.fleck = c.d_fleck; A.sa = c.d_sa; A.radsrc = c.d_radsrc; A.tsurf = c.d_tsurf;
  A.dx = c.d_dx; A.dy = c.d_dy;
  A.ent_e = c.d_ent_e; A.ent_esc = c.d_ent_esc; A.ent_term = c.d_ent_term; A.ent_k = c.d_ent_k;
  A.emitted = c.d_emitted;
  for (int k = 0; k < 16; ++k) A.scales[k] = c.cfg.energyscales[k];
  const unsigned grid = (unsigned)((c.NE + 255) / 256);
#define SRC(P, D) k_source_entries<P, D><<<grid, 256, 0, c.stream>>>(A)
  if (c.prec == 64) { if (c.is2d) SRC(64, true); else SRC(64, false); }
  else if (c.prec == 32) { if (c.is2d) SRC(32, true); else SRC(32, false); }
  else { if (c.is2d) SRC(16, true); else SRC(16, false); }
#undef SRC
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Deterministic FP64 sum: blocks of 256 values reduced by the halving tree
// s[j] += s[j+h], h = 128..1; repeated over block sums until one value remains.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tree256(const double* __restrict__ in, int64_t n,
                                                 double* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double s[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  s[threadIdx.x] = i < n ? in[i] : 0.0;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

hipError_t launch_tree_sum(Ctx& c, const double* in, int64_t n, double* out_scalar) {
  // scratch: two ping-pong areas inside d_red
  const double* src = in;
  int64_t len = n;
  double* bufs[2] = {c.d_red, c.d_red + (c.red_len / 2)};
  int which = 0;
  if (len <= 0) {
    return launch_zero(c, out_scalar, sizeof(double));
  }
  for (;;) {
    const int64_t nb = (len + 255) / 256;
    double* dst = (nb == 1) ? out_scalar : bufs[which];
    k_tree256<<<(unsigned)nb, 256, 0, c.stream>>>(src, len, dst);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    if (nb == 1) break;
    src = dst; len = nb; which ^= 1;
  }
  return hipSuccess;
}

// ---------------------------------------------------------------------------------
// Budget and counts (imc_sourcing.jl:132-157, 240-263).  Counts use FP64:
// x = (e/escale) * n_source / E_tot, n = round-half-even(x), floored at CELLMIN where
// the reference applies it.  ent_e is overwritten by the per-particle energy e/n (W).
// ---------------------------------------------------------------------------------
template <int P, bool TWO_D>
__global__ void k_counts(int nx, int ny, int C, int64_t NE, int quirks, int64_t n_input,
                         int64_t n_max, double cellmin, const Scalars* __restrict__ sc,
                         const double* __restrict__ ent_term, double* __restrict__ ent_e,
                         int64_t* __restrict__ ent_cnt) {
#pragma clang fp contract(off)
  using W = typename Prec<P>::W;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NE) return;
  const int64_t n_census = sc->n_census;
  const double E_tot = sc->E_tot;
  double n_source = (double)n_input;
  const int ndim = TWO_D ? 2 : 1;
  if (n_input + n_census > n_max) n_source = fmax(cellmin, (double)(n_max - n_census - ndim - 1));
  const int nf = TWO_D ? 2 * nx + 2 * ny : 2;
  const double en = ent_e[e];
  const double x = ent_term[e] * n_source / E_tot;
  double cnt;
  if (e < nf) {
    if (TWO_D) cnt = en > 0 ? fmax(rint(x), cellmin) : 0.0;
    else cnt = en > 0 ? rint(x) : 0.0;
  } else if (e < nf + C) {
    cnt = fmax(rint(x), cellmin);
  } else {
    cnt = en > 0 ? fmax(rint(x), cellmin) : 0.0;
  }
  cnt = fmin(cnt, 1e12);           // a non-finite total energy must not overflow the sums
  const int64_t n_div = cnt > 0 ? (int64_t)cnt : 0;
  int64_t n_loop = n_div;
  if (TWO_D && (quirks & IMC_QUIRK_Q1_RADSRC_NBODY) && e >= nf + C && n_div > 0) {
    const int64_t eb = e - C;                                // body entry of the same cell
    const double xb = ent_term[eb] * n_source / E_tot;
    const double cb = fmin(fmax(rint(xb), cellmin), 1e12);
    n_loop = cb > 0 ? (int64_t)cb : 0;
  }
  ent_cnt[e] = n_loop;
  ent_e[e] = n_div > 0 ? (double)((W)en / (W)(double)n_div) : 0.0;
}

// ---------------------------------------------------------------------------------
// Exclusive scan of int64 counts over many blocks: tile sums (k_scan_tile_sums), a
// single-block scan of the tile sums (k_exscan_i64, with the grand total left in entry
// ntiles), then each tile scanned from its offset (k_scan_tiles; in place allowed).
// ---------------------------------------------------------------------------------
constexpr int SCAN_TPB = 256, SCAN_PER = 16, SCAN_TILE = SCAN_TPB * SCAN_PER;

int64_t scan_tiles(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

__device__ __forceinline__ int64_t block_sum_i64(int64_t v, int64_t* wtot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if (lane == 0) wtot[w] = v;
  __syncthreads();
  int64_t s = 0;
  for (int j = 0; j < SCAN_TPB / 64; ++j) s += wtot[j];
  return s;
}

__global__ __launch_bounds__(SCAN_TPB) void k_scan_tile_sums(const int64_t* __restrict__ a, int64_t n,
                                                              int64_t* __restrict__ tsum) {
  __shared__ int64_t wtot[SCAN_TPB / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t v = 0;
#pragma unroll
  for (int j = 0; j < SCAN_PER; ++j) {
    const int64_t i = base + j * SCAN_TPB + threadIdx.x;
    if (i < n) v += a[i];
  }
  const int64_t s = block_sum_i64(v, wtot);
  if (threadIdx.x == 0) tsum[blockIdx.x] = s;
}

// single: one tile, scanned from 0, whose total goes to toff[1] (the tile-sum pass is skipped).
__global__ __launch_bounds__(SCAN_TPB) void k_scan_tiles(const int64_t* a, int64_t n, int64_t* toff, int64_t* out,
                                                          int single) {
  __shared__ int64_t tile[SCAN_TILE];
  __shared__ int64_t wtot[SCAN_TPB / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
  for (int j = 0; j < SCAN_PER; ++j) {                 // coalesced load of the tile
    const int64_t i = base + j * SCAN_TPB + threadIdx.x;
    tile[j * SCAN_TPB + threadIdx.x] = i < n ? a[i] : 0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t* mine = tile + threadIdx.x * SCAN_PER;       // each thread owns SCAN_PER consecutive entries
  int64_t sum = 0;
#pragma unroll
  for (int j = 0; j < SCAN_PER; ++j) sum += mine[j];
  int64_t incl = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) wtot[w] = incl;
  __syncthreads();
  int64_t run = (single ? 0 : toff[blockIdx.x]) + incl - sum;
  for (int j = 0; j < w; ++j) run += wtot[j];
#pragma unroll
  for (int j = 0; j < SCAN_PER; ++j) {
    const int64_t v = mine[j];
    mine[j] = run;
    run += v;
  }
  if (single && threadIdx.x == SCAN_TPB - 1) toff[1] = run;   // the last thread's end is the total
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SCAN_PER; ++j) {
    const int64_t i = base + j * SCAN_TPB + threadIdx.x;
    if (i < n) out[i] = tile[j * SCAN_TPB + threadIdx.x];
  }
}

__global__ __launch_bounds__(1024) void k_exscan_i64(int64_t* __restrict__ a, int64_t n);

// out[i] = a[0] + ... + a[i-1] for i < n, and tiles[scan_tiles(n)] = the total; tiles holds
// scan_tiles(n) + 1 entries.  n >= 1.
hipError_t launch_scan_i64(Ctx& c, const int64_t* a, int64_t n, int64_t* out, int64_t* tiles) {
  const int64_t nt = scan_tiles(n);
  if (nt == 1) {                                             // one launch for a single tile
    k_scan_tiles<<<1, SCAN_TPB, 0, c.stream>>>(a, n, tiles, out, 1);
    return hipGetLastError();
  }
  k_scan_tile_sums<<<(unsigned)nt, SCAN_TPB, 0, c.stream>>>(a, n, tiles);
  k_exscan_i64<<<1, 1024, 0, c.stream>>>(tiles, nt + 1);     // entry nt (any value) ends as the total
  k_scan_tiles<<<(unsigned)nt, SCAN_TPB, 0, c.stream>>>(a, n, tiles, out, 0);
  return hipGetLastError();
}

// Birth total of the scanned counts (imc_sourcing.jl:126-158 budget), with the slot check.
__global__ void k_counts_total(const int64_t* __restrict__ total, Scalars* sc, int64_t cap) {
  if (threadIdx.x == 0) {
    const int64_t running = *total;
    if (running < 0 || running > cap - sc->n_census) {   // more births than slots: none, flagged
      sc->n_births = 0;
      sc->n_total = sc->n_census;
      sc->birth_overflow = 1u;
    } else {
      sc->n_births = running;
      sc->n_total = sc->n_census + running;
      sc->birth_overflow = 0u;
    }
  }
}

hipError_t launch_counts_and_scan(Ctx& c, int64_t n_input, int64_t n_max, double cellmin, int64_t cap) {
  const unsigned grid = (unsigned)((c.NE + 255) / 256);
#define CNT(P, D)                                                                               \
  k_counts<P, D><<<grid, 256, 0, c.stream>>>(c.nx, c.ny, c.C, c.NE, c.cfg.quirks, n_input,     \
                                             n_max, cellmin, c.d_sc, c.d_ent_term, c.d_ent_e,    \
                                             c.d_ent_cnt)
  if (c.prec == 64) { if (c.is2d) CNT(64, true); else CNT(64, false); }
  else if (c.prec == 32) { if (c.is2d) CNT(32, true); else CNT(32, false); }
  else { if (c.is2d) CNT(16, true); else CNT(16, false); }
#undef CNT
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = launch_scan_i64(c, c.d_ent_cnt, c.NE, c.d_ent_off, c.d_ent_tiles);
  if (err != hipSuccess) return err;
  k_counts_total<<<1, 64, 0, c.stream>>>(c.d_ent_tiles + scan_tiles(c.NE), c.d_sc, cap);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Births (imc_sourcing.jl:161-236 in 1D, :265-366 in 2D).  One thread per birth b;
// its entry is found by binary search over the exclusive prefix of counts.  Uniforms:
// Philox(b_lo, b_hi, TAG_BIRTH, group) keyed by (seed, step); draw j of a birth is
// word j of the groups (53-bit pairs for FP64, 24-bit words otherwise); direction
// redraws use draws 8, 9, ...
// ---------------------------------------------------------------------------------
struct BirthArgs {
  int nx, ny, C, K, quirks;
  int64_t NE, n_births;
  uint32_t seed, step;
  double dt, s;
  const double* dx; const double* dy;
  const double* ent_nrg; const double* ent_esc; const int32_t* ent_k; const int64_t* ent_off;
  const Scalars* sc;
  SoA out;
};

// One birth's state (everything but the slot it goes to).
template <int P> struct Birth {
  typename Prec<P>::T x, y, mu, t;
  uint32_t cell;
};

// Entry of birth b: last e with off[e] <= b (off in LDS when the entries fit, else HBM).
__device__ __forceinline__ int64_t birth_entry(const int64_t* __restrict__ off, int64_t NE, int64_t b) {
  int64_t lo = 0, hi = NE - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (off[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

constexpr int64_t ENT_LDS_MAX = 4096;    // entry offsets staged in LDS up to this many entries

// Stages the entry offsets in LDS (when they fit) and returns the table to search.
__device__ __forceinline__ const int64_t* stage_entries(const BirthArgs& A, int64_t* lds) {
  if (A.NE > ENT_LDS_MAX) return A.ent_off;
  for (int64_t i = threadIdx.x; i < A.NE; i += blockDim.x) lds[i] = A.ent_off[i];
  __syncthreads();
  return lds;
}

template <int P, bool TWO_D>
__device__ __forceinline__ Birth<P> birth_state(const BirthArgs& A, int64_t b, int64_t e) {
#pragma clang fp contract(off)
  using T = typename Prec<P>::T;
  using W = typename Prec<P>::W;
  using Cc = typename Prec<P>::Cc;
  const uint64_t ub = (uint64_t)b;
  const Cc dtc = (Cc)A.dt, sc = (Cc)A.s;
  const double sW = (double)(W)A.s;
  Birth<P> o;
  o.y = Tm<T>::from(0.0);
  if (!TWO_D) {
    const int C = A.C;
    if (e < 2) {                                             // surface births :161-189
      const bool left = (e == 0);
      o.cell = left ? 0u : (uint32_t)(C - 1);
      const double dxe = (double)(W)A.dx[o.cell];
      o.x = Tm<T>::from((left ? 0.01 : 0.99) * dxe * sW);
      Cc m = BirthU<Cc>::sqrt_(BirthU<Cc>::u(ub, 1, A.seed, A.step));
      for (int at = 0; m == (Cc)0; ++at) m = BirthU<Cc>::sqrt_(BirthU<Cc>::u(ub, 8 + at, A.seed, A.step));
      o.mu = Tm<T>::from(left ? (double)m : -(double)m);
      o.t = Tm<T>::from((double)(dtc * BirthU<Cc>::u(ub, 2, A.seed, A.step)));
    } else {                                                 // body :193-214, source :217-236
      o.cell = (uint32_t)((e - 2) % C);
      o.x = Tm<T>::from((double)((Cc)A.dx[o.cell] * BirthU<Cc>::u(ub, 0, A.seed, A.step) * sc));
      Cc m = (Cc)1 - (Cc)2 * BirthU<Cc>::u(ub, 1, A.seed, A.step);
      for (int at = 0; m == (Cc)0; ++at) m = (Cc)1 - (Cc)2 * BirthU<Cc>::u(ub, 8 + at, A.seed, A.step);
      o.mu = Tm<T>::from((double)m);
      o.t = Tm<T>::from((double)(dtc * BirthU<Cc>::u(ub, 2, A.seed, A.step)));
    }
  } else {
    const int nx = A.nx, ny = A.ny;
    const int ei = (int)e;
    int ix, iy;
    if (ei < 2 * nx) {                                       // bottom/top :265-293
      const bool bot = ei < nx;
      ix = ei % nx; iy = bot ? 0 : ny - 1;
      o.t = Tm<T>::from((double)(dtc * BirthU<Cc>::u(ub, 0, A.seed, A.step)));
      o.x = Tm<T>::from((double)((Cc)A.dx[ix] * BirthU<Cc>::u(ub, 1, A.seed, A.step) * sc));
      o.y = Tm<T>::from((bot ? 0.001 : 0.999) * (double)(W)A.dy[iy] * sW);
      o.mu = Tm<T>::from((bot ? PI_D : -PI_D) * (double)BirthU<Cc>::u(ub, 2, A.seed, A.step));
    } else if (ei < 2 * nx + 2 * ny) {                       // left/right :295-323
      const bool lf = ei < 2 * nx + ny;
      iy = (ei - 2 * nx) % ny; ix = lf ? 0 : nx - 1;
      o.t = Tm<T>::from((double)(dtc * BirthU<Cc>::u(ub, 0, A.seed, A.step)));
      o.x = Tm<T>::from((lf ? 0.001 : 0.999) * (double)(W)A.dx[ix] * sW);
      const double wy = (A.quirks & IMC_QUIRK_Q2_LR_FACE_DX) ? A.dx[iy] : A.dy[iy];
      o.y = Tm<T>::from((double)((Cc)wy * BirthU<Cc>::u(ub, 1, A.seed, A.step) * sc));
      const double u2 = (double)BirthU<Cc>::u(ub, 2, A.seed, A.step);
      o.mu = Tm<T>::from(PI_D * (lf ? (0.5 - u2) : (0.5 + u2)));
    } else {                                                 // body / source :326-366
      const int cellc = (int)((e - (2 * nx + 2 * ny)) % A.C);
      ix = cellc % nx; iy = cellc / nx;
      o.x = Tm<T>::from((double)((Cc)A.dx[ix] * BirthU<Cc>::u(ub, 0, A.seed, A.step) * sc));
      o.y = Tm<T>::from((double)((Cc)A.dy[iy] * BirthU<Cc>::u(ub, 1, A.seed, A.step) * sc));
      o.mu = Tm<T>::from(TWO_PI_D * (double)BirthU<Cc>::u(ub, 2, A.seed, A.step));
      o.t = Tm<T>::from((double)(dtc * BirthU<Cc>::u(ub, 3, A.seed, A.step)));
    }
    o.cell = (uint32_t)(ix + nx * iy);
  }
  return o;
}

// Birth order.  The births of a step are stored in a stable order by start-time bucket
// floor(BK t/dt): histories with the most time left (the longest on average) enter the
// transport work queue first, which shortens the tail of the persistent launch.  The
// order changes which slot (and so which event stream) a birth gets, not its state.
constexpr int BK = 16;
constexpr int BIRTH_ITEMS = 8;
constexpr int BIRTH_TILE = 256 * BIRTH_ITEMS;

__device__ __forceinline__ int t_bucket(double t, double dt) {
  int q = (int)(t / dt * (double)BK);
  return q < 0 ? 0 : (q > BK - 1 ? BK - 1 : q);
}

// Per-tile bucket counts, bucket-major: hist[bucket * ntiles + tile].
template <int P, bool TWO_D>
__global__ __launch_bounds__(256) void k_birth_hist(BirthArgs A, int64_t ntiles, int64_t* __restrict__ hist) {
  __shared__ int cnt[BK];
  extern __shared__ int64_t s_off[];
  if (threadIdx.x < BK) cnt[threadIdx.x] = 0;
  const int64_t* off = stage_entries(A, s_off);
  __syncthreads();
  const int64_t tile = blockIdx.x;
  const int64_t nbirths = A.sc->n_births;
#pragma unroll 1
  for (int j = 0; j < BIRTH_ITEMS; ++j) {
    const int64_t b = tile * BIRTH_TILE + (int64_t)j * 256 + threadIdx.x;
    if (b < nbirths) {
      const Birth<P> o = birth_state<P, TWO_D>(A, b, birth_entry(off, A.NE, b));
      atomicAdd(&cnt[t_bucket(Tm<typename Prec<P>::T>::d(o.t), A.dt)], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < BK) hist[(int64_t)threadIdx.x * ntiles + tile] = cnt[threadIdx.x];
}

// Single-block exclusive scan of n int64 values in place: each thread owns a contiguous
// chunk (serial sum, block scan of the chunk sums, serial rewrite).
__global__ __launch_bounds__(1024) void k_exscan_i64(int64_t* __restrict__ a, int64_t n) {
  __shared__ int64_t wtot[16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t per = (n + 1023) / 1024;
  const int64_t lo = (int64_t)tid * per, hi = lo + per < n ? lo + per : n;
  int64_t sum = 0;
  for (int64_t i = lo; i < hi; ++i) sum += a[i];
  int64_t incl = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) wtot[w] = incl;
  __syncthreads();
  int64_t run = incl - sum;
  for (int j = 0; j < w; ++j) run += wtot[j];
  for (int64_t i = lo; i < hi; ++i) {
    const int64_t v = a[i];
    a[i] = run;
    run += v;
  }
}

// One block per tile of BIRTH_TILE births (imc_sourcing.jl:161-236 in 1D, :265-366 in 2D).
// Birth b's uniforms: Philox(b_lo, b_hi, TAG_BIRTH, group) keyed by (seed, step); draw j
// of a birth is word j of the groups (53-bit pairs for FP64, 24-bit words otherwise);
// direction redraws use draws 8, 9, ...  Its slot is n_census + its stable rank by
// start-time bucket (off = scanned k_birth_hist counts).
template <int P, bool TWO_D>
__global__ __launch_bounds__(256) void k_births(BirthArgs A, int64_t ntiles, const int64_t* __restrict__ off) {
  using T = typename Prec<P>::T;
  using W = typename Prec<P>::W;
  __shared__ int run[BK];
  __shared__ int wc[4][BK];
  extern __shared__ int64_t s_off[];
  const int64_t* eoff = stage_entries(A, s_off);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  if (tid < BK) run[tid] = 0;
  const int64_t tile = blockIdx.x;
  const int64_t n_census = A.sc->n_census;
  const int64_t nbirths = A.sc->n_births;
  T* X = (T*)A.out.x; T* MU = (T*)A.out.mu; T* TT = (T*)A.out.t;
  W* EE = (W*)A.out.E; W* E0 = (W*)A.out.E0;
#pragma unroll 1
  for (int j = 0; j < BIRTH_ITEMS; ++j) {
    const int64_t b = tile * BIRTH_TILE + (int64_t)j * 256 + tid;
    const bool valid = b < nbirths;
    Birth<P> o;
    int64_t e = 0;
    int bk = 0;
    if (valid) {
      e = birth_entry(eoff, A.NE, b);
      o = birth_state<P, TWO_D>(A, b, e);
      bk = t_bucket(Tm<T>::d(o.t), A.dt);
    }
    // lanes of this wave with the same bucket, and this lane's rank among them
    uint64_t same = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) {
      const uint64_t on = __ballot(valid && ((bk >> bit) & 1));
      same &= ((bk >> bit) & 1) ? on : ~on;
    }
    const int rank = __popcll(same & lt_mask);
    if (tid < 4 * BK) wc[tid / BK][tid % BK] = 0;
    __syncthreads();
    if (valid && rank == 0) wc[w][bk] = __popcll(same);
    __syncthreads();
    if (valid) {
      int pre = run[bk];
      for (int v = 0; v < w; ++v) pre += wc[v][bk];
      const int64_t slot = n_census + off[(int64_t)bk * ntiles + tile] + pre + rank;
      const W nrg = (W)A.ent_nrg[e];
      X[slot] = o.x; MU[slot] = o.mu; TT[slot] = o.t;
      if (TWO_D) ((T*)A.out.y)[slot] = o.y;
      A.out.cellk[slot] = o.cell | ((uint32_t)A.ent_k[e] << KSHIFT);
      EE[slot] = nrg;
      E0[slot] = nrg;
    }
    __syncthreads();
    if (tid < BK) run[tid] += wc[0][tid] + wc[1][tid] + wc[2][tid] + wc[3][tid];
    __syncthreads();
  }
}

// n_births: the step's birth count or an upper bound of it (the kernels read the count from
// the device scalars; tiles past it do nothing).
hipError_t launch_births(Ctx& c, int64_t n_births) {
  if (n_births <= 0) return hipSuccess;
  BirthArgs A;
  A.nx = c.nx; A.ny = c.ny; A.C = c.C; A.K = c.K; A.quirks = c.cfg.quirks;
  A.NE = c.NE; A.n_births = n_births;
  A.seed = c.cfg.seed; A.step = (uint32_t)c.step;
  A.dt = c.dt; A.s = c.cfg.distancescale;
  A.dx = c.d_dx; A.dy = c.d_dy;
  A.ent_nrg = c.d_ent_e; A.ent_esc = c.d_ent_esc; A.ent_k = c.d_ent_k; A.ent_off = c.d_ent_off;
  A.sc = c.d_sc; A.out = c.buf[c.cur];
  const int64_t ntiles = (n_births + BIRTH_TILE - 1) / BIRTH_TILE;
  if ((int64_t)BK * ntiles > c.bhist_len) return hipErrorInvalidValue;   // sized by the caller (ensure_bhist)
  const unsigned grid = (unsigned)ntiles;
  const size_t elds = c.NE <= ENT_LDS_MAX ? sizeof(int64_t) * (size_t)c.NE : 0;
  hipError_t err = hipSuccess;
#define BIR(P, D)                                                                               \
  do {                                                                                          \
    k_birth_hist<P, D><<<grid, 256, elds, c.stream>>>(A, ntiles, c.d_bhist);                    \
    err = launch_scan_i64(c, c.d_bhist, (int64_t)BK * ntiles, c.d_bhist, c.d_bhist_tiles);      \
    if (err != hipSuccess) break;                                                               \
    k_births<P, D><<<grid, 256, elds, c.stream>>>(A, ntiles, c.d_bhist);                        \
    err = hipGetLastError();                                                                    \
  } while (0)
  if (c.prec == 64) { if (c.is2d) BIR(64, true); else BIR(64, false); }
  else if (c.prec == 32) { if (c.is2d) BIR(32, true); else BIR(32, false); }
  else { if (c.is2d) BIR(16, true); else BIR(16, false); }
#undef BIR
  return err;
}

int64_t birth_hist_len(int64_t n_births) {
  return (int64_t)BK * ((n_births + BIRTH_TILE - 1) / BIRTH_TILE);
}

// ---------------------------------------------------------------------------------
// Tally flush helpers shared by transport and clean.
// ---------------------------------------------------------------------------------
template <class Acc, int TM>
__device__ __forceinline__ void tally_init(Acc* lds, int NB, int hbits = 0) {
  if (TM == TM_HASH) {
    const int slots = 1 << hbits;
    int* tags = (int*)(lds + slots);
    for (int i = threadIdx.x; i < slots; i += blockDim.x) { lds[i] = (Acc)0; tags[i] = -1; }
  } else if (TM == TM_WAVEROW) {
    for (int i = threadIdx.x; i < NWAVES * NB * 2; i += blockDim.x) lds[i] = (Acc)0;
  } else if (TM == TM_BLOCKROW) {
    for (int i = threadIdx.x; i < NB; i += blockDim.x) lds[i] = (Acc)0;
  }
  __syncthreads();
}

// Writes this block's (sum, compensation) per bin to partial[blockIdx.x][bin][2].
template <class Acc, int TM>
__device__ __forceinline__ void tally_flush(const Acc* lds, int NB, Acc* partial) {
  if (TM == TM_GLOBAL || TM == TM_HASH) return;
  __syncthreads();
  Acc* out = partial + (size_t)blockIdx.x * NB * 2;
  for (int j = threadIdx.x; j < NB; j += blockDim.x) {
    Acc s = (Acc)0, cc = (Acc)0;
    if (TM == TM_WAVEROW) {
#pragma unroll
      for (int w = 0; w < NWAVES; ++w) {
        neumaier(s, cc, lds[(size_t)w * NB * 2 + 2 * j]);
        cc += lds[(size_t)w * NB * 2 + 2 * j + 1];
      }
    } else {
      s = lds[j];
    }
    out[2 * j] = s;
    out[2 * j + 1] = cc;
  }
}

// Cross-block reduction, one block per bin: each thread folds a strided subset of the
// block partials with Neumaier compensation, then a fixed halving tree merges the
// threads' (sum, comp) pairs.  The order is fixed, so the result is deterministic for
// given partials.
template <class Acc>
__global__ __launch_bounds__(256) void k_reduce_partials(const Acc* __restrict__ partial, int nblocks,
                                                          int NB, double* __restrict__ out) {
  __shared__ Acc ss[256], sc[256];
  const int j = blockIdx.x;
  Acc s = (Acc)0, cc = (Acc)0;
  for (int g = threadIdx.x; g < nblocks; g += 256) {
    neumaier(s, cc, partial[(size_t)g * NB * 2 + 2 * j]);
    cc += partial[(size_t)g * NB * 2 + 2 * j + 1];
  }
  ss[threadIdx.x] = s;
  sc[threadIdx.x] = cc;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      Acc a = ss[threadIdx.x], ac = sc[threadIdx.x];
      neumaier(a, ac, ss[threadIdx.x + h]);
      ac += sc[threadIdx.x + h];
      ss[threadIdx.x] = a;
      sc[threadIdx.x] = ac;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = (double)ss[0] + (double)sc[0];
}

hipError_t launch_reduce_partials(Ctx& c, int nblocks, int tally_mode, int nbins, double* out) {
  if (tally_mode == TM_GLOBAL || tally_mode == TM_HASH) return hipSuccess;   // no per-block partials
  if (c.prec == 64)
    k_reduce_partials<double><<<nbins, 256, 0, c.stream>>>((const double*)c.d_partial, nblocks, nbins, out);
  else
    k_reduce_partials<float><<<nbins, 256, 0, c.stream>>>((const float*)c.d_partial, nblocks, nbins, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Transport (imc_transport.jl:46-197 in 1D, :515-720 in 2D).
// Persistent blocks; each wave pulls CHUNK-slot (CHUNK_SURR with the surrogate) ranges from a
// device work counter and refills a lane as soon as its history ends.  History in slot p uses Philox counters
// (p, event, TAG_EVENT, redraw) keyed by (seed, step).
//
// SURR (1D): a lane whose cell has collision optical thickness sigma_c*dx inside the
// surrogate's training box takes one network evaluation instead of its in-cell random
// walk (the slot of Transport.MC_RW, imc_transport.jl:286-372): with the predicted
// P(stay) it stays until its path budget ends (census with weight E exp(-f sigma_a c(dt-t)),
// or death by the 1% cut-off), otherwise it leaves through the right face with
// P(right | exit), after a path drawn from the inverse-Gaussian law that matches the
// predicted mean path and expected deposit fraction (weight and clock both follow from that
// path), in an outward direction drawn from mu (1 + 1.5 mu).  Only particles at least
// depth_min mean free paths from both faces use it; plain events resolve the layer next to
// the faces.  Deposits are E_in - E_out, so every history conserves energy exactly.
// ---------------------------------------------------------------------------------
struct SurrArgs {
  const uint16_t* wbf;
  const float* bias;
  const float4* cells;   // per cell {tau_cell if inside the box else -1, sigma_c, dx s, f sigma_a}
  SurrBox box;
  SurrMap map;
  int tally_bytes;
  const float* rw_table; // random walk: inverse CDF of the exit time (RW_TABLE_POINTS values)
  float rw_min;          // random walk: smallest half-width, in collision mean free paths
};

// Per-cell inputs of the surrogate decision, formed once per transport launch.
template <int P>
__global__ void k_surr_cells(int C, SurrBox bx, const typename Prec<P>::T* __restrict__ dxs,
                             const typename Prec<P>::T* __restrict__ sigcol,
                             const typename Prec<P>::W* __restrict__ sigdep, float4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  const float scf = (float)sigcol[i], Lf = (float)dxs[i];
  const float tau = scf * Lf;
  const bool in = tau >= bx.tau_cell_min && tau <= bx.tau_cell_max;
  out[i] = make_float4(in ? tau : -1.0f, scf, Lf, (float)sigdep[i]);
}

constexpr uint32_t TAG_SURR = 3u;
constexpr int SURR_BATCH = 16;     // lanes that trigger one network evaluation (one full MFMA pass;
                                   // 4: -4%, 8: -2%, 32: -7% on the thick/thin blocks)
// Event cap per history.  A history still running after EV_MAX events (only a non-finite
// state can do that) is ended like the 1% cut-off and counted in Scalars::truncated, so
// every lane, and with it every wave, finishes.
constexpr uint32_t EV_MAX = 1u << 24;

// The four per-cell values a 1D event reads, staged once per block in LDS when the
// table is small (one 16-byte LDS read per event instead of four HBM/L2 gathers).
template <int P> struct alignas(16) CellRec {
  typename Prec<P>::T dxs, isig;
  typename Prec<P>::W sigdep, idxw;
};
constexpr size_t CELL_LDS_MAX = 16 * 1024;
constexpr size_t SURR_LDS = (size_t)BF_TOTAL * 2 + 544 + (size_t)NWAVES * sizeof(MlpScratch);

// Exit path s of a surrogate exit, given the network's mean path m = E[s | exit] and mean
// deposit d = E[1 - exp(-sd s) | exit] (sd = f sigma_a): drawn from the inverse-Gaussian
// (first-passage-time) law IG(m, phi m), whose Laplace transform -ln E[exp(-sd s)] =
// phi (sqrt(1 + 2 sd m / phi) - 1) gives back d when phi = L^2 / (2 (sd m - L)), L = -ln(1 - d)
// (Jensen: L < sd m).  Weight and clock then follow from the same s, which keeps the positive
// correlation of remaining weight and remaining time: taking their means separately
// understated what exiting particles deposit downstream (thick interiors ran 0.3-0.5% cool).
// Sampler: Michael, Schucany and Haas (1976), one normal and one uniform, no rejection; the
// smaller root is written without cancellation.
__device__ __forceinline__ float surr_exit_path(float m, float sd, float d, const u4& r) {
  const float lam = sd * m;
  const float L = -__logf(fmaxf(1.0f - d, 0x1p-126f));
  const float gap = lam - L;
  if (!(gap > 0.f) || !(L > 0.f)) return m;                 // degenerate law: the mean path
  const float half_cv2 = fminf(gap / (L * L), 50.0f);       // 1 / (2 phi); CV^2 capped at 100
  const float u1 = (float)((r.r[0] >> 8) + 1) * 0x1p-24f;
  const float z = sqrtf(-2.0f * __logf(u1)) * __cosf(6.2831853f * u24(r.r[1]));
  const float w = z * z * half_cv2;
  const float xr = m / (1.0f + w + sqrtf(w * (w + 2.0f)));
  return u24(r.r[2]) * (m + xr) <= m ? xr : m * m / xr;
}

// Outgoing direction cosine from a diffusive medium, p(mu) = mu (1 + 1.5 mu) on (0, 1]: the
// CDF is mu^2 (1 + mu) / 2; two Newton steps on mu^3 + mu^2 = 2u from sqrt(u) leave the CDF
// within 1e-4 of u.
__device__ __forceinline__ float surr_exit_mu(float u) {
  u = fmaxf(u, 0x1p-24f);
  float m = sqrtf(u);
#pragma unroll
  for (int i = 0; i < 2; ++i) m -= __fdividef(m * m * (m + 1.0f) - 2.0f * u, m * (3.0f * m + 2.0f));
  return fminf(m, 1.0f);
}

// ---------------------------------------------------------------------------------
// RW (1D): the analytic random walk of the Transport.MC_RW slot (imc_transport.jl:212-479,
// corrected).  A lane whose distance R0 to the nearest face is at least rw_min collision mean
// free paths replaces its walk inside [x - R0, x + R0] by the diffusion limit of effective
// scattering: D = c / (3 sigma_c) with the collision cross section sigma_c = sigma_a (1 - f) +
// sigma_s (the reference omitted sigma_s).  In path units the dimensionless time is
// a = s / (3 sigma_c R0^2); the first exit from the interval has the law of one-dimensional
// diffusion started at its centre (an exact table of its inverse CDF, with the closed-form tail
// a = -(4/pi^2) ln(pi (1 - u) / 4)), and the walker leaves at x - R0 or x + R0 with equal
// probability -- one of them is the cell face -- in an outward direction from mu (1 + 1.5 mu).
// A walker still inside at census is placed by the interval's conditional density: the
// fundamental mode cos(pi y / 2) for a >= 0.2, else a Gaussian of variance 2a with the
// first-order images, accepted with probability 1 - e^{(y-1)/a} - e^{-(y+1)/a}.  The weight
// follows the exact path length (E exp(-f sigma_a s)), the deposit is E_in - E_out, and the 1%
// cut-off applies at its own path length, so every history conserves energy.
// ---------------------------------------------------------------------------------
constexpr uint32_t TAG_RW = 5u;

__device__ __forceinline__ float rw_exit_a(const float* __restrict__ tab, float u) {
  const float f = u * (float)RW_TABLE_POINTS;
  const int i = (int)f;
  if (i >= RW_TABLE_POINTS - 1) return -0.40528473f * __logf(0.78539816f * fmaxf(1.0f - u, 0x1p-30f));
  const float w = f - (float)i;
  return tab[i] + w * (tab[i + 1] - tab[i]);
}

__device__ __forceinline__ float rw_census_y(float a, float u, uint32_t slot, uint32_t ev, uint32_t seed,
                                             uint32_t step) {
  if (a >= 0.2f) return 0.63661977f * asinf(2.0f * u - 1.0f);
  const float sg = sqrtf(2.0f * a);
  float y = 0.0f;
  for (uint32_t at = 1; at <= 16; ++at) {
    const u4 r = philox(slot, ev, TAG_RW, at, seed, step);
    const float z = sqrtf(-2.0f * __logf((float)((r.r[0] >> 8) + 1) * 0x1p-24f)) * __cosf(6.2831853f * u24(r.r[1]));
    const float yc = sg * z;
    if (fabsf(yc) >= 1.0f) continue;
    y = yc;
    if (u24(r.r[2]) < 1.0f - __expf((yc - 1.0f) / a) - __expf(-(yc + 1.0f) / a)) break;
  }
  return y;
}

// ---------------------------------------------------------------------------------
// RW (2D): the same walk in the plane.  Directions are angles in the plane (imc_sourcing.jl
// :272-362, imc_transport.jl:710), so effective scattering diffuses with D = c / (2 sigma_c),
// and the walk is taken over the disk of radius R0 = the distance to the nearest face: in path
// units a = s / (2 sigma_c R0^2).  The first exit from the disk's centre has
// P(T <= a) = 1 - sum_k 2 e^{-j_k^2 a} / (j_k J1(j_k)) (j_k the zeros of J0; an exact table of its
// inverse CDF, with the one-term tail a = ln(c1 / (1 - u)) / j1^2), at a uniform point of the
// circle, in a direction psi from the outward normal with density cos psi (1 + (4/pi) cos psi) / 4
// (the P1 emergent law of the plane, as mu (1 + 1.5 mu) is in 1D), set as the normal turned by psi.  A walker still inside at
// census is placed by the disk's conditional density: radius from the fundamental mode
// rho J0(j1 rho) (a table of its inverse CDF rho J1(j1 rho) / J1(j1)) for a >= 0.2, else a
// Gaussian of variance 2a per axis with the first-order image of the nearest boundary, accepted
// with probability 1 - e^{-(1 - rho) / a}; the angle is uniform.  The circle lies inside the cell,
// so the walk never crosses a face.  Weight, deposit and cut-off as in 1D.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ float rw2_exit_a(const float* __restrict__ tab, float u) {
  const float f = u * (float)RW_TABLE_POINTS;
  const int i = (int)f;
  // 1 - u = c1 e^{-j1^2 a}, c1 = 2 / (j1 J1(j1)) = 1.6019747, 1 / j1^2 = 0.17291507
  if (i >= RW_TABLE_POINTS - 1) return 0.17291507f * __logf(1.6019747f / fmaxf(1.0f - u, 0x1p-30f));
  const float w = f - (float)i;
  return tab[i] + w * (tab[i + 1] - tab[i]);
}

// Offset of a census walker from the disk's centre, in units of R0.
__device__ __forceinline__ float2 rw2_census(float a, const float* __restrict__ rad, float u, float v, uint32_t slot,
                                             uint32_t ev, uint32_t seed, uint32_t step) {
  if (a >= 0.2f) {
    const float f = u * (float)(RW_RADIUS_POINTS - 1);
    const int i = min((int)f, RW_RADIUS_POINTS - 2);
    const float rho = rad[i] + (f - (float)i) * (rad[i + 1] - rad[i]);
    float sn, cs;
    __sincosf(6.2831853f * v, &sn, &cs);
    return make_float2(rho * cs, rho * sn);
  }
  const float sg = sqrtf(2.0f * a);
  float2 p = make_float2(0.0f, 0.0f);
  for (uint32_t at = 1; at <= 16; ++at) {
    const u4 r = philox(slot, ev, TAG_RW, at, seed, step);
    const float m = sg * sqrtf(-2.0f * __logf((float)((r.r[0] >> 8) + 1) * 0x1p-24f));
    float sn, cs;
    __sincosf(6.2831853f * u24(r.r[1]), &sn, &cs);
    const float rho = m;                                  // |(m cos, m sin)|
    if (rho >= 1.0f) continue;
    p = make_float2(m * cs, m * sn);
    if (u24(r.r[2]) < 1.0f - __expf(-(1.0f - rho) / a)) break;
  }
  return p;
}

// Exit direction from the outward normal as (sin psi, cos psi), psi with density
// cos psi (1 + (4/pi) cos psi) / 4 on (-pi/2, pi/2): proposals from the cosine law (sin psi = 2u - 1)
// accepted with probability (1 + (4/pi) cos psi) / (1 + 4/pi) (0.88 of them); no trigonometry.
// The first proposal uses (u, v); later ones Philox blocks 18, 19, ...  After 8 rejections the
// last proposal is kept (probability 5e-8, a cosine-law direction).
__device__ __forceinline__ float2 rw2_exit_dir(float u, float v, uint32_t slot, uint32_t ev, uint32_t seed,
                                               uint32_t step) {
  float sn = 2.0f * u - 1.0f, cs = sqrtf(fmaxf(1.0f - sn * sn, 0.0f));
  for (uint32_t at = 0; at < 8u; ++at) {
    if (v * 2.2732395f < 1.0f + 1.2732395f * cs) break;
    const u4 r = philox(slot, ev, TAG_RW, 18u + at, seed, step);
    sn = 2.0f * u24(r.r[0]) - 1.0f;
    cs = sqrtf(fmaxf(1.0f - sn * sn, 0.0f));
    v = u24(r.r[1]);
  }
  return make_float2(sn, cs);
}

template <int P, int TM, bool TWO_D, int ACC, bool CELL>
__global__ __launch_bounds__(TPB) void k_transport(SoA s, Geo g, Scalars* sc, void* partial_v,
                                                   double* gbins, SurrArgs sa) {
  using T = typename Prec<P>::T;
  using W = typename Prec<P>::W;
  using Acc = typename Prec<P>::Acc;
  constexpr bool SURR = ACC == 1, RW = ACC == 2;       // learned surrogate / analytic random walk
  static_assert(!(SURR && TWO_D), "the surrogate is 1D");
  static_assert(!(CELL && TWO_D), "the LDS cell table is 1D");
  extern __shared__ __align__(16) unsigned char smem[];
  Acc* lds = (Acc*)smem;
  const int NB = g.NB, C = g.C, nx = g.nx, ny = g.ny;
  uint16_t* s_w = nullptr;
  float* s_b = nullptr;
  MlpScratch* s_scr = nullptr;
  if (SURR) {
    unsigned char* base = smem + sa.tally_bytes;
    s_w = (uint16_t*)base;
    s_b = (float*)(base + (size_t)BF_TOTAL * 2);
    s_scr = (MlpScratch*)(base + (size_t)BF_TOTAL * 2 + 544);
    for (int i = threadIdx.x; i < BF_TOTAL; i += blockDim.x) s_w[i] = sa.wbf[i];
    for (int i = threadIdx.x; i < BIAS_TOTAL; i += blockDim.x) s_b[i] = sa.bias[i];
  }
  const SurrMap& smap = sa.map;
  const CellRec<P>* s_cell = nullptr;
  if (CELL) {
    CellRec<P>* w = (CellRec<P>*)(smem + g.cell_off);
    for (int i = threadIdx.x; i < C; i += blockDim.x) {
      CellRec<P> r;
      r.dxs = ((const T*)g.dxs)[i]; r.isig = ((const T*)g.isig)[i];
      r.sigdep = ((const W*)g.sigdep)[i]; r.idxw = ((const W*)g.idxw)[i];
      w[i] = r;
    }
    s_cell = w;
  }
  tally_init<Acc, TM>(lds, NB, g.hash_bits);   // its barrier also publishes the cell table
  Tally<Acc, TM> tal{lds, (TM == TM_GLOBAL || TM == TM_HASH) ? gbins + (size_t)(blockIdx.x % g.gcopies) * NB : gbins,
                     NB, g.hash_bits};

  const T* __restrict__ dxs = (const T*)g.dxs;
  const T* __restrict__ dys = (const T*)g.dys;
  const W* __restrict__ sigdep = (const W*)g.sigdep;
  const T* __restrict__ isig = (const T*)g.isig;
  const W* __restrict__ idxw = (const W*)g.idxw;
  const W* __restrict__ idyw = (const W*)g.idyw;
  uint32_t* __restrict__ CK = s.cellk;
  T* __restrict__ X = (T*)s.x;
  T* __restrict__ Y = (T*)s.y;
  T* __restrict__ MU = (T*)s.mu;
  T* __restrict__ TT = (T*)s.t;
  W* __restrict__ EE = (W*)s.E;
  const W* __restrict__ E0 = (const W*)s.E0;

  const uint32_t n = (uint32_t)sc->n_total;
  const int lane = lane_id();
  const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const T zero = (T)0, one = (T)1, two = (T)2;
  const T cT = Tm<T>::from(g.c), sT = Tm<T>::from(g.s), dtT = Tm<T>::from(g.dt);
  const T icsT = Tm<T>::from(g.ics);   // (d/s)/c of imc_transport.jl:126 / :618 as d * (1/(c s))
  auto dx_of = [&](int i) -> T {
    if (CELL) return s_cell[i].dxs;
    return dxs[i];
  };

  bool has = false;
  uint32_t slot = 0, ev = 0;
  int ix = 0, iy = 0, k = 0;
  T x = zero, y = zero, mu = zero, t = zero, ox = zero, oy = zero;
  W E = (W)0, Emin = (W)0;
  uint32_t cur = 0, cend = 0;
  bool done = false;
  unsigned long long segs = 0, cens = 0, nsurr = 0;
  uint32_t ntrunc = 0;
  // Without the LDS cell table, the coefficients of the lane's current cell stay in registers
  // and are reloaded only when the lane enters another cell: a particle in an optically thick
  // cell takes many events per visit, and each reload is a scattered load from a mesh-sized
  // array.  The values are the same; only the loads are skipped.
  int ccell = -1;
  T c_dx = zero, c_is = zero, c_dy = zero;
  W c_sd = (W)0, c_idx = (W)0, c_idy = (W)0;
  // Each lane keeps the deposits to its current bin in an FP64 register sum and hands it to the
  // tally only when it deposits into another bin (and once after the loop).  A particle in an
  // optically thick cell deposits into the same bin for many events per visit; the tally then
  // sees one add per visit instead of one per event (on large meshes each add is a scattered
  // global atomic).  The sum is rounded once to the tally type per visit.
  int abin = -1;
  double aval = 0.0;
  // 2D walk: an exit sets the direction (ox, oy) without forming its angle; the angle is formed
  // (atan2) only if the particle reaches census before its next scattering or reflection.
  bool mu_stale = false;

  for (;;) {
    // ---- lane refill from the wave's chunk ----
    uint64_t need = __ballot(!has);
    while (need != 0ull && !done) {
      if (cur >= cend) {
        uint32_t base = 0;
        constexpr uint32_t chunk = (SURR || (RW && !TWO_D)) ? CHUNK_SURR : CHUNK;   // measured per kernel
        if (lane == 0) base = atomicAdd(&sc->work, chunk);
        base = __shfl(base, 0, 64);
        if (base >= n) { done = true; break; }
        cur = base;
        cend = min(base + chunk, n);
      }
      const uint32_t avail = cend - cur;
      const uint32_t rank = (uint32_t)__popcll(need & lt_mask);
      if (!has && rank < avail) {
        slot = cur + rank;
        const uint32_t ck = CK[slot];
        const int cell = (int)(ck & CELLMASK);
        k = (int)(ck >> KSHIFT);
        if (TWO_D) { ix = cell % nx; iy = cell / nx; y = Y[slot]; }
        else ix = cell;
        x = X[slot]; mu = MU[slot]; t = TT[slot]; E = EE[slot];
        Emin = (W)(0.01 * (double)E0[slot]);                 // :61 / :531
        if (TWO_D) { ox = Tm<T>::cos_(mu); oy = Tm<T>::sin_(mu); }
        ev = 0;
        has = true;
        if (RW && TWO_D) mu_stale = false;
      }
      cur += min((uint32_t)__popcll(need), avail);
      need = __ballot(!has);
    }
    if (!__any(has)) break;

    int bin = -1, lbin = -1;
    Acc val = (Acc)0, lval = (Acc)0;

    // ---- surrogate decision (1D, thick cells inside the training box) ----
    bool want = false;
    float feat[MLP_IN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float sout[MLP_OUT] = {0.f, 0.f, 0.f, 0.f};
    float s_ST = 0.f, s_Scut = 0.f, s_SB = 0.f, s_sc = 0.f, s_sd = 0.f;
    if (SURR) {
      if (has) {
        const float4 cellv = sa.cells[ix];                   // {tau_cell or -1, sigma_c, dx s, f sigma_a}
        const float tau_cell = cellv.x;
        if (tau_cell > 0.f) {
          const float scf = cellv.y, Lf = cellv.z, sdf = cellv.w;
          const float ST = (float)(cT * (dtT - t) * sT);
          const float Scut = sdf > 0.f ? __logf((float)E / (float)Emin) / sdf : INFINITY;
          const float SB = fminf(ST, Scut);
          const float tau_b = scf * SB, kappa = sdf / scf;
          const float depth = scf * fminf((float)x, Lf - (float)x);
          if (depth >= sa.box.depth_min && surr_in_box(sa.box, tau_cell, tau_b, kappa)) {
            want = true;
            s_ST = ST; s_Scut = Scut; s_SB = SB; s_sc = scf; s_sd = sdf;
            surr_features(smap, tau_cell, (float)x / Lf, (float)mu, tau_b, kappa, feat);
          }
        }
      }
      // Batch the network: evaluate once SURR_BATCH lanes want it, or when no lane of the
      // wave has plain work left; until then wanting lanes wait (they do not advance).
      const int nwant = __popcll(__ballot(want));
      if (nwant > 0 && (nwant >= SURR_BATCH || nwant == __popcll(__ballot(has)))) {
        mlp_forward_wave(s_w, s_b, s_scr + (threadIdx.x >> 6), want, feat, sout);
      } else if (want) {
        want = false;
        has = false;          // sits this pass out (restored below)
        s_ST = -1.f;
      }
    }

    // ---- analytic random walk (1D, at least rw_min mean free paths from both faces) ----
    bool rw = false;
    T rwR0 = zero;
    if (RW && has) {
      const T cdx_ = dx_of(ix);
      const T cis_ = CELL ? s_cell[ix].isig : isig[TWO_D ? ix + nx * iy : ix];
      rwR0 = (x < cdx_ - x) ? x : cdx_ - x;
      if (TWO_D) {
        const T cdy_ = dys[iy];
        const T ry = (y < cdy_ - y) ? y : cdy_ - y;
        rwR0 = (ry < rwR0) ? ry : rwR0;
      }
      rw = (float)rwR0 > 0.0f && (float)rwR0 >= sa.rw_min * (float)cis_;
    }

    if (RW && TWO_D && rw) {
      ++segs;
      ++nsurr;
      const u4 q = philox(slot, ev, TAG_RW, 0u, g.seed, g.step);
      const int cell = ix + nx * iy;
      bin = k * C + cell;
      const float isf = (float)isig[cell], sdf = (float)sigdep[cell];
      const W cidxy = idxw[ix] * idyw[iy];
      const float R0 = (float)rwR0;
      const float spa = 2.0f * R0 * R0 / isf;                  // path length per unit of a
      const float s_e = rw2_exit_a(sa.rw_table, u24(q.r[0])) * spa;
      const float s_rem = (float)(cT * (dtT - t) * sT);
      const float s_cut = sdf > 0.f ? __logf((float)E / (float)Emin) / sdf : INFINITY;
      if (s_cut < fminf(s_e, s_rem)) {                        // 1% cut-off first: all deposited
        val = (Acc)(E * cidxy);
        CK[slot] = DEAD;
        has = false;
      } else if (s_e < s_rem) {                                // leaves the disk, inside the cell
        const W Enew = E * (W)__expf(-sdf * s_e);
        val = (Acc)((E - Enew) * cidxy);
        E = Enew;
        t = t + Tm<T>::from((double)s_e) * icsT;
        const float phi = 6.2831853f * u24(q.r[1]);
        float sph, cph;
        __sincosf(phi, &sph, &cph);
        const T cdx_ = dxs[ix], cdy_ = dys[iy];
        const T xn = x + Tm<T>::from((double)(R0 * cph)), yn = y + Tm<T>::from((double)(R0 * sph));
        x = xn < zero ? zero : (xn > cdx_ ? cdx_ : xn);
        y = yn < zero ? zero : (yn > cdy_ ? cdy_ : yn);
        const float2 sc2 = rw2_exit_dir(u24(q.r[2]), u24(q.r[3]), slot, ev, g.seed, g.step);   // (sin, cos) psi
        ox = Tm<T>::from((double)(cph * sc2.y - sph * sc2.x));   // the normal turned by psi
        oy = Tm<T>::from((double)(sph * sc2.y + cph * sc2.x));
        mu_stale = true;
      } else {                                                 // census inside the disk
        const W Enew = E * (W)__expf(-sdf * s_rem);
        val = (Acc)((E - Enew) * cidxy);
        E = Enew;
        const float2 o = rw2_census(s_rem / spa, sa.rw_table + RW_TABLE_POINTS, u24(q.r[2]), u24(q.r[3]),
                                    slot, ev, g.seed, g.step);
        const float xn = fminf(fmaxf((float)x + R0 * o.x, 0.0f), (float)dxs[ix]);
        const float yn = fminf(fmaxf((float)y + R0 * o.y, 0.0f), (float)dys[iy]);
        const u4 q2 = philox(slot, ev, TAG_RW, 17u, g.seed, g.step);
        CK[slot] = (uint32_t)cell | ((uint32_t)k << KSHIFT);
        X[slot] = Tm<T>::from((double)xn);
        Y[slot] = Tm<T>::from((double)yn);
        MU[slot] = Tm<T>::from(TWO_PI_D * (double)u24(q2.r[0]));
        TT[slot] = zero;
        EE[slot] = E;
        ++cens;
        has = false;
      }
      ++ev;
    } else if (RW && !TWO_D && rw) {
      ++segs;
      ++nsurr;
      const u4 q = philox(slot, ev, TAG_RW, 0u, g.seed, g.step);
      bin = k * C + ix;
      const T cdx_ = dx_of(ix);
      float isf, sdf;
      W cidx_;
      if (CELL) {
        const CellRec<P> r = s_cell[ix];
        isf = (float)r.isig; sdf = (float)r.sigdep; cidx_ = r.idxw;
      } else {
        isf = (float)isig[ix]; sdf = (float)sigdep[ix]; cidx_ = idxw[ix];
      }
      const float R0 = (float)rwR0;
      const float spa = 3.0f * R0 * R0 / isf;                  // path length per unit of a
      const float s_e = rw_exit_a(sa.rw_table, u24(q.r[0])) * spa;
      const float s_rem = (float)(cT * (dtT - t) * sT);
      const float s_cut = sdf > 0.f ? __logf((float)E / (float)Emin) / sdf : INFINITY;
      if (s_cut < fminf(s_e, s_rem)) {                        // 1% cut-off first: all deposited
        val = (Acc)(E * cidx_);
        CK[slot] = DEAD;
        has = false;
      } else if (s_e < s_rem) {                                // leaves the interval
        const W Enew = E * (W)__expf(-sdf * s_e);
        val = (Acc)((E - Enew) * cidx_);
        E = Enew;
        t = t + Tm<T>::from((double)s_e) * icsT;
        const bool right = u24(q.r[1]) < 0.5f;
        const float m = surr_exit_mu(u24(q.r[2]));
        mu = Tm<T>::from(right ? (double)m : -(double)m);
        const bool face = right ? (rwR0 == cdx_ - x) : (rwR0 == x);
        if (!face) {
          x = right ? x + rwR0 : x - rwR0;                     // the interior end of the interval
        } else {                                               // the cell face: crossing as :130-170
          bool dead = false;
          x = right ? cdx_ : zero;
          if (mu > zero) {
            if (ix == C - 1) {
              if (g.bc[1] == IMC_REFLECT) mu = -mu; else dead = true;
            }
            if (!dead) { ix += 1; x = zero; }
          }
          if (!dead && mu < zero) {
            if (ix == 0) {
              if (g.bc[0] == IMC_REFLECT) mu = -mu; else dead = true;
            } else {
              ix -= 1; x = dx_of(ix);
            }
          }
          if (dead) {
            lbin = NB - 1;
            lval = (Acc)(E / (W)g.scales[k]);
            CK[slot] = DEAD;
            has = false;
          }
        }
      } else {                                                 // census inside the interval
        const W Enew = E * (W)__expf(-sdf * s_rem);
        val = (Acc)((E - Enew) * cidx_);
        E = Enew;
        const float y = rw_census_y(s_rem / spa, u24(q.r[2]), slot, ev, g.seed, g.step);
        const float xn = fminf(fmaxf((float)x + R0 * y, 0.0f), (float)cdx_);
        float m = 1.0f - 2.0f * u24(q.r[3]);
        if (m == 0.0f) m = 0x1p-24f;
        CK[slot] = (uint32_t)ix | ((uint32_t)k << KSHIFT);
        X[slot] = Tm<T>::from((double)xn);
        MU[slot] = Tm<T>::from((double)m);
        TT[slot] = zero;
        EE[slot] = E;
        ++cens;
        has = false;
      }
      ++ev;
    } else if (SURR && want) {
      ++segs;
      ++nsurr;
      const u4 q = philox(slot, ev, TAG_SURR, 0u, g.seed, g.step);
      bin = k * C + ix;
      if (u24(q.r[0]) < sout[0]) {                           // stays until the budget ends
        if (s_Scut < s_ST) {                                  // 1% cut-off first: all deposited
          val = (Acc)(E * idxw[ix]);
          CK[slot] = DEAD;
          has = false;
        } else {                                              // census inside the cell
          const W Enew = E * Tm<W>::exp_(-(sigdep[ix] * (W)(cT * (dtT - t) * sT)));
          val = (Acc)((E - Enew) * idxw[ix]);
          E = Enew;
          // position of a walker that has not left the cell: diffusion from x0 with absorbing
          // faces, spread sigma^2 = 2 S / (3 sigma_c).  Narrow spread: rejection against the
          // first-order image terms, p(x) ~ G(x-x0) - G(x+x0) - G(x-2L+x0); wide spread: the
          // fundamental mode sin(pi x / L).
          const float Lf = (float)dxs[ix], x0 = (float)x;
          const float sig = sqrtf(2.0f * s_ST / (3.0f * s_sc));
          float xn;
          if (sig < 0.5f * Lf) {
            xn = x0;
            for (uint32_t at = 1; at <= 16; ++at) {
              const u4 r = philox(slot, ev, TAG_SURR, at, g.seed, g.step);
              const float gz = sqrtf(-2.0f * __logf((float)((r.r[0] >> 8) + 1) * 0x1p-24f)) *
                               __cosf(6.2831853f * u24(r.r[1]));
              const float xc = x0 + sig * gz;
              if (xc < 0.f || xc > Lf) continue;
              const float inv = 1.0f / (2.0f * sig * sig);
              const float e0 = (xc - x0) * (xc - x0) * inv;
              const float im = __expf(e0 - (xc + x0) * (xc + x0) * inv) +
                               __expf(e0 - (xc - 2.0f * Lf + x0) * (xc - 2.0f * Lf + x0) * inv);
              xn = xc;
              if (u24(r.r[2]) < 1.0f - im) break;
            }
          } else {
            xn = Lf * 0.318309886f * acosf(1.0f - 2.0f * u24(q.r[2]));
          }
          xn = fminf(fmaxf(xn, 0.0f), Lf);
          float m = 1.0f - 2.0f * u24(q.r[1]);
          if (m == 0.0f) m = 0x1p-24f;
          CK[slot] = (uint32_t)ix | ((uint32_t)k << KSHIFT);
          X[slot] = Tm<T>::from((double)xn);
          MU[slot] = Tm<T>::from((double)m);
          TT[slot] = zero;
          EE[slot] = E;
          ++cens;
          has = false;
        }
      } else {                                                // leaves through a face
        const bool right = u24(q.r[1]) < sout[1];
        // one exit path drawn for both the weight and the clock (see surr_exit_path)
        const float sx = fminf(surr_exit_path(sout[2] * s_SB, s_sd, sout[3],
                                              philox(slot, ev, TAG_SURR, 33u, g.seed, g.step)), s_SB);
        const W Enew = E * (W)__expf(-s_sd * sx);
        val = (Acc)((E - Enew) * idxw[ix]);
        E = Enew;
        t = t + Tm<T>::from((double)sx) * icsT;
        const float m = surr_exit_mu(u24(q.r[2]));
        bool dead = false;
        if (E <= Emin) {                                      // below the cut-off: deposit here
          val += (Acc)(E * idxw[ix]);
          CK[slot] = DEAD;
          has = false;
        } else {
          mu = Tm<T>::from(right ? (double)m : -(double)m);
          x = right ? dx_of(ix) : zero;
          if (mu > zero) {                                    // face crossing as :130-170
            if (ix == C - 1) {
              if (g.bc[1] == IMC_REFLECT) mu = -mu; else dead = true;
            }
            if (!dead) { ix += 1; x = zero; }
          }
          if (!dead && mu < zero) {
            if (ix == 0) {
              if (g.bc[0] == IMC_REFLECT) mu = -mu; else dead = true;
            } else {
              ix -= 1; x = dx_of(ix);
            }
          }
          if (dead) {
            lbin = NB - 1;
            lval = (Acc)(E / (W)g.scales[k]);
            CK[slot] = DEAD;
            has = false;
          }
        }
      }
      ++ev;
    } else if (has) {
      ++segs;
      const u4 q = philox(slot, ev, TAG_EVENT, 0u, g.seed, g.step);
      const int cell = TWO_D ? ix + nx * iy : ix;
      T cdx, cis, cdy = zero;
      W csd, cidx, cidy = (W)1;
      if (CELL) {
        const CellRec<P> r = s_cell[ix];
        cdx = r.dxs; cis = r.isig; csd = r.sigdep; cidx = r.idxw;
      } else {
        if (cell != ccell) {
          ccell = cell;
          c_dx = dxs[ix]; c_is = isig[cell]; c_sd = sigdep[cell]; c_idx = idxw[ix];
          if (TWO_D) { c_dy = dys[iy]; c_idy = idyw[iy]; }
        }
        cdx = c_dx; cis = c_is; csd = c_sd; cidx = c_idx; cdy = c_dy; cidy = c_idy;
      }
      T db, dbx = zero, dby = zero;
      if (!TWO_D) {
        // :77-83 with one division for both signs of mu (same values as the two branches)
        const T qb = ((mu > zero) ? (cdx - x) : x) / mu;
        db = (mu > zero) ? qb : Tm<T>::abs_(qb);
      } else {
        dbx = Tm<T>::abs_(((ox > zero) ? (cdx - x) : x) / ox);                    // :538-542
        dby = Tm<T>::abs_(((oy > zero) ? (cdy - y) : y) / oy);                    // :544-548
        if (Tm<T>::nan_(dbx)) db = dby;                                       // :551-557
        else if (Tm<T>::nan_(dby)) db = dbx;
        else db = jmin(dbx, dby);
      }
      const T dc = Tm<T>::expvar(q) * cis;                                    // :87 / :561
      const T dcen = cT * (dtT - t) * sT;                                     // :89 / :569
      const T d = jmin(jmin(db, dc), dcen);                                   // :92 / :571
      const W tau = csd * (W)d;
      const W Enew = E * Tm<W>::exp_(-tau);                                   // :95 / :580
      const W Evol = TWO_D ? (E * cidx) * cidy : E * cidx;
      bin = k * C + cell;
      const bool capped = ev >= EV_MAX;
      ntrunc += capped ? 1u : 0u;
      if (Enew <= Emin || capped) {                                           // :97-106 / :586-595
        val = (Acc)Evol;
        CK[slot] = DEAD;
        has = false;
      } else {
        val = (Acc)(-Evol * Tm<W>::expm1_(-tau));                              // :120 / :607
        E = Enew;
        t = t + d * icsT;
        bool dead = false, face = false;
        if (!TWO_D) {
          x = x + mu * d;                                                     // :124
          if (d == db) {                                                      // :130-170
            if (mu > zero) {
              if (ix == C - 1) {
                if (g.bc[1] == IMC_REFLECT) mu = -mu; else dead = true;
              }
              if (!dead) { ix += 1; x = zero; }
            }
            if (!dead && mu < zero) {
              if (ix == 0) {
                if (g.bc[0] == IMC_REFLECT) mu = -mu; else dead = true;
              } else {
                ix -= 1; x = dx_of(ix);
              }
            }
          }
        } else {
          x = x + d * ox;                                                     // :615-616
          y = y + d * oy;
          if (d == dbx || d == dby) {                                         // :621-704
            face = true;
            bool turned = false;
            if (dbx < dby) {
              if (ox > zero) {
                if (ix == nx - 1) {
                  if (g.bc[1] == IMC_REFLECT) { mu = Tm<T>::atan2_(oy, -ox); turned = true; } else dead = true;
                } else { ix += 1; x = zero; }
              } else {
                if (ix == 0) {
                  if (g.bc[0] == IMC_REFLECT) { mu = Tm<T>::atan2_(oy, -ox); turned = true; } else dead = true;
                } else { ix -= 1; x = dxs[ix]; }
              }
            } else {
              if (oy > zero) {
                if (iy == ny - 1) {
                  if (g.bc[2] == IMC_REFLECT) { mu = Tm<T>::atan2_(-oy, ox); turned = true; } else dead = true;
                } else { iy += 1; y = zero; }
              } else {
                if (iy == 0) {
                  if (g.bc[3] == IMC_REFLECT) { mu = Tm<T>::atan2_(-oy, ox); turned = true; } else dead = true;
                } else { iy -= 1; y = dys[iy]; }
              }
            }
            if (turned) { ox = Tm<T>::cos_(mu); oy = Tm<T>::sin_(mu); if (RW) mu_stale = false; }
          }
        }
        if (dead) {
          lbin = NB - 1;
          lval = (Acc)(E / (W)g.scales[k]);
          CK[slot] = DEAD;
          has = false;
        } else if (!face) {
          if (d == dc) {                                                      // :174-183 / :706-710
            if (!TWO_D) {
              T m = one - two * Tm<T>::uscat(q, true);
              for (uint32_t at = 1; m == zero; ++at)
                m = one - two * Tm<T>::uscat(philox(slot, ev, TAG_EVENT, at, g.seed, g.step), false);
              mu = m;
            } else {
              mu = Tm<T>::from(TWO_PI_D * Tm<T>::d(Tm<T>::uscat(q, true)));
              ox = Tm<T>::cos_(mu); oy = Tm<T>::sin_(mu);
              if (RW) mu_stale = false;
            }
          }
          if (d == dcen) {                                                    // :185-193 / :712-717
            const int cl = TWO_D ? ix + nx * iy : ix;
            CK[slot] = (uint32_t)cl | ((uint32_t)k << KSHIFT);
            X[slot] = x;
            if (TWO_D) Y[slot] = y;
            MU[slot] = (RW && TWO_D && mu_stale) ? Tm<T>::atan2_(oy, ox) : mu;
            TT[slot] = zero;
            EE[slot] = E;
            ++cens;
            has = false;
          }
        }
      }
      ++ev;
    }
    if (SURR && s_ST < 0.f) has = true;     // a parked lane keeps its particle
    int fbin = -1;
    Acc fval = (Acc)0;
    if (bin >= 0) {
      if (bin == abin) {
        aval += (double)val;
      } else {
        fbin = abin; fval = (Acc)aval;
        abin = bin; aval = (double)val;
      }
    }
    tal.add(fbin, fval);
    tal.add(lbin, lval);
  }
  tal.add(abin, (Acc)aval);

  segs = wave_sum(segs);
  cens = wave_sum(cens);
  if (SURR || RW) nsurr = wave_sum(nsurr);
  ntrunc = (uint32_t)wave_sum((unsigned long long)ntrunc);
  if (lane == 0) {
    atomicAdd(&sc->segments, segs);
    atomicAdd(&sc->census_out, cens);
    if (ntrunc) atomicAdd(&sc->truncated, ntrunc);
    if (SURR || RW) atomicAdd(&sc->surr_events, nsurr);
  }
  if (TM == TM_HASH) {                    // claimed slots go to the block's global copy once
    __syncthreads();
    const int slots = 1 << g.hash_bits;
    const int* tags = (const int*)(lds + slots);
    for (int j = threadIdx.x; j < slots; j += blockDim.x)
      if (tags[j] >= 0) atomicAdd(&tal.gbins[tags[j]], (double)lds[j]);
  }
  tally_flush<Acc, TM>(lds, NB, (Acc*)partial_v);
}

// log2 of the hashed tally's slots: cfg.reserved[2] when set, else 12.
int hash_bits(const Ctx& c) {
  const int r = c.cfg.reserved[2];
  return r >= HASH_BITS_MIN && r <= HASH_BITS_MAX ? r : 12;
}

// Tally mode: cfg.reserved[0] (0 auto, 1 WAVEROW, 2 BLOCKROW, 3 GLOBAL, 4 HASH).  Auto keeps
// a block row in LDS when it fits in 64 KiB and otherwise uses the hashed LDS table, whose
// claimed slots flush once per block and whose colliding bins add into global copies.
static void pick_tally(const Ctx& c, int nbins, int* tally_mode, size_t* lds_bytes) {
  const size_t acc = c.prec == 64 ? 8 : 4;
  const size_t wave_rows = (size_t)NWAVES * nbins * 2 * acc, block_row = (size_t)nbins * acc;
  int req = c.cfg.reserved[0];
  if (req == 1 && wave_rows > 64 * 1024) req = 0;
  if (req == 2 && block_row > 64 * 1024) req = 0;
  const size_t hash_bytes = ((size_t)1 << hash_bits(c)) * (acc + 4);
  if (req == 1) { *tally_mode = TM_WAVEROW; *lds_bytes = wave_rows; }
  else if (req == 3) { *tally_mode = TM_GLOBAL; *lds_bytes = 0; }
  else if (req == 4) { *tally_mode = TM_HASH; *lds_bytes = hash_bytes; }
  else if (block_row <= 64 * 1024) { *tally_mode = TM_BLOCKROW; *lds_bytes = block_row; }
  else { *tally_mode = TM_HASH; *lds_bytes = hash_bytes; }
  *lds_bytes = (*lds_bytes + 15) & ~(size_t)15;
}

static bool use_surrogate(const Ctx& c) { return c.surr_enabled && !c.is2d; }
static bool use_random_walk(const Ctx& c) { return c.rw_enabled && !use_surrogate(c); }

// Bytes of the 1D LDS cell table, 0 when it is not used (2D, or more than CELL_LDS_MAX).
static size_t cell_table_bytes(const Ctx& c) {
  if (c.is2d) return 0;
  const size_t rec = c.prec == 64 ? sizeof(CellRec<64>) : c.prec == 32 ? sizeof(CellRec<32>) : sizeof(CellRec<16>);
  const size_t b = (size_t)c.C * rec;
  return b <= CELL_LDS_MAX ? b : 0;
}

// Dispatch over (precision, tally mode, geometry, acceleration, LDS cell table).  F(P, TM, D, A, L)
// expands to a statement using the kernel k_transport<P, TM, D, A, L>; A is 0 (plain events),
// 1 (learned surrogate) or 2 (analytic random walk).  The accelerated variants exist for the
// block-row and global modes only.
#define IMC_TM_DISPATCH(F, P, D, A, L)                                                   \
  do {                                                                                   \
    if (tm == TM_WAVEROW) F(P, TM_WAVEROW, D, A, L);                                     \
    else if (tm == TM_BLOCKROW) F(P, TM_BLOCKROW, D, A, L);                              \
    else if (tm == TM_HASH) F(P, TM_HASH, D, A, L);                                      \
    else F(P, TM_GLOBAL, D, A, L);                                                       \
  } while (0)
#define IMC_TM_DISPATCH_ACC(F, P, A, L)                                                  \
  do {                                                                                   \
    if (tm == TM_BLOCKROW) F(P, TM_BLOCKROW, false, A, L);                               \
    else F(P, TM_GLOBAL, false, A, L);                                                   \
  } while (0)
#define IMC_TM_DISPATCH_RW2(F, P)                                                        \
  do {                                                                                   \
    if (tm == TM_BLOCKROW) F(P, TM_BLOCKROW, true, 2, false);                            \
    else if (tm == TM_HASH) F(P, TM_HASH, true, 2, false);                               \
    else F(P, TM_GLOBAL, true, 2, false);                                                \
  } while (0)
#define IMC_TRANSPORT_DISPATCH_L(F, P, L)                                                \
  do {                                                                                   \
    if (surr_) IMC_TM_DISPATCH_ACC(F, P, 1, L);                                          \
    else if (rw_) IMC_TM_DISPATCH_ACC(F, P, 2, L);                                       \
    else IMC_TM_DISPATCH(F, P, false, 0, L);                                             \
  } while (0)
#define IMC_TRANSPORT_DISPATCH_P(F, P)                                                   \
  do {                                                                                   \
    if (c.is2d && rw_) IMC_TM_DISPATCH_RW2(F, P);                                        \
    else if (c.is2d) IMC_TM_DISPATCH(F, P, true, 0, false);                              \
    else if (cell_) IMC_TRANSPORT_DISPATCH_L(F, P, true);                                \
    else IMC_TRANSPORT_DISPATCH_L(F, P, false);                                          \
  } while (0)
#define IMC_TRANSPORT_DISPATCH(F)                                                        \
  do {                                                                                   \
    const bool surr_ = use_surrogate(c);                                                 \
    const bool rw_ = use_random_walk(c);                                                 \
    const bool cell_ = cell_table_bytes(c) > 0;                                          \
    if (c.prec == 64) IMC_TRANSPORT_DISPATCH_P(F, 64);                                   \
    else if (c.prec == 32) IMC_TRANSPORT_DISPATCH_P(F, 32);                              \
    else IMC_TRANSPORT_DISPATCH_P(F, 16);                                                \
  } while (0)

int transport_choose(const Ctx& c, int* grid_blocks, int* tally_mode, size_t* lds_bytes) {
  pick_tally(c, c.NB, tally_mode, lds_bytes);
  const bool acc1d = use_surrogate(c) || (use_random_walk(c) && !c.is2d);
  if (acc1d && (*tally_mode == TM_WAVEROW || *tally_mode == TM_HASH)) {
    *tally_mode = TM_GLOBAL;
    *lds_bytes = 0;
  } else if (use_random_walk(c) && c.is2d && *tally_mode == TM_WAVEROW) {   // 2D walk: no wave rows
    *tally_mode = TM_BLOCKROW;
    *lds_bytes = ((size_t)c.NB * (c.prec == 64 ? 8 : 4) + 15) & ~(size_t)15;
  }
  *lds_bytes += cell_table_bytes(c);   // LDS layout: [tally][cell table][surrogate]
  const size_t total = *lds_bytes + (use_surrogate(c) ? SURR_LDS : 0);
  const int tm = *tally_mode;
  int per_cu = 0;
  hipError_t err = hipSuccess;
#define OCC(P, TM_, D, S, L) \
  err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_transport<P, TM_, D, S, L>, TPB, total)
  IMC_TRANSPORT_DISPATCH(OCC);
#undef OCC
  if (err != hipSuccess || per_cu <= 0) per_cu = 1;
  // 5 blocks (20 waves) per CU saturate the SIMDs on gfx950 while leaving each lane more
  // histories, which shortens the launch's tail; reserved[1] overrides the bound.
  per_cu = std::min(per_cu, c.cfg.reserved[1] > 0 ? (int)c.cfg.reserved[1] : 5);
  *grid_blocks = per_cu * c.num_cu;
  return 0;
}

hipError_t launch_transport(Ctx& c, int grid, int tm, size_t lds) {
  Geo g = make_geo(c);
  const size_t cb = cell_table_bytes(c);
  g.cell_off = cb ? (int)(lds - cb) : -1;
  SoA s = c.buf[c.cur];
  SurrArgs sa;
  sa.wbf = c.d_surr_wbf;
  sa.bias = c.d_surr_bias;
  sa.box = c.surr_box;
  sa.map = surr_map(c.surr_box);
  sa.tally_bytes = (int)lds;
  sa.cells = (const float4*)c.d_surr_cells;
  sa.rw_table = c.d_rw_table;
  sa.rw_min = (float)c.rw_min;
  if (use_surrogate(c)) {
    const unsigned gcell = (unsigned)((c.C + 255) / 256);
#define SC(P) k_surr_cells<P><<<gcell, 256, 0, c.stream>>>(c.C, c.surr_box, (const Prec<P>::T*)c.d_dxs,            \
                                                           (const Prec<P>::T*)c.d_sigcol, (const Prec<P>::W*)c.d_sigdep, \
                                                           (float4*)c.d_surr_cells)
    if (c.prec == 64) SC(64);
    else if (c.prec == 32) SC(32);
    else SC(16);
#undef SC
  }
  const size_t total = lds + (use_surrogate(c) ? SURR_LDS : 0);
#define TR(P, TM_, D, S, L) \
  k_transport<P, TM_, D, S, L><<<grid, TPB, total, c.stream>>>(s, g, c.d_sc, c.d_partial, c.d_gtally, sa)
  IMC_TRANSPORT_DISPATCH(TR);
#undef TR
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Clean (imc_clean.jl:6-19) fused with the census tally (imc_tally.jl:157-190).
// The launch covers n slots, the population or an upper bound of it; the kernels stop at the
// device's population (n_total for the count, n_cleaned -- the same value, kept by the scan,
// which replaces n_total with the survivors -- for the scatter).
//   k_clean_count  : alive count per tile of CLEAN_TILE slots
//   k_clean_scan   : exclusive scan of tile counts (single block); n_alive
//   k_clean_scatter: stable scatter of the live particles to the other buffer while
//                    tallying E/(dx*scale) (1D) or E/(dx*dy*scale) (2D) per cell
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_clean_count(const uint32_t* __restrict__ ck, int64_t n_bound,
                                                     const Scalars* __restrict__ sc, int64_t* __restrict__ tile_cnt) {
  __shared__ int wsum[NWAVES];
  const int64_t n = min(n_bound, sc->n_total);
  const int64_t base = (int64_t)blockIdx.x * CLEAN_TILE;
  int cnt = 0;
#pragma unroll
  for (int j = 0; j < CLEAN_ITEMS; ++j) {
    const int64_t i = base + j * TPB + threadIdx.x;
    if (i < n && ck[i] != DEAD) ++cnt;
  }
  cnt = wave_sum(cnt);
  if (lane_id() == 0) wsum[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < NWAVES; ++w) t += wsum[w];
    tile_cnt[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(1024) void k_clean_scan(int64_t* __restrict__ tiles, int64_t ntiles,
                                                     Scalars* sc) {
  __shared__ int64_t wtot[16];
  __shared__ int64_t running;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) running = 0;
  __syncthreads();
  for (int64_t base = 0; base < ntiles; base += 1024) {
    const int64_t i = base + tid;
    const int64_t v = i < ntiles ? tiles[i] : 0;
    int64_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int64_t o = __shfl_up(incl, d, 64);
      if (lane >= d) incl += o;
    }
    if (lane == 63) wtot[w] = incl;
    __syncthreads();
    int64_t wpre = 0;
    for (int j = 0; j < w; ++j) wpre += wtot[j];
    const int64_t r0 = running;
    if (i < ntiles) tiles[i] = r0 + wpre + incl - v;
    __syncthreads();
    if (tid == 1023) running = r0 + wpre + incl;
    __syncthreads();
  }
  if (tid == 0) {
    sc->n_cleaned = sc->n_total;
    sc->n_alive = running;
    sc->n_census = running;
    sc->n_births = 0;
    sc->n_total = running;
  }
}

template <int P, int TM, bool TWO_D>
__global__ __launch_bounds__(TPB) void k_clean_scatter(SoA src, SoA dst, int64_t n_bound, int64_t ntiles,
                                                       const int64_t* __restrict__ tile_off, Geo g,
                                                       void* partial_v, double* gbins, const Scalars* __restrict__ sc) {
  const int64_t n = min(n_bound, sc->n_cleaned);   // k_clean_scan has already set n_total to the survivors
  using T = typename Prec<P>::T;
  using W = typename Prec<P>::W;
  using Acc = typename Prec<P>::Acc;
  extern __shared__ __align__(16) unsigned char smem[];
  Acc* lds = (Acc*)smem;
  __shared__ int wcnt[NWAVES];
  const int NB = g.C;
  tally_init<Acc, TM>(lds, NB);
  Tally<Acc, TM> tal{lds, gbins, NB};
  const W* __restrict__ dxw = (const W*)g.dxw;
  const W* __restrict__ dyw = (const W*)g.dyw;
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    int64_t out = tile_off[tile];
    const int64_t base = tile * CLEAN_TILE;
    for (int j = 0; j < CLEAN_ITEMS; ++j) {
      const int64_t i = base + j * TPB + threadIdx.x;
      const uint32_t ck = i < n ? src.cellk[i] : DEAD;
      const bool alive = ck != DEAD;
      const uint64_t bal = __ballot(alive);
      if (lane == 0) wcnt[w] = __popcll(bal);
      __syncthreads();
      int pre = 0, tot = 0;
#pragma unroll
      for (int q = 0; q < NWAVES; ++q) { pre += q < w ? wcnt[q] : 0; tot += wcnt[q]; }
      int bin = -1;
      Acc val = (Acc)0;
      if (alive) {
        const int64_t o = out + pre + __popcll(bal & lt_mask);
        const W E = ((const W*)src.E)[i];
        dst.cellk[o] = ck;
        ((T*)dst.x)[o] = ((const T*)src.x)[i];
        if (TWO_D) ((T*)dst.y)[o] = ((const T*)src.y)[i];
        ((T*)dst.mu)[o] = ((const T*)src.mu)[i];
        ((T*)dst.t)[o] = ((const T*)src.t)[i];
        ((W*)dst.E)[o] = E;
        ((W*)dst.E0)[o] = ((const W*)src.E0)[i];
        const int cell = (int)(ck & CELLMASK);
        const W sc = (W)g.scales[ck >> KSHIFT];
        bin = cell;
        if (TWO_D) val = (Acc)(E / ((dxw[cell % g.nx] * dyw[cell / g.nx]) * sc));
        else val = (Acc)(E / (dxw[cell] * sc));
      }
      tal.add(bin, val);
      out += tot;
      __syncthreads();
    }
  }
  tally_flush<Acc, TM>(lds, NB, (Acc*)partial_v);
}

int clean_choose(const Ctx& c, int* grid_blocks, int* tally_mode, size_t* lds_bytes) {
  pick_tally(c, c.C, tally_mode, lds_bytes);
  if (*tally_mode == TM_HASH) { *tally_mode = TM_GLOBAL; *lds_bytes = 0; }   // census: global atomics
  // 8 blocks per CU keep enough loads in flight (the clean of the Marshak wave: 73 -> 46 us per
  // step against 2); each block leaves a partial row of C bins, so large meshes keep 2.
  *grid_blocks = (c.C <= 4096 ? 8 : 2) * c.num_cu;
  return 0;
}

hipError_t launch_clean(Ctx& c, int64_t n_total) {
  const int64_t ntiles = (n_total + CLEAN_TILE - 1) / CLEAN_TILE;
  if (ntiles > 0) {
    k_clean_count<<<(unsigned)ntiles, TPB, 0, c.stream>>>(c.buf[c.cur].cellk, n_total, c.d_sc, c.d_tile);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  k_clean_scan<<<1, 1024, 0, c.stream>>>(c.d_tile, ntiles, c.d_sc);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  int grid, tm; size_t lds;
  clean_choose(c, &grid, &tm, &lds);
  // no more blocks than tiles (fewer partial rows to reduce); the partials are sized for the
  // full grid
  grid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, ntiles));
  if (tm == TM_GLOBAL) {
    err = launch_zero(c, c.d_rad, sizeof(double) * c.C);
    if (err != hipSuccess) return err;
  }
  const Geo g = make_geo(c);
  SoA src = c.buf[c.cur], dst = c.buf[c.cur ^ 1];
#define CL(P, TM_, D)                                                                                \
  k_clean_scatter<P, TM_, D><<<grid, TPB, lds, c.stream>>>(src, dst, n_total, ntiles, c.d_tile, g, \
                                                          c.d_partial, c.d_rad, c.d_sc)
#define CLD(P, TM_) do { if (c.is2d) CL(P, TM_, true); else CL(P, TM_, false); } while (0)
#define CLT(P) do { if (tm == TM_WAVEROW) CLD(P, TM_WAVEROW);                                     \
                    else if (tm == TM_BLOCKROW) CLD(P, TM_BLOCKROW);                              \
                    else CLD(P, TM_GLOBAL); } while (0)
  if (c.prec == 64) CLT(64);
  else if (c.prec == 32) CLT(32);
  else CLT(16);
#undef CLT
#undef CLD
#undef CL
  err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = launch_reduce_partials(c, grid, tm, c.C, c.d_rad);
  if (err != hipSuccess) return err;
  c.cur ^= 1;
  return hipSuccess;
}

}  // namespace imc
