// imc_api.cpp -- C ABI of libimc_hip: context, device memory and the entry points that
// replace Sourcing.sourcing / Transport.MC* / Clean.clean / the census tally of
// MixedPrecisionIMC.jl.  Kernels live in imc_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <functional>
#include <new>
#include <random>
#include <vector>

#include "imc_device.h"
#include "imc_internal.h"
#include "imc_mlp.h"

using namespace imc;

namespace {

int fail(Ctx& c, int code, const char* what) {
  c.err = what;
  return code;
}
int hipfail(Ctx& c, hipError_t e, const char* where) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  c.err = buf;
  return IMC_EHIP;
}

#define HIPCHK(expr, where)                          \
  do {                                               \
    hipError_t e_ = (expr);                          \
    if (e_ != hipSuccess) return hipfail(c, e_, where); \
  } while (0)

size_t tsize(int prec) { return prec == 64 ? 8 : (prec == 32 ? 4 : 2); }
size_t wsize(int prec) { return prec == 64 ? 8 : 4; }

template <class P>
hipError_t dalloc(P** p, size_t bytes) {
  return hipMalloc((void**)p, std::max<size_t>(bytes, 8));
}

void free_soa(SoA& s) {
  void* ps[] = {s.cellk, s.x, s.y, s.mu, s.t, s.E, s.E0};
  for (void* p : ps) (void)hipFree(p);
  s = SoA{};
}

hipError_t alloc_soa(Ctx& c, SoA& s, int64_t cap) {
  const size_t T = tsize(c.prec), W = wsize(c.prec), n = (size_t)std::max<int64_t>(cap, 1);
  hipError_t e;
  if ((e = dalloc(&s.cellk, n * 4)) != hipSuccess) return e;
  if ((e = dalloc(&s.x, n * T)) != hipSuccess) return e;
  if (c.is2d && (e = dalloc(&s.y, n * T)) != hipSuccess) return e;
  if ((e = dalloc(&s.mu, n * T)) != hipSuccess) return e;
  if ((e = dalloc(&s.t, n * T)) != hipSuccess) return e;
  if ((e = dalloc(&s.E, n * W)) != hipSuccess) return e;
  if ((e = dalloc(&s.E0, n * W)) != hipSuccess) return e;
  s.cap = cap;
  return hipSuccess;
}

hipError_t copy_soa(Ctx& c, SoA& dst, const SoA& src, int64_t n) {
  if (n <= 0) return hipSuccess;
  const size_t T = tsize(c.prec), W = wsize(c.prec);
  hipError_t e;
  if ((e = hipMemcpyAsync(dst.cellk, src.cellk, n * 4, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(dst.x, src.x, n * T, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return e;
  if (c.is2d && (e = hipMemcpyAsync(dst.y, src.y, n * T, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(dst.mu, src.mu, n * T, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(dst.t, src.t, n * T, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(dst.E, src.E, n * W, hipMemcpyDeviceToDevice, c.stream)) != hipSuccess) return e;
  return hipMemcpyAsync(dst.E0, src.E0, n * W, hipMemcpyDeviceToDevice, c.stream);
}

// Grows both particle buffers to hold at least `need` particles, keeping [0, keep) of
// the current buffer.
int grow(Ctx& c, int64_t need, int64_t keep) {
  if (need <= c.buf[c.cur].cap) return IMC_OK;
  if (need > (int64_t)0xFFFFFFF0ll) return fail(c, IMC_EINVAL, "particle population exceeds 2^32 slots");
  const int64_t cap = std::max<int64_t>(need + need / 4, 2 * c.buf[c.cur].cap);
  SoA fresh;
  hipError_t e = alloc_soa(c, fresh, cap);
  if (e != hipSuccess) { free_soa(fresh); return fail(c, IMC_ENOMEM, "particle buffer allocation failed"); }
  HIPCHK(copy_soa(c, fresh, c.buf[c.cur], keep), "grow copy");
  HIPCHK(hipStreamSynchronize(c.stream), "grow sync");
  free_soa(c.buf[c.cur]);
  c.buf[c.cur] = fresh;
  free_soa(c.buf[c.cur ^ 1]);
  e = alloc_soa(c, c.buf[c.cur ^ 1], cap);
  if (e != hipSuccess) return fail(c, IMC_ENOMEM, "particle buffer allocation failed");
  (void)hipFree(c.d_tile);
  c.tile_len = cap / 2048 + 2;
  e = dalloc(&c.d_tile, sizeof(int64_t) * c.tile_len);
  if (e != hipSuccess) return fail(c, IMC_ENOMEM, "tile buffer allocation failed");
  return IMC_OK;
}

int fetch_scalars(Ctx& c) {
  HIPCHK(hipMemcpyAsync(&c.h_sc, c.d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, c.stream), "scalars D2H");
  HIPCHK(hipStreamSynchronize(c.stream), "scalars sync");
  return IMC_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  (void)hipEventSynchronize(b);
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

}  // namespace

namespace imc {

int state_bytes(int prec, bool is2d) {
  const int T = (int)tsize(prec), W = (int)wsize(prec);
  return 4 + (is2d ? 4 : 3) * T + 2 * W;
}

// Bytes of per-block tally partials: (sum, compensation) rows per block for the tally modes
// that flush through them (per-wave rows, block row); the global and hashed modes add into
// global bins and need none.
size_t partials_needed(const Ctx& c) {
  int gt, tmt, gc, tmc; size_t lt, lc;
  transport_choose(c, &gt, &tmt, &lt);
  clean_choose(c, &gc, &tmc, &lc);
  const size_t acc = c.prec == 64 ? 8 : 4;
  const auto rows = [](int tm) { return tm == TM_WAVEROW || tm == TM_BLOCKROW; };
  size_t need = 256;
  if (rows(tmt)) need = std::max(need, (size_t)gt * c.NB * 2 * acc);
  if (rows(tmc)) need = std::max(need, (size_t)gc * c.C * 2 * acc);
  return need;
}

// Per-tile bucket counts of the birth order (hl entries, with a quarter's headroom), followed in
// the same allocation by the tile sums of their scan.
int ensure_bhist(Ctx& c, int64_t hl) {
  if (hl <= c.bhist_len) return IMC_OK;
  (void)hipFree(c.d_bhist);
  c.d_bhist = c.d_bhist_tiles = nullptr;
  c.bhist_len = 0;
  const int64_t len = hl + hl / 4;
  if (dalloc(&c.d_bhist, sizeof(int64_t) * (size_t)(len + scan_tiles(len) + 1)) != hipSuccess)
    return fail(c, IMC_ENOMEM, "birth-order scratch allocation failed");
  c.d_bhist_tiles = c.d_bhist + len;
  c.bhist_len = len;
  return IMC_OK;
}

Geo make_geo(const Ctx& c) {
  Geo g;
  g.nx = c.nx; g.ny = c.ny; g.C = c.C; g.K = c.K; g.NB = c.NB;
  for (int i = 0; i < 4; ++i) g.bc[i] = c.cfg.bc[i];
  g.seed = c.cfg.seed; g.step = (uint32_t)c.step;
  g.dt = c.dt; g.c = c.cfg.phys_c; g.s = c.cfg.distancescale;
  g.dxs = c.d_dxs; g.dys = c.d_dys; g.dxw = c.d_dxw; g.dyw = c.d_dyw;
  g.sigcol = c.d_sigcol; g.sigdep = c.d_sigdep;
  g.isig = c.d_isig; g.idxw = c.d_idxw; g.idyw = c.d_idyw;
  g.ics = 1.0 / (c.cfg.phys_c * c.cfg.distancescale);
  g.cell_off = -1;
  g.gcopies = gtally_copies(c.NB);
  g.hash_bits = hash_bits(c);
  for (int k = 0; k < 16; ++k) g.scales[k] = c.cfg.energyscales[k];
  return g;
}

}  // namespace imc

extern "C" {

int imc_create(imc_ctx** out, const imc_config* cfg) {
  if (!out || !cfg) return IMC_EINVAL;
  *out = nullptr;
  imc_ctx* h = new (std::nothrow) imc_ctx();
  if (!h) return IMC_ENOMEM;
  Ctx& c = h->c;
  c.cfg = *cfg;
  *out = h;
  if (cfg->geometry != 1 && cfg->geometry != 2) return fail(c, IMC_EINVAL, "geometry must be 1 or 2");
  if (cfg->precision != 64 && cfg->precision != 32 && cfg->precision != 16)
    return fail(c, IMC_EINVAL, "precision must be 64, 32 or 16");
  if (cfg->nx < 1 || (cfg->geometry == 2 && cfg->ny < 1)) return fail(c, IMC_EINVAL, "nx/ny must be >= 1");
  if (cfg->n_scales < 1 || cfg->n_scales > 16) return fail(c, IMC_EINVAL, "n_scales must be in 1..16");
  for (int i = 0; i < 4; ++i)
    if (cfg->bc[i] != IMC_VACUUM && cfg->bc[i] != IMC_REFLECT) return fail(c, IMC_EINVAL, "bc must be VACUUM or REFLECT");
  c.prec = cfg->precision;
  c.is2d = cfg->geometry == 2;
  c.nx = cfg->nx;
  c.ny = c.is2d ? cfg->ny : 1;
  c.C = c.nx * c.ny;
  if ((int64_t)c.nx * c.ny >= (1ll << KSHIFT)) return fail(c, IMC_EINVAL, "too many cells (limit 2^28)");
  c.K = cfg->n_scales;
  c.NB = c.C * c.K + 1;
  c.NE = c.is2d ? (int64_t)2 * c.nx + 2 * c.ny + 2 * (int64_t)c.C : 2 + 2 * (int64_t)c.C;
  c.device = cfg->device;
  HIPCHK(hipSetDevice(c.device), "hipSetDevice");
  HIPCHK(hipDeviceGetAttribute(&c.num_cu, hipDeviceAttributeMultiprocessorCount, c.device), "CU count");
  HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "stream");
  HIPCHK(hipEventCreate(&c.ev0), "event"); HIPCHK(hipEventCreate(&c.ev1), "event");
  HIPCHK(hipEventCreate(&c.evk0), "event"); HIPCHK(hipEventCreate(&c.evk1), "event");
  for (hipEvent_t& e : c.evs) HIPCHK(hipEventCreate(&e), "event");
  const size_t T = tsize(c.prec), W = wsize(c.prec), C = (size_t)c.C;
  const size_t nx = (size_t)c.nx, ny = (size_t)c.ny;
  bool ok = true;
  ok &= dalloc(&c.d_dx, 8 * nx) == hipSuccess;
  ok &= dalloc(&c.d_dy, 8 * ny) == hipSuccess;
  ok &= dalloc(&c.d_dxs, T * nx) == hipSuccess;
  ok &= dalloc(&c.d_dys, T * ny) == hipSuccess;
  ok &= dalloc(&c.d_dxw, W * nx) == hipSuccess;
  ok &= dalloc(&c.d_dyw, W * ny) == hipSuccess;
  // step fields, one allocation in upload order: sigma_a, sigma_s, fleck, temp, radsource, t_surf
  c.fields_len = 5 * C + (c.is2d ? 2 * nx + 2 * ny : 2);
  ok &= dalloc(&c.d_fields, 8 * c.fields_len) == hipSuccess;
  ok &= hipHostMalloc((void**)&c.h_fields, 8 * c.fields_len, hipHostMallocDefault) == hipSuccess;
  ok &= hipEventCreateWithFlags(&c.ev_fields, hipEventDisableTiming) == hipSuccess;
  if (ok) {
    c.d_sa = c.d_fields; c.d_ss = c.d_sa + C; c.d_fleck = c.d_ss + C; c.d_temp = c.d_fleck + C;
    c.d_radsrc = c.d_temp + C; c.d_tsurf = c.d_radsrc + C;
  }
  ok &= dalloc(&c.d_sigcol, T * C) == hipSuccess;
  ok &= dalloc(&c.d_sigdep, W * C) == hipSuccess;
  ok &= dalloc(&c.d_isig, T * C) == hipSuccess;
  ok &= dalloc(&c.d_idxw, W * nx) == hipSuccess;
  ok &= dalloc(&c.d_idyw, W * ny) == hipSuccess;
  const size_t NE = (size_t)c.NE;
  ok &= dalloc(&c.d_ent_e, 8 * NE) == hipSuccess;
  ok &= dalloc(&c.d_ent_esc, 8 * NE) == hipSuccess;
  ok &= dalloc(&c.d_ent_term, 8 * NE) == hipSuccess;
  ok &= dalloc(&c.d_ent_k, 4 * NE) == hipSuccess;
  ok &= dalloc(&c.d_ent_cnt, 8 * NE) == hipSuccess;
  ok &= dalloc(&c.d_ent_off, 8 * NE) == hipSuccess;
  ok &= dalloc(&c.d_ent_tiles, 8 * (size_t)(scan_tiles(c.NE) + 1)) == hipSuccess;
  c.red_len = 2 * ((int64_t)NE / 256 + 4);
  ok &= dalloc(&c.d_red, 8 * (size_t)c.red_len) == hipSuccess;
  ok &= dalloc(&c.d_emitted, 8 * C * c.K) == hipSuccess;
  ok &= dalloc(&c.d_dep, 8 * (size_t)c.NB) == hipSuccess;
  ok &= dalloc(&c.d_gtally, 8 * (size_t)c.NB * gtally_copies(c.NB)) == hipSuccess;
  ok &= dalloc(&c.d_rad, 8 * C) == hipSuccess;
  ok &= dalloc(&c.d_sc, sizeof(Scalars)) == hipSuccess;
  if (!ok) return fail(c, IMC_ENOMEM, "device allocation failed");
  c.partial_bytes = partials_needed(c);
  if (dalloc(&c.d_partial, c.partial_bytes) != hipSuccess) return fail(c, IMC_ENOMEM, "tally partial allocation failed");
  const int64_t cap = std::max<int64_t>(cfg->capacity, 4096);
  if (alloc_soa(c, c.buf[0], cap) != hipSuccess || alloc_soa(c, c.buf[1], cap) != hipSuccess)
    return fail(c, IMC_ENOMEM, "particle buffer allocation failed");
  c.tile_len = cap / 2048 + 2;
  if (dalloc(&c.d_tile, sizeof(int64_t) * c.tile_len) != hipSuccess) return fail(c, IMC_ENOMEM, "tile allocation failed");
  std::memset(&c.h_sc, 0, sizeof(Scalars));
  HIPCHK(hipMemcpyAsync(c.d_sc, &c.h_sc, sizeof(Scalars), hipMemcpyHostToDevice, c.stream), "scalars H2D");
  HIPCHK(hipStreamSynchronize(c.stream), "create sync");
  return IMC_OK;
}

int imc_destroy(imc_ctx* h) {
  if (!h) return IMC_OK;
  Ctx& c = h->c;
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  free_soa(c.buf[0]); free_soa(c.buf[1]);
  void* ptrs[] = {c.d_dx, c.d_dy, c.d_dxs, c.d_dys, c.d_dxw, c.d_dyw, c.d_fields, c.d_sigcol, c.d_sigdep, c.d_isig, c.d_idxw, c.d_idyw, c.d_ent_e, c.d_ent_esc, c.d_ent_term,
                  c.d_ent_k, c.d_ent_cnt, c.d_ent_off, c.d_red, c.d_emitted, c.d_dep, c.d_gtally, c.d_rad, c.d_sc,
                  c.d_partial, c.d_tile, c.d_bhist, c.d_ent_tiles, c.d_surr_feat, c.d_surr_targ, c.d_surr_params, c.d_surr_grad,
                  c.d_surr_m, c.d_surr_v, c.d_surr_wbf, c.d_surr_bias, c.d_surr_loss, c.d_surr_events, c.d_surr_cells,
                  c.d_mat, c.d_rw_table};
  for (void* p : ptrs) (void)hipFree(p);
  if (c.h_fields) (void)hipHostFree(c.h_fields);
  if (c.h_out) (void)hipHostFree(c.h_out);
  hipEvent_t evs[] = {c.ev0, c.ev1, c.evk0, c.evk1, c.ev_fields, c.evs[0], c.evs[1], c.evs[2], c.evs[3]};
  for (hipEvent_t e : evs)
    if (e) (void)hipEventDestroy(e);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete h;
  return IMC_OK;
}

const char* imc_last_error(const imc_ctx* h) { return h ? h->c.err.c_str() : "null context"; }

int imc_set_geometry(imc_ctx* h, const double* dx, const double* dy) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  if (!dx || (c.is2d && !dy)) return fail(c, IMC_EINVAL, "dx (and dy in 2D) required");
  for (int i = 0; i < c.nx; ++i)
    if (!(dx[i] > 0.0) || !std::isfinite(dx[i])) return fail(c, IMC_EINVAL, "dx must be positive and finite");
  std::vector<double> ones(c.ny, 1.0);
  const double* ydata = c.is2d ? dy : ones.data();
  for (int j = 0; j < c.ny; ++j)
    if (!(ydata[j] > 0.0) || !std::isfinite(ydata[j])) return fail(c, IMC_EINVAL, "dy must be positive and finite");
  HIPCHK(hipMemcpyAsync(c.d_dx, dx, 8 * c.nx, hipMemcpyHostToDevice, c.stream), "dx H2D");
  HIPCHK(hipMemcpyAsync(c.d_dy, ydata, 8 * c.ny, hipMemcpyHostToDevice, c.stream), "dy H2D");
  HIPCHK(hipStreamSynchronize(c.stream), "geometry sync");
  c.geo_set = true;
  return IMC_OK;
}

int imc_set_step_fields(imc_ctx* h, int64_t step, double dt, const double* sigma_a, const double* sigma_s,
                        const double* fleck, const double* temp, const double* radsource,
                        const double* const* t_surf) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  c.step_counters_clear = false;      // this call leaves the scalar counters in use
  if (!c.geo_set) return fail(c, IMC_EINVAL, "imc_set_geometry must be called first");
  if (!sigma_a || !sigma_s || !fleck || !temp || !radsource || !t_surf)
    return fail(c, IMC_EINVAL, "null step field");
  if (step < 0 || step > 0xFFFFFFFFll) return fail(c, IMC_EINVAL, "step out of range");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, IMC_EINVAL, "dt must be positive and finite");
  for (int i = 0; i < c.C; ++i) {
    if (!std::isfinite(sigma_a[i]) || !std::isfinite(sigma_s[i]) || !std::isfinite(fleck[i]) ||
        !std::isfinite(temp[i]) || !std::isfinite(radsource[i]) || sigma_a[i] < 0.0 || sigma_s[i] < 0.0)
      return fail(c, IMC_ENUMERIC, "non-finite or negative opacity / Fleck factor / temperature");
    if (!(fleck[i] >= 0.0 && fleck[i] <= 1.0)) return fail(c, IMC_ENUMERIC, "Fleck factor outside [0, 1]");
  }
  std::vector<double> ts;
  if (c.is2d) {
    for (int q = 0; q < 4; ++q) {
      const int len = q < 2 ? c.nx : c.ny;
      if (!t_surf[q]) return fail(c, IMC_EINVAL, "null t_surf entry");
      ts.insert(ts.end(), t_surf[q], t_surf[q] + len);
    }
  } else {
    if (!t_surf[0] || !t_surf[1]) return fail(c, IMC_EINVAL, "null t_surf entry");
    ts = {t_surf[0][0], t_surf[1][0]};
  }
  c.step = step;
  c.dt = dt;
  const size_t C = (size_t)c.C;
  // The caller's arrays are copied into the pinned staging block before returning, and one
  // asynchronous copy uploads it; the previous upload from that block is waited for first.
  HIPCHK(hipEventSynchronize(c.ev_fields), "staging wait");
  double* hf = c.h_fields;
  const double* src[5] = {sigma_a, sigma_s, fleck, temp, radsource};
  for (int q = 0; q < 5; ++q) std::memcpy(hf + q * C, src[q], 8 * C);
  std::memcpy(hf + 5 * C, ts.data(), 8 * ts.size());
  HIPCHK(hipMemcpyAsync(c.d_fields, hf, 8 * c.fields_len, hipMemcpyHostToDevice, c.stream), "step fields H2D");
  HIPCHK(hipEventRecord(c.ev_fields, c.stream), "event");
  HIPCHK(launch_zero(c, (char*)c.d_sc + offsetof(Scalars, prep_underflow), sizeof(uint32_t)), "zero prep flag");
  HIPCHK(launch_prep(c), "prep kernel");
  uint32_t underflow = 0;
  if (c.prec == 16) {                   // only binary16 can round a mean free path to 0
    HIPCHK(hipMemcpyAsync(&underflow, (char*)c.d_sc + offsetof(Scalars, prep_underflow), sizeof(uint32_t),
                          hipMemcpyDeviceToHost, c.stream), "prep flag D2H");
    HIPCHK(hipStreamSynchronize(c.stream), "step fields sync");
  }
  if (underflow)
    return fail(c, IMC_ENUMERIC, (std::to_string(underflow) + " cell(s) have a collision mean free path "
                "1/sigma_c that rounds to 0 in the tracking precision; use a wider PRECISION").c_str());
  c.step_set = true;
  return IMC_OK;
}

int imc_source(imc_ctx* h, int64_t n_input, int64_t n_max, double cellmin, double* emittedenergy,
               double* totalenergy, int64_t* n_alive) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  c.step_counters_clear = false;      // this call leaves the scalar counters in use
  if (!c.step_set) return fail(c, IMC_EINVAL, "imc_set_step_fields must be called before imc_source");
  if (n_input < 0 || n_max < 0 || !(cellmin >= 0.0)) return fail(c, IMC_EINVAL, "negative particle budget");
  HIPCHK(hipEventRecord(c.ev0, c.stream), "event");
  HIPCHK(launch_source_entries(c), "source entries kernel");
  HIPCHK(launch_tree_sum(c, c.d_ent_term, c.NE, (double*)((char*)c.d_sc + offsetof(Scalars, E_tot))),
         "energy sum kernel");
  // Births per entry are max(round(share), cellmin) with shares summing to the budget
  // (imc_sourcing.jl:139-156), so budget + NE (cellmin + 1) bounds them (twice the budget
  // with quirk Q1).  When that bound fits the buffers, the births are launched over it
  // without a host round trip; the count kernel still refuses births that do not fit.
  const int64_t n_census = c.h_sc.n_census;
  const double budget = std::ceil(std::max((double)n_input, cellmin));
  const double bound_d = budget * ((c.cfg.quirks & IMC_QUIRK_Q1_RADSRC_NBODY) ? 2.0 : 1.0) +
                         (double)c.NE * (std::ceil(cellmin) + 1.0) + 16.0;
  const int64_t bound = bound_d < 4e18 ? (int64_t)bound_d : INT64_MAX / 4;
  const bool direct = n_census + bound <= c.buf[c.cur].cap && birth_hist_len(bound) <= c.bhist_len;
  HIPCHK(launch_counts_and_scan(c, n_input, n_max, cellmin, direct ? c.buf[c.cur].cap : INT64_MAX), "count kernels");
  int rc;
  if (direct) {
    HIPCHK(launch_births(c, bound), "birth kernel");
  } else {
    rc = fetch_scalars(c);
    if (rc) return rc;
    if (!std::isfinite(c.h_sc.E_tot) || !(c.h_sc.E_tot > 0.0))
      return fail(c, IMC_ENUMERIC, "total source energy is zero or not finite");
    rc = grow(c, c.h_sc.n_census + c.h_sc.n_births, c.h_sc.n_census);
    if (rc) return rc;
    rc = ensure_bhist(c, birth_hist_len(c.h_sc.n_births));
    if (rc) return rc;
    HIPCHK(launch_births(c, c.h_sc.n_births), "birth kernel");
  }
  if (emittedenergy)
    HIPCHK(hipMemcpyAsync(emittedenergy, c.d_emitted, 8 * (size_t)c.C * c.K, hipMemcpyDeviceToHost, c.stream), "emitted D2H");
  HIPCHK(hipMemcpyAsync(&c.h_sc, c.d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, c.stream), "scalars D2H");
  HIPCHK(hipEventRecord(c.ev1, c.stream), "event");
  HIPCHK(hipStreamSynchronize(c.stream), "source sync");
  c.timing.ms_source += elapsed(c.ev0, c.ev1);
  if (!std::isfinite(c.h_sc.E_tot) || !(c.h_sc.E_tot > 0.0))
    return fail(c, IMC_ENUMERIC, "total source energy is zero or not finite");
  if (c.h_sc.birth_overflow) return fail(c, IMC_ENOMEM, "births exceed the particle buffer");
  if (totalenergy) *totalenergy = c.h_sc.E_tot;
  if (n_alive) *n_alive = c.h_sc.n_total;
  return IMC_OK;
}

int imc_transport(imc_ctx* h, double* energydep, double* lostenergy, int64_t* segments) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  c.step_counters_clear = false;      // this call leaves the scalar counters in use
  if (!c.step_set) return fail(c, IMC_EINVAL, "imc_set_step_fields must be called before imc_transport");
  int grid, tm; size_t lds;
  transport_choose(c, &grid, &tm, &lds);
  HIPCHK(hipEventRecord(c.ev0, c.stream), "event");
  HIPCHK(launch_zero(c, (char*)c.d_sc + offsetof(Scalars, segments),
                     offsetof(Scalars, prep_underflow) - offsetof(Scalars, segments)), "zero scalars");
  if (tm == TM_GLOBAL || tm == TM_HASH) HIPCHK(launch_zero(c, c.d_gtally, 8 * (size_t)c.NB * gtally_copies(c.NB)), "zero tallies");
  HIPCHK(hipEventRecord(c.evk0, c.stream), "event");
  HIPCHK(launch_transport(c, grid, tm, lds), "transport kernel");
  HIPCHK(hipEventRecord(c.evk1, c.stream), "event");
  if (tm == TM_GLOBAL || tm == TM_HASH) HIPCHK(launch_sum_copies(c, c.d_gtally, gtally_copies(c.NB), c.NB, c.d_dep), "tally copies");
  else HIPCHK(launch_reduce_partials(c, grid, tm, c.NB, c.d_dep), "tally reduction");
  c.h_dep.resize((size_t)c.NB);                 // C*K bins + lost energy, one copy
  HIPCHK(hipMemcpyAsync(c.h_dep.data(), c.d_dep, 8 * (size_t)c.NB, hipMemcpyDeviceToHost, c.stream), "energydep D2H");
  HIPCHK(hipMemcpyAsync(&c.h_sc, c.d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, c.stream), "scalars D2H");
  HIPCHK(hipEventRecord(c.ev1, c.stream), "event");
  HIPCHK(hipStreamSynchronize(c.stream), "transport sync");
  if (energydep) std::memcpy(energydep, c.h_dep.data(), 8 * (size_t)c.C * c.K);
  const double lost = c.h_dep[(size_t)c.NB - 1];
  c.timing.ms_transport += elapsed(c.ev0, c.ev1);
  c.timing.ms_transport_kernel += elapsed(c.evk0, c.evk1);
  c.timing.n_transport_launches += 1;
  c.timing.histories += c.h_sc.n_total;
  c.timing.census_out += (int64_t)c.h_sc.census_out;
  c.timing.segments += (int64_t)c.h_sc.segments;
  c.timing.surrogate_events += (int64_t)c.h_sc.surr_events;
  c.timing.truncated += (int64_t)c.h_sc.truncated;
  c.timing.transport_tally_mode = tm;
  c.timing.transport_grid = grid;
  if (lostenergy) *lostenergy = lost;
  if (segments) *segments = (int64_t)c.h_sc.segments;
  if (!std::isfinite(lost)) return fail(c, IMC_ENUMERIC, "non-finite lost energy");
  return IMC_OK;
}

int imc_clean(imc_ctx* h, int64_t* n_alive) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  c.step_counters_clear = false;      // this call leaves the scalar counters in use
  if (!c.step_set) return fail(c, IMC_EINVAL, "imc_set_step_fields must be called before imc_clean");
  HIPCHK(hipEventRecord(c.ev0, c.stream), "event");
  HIPCHK(launch_clean(c, c.h_sc.n_total), "clean kernels");
  // the census tally is computed by the clean kernels; bring it back in the same wait
  c.h_rad.resize((size_t)c.C);
  HIPCHK(hipMemcpyAsync(c.h_rad.data(), c.d_rad, 8 * (size_t)c.C, hipMemcpyDeviceToHost, c.stream), "rad D2H");
  c.rad_valid = true;
  int rc = fetch_scalars(c);
  if (rc) return rc;
  HIPCHK(hipEventRecord(c.ev1, c.stream), "event");
  c.timing.ms_clean += elapsed(c.ev0, c.ev1);
  if (n_alive) *n_alive = c.h_sc.n_alive;
  return IMC_OK;
}

int imc_tally_census(imc_ctx* h, double* radenergydens) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  if (!radenergydens) return fail(c, IMC_EINVAL, "radenergydens required");
  if (c.rad_valid) {                              // copied back by imc_clean
    std::memcpy(radenergydens, c.h_rad.data(), 8 * (size_t)c.C);
    return IMC_OK;
  }
  HIPCHK(hipEventRecord(c.ev0, c.stream), "event");
  HIPCHK(hipMemcpyAsync(radenergydens, c.d_rad, 8 * (size_t)c.C, hipMemcpyDeviceToHost, c.stream), "rad D2H");
  HIPCHK(hipEventRecord(c.ev1, c.stream), "event");
  HIPCHK(hipStreamSynchronize(c.stream), "tally sync");
  c.timing.ms_tally += elapsed(c.ev0, c.ev1);
  return IMC_OK;
}

int imc_particle_count(imc_ctx* h, int64_t* n) {
  if (!h || !n) return IMC_EINVAL;
  *n = h->c.h_sc.n_total;
  return IMC_OK;
}

int imc_export_particles(imc_ctx* h, double* rec, int64_t cap, int64_t* nout) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  const int64_t n = c.h_sc.n_total;
  if (nout) *nout = n;
  if (!rec || cap < n) return fail(c, IMC_EINVAL, "record buffer too small");
  if (n == 0) return IMC_OK;
  const size_t T = tsize(c.prec), W = wsize(c.prec);
  const SoA& s = c.buf[c.cur];
  std::vector<uint32_t> ck(n);
  std::vector<unsigned char> x(n * T), y(c.is2d ? n * T : 1), mu(n * T), t(n * T), E(n * W), E0(n * W);
  HIPCHK(hipMemcpyAsync(ck.data(), s.cellk, 4 * n, hipMemcpyDeviceToHost, c.stream), "export");
  HIPCHK(hipMemcpyAsync(x.data(), s.x, T * n, hipMemcpyDeviceToHost, c.stream), "export");
  if (c.is2d) HIPCHK(hipMemcpyAsync(y.data(), s.y, T * n, hipMemcpyDeviceToHost, c.stream), "export");
  HIPCHK(hipMemcpyAsync(mu.data(), s.mu, T * n, hipMemcpyDeviceToHost, c.stream), "export");
  HIPCHK(hipMemcpyAsync(t.data(), s.t, T * n, hipMemcpyDeviceToHost, c.stream), "export");
  HIPCHK(hipMemcpyAsync(E.data(), s.E, W * n, hipMemcpyDeviceToHost, c.stream), "export");
  HIPCHK(hipMemcpyAsync(E0.data(), s.E0, W * n, hipMemcpyDeviceToHost, c.stream), "export");
  HIPCHK(hipStreamSynchronize(c.stream), "export sync");
  auto tv = [&](const std::vector<unsigned char>& v, int64_t i) -> double {
    if (c.prec == 64) return ((const double*)v.data())[i];
    if (c.prec == 32) return (double)((const float*)v.data())[i];
    return (double)(float)((const _Float16*)v.data())[i];
  };
  auto wv = [&](const std::vector<unsigned char>& v, int64_t i) -> double {
    if (c.prec == 64) return ((const double*)v.data())[i];
    return (double)((const float*)v.data())[i];
  };
  for (int64_t i = 0; i < n; ++i) {
    const bool dead = ck[i] == DEAD;
    const int cell = dead ? 0 : (int)(ck[i] & CELLMASK);
    const int k = dead ? 0 : (int)(ck[i] >> KSHIFT);
    const double scale = c.cfg.energyscales[k];
    if (!c.is2d) {
      double* r = rec + 9 * i;
      r[0] = 0.0; r[1] = tv(t, i); r[2] = dead ? 0.0 : cell + 1; r[3] = tv(x, i); r[4] = tv(mu, i);
      r[5] = 1.0; r[6] = wv(E, i); r[7] = dead ? -1.0 : wv(E0, i); r[8] = scale;
    } else {
      double* r = rec + 10 * i;
      r[0] = tv(t, i); r[1] = dead ? 0.0 : cell % c.nx + 1; r[2] = dead ? 0.0 : cell / c.nx + 1;
      r[3] = tv(x, i); r[4] = tv(y, i); r[5] = tv(mu, i); r[6] = 1.0;
      r[7] = dead ? -1.0 : wv(E, i); r[8] = wv(E0, i); r[9] = scale;
    }
  }
  return IMC_OK;
}

int imc_import_particles(imc_ctx* h, const double* rec, int64_t n) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  c.step_counters_clear = false;      // this call leaves the scalar counters in use
  if (n < 0 || (n > 0 && !rec)) return fail(c, IMC_EINVAL, "bad record buffer");
  int rc = grow(c, n, 0);
  if (rc) return rc;
  const size_t T = tsize(c.prec), W = wsize(c.prec);
  std::vector<uint32_t> ck(n);
  std::vector<unsigned char> x(n * T + 1), y(n * T + 1), mu(n * T + 1), t(n * T + 1), E(n * W + 1), E0(n * W + 1);
  auto put_t = [&](std::vector<unsigned char>& v, int64_t i, double val) {
    if (c.prec == 64) ((double*)v.data())[i] = val;
    else if (c.prec == 32) ((float*)v.data())[i] = (float)val;
    else ((_Float16*)v.data())[i] = (_Float16)(float)val;
  };
  auto put_w = [&](std::vector<unsigned char>& v, int64_t i, double val) {
    if (c.prec == 64) ((double*)v.data())[i] = val;
    else ((float*)v.data())[i] = (float)val;
  };
  for (int64_t i = 0; i < n; ++i) {
    const double* r = rec + (c.is2d ? 10 : 9) * i;
    const double scale = c.is2d ? r[9] : r[8];
    int k = 0;
    for (int q = 0; q < c.K; ++q)
      if (c.cfg.energyscales[q] == scale) { k = q; break; }
    if (!c.is2d) {
      const int cell = (int)r[2] - 1;
      if (r[7] != -1.0 && (cell < 0 || cell >= c.C)) return fail(c, IMC_EINVAL, "record cell out of range");
      ck[i] = r[7] == -1.0 ? DEAD : ((uint32_t)cell | ((uint32_t)k << KSHIFT));
      put_t(t, i, r[1]); put_t(x, i, r[3]); put_t(mu, i, r[4]); put_w(E, i, r[6]); put_w(E0, i, r[7]);
    } else {
      const int ix = (int)r[1] - 1, iy = (int)r[2] - 1;
      if (r[7] != -1.0 && (ix < 0 || ix >= c.nx || iy < 0 || iy >= c.ny))
        return fail(c, IMC_EINVAL, "record cell out of range");
      ck[i] = r[7] == -1.0 ? DEAD : ((uint32_t)(ix + c.nx * iy) | ((uint32_t)k << KSHIFT));
      put_t(t, i, r[0]); put_t(x, i, r[3]); put_t(y, i, r[4]); put_t(mu, i, r[5]);
      put_w(E, i, r[7]); put_w(E0, i, r[8]);
    }
  }
  const SoA& s = c.buf[c.cur];
  if (n > 0) {
    HIPCHK(hipMemcpyAsync(s.cellk, ck.data(), 4 * n, hipMemcpyHostToDevice, c.stream), "import");
    HIPCHK(hipMemcpyAsync(s.x, x.data(), T * n, hipMemcpyHostToDevice, c.stream), "import");
    if (c.is2d) HIPCHK(hipMemcpyAsync(s.y, y.data(), T * n, hipMemcpyHostToDevice, c.stream), "import");
    HIPCHK(hipMemcpyAsync(s.mu, mu.data(), T * n, hipMemcpyHostToDevice, c.stream), "import");
    HIPCHK(hipMemcpyAsync(s.t, t.data(), T * n, hipMemcpyHostToDevice, c.stream), "import");
    HIPCHK(hipMemcpyAsync(s.E, E.data(), W * n, hipMemcpyHostToDevice, c.stream), "import");
    HIPCHK(hipMemcpyAsync(s.E0, E0.data(), W * n, hipMemcpyHostToDevice, c.stream), "import");
  }
  c.h_sc.n_census = n; c.h_sc.n_births = 0; c.h_sc.n_total = n; c.h_sc.n_alive = n;
  HIPCHK(hipMemcpyAsync(c.d_sc, &c.h_sc, sizeof(Scalars), hipMemcpyHostToDevice, c.stream), "import scalars");
  HIPCHK(hipStreamSynchronize(c.stream), "import sync");
  return IMC_OK;
}

int imc_get_timing(imc_ctx* h, imc_timing* out) {
  if (!h || !out) return IMC_EINVAL;
  *out = h->c.timing;
  return IMC_OK;
}

int imc_reset_timing(imc_ctx* h) {
  if (!h) return IMC_EINVAL;
  std::memset(&h->c.timing, 0, sizeof(imc_timing));
  return IMC_OK;
}

int imc_synchronize(imc_ctx* h) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  HIPCHK(hipStreamSynchronize(c.stream), "synchronize");
  return IMC_OK;
}

int imc_state_bytes(const imc_ctx* h, int32_t* bytes) {
  if (!h || !bytes) return IMC_EINVAL;
  *bytes = state_bytes(h->c.prec, h->c.is2d);
  return IMC_OK;
}

// ---------------------------------------------------------------------------------
// device-resident step
// ---------------------------------------------------------------------------------
int imc_set_material(imc_ctx* h, const imc_material* m) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  if (!m) return fail(c, IMC_EINVAL, "material required");
  if (!c.geo_set) return fail(c, IMC_EINVAL, "imc_set_geometry must be called first");
  if (!m->sigma_a_coef || !m->sigma_a_power || !m->sigma_s_coef || !m->sigma_s_power || !m->temp ||
      !m->radsource || (!m->linearized && !m->bee))
    return fail(c, IMC_EINVAL, "null material array");
  if (!std::isfinite(m->alpha)) return fail(c, IMC_EINVAL, "alpha must be finite");
  const size_t C = (size_t)c.C;
  for (size_t i = 0; i < C; ++i) {
    const double v[6] = {m->sigma_a_coef[i], m->sigma_a_power[i], m->sigma_s_coef[i], m->sigma_s_power[i],
                         m->temp[i], m->radsource[i]};
    for (double x : v)
      if (!std::isfinite(x)) return fail(c, IMC_ENUMERIC, "non-finite material coefficient or temperature");
    if (!m->linearized && !(std::isfinite(m->bee[i]) && m->bee[i] != 0.0))
      return fail(c, IMC_ENUMERIC, "bee must be finite and non-zero");
  }
  std::vector<double> ts;
  if (c.is2d) {
    for (int q = 0; q < 4; ++q) {
      if (!m->t_surf[q]) return fail(c, IMC_EINVAL, "null t_surf entry");
      ts.insert(ts.end(), m->t_surf[q], m->t_surf[q] + (q < 2 ? c.nx : c.ny));
    }
  } else {
    if (!m->t_surf[0] || !m->t_surf[1]) return fail(c, IMC_EINVAL, "null t_surf entry");
    ts = {m->t_surf[0][0], m->t_surf[1][0]};
  }
  if (!c.d_mat) {
    if (dalloc(&c.d_mat, 8 * MAT_N * C + sizeof(Scalars) + step_sums_scratch_bytes(c)) != hipSuccess)
      return fail(c, IMC_ENOMEM, "material allocation failed");
    if (hipHostMalloc((void**)&c.h_out, 8 * MAT_OUT * C + sizeof(Scalars), hipHostMallocDefault) != hipSuccess) {
      c.h_out = nullptr;
      return fail(c, IMC_ENOMEM, "pinned result allocation failed");
    }
  }
  std::vector<double> st(MAT_N * C, 0.0);
  const double* src[6] = {m->sigma_a_coef, m->sigma_a_power, m->sigma_s_coef, m->sigma_s_power,
                          m->linearized ? nullptr : m->bee, m->temp};
  const int slot[6] = {MAT_SA_COEF, MAT_SA_POW, MAT_SS_COEF, MAT_SS_POW, MAT_BEE, MAT_TEMP};
  for (int q = 0; q < 6; ++q)
    if (src[q]) std::memcpy(st.data() + slot[q] * C, src[q], 8 * C);
  HIPCHK(hipStreamSynchronize(c.stream), "material sync");
  HIPCHK(hipMemcpy(c.d_mat, st.data(), 8 * MAT_N * C, hipMemcpyHostToDevice), "material H2D");
  HIPCHK(hipMemcpy(c.d_radsrc, m->radsource, 8 * C, hipMemcpyHostToDevice), "radsource H2D");
  HIPCHK(hipMemcpy(c.d_tsurf, ts.data(), 8 * ts.size(), hipMemcpyHostToDevice), "t_surf H2D");
  c.step_counters_clear = false;
  c.mat_linearized = m->linearized != 0;
  c.mat_marshak = m->marshak != 0;
  c.mat_alpha = m->alpha;
  c.mat_set = true;
  return IMC_OK;
}

int imc_step(imc_ctx* h, int64_t step, double dt, int32_t first, int64_t n_input, int64_t n_max, double cellmin,
             double* temp, double* radenergydens, double* matenergydens, imc_step_result* res) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  if (!c.mat_set) return fail(c, IMC_EINVAL, "imc_set_material must be called before imc_step");
  if (step < 0 || step > 0xFFFFFFFFll) return fail(c, IMC_EINVAL, "step out of range");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, IMC_EINVAL, "dt must be positive and finite");
  if (n_input < 0 || n_max < 0 || !(cellmin >= 0.0)) return fail(c, IMC_EINVAL, "negative particle budget");
  c.step = step;
  c.dt = dt;
  const size_t C = (size_t)c.C;
  // Update.update on the device, then the per-cell tables of the tracking type
  HIPCHK(hipEventRecord(c.evs[0], c.stream), "event");
  // the previous imc_step's last kernel leaves the flags and transport counters at zero
  const bool clear = c.step_counters_clear;
  c.step_counters_clear = false;
  if (!clear)
    HIPCHK(launch_zero(c, (char*)c.d_sc + offsetof(Scalars, prep_underflow),
                       offsetof(Scalars, pad_) - offsetof(Scalars, prep_underflow)), "zero flags");
  HIPCHK(launch_update(c, first != 0), "update kernel");
  HIPCHK(launch_prep(c), "prep kernel");
  if (c.prec == 16) {                   // binary16 can round 1/sigma_c to 0: refuse before sourcing
    int rc = fetch_scalars(c);
    if (rc) return rc;
    if (c.h_sc.field_error) return fail(c, IMC_ENUMERIC, "non-finite or out-of-range coefficients after the update");
    if (c.h_sc.prep_underflow)
      return fail(c, IMC_ENUMERIC, (std::to_string(c.h_sc.prep_underflow) + " cell(s) have a collision mean free "
                  "path 1/sigma_c that rounds to 0 in the tracking precision; use a wider PRECISION").c_str());
  }
  c.step_set = true;
  // sourcing, as imc_source, without its closing wait when the birth bound fits
  HIPCHK(hipEventRecord(c.ev0, c.stream), "event");
  HIPCHK(launch_source_entries(c), "source entries kernel");
  HIPCHK(launch_tree_sum(c, c.d_ent_term, c.NE, (double*)((char*)c.d_sc + offsetof(Scalars, E_tot))),
         "energy sum kernel");
  const int64_t n_census = c.h_sc.n_census;
  const double budget = std::ceil(std::max((double)n_input, cellmin));
  const double bound_d = budget * ((c.cfg.quirks & IMC_QUIRK_Q1_RADSRC_NBODY) ? 2.0 : 1.0) +
                         (double)c.NE * (std::ceil(cellmin) + 1.0) + 16.0;
  const int64_t bound = bound_d < 4e18 ? (int64_t)bound_d : INT64_MAX / 4;
  const bool direct = n_census + bound <= c.buf[c.cur].cap && birth_hist_len(bound) <= c.bhist_len;
  HIPCHK(launch_counts_and_scan(c, n_input, n_max, cellmin, direct ? c.buf[c.cur].cap : INT64_MAX), "count kernels");
  int64_t n_bound;                      // the population transport and clean will see, or a bound of it
  if (direct) {
    HIPCHK(launch_births(c, bound), "birth kernel");
    n_bound = n_census + bound;
  } else {
    int rc = fetch_scalars(c);
    if (rc) return rc;
    if (c.h_sc.field_error) return fail(c, IMC_ENUMERIC, "non-finite or out-of-range coefficients after the update");
    if (!std::isfinite(c.h_sc.E_tot) || !(c.h_sc.E_tot > 0.0))
      return fail(c, IMC_ENUMERIC, "total source energy is zero or not finite");
    rc = grow(c, c.h_sc.n_census + c.h_sc.n_births, c.h_sc.n_census);
    if (rc) return rc;
    rc = ensure_bhist(c, birth_hist_len(c.h_sc.n_births));
    if (rc) return rc;
    HIPCHK(launch_births(c, c.h_sc.n_births), "birth kernel");
    n_bound = c.h_sc.n_census + c.h_sc.n_births;
  }
  HIPCHK(hipEventRecord(c.ev1, c.stream), "event");
  // transport
  int grid, tm; size_t lds;
  transport_choose(c, &grid, &tm, &lds);
  if (!clear)
    HIPCHK(launch_zero(c, (char*)c.d_sc + offsetof(Scalars, segments),
                       offsetof(Scalars, prep_underflow) - offsetof(Scalars, segments)), "zero scalars");
  if (tm == TM_GLOBAL || tm == TM_HASH)
    HIPCHK(launch_zero(c, c.d_gtally, 8 * (size_t)c.NB * gtally_copies(c.NB)), "zero tallies");
  HIPCHK(hipEventRecord(c.evk0, c.stream), "event");
  HIPCHK(launch_transport(c, grid, tm, lds), "transport kernel");
  HIPCHK(hipEventRecord(c.evk1, c.stream), "event");
  if (tm == TM_GLOBAL || tm == TM_HASH)
    HIPCHK(launch_sum_copies(c, c.d_gtally, gtally_copies(c.NB), c.NB, c.d_dep), "tally copies");
  else HIPCHK(launch_reduce_partials(c, grid, tm, c.NB, c.d_dep), "tally reduction");
  HIPCHK(hipEventRecord(c.evs[1], c.stream), "event");
  // clean with the census tally, then the material update and the energy-check sums
  HIPCHK(launch_clean(c, n_bound), "clean kernels");
  HIPCHK(hipEventRecord(c.evs[2], c.stream), "event");
  HIPCHK(launch_temperature(c), "temperature kernel");
  HIPCHK(launch_step_sums(c), "energy sums kernel");
  // temperature, radenergydens, matenergydens, nrg_inc and the scalar snapshot: one copy
  HIPCHK(hipMemcpyAsync(c.h_out, c.d_mat + (size_t)MAT_TEMP * C, 8 * MAT_OUT * C + sizeof(Scalars),
                        hipMemcpyDeviceToHost, c.stream), "results D2H");
  HIPCHK(hipEventRecord(c.evs[3], c.stream), "event");
  HIPCHK(hipStreamSynchronize(c.stream), "step sync");
  std::memcpy(&c.h_sc, c.h_out + MAT_OUT * C, sizeof(Scalars));
  c.step_counters_clear = true;
  c.rad_valid = false;
  c.timing.ms_source += elapsed(c.evs[0], c.ev1);
  c.timing.ms_transport += elapsed(c.ev1, c.evs[1]);
  c.timing.ms_transport_kernel += elapsed(c.evk0, c.evk1);
  c.timing.ms_clean += elapsed(c.evs[1], c.evs[2]);
  c.timing.ms_tally += elapsed(c.evs[2], c.evs[3]);
  c.timing.n_transport_launches += 1;
  c.timing.histories += c.h_sc.n_cleaned;
  c.timing.census_out += (int64_t)c.h_sc.census_out;
  c.timing.segments += (int64_t)c.h_sc.segments;
  c.timing.surrogate_events += (int64_t)c.h_sc.surr_events;
  c.timing.truncated += (int64_t)c.h_sc.truncated;
  c.timing.transport_tally_mode = tm;
  c.timing.transport_grid = grid;
  if (c.h_sc.field_error) return fail(c, IMC_ENUMERIC, "non-finite or out-of-range coefficients after the update");
  if (c.h_sc.prep_underflow)
    return fail(c, IMC_ENUMERIC, "a collision mean free path rounds to 0 in the tracking precision");
  if (!std::isfinite(c.h_sc.E_tot) || !(c.h_sc.E_tot > 0.0))
    return fail(c, IMC_ENUMERIC, "total source energy is zero or not finite");
  if (c.h_sc.birth_overflow) return fail(c, IMC_ENOMEM, "births exceed the particle buffer");
  double* outs[4] = {temp, radenergydens, matenergydens, nullptr};
  for (int q = 0; q < 3; ++q)
    if (outs[q]) std::memcpy(outs[q], c.h_out + q * C, 8 * C);
  if (res) {
    res->totalenergy = c.h_sc.E_tot;
    res->totalenergydep = c.h_sc.E_dep;
    res->radenergy = c.h_sc.E_rad;
    res->lostenergy = c.h_sc.E_lost;
    res->histories = c.h_sc.n_cleaned;
    res->n_alive = c.h_sc.n_alive;
    res->segments = (int64_t)c.h_sc.segments;
  }
  return IMC_OK;
}

int imc_step_energy_increase(imc_ctx* h, double* energyincrease) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  if (!c.mat_set || !c.h_out) return fail(c, IMC_EINVAL, "no device-resident step has run");
  if (!energyincrease) return fail(c, IMC_EINVAL, "energyincrease required");
  std::memcpy(energyincrease, c.h_out + 3 * (size_t)c.C, 8 * (size_t)c.C);
  return IMC_OK;
}

// ---------------------------------------------------------------------------------
// learned thick-cell surrogate
// ---------------------------------------------------------------------------------
static int surr_alloc(Ctx& c) {
  if (c.d_surr_params) return IMC_OK;
  bool ok = true;
  ok &= dalloc(&c.d_surr_params, 4 * (size_t)MLP_NPARAM) == hipSuccess;
  ok &= dalloc(&c.d_surr_grad, 4 * (size_t)MLP_NPARAM) == hipSuccess;
  ok &= dalloc(&c.d_surr_m, 4 * (size_t)MLP_NPARAM) == hipSuccess;
  ok &= dalloc(&c.d_surr_v, 4 * (size_t)MLP_NPARAM) == hipSuccess;
  ok &= dalloc(&c.d_surr_wbf, 2 * (size_t)BF_TOTAL) == hipSuccess;
  ok &= dalloc(&c.d_surr_bias, 4 * (size_t)BIAS_TOTAL) == hipSuccess;
  ok &= dalloc(&c.d_surr_loss, 8) == hipSuccess;
  ok &= dalloc(&c.d_surr_events, 8) == hipSuccess;
  ok &= dalloc(&c.d_surr_cells, 16 * (size_t)c.C) == hipSuccess;
  if (!ok) return fail(c, IMC_ENOMEM, "surrogate allocation failed");
  HIPCHK(launch_zero(c, c.d_surr_params, 4 * (size_t)MLP_NPARAM), "zero params");
  HIPCHK(launch_zero(c, c.d_surr_grad, 4 * (size_t)MLP_NPARAM), "zero grad");
  HIPCHK(launch_zero(c, c.d_surr_m, 4 * (size_t)MLP_NPARAM), "zero m");
  HIPCHK(launch_zero(c, c.d_surr_v, 4 * (size_t)MLP_NPARAM), "zero v");
  HIPCHK(surr_make_weights(c), "surrogate weights");
  HIPCHK(hipStreamSynchronize(c.stream), "surrogate alloc sync");
  return IMC_OK;
}

int imc_surrogate_configure(imc_ctx* h, const imc_surrogate_box* box) {
  if (!h || !box) return IMC_EINVAL;
  Ctx& c = h->c;
  const imc_surrogate_box& b = *box;
  if (!(b.tau_cell_min > 0 && b.tau_cell_max > b.tau_cell_min && b.tau_b_min > 0 && b.tau_b_max > b.tau_b_min &&
        b.kappa_min > 0 && b.kappa_max > b.kappa_min && b.depth_min >= 0 && 2 * b.depth_min < b.tau_cell_min))
    return fail(c, IMC_EINVAL, "surrogate box bounds must be positive and increasing (2 depth_min < tau_cell_min)");
  c.surr_box = SurrBox{b.tau_cell_min, b.tau_cell_max, b.tau_b_min, b.tau_b_max, b.kappa_min, b.kappa_max, b.depth_min};
  return surr_alloc(c);
}

int imc_surrogate_set_params(imc_ctx* h, const float* params) {
  if (!h || !params) return IMC_EINVAL;
  Ctx& c = h->c;
  int rc = surr_alloc(c);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c.d_surr_params, params, 4 * (size_t)MLP_NPARAM, hipMemcpyHostToDevice, c.stream), "params H2D");
  HIPCHK(surr_make_weights(c), "surrogate weights");
  HIPCHK(hipStreamSynchronize(c.stream), "params sync");
  return IMC_OK;
}

int imc_surrogate_get_params(imc_ctx* h, float* params) {
  if (!h || !params) return IMC_EINVAL;
  Ctx& c = h->c;
  int rc = surr_alloc(c);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(params, c.d_surr_params, 4 * (size_t)MLP_NPARAM, hipMemcpyDeviceToHost, c.stream), "params D2H");
  HIPCHK(hipStreamSynchronize(c.stream), "params sync");
  return IMC_OK;
}

// He-normal initialisation of the weights, zero biases; resets the Adam state.
int imc_surrogate_init(imc_ctx* h, uint32_t seed) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  int rc = surr_alloc(c);
  if (rc) return rc;
  std::vector<float> p(MLP_NPARAM, 0.0f);
  std::mt19937 gen(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  const float s1 = std::sqrt(2.0f / MLP_IN), s2 = std::sqrt(2.0f / MLP_H), s3 = std::sqrt(1.0f / MLP_H);
  for (int i = 0; i < MLP_IN * MLP_H; ++i) p[OFF_W1 + i] = s1 * nd(gen);
  for (int i = 0; i < MLP_H * MLP_H; ++i) p[OFF_W2 + i] = s2 * nd(gen);
  for (int i = 0; i < MLP_H * MLP_OUT; ++i) p[OFF_W3 + i] = s3 * nd(gen);
  HIPCHK(launch_zero(c, c.d_surr_m, 4 * (size_t)MLP_NPARAM), "zero m");
  HIPCHK(launch_zero(c, c.d_surr_v, 4 * (size_t)MLP_NPARAM), "zero v");
  HIPCHK(launch_zero(c, c.d_surr_grad, 4 * (size_t)MLP_NPARAM), "zero grad");
  c.surr_step = 0;
  return imc_surrogate_set_params(h, p.data());
}

int imc_surrogate_generate(imc_ctx* h, int64_t n, uint32_t seed, int64_t* events) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  if (n <= 0) return fail(c, IMC_EINVAL, "n_samples must be positive");
  int rc = surr_alloc(c);
  if (rc) return rc;
  if (n > c.surr_cap) {
    (void)hipFree(c.d_surr_feat);
    (void)hipFree(c.d_surr_targ);
    c.d_surr_feat = nullptr; c.d_surr_targ = nullptr; c.surr_cap = 0;
    if (dalloc(&c.d_surr_feat, 4 * (size_t)n * MLP_IN) != hipSuccess ||
        dalloc(&c.d_surr_targ, 4 * (size_t)n * 4) != hipSuccess)
      return fail(c, IMC_ENOMEM, "surrogate dataset allocation failed");
    c.surr_cap = n;
  }
  HIPCHK(launch_zero(c, c.d_surr_events, 8), "zero events");
  HIPCHK(surr_generate(c, n, seed, c.d_surr_events), "generate kernel");
  unsigned long long ev = 0;
  HIPCHK(hipMemcpyAsync(&ev, c.d_surr_events, 8, hipMemcpyDeviceToHost, c.stream), "events D2H");
  HIPCHK(hipStreamSynchronize(c.stream), "generate sync");
  c.surr_n = n;
  if (events) *events = (int64_t)ev;
  return IMC_OK;
}

int imc_surrogate_dataset(imc_ctx* h, float* features, float* targets, int64_t n) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  if (n > c.surr_n) return fail(c, IMC_EINVAL, "more samples requested than generated");
  if (features)
    HIPCHK(hipMemcpyAsync(features, c.d_surr_feat, 4 * (size_t)n * MLP_IN, hipMemcpyDeviceToHost, c.stream), "D2H");
  if (targets)
    HIPCHK(hipMemcpyAsync(targets, c.d_surr_targ, 4 * (size_t)n * 4, hipMemcpyDeviceToHost, c.stream), "D2H");
  HIPCHK(hipStreamSynchronize(c.stream), "dataset sync");
  return IMC_OK;
}

// Mini-batch Adam over the generated set, cosine learning-rate decay to lr/20.
int imc_surrogate_train(imc_ctx* h, int32_t epochs, int32_t batch, float lr, float* loss_per_epoch) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  if (c.surr_n <= 0) return fail(c, IMC_EINVAL, "imc_surrogate_generate must be called first");
  if (epochs <= 0 || batch < 64 || (batch % 64) != 0 || !(lr > 0))
    return fail(c, IMC_EINVAL, "epochs > 0, batch a positive multiple of 64, lr > 0 required");
  const int64_t steps_per_epoch = c.surr_n / batch;
  if (steps_per_epoch <= 0) return fail(c, IMC_EINVAL, "batch larger than the data set");
  const int64_t total = steps_per_epoch * epochs;
  int64_t it = 0;
  for (int e = 0; e < epochs; ++e) {
    HIPCHK(launch_zero(c, c.d_surr_loss, 8), "zero loss");
    for (int64_t b = 0; b < steps_per_epoch; ++b, ++it) {
      HIPCHK(surr_grad(c, c.d_surr_feat, c.d_surr_targ, b * batch, batch, 1.0f / batch, c.d_surr_grad,
                       c.d_surr_loss), "gradient kernel");
      const float frac = (float)it / (float)std::max<int64_t>(total - 1, 1);
      const float lr_t = lr * (0.05f + 0.95f * 0.5f * (1.0f + std::cos(3.14159265f * frac)));
      HIPCHK(surr_adam(c, lr_t, ++c.surr_step), "adam kernel");
      HIPCHK(surr_make_weights(c), "surrogate weights");
    }
    float lsum = 0.f;
    HIPCHK(hipMemcpyAsync(&lsum, c.d_surr_loss, 4, hipMemcpyDeviceToHost, c.stream), "loss D2H");
    HIPCHK(hipStreamSynchronize(c.stream), "train sync");
    if (loss_per_epoch) loss_per_epoch[e] = lsum / (float)(steps_per_epoch * batch);
  }
  return IMC_OK;
}

int imc_surrogate_forward(imc_ctx* h, const float* features, int64_t n, float* outputs) {
  if (!h || !features || !outputs || n <= 0) return IMC_EINVAL;
  Ctx& c = h->c;
  int rc = surr_alloc(c);
  if (rc) return rc;
  float *df = nullptr, *dout = nullptr;
  if (dalloc(&df, 4 * (size_t)n * MLP_IN) != hipSuccess || dalloc(&dout, 4 * (size_t)n * MLP_OUT) != hipSuccess) {
    (void)hipFree(df);
    return fail(c, IMC_ENOMEM, "forward buffers");
  }
  hipError_t e = hipMemcpyAsync(df, features, 4 * (size_t)n * MLP_IN, hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess) e = surr_forward(c, df, n, dout);
  if (e == hipSuccess) e = hipMemcpyAsync(outputs, dout, 4 * (size_t)n * MLP_OUT, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  (void)hipFree(df);
  (void)hipFree(dout);
  if (e != hipSuccess) return hipfail(c, e, "surrogate forward");
  return IMC_OK;
}

// Gradient of the mean training loss over n samples (no parameter update).
int imc_surrogate_gradient(imc_ctx* h, const float* features, const float* targets, int64_t n, float* grad,
                           float* loss) {
  if (!h || !features || !targets || !grad || n <= 0) return IMC_EINVAL;
  Ctx& c = h->c;
  int rc = surr_alloc(c);
  if (rc) return rc;
  float *df = nullptr, *dt = nullptr, *dg = nullptr, *dl = nullptr;
  if (dalloc(&df, 4 * (size_t)n * MLP_IN) != hipSuccess || dalloc(&dt, 4 * (size_t)n * 4) != hipSuccess ||
      dalloc(&dg, 4 * (size_t)MLP_NPARAM) != hipSuccess || dalloc(&dl, 8) != hipSuccess) {
    (void)hipFree(df); (void)hipFree(dt); (void)hipFree(dg);
    return fail(c, IMC_ENOMEM, "gradient buffers");
  }
  float lsum = 0.f;
  hipError_t e = hipMemcpyAsync(df, features, 4 * (size_t)n * MLP_IN, hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dt, targets, 4 * (size_t)n * 4, hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess) e = launch_zero(c, dg, 4 * (size_t)MLP_NPARAM);
  if (e == hipSuccess) e = launch_zero(c, dl, 8);
  if (e == hipSuccess) e = surr_grad(c, df, dt, 0, n, 1.0f / (float)n, dg, dl);
  if (e == hipSuccess) e = hipMemcpyAsync(grad, dg, 4 * (size_t)MLP_NPARAM, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&lsum, dl, 4, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  (void)hipFree(df); (void)hipFree(dt); (void)hipFree(dg); (void)hipFree(dl);
  if (e != hipSuccess) return hipfail(c, e, "surrogate gradient");
  if (loss) *loss = lsum / (float)n;
  return IMC_OK;
}

static int fit_partials(Ctx& c);

int imc_surrogate_enable(imc_ctx* h, int32_t on) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  if (on && c.is2d) return fail(c, IMC_EINVAL, "the surrogate is available for 1D geometry");
  int rc = surr_alloc(c);
  if (rc) return rc;
  c.surr_enabled = on != 0;
  if (on) c.rw_enabled = false;
  return fit_partials(c);
}

// First-exit law of one-dimensional diffusion from (-1, 1) started at 0, in the dimensionless
// time a = D t / R0^2: F(a) = P(exit by a).  Small a: F = 2 sum_n (-1)^n erfc((2n+1) / (2 sqrt a))
// (images of the two absorbing ends); large a: F = 1 - (4/pi) sum_k (-1)^k/(2k+1)
// exp(-(2k+1)^2 pi^2 a / 4) (eigenfunction expansion).  Both converge fast on their side of 0.25.
static double rw_exit_cdf(double a) {
  if (a <= 0.0) return 0.0;
  double f = 0.0;
  if (a < 0.25) {
    const double r = 1.0 / (2.0 * std::sqrt(a));
    for (int n = 0; n < 8; ++n) f += (n & 1 ? -2.0 : 2.0) * std::erfc((2 * n + 1) * r);
  } else {
    const double pi = 3.14159265358979323846;
    double s = 0.0;
    for (int k = 0; k < 40; ++k) {
      const double m = 2 * k + 1;
      s += (k & 1 ? -1.0 : 1.0) / m * std::exp(-m * m * pi * pi * a / 4.0);
    }
    f = 1.0 - 4.0 / pi * s;
  }
  return std::min(std::max(f, 0.0), 1.0);
}

int imc_sample_planck(imc_ctx* h, int64_t n, const double* temp, uint32_t stream, double* freq) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  if (n < 0 || n > (int64_t)INT32_MAX * 256) return fail(c, IMC_EINVAL, "sample count out of range");
  if (n == 0) return IMC_OK;
  if (!temp || !freq) return fail(c, IMC_EINVAL, "null temperature or frequency array");
  for (int64_t i = 0; i < n; ++i)
    if (!(std::isfinite(temp[i]) && temp[i] >= 0.0)) return fail(c, IMC_EINVAL, "temperatures must be finite and >= 0");
  double *dt = nullptr, *df = nullptr;
  if (dalloc(&dt, 8 * (size_t)n) != hipSuccess || dalloc(&df, 8 * (size_t)n) != hipSuccess) {
    (void)hipFree(dt);
    return fail(c, IMC_ENOMEM, "Planck sample buffers");
  }
  hipError_t e = hipMemcpyAsync(dt, temp, 8 * (size_t)n, hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess) e = launch_planck(c, n, dt, df, stream);
  if (e == hipSuccess) e = hipMemcpyAsync(freq, df, 8 * (size_t)n, hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  (void)hipFree(dt);
  (void)hipFree(df);
  if (e != hipSuccess) return hipfail(c, e, "Planck sampling");
  return IMC_OK;
}

// First exit of two-dimensional diffusion from the centre of the unit disk, in a = D t / R0^2:
// F(a) = 1 - sum_k c_k e^{-j_k^2 a}, c_k = 2 / (j_k J1(j_k)), j_k the zeros of J0 (eigenfunction
// expansion of the disk with an absorbing rim).  DISK_TERMS terms converge for a >= 3.2e-5, below
// which F < 1e-300 is taken as 0.
constexpr int DISK_TERMS = 400;

struct DiskSeries {
  double j[DISK_TERMS], c[DISK_TERMS];
  DiskSeries() {
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < DISK_TERMS; ++k) {
      const double beta = (k + 0.75) * pi;                   // McMahon's estimate of the (k+1)-th zero
      double x = beta + 1.0 / (8.0 * beta);
      for (int it = 0; it < 8; ++it) x += std::cyl_bessel_j(0.0, x) / std::cyl_bessel_j(1.0, x);   // J0' = -J1
      j[k] = x;
      c[k] = 2.0 / (x * std::cyl_bessel_j(1.0, x));
    }
  }
  double exit_cdf(double a) const {
    if (a < 3.2e-5) return 0.0;
    double s = 0.0;
    for (int k = 0; k < DISK_TERMS; ++k) s += c[k] * std::exp(-j[k] * j[k] * a);
    return std::min(std::max(1.0 - s, 0.0), 1.0);
  }
};

// Inverse of a CDF on [lo, hi] at u by bisection.
static double invert_cdf(const std::function<double(double)>& cdf, double u, double lo, double hi) {
  for (int it = 0; it < 100; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (cdf(mid) < u) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

// Transport.MC_RW (imc_transport.jl:212-479) with the random-walk table of :786-797, as a
// correct analytic walk (see the RW notes in imc_kernels.hip).  Tables: the exit time's inverse
// CDF at u = i / RW_TABLE_POINTS (the tail past (N-1)/N uses the closed form); in 2D followed by
// the census radius's inverse CDF rho J1(j1 rho) / J1(j1) at u = i / (RW_RADIUS_POINTS - 1).
int imc_random_walk_enable(imc_ctx* h, int32_t on, double r_min) {
  if (!h) return IMC_EINVAL;
  Ctx& c = h->c;
  if (on && !(r_min >= 1.0 && std::isfinite(r_min)))
    return fail(c, IMC_EINVAL, "r_min must be at least one mean free path");
  if (on && !c.d_rw_table) {
    const size_t n = RW_TABLE_POINTS + (c.is2d ? RW_RADIUS_POINTS : 0);
    std::vector<float> tab(n);
    if (!c.is2d) {
      for (int i = 0; i < RW_TABLE_POINTS; ++i)
        tab[i] = (float)invert_cdf(rw_exit_cdf, (double)i / RW_TABLE_POINTS, 0.0, 60.0);
    } else {
      const DiskSeries d;
      for (int i = 0; i < RW_TABLE_POINTS; ++i)
        tab[i] = (float)invert_cdf([&](double a) { return d.exit_cdf(a); }, (double)i / RW_TABLE_POINTS, 0.0, 40.0);
      const double j1 = d.j[0], J1j1 = std::cyl_bessel_j(1.0, j1);
      const auto rad_cdf = [&](double r) { return r * std::cyl_bessel_j(1.0, j1 * r) / J1j1; };
      for (int i = 0; i < RW_RADIUS_POINTS; ++i)
        tab[RW_TABLE_POINTS + i] = (float)invert_cdf(rad_cdf, (double)i / (RW_RADIUS_POINTS - 1), 0.0, 1.0);
    }
    if (dalloc(&c.d_rw_table, sizeof(float) * n) != hipSuccess)
      return fail(c, IMC_ENOMEM, "random-walk table allocation failed");
    HIPCHK(hipMemcpy(c.d_rw_table, tab.data(), sizeof(float) * n, hipMemcpyHostToDevice), "random-walk table H2D");
  }
  c.rw_enabled = on != 0;
  if (on) {
    c.rw_min = r_min;
    c.surr_enabled = false;
  }
  return fit_partials(c);
}

// The tally partial buffer must cover the transport grid of the active configuration.
static int fit_partials(Ctx& c) {
  const size_t need = partials_needed(c);
  if (need > c.partial_bytes) {
    (void)hipFree(c.d_partial);
    c.d_partial = nullptr;
    if (dalloc(&c.d_partial, need) != hipSuccess) return fail(c, IMC_ENOMEM, "tally partial allocation failed");
    c.partial_bytes = need;
  }
  return IMC_OK;
}

}  // extern "C"
