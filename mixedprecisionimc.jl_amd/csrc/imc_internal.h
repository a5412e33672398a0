// imc_internal.h -- library-private context and kernel-launcher declarations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/imc_hip.h"

namespace imc {

// The global-atomic transport tally is spread over copies (block b adds into copy
// b mod ncopies) so that hot bins do not serialise on one address; the copies are summed in
// a fixed order after the launch.  ncopies = min(GTALLY_MAX_COPIES, GTALLY_BUDGET / (8 NB)).
constexpr int GTALLY_MAX_COPIES = 64;
constexpr size_t GTALLY_BUDGET = (size_t)256 << 20;
inline int gtally_copies(int64_t NB) {
  const int64_t k = (int64_t)(GTALLY_BUDGET / (8 * (size_t)NB));
  return (int)(k < 1 ? 1 : (k > GTALLY_MAX_COPIES ? GTALLY_MAX_COPIES : k));
}

// Structure-of-arrays particle store in HBM.  Element types follow the precision:
// x/y/mu(theta)/t in T, E/E0 in W, cellk = linear cell | scale index << 28 (DEAD if killed).
struct SoA {
  uint32_t* cellk = nullptr;
  void* x = nullptr;
  void* y = nullptr;      // 2D only
  void* mu = nullptr;     // 1D: direction cosine; 2D: angle theta
  void* t = nullptr;
  void* E = nullptr;
  void* E0 = nullptr;
  int64_t cap = 0;
};

// Device-resident scalars of one step.
struct Scalars {
  int64_t n_census;            // particles in [0, n_census) before sourcing
  int64_t n_births;            // births appended by the last sourcing
  int64_t n_total;             // transport range [0, n_total)
  int64_t n_alive;             // after clean
  double E_tot;                // mesh.totalenergy
  unsigned long long segments; // event-loop passes of the last transport
  unsigned long long census_out;
  unsigned long long surr_events;  // surrogate evaluations of the last transport
  uint32_t work;               // transport work-queue cursor
  uint32_t truncated;          // histories ended by the per-history event cap
  uint32_t prep_underflow;     // cells whose 1/sigma_c rounds to 0 in the tracking type
  uint32_t birth_overflow;     // the step's births did not fit the particle buffer
  // device-resident step (imc_step): written by its kernels, not cleared by transport
  uint32_t field_error;        // cells whose updated coefficients are non-finite or out of range
  uint32_t pad_;
  double E_dep;                // mesh.totalenergydep
  double E_rad;                // radiation energy (sum of radenergydens * volume)
  double E_lost;               // the step's vacuum loss
  int64_t n_cleaned;           // population handed to the last clean (the step's histories)
};

constexpr int RW_TABLE_POINTS = 4096;   // analytic random walk: inverse-CDF table of the exit time
constexpr int RW_RADIUS_POINTS = 1024;  // 2D walk: inverse CDF of the census radius (fundamental mode)

struct SurrBox {               // training box of the thick-cell surrogate (see imc_mlp.h)
  float tau_cell_min, tau_cell_max;
  float tau_b_min, tau_b_max;
  float kappa_min, kappa_max;
  float depth_min;             // distance to the nearest face, in mean free paths
};

struct Geo {                   // kernel view of the geometry and step coefficients
  int nx, ny, C, K, NB;
  int bc[4];
  uint32_t seed, step;
  double dt, c, s;             // dt, phys_c, distancescale
  const void* dxs;             // T: dx*s (1D: per cell; 2D: per column)
  const void* dys;             // T: dy*s per row (2D)
  const void* dxw;             // W: dx
  const void* dyw;             // W: dy
  const void* sigcol;          // T: sigma_a(1-f)+sigma_s per cell
  const void* sigdep;          // W: sigma_a*f per cell
  const void* isig;            // T: 1/(sigma_a(1-f)+sigma_s) per cell
  const void* idxw;            // W: 1/dx per column
  const void* idyw;            // W: 1/dy per row (2D)
  double ics;                  // 1/(phys_c * distancescale)
  int cell_off;                // 1D: byte offset of the LDS cell table, -1 when cells are read from HBM
  int gcopies;                 // rows of the global-atomic tally
  int hash_bits;               // log2 of the hashed tally's LDS slots
  double scales[16];
};

struct Ctx {
  imc_config cfg;
  int prec = 64, nx = 1, ny = 1, C = 1, K = 1, NB = 2;
  bool is2d = false, geo_set = false, step_set = false;
  std::string err;
  hipStream_t stream = nullptr;
  int device = 0, num_cu = 256;
  int64_t step = 0;
  double dt = 0.0;

  // geometry / coefficients
  double* d_dx = nullptr; double* d_dy = nullptr;          // FP64 copies
  void* d_dxs = nullptr; void* d_dys = nullptr;            // T
  void* d_dxw = nullptr; void* d_dyw = nullptr;            // W
  double* d_fields = nullptr;                              // the step fields below, one allocation
  double* d_temp = nullptr; double* d_fleck = nullptr; double* d_sa = nullptr; double* d_ss = nullptr;
  double* d_radsrc = nullptr; double* d_tsurf = nullptr;   // tsurf packed (1D: 2, 2D: 2nx+2ny)
  double* h_fields = nullptr;                              // pinned staging block of the upload
  size_t fields_len = 0;
  hipEvent_t ev_fields = nullptr;                          // end of the last upload from h_fields
  void* d_sigcol = nullptr; void* d_sigdep = nullptr;
  void* d_isig = nullptr; void* d_idxw = nullptr; void* d_idyw = nullptr;   // reciprocals

  // sourcing entries (faces, body cells, radiation-source cells)
  int64_t NE = 0;
  double* d_ent_e = nullptr;       // energy (W value)
  double* d_ent_esc = nullptr;     // energy scale value
  double* d_ent_term = nullptr;    // e/escale (W) as FP64
  int32_t* d_ent_k = nullptr;      // scale index
  int64_t* d_ent_cnt = nullptr;    // births
  int64_t* d_ent_off = nullptr;    // exclusive prefix of births
  double* d_emitted = nullptr;     // C x K
  double* d_red = nullptr;         // tree-reduction scratch
  int64_t red_len = 0;

  // particles (ping-pong)
  SoA buf[2];
  int cur = 0;
  Scalars* d_sc = nullptr;
  Scalars h_sc{};                  // last copy of the scalars

  // tallies
  void* d_partial = nullptr;       // per-block partial sums
  size_t partial_bytes = 0;
  double* d_dep = nullptr;         // NB = C*K + 1 (last bin: lost energy)
  double* d_gtally = nullptr;      // global-atomic tally: gtally_copies(NB) rows of NB bins
  double* d_rad = nullptr;         // C
  int64_t* d_tile = nullptr;       // clean: per-tile counts/offsets
  int64_t* d_bhist = nullptr;      // births: per-tile start-time bucket counts/offsets
  int64_t* d_bhist_tiles = nullptr;  // scan tile sums of d_bhist (inside the same allocation)
  int64_t* d_ent_tiles = nullptr;  // scan tile sums of the birth counts (scan_tiles(NE) + 1)
  int64_t bhist_len = 0;
  int64_t tile_len = 0;

  // learned thick-cell surrogate
  SurrBox surr_box{2.5f, 200.f, 2.f, 1e4f, 1e-5f, 4.f, 0.7f};   // the tuned default box
  bool surr_enabled = false;
  int64_t surr_cap = 0, surr_n = 0;
  float* d_surr_feat = nullptr;      // [surr_cap][6]
  float* d_surr_targ = nullptr;      // [surr_cap][4]
  float* d_surr_params = nullptr;    // 4868 FP32 master parameters
  float* d_surr_grad = nullptr;
  float* d_surr_m = nullptr;
  float* d_surr_v = nullptr;
  uint16_t* d_surr_wbf = nullptr;    // BF16 inference copy
  float* d_surr_bias = nullptr;      // FP32 biases
  float* d_surr_loss = nullptr;
  unsigned long long* d_surr_events = nullptr;
  void* d_surr_cells = nullptr;      // [C] float4 per-cell decision inputs
  int surr_step = 0;

  // analytic random walk (imc_random_walk_enable); table of RW_TABLE_POINTS values
  bool rw_enabled = false;
  double rw_min = 8.0;
  float* d_rw_table = nullptr;       // [4096] inverse CDF of the dimensionless exit time

  // host copies of the last transport tallies and census tally
  std::vector<double> h_dep, h_rad;
  bool rad_valid = false;

  // device-resident step (imc_set_material / imc_step): FP64 material state per cell
  bool mat_set = false;
  int mat_linearized = 0, mat_marshak = 0;
  double mat_alpha = 0.0;
  double* d_mat = nullptr;         // MAT_N x C (MatSlot), then a Scalars snapshot of the step
  double* h_out = nullptr;         // pinned copy of MAT_TEMP .. MAT_INC and the snapshot
  bool step_counters_clear = false;  // the last imc_step cleared the counters the next one needs at 0

  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evk0 = nullptr, evk1 = nullptr;
  hipEvent_t evs[4] = {nullptr, nullptr, nullptr, nullptr};   // imc_step stage boundaries
  imc_timing timing{};
};

// Material state offsets inside Ctx::d_mat (multiples of C).
// MAT_TEMP .. MAT_INC are the step's results, in the order imc_step returns them, followed in
// the same allocation by the Scalars snapshot: one copy brings them all back.
enum MatSlot { MAT_SA_COEF = 0, MAT_SA_POW, MAT_SS_COEF, MAT_SS_POW, MAT_BEE, MAT_TEMP, MAT_RAD, MAT_E, MAT_INC, MAT_N };
constexpr int MAT_OUT = MAT_N - MAT_TEMP;     // result arrays per step

Geo make_geo(const Ctx& c);
int state_bytes(int prec, bool is2d);

// Launchers (imc_kernels.hip).  Each returns a hipError_t.
hipError_t launch_prep(Ctx& c);
hipError_t launch_source_entries(Ctx& c);
hipError_t launch_tree_sum(Ctx& c, const double* in, int64_t n, double* out_scalar);
hipError_t launch_counts_and_scan(Ctx& c, int64_t n_input, int64_t n_max, double cellmin, int64_t cap);
hipError_t launch_births(Ctx& c, int64_t n_births);
int64_t birth_hist_len(int64_t n_births);
hipError_t launch_transport(Ctx& c, int grid_blocks, int tally_mode, size_t lds_bytes);
hipError_t launch_clean(Ctx& c, int64_t n_total);   // n_total: the count or an upper bound of it
hipError_t launch_reduce_partials(Ctx& c, int nblocks, int tally_mode, int nbins, double* out);
hipError_t launch_zero(Ctx& c, void* p, size_t bytes);
hipError_t launch_sum_copies(Ctx& c, const double* copies, int ncopies, int64_t n, double* out);
int hash_bits(const Ctx& c);
int transport_choose(const Ctx& c, int* grid_blocks, int* tally_mode, size_t* lds_bytes);
int clean_choose(const Ctx& c, int* grid_blocks, int* tally_mode, size_t* lds_bytes);
// Device-resident step (imc_step.hip).
hipError_t launch_update(Ctx& c, int first);
hipError_t launch_temperature(Ctx& c);
int64_t scan_tiles(int64_t n);
hipError_t launch_scan_i64(Ctx& c, const int64_t* a, int64_t n, int64_t* out, int64_t* tiles);
hipError_t launch_planck(Ctx& c, int64_t n, const double* d_temp, double* d_freq, uint32_t stream);
hipError_t launch_step_sums(Ctx& c);
size_t step_sums_scratch_bytes(const Ctx& c);   // after the scalar snapshot in d_mat
hipError_t surr_generate(Ctx& c, int64_t n, uint32_t seed, unsigned long long* d_events);
hipError_t surr_make_weights(Ctx& c);
hipError_t surr_grad(Ctx& c, const float* feat, const float* targ, int64_t n0, int64_t n, float inv_batch,
                     float* grad, float* loss);
hipError_t surr_adam(Ctx& c, float lr, int step);
hipError_t surr_forward(Ctx& c, const float* feat, int64_t n, float* out);

}  // namespace imc

struct imc_ctx {
  imc::Ctx c;
};
