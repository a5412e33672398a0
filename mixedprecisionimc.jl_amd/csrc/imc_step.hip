// imc_step.hip -- the FP64 work around the particle loop, for the device-resident step:
//   k_update      : Update.update (imc_update.jl:12-70) per cell, writing the step fields that
//                   k_prep and the sourcing kernels read, and at t = 0 the initial material
//                   energy density (imc_tally.jl:29-34);
//   k_temperature : the material part of Tally.tally (imc_tally.jl:44-75);
//   k_step_sums   : the totals EnergyCheck.energychecker needs (imc_energycheck.jl:10-38).
// Each follows the host restatement operation for operation (same products, same order): where
// no transcendental is involved, the same inputs give the host's values bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "imc_device.h"
#include "imc_internal.h"

namespace imc {

struct UpdateArgs {
  const double* sa_coef;
  const double* sa_pow;
  const double* ss_coef;
  const double* ss_pow;
  double* bee;             // written when linearized (bee = 4 a T^3)
  const double* temp;      // material temperature, device-resident
  double* matE;            // material energy density (formed when first)
  double* f_sa;            // step fields read by k_prep and sourcing
  double* f_ss;
  double* f_fleck;
  double* f_temp;
  const double* radsrc;
  double phys_a, phys_c, alpha, s, dt;
  int C, linearized, marshak, first;
  Scalars* sc;
};

__global__ __launch_bounds__(256) void k_update(UpdateArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.C) return;
  const double T = a.temp[i];
  const double T3 = T * T * T;
  double beta;
  if (a.linearized) {                                       // imc_update.jl:18-21
    a.bee[i] = 4.0 * a.phys_a * T3;
    beta = 1.0;
  } else {
    beta = (4.0 * a.phys_a * T3) / a.bee[i];                // :22-23
  }
  double sa = a.sa_coef[i] * pow(T, a.sa_pow[i]);           // :28-31
  if (a.marshak) sa = a.sa_coef[i] / T / T / T;             // :32-34
  const double ss = a.ss_coef[i] * pow(T, a.ss_pow[i]);
  const double one = 1.0;
  double prod, esc;
  const double fv[6] = {a.s, a.alpha, beta, a.phys_c, a.dt, sa};
  sorter_dev<double, 6>(fv, true, &one, 1, prod, esc);      // :39 (Fleck factor)
  const double f = 1.0 / (1.0 + prod);
  a.f_sa[i] = sa;
  a.f_ss[i] = ss;
  a.f_fleck[i] = f;
  a.f_temp[i] = T;
  if (a.first) {                                            // imc_tally.jl:29-34
    const double mv[10] = {f, sa, a.phys_a, a.phys_c, T, T, T, T, a.dt, a.s};
    sorter_dev<double, 10>(mv, true, &one, 1, prod, esc);
    a.matE[i] = prod;
  }
  // the checks imc_set_step_fields applies to host-supplied fields
  const bool bad = !isfinite(sa) || !isfinite(ss) || !isfinite(f) || !isfinite(T) || !isfinite(a.radsrc[i]) ||
                   sa < 0.0 || ss < 0.0 || !(f >= 0.0 && f <= 1.0);
  if (bad) atomicAdd(&a.sc->field_error, 1u);
}

hipError_t launch_update(Ctx& c, int first) {
  UpdateArgs a;
  const size_t C = (size_t)c.C;
  a.sa_coef = c.d_mat + MAT_SA_COEF * C;
  a.sa_pow = c.d_mat + MAT_SA_POW * C;
  a.ss_coef = c.d_mat + MAT_SS_COEF * C;
  a.ss_pow = c.d_mat + MAT_SS_POW * C;
  a.bee = c.d_mat + MAT_BEE * C;
  a.temp = c.d_mat + MAT_TEMP * C;
  a.matE = c.d_mat + MAT_E * C;
  a.f_sa = c.d_sa;
  a.f_ss = c.d_ss;
  a.f_fleck = c.d_fleck;
  a.f_temp = c.d_temp;
  a.radsrc = c.d_radsrc;
  a.phys_a = c.cfg.phys_a;
  a.phys_c = c.cfg.phys_c;
  a.alpha = c.mat_alpha;
  a.s = c.cfg.distancescale;
  a.dt = c.dt;
  a.C = c.C;
  a.linearized = c.mat_linearized;
  a.marshak = c.mat_marshak;
  a.first = first;
  a.sc = c.d_sc;
  k_update<<<(unsigned)((C + 255) / 256), 256, 0, c.stream>>>(a);
  return hipGetLastError();
}

// nrg_inc = sum_k (energydep_k - emitted_k) / scale_k; matE += nrg_inc; T = matE^(1/4) when
// linearized, else T += nrg_inc / bee (imc_tally.jl:46-75).  Also stages radenergydens next
// to the other results.
__global__ __launch_bounds__(256) void k_temperature(int C, int K, Geo g, const double* __restrict__ dep,
                                                     const double* __restrict__ emitted, const double* __restrict__ bee,
                                                     const double* __restrict__ rad, double* __restrict__ mat,
                                                     int linearized) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  double inc = 0.0;
  for (int k = 0; k < K; ++k) inc = inc + (dep[(size_t)k * C + i] - emitted[(size_t)k * C + i]) / g.scales[k];
  double* temp = mat + (size_t)MAT_TEMP * C;
  const double m = mat[(size_t)MAT_E * C + i] + inc;
  mat[(size_t)MAT_E * C + i] = m;
  mat[(size_t)MAT_INC * C + i] = inc;
  mat[(size_t)MAT_RAD * C + i] = rad[i];
  temp[i] = linearized ? pow(m, 0.25) : temp[i] + inc / bee[i];
}

hipError_t launch_temperature(Ctx& c) {
  const size_t C = (size_t)c.C;
  const Geo g = make_geo(c);
  k_temperature<<<(unsigned)((C + 255) / 256), 256, 0, c.stream>>>(
      c.C, c.K, g, c.d_dep, c.d_emitted, c.d_mat + MAT_BEE * C, c.d_rad, c.d_mat, c.mat_linearized);
  return hipGetLastError();
}

// One block: E_dep = sum_k sum_i energydep_k[i] * vol_i / scale_k (summed per scale, then over
// scales in order, as imc_tally.jl:47-57), E_rad = sum_i radenergydens[i] * vol_i, and the lost
// bin.  Strided per-thread partials and a fixed tree: the result does not depend on timing.
// Two passes: SUMS_BLOCKS blocks each reduce a grid-strided share of the cells for every scale
// and for the radiation energy; one block then adds the block partials in a fixed order (the
// sums are deterministic) and writes the scalars.
constexpr int SUMS_BLOCKS = 256, SUMS_TPB = 256;

size_t step_sums_scratch_bytes(const Ctx& c) { return 8 * (size_t)(c.K + 1) * SUMS_BLOCKS; }

__device__ __forceinline__ double block_tree_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int h = SUMS_TPB / 2; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  const double s = red[0];
  __syncthreads();
  return s;
}

__global__ __launch_bounds__(SUMS_TPB) void k_step_sums_part(int C, int K, int nx, Geo g, const double* __restrict__ dx,
                                                             const double* __restrict__ dy,
                                                             const double* __restrict__ dep,
                                                             const double* __restrict__ rad, double* __restrict__ part) {
  __shared__ double red[SUMS_TPB];
  const int stride = gridDim.x * SUMS_TPB;
  for (int k = 0; k <= K; ++k) {
    double acc = 0.0;
    for (int i = blockIdx.x * SUMS_TPB + threadIdx.x; i < C; i += stride) {
      const double dxi = dx[i % nx], dyi = dy[i / nx];       // host: dep * (dx dy) / scale, rad * dx * dy
      acc += k < K ? dep[(size_t)k * C + i] * (dxi * dyi) / g.scales[k] : rad[i] * dxi * dyi;
    }
    const double s = block_tree_sum(acc, red);
    if (threadIdx.x == 0) part[(size_t)k * gridDim.x + blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(SUMS_TPB) void k_step_sums(int C, int K, int nblocks, const double* __restrict__ part,
                                                        const double* __restrict__ dep, Scalars* sc, Scalars* snap) {
  __shared__ double red[SUMS_TPB];
  double edep = 0.0, erad = 0.0;
  for (int k = 0; k <= K; ++k) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += SUMS_TPB) acc += part[(size_t)k * nblocks + b];
    const double s = block_tree_sum(acc, red);
    if (k < K) edep += s; else erad = s;
  }
  if (threadIdx.x == 0) {
    sc->E_dep = edep;
    sc->E_rad = erad;
    sc->E_lost = dep[(size_t)K * C];
    *snap = *sc;               // the step's scalars, next to its arrays for one copy back
    // and the counters the next step needs at zero (no fill launches there)
    sc->segments = 0; sc->census_out = 0; sc->surr_events = 0; sc->work = 0; sc->truncated = 0;
    sc->prep_underflow = 0; sc->field_error = 0;
  }
}

hipError_t launch_step_sums(Ctx& c) {
  const Geo g = make_geo(c);
  Scalars* snap = (Scalars*)(c.d_mat + (size_t)MAT_N * c.C);
  double* part = (double*)(snap + 1);
  const int nblocks = (int)std::min<int64_t>(SUMS_BLOCKS, (c.C + SUMS_TPB - 1) / SUMS_TPB);
  k_step_sums_part<<<nblocks, SUMS_TPB, 0, c.stream>>>(c.C, c.K, c.nx, g, c.d_dx, c.d_dy, c.d_dep, c.d_rad, part);
  k_step_sums<<<1, SUMS_TPB, 0, c.stream>>>(c.C, c.K, nblocks, part, c.d_dep, c.d_sc, snap);
  return hipGetLastError();
}

}  // namespace imc
