/*
 * imc_oracle.h -- CPU restatement of MixedPrecisionIMC.jl's particle engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for libimc_hip: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product (mixedprecisionimc.jl_amd/) never links, imports or calls it.
 *
 * Every function follows the reference Julia source line by line; the citation
 * is given at each definition in imc_oracle.cpp.  Random numbers are NOT Julia's
 * global MersenneTwister/Xoshiro stream (that cannot be reproduced without
 * Julia): every draw is Philox4x32-10 keyed by (seed, step) with a counter made
 * of (particle slot or birth index, event index, tag, group).  The same keying is
 * the specification the GPU kernels implement, so per-history replay is possible.
 *
 * Particle records use the reference's AoS layout (doubles holding values of the
 * run precision):
 *   1D: [origin, t, cell, x, mu, nu, E, E0, scale]        (imc_sourcing.jl:173)
 *   2D: [t, ix, iy, x, y, theta, nu, E, E0, scale]        (imc_sourcing.jl:276)
 * Cell indices are 1-based as in Julia.  The kill flag is slot 8 (1-based) set to
 * -1.0 (imc_transport.jl:104 / :593), i.e. E0 in 1D and E in 2D.
 */
#ifndef IMC_ORACLE_H
#define IMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_VACUUM = 0, ORC_REFLECT = 1 };
enum { ORC_QUIRK_Q1_RADSRC_NBODY = 1, ORC_QUIRK_Q2_LR_FACE_DX = 2 };

typedef struct {
  int32_t geometry;      /* 1 or 2 */
  int32_t prec;          /* 64: f64 track/f64 weight; 32: f32/f32; 16: f16 track / f32 weight */
  int32_t nx, ny;        /* ny = 1 in 1D */
  int32_t K;             /* number of energy scales (<= 16) */
  int32_t quirks;        /* ORC_QUIRK_* bits: reproduce reference quirks Q1/Q2 */
  int32_t bc[4];         /* (LEFT, RIGHT, TOP, BOTTOM): ORC_VACUUM / ORC_REFLECT */
  uint32_t seed;
  int32_t pad_;
  double scales[16];     /* descending */
  double phys_c, phys_a, distancescale;
} orc_problem;

typedef struct {
  double dt;
  const double* dx;        /* nx */
  const double* dy;        /* ny (ignored in 1D) */
  const double* temp;      /* C = nx*ny, column-major (ix fastest) */
  const double* fleck;     /* C */
  const double* sigma_a;   /* C, already divided by distancescale */
  const double* sigma_s;   /* C */
  const double* radsource; /* C */
  const double* tsurf[4];  /* 1D: [0][0]=left, [1][0]=right.  2D: bottom(nx), top(nx), left(ny), right(ny) */
} orc_step;

void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
double orc_round_half(double x);
int orc_sorter(int prec, const double* vals, int n, int promote64, const double* scales, int K,
               double* product, double* scale);

int64_t orc_source(const orc_problem* pb, const orc_step* st, uint32_t step, int64_t n_census,
                   int64_t n_input, int64_t n_max, double cellmin, double* rec, int64_t cap,
                   double* emitted, double* etot);
int orc_transport(const orc_problem* pb, const orc_step* st, uint32_t step, double* rec, int64_t n,
                  double* energydep, double* lost, int64_t* segments, int nthreads);
int64_t orc_clean(const orc_problem* pb, double* rec, int64_t n);
void orc_census_tally(const orc_problem* pb, const orc_step* st, const double* rec, int64_t n,
                      double* radenergydens);

typedef struct {
  orc_problem pb;
  const double* dx; const double* dy;
  const double* sa0; const double* sa_pow;   /* sigma_a coefficient (already / distancescale), power */
  const double* ss0; const double* ss_pow;
  const double* bee; const double* radsource;
  const double* tsurf[4];
  double t_init, alpha;
  int32_t linearized, marshak, timestepping; /* timestepping: 0 CONSTANT, 1 RAMP */
  int32_t nthreads;
  double dt, t_end, dt0, k, dtmax;
  int64_t n_input, n_max;
  double cellmin;
  int64_t max_steps;                         /* 0: run to t_end */
  int64_t time_from;                         /* steps before this index are run but not counted */
} orc_run_cfg;

typedef struct {
  double* temp;            /* C */
  double* radenergydens;   /* C */
  double* matenergydens;   /* C */
  double* energy_error;    /* per step, capacity err_cap */
  int64_t err_cap;
  int64_t steps_done;
  int64_t histories;
  int64_t segments;
  double t_hot;            /* wall seconds in source+transport+clean+census tally, steps >= time_from */
  double t_final;
  int64_t n_alive;
} orc_run_out;

int orc_run(const orc_run_cfg* cfg, orc_run_out* out);

// Planck frequencies nu_i = temp_i * x_i, x from the normalised Planck spectrum (imc_sourcing.jl:372-399).
void orc_planck(int64_t n, const double* temp, double* freq, uint32_t seed, uint32_t stream);
// Mixture terms n of samples 0 .. n-1; u_force >= 0 replaces the term uniform.
void orc_planck_terms(int64_t n, uint32_t seed, uint32_t stream, double u_force, int64_t* terms);

#ifdef __cplusplus
}
#endif
#endif
