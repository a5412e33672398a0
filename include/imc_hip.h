/*
 * imc_hip.h -- C ABI of libimc_hip, the MI355X (gfx950) particle engine that stands in
 * for the per-timestep particle loop of MixedPrecisionIMC.jl.
 *
 * The reference has no FFI layer: its hot path is five module functions that main()
 * calls positionally, each mutating its arguments (src/MixedPrecisionIMC.jl:139-146 in
 * 1D, :168-171 in 2D).  Each entry point below replaces one of them; the citation is
 * given on each declaration.  Conventions:
 *   - every function returns IMC_OK (0) or a negative IMC_E* status; imc_last_error()
 *     gives the message of the last failure on that context;
 *   - host arrays are FP64 and Julia column-major (cell (ix,iy) at ix + nx*iy, energy
 *     scale k at + nx*ny*k), copied in and out; the library owns all device memory
 *     (particle structure-of-arrays, Philox key state, tally scratch);
 *   - one context per host thread; calls are synchronous on the context's HIP stream.
 */
#ifndef IMC_HIP_H
#define IMC_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMC_OK 0
#define IMC_EINVAL (-1)     /* bad argument or call order */
#define IMC_EHIP (-2)       /* HIP runtime failure */
#define IMC_ENOMEM (-3)     /* device allocation failed */
#define IMC_ENUMERIC (-4)   /* non-finite value detected (NaN/Inf opacity, energy) */

#define IMC_VACUUM 0
#define IMC_REFLECT 1

/* Reference quirks that can be reproduced on request (product default: fixed). */
#define IMC_QUIRK_Q1_RADSRC_NBODY 1  /* 2D radiation-source births counted with n_body (imc_sourcing.jl:357) */
#define IMC_QUIRK_Q2_LR_FACE_DX 2    /* 2D left/right face births use dx[j] for y (imc_sourcing.jl:301,316) */

typedef struct imc_ctx imc_ctx;

typedef struct {
  int32_t geometry;          /* 1 or 2 (deck GEOMETRY) */
  int32_t precision;         /* 64: FP64 tracking+weights+tallies; 32: FP32 tracking/weights with
                                compensated FP32 tallies; 16: FP16 tracking, FP32 weights,
                                compensated FP32 tallies (deck PRECISION) */
  int32_t nx, ny;            /* cells; ny = 1 in 1D */
  int32_t n_scales;          /* K = length(energyscales), 1..16 */
  int32_t quirks;            /* IMC_QUIRK_* */
  int32_t bc[4];             /* LEFTBC, RIGHTBC, TOPBC, BOTTOMBC (MixedPrecisionIMC.jl:121-122,158-162) */
  uint32_t seed;             /* deck SEED: Philox key word 0 */
  int32_t device;            /* HIP device ordinal */
  double energyscales[16];   /* descending (imc_mesh.jl:130) */
  double phys_c, phys_a, distancescale;  /* constants.jl:9-12, imc_mesh.jl:131 */
  int64_t capacity;          /* initial particle capacity; grows on demand */
  /* Tuning, 0 = automatic.  reserved[0]: transport tally mode (1 per-wave LDS rows,
     2 per-block LDS row, 3 global FP64 atomics, 4 hashed per-block LDS table).  reserved[1]: transport blocks per CU
     (upper bound on the occupancy-derived persistent grid; 0 means 5).  reserved[2]: log2
     of the hashed tally's LDS slots, 8..13 (0 means 12).  Others must be 0. */
  int32_t reserved[8];
} imc_config;

/* Per-call timing (HIP events on the context stream), accumulated since the last reset. */
typedef struct {
  double ms_source, ms_transport, ms_clean, ms_tally;
  double ms_transport_kernel;     /* the transport kernel launches alone */
  int64_t n_transport_launches;
  int64_t histories;              /* particles handed to transport */
  int64_t census_out;             /* particles written back to census */
  int64_t segments;               /* event-loop passes (the reference's `iterations`) */
  int32_t transport_tally_mode;   /* 0 wave rows, 1 block row (LDS atomics), 2 global atomics, 3 hashed */
  int32_t transport_grid;         /* persistent transport blocks of the last launch */
  int64_t surrogate_events;       /* thick-cell walks replaced by the surrogate */
  int64_t truncated;              /* histories ended by the 2^24-event cap (non-finite state) */
} imc_timing;

/* Creates a context: MixedPrecisionIMC.jl:117 (particles = Vector{Vector{}}()) plus the
 * simvars/mesh scalars the hot path reads. */
int imc_create(imc_ctx** ctx, const imc_config* cfg);
int imc_destroy(imc_ctx* ctx);
const char* imc_last_error(const imc_ctx* ctx);

/* Cell widths (mesh.dx, mesh.dy from imc_mesh.jl:42-173).  dy is ignored in 1D. */
int imc_set_geometry(imc_ctx* ctx, const double* dx, const double* dy);

/* Per-step coefficient upload: the outputs of Update.update (imc_update.jl:12-70) that
 * the hot path reads -- sigma_a[:,1], sigma_s[:,1], fleck -- plus mesh.temp, mesh.radsource
 * and mesh.temp_surf.  step is the 0-based timestep index (Philox key word 1).
 * t_surf: 1D -> t_surf[0][0] left, t_surf[1][0] right; 2D -> bottom(nx), top(nx),
 * left(ny), right(ny) as in imc_mesh.jl:222-259. */
int imc_set_step_fields(imc_ctx* ctx, int64_t step, double dt, const double* sigma_a,
                        const double* sigma_s, const double* fleck, const double* temp,
                        const double* radsource, const double* const* t_surf);

/* Sourcing.sourcing(mesh, simvars, particles) -- imc_sourcing.jl:12-370.
 * Writes mesh.emittedenergy (C x K), mesh.totalenergy and the population after births. */
int imc_source(imc_ctx* ctx, int64_t n_input, int64_t n_max, double cellmin,
               double* emittedenergy, double* totalenergy, int64_t* n_alive);

/* Transport.MC / Transport.MC2D(mesh, simvars, particles) -- imc_transport.jl:13-210 and
 * :483-732.  energydep (C x K, per unit length/area, scaled) is overwritten; *lostenergy
 * receives this call's vacuum loss (caller adds, as mesh.lostenergy +=); *segments the
 * number of event-loop passes (simvars.iterations increment). */
int imc_transport(imc_ctx* ctx, double* energydep, double* lostenergy, int64_t* segments);

/* Clean.clean(particles) -- imc_clean.jl:6-19: order-preserving removal of killed
 * particles.  Also forms the census radiation energy density for imc_tally_census. */
int imc_clean(imc_ctx* ctx, int64_t* n_alive);

/* Census part of Tally.tally -- imc_tally.jl:157-190: radenergydens (C). */
int imc_tally_census(imc_ctx* ctx, double* radenergydens);

/* Particle records in the reference's AoS layout (imc_sourcing.jl:173 / :276):
 * 1D [origin, t, cell, x, mu, nu, E, E0, scale], 2D [t, ix, iy, x, y, theta, nu, E, E0, scale],
 * 1-based cells.  Export writes at most cap records; import replaces the population. */
int imc_particle_count(imc_ctx* ctx, int64_t* n);
int imc_export_particles(imc_ctx* ctx, double* records, int64_t cap, int64_t* n);
int imc_import_particles(imc_ctx* ctx, const double* records, int64_t n);

int imc_get_timing(imc_ctx* ctx, imc_timing* out);
int imc_reset_timing(imc_ctx* ctx);
int imc_synchronize(imc_ctx* ctx);

/* Bytes of particle state per particle held between kernels (S_state). */
int imc_state_bytes(const imc_ctx* ctx, int32_t* bytes);

/* ---- Device-resident step ----------------------------------------------------------
 * The whole main-loop body (MixedPrecisionIMC.jl:137-149 in 1D, :166-174 in 2D) on the GPU in
 * FP64 around the particle loop: Update.update (imc_update.jl:12-70), sourcing, transport,
 * clean, Tally.tally (imc_tally.jl:11-75 and the census part :157-190) and the sums of
 * EnergyCheck.energychecker (imc_energycheck.jl:10-38).  The material state (temperature,
 * material energy density) stays on the device between steps; one wait per step returns it.
 *
 * imc_set_material gives the per-cell material description once (column-major, C cells):
 * sigma = coef * T^power for absorption and scattering (mesh.sigma_a[:,2:3], imc_mesh.jl:136-142),
 * bee (mesh.bee; unused when linearized), the initial temperature, radsource and the surface
 * temperatures (laid out as in imc_set_step_fields).  It resets the step sequence. */
typedef struct {
  const double* sigma_a_coef;
  const double* sigma_a_power;
  const double* sigma_s_coef;
  const double* sigma_s_power;
  const double* bee;
  const double* temp;
  const double* radsource;
  const double* t_surf[4];
  double alpha;               /* constants.jl:11 */
  int32_t linearized;         /* LINEARIZED = TRUE: bee = 4aT^3, beta = 1 (imc_update.jl:18-23) */
  int32_t marshak;            /* 1D "MARSHAK WAVE": sigma_a = coef / T^3 (imc_update.jl:32-34) */
} imc_material;

/* Results of one device-resident step (the scalars of the energy check). */
typedef struct {
  double totalenergy;         /* mesh.totalenergy (emitted this step) */
  double totalenergydep;      /* sum over cells and scales of energydep * volume / scale */
  double radenergy;           /* sum over cells of radenergydens * volume */
  double lostenergy;          /* this step's vacuum loss */
  int64_t histories;          /* particles transported */
  int64_t n_alive;            /* census population after clean */
  int64_t segments;           /* event-loop passes (simvars.iterations increment) */
} imc_step_result;

int imc_set_material(imc_ctx* ctx, const imc_material* material);

/* One step: step index (Philox key), dt, and the sourcing budget as in imc_source.  first != 0
 * forms the initial material energy density first (imc_tally.jl:29-34, at t = 0).  On return temp,
 * radenergydens and matenergydens (C each, may be NULL) hold the step's new values. */
int imc_step(imc_ctx* ctx, int64_t step, double dt, int32_t first, int64_t n_input, int64_t n_max,
             double cellmin, double* temp, double* radenergydens, double* matenergydens,
             imc_step_result* result);

/* The last imc_step's energy increase nrg_inc (C), which Tally.tally appends to
 * mesh.energyincrease_saved (imc_tally.jl:46-58). */
int imc_step_energy_increase(imc_ctx* ctx, double* energyincrease);

/* ---- Learned thick-cell surrogate -------------------------------------------------
 * Fills the slot of Transport.MC_RW (imc_transport.jl:212-479, deck RANDOMWALK = TRUE):
 * a 6 -> 64 -> 64 -> 4 network (BF16 matrix-core inference and training, FP32
 * accumulation and master weights) replaces the random walk of a particle inside a 1D
 * cell whose collision optical thickness lies in the training box.  Training data are
 * plain single-cell histories generated on the GPU.  Features and outputs are described
 * in mixedprecisionimc.jl_amd/csrc/imc_mlp.h.  Parameter vector layout (4868 floats):
 * W1[6][64], b1[64], W2[64][64], b2[64], W3[64][4], b3[4] (weights stored in x out). */
#define IMC_SURROGATE_NPARAM 4868
typedef struct {
  float tau_cell_min, tau_cell_max;   /* collision optical thickness sigma_c*dx of the cell */
  float tau_b_min, tau_b_max;         /* path budget in mean free paths */
  float kappa_min, kappa_max;         /* f*sigma_a / sigma_c */
  float depth_min;                    /* applied only this many mean free paths from both faces */
} imc_surrogate_box;   /* default (before any configure): {2.5, 200, 2, 1e4, 1e-5, 4, 0.7} */

int imc_surrogate_configure(imc_ctx* ctx, const imc_surrogate_box* box);
int imc_surrogate_init(imc_ctx* ctx, uint32_t seed);
int imc_surrogate_generate(imc_ctx* ctx, int64_t n_samples, uint32_t seed, int64_t* events);
int imc_surrogate_dataset(imc_ctx* ctx, float* features, float* targets, int64_t n);
int imc_surrogate_train(imc_ctx* ctx, int32_t epochs, int32_t batch, float lr, float* loss_per_epoch);
int imc_surrogate_set_params(imc_ctx* ctx, const float* params);
int imc_surrogate_get_params(imc_ctx* ctx, float* params);
int imc_surrogate_forward(imc_ctx* ctx, const float* features, int64_t n, float* outputs);
int imc_surrogate_gradient(imc_ctx* ctx, const float* features, const float* targets, int64_t n,
                           float* grad, float* loss);
int imc_surrogate_enable(imc_ctx* ctx, int32_t on);

/* ---- Analytic random walk ----------------------------------------------------------
 * The other filling of the Transport.MC_RW slot (imc_transport.jl:212-479, with its
 * tables P_r / randomwalk_table at :734-797), corrected: in a cell, a particle whose distance
 * R0 to the nearest face is at least r_min collision mean free paths (the reference: one
 * total mean free path) replaces its walk inside [x - R0, x + R0] by the first exit of
 * one-dimensional diffusion with D = c / (3 sigma_c), sigma_c = sigma_a (1 - f) + sigma_s
 * (the reference used a sphere law and dropped sigma_s).  The weight follows the exact path
 * length and every deposit is E_in - E_out, so energy is conserved (the reference's is not,
 * quirk Q6).  In 2D (directions in the plane, D = c / (2 sigma_c)) the walk is taken over the
 * disk of radius R0 = the distance to the nearest of the four faces, with the exit-time law of
 * planar diffusion from the disk's centre; the disk lies inside the cell.  Enabling it
 * disables the surrogate and the other way round. */
int imc_random_walk_enable(imc_ctx* ctx, int32_t on, double r_min);

/* ---- Planck frequency sampling -----------------------------------------------------
 * Sourcing.sample_planck (imc_sourcing.jl:372-399; unused by the reference's gray births,
 * which set nu = 1 and leave `T * sample_planck()` commented out at imc_sourcing.jl:171,186,
 * 209,232,342,362).  freq[i] = temp[i] * x_i with x_i drawn from the
 * normalised Planck spectrum (15 / pi^4) x^3 / (e^x - 1), by the same mixture of Gamma(4, 1/n)
 * laws with weights 90 / (pi^4 n^4); computed in FP64 on the GPU, every draw ends (the
 * reference's loop does not end for a first uniform above 90 sum(n^-4) / pi^4 with the sum
 * rounded to Float32: about one draw in 1e6).  Draw i is keyed by (the context's seed,
 * `stream`, i): the same triple gives the same frequency.  Host arrays of n values; n = 0 is a
 * no-op; every temp must be finite and >= 0 (IMC_EINVAL otherwise). */
int imc_sample_planck(imc_ctx* ctx, int64_t n, const double* temp, uint32_t stream, double* freq);

#ifdef __cplusplus
}
#endif
#endif
