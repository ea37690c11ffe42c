/*
 * stophys_gpu.h -- C ABI of the MI355X (gfx950) ensemble SDE engine for the
 * StochasticPhysics particle simulator.
 *
 * This is the drop-in boundary that a GFX950 backend of the reference's
 * SimulationManager binds.  The reference has no C ABI of its own: its entry
 * point is the C++ template SimulationManager<prec, InstructionSet>
 * (Particle_Simulation/Simulator/SimulationManager.h:545-615), which builds one
 * SingleParticleSimulator per particle (SingleParticleSimulator.h:129-159) on a
 * host thread pool and merges the per-particle trajectories into a moment
 * weighted mean (Results/MeanSimulationResult.h:87-119).  Every entry point
 * below names the reference interface it replaces.
 *
 * Conventions: no exceptions cross this boundary; every function returns an
 * sp_status and sp_last_error() gives a thread-local message.  Host buffers
 * belong to the caller; device memory belongs to the library.
 */
#ifndef STOPHYS_GPU_H
#define STOPHYS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SP_ABI_VERSION 8

/* ---- status codes (the reference throws std::runtime_error instead,
 *      e.g. Settings/SolverSettings.h:139, Fields/LissajousField.h:56) ---- */
typedef enum sp_status {
    SP_OK = 0,
    SP_ERR_INVALID_ARG = 1,
    SP_ERR_UNSUPPORTED = 2,      /* problem/solver/anisotropy combination not available */
    SP_ERR_OUT_OF_MEMORY = 3,
    SP_ERR_DEVICE = 4,
    SP_ERR_NOT_CONVERGED = 5,    /* ImplicitMidpoint Newton hit Max_Iterations */
    SP_ERR_CANCELLED = 6,
    SP_ERR_TIMEOUT = 7,
    SP_ERR_NOT_FOUND = 8,
    SP_ERR_NOT_READY = 9
} sp_status;

/* ---- selector enums: numeric values equal the reference's enums ---- */
/* Settings::IProblem (Settings/ProblemSettings.h:55-70) */
enum { SP_PROBLEM_NEEL = 1, SP_PROBLEM_NEEL_SPHERICAL = 2, SP_PROBLEM_BROWN_AND_NEEL = 3,
       SP_PROBLEM_BROWN_AND_NEEL_EULER_SPHERICAL = 4 };
/* Settings::ISolver (Settings/SolverSettings.h:34-47) */
enum { SP_SOLVER_EULER_MARUYAMA = 1, SP_SOLVER_EULER_MARUYAMA_NORMALIZED = 2, SP_SOLVER_IMPLICIT_MIDPOINT = 3,
       SP_SOLVER_MILLSTEIN = 6, SP_SOLVER_HEUN_STRONG = 7,
       SP_SOLVER_HEUN_NOT_CONSISTENT = 8 /* two noise draws per step (SDEFramework/Solver/Heun_NotConsistent.inl:34-45);
                                            runs in FP64 whatever the precision asked for */ };
/* Properties::IField (Fields/FieldList.h:48-59) */
enum { SP_FIELD_ZERO = 1, SP_FIELD_CONSTANT = 2, SP_FIELD_SINUSOIDAL = 3, SP_FIELD_LISSAJOUS = 4,
       SP_FIELD_TRIANGULAR = 5, SP_FIELD_RECTANGULAR = 6, SP_FIELD_SINC = 7 };
/* Properties::IAnisotropy (Problems/Anisotropy/AnisotropyList.h:45-54); the reference's app dispatches
 * uniaxial, cubic and uniaxialcubic (Simulator/SimulationManager.h:454-476), not mixed (3) */
enum { SP_ANISO_UNIAXIAL = 1, SP_ANISO_CUBIC = 2, SP_ANISO_UNIAXIALCUBIC = 4 };

/* precision of the bulk arithmetic (no reference counterpart: PREC double,
 * General/GeneralDefines.h:3) */
enum { SP_PREC_FP64 = 0, SP_PREC_MIXED = 1 };
/* noise source: in-register Philox4x32-10 (replaces SDEFramework/NoiseField.inl:24-44),
 * or replay of caller-supplied standard normals (parity testing) */
enum { SP_NOISE_PHILOX = 0, SP_NOISE_REPLAY = 1 };

/* request flags */
enum {
    SP_FLAG_BN_SIMPLE_MODEL = 1u << 0,   /* BrownAndNeel Use_simple_model (ProblemSettings.h:126-130) */
    SP_FLAG_COORD_TRANSFORM = 1u << 1,   /* NeelSpherical Use_coordinate_transformation (ProblemSettings.h:192-205) */
    SP_FLAG_KEEP_FINAL_STATE = 1u << 2,  /* copy per-particle final state back */
    SP_FLAG_NO_FALLBACK = 1u << 3,       /* mixed precision: report the check error but never re-run in FP64 */
    SP_FLAG_SKIP_CHECK = 1u << 4,        /* mixed precision: no FP64 shadow check */
    SP_FLAG_RENORMALIZE = 1u << 5,       /* apply the problem's normalize() (NeelRelaxation.h:204-208,
                                            BrownAndNeelRelaxation.h:168-172) after every step of any
                                            solver, in FP64; EulerMaruyama_Normalized always does */
    SP_FLAG_EULER_TRANSFORM = 1u << 6    /* BrownAndNeelEulerSpherical Use_euler_coordinate_transformation
                                            (ProblemSettings.h:146-159); its Use_spherical_coordinate_transformation
                                            is SP_FLAG_COORD_TRANSFORM with min_angle */
};

/*
 * One simulation request: the content of one reference parameter file
 * (Settings/SimulationManagerSettings.h:34-150) for one ensemble whose
 * particles all share problem, solver, field and anisotropy type.
 */
typedef struct sp_request {
    int32_t problem;          /* SP_PROBLEM_* */
    int32_t solver;           /* SP_SOLVER_* */
    int32_t field;            /* SP_FIELD_* */
    int32_t anisotropy;       /* SP_ANISO_* */
    int32_t precision;        /* SP_PREC_* */
    int32_t noise;            /* SP_NOISE_* */
    uint32_t flags;           /* SP_FLAG_* */
    uint32_t newton_max_iter; /* Implicit_Midpoint Max_Iterations (SolverSettings.h:121) */
    uint64_t n_particles;     /* Number_of_simulations (SimulationSettings.h:67-79) */
    uint64_t n_steps;         /* Number_of_steps; (n_steps/oversampling - 1)*oversampling steps run */
    uint64_t oversampling;    /* Oversampling */
    uint64_t seed;            /* Philox key */
    double dt;                /* Timestep [s] */
    double newton_tol;        /* Implicit_Midpoint Accuracy_Goal (SolverSettings.h:122) */
    double min_angle;         /* NeelSpherical Min_angle_transformation [rad]; BrownAndNeelEulerSpherical
                                 Neel_Min_angle_before_transformation */
    double check_tol;         /* mixed precision: tolerated bias of <m>_mixed - <m>_fp64 (0: 1e-3; a negative
                                 value fails every check, which forces the FP64 re-run) */
    /* field parameters (Properties/Fields/*.hpp): Offset_Field, Amplitude for every field but zero;
     * sinusoidal: H = amplitude*sin(2 pi freq[0] t + phase[0]) + offset
     * lissajous:  H_i = offset_i + amplitude_i*sin(2 pi freq_i t + phase_i)
     * constant: H = offset; triangular, rectangular, sinc: see the period/time-offset/tau fields below */
    double field_offset[3];
    double field_amplitude[3];
    double field_freq[3];
    double field_phase[3];
    /* SP_NOISE_REPLAY only: standard normals z[k*n_particles + p], k = step*dim_W + j (k = step*2*dim_W + j for
     * Heun_NotConsistent: dW then dW2) */
    const double* replay_normals;
    uint64_t replay_count;    /* number of k values available */
    double field_period;      /* triangular, rectangular, sinc: Periode [s] */
    double field_time_offset; /* triangular, rectangular: TimeOffset [s] */
    double field_tau;         /* rectangular: Tau [s], time constant of the edges */
    uint32_t field_alternating; /* rectangular: Alternating (0/1) */
    uint32_t save_interval;   /* Save_single_simulations with Save_intervall (Settings/ResultSettings.h:93-94,
                                 Results/SimulationResultManager.h:115-149): 0 = off; k > 0 keeps the window
                                 means of every particle p with (p + 1) % k == 0 (the reference saves every k-th
                                 finished simulation, in completion order) */
    double brown_min_angle;   /* BrownAndNeelEulerSpherical Brown_Min_angle_before_transformation [rad] */
} sp_request;

/*
 * Per-particle properties, structure of arrays, n_particles entries each
 * (Properties/MagneticProperties.h:100-267, HydrodynamicProperties.h:31-83,
 * ParticleProperties.h:32-127).  The library evaluates the reference's
 * parameter calculators (Problems/Helpers/ParameterCalculator*.h) from them.
 */
typedef struct sp_particles {
    const double* magnetic_radius;           /* [n] m */
    const double* saturation_magnetisation;  /* [n] A/m */
    const double* damping;                   /* [n] */
    const double* gyromagnetic_ratio;        /* [n] rad/(s T) */
    const double* temperature;               /* [n] K */
    const double* anisotropy_constant;       /* [n] K_uniaxial (uniaxial, uniaxialcubic) or K_cubic (cubic), J/m^3 */
    const double* hydrodynamic_radius;       /* [n] m (BrownAndNeel only, else NULL) */
    const double* viscosity;                 /* [n] Pa s (BrownAndNeel only, else NULL) */
    const double* axes;                      /* [n][9] body axes x,y,z (Neel/NeelSpherical; NULL = identity) */
    const double* initial_state;             /* [n][dim_y]: m (Neel), (theta,phi) (NeelSpherical), (n,m) (BrownAndNeel),
                                                (phi, theta, psi, theta_N, phi_N) (BrownAndNeelEulerSpherical) */
    /* uniaxialcubic only (Properties/Anisotropy/UniaxialCubic.hpp:20-51), else NULL */
    const double* anisotropy_constant_cubic; /* [n] K_cubic of the cubic part, J/m^3 */
    const double* uniaxial_rotation;         /* [n][3] Euler angles of the uniaxial part (Uniaxial_Rotation): its easy
                                                axis is R^T z, R the rotation of UniaxialCubicAnisotropy.hpp:61-90, :114 */
} sp_particles;

/*
 * Output of one request; the caller allocates the arrays (sizes from
 * sp_request_shape).  Mirrors Mean_Results {Time, Field, Results, Weight}
 * (Results/MeanSimulationResult.h:131-141).
 */
typedef struct sp_result {
    double* time;         /* [n_out] window centre times (SingleParticleSimulator.h:236-273) */
    double* field;        /* [n_out*3] window mean field */
    double* mean;         /* [n_out*dim_out] weighted ensemble mean */
    double* weight;       /* [dim_out] sum of weights */
    double* final_state;  /* [n_particles*dim_y] or NULL */
    double* trajectories; /* save_interval > 0: [n_out][dim_out][n_saved] unweighted window means of the saved
                             particles (SingleSimulationResult "Results", Results/SingleSimulationResult.h:96-104),
                             n_saved = n_particles / save_interval; or NULL */
    /* filled by the library */
    uint64_t n_saved;     /* particles whose trajectories are kept */
    uint64_t n_out;
    uint32_t dim_out;
    uint32_t dim_y;
    double device_ms;          /* device time spent on this request's kernels (batch share) */
    double latency_ms;         /* submit -> result ready */
    double particle_steps;     /* particles * steps executed */
    double check_error;        /* mixed: max over outputs of |<m>_mixed - <m>_FP64| on the check subset, the
                                  request's first 1/8 (512..4096 particles) re-run in FP64 (-1 if not checked) */
    double check_noise;        /* mixed: max over outputs of 3 standard errors of that difference, from its
                                  spread over the subset's 64-particle tiles.  The check fails, and the request
                                  is re-run in FP64, where the time-averaged difference exceeds check_tol by
                                  more than its 99.9% margin, or where the mixed result is not finite */
    uint32_t used_fallback;    /* 1 if the mixed result failed the check and was recomputed in FP64 */
    uint32_t precision_used;   /* SP_PREC_* of the returned result */
    uint64_t newton_failures;  /* ImplicitMidpoint steps that hit Max_Iterations */
    double particle_means[8];  /* means over the ensemble's particles of Temperature, Magnetic_radius,
                                  Saturation_magnetisation, Damping_constant, Gyromagnetic_ratio,
                                  Anisotropy_constant, Hydrodynamic_Radius, Viscosity: the Mean_Properties of
                                  Mean_Results (MeanSimulationResult.h:131-141) */
    double check_bias;         /* mixed: the figure the check accepts on, max over components of the time-averaged
                                  (windows >= 1) difference |<m>_mixed - <m>_FP64| on the check subset (-1 if not checked) */
    double check_bias_se;      /* mixed: the largest standard error of that average, from the spread of the subset's
                                  64-particle tiles; the check fails where |bias| - t_0.999 * se > check_tol */
} sp_result;

/* execution knobs chosen by the autotuner (or by hand) */
typedef struct sp_tuning {
    uint32_t block_size;       /* work-group size: 64, 128, 192 or 256 (a launch of fewer 64-particle tiles than the
                                  device has SIMDs uses 64, one wave per work-group) */
    uint32_t chunk_steps;      /* steps advanced per request per scheduler iteration / launch */
    uint32_t use_graph;        /* capture chunk launches in a hipGraph (static batches) */
    uint32_t persistent_waves_per_simd; /* persistent grid = CUs * 4 * this (1..8) */
} sp_tuning;

typedef struct sp_ctx sp_ctx;

/* ---- errors / info ---- */
const char* sp_last_error(void);
int sp_abi_version(void);

/* ---- selector strings (reference maps: SolverSettings.cpp:10-20, ProblemSettings.h:63-70,
 *      FieldList.h:51-59, AnisotropyList.h:48-54); CPU only ---- */
int sp_parse_enum(const char* kind, const char* name, int32_t* value);
const char* sp_enum_name(const char* kind, int32_t value);

/* ---- shapes; CPU only ---- */
int sp_request_shape(const sp_request* req, uint64_t* n_out, uint32_t* dim_y,
                     uint32_t* dim_w, uint32_t* dim_out);

/* Parameter calculators of the reference (ParameterCalculatorNeel.h:44-92,
 * ParameterCalculatorBrown.h:39-58, ParameterCalculatorBrownAndNeel.h:67-76,
 * UniaxialAnisotropy.h:71-85, CubicAnisotropy.h:92-94).  Writes the SoA
 * parameter block the kernels load, param-major: out[k*n + p].  CPU only. */
int sp_compute_particle_params(const sp_request* req, const sp_particles* parts,
                               double* out_params, uint32_t* n_params);

/* ---- context: one per GPU (replaces the ThreadManager of SimulationManager.h:109,546) ---- */
int sp_ctx_create(int device, sp_ctx** out);
void sp_ctx_destroy(sp_ctx* ctx);
int sp_ctx_set_tuning(sp_ctx* ctx, int problem, int solver, int precision, const sp_tuning* t);
int sp_ctx_get_tuning(sp_ctx* ctx, int problem, int solver, int precision, sp_tuning* t);
/* device memory budget for admitted requests (state + params + outputs; default half of the free device memory)
 * and the most requests admitted at once (default 512) */
int sp_ctx_set_memory_budget(sp_ctx* ctx, uint64_t bytes, uint32_t max_active);

/* plain path (sp_run, sp_execute): the largest partial-sum buffer used to reduce a whole pass at once
 * (default 4 GiB, and at most a quarter of the free device memory); a pass whose per-window partial sums
 * exceed it, or whose buffer cannot be allocated, is reduced chunk by chunk in a buffer of one chunk's rows.
 * Results do not depend on it.  No reference counterpart (the reference merges per-particle trajectories
 * under a mutex, SimulationManager.h:170-175). */
int sp_ctx_set_reduction_limit(sp_ctx* ctx, uint64_t bytes);
/* server (sp_submit): each kernel class's steps per scheduler iteration are cut, from its tuned chunk_steps, so
 * that its launch takes about `ms` by the class's measured cost per particle-step: requests of cheap classes then
 * finish without waiting on long launches of expensive ones (Implicit_Midpoint requests iterate on a lane of their
 * own in any case).  0 keeps the tuned chunk; default 5. */
int sp_ctx_set_iteration_budget(sp_ctx* ctx, double ms);

/* statistics of the last sp_run on this context: [0] kernel launches, [1] kernel ms,
 * [2] reduction + copy ms, [3] wall ms, [4] particle-steps, [5] VGPRs of the kernel,
 * [6] grid size (work-groups), [7] work-group size */
int sp_ctx_last_run_stats(sp_ctx* ctx, double* out /* [8] */);

/* ---- plain path: one request, synchronous.  Replaces SimulationManager::
 *      StartSimulationManager + waitUntilFinsihed for one ensemble
 *      (SimulationManager.h:549-584) ---- */
int sp_run(sp_ctx* ctx, const sp_request* req, const sp_particles* parts, sp_result* res);

/* ---- plain path with the inputs resident in HBM: sp_prepare uploads parameters and
 *      initial state once; every sp_execute restarts from that state on the device ---- */
typedef struct sp_prepared sp_prepared;
int sp_prepare(sp_ctx* ctx, const sp_request* req, const sp_particles* parts, sp_prepared** out);
int sp_execute(sp_prepared* h, sp_result* res);
void sp_release(sp_prepared* h);

/* ---- allocation-free execution.  No counterpart in the reference: for callers that cannot allocate
 *      device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Runs one request like sp_execute, asynchronously on `stream` (a hipStream_t; NULL = the null stream), on
 * caller-owned device buffers: d_params [NPARAM = 23][n_particles] (param-major, the rows of
 * sp_compute_particle_params), d_state [dim_y][n_particles] (particle index fastest; advanced in place to the
 * final state), d_mean [n_out][dim_out] (the weighted ensemble means, equal to sp_run's bit for bit).  It
 * allocates nothing: the request descriptor, the partial sums of the output windows and the weight sums live
 * in a 32 MiB device array that each GPU holds from the first call on it until the process ends.  A request
 * whose partial sums for one chunk, (lcm(Oversampling, 4) / Oversampling + 1) * dim_out * ceil(n_particles / 64)
 * * 8 bytes, exceed 32 MiB - 4 KiB is rejected with SP_ERR_INVALID_ARG (about 22 million BrownAndNeel particles
 * at Oversampling 50).  Calls that overlap on one GPU share that array and race.  Philox noise only, no
 * single-simulation outputs; mixed precision runs without its FP64 check; ImplicitMidpoint's Newton failures
 * are not reported. */
int sp_execute_noalloc(int device, void* stream, const sp_request* req, const double* d_params,
                       double* d_state, double* d_mean);

/* ---- batched path: many uneven requests in one persistent launch per chunk ---- */
int sp_run_batch(sp_ctx* ctx, uint32_t n_req, const sp_request* reqs,
                 const sp_particles* parts, sp_result* res);

/* ---- serving: continuous batching request loop.  sp_submit copies the inputs;
 *      sp_wait blocks; sp_result copies the outputs and releases the request.  A worker thread runs the
 *      requests in two lanes of iterations that overlap on the device: Implicit_Midpoint requests in one,
 *      every other request in the other ---- */
int sp_submit(sp_ctx* ctx, const sp_request* req, const sp_particles* parts, uint64_t* id);
int sp_wait(sp_ctx* ctx, uint64_t id, int timeout_ms);
int sp_result_get(sp_ctx* ctx, uint64_t id, sp_result* res);
int sp_cancel(sp_ctx* ctx, uint64_t id);   /* replaces SimulationManager::abort (SimulationManager.h:595-599) */
/* fraction of a submitted request's steps done, updated after every chunk (1 once its result is ready); replaces
 * the per-finished-simulation count of SimulationManager::finishSimulationTask (SimulationManager.h:177-214), which a
 * GPU ensemble that advances all its particles together cannot give */
int sp_progress(sp_ctx* ctx, uint64_t id, double* fraction);
int sp_server_stats(sp_ctx* ctx, double* out /* [8]: completed, active, pending, iterations (both lanes),
                                                 busy_ms (device time of the iterations, summed over the
                                                 lanes, so above wall_ms while they overlap), wall_ms,
                                                 particle_steps, fallbacks */);

/* ---- parameter files: the reference's .ini schema (Settings/SimulationManagerSettings.h:34-186,
 *      Provider/ParticleProvider.h:230-295) read into one request; CPU only ---- */
typedef struct sp_settings sp_settings;
int sp_settings_load(const char* ini_path, sp_settings** out);
/* fills req and, when parts is not NULL, points parts at per-particle arrays owned by the settings object
 * (materialised on the host at the first such call) */
int sp_settings_request(const sp_settings* s, sp_request* req, sp_particles* parts);
const char* sp_settings_savefile(const sp_settings* s);
/* writes Mean_Results {Time, Field, Results, Weight, Mean_Properties} as JSON
 * (replaces the SerAr MAT/HDF5 writer, Results/SimulationResultManager.h:97-113) */
int sp_settings_write_result(const sp_settings* s, const sp_result* res, const char* path);
/* Single_simulations_savefile of [Result_Settings] ("" when unset) */
const char* sp_settings_single_savefile(const sp_settings* s);
/* writes the single-simulation results of a request run with save_interval > 0 as JSON, one named entry
 * per kept particle {UsedProperties, Time, Weight, Field, Results} (replaces the SerAr writer of
 * SimulationResultManager::addSingleResult, Results/SimulationResultManager.h:115-149) */
int sp_settings_write_single_results(const sp_settings* s, const sp_result* res, const char* path);
void sp_settings_free(sp_settings* s);

/* ---- device-side particle provider: the particles of a parameter file drawn on the GPU ----
 * Replace ParticleProvider::getProvidedObject and ParticleSimulationSettings::applySettingsToParticleProperties
 * (Provider/ParticleProvider.h:230-295, Settings/ParticleSimulationSettings.h:51-152) with the start states and
 * parameter calculators after them, for ensembles too large to materialise on the host: one kernel writes
 * every particle's parameters and initial state straight into HBM when the request's buffers are allocated.
 * The particles are those sp_settings_request materialises on the host (same type selection, same Philox
 * draws; values equal up to the last-bit rounding of exp, sin, cos, acos and atan2), and no per-particle array
 * is built on the host.  sp_settings_prepare gives an sp_prepared for sp_execute; sp_settings_submit queues the
 * ensemble on the request server like sp_submit. */
int sp_settings_prepare(sp_ctx* ctx, const sp_settings* s, sp_prepared** out);
int sp_settings_submit(sp_ctx* ctx, const sp_settings* s, uint64_t* id);
/* copies a prepared request's resident inputs to the host: params [NPARAM = 23][n_particles] (param-major,
 * the rows of sp_compute_particle_params), initial_state [n_particles][dim_y]; either may be NULL.  No
 * counterpart in the reference (its particles live in its Problem objects): it lets a caller inspect what the
 * device provider drew */
int sp_prepared_inputs(const sp_prepared* h, double* params, double* initial_state);

/* ---- autotuner: times candidate tunings (work-group size, persistent waves per SIMD, chunk length) on a shortened
 * copy of the request, the four fastest again on a copy up to four times longer, and stores the fastest for the
 * request's class; best_rate = its device rate in particle-steps/s (launches only) ---- */
int sp_autotune(sp_ctx* ctx, const sp_request* req, const sp_particles* parts,
                sp_tuning* best, double* best_rate);

/* ---- device-side test hooks (parity of the noise stream) ---- */
/* raw Philox4x32-10 blocks: out[4*i..] = philox(ctr = (ctr0 + i, ctr1, ctr2, ctr3), key) */
int sp_dev_philox(sp_ctx* ctx, const uint32_t ctr[4], const uint32_t key[2], uint64_t count, uint32_t* out);
/* standard normals z_k, k in [first, first+count), of particle p's stream under seed */
int sp_dev_normals(sp_ctx* ctx, uint64_t seed, uint64_t particle, uint64_t first,
                   uint64_t count, int precision, double* out);
/* host copy of the same generator (the library's own host path) */
int sp_host_normals(uint64_t seed, uint64_t particle, uint64_t first, uint64_t count, double* out);
/* Box-Muller transform of caller-given Philox words: out[4i..4i+3] from words[4i..4i+3], in the
 * FP64 (bit-identical to sp_host_box_muller) or the mixed-precision (FP32 hardware) form */
int sp_dev_box_muller(sp_ctx* ctx, const uint32_t* words, uint64_t n_blocks, int precision, double* out);
int sp_host_box_muller(const uint32_t* words, uint64_t n_blocks, double* out);
/* the request server's next n_iterations scheduler iterations fail on the host, before their launches, with
 * SP_ERR_DEVICE: exercises the error path of sp_submit/sp_wait/sp_result_get without a device fault */
int sp_ctx_inject_failures(sp_ctx* ctx, uint32_t n_iterations);

#ifdef __cplusplus
}
#endif
#endif /* STOPHYS_GPU_H */
