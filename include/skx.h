/*
 * skx.h -- C ABI of the MI355X (gfx950) engine for the skyrmion_solver hot path.
 *
 * The engine replaces the per-step Numba-CUDA kernels of the reference
 * (Paulnleask/skyrmion_solver) behind its Python `Simulation` API.  Each entry
 * point below names the reference interface it replaces (file:line relative
 * to the reference's `skyrmion_solver/` package).
 *
 * Conventions
 *   - Return 0 (SKX_OK) on success, a negative SKX_E* code on failure;
 *     skx_last_error() then holds a message (thread-local).  Nothing aborts.
 *   - Host arrays are always FP64 in the reference layout
 *     idx = z + Z*(y + Y*x) + N*a  (core/utils.py:48); the engine converts to
 *     its storage precision.  Host pointers are borrowed for the call only.
 *   - A handle owns all of its device memory and one HIP stream.  A handle is
 *     not re-entrant; distinct handles may be driven from different threads.
 */
#ifndef SKX_H
#define SKX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SKX_VERSION 1

enum skx_status_code {
  SKX_OK = 0,
  SKX_EINVAL = -1,      /* invalid argument / parameter array */
  SKX_ENOMEM = -2,      /* device allocation failed */
  SKX_EHIP = -3,        /* HIP runtime error (message in skx_last_error) */
  SKX_ENONFINITE = -4,  /* a non-finite energy or gradient was produced */
  SKX_EBUSY = -5,       /* resource busy (admission refused) */
  SKX_ENOTSUP = -6      /* theory / option not supported by this build */
};

/* Theories: the plugins of theories/registry.py that the engine accelerates. */
enum skx_theory {
  SKX_THEORY_PION = 0,    /* theories/nuclear_skyrme  ("Nuclear Skyrme model") */
  SKX_THEORY_OMEGA = 1,   /* theories/omega_skyrme    ("Omega-Skyrme model")   */
  SKX_THEORY_COULOMB = 2, /* theories/coulomb_skyrme  ("Coulomb-Skyrme model") */
  SKX_THEORY_RHO = 3,     /* theories/rho_skyrme      ("Rho-Skyrme model")     */
  SKX_THEORY_BERGER = 4,  /* theories/berger_skyrme   ("Berger Skyrme model"): p_f[12] = alpha */
  SKX_THEORY_LIGHTLY_BOUND = 5 /* theories/lightly_bound_skyrme ("Lightly bound Skyrme model"): p_f[9] = alpha */
};

enum skx_precision {
  SKX_FP64 = 0, /* plain form: FP64 storage and arithmetic (reference arithmetic) */
  SKX_FP32 = 1  /* FP32 storage and stencils; FP64 reductions, RK4 combine and S^3 renormalisation */
};

enum skx_buffer { SKX_FIELD = 0, SKX_VELOCITY = 1 };

enum skx_density_kind {
  SKX_DENSITY_ENERGY = 0,    /* compute_energy_kernel (theories/nuclear_skyrme/kernels.py:349-382) */
  SKX_DENSITY_BARYON = 1,    /* compute_skyrmion_number_kernel (:449-482) */
  SKX_DENSITY_POTENTIAL = 2, /* compute_potential_kernel (theories/omega_skyrme/kernels.py:412-441) */
  SKX_DENSITY_RHO = 3        /* compute_rho_meson_kernel (theories/rho_skyrme/kernels.py:384-412) */
};

enum skx_flags {
  SKX_FLAG_NO_GRAPH = 1, /* launch stage kernels directly instead of replaying a captured hipGraph */
  SKX_FLAG_TRACE = 2,    /* time every stage kernel with a pair of HIP events on the handle's stream
                            (runs then launch directly, without graphs) */
  SKX_FLAG_ONE_LANE = 4  /* FP64 pion: RK stages 2-4 with one thread per site instead of two (the
                            default wherever N + Y*Z < 2^28 sites); results agree to rounding */
};

/* Lattice + theory description.  p_i / p_f are the reference's packed
 * parameter arrays verbatim (core/params.py:248-268 plus the theory's
 * pack_device_params, e.g. theories/nuclear_skyrme/params.py:203-245). */
typedef struct skx_desc {
  int32_t theory;       /* enum skx_theory */
  int32_t precision;    /* enum skx_precision */
  const int32_t* p_i;
  int32_t n_i;
  const double* p_f;
  int32_t n_f;
  int32_t device;       /* HIP device ordinal */
  uint32_t flags;       /* enum skx_flags */
  int32_t tile_z;       /* 0 = engine default; else 16, 32 or 64 */
  int32_t x_chunk;      /* 0 = engine default; planes per block along x */
} skx_desc;

typedef struct skx_engine* skx_handle;

/* Result of a device-resident run. */
typedef struct skx_run_stats {
  int32_t steps;        /* ANF steps taken */
  int32_t converged;    /* 1 if the last stage-4 max|g| < tol */
  double energy;        /* energy of the final field (what the last step() would return) */
  double err_signed;    /* reference `err`: signed max of the last stage-4 gradient */
  double err_abs;       /* max |g| of the last stage-4 gradient */
  double baryon;        /* baryon number of the final field */
  double seconds;       /* wall time of the run on the host clock */
  int32_t issued;       /* steps enqueued on the device: never more than max_steps; steps after the
                           stop test fired are enqueued but skip their work */
  int32_t graph_launches; /* captured chunks replayed (the remainder runs as direct launches) */
} skx_run_stats;

/* One record per step of a run (when a history buffer is attached). */
typedef struct skx_step_record {
  double energy;        /* energy of the field entering the step */
  double err_signed;
  double err_abs;
  double baryon;        /* baryon number of the field entering the step */
} skx_step_record;

/* Replaces Simulation.__init__ buffer allocation + set_params
 * (core/simulation.py:46-122, :175-197). */
int skx_create(const skx_desc* desc, skx_handle* out);
int skx_destroy(skx_handle h);

/* Bytes of device memory skx_create would allocate for this description
 * (used by admission control; no device work). */
int skx_footprint_bytes(const skx_desc* desc, size_t* bytes);

/* Replaces cuda.to_device / DeviceNDArray.copy_to_device of Simulation.Field /
 * .Velocity (core/simulation.py:94-95) and .copy_to_host (:335). n = nf*N. */
int skx_upload(skx_handle h, int32_t which, const double* host, size_t n);
int skx_download(skx_handle h, int32_t which, double* host, size_t n);

/* Replaces Simulation.step -> do_arrested_newton_flow
 * (core/simulation.py:287-308, core/integrator.py:284-410): one RK4 ANF step
 * with the reference's exact semantics; returns the new energy, the
 * reference's signed-max `err`, and the true max|g|.  Synchronous. */
int skx_anf_step(skx_handle h, double e_prev, double* e_new, double* err_signed, double* err_abs);

/* Device-resident loop of skx_anf_step calls (the loop the reference's
 * viewer drives, theories/nuclear_skyrme/render_gl.py:944-973, with a
 * convergence test): stops after max_steps or when the stage-4 max|g| < tol.
 * The arrest decision of each step is taken on the device in FP64.
 * check_every = steps between host polls (0 = engine default). */
int skx_anf_run(skx_handle h, int32_t max_steps, double tol, int32_t check_every, skx_run_stats* stats);

/* Capture the step graphs skx_anf_run replays for this check_every (one per
 * ping-pong parity), without running a step, so that a later timed or
 * latency-bound run does no capture.  No reference counterpart (the
 * reference launches every kernel from the host); for callers that time
 * or serve runs.  A no-op with SKX_FLAG_NO_GRAPH or SKX_FLAG_TRACE. */
int skx_anf_prepare(skx_handle h, int32_t check_every);

/* Plain RK4 steps (killkinen forced off), n steps. */
int skx_rk4_steps(skx_handle h, int32_t n, skx_run_stats* stats);

/* Replaces observables.compute_energy / compute_skyrmion_number
 * (theories/nuclear_skyrme/observables.py:114-180). */
int skx_energy(skx_handle h, double* energy);
int skx_baryon(skx_handle h, double* baryon);

/* Replaces Simulation.compute_*_density (core/simulation.py:219-285): writes
 * N doubles (density times cell volume for ENERGY / BARYON). */
int skx_density(skx_handle h, int32_t kind, double* host, size_t n);

/* Parity probe: evaluate one gradient stage (gradient_step_kernel,
 * core/integrator.py:37-46) on the given stage field `temp` without touching
 * the solver state.  eg = projected EnergyGradient, l = -dt*eg; nf*N each
 * (either may be NULL). */
int skx_eval_gradient(skx_handle h, const double* temp, double* eg, double* l, size_t n);

/* ---------------------------------------------------------------- observables
 * Post-convergence observables: one FP64 pass over the lattice evaluates the
 * per-site integrand the reference's observable kernel writes and returns its
 * lattice sums (the reference reduces the same per-site arrays,
 * theories/nuclear_skyrme/observables.py:182-486):
 *   COM         4: sum b, sum x b, sum y b, sum z b   compute_center_of_mass_kernel (kernels.py:678-718)
 *   RMS         1: sum |r|^2 |b|                      compute_rms_radius_kernel (:720-759)
 *   QUADRUPOLE  9: sum (3 r_i r_j - d_ij r^2) b / 2   compute_quadrupole_tensor_kernel (:761-812)
 *   U           9: sum 2 U0_ij vol                    compute_U_tensor_kernel (:814-851)
 *   V           9: sum 2 e_ilm e_jnp r_l r_n A_mp vol compute_V_tensor_kernel (:853-905)
 *   W           9: sum 2 e_jlm r_l B_im vol           compute_W_tensor_kernel (:907-975)
 *   D           1: sum (r_i r_j T_ij - r^2 T_kk / 3) vol  compute_D_integrand_kernel (:977-1033)
 * b is the baryon density times the cell volume, the grid is x = lsx * i
 * (create_grid_kernel, kernels.py:50-56) and r = x - com; tensors are
 * row-major (3 i + j).  `com` (3 doubles) is read by RMS, QUADRUPOLE, V, W
 * and D only.  COM and RMS exist for every theory (e.g.
 * theories/omega_skyrme/kernels.py:503,545); the others for the nuclear
 * theory only (SKX_EINVAL otherwise).  n must be at least the count above. */
enum skx_observable_kind {
  SKX_OBS_COM = 0,
  SKX_OBS_RMS = 1,
  SKX_OBS_QUADRUPOLE = 2,
  SKX_OBS_U = 3,
  SKX_OBS_V = 4,
  SKX_OBS_W = 5,
  SKX_OBS_D = 6
};
int skx_observable(skx_handle h, int32_t kind, const double* com, double* out, int32_t n);

/* skx_observable without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture set-up, tight serving loops, harnesses that forbid
 * hipMalloc).  skx_observable allocates its block partials on a handle's
 * first observable call; this variant keeps them in the library's static
 * per-device array instead (2048 blocks x 9 doubles = 147456 bytes on each
 * device, held until the process ends; a pass needing more is refused with
 * SKX_EINVAL).  Same kernel, same fold order: results equal skx_observable's
 * bit for bit.  Every call on one device shares that array, so calls that
 * overlap on one device (from several threads) race; serialise them. */
int skx_observable_noalloc(skx_handle h, int32_t kind, const double* com, double* out, int32_t n);

/* ---------------------------------------------------------------- initial data
 * Device-side initial configurations.  Replaces the host-built field of
 * Simulation.initialize -> theory initialize (core/simulation.py:124-173) with
 * the kernels of theories/nuclear_skyrme/initial_config.py:701-1126
 * (hedgehog_field, create_initial_configuration_kernel's rational-map branch,
 * product_ansatz_kernel) and, for the omega theory, create_omega_field_kernel
 * (theories/omega_skyrme/initial_config.py:1495-1532). */
enum skx_lump_kind { SKX_LUMP_HEDGEHOG = 0, SKX_LUMP_RATIONAL_MAP = 1 };
enum skx_profile_kind {
  SKX_PROFILE_EXP = 0,   /* f = pi exp(-r / a),      a = profile_param */
  SKX_PROFILE_GAUSS = 1, /* f = pi exp(-(r / a)^2),  a = profile_param */
  SKX_PROFILE_FIT = 2    /* profilefun(r, B, m0, 0) (initial_config.py:343-442), m0 = profile_param */
};
#define SKX_MAX_LUMPS 16

typedef struct skx_lump {
  int32_t kind;          /* skx_lump_kind */
  int32_t baryon_number; /* rational map R(z) for B = 1..8 (initial_config.py:446-549), z^B above */
  int32_t profile;       /* skx_profile_kind */
  int32_t pad_;
  double profile_param;
  double center[3];      /* lump centre relative to the lattice centre */
  double space_rot[9];   /* row-major S: position r -> S (r - center) (rational map only) */
  double iso_rot[9];     /* row-major R: unit isovector n -> R n */
} skx_lump;

typedef struct skx_init_desc {
  int32_t n_lumps;           /* 0: vacuum; 1: one lump; > 1: quaternion product of the lumps, in order,
                                renormalised after each product */
  int32_t vacuum_band;       /* write the vacuum into the halo-wide boundary band */
  double noise;              /* N(0, noise) added to each sigma-model component before renormalising
                                (counter-based generator keyed by noise_seed) */
  uint64_t noise_seed;
  double omega_cs;           /* omega theory: omega = -omega_cs * CS of the new pion field (0: omega = 0) */
  skx_lump lumps[SKX_MAX_LUMPS];
} skx_init_desc;

/* Set the field to the described configuration (components beyond the 4
 * sigma-model ones 0, or omega = -c CS) and the velocity to 0. */
int skx_init_field(skx_handle h, const skx_init_desc* desc);

/* Attach a device history of up to `capacity` skx_step_record entries for the
 * next runs (0 detaches); skx_history_read copies min(n, steps) records. */
int skx_history_attach(skx_handle h, int32_t capacity);
int skx_history_read(skx_handle h, skx_step_record* out, int32_t n, int32_t* n_read);

/* Stage-kernel timing (SKX_FLAG_TRACE): per stage kind 0..4 (0 = energy
 * pass, 1..4 = RK4 stages), accumulated milliseconds and launch counts.  For
 * the rho theory a stage's time covers its curl pass and stage kernel. */
int skx_trace_read(skx_handle h, double ms[5], int64_t counts[5]);
int skx_trace_reset(skx_handle h);

/* Launch shape the engine chose for stage kind 0..4 (no counterpart in the
 * reference; for the autotuner and diagnostics): shape[0] x shape[1] = the
 * (z, y) tile in sites, shape[2] = planes per block along x, shape[3] =
 * threads per site. */
int skx_stage_shape(skx_handle h, int32_t stage, int32_t shape[4]);

/* Wait for all work queued on the handle's stream. */
int skx_sync(skx_handle h);

/* ---------------------------------------------------------------- batches
 * Many lattices of one theory / precision minimised together: one launch
 * per RK4 stage covers every lattice (no reference counterpart; the
 * reference runs one Simulation at a time).  Lattices join and leave between
 * run calls; each keeps the exact per-lattice semantics of skx_anf_run. */
typedef struct skx_batch* skx_batch_handle;

typedef struct skx_batch_desc {
  int32_t theory;       /* enum skx_theory */
  int32_t precision;    /* enum skx_precision */
  int32_t device;
  uint32_t flags;       /* SKX_FLAG_NO_GRAPH */
  int32_t tile_z;       /* 16, 32 or 64: one tile for every stage; 0: a tile per stage as for one
                           lattice (FP32: 32 x 16 for the stages that run faster on it, 32 x 8
                           elsewhere) */
  int32_t max_slots;    /* lattices resident at once */
  int32_t max_blocks;   /* blocks of every batched launch (sum over lattices; with a tile
                           per stage, each of its two block maps holds this many) */
} skx_batch_desc;

typedef struct skx_slot_status {
  int32_t slot, active, steps, done, converged, nonfinite;
  double energy;        /* energy of the last field reduced (final field after skx_batch_remove) */
  double err_signed, err_abs, baryon;
} skx_slot_status;

int skx_batch_create(const skx_batch_desc* desc, skx_batch_handle* out);
int skx_batch_destroy(skx_batch_handle b);
/* Add a lattice (F0 / V0 host FP64; F0 NULL = the vacuum, V0 NULL = 0); it runs until its
 * stage-4 max|g| < tol or max_steps.  lattice->x_chunk = planes per block
 * along x (0: the whole lattice, one block per (y, z) tile); lattice->tile_z
 * is ignored (the batch's tile applies).  SKX_EBUSY when no slot / blocks are
 * free now; SKX_EINVAL when the lattice needs more blocks than max_blocks. */
int skx_batch_add(skx_batch_handle b, const skx_desc* lattice, const double* F0, const double* V0, int32_t max_steps,
                  double tol, int32_t* slot);
/* Advance every active lattice by `steps` ANF steps (rounded up to even).
 * skx_batch_run = skx_batch_launch (enqueue on the batch's stream, return at
 * once) + skx_batch_wait; batches on different streams overlap on the device. */
int skx_batch_run(skx_batch_handle b, int32_t steps);
int skx_batch_launch(skx_batch_handle b, int32_t steps);
int skx_batch_wait(skx_batch_handle b);
int skx_batch_poll(skx_batch_handle b, skx_slot_status* out, int32_t n, int32_t* n_out);
/* Finish a lattice (final energy, pending arrest) and copy its field and
 * velocity out (either may be NULL); frees its slot. */
int skx_batch_remove(skx_batch_handle b, int32_t slot, double* F_out, double* V_out, skx_slot_status* final_status);
/* skx_init_field on the lattice in `slot` (add it with F0 = NULL, the vacuum, then initialise it on the
 * device). */
int skx_batch_init(skx_batch_handle b, int32_t slot, const skx_init_desc* desc);

/* Diagnostic (no reference counterpart): time `reps` launches of a plain
 * streaming copy of `bytes` bytes on `device`; returns read+write GB/s and
 * milliseconds per copy.  Used as the measured copy peak and to calibrate
 * HBM byte counters against a kernel of known traffic. */
int skx_diag_stream_copy(int32_t device, size_t bytes, int32_t reps, double* gbs, double* ms_per_copy);

/* Device identity (no reference counterpart): architecture name (e.g.
 * "gfx950", NUL-terminated, truncated to n bytes) and compute-unit count.
 * Keys the launch-shape autotuner's cache. */
int skx_device_info(int32_t device, char* arch, int32_t n, int32_t* cu_count);

/* ---------------------------------------------------------------- text output
 * Host-side writer of the reference's .dat text format: output_field_dat
 * (core/io.py:52-92) when component_break = 1, output_density_data_dat
 * (:123-160) when ncomp = 1 and component_break = 0.  `data` holds ncomp
 * grids of X*Y*Z doubles in the reference layout; every value is printed as
 * Python's format(v, ".{precision}g") prints it, precision 0..100.  No GPU work. */
int skx_write_dat(const char* path, const double* data, int32_t X, int32_t Y, int32_t Z, int32_t ncomp,
                  int32_t precision, int32_t component_break);

const char* skx_last_error(void);
int skx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SKX_H */
