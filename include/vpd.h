/*
 * vpd.h — C ABI of the MI355X (gfx950) explicit correspondence-PD increment with the
 * Johnson-Cook VUMAT. Plain C: opaque handles, plain pointers and sizes, status codes.
 *
 * One vpd_context owns one particle block on one GPU and one HIP stream; calls on one
 * context are not reentrant (as the reference's objects are not thread-safe for
 * concurrent callers, docs/architecture.md:28). Every call returns VPD_OK (0) or an error
 * status; vpd_last_error() gives the message, worded as the reference's exceptions are,
 * so a C++ wrapper can rethrow std::runtime_error(message) unchanged.
 *
 * Host arrays crossing this boundary use the reference's in-memory layouts:
 *   points     double[n][3]   (std::vector<Vec3>,    include/PdTypes.hpp:9)
 *   tensors    double[n][9]   row-major 3x3 (std::vector<Tensor3>, include/PdTypes.hpp:10)
 *   neighbours CSR: int32 rowptr[n+1], int32 col[nnz], rows ascending
 *              (std::vector<std::vector<int>>, include/PdTypes.hpp:18)
 *   VUMAT      column-major blocks [k + nblock*c] (src/material/VumatMaterialAdapter.cpp:77-80)
 * On the device every particle field is SoA (one plane per component).
 */
#ifndef VPD_H
#define VPD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VPD_ABI_VERSION 3

enum vpd_status {
  VPD_OK = 0,
  VPD_ERR_ARGUMENT = 1,  /* invalid argument / size mismatch (std::runtime_error in the reference) */
  VPD_ERR_NUMERIC = 2,   /* singular tensor, bad determinant, non-finite coordinate */
  VPD_ERR_STATE = 3,     /* call out of order (e.g. step before initialisation) */
  VPD_ERR_DEVICE = 4,    /* HIP runtime failure or no device */
  VPD_ERR_UNSUPPORTED = 5
};

/* Return-map solver of the Johnson-Cook stage. */
enum vpd_return_map {
  VPD_RETURN_MAP_BISECTION = 0,       /* exact bracketed bisection of JohnsonCook.f90:114-136 */
  VPD_RETURN_MAP_BISECTION_TIGHT = 1, /* the same scheme to |r| <= rtol·q_trial */
  VPD_RETURN_MAP_NEWTON = 2,          /* safeguarded Newton seeded with state 5 */
  VPD_RETURN_MAP_LEARNED = 3,         /* safeguarded Newton seeded by state 5 where it is
                                         positive and by the warm-start network elsewhere */
  /* vpd_return_map_device only: mode 0 with the FP64 yield evaluated at every iteration (modes
   * 0 and 1 evaluate it only where an FP32 estimate with an error bound cannot decide the
   * iteration's branch, with the same iterates and results bit for bit); for checking. */
  VPD_RETURN_MAP_BISECTION_UNSCREENED = 4,
  /* vpd_return_map_device only: mode 0 on lane groups (32 lanes per point, each round
   * evaluating every iterate of the next five iterations at once), the form vpd_step uses
   * for short plastic lists of blocks of up to 32,768 particles; the same bits as mode 0. */
  VPD_RETURN_MAP_BISECTION_GROUPED = 5
};

/* The VUMAT the material stage runs (the plugin of VumatMaterialConfig::pluginPath). */
enum vpd_material_model {
  VPD_MATERIAL_JOHNSON_COOK = 0, /* materials/examples/JohnsonCook.f90: 14 props, >= 6 states */
  VPD_MATERIAL_ELASTIC = 1       /* materials/examples/Elastic.f: props E, nu; state untouched */
};

enum vpd_damage_mode {
  VPD_DAMAGE_NONE = 0,
  VPD_DAMAGE_CRITICAL_STRETCH = 1 /* src/damage/PdSimpleDamage.cpp:63-84 */
};

/* Constraint code per particle and axis (BoundaryConditions.cpp:81-115): bits 2d..2d+1
 * of codes[i] for axis d: 0 free, 1 held at zero (u = v = a = 0), 2 follows the ramp. */
enum vpd_constraint { VPD_FREE = 0, VPD_ZERO = 1, VPD_RAMP = 2 };

/* Field identifiers for vpd_upload / vpd_download (host layouts above). */
enum vpd_field {
  VPD_FIELD_REFERENCE_POSITION = 0, /* n×3 (read-only after vpd_set_particles) */
  VPD_FIELD_VOLUME = 1,             /* n */
  VPD_FIELD_DISPLACEMENT = 2,       /* n×3 ParticleBlock::displacement */
  VPD_FIELD_VELOCITY = 3,           /* n×3 ExplicitDynamicState::velocity */
  VPD_FIELD_ACCELERATION = 4,       /* n×3 ExplicitDynamicState::acceleration */
  VPD_FIELD_DEFGRAD_N = 5,          /* n×9 MaterialState::deformationGradientN */
  VPD_FIELD_LEFT_STRETCH_N = 6,     /* n×9 MaterialState::leftStretchN */
  VPD_FIELD_ROTATION_N = 7,         /* n×9 MaterialState::rotationN */
  VPD_FIELD_STRESS_N = 8,           /* n×9 MaterialState::rawStressN (corotational, symmetric) */
  VPD_FIELD_STATEV = 9,             /* n×nstatev MaterialState::vumatStateN */
  VPD_FIELD_TEMPERATURE = 10,       /* n */
  VPD_FIELD_INTERNAL_ENERGY = 11,   /* n */
  VPD_FIELD_INELASTIC_ENERGY = 12,  /* n */
  VPD_FIELD_STRESS = 13,            /* n×9 MaterialUpdateResult::rawStress (spatial; download) */
  VPD_FIELD_STRAIN = 14,            /* n×9 MaterialUpdateResult::strain (Green-Lagrange; download) */
  VPD_FIELD_FORCE_DENSITY = 15,     /* n×3 computeForceDensity output (download) */
  VPD_FIELD_BOND_DAMAGE = 16,       /* nnz doubles 0/1 PdDamageState::bondDamageN */
  VPD_FIELD_DAMAGE_INDEX = 17,      /* n PdDamageResult::damageIndex (download) */
  VPD_FIELD_SHAPE_INVERSE = 18,     /* n×9 CorrespondenceState::shapeTensorInverse */
  VPD_FIELD_DEFGRAD = 19,           /* n×9 computeDeformationGradients output */
  VPD_FIELD_RETURN_MAP_ITERATIONS = 20, /* n int32, iterations of the last return map */
  VPD_FIELD_STRAIN_INCREMENT = 21,      /* n×6 Voigt strainInc = D̄·dt of the last update */
  VPD_FIELD_RETURN_MAP_FALLBACKS = 22   /* n int32, Newton steps replaced by bisection */
};

typedef struct vpd_config {
  /* PdCorrespondenceSkeleton(horizon), include/PdCorrespondenceSkeleton.hpp:12 */
  double horizon;
  /* ExplicitTimeIntegrator(density), include/ExplicitTimeIntegrator.hpp:17 (N·ms²/mm⁴) */
  double density;
  /* VumatPdCorrespondenceConfig, include/VumatPdCorrespondenceMaterial.hpp:13-19, with the
   * Johnson-Cook props of materials/examples/JohnsonCook.f90:9 (14 values). */
  int nprops;
  const double *props;
  int nstatev;                 /* >= 6 */
  const double *initial_state; /* nstatev values, or NULL for zeros */
  double reference_temperature;
  double char_length;
  double material_density; /* VumatMaterialConfig::density (kg/m³, passed to the VUMAT) */
  /* SimpleDamageConfig, include/SimpleDamageModel.hpp:14-23 (critical_stretch only) */
  int damage_mode;
  double damage_initiation;
  double damage_failure;
  int return_map; /* enum vpd_return_map (Johnson-Cook only) */
  int material_model; /* enum vpd_material_model; 0 (zero-initialised) is Johnson-Cook */
} vpd_config;

typedef struct vpd_context vpd_context;

/* Per-kernel timing collected with HIP events on the context stream. */
typedef struct vpd_kernel_time {
  const char *name;
  double total_ms;
  int64_t launches;
} vpd_kernel_time;

typedef struct vpd_step_args {
  double dt;
  double time;       /* step·dt: constraints are applied at this time */
  double step_time;  /* VUMAT stepTime = totalTime = (step−1)·dt (BoundaryValueCaseRunner.cpp:218-220) */
  double ramp[3];    /* loaded-band (u, v, a) at `time` (RegionalConstraintSystem::ramp) */
  int fuse_next;     /* 1: also predict the next step and constrain it with ramp_next */
  double ramp_next[3];
  int want_fields;   /* 1: keep spatial stress, strain and force density for download */
} vpd_step_args;

/* ---- lifetime ------------------------------------------------------------------- */
int vpd_abi_version(void);
int vpd_device_count(int *count);
/* Replaces the constructors of PdCorrespondenceSkeleton (src/pd/PdCorrespondenceSkeleton.cpp:144-147),
 * VumatPdCorrespondenceMaterial (src/material/VumatPdCorrespondenceMaterial.cpp:33-46),
 * PdSimpleDamage (src/damage/PdSimpleDamage.cpp:39-40) and ExplicitTimeIntegrator
 * (src/pd/ExplicitTimeIntegrator.cpp:9-12), with their argument checks. */
int vpd_create(const vpd_config *config, vpd_context **out);
void vpd_destroy(vpd_context *ctx);
const char *vpd_last_error(const vpd_context *ctx); /* ctx may be NULL: last vpd_create error */

/* ---- the particle block (include/PdTypes.hpp:14-19) ----------------------------- */
int vpd_set_particles(vpd_context *ctx, const double *reference_position, const double *volume,
                      int64_t n);
/* buildNeighborList (src/pd/PdCorrespondenceSkeleton.cpp:95-129) on the GPU. */
int vpd_build_neighbors(vpd_context *ctx, int64_t *nnz);
/* ParticleBlock::neighbors supplied by the caller (validated: 0 <= j < n). */
int vpd_set_neighbors(vpd_context *ctx, const int32_t *rowptr, const int32_t *col, int64_t nnz);
int vpd_get_neighbors(vpd_context *ctx, int32_t *rowptr, int32_t *col);
int vpd_set_constraints(vpd_context *ctx, const uint8_t *codes); /* n packed codes, or NULL */

/* ---- initialisation -------------------------------------------------------------- */
/* PdCorrespondenceSkeleton::initialize (src/pd/PdCorrespondenceSkeleton.cpp:149-170). */
int vpd_initialize_shape(vpd_context *ctx);
/* VumatPdCorrespondenceMaterial::initialize (src/material/VumatPdCorrespondenceMaterial.cpp:48-70)
 * with F₀ = computeDeformationGradients at the current displacement. */
int vpd_initialize_material(vpd_context *ctx);
/* ExplicitTimeIntegrator::initialize (src/pd/ExplicitTimeIntegrator.cpp:14-21) and
 * PdSimpleDamage::initialize (src/damage/PdSimpleDamage.cpp:42-49). */
int vpd_initialize_dynamics(vpd_context *ctx);

/* ---- state transfer ------------------------------------------------------------- */
int vpd_upload(vpd_context *ctx, int field, const void *host);
int vpd_download(vpd_context *ctx, int field, void *host);

/* ---- the increment, piece by piece (for per-function parity) --------------------- */
/* ExplicitTimeIntegrator::predictDisplacement (src/pd/ExplicitTimeIntegrator.cpp:23-34)
 * followed by RegionalConstraintSystem::apply (src/host/BoundaryConditions.cpp:81-115). */
int vpd_predict(vpd_context *ctx, double dt, const double ramp[3]);
int vpd_apply_constraints(vpd_context *ctx, const double ramp[3]);
/* PdCorrespondenceSkeleton::computeDeformationGradients (src/pd/PdCorrespondenceSkeleton.cpp:172-197);
 * with damage on, also PdSimpleDamage::update's bond loop (src/damage/PdSimpleDamage.cpp:63-84)
 * on the same positions. */
int vpd_compute_deformation_gradient(vpd_context *ctx);
/* VumatPdCorrespondenceMaterial::update (src/material/VumatPdCorrespondenceMaterial.cpp:72-176)
 * on the last computed F, plus cauchyToFirstPiola·K⁻¹ (src/pd/PdCorrespondenceSkeleton.cpp:209-210). */
int vpd_update_material(vpd_context *ctx, double dt, double step_time, double total_time);
/* computeForceDensity's gather (src/pd/PdCorrespondenceSkeleton.cpp:213-227) and
 * ExplicitTimeIntegrator::correctVelocity (src/pd/ExplicitTimeIntegrator.cpp:36-52), then the
 * constraints at `ramp`. */
int vpd_compute_force_and_correct(vpd_context *ctx, double dt, const double ramp[3]);

/* PdCorrespondenceSkeleton::computeForceDensity (src/pd/PdCorrespondenceSkeleton.cpp:199-230)
 * for a caller-supplied spatial Cauchy stress (n×9), with the context's last F (upload
 * VPD_FIELD_DEFGRAD or compute it), K⁻¹ and bond damage. Writes force density (n×3). */
int vpd_compute_force_density(vpd_context *ctx, const double *cauchy, double *force);

/* ExplicitTimeIntegrator::correctVelocity (src/pd/ExplicitTimeIntegrator.cpp:36-52) for a
 * caller-supplied force density (n×3). */
int vpd_correct_velocity(vpd_context *ctx, double dt, const double *force);
/* PdSimpleDamage::update, critical_stretch bond loop (src/damage/PdSimpleDamage.cpp:63-84), on
 * the context's current displacement; no-op for VPD_DAMAGE_NONE. */
int vpd_update_damage(vpd_context *ctx);

/* ---- per-point kinematics on host arrays (computed on the GPU) --------------------- */
/* updateCorotationalKinematics (src/pd/CorotationalKinematics.cpp:61-135) for n points:
 * velocity gradient, unrotated rate of deformation, relative spin increment, R_{n+1},
 * V_{n+1}; all n×9. */
int vpd_corotational_kinematics(const double *fN, const double *fNP1, const double *vN,
                                const double *rN, double dt, int64_t n, double *L, double *dbar,
                                double *rel_spin, double *r1, double *v1);
/* rotateCorotationalStressToSpatial (:137-139) and greenLagrangeStrain (:141-149) for n
 * points; either pair of pointers may be NULL. */
int vpd_rotate_and_strain(const double *stress, const double *rotation, const double *defgrad,
                          int64_t n, double *spatial_stress, double *strain);

/* ---- the fused increment ---------------------------------------------------------- */
/* One step of the loop body in src/host/BoundaryValueCaseRunner.cpp:207-229 (plus its
 * diagnostics when enabled, see vpd_enable_diagnostics): [predict + constrain unless the previous step fused it] → F (+ damage) →
 * material → force + correct + constrain [+ next predict + constrain]. Asynchronous on the
 * context stream; numeric errors surface at the next vpd_synchronize / download. */
int vpd_step(vpd_context *ctx, const vpd_step_args *args);
int vpd_synchronize(vpd_context *ctx);

/* ---- batched specimens ------------------------------------------------------------ */
/* Many velocity-driven specimens of uneven size in one particle block, stepped by one
 * launch per pass. This replaces the process-level case sweep of
 * benchmarks/tensile3d/run_cases.py:95-96 (a thread pool of single-case runs) for specimens
 * that share the material and damage settings and differ in lattice, horizon, time step,
 * end velocity and step count.
 * Specimen s owns particles [first[s], first[s+1]) (first[0] = 0, first[count] = n); its
 * loaded band moves at u = v·t, v, a = 0 with t = step·dt[s], v = velocity[s]; it takes
 * steps 1..last_step[s] and then keeps its state while the others go on. Bonds between
 * specimens are refused (place them more than a horizon apart, or give each its own CSR:
 * vpd_set_neighbors accepts the concatenated rows). Call after vpd_initialize_dynamics; it
 * applies every specimen's constraints at t = 0. count 0 returns to one specimen. While a
 * table is set, the scalar-time entries (vpd_step, vpd_predict, vpd_apply_constraints,
 * vpd_compute_deformation_gradient, vpd_update_material, vpd_compute_force_and_correct)
 * are refused; vpd_initialize_material covers every particle. Set the table before enabling
 * run diagnostics: they then keep one record per specimen (see vpd_enable_diagnostics).
 * Each specimen's results equal those of the same specimen run alone through vpd_step with
 * ramp (v·t, v, 0), bit for bit. */
int vpd_set_specimens(vpd_context *ctx, int count, const int64_t *first, const double *dt,
                      const double *velocity, const int32_t *last_step);
/* Step `step` (1-based, consecutive while fuse_next chains) of every specimen that has not
 * finished. With fuse_next the next step's predictor runs inside the force pass; a
 * specimen's final step is never followed by a prediction. */
/* Prescribed band kinematics per specimen, e.g. the half-cosine ramp of
 * RegionalConstraintSystem::ramp (src/host/BoundaryConditions.cpp:70-79) that the reference's
 * boundary cases use. Specimen s takes rows offsets[s] .. offsets[s+1]−1 of `table` (3 doubles
 * per row: the loaded band's u, v, a), one row per step 0..last_step[s], in place of its
 * velocity band; a specimen with no rows keeps the velocity band. Call after
 * vpd_set_specimens and before the first step; it applies every specimen's constraints at
 * t = 0 again. offsets NULL removes the tables. The rows are the caller's own values (e.g.
 * vpd_smooth_ramp at step·dt[s]), so each specimen's results equal its run alone through
 * vpd_step with the same ramp values, bit for bit. No counterpart in the reference (its cases
 * run one per process, benchmarks/tensile3d/run_cases.py:95-96). */
int vpd_set_specimen_bands(vpd_context *ctx, const int64_t *offsets, const double *table);
int vpd_step_specimens(vpd_context *ctx, int32_t step, int fuse_next, int want_fields);
/* Steps first_step .. first_step+count−1 of every unfinished specimen as `count` replays of
 * one captured HIP graph of a fused step (F + damage, material, force + correct + next
 * prediction, then a device step counter moves on), so the host issues one launch per step
 * instead of one per kernel. Results equal vpd_step_specimens with fuse_next, bit for bit.
 * Specimens still running afterwards hold their prediction for the next step (read state
 * after a vpd_step_specimens with fuse_next = 0, or once every specimen has finished).
 * Returns when the steps are done. No counterpart in the reference (its step loop is serial
 * host code, BoundaryValueCaseRunner.cpp:207-229). */
int vpd_run_specimens(vpd_context *ctx, int32_t first_step, int32_t count);

/* ---- return-map solver of the material stage -------------------------------------- */
/* Selects the solver (enum vpd_return_map) and the fast modes' relative residual bound
 * rtol (|r| <= rtol·q_trial; 1e-12 by default). Bisection is the parity mode; the others
 * change state 6 (iteration count) and Δε̄ᵖ within the tolerance. */
int vpd_set_return_map(vpd_context *ctx, int mode, double rtol);
/* The network used by VPD_RETURN_MAP_LEARNED (owned by the caller; NULL detaches). */
typedef struct vpd_mlp vpd_mlp;
int vpd_attach_warm_start(vpd_context *ctx, vpd_mlp *mlp);

/* ---- launch tuning ---------------------------------------------------------------- */
/* Knobs per bond pass (no reference counterpart):
 *  - sweeps: the pass walks the particle range in this many consecutive sweeps, so the
 *    neighbour data an XCD touches at once (its stretch of the sweep plus a horizon either
 *    side) can fit its L2;
 *  - LDS bytes reserved per 256-thread workgroup, capping workgroups per CU;
 *  - gather window: 0 gathers neighbour data straight from memory; W in {64, 80, 96, 128,
 *    160} stages runs of W consecutive particles' data in each wavefront's LDS and serves
 *    the bonds that fall in them from there (the LDS reservation is then at least the
 *    windows'); W in {256, 384, 480, 512, 640} does so once per 256-thread workgroup
 *    (form 3); with form 4, shared rows, W in {16, 32, 64} is the lanes per particle, each
 *    computing one bond's term per W-bond chunk, summed in bond order (for small blocks,
 *    where one lane per particle leaves the GPU idle);
 *  - bonds per window (windowed gathers): the most bonds a particle takes from one window
 *    before the wavefront moves on, 0 for no limit, at most 1024;
 *  - window form (windowed gathers): 0 walks the bond lists during the pass; 1 replays a
 *    window plan (built once per bond structure, window and limit: the windows of every
 *    wavefront and, per window and particle, where its bonds' neighbours sit), 2 does that
 *    and stages the next window while summing the current one; 3 replays a plan of windows
 *    shared by the workgroup's 256 particles (per window and particle the neighbours' LDS
 *    offsets, ≈ 2 B per bond). Forms 1 and 2 need a window of 64, 80, 96 or 128 and 1 to 8
 *    bonds per window, form 3 a window of 256, 384, 480, 512 or 640 and no limit (0),
 *    form 4 a window of 16, 32 or 64 and no limit, form 0 a window of 64 to 160; a step
 *    with another combination fails with
 *    VPD_ERR_ARGUMENT.
 * The sum order, and so every result bit, is the same for every setting (sweeps do not apply
 * to form 4). vpd_autotune times both passes over the grid of these on the context's own
 * block (form 4 on blocks of at most 262,144 particles) and keeps the fastest;
 * tuned[] (VPD_TUNE_COUNT ints) receives the values indexed by enum vpd_tuning. */
enum vpd_tuning {
  VPD_TUNE_DEFGRAD_LDS = 0,
  VPD_TUNE_FORCE_LDS = 1,
  VPD_TUNE_DEFGRAD_ROUNDS = 2,
  VPD_TUNE_FORCE_ROUNDS = 3,
  VPD_TUNE_DEFGRAD_WINDOW = 4,
  VPD_TUNE_FORCE_WINDOW = 5,
  VPD_TUNE_DEFGRAD_WINDOW_BONDS = 6,
  VPD_TUNE_FORCE_WINDOW_BONDS = 7,
  VPD_TUNE_DEFGRAD_WINDOW_PLAN = 8,
  VPD_TUNE_FORCE_WINDOW_PLAN = 9,
  VPD_TUNE_COUNT = 10
};
int vpd_set_tuning(vpd_context *ctx, int knob, int value);
int vpd_get_tuning(vpd_context *ctx, int knob, int *value);
int vpd_autotune(vpd_context *ctx, int reps, int *tuned);
/* The settings of bond pass `pass` (0 F, 1 force) that the context's last vpd_autotune timed
 * again in its final stage (up to 4), fastest first: `*count` (at most `max`) full tunings of
 * VPD_TUNE_COUNT ints in knobs[], each the context's current tuning with that pass's five
 * knobs replaced, and their mean trial times in ms[] (may be NULL). *count is 0 before any
 * autotuning. A caller can time these in its own steps (the trials time each pass alone, and
 * settings within a few percent of each other there can rank otherwise in a step). No
 * reference counterpart. */
int vpd_autotune_candidates(vpd_context *ctx, int pass, int max, int *knobs, float *ms,
                            int *count);

/* ---- run diagnostics ------------------------------------------------------------- */
/* The per-step observables of runMaterial (src/host/BoundaryValueCaseRunner.cpp:231-305),
 * reduced on the device inside the material and force kernels of every vpd_step while
 * enabled. Each step appends one record of VPD_DIAG_STRIDE doubles (slots below) to a
 * device buffer of `capacity` records; vpd_fetch_diagnostics copies the pending records out
 * in step order and empties the buffer (vpd_step fails while it is full). Sums run in a
 * fixed order: identical inputs and particle count give identical records, whatever the
 * launch tuning; they differ from the CPU's serial sums by rounding only. With batched
 * specimens (vpd_set_specimens before this call) vpd_step_specimens appends, per step, one
 * record per specimen (specimen-minor: `capacity` steps hold capacity × count records); each
 * specimen's chunks are laid out from its first particle, so its records equal those of its
 * run alone bit for bit. The configuration (axes, weights, state indices) is shared: weights
 * and face sets are per particle of the whole block. */
enum vpd_diag_slot {
  VPD_DIAG_INTERNAL_ENERGY = 0,  /* Σ contraction(½(σ̄_n + σ_{n+1}), E_{n+1} − E_n)·V (:231-237) */
  VPD_DIAG_OBSERVED_STRESS = 1,  /* Σ σ[axis][axis]·w over the observation weights (:250-253) */
  VPD_DIAG_OBSERVED_STRAIN = 2,  /* Σ E[axis][axis]·w */
  VPD_DIAG_MAX_EQPS = 3,         /* max(0, statev[eqps_index]); 0 if eqps_index < 0 */
  VPD_DIAG_MAX_INDICATOR = 4,    /* max(0, internalIndicator(statev)) (:95-110) */
  VPD_DIAG_NONFINITE_MATERIAL = 5, /* 1 if any stress, strain or state value is not finite */
  VPD_DIAG_NET_FORCE_X = 6,      /* Σ f·V (:248-249) */
  VPD_DIAG_NET_FORCE_Y = 7,
  VPD_DIAG_NET_FORCE_Z = 8,
  VPD_DIAG_REACTION = 9,         /* Σ over the load band (code RAMP on load_axis) of
                                    (density·ramp[2] − f[axis])·V (:268-271) */
  VPD_DIAG_KINETIC = 10,         /* Σ ½·density·V·|v|² after the step (:75-83) */
  VPD_DIAG_LATERAL_NEG = 11,     /* Σ u[lateral_axis] over particles with lateral_sets bit 0 */
  VPD_DIAG_LATERAL_POS = 12,     /* Σ u[lateral_axis] over particles with lateral_sets bit 1 */
  VPD_DIAG_NONFINITE_FORCE = 13, /* 1 if any force component is not finite */
  VPD_DIAG_STRIDE = 16
};
typedef struct vpd_diag_config {
  int load_axis;                     /* 0 x, 1 y, 2 z */
  int lateral_axis;
  const double *observation_weight;  /* host, n; NULL: no observation window */
  const uint8_t *lateral_sets;       /* host, n; bit 0 negative face, bit 1 positive face */
  int eqps_index;                    /* state index of ε̄ᵖ, or -1 */
  int indicator_count;               /* 0..8 state indices of the internal indicator */
  int indicator_indices[8];
  int indicator_min;                 /* 1: min over the indices; 0: max */
  int indicator_one_minus;           /* 1: report 1 − value */
  int capacity;                      /* step records held between fetches */
} vpd_diag_config;
/* NULL cfg disables. The configuration is copied; the arrays may be freed afterwards. */
int vpd_enable_diagnostics(vpd_context *ctx, const vpd_diag_config *cfg);
/* Synchronises; writes the pending records (count × VPD_DIAG_STRIDE doubles). */
int vpd_fetch_diagnostics(vpd_context *ctx, double *records, int capacity, int *count);

/* ---- per-kernel timing ------------------------------------------------------------ */
/* enable = 0: off. enable = k >= 1: HIP events bracket the kernels of every k-th vpd_step /
 * vpd_step_specimens call (counting from the next one) and every launch of the other entries;
 * k > 1 keeps the events' own cost (a few microseconds per kernel) out of most steps. No
 * reference counterpart. */
int vpd_set_timing(vpd_context *ctx, int enable);
/* Synchronises, sums the recorded launches per kernel and clears them. Writes up to
 * `capacity` entries and returns their count in *count. */
int vpd_collect_timing(vpd_context *ctx, vpd_kernel_time *out, int capacity, int *count);

/* ---- the VUMAT on device arrays ---------------------------------------------------- */
/* vumat_adapter_call (materials/bridge/vumat_c_wrapper_jinfo.f90:1-44; typedef in
 * include/VumatMaterialAdapter.hpp:35-39) running the Johnson-Cook routine
 * (materials/examples/JohnsonCook.f90:12-172) on the GPU: the same 33 arguments, scalars
 * (nblock … lanneal, stepTime, totalTime, dt, cmname) on the host, every array a device
 * pointer in column-major [k + nblock*c] layout. Synchronous. Dimension errors that make
 * the Fortran routine STOP return VPD_ERR_ARGUMENT instead. */
int vpd_vumat_adapter_call_device(int *nblock, int *ndir, int *nshr, int *nstatev, int *nfieldv,
                                  int *nprops, int *lanneal, double *stepTime, double *totalTime,
                                  double *dt, char *cmname, double *coordMp, double *charLength,
                                  double *props, double *density, double *strainInc,
                                  double *relSpinInc, double *tempOld, double *stretchOld,
                                  double *defgradOld, double *fieldOld, double *stressOld,
                                  double *stateOld, double *enerInternOld, double *enerInelasOld,
                                  double *tempNew, double *stretchNew, double *defgradNew,
                                  double *fieldNew, double *stressNew, double *stateNew,
                                  double *enerInternNew, double *enerInelasNew);

/* ---- device memory for callers without a HIP runtime of their own ------------------ */
/* Plumbing for the device VUMAT entry (no reference counterpart): allocate, free and copy
 * device buffers through this library's HIP runtime. to_device 1: host → device. */
int vpd_device_malloc(uint64_t bytes, void **out);
int vpd_device_free(void *ptr);
int vpd_memcpy(void *dst, const void *src, uint64_t bytes, int to_device);

/* The neighbour build of vpd_build_neighbors (buildNeighborList,
 * src/pd/PdCorrespondenceSkeleton.cpp:95-129: the same rows, bit for bit) on caller device
 * arrays, allocating nothing. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (serving loops and harnesses that forbid hipMalloc). Its work memory
 * is one static device array of about 44 MiB that the library holds on each GPU it runs on
 * until the process ends; it limits a call to n <= 2097152 particles and a cell grid of at
 * most 1048576 cells (cells of width >= horizon over the bounding box), and larger inputs are
 * refused (VPD_ERR_ARGUMENT / VPD_ERR_UNSUPPORTED). Every call on one GPU shares that array:
 * calls that overlap on one GPU (other host threads, other streams) race and must be
 * serialised by the caller.
 *   x        device, three planes x[d*n + i] (d = 0, 1, 2)
 *   rowptr   device int32[n + 1], written
 *   col      device int32[col_capacity], written (may be NULL when col_capacity is 0)
 *   *nnz     host: the neighbour count, also when col_capacity is too small (then the call
 *            returns VPD_ERR_ARGUMENT with rowptr written and col untouched: size col and call
 *            again)
 *   stream   a hipStream_t, or NULL for the default stream. The call synchronises it (it reads
 *            the coordinate bounds and the counts back), so it cannot be graph-captured. */
int vpd_build_neighbors_noalloc(const double *x, int64_t n, double horizon, int32_t *rowptr,
                                int32_t *col, int64_t col_capacity, int64_t *nnz, void *stream);

/* The Johnson-Cook return map alone on device arrays (column-major [k + nblock*c]), for
 * comparing solvers on identical inputs. No reference counterpart beyond the solver of
 * materials/examples/JohnsonCook.f90:114-136. mode: 0 the reference bisection (parity),
 * 1 the same bisection run to |r| <= rtol·q, 2 safeguarded Newton seeded with state 5,
 * 3 safeguarded Newton seeded with `seed` (the learned warm start), 4 mode 0 evaluating the
 * FP64 yield at every iteration. iterations/fallbacks
 * (per point, may be NULL) report the solver work; *elapsed_ms the kernel time. */
int vpd_return_map_device(int mode, int64_t nblock, int nstatev, const double *props, int nprops,
                          double dt, double rtol, const double *strain_inc,
                          const double *stress_old, const double *state_old,
                          const double *ener_intern_old, const double *ener_inelas_old,
                          const double *seed, double *stress_new, double *state_new,
                          double *ener_intern_new, double *ener_inelas_new, int32_t *iterations,
                          int32_t *fallbacks, double *elapsed_ms);

/* The largest errors of the FP32 hardware log2 and exp2 over every input the bisection's
 * FP32 yield screen can give them (every float in [0.5, 1] for log2, absolute error; every
 * float in [0, 1] for exp2, relative error), each against the FP64 function. The screen's
 * error bound is built on these figures; this call measures them on the device at hand. No
 * reference counterpart. */
int vpd_screen_accuracy(double *log2_abs_error, double *exp2_rel_error);

/* The learned return map (mode 3) on device arrays with the network attached instead of a
 * seed array: a point with a positive state 5 is seeded with it and the others with the
 * network's estimate, which runs in dense tiles on a list of those points (the seeds of
 * vpd_mlp_seed_device), then the Newton solve. seed_out (device, nblock doubles, may be NULL)
 * receives the seeds; without it only the network's seeds are written (to the network's work
 * buffer) and the solve reads state 5 itself, the same seeds and results with less traffic.
 * *elapsed_ms covers seeding and solve.
 * Uses the network's work buffers: calls with one network must not overlap. Other arguments
 * as for vpd_return_map_device. No reference counterpart. */
int vpd_return_map_learned_device(vpd_mlp *mlp, int64_t nblock, int nstatev,
                                  const double *props, int nprops, double dt, double rtol,
                                  const double *strain_inc, const double *stress_old,
                                  const double *state_old, const double *ener_intern_old,
                                  const double *ener_inelas_old, double *seed_out,
                                  double *stress_new, double *state_new, double *ener_intern_new,
                                  double *ener_inelas_new, int32_t *iterations,
                                  int32_t *fallbacks, double *elapsed_ms);

/* ---- the return-map warm start (no reference counterpart) -------------------------- */
/* An 8 → 64 → 64 → 64 → 1 ReLU network mapping (6 Voigt strain increments, ε̄ᵖ, T) to a
 * Δε̄ᵖ estimate that seeds the Newton return map (vpd_return_map_device mode 3). Trained on
 * the GPU (forward, backward, Adam; FP32 master weights and optimiser state) with bf16 MFMA
 * products and FP32 accumulation by default, or with plain FP32 kernels
 * (vpd_mlp_set_training_form); inference runs as bf16 MFMA with FP32 accumulation.
 * theta: 8961 FP32 values, W1 (64×8), b1, W2 (64×64), b2, W3 (64×64), b3, W4 (1×64), b4;
 * norm: input mean[8], std[8], output mean, std. Gradient sums run in a fixed order, so a
 * training run is deterministic in either form. */
enum vpd_mlp_training_form {
  VPD_MLP_TRAIN_FP32 = 0, /* one thread per output, FP32 arithmetic throughout */
  VPD_MLP_TRAIN_MFMA = 1  /* bf16 operands on the matrix cores, FP32 accumulation (default) */
};
int vpd_mlp_create(int max_batch, vpd_mlp **out);
void vpd_mlp_destroy(vpd_mlp *mlp);
const char *vpd_mlp_last_error(const vpd_mlp *mlp);
int vpd_mlp_set_params(vpd_mlp *mlp, const float *theta, const float *norm);
int vpd_mlp_get_params(vpd_mlp *mlp, float *theta, float *norm);
/* MSE loss and its gradient over one batch of normalised inputs x[B][8], targets t[B]. */
int vpd_mlp_loss_and_gradient(vpd_mlp *mlp, const float *x, const float *t, int batch,
                              float *loss, float *grad);
int vpd_mlp_set_dataset(vpd_mlp *mlp, const float *x, const float *t, int64_t n, uint64_t seed);
/* `steps` Adam steps on minibatches of the dataset; *last_loss = loss of the last batch. */
int vpd_mlp_train(vpd_mlp *mlp, int steps, int batch, float lr, float *last_loss);
/* The arithmetic of vpd_mlp_loss_and_gradient and vpd_mlp_train (enum vpd_mlp_training_form). */
int vpd_mlp_set_training_form(vpd_mlp *mlp, int form);
/* Raw network outputs (normalised units) for normalised inputs, through the MFMA kernel. */
int vpd_mlp_predict(vpd_mlp *mlp, const float *x, int64_t n, float *y);
/* Return-map seeds Δε̄ᵖ for n points from device planes: inc[c*inc_stride + i] (c < 6) and
 * the JC state planes state[c*state_stride + i] (ε̄ᵖ = c 1, T = c 2, state 5 = c 4). A point
 * with a positive state 5 (its previous plastic increment, the reference's own seed,
 * JohnsonCook.f90:135) is seeded with it; the others (first yield) get the network's
 * estimate, computed for those points only. *elapsed_ms: both passes. */
int vpd_mlp_seed_device(vpd_mlp *mlp, const double *inc, int64_t inc_stride, const double *state,
                        int64_t state_stride, int64_t n, double *seed, double *elapsed_ms);

/* ---- host helpers of the case runner (no GPU work) --------------------------------- */
/* makeParticlePatch (src/host/Geometry.cpp:10-34): x[n][3], volume[n], n = nx·ny·nz. */
int vpd_make_patch(int nx, int ny, int nz, double spacing, double *x, double *volume);
/* RegionalConstraintSystem (src/host/BoundaryConditions.cpp:54-68, 81-115) reduced to per-
 * particle codes. component 0: support left / load right along x; 1: bottom / top along y. */
int vpd_regional_constraint_codes(const double *x, int64_t n, int component, int region_layers,
                                  double spacing, int plane_strain, uint8_t *codes,
                                  int64_t *anchor, int64_t *rotation_pin);
/* RegionalConstraintSystem::ramp (src/host/BoundaryConditions.cpp:70-79). */
void vpd_smooth_ramp(double amplitude, double rise_time, double time, double out[3]);
/* selectBoundaryRegion (src/host/ParticleSets.cpp:12-50): mask[i] = 1 for the particles within
 * `width` of the "left"/"right" (x) or "bottom"/"top" (y) extreme. Returns the count, or
 * -VPD_ERR_ARGUMENT for an unknown side, a width <= 0 or an empty selection. */
int64_t vpd_select_boundary_region(const double *x, int64_t n, const char *side, double width,
                                   uint8_t *mask);

/* ---- the JSON case runner ----------------------------------------------------------- */
/* runDemoCaseFromJson (src/host/DemoCaseRunner.cpp:435-458) on the GPU, with the same case
 * and material JSON schema, output files and stdout lines as the reference:
 *  - "vumat_pd_boundary_case" (runBoundaryValueCase, BoundaryValueCaseRunner.cpp:309-440):
 *    timeseries.csv, field_snapshots.csv, metadata.json; Johnson-Cook materials;
 *  - "vumat_pd_demo_case" (the affine host, DemoCaseRunner.cpp:295-431): timeseries.csv,
 *    field_final.csv, metadata.json; Johnson-Cook or linear elastic (Elastic.f) materials.
 * Damage "none" or "critical_stretch". "vumat_material_point_case" files are refused.
 * Returns 0 when every material passes its acceptance checks, 2 when one fails, 1 on an
 * error (the message goes to stderr as the reference prints it). */
int vpd_run_case_json(const char *case_path, const char *plugin_directory);
/* Several "vumat_pd_boundary_case" files run as one batch: every material of every case is a
 * specimen of one particle block (vpd_set_specimens, vpd_set_specimen_bands with each case's
 * half-cosine ramp, per-specimen run diagnostics), stepped together; each case's output files
 * are those vpd_run_case_json writes for it (runtimeSeconds is the batch's wall time), its
 * rows bit for bit. Replaces the process pool of benchmarks/tensile3d/run_cases.py:95-96
 * (and the serial runs of time_step_study.py:132-148) for cases that share the material,
 * solver density, damage settings, load axis and state indices; other combinations are
 * refused. Returns 0 when every material passes, 2 when one fails, 1 on an error. */
int vpd_run_cases_json(const char *const *case_paths, int count, const char *plugin_directory);

#ifdef __cplusplus
}
#endif

#endif /* VPD_H */
