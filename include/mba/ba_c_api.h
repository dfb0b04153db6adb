/*
 * ba_c_api.h -- C ABI of the MI355X bundle-adjustment engine (libslamplay_ba.so).
 *
 * Drop-in boundary for slamplay's BA path (backend/slam_ba_simple_g2o_ceres):
 *   - ba_problem_view maps a slamplay::BALProblem with zero copies (BALProblem.h:59-98 accessors:
 *     num_cameras/num_points/num_observations, camera_index, point_index, observations,
 *     mutable_cameras, mutable_points);
 *   - ba_solve replaces `void SolveBA(BALProblem&)` (test_bundle_adjustment_ceres.cpp:35,58-101):
 *     one Snavely residual per observation with HuberLoss(1.0) (:75-86), trust-region
 *     Levenberg-Marquardt with the Schur complement (:96-100), parameters updated in place (:83-84);
 *   - ba_eval_residual_jacobian replaces the per-observation
 *     AutoDiffCostFunction<SnavelyReprojectionError,2,9,3> evaluation
 *     (utils/camera/SnavelyReprojectionError.h:42-92).
 *
 * All entry points are synchronous, return a ba_status (0 = OK) and never abort; the message of
 * the last failure on the calling thread is returned by ba_last_error_string().  Host arrays are
 * owned by the caller; device buffers are owned by the library and freed by the *_destroy calls.
 * No allocation crosses the ABI.
 */
#ifndef MBA_BA_C_API_H
#define MBA_BA_C_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BA_ABI_VERSION 2

typedef enum ba_status {
    BA_OK = 0,
    BA_ERR_INVALID_ARGUMENT = 1,
    BA_ERR_OUT_OF_MEMORY = 2,
    BA_ERR_NON_FINITE = 3,
    BA_ERR_HIP = 4,
    BA_ERR_NO_DEVICE = 5,
    BA_ERR_CAPACITY = 6,
    BA_ERR_IO = 7,
    BA_ERR_NOT_FOUND = 8,
    BA_ERR_TIMEOUT = 9 /* the serving loop made no progress within its deadline (state in the message) */
} ba_status;

typedef enum ba_precision {
    BA_PRECISION_FP64 = 0,  /* every stage in FP64: the plain drop-in */
    BA_PRECISION_MIXED = 1  /* FP32 residual/Jacobian/Hessian blocks and PCG matvecs; FP64 block
                               inverses, dense-step refinement and the LM accept-test cost */
} ba_precision;

typedef enum ba_linear_solver {
    BA_SOLVER_AUTO = 0,            /* dense Schur up to dense_schur_max_cameras, else iterative */
    BA_SOLVER_DENSE_SCHUR = 1,     /* explicit reduced camera system + dense Cholesky */
    BA_SOLVER_ITERATIVE_SCHUR = 2  /* implicit reduced camera system + block-Jacobi PCG */
} ba_linear_solver;

typedef enum ba_loss { BA_LOSS_NONE = 0, BA_LOSS_HUBER = 1 } ba_loss;

typedef enum ba_mode {
    BA_MODE_CERES = 0, /* test_bundle_adjustment_ceres.cpp's SolveBA (Ceres trust-region LM) */
    BA_MODE_G2O = 1    /* test_bundle_adjustment_g2o.cpp's SolveBA (g2o Levenberg, SO(3) rotation update) */
} ba_mode;

/* ceres::TerminationType's values; BA_USER_FAILURE (Ceres: a callback aborted the solve) marks a
 * serving request ended by ba_server_cancel. */
typedef enum ba_termination {
    BA_CONVERGENCE = 0,
    BA_NO_CONVERGENCE = 1,
    BA_FAILURE = 2,
    BA_USER_FAILURE = 4
} ba_termination;

/* Zero-copy view of a BAL problem.  cameras: 9 doubles per camera (angle-axis[3], t[3], f, k1, k2),
 * points: 3 doubles per point; both updated in place by ba_solve.  observations: x,y interleaved. */
typedef struct ba_problem_view {
    int32_t num_cameras;
    int32_t num_points;
    int32_t num_observations;
    int32_t camera_block_size; /* must be 9 (BALProblem::camera_block_size without quaternions) */
    const int32_t* camera_index;
    const int32_t* point_index;
    const double* observations;
    double* cameras;
    double* points;
} ba_problem_view;

/* Solver options.  Defaults (ba_default_options) are Ceres Solver::Options defaults, which is what
 * the reference's SolveBA uses (it only sets SPARSE_SCHUR and progress output). */
typedef struct ba_options {
    int32_t max_num_iterations;                /* 50 */
    int32_t max_num_consecutive_invalid_steps; /* 5 */
    int32_t jacobi_scaling;                    /* 1 */
    int32_t loss;                              /* BA_LOSS_HUBER */
    int32_t linear_solver;                     /* BA_SOLVER_AUTO */
    int32_t pcg_max_iterations;                /* 50 */
    int32_t precision;                         /* BA_PRECISION_FP64 */
    int32_t dense_schur_max_cameras;           /* 300 */
    int32_t device;                            /* HIP device ordinal, 0 */
    int32_t verbose;                           /* print a Ceres-style iteration table to stdout */
    int32_t refine_steps;                      /* FP64 refinement steps of the dense mixed solve, 2; batched
                                                  windows run up to 10x as many preconditioned CG steps */
    int32_t launch_mode;                       /* 0: each LM iteration replayed as one hipGraph (default);
                                                  1: plain stream launches (always used while profiling) */
    double loss_scale;                  /* 1.0 (HuberLoss(1.0)) */
    double function_tolerance;          /* 1e-6 */
    double gradient_tolerance;          /* 1e-10 */
    double parameter_tolerance;         /* 1e-8 */
    double initial_trust_region_radius; /* 1e4 */
    double max_trust_region_radius;     /* 1e16 */
    double min_trust_region_radius;     /* 1e-32 */
    double min_relative_decrease;       /* 1e-3 */
    double min_lm_diagonal;             /* 1e-6 */
    double max_lm_diagonal;             /* 1e32 */
    double pcg_tolerance;               /* 1e-6: stop when ||r|| <= tol * ||b|| */
    /* BA_MODE_CERES (default): the Ceres driver's trust-region LM (test_bundle_adjustment_ceres.cpp:96-100).
     * BA_MODE_G2O: the g2o driver's (test_bundle_adjustment_g2o.cpp:166-218): camera rotation updated as
     * exp(delta) R, lambda-LM on H + lambda I without column scaling, lambda_0 = g2o_tau * max diag(J'J),
     * at most max_num_iterations iterations of up to g2o_max_trials trials; costs reported as g2o's chi2
     * (sum of rho, no 1/2).  Single-problem solves only (the batch and server entry points reject it). */
    int32_t mode;
    int32_t g2o_max_trials; /* 10 */
    double g2o_tau;         /* 1e-5 */
} ba_options;

/* One row of the per-iteration trace (the Ceres progress table's columns). */
typedef struct ba_iteration {
    int32_t iteration;
    int32_t step_is_successful;
    int32_t step_is_valid;
    int32_t linear_solver_iterations;
    double cost;
    double cost_change;
    double gradient_max_norm;
    double step_norm;
    double relative_decrease;
    double trust_region_radius;
    double model_cost_change;
    double iteration_time_in_seconds;  /* device time since the previous row (row 0: 0) */
    double cumulative_time_in_seconds; /* device time since the first evaluation */
} ba_iteration;

/* Which test ended the solve (Ceres's termination messages, Solver::Summary::message). */
typedef enum ba_termination_reason {
    BA_REASON_NONE = 0,
    BA_REASON_MAX_ITERATIONS = 1,     /* NO_CONVERGENCE: max_num_iterations reached */
    BA_REASON_FUNCTION_TOLERANCE = 2, /* CONVERGENCE: |cost_change| / cost <= function_tolerance */
    BA_REASON_GRADIENT_TOLERANCE = 3, /* CONVERGENCE: max |gradient| <= gradient_tolerance */
    BA_REASON_PARAMETER_TOLERANCE = 4,/* CONVERGENCE: |step| <= parameter_tolerance (|x| + parameter_tolerance) */
    BA_REASON_MIN_RADIUS = 5,         /* CONVERGENCE: trust region radius below min_trust_region_radius */
    BA_REASON_INVALID_STEPS = 6,      /* FAILURE: max_num_consecutive_invalid_steps invalid steps in a row */
    BA_REASON_NON_FINITE_COST = 7,    /* FAILURE: the cost is not finite */
    BA_REASON_G2O_STOP = 8            /* g2o mode: a trial budget exhausted or rho = 0 (g2o stops) */
} ba_termination_reason;

typedef struct ba_summary {
    double initial_cost;
    double final_cost;
    int32_t num_iterations; /* trace rows written, iteration 0 included */
    int32_t num_successful_steps;
    int32_t num_unsuccessful_steps;
    int32_t termination_type; /* ba_termination */
    int32_t linear_solver_used; /* BA_SOLVER_DENSE_SCHUR or BA_SOLVER_ITERATIVE_SCHUR */
    int32_t total_linear_solver_iterations;
    double total_time_ms;      /* wall time of the solve, device work included */
    ba_iteration* trace;       /* optional caller-owned array of trace_capacity rows */
    int32_t trace_capacity;
    int32_t termination_reason; /* ba_termination_reason */
    double termination_value;   /* the tested quantity (|cost_change| / cost, max |gradient|, ...) */
} ba_summary;

int32_t ba_abi_version(void);
const char* ba_last_error_string(void);
void ba_default_options(ba_options* options);
int ba_device_count(int32_t* count);

/* Raw (uncorrected) residuals r (2 per observation), Jacobians jac_camera (2x9 row-major per
 * observation) and jac_point (2x3 row-major) in observation order, evaluated on the GPU in the
 * requested precision (FP32 results are widened to double).  Replaces the AutoDiffCostFunction
 * evaluation of SnavelyReprojectionError.h:42-92.  jac_* may be NULL. */
int ba_eval_residual_jacobian(const ba_problem_view* problem, int32_t precision, double* residuals,
                              double* jac_camera, double* jac_point);

/* 1/2 sum_i rho(|r_i|^2) on the GPU in FP64 (loss = ba_loss, scale = Huber a). */
int ba_eval_cost(const ba_problem_view* problem, int32_t loss, double loss_scale, double* cost);

/* The text of Ceres's Solver::Summary::FullReport() for a solve of `num_cameras` cameras,
 * `num_points` points and `num_observations` observations with `options` (the reference driver
 * prints it after ceres::Solve, test_bundle_adjustment_ceres.cpp:101): problem size, minimizer and
 * linear solver, costs, step counts, times and the termination message.  Writes at most `capacity`
 * bytes (NUL-terminated) into `out`; *length (optional) receives the full length. */
int ba_summary_full_report(const ba_summary* summary, int32_t num_cameras, int32_t num_points,
                           int32_t num_observations, const ba_options* options, char* out, int64_t capacity,
                           int64_t* length);

/* ba_eval_cost without any device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Every pointer of `problem` and `cost` (one double) is DEVICE memory on
 * the current device; the kernels of ba_eval_cost are enqueued on `stream` (a hipStream_t, NULL for
 * the null stream) and the call returns without waiting.  Work space is a fixed per-GPU buffer of
 * the library (about 1.8 MB, held until the process ends): at most 4096 cameras and 2^24
 * observations (BA_ERR_CAPACITY beyond).  Indices are not checked on the host; an out-of-range
 * index makes the cost NaN.  Calls that overlap on one GPU share the buffer and race. */
int ba_eval_cost_noalloc(const ba_problem_view* problem, int32_t loss, double loss_scale, double* cost,
                         void* stream);

/* Full solve: upload, device-resident LM loop, write parameters back in place.
 * Replaces SolveBA(BALProblem&) of test_bundle_adjustment_ceres.cpp:58-101. summary may be NULL. */
int ba_solve(const ba_problem_view* problem, const ba_options* options, ba_summary* summary);

/* ---------------------------------------------------------------------------------------------
 * Stereo local bundle adjustment of full_slam (full_slam/src/backend.cpp:67-199, the g2o types of
 * full_slam/include/myslam/g2o_types.h:48-175): SE3 keyframe poses observe world points through
 * the left or right camera of a calibrated rig; e = measurement - pi(K E_side T X); pose update
 * T <- exp(xi) T with the edge's own Jacobian; g2o Levenberg (ba_stereo_default_options: mode
 * BA_MODE_G2O, 10 iterations, Huber 5.991, dense Schur); afterwards the reference's outlier rule:
 * chi2 = |e|^2 > threshold, the threshold starting at loss_scale and doubled (at most 5 times)
 * while the inlier ratio is <= 0.5. */
typedef struct ba_stereo_view {
    int32_t num_poses;
    int32_t num_points;
    int32_t num_observations;
    const int32_t* pose_index;
    const int32_t* point_index;
    const int32_t* side;         /* 0: left camera, 1: right camera (backend.cpp:117-121) */
    const double* observations;  /* u, v pixels per observation */
    double* poses;               /* 6 per pose: angle-axis and translation of T (world to body) */
    double* points;              /* 3 per point, world frame */
    double K[4];                 /* fx, fy, cx, cy (backend.cpp:97) */
    double extrinsics[2][6];     /* left, right camera: angle-axis and translation of E (backend.cpp:98-99) */
} ba_stereo_view;

void ba_stereo_default_options(ba_options* options);
/* Solves in place; outlier (optional, num_observations bytes) receives 1 for outliers, chi2_threshold
 * (optional) the final classification threshold. */
int ba_stereo_solve(const ba_stereo_view* problem, const ba_options* options, ba_summary* summary, uint8_t* outlier,
                    double* chi2_threshold);
/* Many full_slam windows at once (the local BA of backend.cpp:67-199 for several maps, robots or
 * sliding windows): every window is solved as ba_stereo_solve would (same model, options and outlier
 * rule) by the batched window engine (all windows advanced together, one launch sequence per LM
 * trial; see ba_batch_create), in place.  summaries (optional) has `count` entries; outliers
 * (optional) `count` pointers, each NULL or to num_observations bytes of that window;
 * chi2_thresholds (optional) `count` doubles.  options NULL = ba_stereo_default_options. */
int ba_stereo_batch_solve(const ba_stereo_view* windows, int32_t count, const ba_options* options,
                          ba_summary* summaries, uint8_t* const* outliers, double* chi2_thresholds);

/* Device-resident solver: create uploads the problem and builds the device layout once; run solves
 * from the parameters currently held on the device; reset restores the parameters given at create;
 * download copies the device parameters to host AoS arrays. */
typedef struct ba_solver ba_solver;
int ba_solver_create(const ba_problem_view* problem, const ba_options* options, ba_solver** out);
int ba_solver_run(ba_solver* solver, ba_summary* summary);
int ba_solver_reset(ba_solver* solver);
int ba_solver_download(ba_solver* solver, double* cameras, double* points);
int ba_solver_destroy(ba_solver* solver);

/* BALProblem::Normalize (utils/datasets/BALProblem.cpp:263-303) on the parameters the solver holds
 * on the device: per-axis upper medians, the upper median of the L1 distances to them, then points
 * and camera centres scaled to that median distance 100.  An empty problem is left as it is. */
int ba_solver_normalize(ba_solver* solver);
/* BALProblem::Perturb (BALProblem.cpp:305-336) on the device-held parameters, with the normals
 * (RandNormal, :39-55) that glibc's rand() gives after srand(seed); seed 1 is the stream of a
 * process that never calls srand, as the reference's drivers.  Sigmas must be >= 0. */
int ba_solver_perturb(ba_solver* solver, uint32_t seed, double rotation_sigma, double translation_sigma,
                      double point_sigma);

/* Per-kernel timing collected with HIP events on the solver's stream while enabled. */
typedef struct ba_kernel_time {
    char name[48];
    int64_t launches;
    double total_ms;
    double bytes_per_launch; /* algorithmic bytes of one launch (stated per kernel) */
    double flops_per_launch;
} ba_kernel_time;
int ba_solver_set_profiling(ba_solver* solver, int32_t enabled);
int ba_solver_kernel_times(ba_solver* solver, ba_kernel_time* out, int32_t capacity, int32_t* count);
/* Observations in the solver's device layout (after dropping nothing) and device bytes held. */
int ba_solver_info(ba_solver* solver, int64_t* device_bytes, int32_t* linear_solver);

/* ---- Batched local BA: many independent small problems ("windows") ----
 * The caller of windowed BA is full_slam's local bundle adjustment (full_slam/src/backend.cpp:67-199,
 * 7-10 keyframes); here every window is a Snavely BA problem solved exactly like ba_solve with the
 * dense Schur solver (per window: explicit reduced camera system + Cholesky), each with its own LM
 * state.  Every sum is formed in a fixed order, so a window's result does not depend on which other
 * windows share the launch nor on the schedule.  Windows compute in FP64 whatever options.precision
 * says (a reduced system of at most 288 unknowns is factored in FP64, as ba_solve does); it is
 * factored in the workgroup's LDS up to 21 cameras and in the window's own device buffer up to 32;
 * larger windows are solved by the single-problem solver (ba_solver, with the requested precision)
 * and reported the same way. */
typedef struct ba_batch ba_batch;
int ba_batch_create(const ba_problem_view* windows, int32_t count, const ba_options* options, ba_batch** out);
/* Solves every window from the parameters held on the device (options.max_num_iterations each).
 * summaries: optional array of `count` (their trace pointers are ignored); total_time_ms of a window
 * is the time from the start of the run to the moment that window finished on the device. */
int ba_batch_run(ba_batch* batch, ba_summary* summaries);
int ba_batch_reset(ba_batch* batch);
int ba_batch_download(ba_batch* batch, int32_t window, double* cameras, double* points);
int ba_batch_destroy(ba_batch* batch);
/* One-shot: create, run, write every window's parameters back in place, destroy. */
int ba_batch_solve(const ba_problem_view* windows, int32_t count, const ba_options* options, ba_summary* summaries);
/* The launch schedule of a batch's window kernels.  No choice in it changes a window's result
 * (each 9x9 block of S is reduced by one wave in pair order, each factor entry by one thread), so
 * ba_batch_create times every schedule on the batch itself and keeps the fastest. */
typedef struct ba_batch_schedule {
    int32_t schur_threads;       /* workgroup of the Schur-product kernel: 256 or 512 */
    int32_t solve_threads;       /* workgroup of the per-window factor and solve: 256 or 512 */
    int32_t blocks_per_wave;     /* 9x9 blocks of S per wave in one task of the Schur-product kernel: 1, 2, 4 or 8 */
    int32_t lds_cameras;         /* windows of at most this many cameras factor in LDS, larger ones in their device
                                    buffer (capped at what the LDS holds, 21; the launch reserves the largest such
                                    window's triangle for every workgroup) */
    int32_t candidates;          /* schedules the autotuner timed (0: set by the caller) */
    double iteration_ms;         /* one LM iteration of the batch with this schedule (0: not timed) */
    double default_iteration_ms; /* the same with the default schedule 256 / 256 / 2 / 21 (0: not timed) */
    int32_t lanes;               /* output only: launch streams the windows are dealt to (largest work first, in
                                    turn) and run concurrently -- one per 32 windows, at most 4; the schedule
                                    applies to every lane and the timings are of all lanes launched together */
} ba_batch_schedule;
/* The schedule in use. */
int ba_batch_get_schedule(ba_batch* batch, ba_batch_schedule* out);
/* Sets the schedule's four choices (timing fields ignored); NULL runs the autotuner again. */
int ba_batch_set_schedule(ba_batch* batch, const ba_batch_schedule* schedule);
/* Workgroup size of both the Schur-product and the factor kernel (256 or 512; 0 = autotune). */
int ba_batch_set_block(ba_batch* batch, int32_t threads);
int ba_batch_info(ba_batch* batch, int32_t* threads, int64_t* device_bytes);

/* ---- Serving: continuous batching of local-BA requests ----
 * Requests are admitted while their device footprint fits in `arena_bytes` and a slot is free,
 * their state stays resident on the device, and every step runs `iterations_per_step` LM
 * iterations for all resident windows in one launch; finished windows are retired and their
 * results queued for ba_server_poll. */
typedef struct ba_server_config {
    int64_t arena_bytes;          /* device memory for resident windows */
    int32_t max_windows;          /* resident window slots */
    int32_t iterations_per_step;  /* LM iterations per step (1 = per-iteration batching) */
    ba_options options;
} ba_server_config;
typedef struct ba_server_result {
    int64_t tag;
    double initial_cost;
    double final_cost;
    int32_t iterations;
    int32_t termination_type;
    double submit_time_s;   /* ba_server clock (seconds since create) */
    double admit_time_s;
    double finish_time_s;
    int32_t num_cameras;
    int32_t num_points;
    const double* cameras;  /* solved parameters, owned by the server, valid until the next poll */
    const double* points;
    int32_t evictions;      /* times the window was evicted (ba_server_evict) and admitted again */
} ba_server_result;
typedef struct ba_server ba_server;
int ba_server_create(const ba_server_config* config, ba_server** out);
/* Copies the window into the server's queue; it is admitted by a later ba_server_step. */
int ba_server_submit(ba_server* server, const ba_problem_view* window, int64_t tag);
/* Admits queued windows that fit, runs one step, retires finished windows.  *resident receives the
 * number of windows on the device after the step, *queued the number still waiting. */
int ba_server_step(ba_server* server, int32_t* resident, int32_t* queued);
/* Cancels request `tag` at the next ba_server_step: a queued or resident window leaves the server
 * with a result of termination_type BA_USER_FAILURE carrying the parameters it has then (those of
 * its last accepted step, the submitted ones if it never ran; costs NaN if it never ran), a large
 * problem still waiting for the single-problem solver likewise.  A tag that names nothing pending
 * then (already finished, unknown, a large problem being solved) is ignored.  Thread-safe. */
int ba_server_cancel(ba_server* server, int64_t tag);
/* Evicts resident window `tag` at the next ba_server_step: its whole device region (parameters,
 * LM state, trust region, working arrays) is copied to host memory, its slot and arena range are
 * freed, and it rejoins the back of the queue.  Admitted again (into any free slot and range), it
 * continues from that state, and its result is bit for bit that of a solve never evicted.  A tag
 * that is not resident then is ignored.  Thread-safe. */
int ba_server_evict(ba_server* server, int64_t tag);
/* Pops up to `capacity` finished results (oldest first). */
int ba_server_poll(ba_server* server, ba_server_result* out, int32_t capacity, int32_t* count);
double ba_server_clock(ba_server* server);
int ba_server_destroy(ba_server* server);

/* Load generator: Poisson arrivals at `rate_hz` over `count` requests drawn round-robin from
 * `pool`, served by a ba_server; per-request results (latency = finish - submit) in `results`. */
int ba_server_simulate(const ba_server_config* config, const ba_problem_view* pool, int32_t pool_count,
                       double rate_hz, int32_t count, uint64_t seed, ba_server_result* results);

/* ---- Host BALProblem (utils/datasets/BALProblem.{h,cpp}) through opaque handles ---- */
typedef struct ba_bal ba_bal;
/* BALProblem::BALProblem(filename, use_quaternions) (BALProblem.cpp:78-139); use_quaternions must
 * be 0 for solving (the Snavely functor is 9-parameter). */
int ba_bal_load(const char* path, int32_t use_quaternions, ba_bal** out);
/* A problem built from caller arrays (copied). */
int ba_bal_from_arrays(int32_t num_cameras, int32_t num_points, int32_t num_observations,
                       const int32_t* camera_index, const int32_t* point_index, const double* observations,
                       const double* cameras, const double* points, ba_bal** out);
int ba_bal_view(ba_bal* bal, ba_problem_view* view);
int ba_bal_normalize(ba_bal* bal);                                                      /* :263-303 */
int ba_bal_perturb(ba_bal* bal, double rotation_sigma, double translation_sigma, double point_sigma); /* :305-336 */
/* BAL writer. legacy_header != 0 reproduces the reference's 4-integer header "Nc Nc Np No"
 * (BALProblem.cpp:149), which its own reader cannot parse back; 0 writes "Nc Np No". */
int ba_bal_write(ba_bal* bal, const char* path, int32_t legacy_header);
int ba_bal_write_ply(ba_bal* bal, const char* path); /* :185-228 */
int ba_bal_destroy(ba_bal* bal);

/* ---- Seeded synthetic problems (the benchmark configurations) ---- */
typedef struct ba_synth_params {
    int32_t kind;        /* 1 circle (config 1), 2 random walk (configs 2-5) */
    int32_t num_cameras;
    int32_t num_points;  /* kind 2: 0 = draw tracks until target_observations is reached */
    int32_t min_track;
    int32_t max_track;
    int32_t target_observations;
    double zipf_s;       /* kind 2 track-length exponent; 0 = uniform on [min_track, max_track] */
    double pixel_noise;  /* observation noise sigma (px) */
    double rotation_sigma, translation_sigma, point_sigma; /* initial perturbation */
    double focal_mean, focal_sigma, distortion_sigma;
    uint64_t seed;
} ba_synth_params;
/* Fill defaults for a benchmark configuration (1..5; 4/5 describe one local-BA window). */
int ba_synth_default(int32_t config, ba_synth_params* params);
/* Generate a perturbed problem; if ground_truth is non-NULL it receives the exact parameters
 * (9*Nc + 3*Np doubles, cameras then points) the observations were projected from. */
int ba_synth_generate(const ba_synth_params* params, ba_bal** out, double* ground_truth, int64_t gt_capacity);

#ifdef __cplusplus
}
#endif
#endif /* MBA_BA_C_API_H */
