/*
 * rednative.h -- C ABI of the MI355X-native red.cuda integration path.
 *
 * Plain C: pointers, sizes and return codes only.  Every function returns
 * RED_OK (0) or a positive RED_ERR_* code and never throws; the message of the
 * last failure on a context is available from red_last_error().
 *
 * The reference is a monolithic C++/CUDA program; its "plugin" seam is the class
 * surface that red.cuda/src/red.cuda/main.cu drives.  Each entry point below names
 * the reference function it replaces.  Units are the reference's internal ones:
 * lengths in AU, masses in solar masses, time t_int = k * t_day and velocity
 * v_int = v_day / k with k = 0.01720209895 (Gaussian units, G = 1;
 * red.cuda/src/red.cuda/pp_disk.cu:1512-1551, integrator.cu:29).
 *
 * Body arrays are in body-file order: star, giant, rocky, proto (self-interacting,
 * SI), super-planetesimal, planetesimal (non-self-interacting, NSI), test particle
 * (non-interacting, NI) (red.cuda/src/red.cuda/number_of_bodies.cpp:106-258).
 * Positions and velocities cross the boundary as n*3 interleaved doubles.
 *
 * Ownership: a context owns all device memory it allocates and one HIP stream.
 * Host arrays passed in are copied; host arrays passed out are caller-owned.
 * Threading: one context per host thread; calls on one context must not overlap.
 */
#ifndef REDNATIVE_H
#define REDNATIVE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------- */
#define RED_OK                 0
#define RED_ERR_ARG            1   /* invalid argument or state                       */
#define RED_ERR_HIP            2   /* HIP runtime failure (launch, copy, allocation)  */
#define RED_ERR_EVENT_OVERFLOW 3   /* more pair events in one attempt than capacity   */
#define RED_ERR_NAN            4   /* non-finite error estimate                       */
#define RED_ERR_UNSUPPORTED    5   /* what the reference throws "not implemented" for */
#define RED_ERR_NO_STAR        6   /* no body of type STAR among the massive bodies   */
#define RED_ERR_IO             7   /* file could not be read or parsed                */
#define RED_ERR_STEP_LIMIT     8   /* attempt loop exceeded its cap                   */

/* ---- enumerations (values match red.cuda/red.type/red_type.h:45-101) ---- */
#define RED_EULER   0
#define RED_RK2     1
#define RED_RK4     2
#define RED_RK5     3   /* Cash-Karp RK5(4) */
#define RED_RKF78   4   /* Runge-Kutta-Fehlberg 7(8), the reference's "rkf8" */

#define RED_FRAME_ASTRO 0
#define RED_FRAME_BARY  1

#define RED_PREC_FP64   0   /* FP64 everywhere, same operation order as the reference */
#define RED_PREC_MIXED  1   /* FP64 differences, FP32 1/r^3, compensated FP64 sums    */

#define RED_EV_NONE            0
#define RED_EV_HIT_CENTRUM     1
#define RED_EV_EJECTION        2
#define RED_EV_CLOSE_ENCOUNTER 3
#define RED_EV_COLLISION       4

/* One event record: the reference's event_data_t (red.cuda/red.type/red_type.h:230-275)
 * as a flat C struct, plus the detection key (attempt, stage, sink, source) that
 * fixes a canonical order for records appended concurrently on the GPU. */
typedef struct red_event {
    int    name;                  /* RED_EV_*                                       */
    int    id1, idx1;             /* survivor id / index                            */
    int    id2, idx2;             /* merger id / index                              */
    int    attempt, stage;        /* attempt within the step, stage index r         */
    int    sink, source;          /* (i, j) of the detecting pair                   */
    int    pad;
    double t, d;                  /* event time (internal units), separation        */
    double p1[4], p2[4], ps[4];   /* mass radius density cd: survivor, merger, after */
    double r1[3], v1[3], r2[3], v2[3], rs[3], vs[3];
    double oe1[6], oe2[6], oes[6];/* sma ecc inc peri node mean                     */
} red_event;

/* Run parameters: the reference's parameter file (red.cuda/src/red.cuda/parameter.cpp:32-292)
 * plus the switches this implementation adds. */
typedef struct red_params {
    int    integrator;            /* RED_EULER .. RED_RKF78                         */
    int    adaptive;              /* embedded-error step control (set by tolerance) */
    double tolerance;
    int    error_check_for_tp;    /* include test particles in the error norm       */
    int    close_encounter;       /* report pair events instead of merging          */
    int    inner_steps;           /* detect pair events inside the force sweep      */
    int    frame_center;          /* RED_FRAME_BARY, or RED_FRAME_ASTRO: star = body 0
                                     at the origin, FP64, self-interacting bodies only
                                     (pp_disk.cu:409-454, number_of_bodies.cpp:205-209) */
    double hit_centrum;           /* AU, 0 = off                                    */
    double ejection;              /* AU, 0 = off                                    */
    double radii_enhance_factor;  /* collision when d < f (R_i + R_j), 0 = off      */
    double encounter_hill_factor; /* close encounter when d < f R_Hill, 0 = off     */
    double dt_day;                /* initial step in days (reference: 0.001)        */
    int    precision;             /* RED_PREC_FP64 or RED_PREC_MIXED                */
    int    event_capacity;        /* pair events per attempt, 0 = max(4096, n)      */
    int    use_graph;             /* replay each attempt from a captured HIP graph  */
    int    kahan_pairs;           /* mixed: compensate every pair, not every tile   */
    int    device;                /* HIP device ordinal                             */
    int    block;                 /* gravity launch shape, 0 = autotuned/default    */
    int    sinks_per_thread;      /* 0 = default; strategy 3: sinks per block
                                     (4, 8 or 16)                                   */
    int    strategy;              /* 0 = auto, 1 = scalar-broadcast, 2 = LDS tiles,
                                     3 = FP64 sequential sums (the FP64 default):
                                     producer/consumer, one work-item per sink for
                                     disjoint sweeps of >= 49152 sinks;
                                     4 = as 3 with one work-item per sink for every
                                     disjoint sweep                                 */
    int    fp64_split;            /* FP64: let sums split across source ranges for
                                     occupancy (faster, not bit-identical to the
                                     reference's sequential order); 0 = sequential  */
    int    source_parts;          /* source-range split factor, 0 = automatic       */
    int    host_threads;          /* ensembles: host threads of the per-disk event
                                     passes, 0 = min(8, hardware threads)           */
    int    padded_storage;        /* -ups (options.cpp:114): the padded kernel's
                                     semantics (pp_disk.cu:144-215): no id tests,
                                     inactive bodies keep pulling, being pulled and
                                     being tested for collisions until compaction   */
} red_params;

/* Bodies in body-file order (host arrays, read during the call only). */
typedef struct red_bodies {
    int           counts[7];      /* ns ngp nrp npp nspl npl ntp                    */
    const int*    id;
    const int*    type;
    const int*    mig_type;       /* may be NULL */
    const double* mig_stop_at;    /* may be NULL */
    const double* epoch;          /* may be NULL */
    const double* mass;
    const double* radius;
    const double* density;
    const double* cd;
    const double* r;              /* n*3, AU                                        */
    const double* v;              /* n*3, internal units                            */
} red_bodies;

typedef struct red_step_info {
    double dt_did;                /* step taken (internal units)                    */
    double dt_next;               /* next trial step                                */
    double err;                   /* last error estimate (scaled as the reference)  */
    int    tries;                 /* attempts in this step                          */
    int    n_events;              /* raw pair events recorded in the step           */
} red_step_info;

typedef struct red_ctx red_ctx;

/* Context life cycle.  Replaces pp_disk::pp_disk + options::create_integrator
 * (red.cuda/src/red.cuda/pp_disk.cu:1204-1235, options.cpp:197-274): allocates device
 * storage, uploads the state and thresholds (pp_disk.cu:1411-1452).  The state must
 * already be in internal units and in the wanted frame (see red_host_* below). */
int red_ctx_create(const red_bodies* bodies, const red_params* params, red_ctx** out);
int red_ctx_destroy(red_ctx* ctx);
/* Message of the last failure on ctx ("" if none); static text when ctx is NULL. */
const char* red_last_error(const red_ctx* ctx);

/* Counts after removals (number_of_bodies::update_numbers, number_of_bodies.cpp:27-104). */
int red_get_counts(red_ctx* ctx, int counts[7]);
/* Time and trial step, internal units (pp_disk::t, integrator::dt_try). */
int red_get_time(red_ctx* ctx, double* t);
int red_set_time(red_ctx* ctx, double t);
int red_get_dt(red_ctx* ctx, double* dt_try);
int red_set_dt(red_ctx* ctx, double dt_try);

/* State transfer, body-file order.  Replaces pp_disk::copy_to_host / copy_to_device
 * (pp_disk.cu:1411-1439).  Any output pointer may be NULL. */
int red_get_state(red_ctx* ctx, double* r, double* v, double* mass, double* radius,
                  double* density, double* cd, int* id, int* type);
int red_set_state(red_ctx* ctx, const double* r, const double* v);

/* Force evaluation on caller-owned device arrays.  Replaces pp_disk::calc_dydx(1, ...)
 * -> call_kernel_calc_grav_accel (pp_disk.cu:609-681, 998-1041): the SI, NSI and NI
 * sweeps with the fused pair tests.  r_dev, v_dev, a_dev are device pointers to n*4
 * doubles (x, y, z, w per body; w of r_dev carries the mass, w of a_dev is written 0).
 * t is the stage time stamped on events; stage tags them.  Asynchronous on the
 * context stream; red_sync() waits. */
int red_accel(red_ctx* ctx, double t, int stage, const double* r_dev, const double* v_dev, double* a_dev);
/* Device arrays of the current state (n*4 doubles each, as above); valid until the
 * next red_step, red_check_rebuild or red_ctx_destroy. */
int red_state_device(red_ctx* ctx, double** r_dev, double** v_dev);
int red_sync(red_ctx* ctx);

/* One integrator step with step control.  Replaces integrator::step()
 * (int_euler.cu:91-141, int_rungekutta2.cu:112-152, int_rungekutta4.cu:189-273,
 * int_rungekutta5.cu:202-282, int_rungekutta8.cu:704-797). */
int red_step(red_ctx* ctx, red_step_info* info);

/* Fixed-step main-loop iterations on the device.  No single counterpart in the reference:
 * it runs, for fixed-step integrators (Euler, RK2, RK4, and RK5 / RKF7(8) without step
 * control), the iterations of the driver loop main.cu:159-281 in which nothing is left for
 * the host -- the ejection / hit-centrum check finds no body, the step, *ps += |dt|, and no
 * pair event was recorded -- in one work-group of one kernel, with the same operations in
 * the same order as red_step, so the state stays bit-identical to stepping one at a time.
 * Before each iteration it stops (reason) when t > t_stop (RED_FUSED_STOP_TIME),
 * output_interval <= |*ps| (RED_FUSED_STOP_OUTPUT), max_steps steps are done
 * (RED_FUSED_STOP_STEPS) or the ejection / hit-centrum check would find a body
 * (RED_FUSED_STOP_EJECTION: the caller runs that iteration itself); after a step that
 * recorded pair events it stops with RED_FUSED_STOP_EVENTS (the step is done; the caller
 * runs the collision / close-encounter and rebuild checks).  *ps is the driver's
 * accumulated step (in and out).  For disks of at most 4096 bodies and 16384 interacting
 * pairs per force evaluation (sum over the sweeps of sinks x sources; above that, launch-by-
 * launch steps are faster), FP64 sums in the reference order and the barycentric frame;
 * RED_ERR_UNSUPPORTED otherwise. */
#define RED_FUSED_STOP_STEPS    1
#define RED_FUSED_STOP_TIME     2
#define RED_FUSED_STOP_OUTPUT   3
#define RED_FUSED_STOP_EJECTION 4
#define RED_FUSED_STOP_EVENTS   5
int red_run_fixed(red_ctx* ctx, int max_steps, double t_stop, double output_interval, double* ps,
                  int* steps_done, int* stop_reason);

/* Event checks of the driver loop (main.cu:201-255).  *handled = 1 when events were
 * found and handled.  Replace pp_disk::check_for_ejection_hit_centrum (pp_disk.cu:683-715),
 * check_for_collision (:717-738), check_for_close_encounter(true) (:740-770) and
 * check_for_rebuild_vectors (:893-907). */
int red_check_ejection_hit_centrum(red_ctx* ctx, int* handled);
int red_check_collision(red_ctx* ctx, int* handled);
int red_check_close_encounter(red_ctx* ctx, int* handled);
int red_check_rebuild(red_ctx* ctx, int threshold, int* rebuilt);
/* pp_disk::clear_event_counter (pp_disk.cu:1469-1476). */
int red_clear_events(red_ctx* ctx);
/* Raw events since the last clear, canonical order (attempt, stage, sink, source).
 * Writes min(n, cap) records; *n = total, *overflow = 1 if records were dropped. */
int red_raw_events(red_ctx* ctx, red_event* buf, int cap, int* n, int* overflow);
/* Events of the last handling pass as pp_disk::sp_events (pp_disk.cu:1111-1149). */
int red_sp_events(red_ctx* ctx, red_event* buf, int cap, int* n);
/* Index map of the last compaction: old index of each new body (n entries). */
int red_last_rebuild_map(red_ctx* ctx, int* old_index, int cap, int* n);

/* out[8]: tried failed passed, total collisions ejections hit-centrum close-encounters,
 * events since the last rebuild (integrator.cu:258-263, pp_disk.h:129-133). */
int red_counters(red_ctx* ctx, int out[8]);

/* Not in the reference: total energy E = sum 1/2 m v^2 - sum m_i m_j / r_ij over
 * interacting pairs (NSI-NSI excluded) and angular momentum L = sum m r x v, G = 1. */
int red_energy(red_ctx* ctx, double* E, double L[3]);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): red_energy
 * without any allocation or host synchronisation.  Its potential partials live in a static
 * 1 MiB device array per GPU (disks whose n_si ceil(n_massive / 4096) partials exceed 2^17
 * doubles, e.g. more than about 20000 self-interacting bodies, are refused with RED_ERR_ARG);
 * the sums are formed on the device in
 * red_energy's order, so out_dev[4] = {E, Lx, Ly, Lz} equals red_energy's results bit for bit.
 * out_dev is caller-owned device memory; the call is asynchronous on the context stream
 * (red_sync waits).  Calls that overlap on one GPU share the array and race. */
int red_energy_noalloc(red_ctx* ctx, double* out_dev);

/* Host-side transforms (tools.cpp:122-178, pp_disk.cu:1512-1551): barycentric shift
 * (sums run from the last body down, as the reference), and t*k / v/k scaling. */
int red_host_transform_to_bc(int n, const double* mass, double* r, double* v);
/* Astro-centric shift (tools.cpp:180-236): subtracts body 0's position and velocity from
 * every body, from the last down; RED_ERR_NO_STAR unless body 0 has type 0 (STAR). */
int red_host_transform_to_ac(int n, const int* type, double* r, double* v);
int red_host_transform_velocity(int n, double* v);

/* Kernel timing (not in the reference, which has none beyond nvprof): with profiling on,
 * HIP events bracket every launch group on the context stream; red_get_profile returns
 * the accumulated milliseconds and launch counts per class.  Needs use_graph = 0. */
#define RED_PROF_SI     0   /* force sweep: self-interacting sinks x massive sources */
#define RED_PROF_NSI    1   /* force sweep: non-self-interacting sinks x SI sources  */
#define RED_PROF_NI     2   /* force sweep: test particles x massive sources         */
#define RED_PROF_STAGE  3   /* stage combinations and y_(n+1)                        */
#define RED_PROF_ERR    4   /* error estimate + max-reduction                        */
#define RED_PROF_PACK   5   /* mixed-precision source packing                        */
#define RED_PROF_LEARN  6   /* step features and learned step proposal               */
#define RED_PROF_EJECT  7   /* ejection / hit-centrum scan                            */
#define RED_PROF_CLASSES 8
int red_set_profiling(red_ctx* ctx, int on);
int red_get_profile(red_ctx* ctx, double ms[RED_PROF_CLASSES], int launches[RED_PROF_CLASSES], int reset);

/* Device buffers for callers without an allocator of their own, on the context's device
 * and stream (the reference's ALLOCATE_DEVICE_VECTOR / copy_vector_to_device /
 * copy_vector_to_host, red.cuda/src/red.cuda.utility/util.cu:147-330). */
int red_device_alloc(red_ctx* ctx, unsigned long long bytes, void** dev);
int red_device_free(red_ctx* ctx, void* dev);
int red_copy_to_device(red_ctx* ctx, void* dev, const void* host, unsigned long long bytes);
int red_copy_to_host(red_ctx* ctx, void* host, const void* dev, unsigned long long bytes);

/* ---- ensembles of independent disks ---------------------------------------------
 * Not in the reference (which integrates one disk per process): many disks of uneven size
 * share one set of device buffers; every stage of an attempt is one launch over all disks,
 * the attempt is captured once in a HIP graph (params->use_graph) and replayed, and each
 * disk runs the reference's step control (integrator.cu:211-263, int_rungekutta8.cu:704-797)
 * on its own time and trial step, on the device, so it takes exactly the steps it would take
 * alone.
 * Every disk also runs the main loop's event checks in the reference's order
 * (main.cu:201-255) with the single-disk code: the ejection / hit-centrum check before a step,
 * pair tests inside the force launches, the collision or close-encounter pass after a step
 * and the compaction after 8 events, so a member's events, merges and removals are those it
 * would have alone.  FP64, barycentric frame. */
typedef struct red_ensemble red_ensemble;
typedef struct red_ensemble_info {
    int    supersteps;            /* supersteps in which some disk attempted        */
    int    active_disks;          /* disks still integrating after the call         */
    double seconds;               /* wall time of the call                          */
    double event_seconds;         /* of which host event passes and ejection scans  */
    double rebuild_seconds;       /* of which slot compactions                      */
    int    batches;               /* batches launched (one graph replay each)       */
    int    full_batches;          /* batches whose later supersteps idled: their
                                     kept pair records passed half the capacity     */
} red_ensemble_info;
int red_ensemble_create(int n_disks, const red_bodies* disks, const red_params* params, red_ensemble** out);
int red_ensemble_destroy(red_ensemble* ens);
const char* red_ensemble_last_error(const red_ensemble* ens);
int red_ensemble_set_time(red_ensemble* ens, double t);
/* Steps every disk while its t <= t_stop (main.cu:176-255), at most max_attempts attempts. */
int red_ensemble_advance(red_ensemble* ens, double t_stop, int max_attempts, red_ensemble_info* info);
int red_ensemble_get_state(red_ensemble* ens, int disk, double* r, double* v);
/* A disk's counts after removals and its bodies' parameters (arrays of the current count;
 * any pointer may be NULL): the ensemble's pp_disk::copy_to_host (pp_disk.cu:1411-1439). */
int red_ensemble_get_bodies(red_ensemble* ens, int disk, int counts[7], double* mass, double* radius,
                            double* density, double* cd, int* id, int* type);
/* out[8] of one disk, as red_counters: tried failed passed, collisions ejections hit-centrum
 * close-encounters, events since the last rebuild. */
int red_ensemble_disk_counters(red_ensemble* ens, int disk, int out[8]);
/* The records of every event pass a disk has handled (each pass's sp list, in order) since
 * the last call; copies min(n, cap), *n = records available, the copied ones are removed. */
int red_ensemble_disk_events(red_ensemble* ens, int disk, red_event* buf, int cap, int* n);
/* out[3] = tried failed passed of one disk; its time and trial step. */
int red_ensemble_disk_info(red_ensemble* ens, int disk, int out[3], double* t, double* dt_try);
/* Picks the force-launch shape (threads per block x source tile, each in {64,128,256}) by
 * timing `reps` attempts' force launches per shape (every stage, each with its pair-test
 * rule, on the current state); ms[9] receives the timings (block-major). */
int red_ensemble_autotune(red_ensemble* ens, int reps, int* best_block, int* best_jtile, double ms[9]);
/* Supersteps per graph launch, 1..64 (default 8).  Each disk's controller runs on the device
 * (the reference's, int_rungekutta8.cu:704-797 and siblings, with step_pow below), so a
 * batch of supersteps needs the host once: a disk whose step needs the host (a merge, an
 * ejection scan, a compaction that moves bodies) waits out the rest of its batch. */
int red_ensemble_set_batch(red_ensemble* ens, int supersteps);

/* The step controllers' factor x^y (0 < y <= 1/2) rounded to nearest, as the host controller
 * (on_device = 0) or the ensembles' device controller (on_device = 1, on the given device)
 * computes it, for n pairs of host arrays.  The reference calls its C runtime's pow here
 * (int_rungekutta8.cu:777, int_rungekutta4.cu:253), whose last bit depends on the platform. */
int red_step_pow(const double* x, const double* y, double* out, int n, int on_device, int device);

/* ---- calibration --------------------------------------------------------------------
 * No counterpart in the reference: a measurement for the force sweeps' rooflines.  Vector-ALU
 * issue rates of the device from kernels of independent instruction chains with no memory
 * traffic: out[0..5] receive lane-instructions per second of FP32 FMA, FP32 multiply, FP64
 * add, f64 <-> f32 conversion, v_rsq_f32, and of FP32 FMAs and FP64 adds issued by different
 * wavefronts of each SIMD; out[6] receives pairs per second of the mixed-precision sweep's
 * pair sequence with its sources in cache, out[7] the same for the FP64 sweep's pair term
 * added in ascending j.  Allocates and frees its own buffers. */
#define RED_VALU_RATES 8
int red_valu_rates(int device, double* out, int n_out);

/* ---- learned step-size proposer -----------------------------------------------------
 * Not in the reference, whose only step control is the elementary controller
 * dt <- 0.9 dt (tol/err)^(1/order) (int_rungekutta8.cu:776-777 and its siblings).
 *
 * Step records.  With recording on, every step of an adaptive method writes one record of
 * RED_HIST_FIELDS floats once its first attempt has finished.  eta denotes the
 * tolerance-matching step of an attempt, eta = dt (tol/err)^(1/order); step n is the last
 * accepted step, dt_n its step and err_n its error.
 *   [0] log10 dt_n              [1] log10(err_n / tol)       [2] ln(eta_n / eta_(n-1))
 *   [3] ln(eta_(n-1)/eta_(n-2)) [4] rejected attempts of n   [5] ln(1 + pair events of n)
 *   [6] ln(eta_(n-2)/eta_(n-3)) [7] log10 min d/R_H           [8] change of [7] since step n
 *   [9] log10 max |a|           [10] log10 max |a - a_kepler|/|a_kepler|  [11] change of [10]
 *   [12] label ln(eta_(n+1) / eta_n) from the first attempt of step n+1
 *   [13] network output used for the step (0 without one)   [14] ln(dt_first / eta_n)
 *   [15] 1 if the first attempt was accepted, else 0
 * d/R_H: the smallest ratio of separation to mutual Hill radius
 * ((m_i+m_j)/(3 m_star))^(1/3) (r_i+r_j)/2 (heliocentric r) over the interacting pairs of
 * massive bodies other than the star; a from the step's first force evaluation.
 * Features [0..11] are the network input. */
#define RED_HIST_FIELDS   16
#define RED_MLP_FEATURES  12
typedef struct red_mlp red_mlp;
/* 12 -> 64 -> 64 -> 1 GELU network, BF16 weights, FP32 accumulation, on `device`.
 * params: red_mlp_param_count() FP32 values W1[64][12] b1[64] W2[64][64] b2[64] W3[64] b3
 * (NULL = zeros).  The input is standardised as (x - mean) * inv_std (default 0 and 1). */
int red_mlp_param_count(void);
int red_mlp_create(int device, const float* params, red_mlp** out);
int red_mlp_destroy(red_mlp* net);
const char* red_mlp_last_error(const red_mlp* net);
int red_mlp_set_normalizer(red_mlp* net, const float* mean, const float* inv_std);
/* FP32 master parameters, their BF16 image and the normaliser (mean[12] then inv_std[12]);
 * any pointer may be NULL. */
int red_mlp_get_params(red_mlp* net, float* params, unsigned short* bf16, float* norm);
int red_mlp_predict(red_mlp* net, const float* x, int n, float* y);
/* Training loss: mean squared error (default) or the pinball loss of level tau in (0, 1),
 * mean(max(tau (t - y), (tau - 1) (t - y))), whose minimiser is the tau-quantile. */
#define RED_MLP_LOSS_MSE      0
#define RED_MLP_LOSS_PINBALL  1
int red_mlp_set_loss(red_mlp* net, int kind, float tau);
/* Training loss over (x[n][12], target[n]) and its gradient (BF16 weights, FP32 sums). */
int red_mlp_gradient(red_mlp* net, const float* x, const float* target, int n, float* grad, float* loss);
/* Adam (beta 0.9, 0.999, eps 1e-8) over minibatches of `batch` records in the given order,
 * `epochs` passes; losses[epoch] receives the mean minibatch loss (may be NULL). */
int red_mlp_train(red_mlp* net, const float* x, const float* target, int n, int epochs, int batch, float lr,
                  float* losses);
int red_mlp_adam_state(red_mlp* net, float* m, float* v, long* step);

/* Records steps into a buffer of `capacity` records (0 switches recording off). */
int red_history_enable(red_ctx* ctx, int capacity);
/* Copies up to `cap` records, oldest first, and clears the buffer; *n = records copied. */
int red_history_get(red_ctx* ctx, float* records, int cap, int* n);
/* Lets `net` propose the first trial step of every step once four accepted steps are known:
 * dt = safety * eta_n * exp(clamp(y, -max_log_change, max_log_change)).  Rejected attempts
 * are retried with the reference controller's step.  net = NULL restores the reference
 * controller.  The network must live on the context's device and outlive its use. */
int red_set_proposer(red_ctx* ctx, red_mlp* net, double safety, double max_log_change);
/* out[0] proposals made, out[1] first attempts accepted after a proposal. */
int red_proposer_stats(red_ctx* ctx, int out[2]);

/* Library version string. */
const char* red_version(void);

#ifdef __cplusplus
}
#endif
#endif /* REDNATIVE_H */
