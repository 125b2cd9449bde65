/*
 * esthera.h - C ABI of the MI355X (gfx950) particle-filter engine.
 *
 * The engine runs the island-model filter step of the reference
 * (src/partfilt/opencl/pf_generic.c:438-808):
 *
 *   MTGP32 words -> Box-Muller -> sample+weight (model hook) -> per-sub-filter
 *   bitonic sort + gather -> argmax of sub-filter winners -> exchange of best
 *   particles (ring / 2-D torus / all-to-all) -> per-sub-filter resampling
 *   (prefix-sum multinomial, systematic, or Vose alias)
 *
 * as hand-written HIP kernels.  Everything crossing this boundary is plain C:
 * integers, floats, host pointers and device pointers obtained from
 * esth_dev_alloc().  Every function returns an int status (ESTH_OK == 0) and
 * never exits the process; esth_last_error() returns a thread-local message.
 *
 * Threading: a filter, batch or server handle is used by one host thread at a
 * time; distinct handles may be driven from different threads.
 *
 * Citations name the reference interface each entry point replaces.
 */
#ifndef ESTHERA_H
#define ESTHERA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ESTH_VERSION 100

/* ---------------------------------------------------------------- status */
#define ESTH_OK               0
#define ESTH_ERR_INVALID     -1  /* bad argument or descriptor */
#define ESTH_ERR_NOMEM       -2  /* host or device allocation failed */
#define ESTH_ERR_DEVICE      -3  /* HIP runtime error */
#define ESTH_ERR_BUDGET      -4  /* request exceeds the engine's HBM budget */
#define ESTH_ERR_UNSUPPORTED -5  /* combination not built */
#define ESTH_ERR_IO          -6  /* trace file error */
#define ESTH_ERR_STATE       -7  /* handle in the wrong state for the call */
#define ESTH_ERR_AGAIN       -8  /* nothing ready yet (polling) */

/* Thread-local message for the last failing call of this thread. */
const char* esth_last_error(void);
int esth_version(void);

/* ------------------------------------------------------------ limits
 * Reference limits: pf_model.h:22-24 (MAX_NUM_PARTICLES 1024,
 * MAX_NUM_BLOCKS 65536, MAX_NUM_TRANSFER = 512), checked at pf_generic.c:1333-1342. */
#define ESTH_MAX_NUM_PARTICLES 1024
#define ESTH_MAX_NUM_BLOCKS    65536
#define ESTH_MAX_NUM_TRANSFER  512
#define ESTH_MAX_STATE_DIM     16
#define ESTH_FRAME_FLOATS      336   /* per-step model frame, see esth_model_build_frame */

/* ---------------------------------------------------------------- models
 * Compiled-in model plug-ins.  Each provides the reference's two hooks:
 * particle_set (copy one particle, pf_model.cl:33-48) and sampling_iw
 * (propagate with noise + importance weight, pf_model.cl:52-131).
 *
 *  robotarm5 : the reference's planar 5-joint arm on a moving base
 *              (src/models/robotarm/opencl/pf_model.{h,cl}); state
 *              a0..a4, x, y, vx, vy (9); 9 normals per particle.
 *  unicycle  : unicycle with range+bearing landmarks; state x, y, theta (3);
 *              3 normals per particle.
 *  dh6       : 6-joint Denavit-Hartenberg arm, constant-velocity dynamics;
 *              state q0..q5, qd0..qd5 (12); 6 normals per particle.
 *  unicycle10: the reference's unicycle (src/models/unicycle/particle_filter.cu:
 *              60-74,313-362): velocity motion model with noisy speed, turn
 *              rate and heading drift, 10 fixed range sensors above the
 *              plane; state x, y, theta (3); 3 normals per particle.
 *  robotarm5_camera: robotarm5 observed through a calibrated camera
 *              (src/models/robotarm/cuda/pf_model_soa_packed_code.cuh:33-43,
 *              57-143): the base point is projected to pixels by the pinhole
 *              + radial/tangential distortion model before the comparison;
 *              noise in the CUDA order (angles first); state as robotarm5.
 *  randomwalk3: a 3-D random walk observed directly (p += v dt + sigma_q z,
 *              log w = -|m - p|^2 / (2 sigma_r^2)); state x, y, z (3); 3
 *              normals.  No reference counterpart: the example of adding a
 *              model to the engine (its hook and frame builder in one header
 *              under esthera_amd/csrc/models/, one line in the model list).
 */
#define ESTH_MODEL_ROBOTARM5  0
#define ESTH_MODEL_UNICYCLE   1
#define ESTH_MODEL_DH6        2
#define ESTH_MODEL_UNICYCLE10 3
#define ESTH_MODEL_ROBOTARM5_CAMERA 4
#define ESTH_MODEL_RANDOMWALK3 5
#define ESTH_NUM_MODELS       6

/* robotarm5 (pf_model.h:34-56) */
typedef struct { float angles[5]; } esth_robotarm5_control;
typedef struct { float angles[5]; float x, y; } esth_robotarm5_measurement;
/* robotarm5 takes no parameters (its constants are fixed, pf_model.cl:1-13). */

/* unicycle */
#define ESTH_UNICYCLE_MAX_LANDMARKS 64
typedef struct {
    int   num_landmarks;
    float sigma_x, sigma_y, sigma_theta;   /* additive process noise */
    float sigma_range, sigma_bearing;      /* measurement noise */
    float lx[ESTH_UNICYCLE_MAX_LANDMARKS];
    float ly[ESTH_UNICYCLE_MAX_LANDMARKS];
} esth_unicycle_params;
typedef struct { float v, omega; } esth_unicycle_control;
typedef struct {
    float range[ESTH_UNICYCLE_MAX_LANDMARKS];
    float bearing[ESTH_UNICYCLE_MAX_LANDMARKS];
} esth_unicycle_measurement;

/* dh6 */
typedef struct {
    float a[6], d[6], alpha[6];            /* DH link length, offset, twist */
    float sigma_acc, sigma_pos, sigma_enc;
} esth_dh6_params;
typedef struct { float unused; } esth_dh6_control;
typedef struct { float p[3]; float q[6]; } esth_dh6_measurement;

/* unicycle10: control is esth_unicycle_control (velocity, angular velocity) */
#define ESTH_UNICYCLE10_SENSORS 10
typedef struct {
    float sx[ESTH_UNICYCLE10_SENSORS], sy[ESTH_UNICYCLE10_SENSORS], sz[ESTH_UNICYCLE10_SENSORS];
    float noise_v, noise_omega, noise_gamma;
} esth_unicycle10_params;
typedef struct { float range[ESTH_UNICYCLE10_SENSORS]; } esth_unicycle10_measurement;
/* the reference's sensor positions and noise constants (particle_filter.cu:60-74) */
void esth_unicycle10_params_default(esth_unicycle10_params* p);

/* robotarm5_camera: control and measurement are robotarm5's (x, y in pixels);
 * the parameters are the camera's intrinsics: focal lengths, principal point
 * and distortion kc1, kc2 (radial), kc3, kc4 (tangential). */
typedef struct { float fc[2], cc[2], kc[4]; } esth_camera_params;
/* the reference's calibration (pf_model_soa_packed_code.cuh:33-43) */
void esth_camera_params_default(esth_camera_params* p);

/* randomwalk3: sigma_q (process), sigma_r (measurement); control is the
 * velocity, measurement the observed position */
typedef struct { float sigma_q, sigma_r; } esth_randomwalk3_params;
typedef struct { float v[3]; } esth_randomwalk3_control;
typedef struct { float p[3]; } esth_randomwalk3_measurement;

/* Model id for "robotarm5" | "unicycle" | "dh6" | "unicycle10" |
 * "robotarm5_camera" | "randomwalk3"; -1 if unknown. */
int esth_model_lookup(const char* name);
/* Sizes for a model: state floats per particle, normals per particle, and
 * the byte sizes of its params / control / measurement structs. */
int esth_model_info(int model, int* state_dim, int* noise_dim, size_t* params_size,
                    size_t* control_size, size_t* measurement_size);
/* Build the per-step model frame (ESTH_FRAME_FLOATS floats) that the kernels
 * read: the model's params, control, measurement and dt folded into
 * precomputed constants.  Exposed for the stage-level entry points. */
int esth_model_build_frame(int model, const void* params, const void* control,
                           const void* measurement, float dt, float* frame_out);

/* -------------------------------------------------------------- filters */
#define ESTH_TOPO_NONE     0
#define ESTH_TOPO_RING     1   /* exchange_ring, pf_generic.cl:368-383 */
#define ESTH_TOPO_TORUS    2   /* exchange_2dtorus, cuda/pf_generic.cu:352-393 (slot offsets fixed) */
#define ESTH_TOPO_ALL2ALL  3   /* global top-t, cuda/pf_generic2_all2all.cu:404-419,1007 */

#define ESTH_RESAMPLE_MULTINOMIAL 0  /* resampling, pf_generic.cl:385-432 */
#define ESTH_RESAMPLE_SYSTEMATIC  1  /* same scan, one uniform per sub-filter */
#define ESTH_RESAMPLE_ALIAS       2  /* resampling_valias, resampling-valias.cl:214-304 */

#define ESTH_WEIGHTS_LINEAR 0  /* weight = exp(-value) as the reference (pf_model.cl:129) */
#define ESTH_WEIGHTS_LOG    1  /* log-weight; resampling uses exp(lw - max) */

/* RNG ownership.  Both produce MTGP32-11213 words with the reference's
 * layout rule (stream j owns [j*size, (j+1)*size), seed+j+1, parameter set j,
 * pf_generic.c:208-210,282; mtgp32-opencl.c:46-48):
 *   COMPAT     : 96 streams, exactly the reference's buffer.
 *   PER_FILTER : stream count chosen from the filter's own size, so a filter's
 *                random numbers never depend on other filters in a batch;
 *                about 4 rounds of 768 words per stream and step (short
 *                chains: the lowest latency of one filter's words).
 *   PER_FILTER_LONG : the same with about 16 rounds per stream (a quarter of
 *                the streams): less status traffic and fewer work-groups per
 *                word, for throughput over many filters (batches). */
#define ESTH_RNG_MTGP_COMPAT     0
#define ESTH_RNG_MTGP_PER_FILTER 1
#define ESTH_RNG_MTGP_PER_FILTER_LONG 2

/* Device layout of a filter's particle state (S floats per particle, N
 * particles), the three the reference ships as model variants:
 *   SOA    : plane c of particle i at c*N + i (pf_model_soa.cuh: one array
 *            per state variable); the default.
 *   AOS    : i*S + c (pf_model.cuh: an array of state structs).
 *   PACKED : planes in groups of four, each group interleaved per particle
 *            (pf_model_soa_packed.cuh: float4 angles1 / float angles2 /
 *            float4 pos): group q = planes 4q..4q+w-1, w = min(4, S-4q),
 *            at 4q*N + i*w + (c - 4q).  Groups start at plane 0, so the
 *            robot arm's 9 planes pack as 4 + 4 + 1 (the reference's arm
 *            packs 4 + 1 + 4: angles 0-3, angle 4, x y vx vy).
 * The layout changes where the kernels read and write, never a result: a
 * filter's particles, weights and estimates are bit-identical in all three.
 * The host boundary (esth_filter_init, esth_filter_read_particles) stays
 * SOA; esth_state_convert turns buffers the caller holds in one layout into
 * another. */
#define ESTH_LAYOUT_SOA    0
#define ESTH_LAYOUT_AOS    1
#define ESTH_LAYOUT_PACKED 2

typedef struct {
    int      model;                /* ESTH_MODEL_* */
    uint32_t num_particles;        /* m: particles per sub-filter, power of two in [2, 1024] */
    uint32_t num_blocks;           /* G: sub-filters, in [1, 65536] */
    uint32_t num_transfer;         /* t: particles sent per neighbour (0 = no exchange) */
    int      topology;             /* ESTH_TOPO_* */
    int      torus_dimx;           /* 0: reference rule dimx = 2^floor(log2 sqrt G) */
    int      torus_dimy;           /* 0: G / dimx */
    int      resampler;            /* ESTH_RESAMPLE_* */
    float    resampling_threshold; /* gate: resample sub-filter g iff u_g <= thr (pf_generic.cl:403) */
    int      adaptive;             /* alias only: resample iff sum p^2 > thr (resampling-valias.cl:273-284) */
    int      weights;              /* ESTH_WEIGHTS_* */
    int      rng;                  /* ESTH_RNG_* */
    uint32_t seed;                 /* MTGP seed; stream j is seeded with seed + j + 1 */
    int      carry_weights;        /* 1: a sub-filter that is not resampled keeps its sorted,
                                      exchanged particles with their weights, and the next step
                                      multiplies each weight by the new likelihood (adds in log
                                      mode); a resampled sub-filter restarts from equal weights
                                      (the TODO of resampling-valias.cl:302-303).  0: the
                                      reference (weights restart every step; a skipped sub-filter
                                      keeps its unsorted pre-exchange state, pf_generic.cl:403-404) */
    int      state_layout;         /* ESTH_LAYOUT_* of the device-resident state (single filters;
                                      batches and servers keep SOA) */
} esth_filter_desc;

/* Fill a descriptor with the reference's defaults (pf_generic.c:1072-1084):
 * m=256, G=16, t=1, ring, multinomial, threshold 1.0, linear weights. */
void esth_filter_desc_default(esth_filter_desc* d, int model);

typedef struct {
    int      state_dim;
    double   mean[ESTH_MAX_STATE_DIM];   /* posterior mean sum(w x)/sum(w) before resampling */
    float    best[ESTH_MAX_STATE_DIM];   /* best particle of the best sub-filter (pf_generic.c:634-668) */
    float    best_weight;                /* its (log-)weight */
    int      best_block;                 /* its sub-filter */
    double   log_weight_max;             /* max log-weight (log mode; 0 in linear mode) */
    double   weight_sum;                 /* sum of weights scaled by exp(-log_weight_max) */
    uint64_t step;                       /* steps completed */
} esth_estimate;

typedef struct esth_engine esth_engine;
typedef struct esth_filter esth_filter;

/* One engine per GPU (device ordinal).  hbm_budget bytes bounds the device
 * memory the engine hands to filters, batches and servers (0 = 90% of free). */
int esth_engine_create(int device, size_t hbm_budget, esth_engine** out);
int esth_engine_destroy(esth_engine* e);
int esth_engine_memory(const esth_engine* e, size_t* used, size_t* budget);
/* Batches and servers: sample+weight and the block sort of sub-filter size
 * class m (64, 128, 256, 512 or 1024) in one launch (ESTH_FUSION_ON), in two
 * (ESTH_FUSION_OFF), or as the engine decides from the kernels' register use
 * (ESTH_FUSION_AUTO, the default).  ON falls back to two launches where the
 * fused kernel does not fit the model (it would spill).  Takes effect for
 * batches created and server waves formed after the call; esth_autotune
 * sets it from measurements.  No counterpart in the reference (one filter,
 * one launch per stage, pf_generic.c:521-585). */
#define ESTH_FUSION_AUTO (-1)
#define ESTH_FUSION_OFF 0
#define ESTH_FUSION_ON 1
int esth_engine_set_batch_fusion(esth_engine* e, uint32_t m, int mode);
int esth_engine_batch_fusion(const esth_engine* e, uint32_t m, int* mode);
/* Batches and servers: the tiling of the prefix-sum resampling of size class
 * m, as lanes per sub-filter: 32 (m = 64 only: two sub-filters per wave, two
 * positions per lane), 64 (one wave per sub-filter, the default), 128 (m >=
 * 256: two waves per sub-filter); 0 restores the default.  Results are the
 * same bit for bit at every tiling.  Takes effect for batches created and
 * server waves formed after the call; esth_autotune sets it from
 * measurements. */
int esth_engine_set_batch_resample_lanes(esth_engine* e, uint32_t m, int lanes);
int esth_engine_batch_resample_lanes(const esth_engine* e, uint32_t m, int* lanes);

/* Replaces the set-up half of particle_filter() (pf_generic.c:245-436). */
int esth_filter_create(esth_engine* e, const esth_filter_desc* desc, const void* params,
                       size_t params_size, esth_filter** out);
int esth_filter_destroy(esth_filter* f);
/* Set every particle to `state` (state_dim floats), as pf_generic.c:346-364 does. */
int esth_filter_init_uniform(esth_filter* f, const float* state);
/* Set all particles from host SoA data: state_dim planes of m*G floats. */
int esth_filter_init(esth_filter* f, const float* state_soa);
/* One filter step (pf_generic.c:438-808), enqueued on the filter's stream;
 * control/measurement are the model's structs (copied before return). */
int esth_filter_step(esth_filter* f, const void* control, const void* measurement, float dt);
int esth_filter_sync(esth_filter* f);
/* Estimate of the last completed step (waits for it). */
int esth_filter_estimate(esth_filter* f, esth_estimate* out);
/* Current particles (state_dim planes of m*G floats) and per-particle
 * weights as left by the sort/exchange stages (sorted per sub-filter). */
int esth_filter_read_particles(esth_filter* f, float* state_soa, float* weights);
/* Stage timing with hipEvents recorded on the filter's stream.
 * set_timing(f, level) starts a new measurement (eager steps only, up to 4096
 * steps): level 1 brackets only the sample+weight kernel (two events per
 * step; with the fused sample+sort launch, that launch), level 2 every
 * stage.  stage_times returns the mean device time per
 * step in microseconds of [rng wait, sample_weight, sort, argmax,
 * exchange/top-t, resample]; stages not timed are NaN. */
int esth_filter_set_timing(esth_filter* f, int on);
/* With timing on, time one step in `every` (default 1): the events around a
 * kernel cost the stream a few microseconds each, so sparse timing keeps the
 * measured stream close to the untimed one. */
int esth_filter_set_timing_every(esth_filter* f, int every);
int esth_filter_stage_times(esth_filter* f, float us[6]);
/* Host-side copy of `n` particles of `state_dim` floats from layout `from`
 * to layout `to` (ESTH_LAYOUT_*); src and dst must not overlap. */
int esth_state_convert(int state_dim, size_t n, int from, const float* src, int to, float* dst);
/* Checkpoint / resume: particles, weights, RNG status and step counter. */
int esth_filter_snapshot_size(esth_filter* f, size_t* bytes);
int esth_filter_snapshot(esth_filter* f, void* buf, size_t bytes);
int esth_filter_restore(esth_filter* f, const void* buf, size_t bytes);
/* The HIP stream (hipStream_t) the filter's work is enqueued on. */
int esth_filter_stream(esth_filter* f, void** hip_stream);
/* Generate the next step's MTGP32 words on a second stream while the current
 * step runs (default on; results are identical either way). */
int esth_filter_set_prefetch(esth_filter* f, int enable);
/* Sample+weight and the block sort as one launch per sub-filter (the
 * propagated state stays in LDS between them; no counterpart in the
 * reference, whose two kernels are pf_generic.c:521-533 and :572-585).  On by
 * default for 128 <= m <= 1024 where the fused kernel keeps 3 waves per SIMD
 * resident, or as many as the sampling alone; enable = 0 keeps two launches.  *active
 * (may be NULL) receives whether steps use the fused launch.  Results are
 * bit-identical either way.  Timing level 2 always runs the two launches. */
int esth_filter_set_fusion(esth_filter* f, int enable, int* active);
/* How the filter's MTGP32 words are made: *inline_words = 1 when the next
 * step's words come from extra work-groups of the step's sampling launch
 * (chosen where the sampling hides them), 0 when they come from a second
 * stream in blocks of *block_steps steps (mtgp32-opencl.cl:186-242 per step;
 * no counterpart in the reference, which generates them before each step,
 * pf_generic.c:191-232). */
int esth_filter_rng_mode(esth_filter* f, int* inline_words, int* block_steps);
/* One launch per step across the step boundary: step k's resampling, argmax
 * and posterior run at the start of step k + 1's launch (each work-group
 * resamples its sub-filter into LDS and samples and sorts it from there), so
 * resampled particles never go through HBM between the steps.  Taken by
 * default where it applies (the unicycle at m = 512 with its next step's
 * random words generated inline, a scan resampler on a ring or torus with
 * t > 0, SOA state, no carried weights, steps without per-stage timing,
 * eager or graph-replayed); reading the particles or the estimate, a
 * snapshot or a change of settings first runs the pending resampling.  Results are
 * the same bit for bit either way.  enable 0 switches it off; *active
 * (optional) reports whether it is taken.  No counterpart in the reference,
 * whose step ends with the resampling (pf_generic.c:732-757). */
int esth_filter_set_cross_step(esth_filter* f, int enable, int* active);
/* Capture one step into a HIP graph and replay it for later steps with the
 * same control/measurement layout (the frame buffer is updated per call). */
int esth_filter_use_graph(esth_filter* f, int enable);

/* -------------------------------------------------- reference driver
 * int particle_filter(m, G, t, thr, file) (pf_generic.c:245-887) for the
 * robotarm5 model over a recorded trace (pf_model_host.h:30-54).  Writes the
 * mean over steps of estimate_error(best particle, truth) (pf_model_host.h:56-69),
 * the DEBUG_ESTIMATE summary of pf_generic.c:814, to *mean_error. */
int esth_particle_filter(uint32_t m, uint32_t G, uint32_t t, float thr, const char* trace_path,
                         uint32_t seed, int resampler, double* mean_error, int* steps);

/* Trace rows of the robot-arm format (20 columns): meas x y, 5 meas angles,
 * 5 controls, dt, truth x y, 5 truth angles (pf_model_host.h:30-54).
 * truth receives 9 floats per row (a0..a4, x, y, vx=0, vy=0). */
int esth_trace_load_robotarm(const char* path, int max_rows, esth_robotarm5_measurement* meas,
                             esth_robotarm5_control* ctrl, float* dt, float* truth, int* rows);
/* Write `rows` rows in the robot-arm format that esth_trace_load_robotarm
 * reads (20 columns; truth 9 floats per row, of which vx, vy are not
 * written).  Floats are printed with 9 significant digits, so a trace
 * written here reads back bit-identically. */
int esth_trace_write_robotarm(const char* path, int rows, const esth_robotarm5_measurement* meas,
                              const esth_robotarm5_control* ctrl, const float* dt, const float* truth);
/* estimate_error (pf_model_host.h:56-69) on robotarm5 states (9 floats). */
float esth_robotarm5_estimate_error(const float* estimate, const float* actual);

/* Trace rows of the reference unicycle format (16 columns,
 * particle_filter.cu:756-772): 10 sensor ranges, velocity, angular velocity,
 * dt, truth x, y, theta.  truth receives 3 floats per row. */
int esth_trace_load_unicycle10(const char* path, int max_rows, esth_unicycle10_measurement* meas,
                               esth_unicycle_control* ctrl, float* dt, float* truth, int* rows);
/* Write `rows` rows in the reference unicycle format (16 columns, 9
 * significant digits; truth 3 floats per row). */
int esth_trace_write_unicycle10(const char* path, int rows, const esth_unicycle10_measurement* meas,
                                const esth_unicycle_control* ctrl, const float* dt, const float* truth);
/* The reference unicycle program's loop (particle_filter.cu:729-1157) on the
 * engine: particles start at (20, 40, -pi/2) (:655-663); writes the mean over
 * steps of the distance between the best particle and the truth (:1044-1057). */
int esth_unicycle10_filter(uint32_t m, uint32_t G, uint32_t t, float thr, const char* trace_path,
                           uint32_t seed, int resampler, double* mean_error, int* steps);

/* Seeded synthetic scenarios.
 * unicycle: dt, v=1, omega_t = 0.2 sin(0.05 t); landmarks at (+-10,+-10) when
 * num_landmarks == 4 and box == 0, else uniform in [-box, box]^2 from `seed`.
 * truth: steps*3 floats (state after each step, starting from x0 = 0,0,0);
 * measurements use the given noise levels; process noise is added to truth. */
int esth_unicycle_scenario(uint32_t seed, int num_landmarks, float box, int steps, float dt,
                           esth_unicycle_params* params, esth_unicycle_control* controls,
                           esth_unicycle_measurement* meas, float* truth);
/* dh6: DH parameters drawn from `seed` (a in [0.2,0.6], d in [-0.1,0.1],
 * twists pi/2, 0, 0, pi/2, -pi/2, 0); truth (steps+1)*12 floats: row 0 the
 * initial q, qd (also drawn from `seed`), row k the state after step k. */
int esth_dh6_scenario(uint32_t seed, int steps, float dt, esth_dh6_params* params,
                      esth_dh6_measurement* meas, float* truth);
/* robotarm5_camera: the arm's joints start at 0.75 pi and follow smooth
 * controls (|rate| <= 0.2 rad/s) with 0.01 rad process noise per step; the
 * base walks near the origin (|x|, |y| <= 1, speeds <= 0.5) so that the
 * camera sees it; measurements: joint angles with 0.05 rad noise and the
 * base's projection through `params` (NULL: the reference's calibration)
 * with 1 px noise.  truth: (steps+1)*9 floats, row 0 the initial state, row
 * k the state after step k. */
int esth_robotarm5_camera_scenario(uint32_t seed, int steps, float dt, const esth_camera_params* params,
                                   esth_robotarm5_control* controls, esth_robotarm5_measurement* meas,
                                   float* truth);

/* ------------------------------------------------------------ batches
 * Many independent filters stepped together: every stage is one launch over
 * all filters' sub-filters, the whole step captured in one HIP graph.
 * Filters must use ESTH_RNG_MTGP_PER_FILTER and m >= 64. */
typedef struct esth_batch esth_batch;
int esth_batch_create(esth_engine* e, int num_filters, const esth_filter_desc* descs,
                      const void* params, size_t params_stride, esth_batch** out);
int esth_batch_destroy(esth_batch* b);
int esth_batch_init_uniform(esth_batch* b, const float* states /* num_filters x state_dim */);
/* controls / measurements: arrays of the model structs with the given strides. */
int esth_batch_step(esth_batch* b, const void* controls, size_t control_stride,
                    const void* measurements, size_t measurement_stride, const float* dts);
int esth_batch_sync(esth_batch* b);
int esth_batch_estimate(esth_batch* b, int filter, esth_estimate* out);
int esth_batch_read_particles(esth_batch* b, int filter, float* state_soa, float* weights);
/* Replay the step's main-stream work as a captured HIP graph (default on). */
int esth_batch_use_graph(esth_batch* b, int enable);
/* Device time of the last step in microseconds. */
int esth_batch_step_time(esth_batch* b, float* us);
/* Device times of every step since the previous call (at most `max`, up to
 * 8192 kept), microseconds; resets the record. */
int esth_batch_step_times(esth_batch* b, float* us, int max, int* count);
/* Totals of the batch: particles, MTGP32 streams, sub-filters. */
int esth_batch_info(esth_batch* b, uint64_t* particles, int* streams, uint32_t* sub_filters);

/* ------------------------------------------------------------ autotuner
 * Measures, on this device, the batch path's step cost per million particles
 * for sub-filter sizes 64, 128, 256, 512, 1024: a one-class batch of 64
 * filters x 65,536 particles with `base`'s model and settings and the given
 * control/measurement, graph-replayed, median of 10 steps, with sample+weight
 * and the sort as two launches and as one (where the fused kernel fits),
 * then at the faster form with the resampling's other tiling (32 lanes per
 * sub-filter at m = 64, 128 at m >= 256); the cost is the best of these, and
 * the engine's batch fusion and resampling tiling for that size
 * (esth_engine_set_batch_fusion, esth_engine_set_batch_resample_lanes) are
 * set to it.  Sizes the descriptor rejects
 * cost +inf.  esth_autotune_pick then chooses a filter's sub-filter size:
 * among the powers of two in [max(64, min_m), 1024] dividing `particles`,
 * the cheapest (ties: larger); if none is >= min_m, the largest divisor. */
int esth_autotune(esth_engine* e, const esth_filter_desc* base, const void* params, size_t params_size,
                  const void* control, const void* measurement, float dt, float us_per_mparticle[5]);
int esth_autotune_pick(const float us_per_mparticle[5], uint32_t particles, uint32_t min_m, uint32_t* m_out);
/* Measures, on this device, the device time per step in microseconds of a
 * filter shaped as `base` (its m, G, model and settings) with each state
 * layout SOA, AOS, PACKED (10 whole steps after 3 warm-up steps);
 * *best receives the fastest layout. */
int esth_autotune_layout(esth_engine* e, const esth_filter_desc* base, const void* params, size_t params_size,
                         const void* control, const void* measurement, float dt, float us_per_step[3],
                         int* best);
/* Calibration, no counterpart in the reference: the device time in
 * microseconds of the model's sample+weight arithmetic alone (Box-Muller,
 * propagation and log-weight of each particle pair, on states and uniforms
 * made up in registers from the pair index; one 4-byte store per pair and no
 * other memory access) over n particles (even, at most 2^32), with `frame`
 * (ESTH_FRAME_FLOATS floats, esth_model_build_frame) as a kernel argument;
 * the median of `reps` launches after one untimed launch.  This is the time
 * the sample+weight launches' arithmetic takes with every memory access
 * hidden: their floor on this device. */
int esth_engine_time_hook(esth_engine* e, int model, const float* frame, uint64_t n, int reps, float* us);

/* ------------------------------------------------------------ serving
 * A request loop with continuous batching.  Tracks (filters) are opened,
 * updated and closed by requests, queued per track; esth_server_pump() forms
 * a wave from the oldest queued update of every admitted track and launches it
 * while up to `max_inflight_waves` earlier waves are still on the device (the
 * RNG of one wave overlaps the step of the previous one).  Particles, weights
 * and MTGP32 statuses stay resident in the server's HBM arena from admission
 * to close; admission (FIFO) waits for a track slot and arena space, which a
 * close frees once its last wave has completed.  A track's results are
 * bit-identical to the same filter stepped alone (ESTH_RNG_MTGP_PER_FILTER,
 * m >= 64), whatever else shares its waves.  Not thread-safe: one thread
 * drives a server. */
typedef struct esth_server esth_server;
typedef struct {
    uint64_t request_id;
    uint64_t track_id;
    int      status;         /* ESTH_OK, or why the request failed */
    int      kind;           /* 0 open, 1 update, 2 close */
    double   submit_us;      /* host clock (steady, microseconds) at submit */
    double   done_us;        /* host clock when the result was collected */
    esth_estimate estimate;  /* for updates */
} esth_result;
/* hbm_bytes: the arena for resident tracks, taken from the engine's budget. */
int esth_server_create(esth_engine* e, int max_tracks, size_t hbm_bytes, esth_server** out);
int esth_server_destroy(esth_server* s);
/* waves in flight 1..3 (default 2); tracks per wave (default unbounded) */
int esth_server_configure(esth_server* s, int max_inflight_waves, int max_wave_tracks);
/* queue a track; its result (kind 0) comes when it has been started on the device */
int esth_server_open(esth_server* s, const esth_filter_desc* desc, const void* params,
                     size_t params_size, const float* init_state, uint64_t* track_id,
                     uint64_t* request_id);
/* queue one step of a track; the control and measurement are copied */
int esth_server_update(esth_server* s, uint64_t track_id, const void* control,
                       const void* measurement, float dt, uint64_t* request_id);
/* queue the close of a track after its queued updates; further updates fail */
int esth_server_close(esth_server* s, uint64_t track_id, uint64_t* request_id);
/* collect finished waves and launch what is queued and fits (non-blocking) */
int esth_server_pump(esth_server* s);
/* take up to max_results collected results, oldest first */
int esth_server_poll(esth_server* s, esth_result* out, int max_results, int* n_out);
/* pump until everything that can run has completed */
int esth_server_drain(esth_server* s);
int esth_server_stats(esth_server* s, size_t* hbm_used, int* open_tracks, int* queued);
int esth_server_wave_stats(esth_server* s, uint64_t* waves, uint64_t* wave_updates, size_t* hbm_peak);
/* particles of an admitted track (waits for the device) */
int esth_server_read_particles(esth_server* s, uint64_t track_id, float* state_soa, float* weights);

/* A seeded request stream driven against a server (or, with `plain`, against
 * one esth_filter per track with per-request launches and a blocking estimate
 * per update).  rate > 0: Poisson arrivals over duration_s, a share
 * open_fraction opening tracks (unicycle with probability unicycle_fraction,
 * else the robot arm, particle counts log-uniform in [min_particles,
 * max_particles] rounded to multiples of 64), the rest updating a uniformly
 * chosen open track, each track closing after a geometric number of updates
 * of mean mean_lifetime (0: never).  rate == 0: closed loop, initial_tracks
 * tracks each with one update always queued.  Latency of an update: from its
 * arrival (closed loop: its submit) to its collected result.  Unicycle tracks
 * replay eight seeded scenarios with `landmarks` landmarks; robot-arm tracks
 * replay the given trace.  Sub-filter size: esth_autotune_pick over
 * us_per_mparticle when given, else the largest power of two <= 256 dividing
 * the count. */
typedef struct {
    double   rate;
    double   duration_s;
    double   open_fraction;
    double   unicycle_fraction;
    uint32_t min_particles, max_particles;
    double   mean_lifetime;
    int      initial_tracks;
    int      landmarks;
    int      topology, num_transfer, resampler, weights;
    uint32_t seed;
    int      plain;
    size_t   hbm_bytes;
    int      max_tracks;
    int      max_inflight;
    const float* us_per_mparticle;
    const esth_robotarm5_control*     ra_controls;
    const esth_robotarm5_measurement* ra_measurements;
    const float* ra_dt;
    int      ra_steps;
} esth_serve_workload;
typedef struct {
    uint64_t requests, opens, updates, closes, failed, backlog;
    double   wall_s, updates_per_s, particle_updates_per_s, mean_particles;
    double   p50_us, p99_us, mean_us, max_us;
    int      peak_tracks;
    size_t   hbm_peak;
    uint64_t waves;
    double   mean_wave_updates;
} esth_serve_report;
int esth_serve_simulate(esth_engine* e, const esth_serve_workload* w, esth_serve_report* out);

/* ------------------------------------------------- device memory helpers */
int esth_set_device(int device);
int esth_dev_alloc(size_t bytes, void** dptr);
int esth_dev_free(void* dptr);
int esth_memcpy_htod(void* dst, const void* src, size_t bytes);
int esth_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int esth_memset_d(void* dst, int value, size_t bytes);
int esth_dev_sync(void);

/* ------------------------------------------------------------- stages
 * Each stage on caller-provided device buffers, on the null stream, for
 * stage-by-stage parity tests and benchmarks.  N = m*G. */

/* MTGP32-11213 parameter table: 200 sets of 52 words
 * (pos, sh1, sh2, mask, tbl[16], tmp_tbl[16], flt_tmp_tbl[16]). */
const uint32_t* esth_mtgp_params(void);
/* mtgp32_init_state (mtgp32-fast.c:119-137) for parameter set `set`: 351 words. */
int esth_mtgp_seed_state(int set, uint32_t seed, uint32_t* state351);
/* mtgp32_single_kernel / mtgp32_uint32_kernel (mtgp32-opencl.cl:125-242):
 * streams j = 0..nstreams-1 use parameter set first_set + j, state
 * d_status[j*351 .. +351], write `size` words to d_out[j*size ..];
 * size must be a multiple of 768.  float_mode: 1 floats in [1,2), 0 uint32. */
int esth_stage_mtgp(uint32_t* d_status, int nstreams, int first_set, int size, void* d_out,
                    int float_mode);
/* `nsteps` consecutive launches of esth_stage_mtgp in one (how a filter
 * generates a block of steps' words): step s's words of stream j at
 * d_out + (s * nstreams + j) * size, and the status after step s at
 * d_status_out + s * nstreams * 351; d_status_in (nstreams * 351 words) must
 * not overlap d_status_out.  Words and statuses equal nsteps calls of
 * esth_stage_mtgp (mtgp32-opencl.cl:186-242; pf_generic.c:191-232 per step). */
int esth_stage_mtgp_steps(const uint32_t* d_status_in, uint32_t* d_status_out, int nstreams, int first_set,
                          int size, int nsteps, void* d_out, int float_mode);
/* boxmuller (pf_generic.cl:465-483), in place on the first n floats. */
int esth_stage_boxmuller(float* d_data, int n);
/* sampling_importance (pf_generic.cl:206-250) with Box-Muller fused: particle
 * i uses normals d_random[k*N + i] after the reference's in-place transform of
 * the uniform pairs.  d_state: state_dim planes of N; d_w: N (log or linear). */
int esth_stage_sample_weight(int model, float* d_state, uint32_t N, const float* d_random,
                             const float* frame, int weights, float* d_w);
/* block_sort (pf_generic.cl:252-291): sorts d_w per sub-filter descending in
 * place with the reference's bitonic network, gathers state planes
 * d_state_in -> d_state_out, writes d_winner[g] = (w[0], g) as float2 and,
 * if d_perm != NULL, the permutation. */
int esth_stage_sort(float* d_w, const float* d_state_in, float* d_state_out, int state_dim,
                    uint32_t m, uint32_t G, float* d_winner, int* d_perm);
/* Allocation-free variant of esth_stage_sort.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same inputs, outputs and results as esth_stage_sort; the kernel's scratch
 * (per-sub-filter partial sums and, when d_winner is NULL, the winners) lives
 * in a 2 MiB array that each GPU holds for the life of the process, so it
 * needs G * (2 + S) floats (+ 2 G without d_winner) <= 524288 and fails with
 * ESTH_ERR_BUDGET beyond that.  Enqueued on `stream` (a hipStream_t; NULL =
 * the null stream) without waiting; it calls no allocator.  Calls that
 * overlap on one GPU share that GPU's array and race. */
int esth_stage_sort_noalloc(float* d_w, const float* d_state_in, float* d_state_out, int state_dim,
                            uint32_t m, uint32_t G, float* d_winner, int* d_perm, void* stream);
/* reduce (reduce.cl:19-55): argmax over G float2 winners -> d_result float2. */
int esth_stage_argmax(const float* d_winner, uint32_t G, int weights, float* d_result);
/* exchange of best particles on sorted sub-filters (in place on d_state/d_w). */
int esth_stage_exchange(float* d_state, float* d_w, int state_dim, uint32_t m, uint32_t G,
                        uint32_t t, int topology, int dimx, int dimy);
/* resampling (pf_generic.cl:385-432) / systematic: gate uniform
 * d_random[N + g], selection uniforms d_random[i] (values in [1,2)).
 * d_state_src: sorted state; d_state_dst: main state; d_sel (optional):
 * selected source index per particle within its sub-filter. */
int esth_stage_resample(int resampler, float* d_state_dst, const float* d_state_src,
                        int state_dim, const float* d_w, const float* d_random, uint32_t m,
                        uint32_t G, float threshold, int adaptive, int weights, int* d_sel);
/* Vose alias tables only (resampling-valias.cl:111-195) for parity tests:
 * per sub-filter prob[m] (scaled, after pairing) and alias[m]. */
int esth_stage_alias_table(const float* d_w, uint32_t m, uint32_t G, int weights, float* d_prob,
                           uint32_t* d_alias);
/* Plain device copy of `bytes` (multiple of 16): a kernel of known traffic
 * used to calibrate HBM byte counters. */
int esth_stage_copy(void* d_dst, const void* d_src, size_t bytes);
/* Deterministic math functions evaluated on the device (n inputs): fn 0 exp,
 * 1 log, 2 sin, 3 cos, 4 atan2(y=x[i], x=x2[i]), 5 length and 6 |angle| of the
 * vector (x[i], x2[i]) (the unicycle likelihood's division-free forms),
 * 7 reciprocal. */
int esth_stage_math(int fn, const float* d_x, const float* d_x2, float* d_out, int n);

#ifdef __cplusplus
}
#endif

#endif /* ESTHERA_H */
