/* leviathan.h -- C-ABI of the gfx950 LEVIATHAN step library (libleviathan_hip.so).
 *
 * The first block of entry points is the reference's own FFI surface: the
 * extern "C" host API of src/cuda/leviathan_kernel.cu:270-497, which
 * src/bindings/leviathan_pybind11.cpp:10-25 declares and binds.  Names,
 * argument order and argument meaning are unchanged, so a caller of the
 * reference links against this library instead.  Two deliberate differences:
 *
 *   - the reference calls exit(1) on any runtime error (leviathan_kernel.cu:11-19);
 *     here every entry point records a message that leviathan_last_error()
 *     returns (an empty string after a call that succeeded), init returns
 *     NULL and step returns NaN on failure;
 *   - inputs are validated (CSR monotone, columns in range, delays <=
 *     max_delay) instead of being trusted.
 *
 * The second block (leviathan_*_ex, step_n, getters, ensembles, profiling,
 * reservoir) has no reference counterpart; each says so.
 *
 * Threading: a handle is not re-entrant (like the reference's).  Every call
 * runs on the handle's own HIP stream and returns after that stream has
 * finished the call's work.  Plain pointers and sizes only: all arrays are
 * host memory unless a name says otherwise.
 */
#ifndef LEVIATHAN_H_
#define LEVIATHAN_H_

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct LeviathanHandle LeviathanHandle;

/* ---------------------------------------------------------------------------
 * Reference surface (leviathan_kernel.cu:270-497)
 * ------------------------------------------------------------------------- */

/* leviathan_kernel.cu:270-389.  Deep-copies the incoming-edge CSR graph
 * (row_ptr[N+1], col_idx[E], delays[E], E = row_ptr[N]) and the initial state
 * (weights[E], theta[N], theta_hat[N], omega[N]; omega_hat starts as omega)
 * to the GPU, with the parameter defaults of :281-297 and FP32 history.  The
 * history layout is chosen automatically (LEV_LAYOUT_AUTO); it is internal and
 * does not change any result.  Returns NULL on failure. */
LeviathanHandle* leviathan_init(int N, int max_delay, const int* h_row_ptr,
                                const int* h_col_idx, const uint8_t* h_delays,
                                const float* h_weights, const float* h_theta,
                                const float* h_theta_hat, const float* h_omega);

/* leviathan_kernel.cu:391-438.  One step; returns the order parameter r of the
 * new phases (NaN on failure).  The gain controller update of :430-435 runs
 * on the GPU in FP64. */
float leviathan_step(LeviathanHandle* handle, float dt);

/* leviathan_kernel.cu:440-443.  Copies theta[N] to host memory. */
void leviathan_get_theta(LeviathanHandle* handle, float* h_theta);

/* leviathan_kernel.cu:445-453.  Uploads theta[N] and re-fills every history
 * slot with it (the delay memory is erased; theta_hat, omega_hat untouched). */
void leviathan_set_theta(LeviathanHandle* handle, const float* h_theta);

/* leviathan_kernel.cu:455-478. */
void leviathan_free(LeviathanHandle* handle);

/* leviathan_kernel.cu:480, :482, :484-486 (network 0 of an ensemble). */
void leviathan_set_gain(LeviathanHandle* handle, float g);
float leviathan_get_gain(LeviathanHandle* handle);
void leviathan_set_gain_controller(LeviathanHandle* handle, bool enabled);

/* leviathan_kernel.cu:488-490.  Bytes of device memory this handle holds
 * (every allocation counted, unlike :299-310 which misses the partials). */
size_t leviathan_get_vram_usage(LeviathanHandle* handle);

/* leviathan_kernel.cu:492-497.  Uploads weights[E]. */
void leviathan_reset_weights(LeviathanHandle* handle, const float* h_weights);

/* ---------------------------------------------------------------------------
 * Extensions (no counterpart in the reference)
 * ------------------------------------------------------------------------- */

/* Message of the last failed call on this thread; "" after a success. */
const char* leviathan_last_error(void);

enum {
  LEV_HISTORY_F32 = 0, /* history ring stores FP32 phases */
  LEV_HISTORY_U16 = 1  /* exact-wrap uint16 angles, step 2*pi/65536 */
};
enum {
  LEV_LAYOUT_NODE_MAJOR = 0, /* ring[node][slot] */
  LEV_LAYOUT_SLOT_MAJOR = 1, /* ring[slot][node] */
  LEV_LAYOUT_AUTO = 2        /* the default: node-major for the direct schedule on a graph without
                                delay-0 edges (paired row passes then read a node's two delayed
                                phases from one line) and with little index locality (under half
                                of the sources within 256 nodes of their rows); slot-major
                                otherwise */
};

typedef struct {
  /* physics (defaults: leviathan_kernel.cu:281-297, g_rest = 1.5 from :433) */
  float eta, lam, gamma, alpha_eps, kappa_w, w_min, w_max;
  float g0, r0, beta, mu, g_rest, g_min, g_max;
  int32_t gain_controller; /* 0/1 */
  /* storage and schedule */
  int32_t history;         /* LEV_HISTORY_* */
  int32_t layout;          /* LEV_LAYOUT_* */
  int32_t edges_per_thread;/* 4, 8 or 16: a 256-lane block streams up to 256*this edges */
  int32_t graph_chunk;     /* steps per captured HIP graph in leviathan_step_n (0: no graphs) */
  int32_t schedule;        /* LEV_SCHEDULE_* */
  int32_t bucket_edges;    /* binned schedule: max incoming edges per bucket (0: automatic) */
} LeviathanConfig;

enum {
  LEV_SCHEDULE_DIRECT = 0, /* one fused kernel; delayed phases gathered per edge */
  LEV_SCHEDULE_BINNED = 1, /* two phases: source-ordered scatter into ~1 MB destination buckets,
                              then the fused row update reading them (graphs without locality) */
  LEV_SCHEDULE_AUTO = 2,   /* every plan is built; leviathan_autotune picks */
  LEV_SCHEDULE_STAGED = 3  /* two phases: source-ordered scatter into each row block's own
                              range, which the fused row update stages through LDS */
};

/* Plans a handle can hold: direct, binned, staged x 4, 8, 16 edges per lane. */
#define LEV_NUM_PLANS 9

/* Fills cfg with the defaults above (FP32 node-major history, 8 edges per
 * thread, 100-step graph chunks, direct schedule). */
void leviathan_config_default(LeviathanConfig* cfg);

/* Ensemble constructor: n_nets independent networks stored as one
 * block-diagonal CSR (node ranges net_offsets[n_nets+1], net_offsets[0] = 0,
 * net_offsets[n_nets] = N; no edge may cross a range).  Per-network initial
 * gains g0[n_nets] (NULL: cfg->g0 for all) and controller flags
 * controller[n_nets] (NULL: cfg->gain_controller for all).  All networks
 * share max_delay, dt and the physics parameters and advance in one fused
 * launch per step.  n_nets = 1 with net_offsets = {0, N} is leviathan_init. */
LeviathanHandle* leviathan_init_ex(const LeviathanConfig* cfg, int N, int max_delay,
                                   const int* h_row_ptr, const int* h_col_idx,
                                   const uint8_t* h_delays, const float* h_weights,
                                   const float* h_theta, const float* h_theta_hat,
                                   const float* h_omega, int n_nets,
                                   const int* net_offsets, const float* g0,
                                   const int* controller);

/* n steps with one host synchronisation at the end (HIP-graph chunks when
 * cfg.graph_chunk > 0).  r_out[n * n_nets] (may be NULL) receives r of
 * step t for network k at r_out[t * n_nets + k].  Returns 0 or -1. */
int leviathan_step_n(LeviathanHandle* handle, float dt, int n, float* r_out);

/* Ensemble accessors: per-network gain and controller flag. */
int leviathan_set_gain_net(LeviathanHandle* handle, int net, float g);
float leviathan_get_gain_net(LeviathanHandle* handle, int net);
int leviathan_set_gain_controller_net(LeviathanHandle* handle, int net, bool enabled);
int leviathan_num_nets(LeviathanHandle* handle);

/* State export: weights[E]; theta_hat[N] and omega_hat[N]; theta(t - delay)
 * for every node as stored in the history ring (decoded to radians), for
 * 0 <= delay <= max_delay.  Return 0 or -1. */
int leviathan_get_weights(LeviathanHandle* handle, float* h_weights);
int leviathan_get_predictor(LeviathanHandle* handle, float* h_theta_hat, float* h_omega_hat);
int leviathan_get_history(LeviathanHandle* handle, int delay, float* h_out);

/* Sizes and the number of steps taken so far. */
int leviathan_num_nodes(LeviathanHandle* handle);
int64_t leviathan_num_edges(LeviathanHandle* handle);
int64_t leviathan_step_count(LeviathanHandle* handle);

/* Full-state restore (a checkpoint is the getters' output: theta, the
 * predictor, the weights, every history slot, the FP64 gains and the step
 * count).  No counterpart in the reference, whose only state setter is
 * set_theta, which erases the history (leviathan_kernel.cu:445-453).  Restore
 * in this order: set_step_count, reset_weights, set_theta (which fills every
 * history slot with theta), set_predictor, set_history for delays 1..D,
 * set_gains; the engine then continues exactly as the one saved.
 *   leviathan_set_history writes the value theta(t - delay) seen by a delay
 *   (1 <= delay <= max_delay) in the next step, rounded to the history format;
 *   leviathan_get_gains / set_gains move each network's FP64 gain (n_nets
 *   values; get_gain rounds to float);
 *   leviathan_set_step_count sets the step counter, which selects the history
 *   slots. */
int leviathan_set_predictor(LeviathanHandle* handle, const float* h_theta_hat, const float* h_omega_hat);
int leviathan_set_history(LeviathanHandle* handle, int delay, const float* h_values);
int leviathan_get_gains(LeviathanHandle* handle, double* g_out);
int leviathan_set_gains(LeviathanHandle* handle, const double* g);
int leviathan_set_step_count(LeviathanHandle* handle, int64_t step);
int leviathan_num_blocks(LeviathanHandle* handle);

/* Runs n eager steps and times them with HIP events on the handle's stream.
 * out[0] = mean duration of the row pass, the step kernel (ms per launch),
 * out[1] = mean duration of the per-step finalize (chunk_sum_kernel, launched
 *          for a single large graph, plus finalize_kernel; ms per step),
 * out[2] = wall time of the n steps between the first and last event (ms),
 * out[3] = mean duration of the scatter kernel (two-phase schedules; 0 otherwise),
 * out[4] = mean duration of the regroup kernel (staged schedule; 0 otherwise).
 * Returns 0 or -1. */
int leviathan_profile_steps(LeviathanHandle* handle, float dt, int n, double* out);

/* Schedule control (a handle created with LEV_SCHEDULE_BINNED or _STAGED
 * holds that schedule's plans and the direct ones, with _AUTO all of them;
 * each has 4, 8 and 16 edges per lane in the row pass).
 * leviathan_set_schedule switches between them; leviathan_autotune times
 * `steps` steps of every built plan from the same state, restores the state,
 * keeps the fastest and writes the mean milliseconds per step to
 * ms_out[LEV_NUM_PLANS] (direct/4, direct/8, direct/16, binned/4, binned/8,
 * binned/16, staged/4, staged/8, staged/16; NaN for a plan not built).  All plans give
 * the same phases and weights (the per-row arithmetic and summation order
 * are shared); r differs only in the grouping of its FP64 partial sums.
 * Return 0 or -1. */
int leviathan_set_schedule(LeviathanHandle* handle, int schedule, int edges_per_thread);
int leviathan_get_schedule(LeviathanHandle* handle, int* schedule, int* edges_per_thread);
int leviathan_autotune(LeviathanHandle* handle, float dt, int steps, double* ms_out);

/* Multi-step scatter of the two-phase schedules.  With steps = K > 1, the
 * steps run by leviathan_step_n (in whole blocks of K; a remainder, and
 * leviathan_step, run singly) share one scatter pass: the first step of a
 * block gathers the delayed phases of all K steps, and step k of the block
 * first gathers only its edges with delay < k (their phases do not exist
 * before).  The staged schedule regroups the K steps' values in one pass too.
 * Phases, weights and r are bit-identical to K = 1.  K is 1, 2 or 4; the
 * handle keeps K arrays of E values (twice that for the staged schedule).
 * K may not exceed the ring's max_delay + 2 slots.
 * leviathan_get_scatter_steps returns K or -1.
 * leviathan_autotune_scatter_steps times K = 1, 2, 4 on the active plan from
 * the same state (ms per step to ms_out[3]), restores the state and keeps the
 * fastest; with the direct schedule it changes nothing and writes NaN.
 * Return 0 or -1.  (No counterpart in the reference.) */
int leviathan_set_scatter_steps(LeviathanHandle* handle, int steps);
int leviathan_get_scatter_steps(LeviathanHandle* handle);
int leviathan_autotune_scatter_steps(LeviathanHandle* handle, float dt, int steps, double* ms_out);

/* Paired row passes.  With steps = 2, leviathan_step_n runs consecutive steps
 * t, t+1 (t even within a block) as one row pass over the edges and a light
 * pass over the nodes: the row pass runs step t in full and step t+1's
 * prediction, Hebbian update and coupling sum, so each weight is read and
 * written once per two steps; after step t's r and gain update, the node pass
 * integrates step t+1 with the new gain.  Legal when every delay is >= 1
 * (step t+1 then reads no phase step t writes); used for the direct schedule
 * and for two-phase schedules at an even number of steps per scatter pass
 * (otherwise, and for leviathan_step and a remainder, steps run singly).
 * Phases, weights, history, r and gains are bit-identical to single steps.
 * The default is 2 for a graph without delay-0 edges and 1 otherwise; 2 is
 * refused for a graph with delay-0 edges.
 * leviathan_get_row_steps returns the setting or -1.
 * Return 0 or -1.  (No counterpart in the reference.) */
int leviathan_set_row_steps(LeviathanHandle* handle, int steps);
int leviathan_get_row_steps(LeviathanHandle* handle);

/* The history layout a handle holds (LEV_LAYOUT_NODE_MAJOR or
 * LEV_LAYOUT_SLOT_MAJOR; LEV_LAYOUT_AUTO is resolved when the handle is
 * created), or -1.  (No counterpart in the reference.) */
int leviathan_get_layout(LeviathanHandle* handle);

/* Input drive (reservoir): omega_i(t) = omega_i + drive_b[i] * u_t.
 * leviathan_set_drive uploads drive_b[N] (NULL switches the drive off);
 * leviathan_step_n_driven runs n steps with u_t = u[t]. */
int leviathan_set_drive(LeviathanHandle* handle, const float* h_drive_b);
int leviathan_step_n_driven(LeviathanHandle* handle, float dt, int n, const float* h_u, float* r_out);

/* ---------------------------------------------------------------------------
 * Reservoir readout support (no counterpart in the reference:
 * python/reservoir.py:45-60 copies every phase to the host each step and
 * python/reservoir.py:81-82 fits sklearn's Ridge on the CPU).  Features are
 * cos(theta) and sin(theta) of n_nodes chosen nodes, captured on the device
 * after every step into a row-major matrix X[rows][2 * n_nodes] (FP32,
 * cos block first).
 * ------------------------------------------------------------------------- */

/* Chooses the feature nodes and reserves max_rows rows (device memory);
 * clears the captured rows.  Returns 0 or -1. */
int leviathan_features_setup(LeviathanHandle* handle, const int* nodes, int n_nodes, int max_rows);

/* n steps that each append one feature row.  u (may be NULL) is the drive
 * input of each step (leviathan_set_drive must have been called when it is
 * given).  Returns 0 or -1; fails if max_rows would be exceeded. */
int leviathan_step_n_features(LeviathanHandle* handle, float dt, int n, const float* u, float* r_out);

/* Number of captured rows. */
int64_t leviathan_features_rows(LeviathanHandle* handle);

/* Copies rows [row0, row0 + nrows) to out[nrows][2 * n_nodes]. */
int leviathan_features_get(LeviathanHandle* handle, int64_t row0, int64_t nrows, float* out);

/* out[f] = sum_t weights[t] * X[row0 + t][f] over nrows rows, FP64
 * accumulation (weights NULL: all ones).  Column sums, means and X^T y. */
int leviathan_features_colsum(LeviathanHandle* handle, int64_t row0, int64_t nrows, const double* weights,
                              double* out);

/* leviathan_features_colsum without any device allocation: no counterpart in
 * the reference.  For callers that cannot call hipMalloc (serving loops,
 * harnesses that forbid it, the readout's own refinement loop, which calls the
 * moments every iteration).  The staged weights and the output live in a
 * file-scope device array of this library (one per GPU, 576 KiB, held until
 * the process ends) instead of per-call allocations; the partial sums use the
 * handle's buffer from leviathan_features_setup.  Limits: 2 * n_nodes <= 8192
 * features, and nrows <= 65536 when weights are given; larger calls fail with
 * an error naming the limit.  Same results as leviathan_features_colsum (the
 * same kernels).  Every call on one GPU shares its array: calls that overlap
 * on one GPU (two threads, two handles) race; calls one after another do not. */
int leviathan_features_colsum_noalloc(LeviathanHandle* handle, int64_t row0, int64_t nrows, const double* weights,
                                      double* out);

/* out[t] = sum_f (X[row0 + t][f] - mean[f]) * w[f], FP64 accumulation. */
int leviathan_features_matvec(LeviathanHandle* handle, int64_t row0, int64_t nrows, const double* mean,
                              const double* w, double* out);

/* Gram matrix of the centred features over rows [row0, row0 + nrows):
 * gram_out[F][F] = Xc^T Xc with Xc = X - mean, on the matrix cores with FP32
 * accumulation (F = 2 * n_nodes).  terms = 1: Xc rounded to BF16, one BF16
 * product; terms = 3: Xc split into BF16 hi + lo parts and the Gram formed as
 * hi hi^T + hi lo^T + lo hi^T (three BF16 products, ~16-bit inputs).  ms_out
 * (may be NULL) receives the Gram kernel's duration in milliseconds. */
int leviathan_features_gram(LeviathanHandle* handle, int64_t row0, int64_t nrows, const double* mean, int terms,
                            float* gram_out, double* ms_out);

/* Cholesky factor, in FP64 on the device (rocSOLVER), of A = G + (alpha +
 * shift) I, with G the Gram of leviathan_features_gram (same rows, mean and
 * terms 1 or 3), or, for terms = 0, the Gram formed in FP64 (rocBLAS DSYRK on
 * the centred features in FP64), and shift the smallest of 0, 1e-7 s,
 * 1e-6 s, ... (s = the mean of G's diagonal) for which A factorises.  The
 * factor stays in the handle for leviathan_features_precond.  shift_out and
 * gram_ms_out (the Gram's duration) may be NULL. */
int leviathan_features_factor(LeviathanHandle* handle, int64_t row0, int64_t nrows, const double* mean, int terms,
                              double alpha, double* shift_out, double* gram_ms_out);

/* z = A^{-1} r (F doubles in and out, host memory) with the factorisation of
 * the last leviathan_features_factor, on the device: one DSYMV with A^{-1}
 * (formed from the Cholesky factor by rocSOLVER dpotri) for the matrix-core
 * Grams, two triangular solves (dpotrs) for the FP64 Gram. */
int leviathan_features_precond(LeviathanHandle* handle, const double* r, double* z);

#ifdef __cplusplus
}
#endif

#endif /* LEVIATHAN_H_ */
