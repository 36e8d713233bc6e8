/* leviathan_oracle.c -- CPU restatement of LEVIATHAN's discrete step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path links or calls this
 * file: it is the checker for the gfx950 kernels (tests/, smoke(), and the
 * cpu_baseline leg of bench.py) and the FP32 CPU baseline.
 *
 * Three instantiations of one body (oracle_step.inc):
 *   lev_oracle_step_f64  -- FP64 arithmetic, single thread, sequential edge
 *                           sums: the correctness oracle.
 *   lev_oracle_step_f64p -- the same FP64 arithmetic with OpenMP over rows and
 *                           nodes.  Every row's edge sum is still sequential in
 *                           CSR order and every per-node value is computed as
 *                           in lev_oracle_step_f64, so theta, W and g are
 *                           bit-identical to it; only r is summed in a fixed
 *                           blocked order (256-node block sums, then the block
 *                           sums in order), independent of the thread count.
 *                           For workloads too large for one core (the
 *                           10^7-node graph).
 *   lev_oracle_step_f32  -- FP32 arithmetic, OpenMP over rows, mirroring the
 *                           reference's four-kernel FP32 step: the CPU
 *                           baseline ("port") timed beside the GPU.
 *
 * Both follow the reference's kernel order exactly (src/cuda/leviathan_kernel.cu):
 *   prep_kernel                  :93-111   eps, theta_hat, omega_hat
 *   csr_physics_kernel           :114-167  delayed gather, coupling, Hebbian
 *   integrate_and_reduce_kernel  :170-234  phase update, history write, sums
 *   final_reduce_r_kernel        :237-264  r = |sum e^{i theta}| / N
 *   leviathan_step               :391-438  buf_idx advance, host gain control
 * with the history ring of B = max_delay + 1 node-major slots (:277, :327) and
 * the parameter defaults of :281-297 (supplied by the caller).
 *
 * Extension beyond the reference (no counterpart there): an optional input
 * drive omega_i(t) = omega_i + drive_b[i] * u_t, used by the reservoir readout.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

typedef struct {
  double eta, lam, gamma, alpha_eps, kappa_w;
  double w_min, w_max;
  double r0, beta, mu, g_rest, g_min, g_max;
  int32_t gain_controller;
} lev_oracle_params;

#include <omp.h>

/* Threads the OpenMP restatements use from the next call on (the CPU baseline
 * is timed at one thread and at all of them). */
void lev_oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int lev_oracle_max_threads(void) { return omp_get_max_threads(); }

#define PREFETCH 0

/* FP64 instantiation */
#define REAL double
#define SUFFIX f64
#define SIN sin
#define COS cos
#define FLOOR floor
#define SQRT sqrt
#define TWO_PI_R 6.283185307179586476925286766559
#define USE_OMP 0
#include "oracle_step.inc"
#undef REAL
#undef SUFFIX
#undef SIN
#undef COS
#undef FLOOR
#undef SQRT
#undef TWO_PI_R
#undef USE_OMP

/* FP64 instantiation, OpenMP */
#define REAL double
#define SUFFIX f64p
#define SIN sin
#define COS cos
#define FLOOR floor
#define SQRT sqrt
#define TWO_PI_R 6.283185307179586476925286766559
#define USE_OMP 1
#undef PREFETCH
#define PREFETCH 16
#include "oracle_step.inc"
#undef PREFETCH
#undef REAL
#undef SUFFIX
#undef SIN
#undef COS
#undef FLOOR
#undef SQRT
#undef TWO_PI_R
#undef USE_OMP

#define PREFETCH 0

/* FP32 instantiation */
#define REAL float
#define SUFFIX f32
#define SIN sinf
#define COS cosf
#define FLOOR floorf
#define SQRT sqrtf
#define TWO_PI_R 6.283185307179586476925286766559f
#define USE_OMP 1
#include "oracle_step.inc"
