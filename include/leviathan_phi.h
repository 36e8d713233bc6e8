/* leviathan_phi.h -- C-ABI of the hub transition statistics and the
 * bipartition measure phi (part of libleviathan_hip.so).
 *
 * The first block is the reference's own FFI surface: the extern "C" host API
 * of src/cuda/leviathan_phi_kernel.cu:240-369, which
 * src/bindings/leviathan_phi_pybind.cpp:8-16 declares and binds.  Names,
 * argument order and meaning are unchanged.  Differences:
 *
 *   - errors are reported through leviathan_last_error() (the reference calls
 *     exit(1)); phi_init returns NULL, phi_compute NaN and
 *     phi_get_transition_count -1 on failure;
 *   - phi_init also rejects num_hubs < 1 and negative hub indices (the
 *     reference computes 1 << (num_hubs - 1) unchecked);
 *   - phi_compute works in FP64 and returns the float nearest the result.
 *
 * The hub state is H bits, bit i set when hub i's phase, reduced into
 * [0, 2*pi), exceeds pi (FP32, leviathan_phi_kernel.cu:69-79); each call after
 * the first records the transition previous state -> state in an S x S table
 * of int64 counts (S = 2^H).  phi_compute forms the state-by-node table
 * P(hub n on at t+1 | state s at t) (a state never left gets 0.5) and, for
 * every bipartition of the hubs, phi = mean over states of
 * |log p_int - log(p_A * p_B)|, and returns the minimum (the reference's
 * formula, :143-235).  As the reference defines them, p_A * p_B multiplies
 * the same clamped factors as p_int, so phi is zero in exact arithmetic for
 * every bipartition: the reference returns the FP32 rounding residue of the
 * two products (~1e-7), this library the FP64 one (~1e-15).  The counts and
 * the state-by-node table are the information the statistics carry; the
 * extensions below return them.
 *
 * Threading: like LeviathanHandle.  A PhiHandle has its own HIP stream; every
 * call returns after its work is done.  d_ names are device pointers.
 */
#ifndef LEVIATHAN_PHI_H_
#define LEVIATHAN_PHI_H_

#include <stdint.h>

#include "leviathan.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct PhiHandle PhiHandle;

/* ---------------------------------------------------------------------------
 * Reference surface (leviathan_phi_kernel.cu:240-369)
 * ------------------------------------------------------------------------- */

/* leviathan_phi_kernel.cu:242-288.  1 <= num_hubs <= 10 hubs, given by node
 * index; counts start at zero and no state has been seen. */
PhiHandle* phi_init(int num_hubs, int* h_hub_indices);

/* leviathan_phi_kernel.cu:290-299.  Records the hub state of the phase vector
 * d_theta (device memory, indexed by node) and the transition from the
 * previous one.  Rejected while the handle is attached to an engine. */
void phi_accumulate(PhiHandle* handle, const float* d_theta);

/* leviathan_phi_kernel.cu:301-338.  phi of the minimum-information
 * bipartition; 1e30 when there is none (one hub). */
float phi_compute(PhiHandle* handle);

/* leviathan_phi_kernel.cu:340-348.  Zero counts; no state seen; no calls. */
void phi_reset(PhiHandle* handle);

/* leviathan_phi_kernel.cu:350-352.  States recorded since the last reset
 * (phi_accumulate calls plus attached engine steps), as the reference's
 * transition_counter counts them. */
int64_t phi_get_transition_count(PhiHandle* handle);

/* leviathan_phi_kernel.cu:354-368.  Detaches the handle first if attached. */
void phi_free(PhiHandle* handle);

/* ---------------------------------------------------------------------------
 * Extensions (no counterpart in the reference)
 * ------------------------------------------------------------------------- */

/* Attach the handle to an engine: from then on every step of the engine
 * (leviathan_step, leviathan_step_n, captured graphs included) records the
 * hub state of its new phases, on the engine's stream, in the kernel that
 * finishes the step -- the phases never leave the device.  This replaces the
 * reference's per-step host round trip (get_phase_snapshot, then
 * GPUPhiWorker.accumulate_from_host, leviathan_phi_pybind.cpp:35-56, which
 * uploads all N phases).  One handle per engine; phi == NULL detaches.  Hub
 * indices must be < the engine's node count.  Returns 0 on success. */
int leviathan_attach_phi(LeviathanHandle* engine, PhiHandle* phi);

/* GPUPhiWorker.accumulate_from_host (leviathan_phi_pybind.cpp:35-56) for a
 * host phase vector of n nodes: uploads only the num_hubs hub phases and
 * records their state as phi_accumulate does.  Returns 0 on success. */
int phi_accumulate_host(PhiHandle* handle, const float* h_theta, int64_t n);

/* Counts (S*S, row-major by previous state) and row totals (S) to host
 * memory; either pointer may be NULL.  Returns 0 on success. */
int phi_get_counts(PhiHandle* handle, int64_t* h_counts, int64_t* h_row_totals);

/* The last recorded state (-1 if none) and the number of states recorded. */
int phi_get_state(PhiHandle* handle, int* h_state, int64_t* h_calls);

/* Restores a checkpointed state: counts, row totals, last state (-1: none)
 * and the number of states recorded.  Returns 0 on success. */
int phi_set_state(PhiHandle* handle, const int64_t* h_counts, const int64_t* h_row_totals, int state,
                  int64_t calls);

/* phi_compute in FP64, plus the mask of the minimum-information bipartition
 * (hub n in part A when bit n is set; the first minimum; 0 when there is
 * none), the state-by-node table (S*num_hubs doubles, row-major by state) and
 * the per-bipartition values (2^(num_hubs-1) - 1 doubles, bipartition p has
 * mask p + 1).  Output pointers other than h_phi may be NULL. */
int phi_compute_ex(PhiHandle* handle, double* h_phi, int* h_mip_mask, double* h_sbn_tpm,
                   double* h_partitions);

#ifdef __cplusplus
}
#endif

#endif /* LEVIATHAN_PHI_H_ */
