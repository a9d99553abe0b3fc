/*
 * spc2_hip.h -- C ABI of the MI355X (gfx950) force-summation and stepping
 * library for the chimera and vortex-arrays self-propelled particle models.
 *
 * Plain C: pointers, sizes and status codes, no C++ or torch types.  One HIP
 * stream per handle; a handle is used by one host thread at a time, distinct
 * handles are independent.  Host buffers belong to the caller, device memory
 * to the handle.  State arrays use the reference's layout: AoS [x, y, phi] per
 * particle in double (Definitions.hpp:17-21, kS = 3).
 *
 * Each entry point names the reference interface it replaces (file:line in
 * the reference tree).
 */
#ifndef SPC2_HIP_H
#define SPC2_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum spc2_status {
  SPC2_OK = 0,
  SPC2_ERR_INVALID = 1,     /* bad argument or parameter combination */
  SPC2_ERR_HIP = 2,         /* a HIP runtime call failed */
  SPC2_ERR_NOMEM = 3,       /* device allocation failed */
  SPC2_ERR_UNSUPPORTED = 4, /* valid request this library does not provide */
} spc2_status;

typedef enum spc2_system { SPC2_CHIMERA = 0, SPC2_VORTEX = 1 } spc2_system;

/* Arithmetic of the pair sums.
 *  SPC2_FP64: the reference method in FP64 (per-pair sin/cos/atan2, reference
 *             operation order, reference min-image), the plain GPU form.
 *  SPC2_FP32_KAHAN: pair sums in FP32 with Kahan compensation over cell-local
 *             FP32 coordinates; the phase enters through cos/sin of each
 *             particle's phase (angle-addition identity), so the pair loop has
 *             no transcendentals.  State, wrap and reductions stay FP64.
 *  SPC2_FP64_FACTORED: the SPC2_FP32_KAHAN algorithm carried out in FP64. */
typedef enum spc2_precision { SPC2_FP64 = 0, SPC2_FP32_KAHAN = 1, SPC2_FP64_FACTORED = 2 } spc2_precision;

typedef enum spc2_method { SPC2_SE = 0, SPC2_SRK = 1, SPC2_RK4 = 2 } spc2_method;

/* Neighbour search.  AUTO follows the reference dispatch: linked cells iff
 * rho <= L/4, all pairs otherwise (ChimeraSystem.cpp:90-98,
 * VortexArraysSystemPtr.cpp:113-119).  VERLET keeps a list of the pairs within
 * max(rho, xi_r) + verlet_skin of the positions at its build and rebuilds it
 * before any particle can have moved more than verlet_skin / 2 (the
 * reference's dormant path, ChimeraSystemPtr.cpp:650-927); the neighbour sets
 * at every evaluation are those of the linked cells. */
typedef enum spc2_neighbour {
  SPC2_NEIGH_AUTO = 0,
  SPC2_NEIGH_CELLS = 1,
  SPC2_NEIGH_ALL_PAIRS = 2,
  SPC2_NEIGH_VERLET = 3
} spc2_neighbour;

/* Model and run parameters.  The reference hard-codes these in the engine
 * (SimulationEnginePtr.cpp:82-91, :131-139) and in Definitions.hpp:20-22;
 * here they are run time. */
typedef struct spc2_params {
  int32_t system;    /* spc2_system */
  int32_t precision; /* spc2_precision */
  int32_t neighbour; /* spc2_neighbour */
  int32_t reserved;
  int64_t n;         /* particles (kN) */
  double Lx, Ly;     /* periodic box (kL) */
  double rho;        /* interaction radius */
  /* chimera (ChimeraSystemPtr.hpp:16-23); v0 is also the vortex velocity */
  double v0, sigma, alpha, D;
  /* vortex arrays (VortexArraysSystemPtr.hpp:16-25) */
  double mu_plus, mu_minus, xi_a, xi_r, kappa;
  uint64_t seed;     /* Philox key for initial state, natural frequencies and noise */
  double verlet_skin; /* SPC2_NEIGH_VERLET list skin; 0 selects 0.2 * rho */
} spc2_params;

typedef struct spc2_sim *spc2_sim_t;

/* Fills *p with the chimera defaults of the benchmark workloads (v0 = 0.01,
 * sigma = 1, rho = 0.1, alpha = 1.45, D = 0.01, L = 1, N = 1000, seed 42). */
void spc2_default_params(spc2_params *p);

/* Replaces the system constructors ChimeraSystemPtr::ChimeraSystemPtr
 * (ChimeraSystemPtr.cpp:11-93) and VortexArraysSystemPtr::VortexArraysSystemPtr
 * (VortexArraysSystemPtr.cpp:9-88) together with the engine's state allocation
 * (SimulationEnginePtr.cpp:18-24). */
spc2_status spc2_create(const spc2_params *p, spc2_sim_t *out);
spc2_status spc2_destroy(spc2_sim_t h);

/* Message of the last failure on the calling thread ("" if none). */
const char *spc2_last_error(void);

/* Host <-> device copies of the AoS state (3N doubles); synchronous.
 * Replace the engine's ownership of system_state_ (SimulationEnginePtr.cpp:23)
 * and BroadcastVector (ThreadSharedMemory.cpp:331-347). */
spc2_status spc2_set_state(spc2_sim_t h, const double *state);
spc2_status spc2_get_state(spc2_sim_t h, double *state);

/* Philox-keyed random initial state: x, y ~ U[0, L), phi ~ U[0, 2 pi), and
 * natural frequencies omega ~ N(0, omega_std).  Replaces
 * SimulationEnginePtr::InitializeRandomSystemState (SimulationEnginePtr.cpp:31-50)
 * and the frequency draw of StochasticRungeKuttaStepperPtr
 * (StochasticRungeKuttaStepperPtr.hpp:25-35). */
spc2_status spc2_init_random_state(spc2_sim_t h, double omega_std);

/* Natural frequencies omega_i (N doubles), added to the phase drift as
 * (omega + a_phi) dt (StochasticRungeKuttaStepperPtr.hpp:65). */
spc2_status spc2_set_omega(spc2_sim_t h, const double *omega);

/* k-form right-hand side at the stage s + k_coef * dt * k_prev; k_next (3N,
 * in/out) keeps its phi entry where a particle has no partner.  Replaces
 * ChimeraSystemPtr::EvaluateRhs(state, k_prev, k_next, k_coef, dt)
 * (ChimeraSystemPtr.cpp:116-129) and the VortexArraysSystemPtr one
 * (VortexArraysSystemPtr.cpp:107-120).  k_prev may equal k_next. */
spc2_status spc2_eval_rhs(spc2_sim_t h, const double *k_prev, double k_coef, double dt, double *k_next);

/* Ito-form right-hand side: derivative (3N, in/out) and additional (2N, in/out,
 * [i][0], [i][1] per particle).  Replaces ChimeraSystemPtr::EvaluateRhs(state,
 * derivative, additional_derivative, dt) (ChimeraSystemPtr.cpp:131-144) and
 * VortexArraysSystemPtr::EvaluateRhs (VortexArraysSystemPtr.cpp:122-135). */
spc2_status spc2_eval_rhs_ito(spc2_sim_t h, double *derivative, double *additional);

/* Advances the device-resident state nsteps steps of `method` with time step
 * dt, the engine's periodic wrap fused in (SimulationEnginePtr.cpp:106-115).
 * Step k uses noise of global step step0 + k.  noise == NULL draws Philox
 * normals keyed (seed, step, particle); otherwise noise holds nsteps blocks of
 * N (SE) or 2N (SRK, (U1, U2) per particle) normals.  Replaces
 * StochasticEulerStepperPtr::DoStep (StochasticEulerStepperPtr.hpp:32-76),
 * StochasticRungeKuttaStepperPtr::DoStep (StochasticRungeKuttaStepperPtr.hpp:41-97),
 * RungeKutta4StepperPtr::DoStep (RungeKutta4StepperPtr.cpp:25-65) and
 * PeriodicBoundaryConditions::ApplyPeriodicBoundaryConditions
 * (PeriodicBoundaryConditions.cpp:133-141). */
spc2_status spc2_step(spc2_sim_t h, int32_t method, double dt, int64_t step0, int64_t nsteps, const double *noise);

/* spc2_step with flags.  SPC2_STEP_NO_WRAP leaves the positions unwrapped, as
 * the reference's DoStep does (StochasticRungeKuttaStepperPtr.hpp:41-68): the
 * engine then applies ApplyPeriodicBoundaryConditions itself
 * (SimulationEnginePtr.cpp:111).  Used by the drop-in stepper classes. */
enum { SPC2_STEP_NO_WRAP = 1 };
spc2_status spc2_step_ex(spc2_sim_t h, int32_t method, double dt, int64_t step0, int64_t nsteps, const double *noise,
                         int32_t flags);

/* Asynchronous trajectory frame: the state converted to float AoS [x, y, phi]
 * by particle id (3N floats, the payload of a BinaryObserverPtr::SaveSystemState
 * frame, BinaryObserverPtr.cpp:163-168) copied into `dst` on the handle's stream.
 * dst should come from spc2_host_alloc (pinned) for the copy to overlap.
 * spc2_snapshot_wait blocks until the last snapshot has landed. */
spc2_status spc2_snapshot_async(spc2_sim_t h, float *dst);
spc2_status spc2_snapshot_wait(spc2_sim_t h);

/* Pinned host memory for snapshot rings. */
spc2_status spc2_host_alloc(uint64_t bytes, void **out);
spc2_status spc2_host_free(void *p);

/* The Philox normals spc2_step would draw at global step `step`: N (SE) or 2N
 * (SRK) doubles.  Replaces the stepper's std::normal_distribution draws
 * (StochasticRungeKuttaStepperPtr.hpp:49-56). */
spc2_status spc2_noise(spc2_sim_t h, int32_t method, int64_t step, double *out);

/* Each particle's partner counts at the current state, by particle id: within
 * rho (chimera; vortex arrays: n0 + n1, within rho) and within xi_r (vortex
 * cohesion; 0 for the chimera), the neighborhood_cardinality_ of
 * VortexArraysSystemPtr.cpp:459-489 / the n of ChimeraSystemPtr.cpp:429-433.
 * The FP32-compensated form decides pairs near a cutoff in FP64, so the counts
 * are those of the reference method in every precision. */
spc2_status spc2_neighbour_counts(spc2_sim_t h, int32_t *n_rho, int32_t *n_xi_r);

/* Exact checkpoint of a handle (no counterpart in the reference, whose restart
 * reads a float frame: InitializeSystemStateFromPreviousSolution,
 * SimulationEnginePtr.cpp:52-73).  The blob holds the handle's parameters
 * (seed included), the FP64 state, the natural frequencies, the phase
 * components of the RK4 stepper's k_1..k_4 and, for Verlet handles, the
 * current list's build positions and displacement bound.  Loaded into a handle
 * created with the same parameters, the steps that follow (with the same step
 * indices, which key the Philox noise) are bit-identical to those of the
 * handle that saved it.  Loading fails with SPC2_ERR_INVALID for a blob of
 * other parameters. */
spc2_status spc2_checkpoint_size(spc2_sim_t h, uint64_t *bytes);
spc2_status spc2_checkpoint_save(spc2_sim_t h, void *buf, uint64_t bytes);
spc2_status spc2_checkpoint_load(spc2_sim_t h, const void *buf, uint64_t bytes);

/* spc2_noise without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The normals are staged in a
 * fixed 16 MiB array per GPU, held until the process ends, so a call takes at
 * most 2^21 normals (N <= 2097152 for SE, 1048576 for SRK) and fails with
 * SPC2_ERR_INVALID beyond that.  Calls on one GPU share the array: calls that
 * overlap in time on one GPU (from several threads) race. */
spc2_status spc2_noise_noalloc(spc2_sim_t h, int32_t method, int64_t step, double *out);

/* Kuramoto order parameter R e^{i psi} = (1/N) sum e^{i phi_j}, FP64.
 * Replaces the reduction in BinaryObserverPtr::SaveSummaryStatistics
 * (BinaryObserverPtr.cpp:237-247), which sums in float. */
spc2_status spc2_order_parameter(spc2_sim_t h, double *R, double *psi);

/* The linked-cell list of the current state in CSR form: cell_of (N),
 * cell_start (nx*ny+1) and sorted (N, ascending particle index within each
 * cell).  Replaces LinkedList::ConstructLinkedList + GetLinkedList
 * (LinkedList.cpp:92-134).  Fails with SPC2_ERR_INVALID on the all-pairs path. */
spc2_status spc2_get_cell_list(spc2_sim_t h, int32_t *cell_of, int32_t *cell_start, int32_t *sorted);
spc2_status spc2_num_cells(spc2_sim_t h, int32_t *nx, int32_t *ny);

/* Pair counters of the RHS evaluations since the last reset: candidates tested,
 * pairs within rho, cohesion pairs.  Counting is off unless enabled. */
spc2_status spc2_set_counting(spc2_sim_t h, int32_t enable);
spc2_status spc2_get_counts(spc2_sim_t h, int64_t *counts3);

/* HIP-event timing of the pair kernels on the handle's stream. */
spc2_status spc2_set_profiling(spc2_sim_t h, int32_t enable);
spc2_status spc2_get_profile(spc2_sim_t h, double *pair_kernel_ms, int64_t *launches);

/* Verlet-list builds since creation and the entries of the current list
 * (SPC2_ERR_INVALID unless the handle uses SPC2_NEIGH_VERLET). */
spc2_status spc2_verlet_stats(spc2_sim_t h, int64_t *builds, int64_t *entries);

/* Layout of the current Verlet list: lds_staged = 1 when the FP32 passes stage
 * each workgroup's halo of tiles in LDS (capacity lds_capacity particles) and
 * the entries index that halo, 0 for the gather passes.  No counterpart in the
 * reference (it has no neighbour list); for tests and tuning.  Setting the
 * environment variable SPC2_VERLET_LDS=0 before spc2_create keeps the gather
 * passes. */
spc2_status spc2_verlet_layout(spc2_sim_t h, int32_t *lds_staged, int32_t *lds_capacity);

spc2_status spc2_synchronize(spc2_sim_t h);

/* Free and total bytes of the current device (admission budgets). */
spc2_status spc2_device_memory(uint64_t *free_bytes, uint64_t *total_bytes);

/* ---- Batched replicas ---------------------------------------------------
 * Many independent systems of uneven N stepped together, one launch per phase
 * per step for the whole batch.  Replaces the reference's parameter-scan mode,
 * one parameter set per MPI rank (SimulationEngine.cpp:82-88 under
 * MPI_PARAMETER_SCAN, main.cpp:15-18).  Each replica keeps its own parameters,
 * seed and step counter; its trajectory is bit-identical to a single handle
 * created with the same parameters: noise is keyed by the replica's seed, its
 * own step and the particle id within it.  All replicas share one precision and
 * use the reference's neighbour dispatch (linked cells iff rho <= L/4).
 * Methods: Stochastic Euler (StochasticEulerStepperPtr.hpp:32-76) and the
 * stochastic Runge-Kutta step (StochasticRungeKuttaStepperPtr.hpp:32-97). */
typedef struct spc2_batch *spc2_batch_t;

/* Device bytes one replica takes in a batch (admission control). */
uint64_t spc2_batch_footprint(const spc2_params *p);

/* A batch with the Philox initial state of every replica (natural frequencies
 * zero; see spc2_batch_init_random_state). */
spc2_status spc2_batch_create(const spc2_params *replicas, int32_t count, spc2_batch_t *out);

/* A batch whose replica r continues replica prev_index[r] of `prev` (state,
 * natural frequencies and step counter copied on the device) or, where
 * prev_index[r] < 0, starts from its Philox initial state with natural
 * frequencies N(0, omega_std).  Used to re-pack a running batch. */
spc2_status spc2_batch_create_from(const spc2_params *replicas, int32_t count, spc2_batch_t prev,
                                   const int32_t *prev_index, double omega_std, spc2_batch_t *out);
spc2_status spc2_batch_destroy(spc2_batch_t b);
spc2_status spc2_batch_init_random_state(spc2_batch_t b, double omega_std);
spc2_status spc2_batch_set_state(spc2_batch_t b, int32_t replica, const double *state);
spc2_status spc2_batch_get_state(spc2_batch_t b, int32_t replica, double *state);
spc2_status spc2_batch_set_omega(spc2_batch_t b, int32_t replica, const double *omega);

/* Advances every replica by min(nsteps, limit[r]) steps (limit == NULL: nsteps)
 * with time step dt and method SPC2_SE or SPC2_SRK; replica r's step k uses
 * noise of its own step steps_done[r] + k (SPC2_ERR_UNSUPPORTED: RK4). */
spc2_status spc2_batch_step(spc2_batch_t b, int32_t method, double dt, int64_t nsteps, const int64_t *limit);
spc2_status spc2_batch_steps_done(spc2_batch_t b, int64_t *steps);
spc2_status spc2_batch_order_parameters(spc2_batch_t b, double *R, double *psi);

/* Float frames [x, y, phi] by particle id of every replica, concatenated in
 * replica order (replica r starts at 3 * first_particle[r]). */
spc2_status spc2_batch_snapshot_async(spc2_batch_t b, float *dst);
spc2_status spc2_batch_snapshot_wait(spc2_batch_t b);
spc2_status spc2_batch_layout(spc2_batch_t b, int32_t *first_particle, int32_t *particles);

/* Tile size of each linked replica's pair kernel, in particles per
 * workgroup (the autotuner's knob): rows with at least 8 particles per cell
 * are cut into chunks of that many slots, rounded to whole 64-lane
 * wavefronts; sparser rows into segments of about that many particles. */
spc2_status spc2_batch_set_tile_targets(spc2_batch_t b, const double *particles_per_segment);
spc2_status spc2_batch_get_tile_targets(spc2_batch_t b, double *particles_per_segment);

/* Picks the tile target of each bucket of replicas with the same
 * floor(log2 N) by timing trial_steps steps of a copy of the batch for each
 * candidate (coordinate descent over the buckets) and applies the best to b;
 * b's replicas do not move.  *ms_per_step (nullable) gets the tuned step time.
 * No counterpart in the reference (its parameter scan runs one CPU rank per
 * replica). */
spc2_status spc2_batch_autotune(spc2_batch_t b, double dt, int32_t trial_steps, double *ms_per_step);

#ifdef __cplusplus
}
#endif
#endif
