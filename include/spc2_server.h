/* spc2 serving loop: continuous batching of simulation requests on one GPU.
 *
 * The reference runs one simulation per process (SimulationEnginePtr.cpp:75-122,
 * one parameter set per MPI rank in its parameter-scan mode,
 * SimulationEngine.cpp:82-88); there is no request interface to replace.
 * These entry points have no counterpart in the reference: they are the
 * request loop for callers that submit many independent simulations.
 *
 * Requests are admitted first come, first served while their device
 * footprint fits the budget; admitted requests join the running batch at the
 * next super-step boundary and leave it when done, so requests that arrive
 * while others run overlap with them on the device.  A request's result does
 * not depend on what it ran with (see spc2_batch_create).
 *
 * Built into libspc2_host.so, over the C ABI of spc2_hip.h.
 */
#ifndef SPC2_SERVER_H
#define SPC2_SERVER_H

#include <stdint.h>

#include "spc2_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spc2_server_config {
  uint64_t hbm_budget;     /* device bytes the admitted requests may take (0: 80% of free memory) */
  int64_t max_particles;   /* particles in flight (0: no limit) */
  int32_t max_replicas;    /* requests in flight (0: no limit) */
  int32_t super_steps;     /* steps between re-packs (and snapshots); >= 1 */
  int32_t snapshot_every;  /* steps between streamed frames of each request (0: none) */
  int32_t keep_final;      /* keep each request's final state for spc2_server_wait */
  double dt;               /* time step of every request */
  double omega_std;        /* natural frequencies N(0, omega_std) of chimera requests */
} spc2_server_config;

typedef struct spc2_server_stats {
  int64_t submitted, admitted, completed, queued, in_flight;
  int64_t repacks, frames;
  uint64_t frame_bytes;
  double particle_steps;    /* work of the completed requests */
  double particle_steps_executed;  /* every step taken, finished or not */
  double wall_seconds;      /* since creation */
  double busy_seconds;      /* time spent stepping */
  double latency_p50, latency_p99, latency_max;  /* submit -> final state, completed requests */
  double queue_wait_p50, queue_wait_p99;         /* submit -> admission */
  int64_t max_in_flight, max_particles_in_flight;
} spc2_server_stats;

typedef struct spc2_server *spc2_server_t;

void spc2_server_default_config(spc2_server_config *cfg);
spc2_status spc2_server_create(const spc2_server_config *cfg, spc2_server_t *out);
/* Stops the loop (requests not finished are dropped) and frees the device. */
spc2_status spc2_server_destroy(spc2_server_t s);
/* Queues a Stochastic Euler simulation of `steps` steps from the Philox
 * initial state of p (all requests share the precision of the first). */
spc2_status spc2_server_submit(spc2_server_t s, const spc2_params *p, int64_t steps, int64_t *id);
/* Waits up to timeout_s for request id; copies its final state (3N doubles,
 * when kept) and its submit -> done latency.  SPC2_ERR_INVALID on timeout. */
spc2_status spc2_server_wait(spc2_server_t s, int64_t id, double timeout_s, double *state, double *latency_s);
/* Waits until no request is queued or in flight (timeout_s < 0: forever). */
spc2_status spc2_server_drain(spc2_server_t s, double timeout_s);
spc2_status spc2_server_stats_get(spc2_server_t s, spc2_server_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* SPC2_SERVER_H */
