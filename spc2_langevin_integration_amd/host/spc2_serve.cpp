// Serving-loop load generator: Poisson arrivals of chimera and vortex-arrays
// simulation requests into the continuous-batching server, then one JSON line
// with latency percentiles and throughput.
//
//   spc2_serve [rate=200] [duration=600] [n_min=1000] [n_max=200000]
//              [steps_min=1000] [steps_max=20000] [max_particles=4000000]
//              [max_replicas=0] [super_steps=100] [snapshot_every=100] [seed=42]
//              [drain=0]
//
// Requests: N log-uniform in [n_min, n_max] at density 4000 (L = sqrt(N/4000)),
// steps uniform in [steps_min, steps_max], chimera or vortex arrays with equal
// probability, sigma in [0.2, 2], alpha in [1.2, 1.57], D log-uniform in
// [0.001, 0.1], rho = 0.1, FP32-compensated, Stochastic Euler, dt = 0.01.
// Arrivals stop after `duration` seconds; with drain=0 the run ends there and
// latencies cover the requests completed by then.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <random>
#include <string>
#include <thread>

#include "../../include/spc2_server.h"

namespace {

struct Args {
  double rate = 200.0, duration = 600.0;
  double n_min = 1000, n_max = 200000;
  long steps_min = 1000, steps_max = 20000;
  long long max_particles = 4000000;
  int max_replicas = 0, super_steps = 100, snapshot_every = 100;
  unsigned long long seed = 42;
  int drain = 0;
};

bool parse(Args &a, int argc, char **argv) {
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    const size_t eq = s.find('=');
    if (eq == std::string::npos) return false;
    const std::string k = s.substr(0, eq), v = s.substr(eq + 1);
    if (k == "rate") a.rate = std::stod(v);
    else if (k == "duration") a.duration = std::stod(v);
    else if (k == "n_min") a.n_min = std::stod(v);
    else if (k == "n_max") a.n_max = std::stod(v);
    else if (k == "steps_min") a.steps_min = std::stol(v);
    else if (k == "steps_max") a.steps_max = std::stol(v);
    else if (k == "max_particles") a.max_particles = std::stoll(v);
    else if (k == "max_replicas") a.max_replicas = std::stoi(v);
    else if (k == "super_steps") a.super_steps = std::stoi(v);
    else if (k == "snapshot_every") a.snapshot_every = std::stoi(v);
    else if (k == "seed") a.seed = std::stoull(v);
    else if (k == "drain") a.drain = std::stoi(v);
    else return false;
  }
  return a.rate > 0 && a.duration > 0 && a.n_min >= 1 && a.n_max >= a.n_min && a.steps_min >= 0 &&
         a.steps_max >= a.steps_min;
}

}  // namespace

int main(int argc, char **argv) {
  Args a;
  if (!parse(a, argc, argv)) {
    std::cerr << "usage: spc2_serve [rate=R] [duration=S] [n_min=] [n_max=] [steps_min=] [steps_max=] "
                 "[max_particles=] [max_replicas=] [super_steps=] [snapshot_every=] [seed=] [drain=0|1]\n";
    return 2;
  }
  spc2_server_config cfg;
  spc2_server_default_config(&cfg);
  cfg.max_particles = a.max_particles;
  cfg.max_replicas = a.max_replicas;
  cfg.super_steps = a.super_steps;
  cfg.snapshot_every = a.snapshot_every;
  cfg.keep_final = 0;
  spc2_server_t srv = nullptr;
  if (spc2_server_create(&cfg, &srv) != SPC2_OK) {
    std::cerr << "spc2_serve: server: " << spc2_last_error() << std::endl;
    return 1;
  }
  std::mt19937_64 rng(a.seed);
  std::exponential_distribution<double> gap(a.rate);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  std::uniform_int_distribution<long> steps(a.steps_min, a.steps_max);
  const auto t0 = std::chrono::steady_clock::now();
  double t_next = 0.0, t_report = 30.0;
  long long submitted = 0;
  int rc = 0;
  for (;;) {
    t_next += gap(rng);
    if (t_next > a.duration) break;
    std::this_thread::sleep_until(t0 + std::chrono::duration<double>(t_next));
    spc2_params p;
    spc2_default_params(&p);
    p.system = u01(rng) < 0.5 ? SPC2_CHIMERA : SPC2_VORTEX;
    p.precision = SPC2_FP32_KAHAN;
    p.n = (int64_t)std::exp(std::log(a.n_min) + u01(rng) * (std::log(a.n_max) - std::log(a.n_min)));
    p.Lx = p.Ly = std::sqrt((double)p.n / 4000.0);
    p.rho = 0.1;
    p.sigma = 0.2 + 1.8 * u01(rng);
    p.alpha = 1.2 + 0.37 * u01(rng);
    p.D = std::exp(std::log(0.001) + u01(rng) * (std::log(0.1) - std::log(0.001)));
    if (p.system == SPC2_VORTEX) {  // the reference's set (SimulationEnginePtr.cpp:131-139) scaled to rho = 0.05
      p.rho = 0.05;
      p.v0 = 0.025;
      p.mu_plus = p.mu_minus = 100.0;
      p.xi_a = 0.01;
      p.xi_r = 0.005;
      p.kappa = 20.0;
    }
    p.seed = a.seed + (uint64_t)submitted;
    int64_t id = 0;
    if (spc2_server_submit(srv, &p, steps(rng), &id) != SPC2_OK) {
      std::cerr << "spc2_serve: submit failed" << std::endl;
      rc = 1;
      break;
    }
    ++submitted;
    if (t_next > t_report) {
      spc2_server_stats s;
      if (spc2_server_stats_get(srv, &s) != SPC2_OK) {
        std::cerr << "spc2_serve: " << spc2_last_error() << std::endl;
        rc = 1;
        break;
      }
      std::fprintf(stderr, "# t=%.0f s submitted %lld completed %lld in flight %lld queued %lld\n", t_next,
                   (long long)s.submitted, (long long)s.completed, (long long)s.in_flight, (long long)s.queued);
      t_report += 30.0;
    }
  }
  if (rc == 0 && a.drain) {
    if (spc2_server_drain(srv, -1.0) != SPC2_OK) rc = 1;
  }
  spc2_server_stats s;
  const spc2_status sst = spc2_server_stats_get(srv, &s);
  if (sst != SPC2_OK) rc = 1;
  std::printf(
      "{\"metric\": \"serving p50/p99 latency and particle-steps/s\", \"offered_rate_per_s\": %.6g, "
      "\"arrival_seconds\": %.6g, \"wall_seconds\": %.6g, \"submitted\": %lld, \"admitted\": %lld, "
      "\"completed\": %lld, \"queued_at_end\": %lld, \"in_flight_at_end\": %lld, \"max_in_flight\": %lld, "
      "\"max_particles_in_flight\": %lld, \"repacks\": %lld, \"frames\": %lld, \"frame_bytes\": %llu, "
      "\"particle_steps_completed\": %.6g, \"particle_steps_per_s\": %.6g, \"particle_steps_executed_per_s\": %.6g, "
      "\"busy_fraction\": %.4f, "
      "\"latency_p50_s\": %.6g, \"latency_p99_s\": %.6g, \"latency_max_s\": %.6g, \"queue_wait_p50_s\": %.6g, "
      "\"queue_wait_p99_s\": %.6g, \"dtype\": \"fp32k\", \"method\": \"se\", \"drained\": %s}\n",
      a.rate, a.duration, s.wall_seconds, (long long)s.submitted, (long long)s.admitted, (long long)s.completed,
      (long long)s.queued, (long long)s.in_flight, (long long)s.max_in_flight, (long long)s.max_particles_in_flight,
      (long long)s.repacks, (long long)s.frames, (unsigned long long)s.frame_bytes, s.particle_steps,
      s.wall_seconds > 0 ? s.particle_steps / s.wall_seconds : 0.0,
      s.wall_seconds > 0 ? s.particle_steps_executed / s.wall_seconds : 0.0,
      s.wall_seconds > 0 ? s.busy_seconds / s.wall_seconds : 0.0, s.latency_p50, s.latency_p99, s.latency_max,
      s.queue_wait_p50, s.queue_wait_p99, a.drain ? "true" : "false");
  std::fflush(stdout);
  spc2_server_destroy(srv);
  return rc;
}
