// Continuous-batching request loop over the batched-replica ABI
// (include/spc2_server.h).
//
// One worker thread owns the device batch.  Each round it (1) admits queued
// requests first come, first served while the device footprint, particle and
// request caps allow, (2) re-packs the batch when its membership changed
// (finished requests leave, admitted ones join; running ones keep their state
// on the device), (3) steps every request min(super_steps, steps left) steps
// in one batched call, (4) streams frames of the requests that reached a
// snapshot step through a pinned buffer, and (5) copies out the final state of
// finished requests and wakes their waiters.
#include "../../include/spc2_server.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  spc2_params p{};
  int64_t steps = 0;
  uint64_t footprint = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  bool admitted = false, done = false;
  std::vector<double> final_state;
};

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q / 100.0 * (double)(v.size() - 1);
  const size_t lo = (size_t)pos, hi = std::min(v.size() - 1, lo + 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace

struct spc2_server {
  spc2_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::string error;  // first device failure of the loop
  std::map<int64_t, Request> req;  // references stay valid while others are inserted
  std::deque<int64_t> queue;
  // worker thread only: request ids in batch order and their step counts
  std::vector<int64_t> active, active_steps;
  int64_t next_id = 0;
  int32_t precision = -1;
  spc2_batch_t batch = nullptr;
  float *frames = nullptr;
  size_t frames_cap = 0;
  // stats (under mu)
  spc2_server_stats st{};
  std::vector<double> latencies, waits;
  double t0 = 0;
  std::thread worker;

  void loop();
  spc2_status round(std::vector<int64_t> &admit);
};

spc2_status spc2_server::round(std::vector<int64_t> &admit) {
  // membership: running requests that are not done, then the admitted ones
  std::vector<int64_t> next, next_steps;
  std::vector<int32_t> prev_index;
  std::vector<spc2_params> params;
  bool changed = !admit.empty();
  {
    std::lock_guard<std::mutex> lk(mu);
    for (size_t r = 0; r < active.size(); ++r) {
      if (req[active[r]].done) {
        changed = true;
        continue;
      }
      next.push_back(active[r]);
      next_steps.push_back(active_steps[r]);
      prev_index.push_back((int32_t)r);
      params.push_back(req[active[r]].p);
    }
    for (int64_t id : admit) {
      next.push_back(id);
      next_steps.push_back(req[id].steps);
      prev_index.push_back(-1);
      params.push_back(req[id].p);
    }
  }
  if (changed) {
    spc2_batch_t nb = nullptr;
    if (!next.empty()) {
      spc2_status s = spc2_batch_create_from(params.data(), (int32_t)params.size(), batch, prev_index.data(),
                                             cfg.omega_std, &nb);
      if (s) return s;
    }
    if (batch) spc2_batch_destroy(batch);
    batch = nb;
    active = next;
    active_steps = next_steps;
    std::lock_guard<std::mutex> lk(mu);
    st.repacks += 1;
    int64_t parts = 0;
    for (int64_t id : active) parts += req[id].p.n;
    st.in_flight = (int64_t)active.size();
    st.max_in_flight = std::max(st.max_in_flight, st.in_flight);
    st.max_particles_in_flight = std::max(st.max_particles_in_flight, parts);
  }
  if (active.empty()) return SPC2_OK;

  // one batched super-step: each request takes min(super_steps, steps left)
  const int R = (int)active.size();
  std::vector<int64_t> done_before(R), limit(R);
  if (spc2_status s = spc2_batch_steps_done(batch, done_before.data())) return s;
  int64_t kmax = 0;
  for (int r = 0; r < R; ++r) {
    limit[r] = std::min<int64_t>(cfg.super_steps, active_steps[r] - done_before[r]);
    kmax = std::max(kmax, limit[r]);
  }
  const double tb = now_s();
  if (spc2_status s = spc2_batch_step(batch, SPC2_SE, cfg.dt, kmax, limit.data())) return s;
  const double te = now_s();

  // frames of the requests at a snapshot step (or done)
  std::vector<int32_t> first(R), n(R);
  if (spc2_status s = spc2_batch_layout(batch, first.data(), n.data())) return s;
  bool any_frame = false, any_done = false;
  std::vector<char> frame(R, 0), fin(R, 0);
  for (int r = 0; r < R; ++r) {
    const int64_t d = done_before[r] + limit[r];
    fin[r] = d >= active_steps[r];
    any_done = any_done || fin[r];
    if (cfg.snapshot_every > 0 && (d / cfg.snapshot_every > done_before[r] / cfg.snapshot_every || fin[r])) {
      frame[r] = 1;
      any_frame = true;
    }
  }
  double executed = 0.0;
  for (int r = 0; r < R; ++r) executed += (double)n[r] * (double)limit[r];
  uint64_t fbytes = 0;
  int64_t nframes = 0;
  if (any_frame) {
    const size_t need = 3 * (size_t)(first[R - 1] + n[R - 1]);
    if (need > frames_cap) {
      if (frames) spc2_host_free(frames);
      frames = nullptr;
      frames_cap = 0;
      void *p = nullptr;
      if (spc2_status s = spc2_host_alloc(need * sizeof(float), &p)) return s;
      frames = (float *)p;
      frames_cap = need;
    }
    if (spc2_status s = spc2_batch_snapshot_async(batch, frames)) return s;
    if (spc2_status s = spc2_batch_snapshot_wait(batch)) return s;
    for (int r = 0; r < R; ++r)
      if (frame[r]) {
        fbytes += (uint64_t)(1 + 3 * (size_t)n[r]) * sizeof(float);  // frame = float t + 3N floats
        nframes += 1;
      }
  }
  // final states of finished requests
  std::vector<std::vector<double>> finals(R);
  if (any_done && cfg.keep_final) {
    for (int r = 0; r < R; ++r)
      if (fin[r]) {
        finals[r].resize(3 * (size_t)n[r]);
        if (spc2_status s = spc2_batch_get_state(batch, r, finals[r].data())) return s;
      }
  }
  const double tf = now_s();
  {
    std::lock_guard<std::mutex> lk(mu);
    st.busy_seconds += te - tb;
    st.particle_steps_executed += executed;
    st.frames += nframes;
    st.frame_bytes += fbytes;
    for (int r = 0; r < R; ++r) {
      if (!fin[r]) continue;
      Request &q = req[active[r]];
      q.done = true;
      q.t_done = tf;
      q.final_state = std::move(finals[r]);
      st.completed += 1;
      st.in_flight -= 1;  // leaves the batch at the next re-pack
      st.particle_steps += (double)q.p.n * (double)q.steps;
      latencies.push_back(q.t_done - q.t_submit);
    }
  }
  if (any_done) cv_done.notify_all();
  return SPC2_OK;
}

void spc2_server::loop() {
  uint64_t used = 0;
  int64_t parts = 0;
  for (;;) {
    std::vector<int64_t> admit;
    {
      std::unique_lock<std::mutex> lk(mu);
      // footprint of the requests that stay
      used = 0;
      parts = 0;
      int64_t staying = 0;
      for (int64_t id : active)
        if (!req[id].done) {
          used += req[id].footprint;
          parts += req[id].p.n;
          staying += 1;
        }
      if (active.empty() || staying == 0) {
        cv_work.wait(lk, [&] { return stop || !queue.empty(); });
      }
      if (stop) break;
      const double t = now_s();
      while (!queue.empty()) {
        Request &q = req[queue.front()];
        const bool fits = used + q.footprint <= cfg.hbm_budget &&
                          (cfg.max_particles <= 0 || parts + q.p.n <= cfg.max_particles) &&
                          (cfg.max_replicas <= 0 || staying + (int64_t)admit.size() < cfg.max_replicas);
        if (!fits && (staying + (int64_t)admit.size()) > 0) break;  // first come, first served
        q.admitted = true;
        q.t_admit = t;
        waits.push_back(t - q.t_submit);
        used += q.footprint;
        parts += q.p.n;
        admit.push_back(queue.front());
        queue.pop_front();
        st.admitted += 1;
      }
      st.queued = (int64_t)queue.size();
    }
    spc2_status s = round(admit);
    if (s != SPC2_OK) {
      std::lock_guard<std::mutex> lk(mu);
      error = spc2_last_error();
      stop = true;
      cv_done.notify_all();
      break;
    }
  }
}

extern "C" {

void spc2_server_default_config(spc2_server_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->super_steps = 100;
  cfg->snapshot_every = 100;
  cfg->keep_final = 1;
  cfg->dt = 0.01;
  cfg->omega_std = 0.1;
}

spc2_status spc2_server_create(const spc2_server_config *cfg, spc2_server_t *out) {
  if (!cfg || !out) return SPC2_ERR_INVALID;
  *out = nullptr;
  if (cfg->super_steps < 1 || !(cfg->dt > 0.0) || !(cfg->omega_std >= 0.0) || cfg->snapshot_every < 0)
    return SPC2_ERR_INVALID;
  spc2_server *s = new (std::nothrow) spc2_server();
  if (!s) return SPC2_ERR_NOMEM;
  s->cfg = *cfg;
  if (s->cfg.hbm_budget == 0) {
    uint64_t free_b = 0, total_b = 0;
    if (spc2_status e = spc2_device_memory(&free_b, &total_b)) {
      delete s;
      return e;
    }
    s->cfg.hbm_budget = (uint64_t)(0.8 * (double)free_b);
  }
  s->t0 = now_s();
  s->worker = std::thread(&spc2_server::loop, s);
  *out = s;
  return SPC2_OK;
}

spc2_status spc2_server_destroy(spc2_server_t s) {
  if (!s) return SPC2_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  if (s->batch) spc2_batch_destroy(s->batch);
  if (s->frames) spc2_host_free(s->frames);
  delete s;
  return SPC2_OK;
}

spc2_status spc2_server_submit(spc2_server_t s, const spc2_params *p, int64_t steps, int64_t *id) {
  if (!s || !p || !id || steps < 0 || p->n <= 0) return SPC2_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return SPC2_ERR_INVALID;
  if (s->precision < 0) s->precision = p->precision;
  if (p->precision != s->precision || p->neighbour != SPC2_NEIGH_AUTO) return SPC2_ERR_UNSUPPORTED;
  Request q;
  q.p = *p;
  q.steps = steps;
  q.footprint = spc2_batch_footprint(p);
  q.t_submit = now_s();
  *id = s->next_id++;
  s->req.emplace(*id, std::move(q));
  s->queue.push_back(*id);
  s->st.submitted += 1;
  s->cv_work.notify_all();
  return SPC2_OK;
}

spc2_status spc2_server_wait(spc2_server_t s, int64_t id, double timeout_s, double *state, double *latency_s) {
  if (!s) return SPC2_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->req.find(id);
  if (it == s->req.end()) return SPC2_ERR_INVALID;
  const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(std::max(0.0, timeout_s)),
                                      [&] { return it->second.done || s->stop; });
  if (!ok || !it->second.done) return s->error.empty() ? SPC2_ERR_INVALID : SPC2_ERR_HIP;
  if (state) {
    if (it->second.final_state.empty()) return SPC2_ERR_INVALID;
    std::copy(it->second.final_state.begin(), it->second.final_state.end(), state);
  }
  if (latency_s) *latency_s = it->second.t_done - it->second.t_submit;
  return SPC2_OK;
}

spc2_status spc2_server_drain(spc2_server_t s, double timeout_s) {
  if (!s) return SPC2_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto idle = [&] { return s->stop || s->st.completed == s->st.submitted; };
  if (timeout_s < 0) {
    s->cv_done.wait(lk, idle);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), idle)) {
    return SPC2_ERR_INVALID;
  }
  return s->error.empty() ? SPC2_OK : SPC2_ERR_HIP;
}

spc2_status spc2_server_stats_get(spc2_server_t s, spc2_server_stats *out) {
  if (!s || !out) return SPC2_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->queued = (int64_t)s->queue.size();
  out->wall_seconds = now_s() - s->t0;
  out->latency_p50 = percentile(s->latencies, 50.0);
  out->latency_p99 = percentile(s->latencies, 99.0);
  out->latency_max = s->latencies.empty() ? 0.0 : *std::max_element(s->latencies.begin(), s->latencies.end());
  out->queue_wait_p50 = percentile(s->waits, 50.0);
  out->queue_wait_p99 = percentile(s->waits, 99.0);
  return s->error.empty() ? SPC2_OK : SPC2_ERR_HIP;
}

}  // extern "C"
