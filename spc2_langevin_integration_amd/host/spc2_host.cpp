// Implementation of the host layer declared in spc2_host.hpp.
#include "spc2_host.hpp"

#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <random>
#include <sstream>

void Spc2Check(spc2_status st, const char *what) {
  if (st != SPC2_OK) throw Spc2Error(std::string(what) + ": " + spc2_last_error());
}

// ---------------------------------------------------------------------------
// PeriodicBoundaryConditions (PeriodicBoundaryConditions.cpp:7-151)
// ---------------------------------------------------------------------------

PeriodicBoundaryConditions::PeriodicBoundaryConditions(Real x_size, Real y_size)
    : x_size_(x_size), y_size_(y_size), x_rsize_(1.0 / x_size), y_rsize_(1.0 / y_size) {}

// signed minimum image, truncation toward zero (PeriodicBoundaryConditions.cpp:22-34)
void PeriodicBoundaryConditions::ClassAEffectiveParticleDistance(Real x_i, Real y_i, Real x_j, Real y_j, Real &dx,
                                                                 Real &dy) const {
  dx = x_j - x_i;
  dx -= (Real)(int)(dx * 2.0 * x_rsize_) * x_size_;
  dy = y_j - y_i;
  dy -= (Real)(int)(dy * 2.0 * y_rsize_) * y_size_;
}

void PeriodicBoundaryConditions::ApplyPeriodicBoundaryConditions(Real x, Real y, Real &x_pbc, Real &y_pbc) const {
  x_pbc = x - std::floor(x * x_rsize_) * x_size_;
  y_pbc = y - std::floor(y * y_rsize_) * y_size_;
}

void PeriodicBoundaryConditions::ApplyPeriodicBoundaryConditions(std::vector<Real> &system_state) const {
  ApplyPeriodicBoundaryConditions(system_state.data(), (long)system_state.size());
}

// x and y of every particle; phi is left as it is (PeriodicBoundaryConditions.cpp:133-141)
void PeriodicBoundaryConditions::ApplyPeriodicBoundaryConditions(Real *const system_state, long size) const {
  for (long ii = 0; ii + 2 < size; ii += kS) {
    system_state[ii] -= std::floor(system_state[ii] * x_rsize_) * x_size_;
    system_state[ii + 1] -= std::floor(system_state[ii + 1] * y_rsize_) * y_size_;
  }
}

// ---------------------------------------------------------------------------
// Systems
// ---------------------------------------------------------------------------

DeviceSystem::DeviceSystem(const spc2_params &p) : p_(p) { Spc2Check(spc2_create(&p_, &h_), "spc2_create"); }

DeviceSystem::~DeviceSystem() {
  if (h_) spc2_destroy(h_);
}

void DeviceSystem::Upload(const Real *state) { Spc2Check(spc2_set_state(h_, state), "spc2_set_state"); }

void DeviceSystem::Download(Real *state) const { Spc2Check(spc2_get_state(h_, state), "spc2_get_state"); }

void DeviceSystem::SetNaturalFrequencies(const std::vector<Real> &omega) {
  if ((long)omega.size() != n()) throw Spc2Error("natural frequencies: expected N values");
  Spc2Check(spc2_set_omega(h_, omega.data()), "spc2_set_omega");
}

namespace {

spc2_params ChimeraParams(Real v_0, Real sigma, Real rho, Real alpha, Real D_phi, const PeriodicBoundaryConditions &pbc,
                          long n, Precision precision, uint64_t seed, NeighbourMode neighbour, Real verlet_skin) {
  spc2_params p;
  spc2_default_params(&p);
  p.neighbour = (int32_t)neighbour;
  p.verlet_skin = verlet_skin;
  p.system = SPC2_CHIMERA;
  p.precision = (int32_t)precision;
  p.n = n;
  p.Lx = pbc.GetXSize();
  p.Ly = pbc.GetYSize();
  p.rho = rho;
  p.v0 = v_0;
  p.sigma = sigma;
  p.alpha = alpha;
  p.D = D_phi;
  p.seed = seed;
  return p;
}

spc2_params VortexParams(Real velocity, Real mu_plus, Real mu_minus, Real xi_a, Real xi_r, Real D_phi, Real kappa,
                         Real rho, const PeriodicBoundaryConditions &pbc, long n, Precision precision, uint64_t seed,
                         NeighbourMode neighbour, Real verlet_skin) {
  spc2_params p;
  spc2_default_params(&p);
  p.neighbour = (int32_t)neighbour;
  p.verlet_skin = verlet_skin;
  p.system = SPC2_VORTEX;
  p.precision = (int32_t)precision;
  p.n = n;
  p.Lx = pbc.GetXSize();
  p.Ly = pbc.GetYSize();
  p.rho = rho;
  p.v0 = velocity;
  p.mu_plus = mu_plus;
  p.mu_minus = mu_minus;
  p.xi_a = xi_a;
  p.xi_r = xi_r;
  p.D = D_phi;
  p.kappa = kappa;
  p.seed = seed;
  return p;
}

// k-form EvaluateRhs shared by both systems: upload the host state, evaluate
// at s + k_coef dt k_prev, k_next in/out.
void EvalK(DeviceSystem &sys, const Real *state, const std::vector<Real> &k_prev, std::vector<Real> &k_next, Real k_coef,
           Real dt) {
  const size_t m = (size_t)kS * (size_t)sys.n();
  if (k_prev.size() != m || k_next.size() != m) throw Spc2Error("EvaluateRhs: k vectors must hold 3N values");
  sys.Upload(state);
  Spc2Check(spc2_eval_rhs(sys.handle(), k_prev.data(), k_coef, dt, k_next.data()), "spc2_eval_rhs");
}

// Ito-form EvaluateRhs shared by both systems: derivative (3N) and the N x 2
// additional derivative, both in/out as in the reference.
void EvalIto(DeviceSystem &sys, const Real *state, std::vector<Real> &derivative,
             std::vector<std::vector<Real>> &additional_derivative) {
  const long nn = sys.n();
  if ((long)derivative.size() != kS * nn || (long)additional_derivative.size() != nn)
    throw Spc2Error("EvaluateRhs: derivative must hold 3N values and additional_derivative N rows");
  std::vector<Real> add(2 * (size_t)nn);
  for (long i = 0; i < nn; ++i) {
    if (additional_derivative[i].size() != 2) throw Spc2Error("EvaluateRhs: additional_derivative rows hold 2 values");
    add[2 * i] = additional_derivative[i][0];
    add[2 * i + 1] = additional_derivative[i][1];
  }
  sys.Upload(state);
  Spc2Check(spc2_eval_rhs_ito(sys.handle(), derivative.data(), add.data()), "spc2_eval_rhs_ito");
  for (long i = 0; i < nn; ++i) {
    additional_derivative[i][0] = add[2 * i];
    additional_derivative[i][1] = add[2 * i + 1];
  }
}

}  // namespace

ChimeraSystem::ChimeraSystem(Real v_0, Real sigma, Real rho, Real alpha, Real D_phi, Real rho_0,
                             PeriodicBoundaryConditions &pbc_config, long n, Precision precision, uint64_t seed,
                             NeighbourMode neighbour, Real verlet_skin)
    : DeviceSystem(
          ChimeraParams(v_0, sigma, rho, alpha, D_phi, pbc_config, n, precision, seed, neighbour, verlet_skin)) {
  (void)rho_0;  // density label of the reference's output file names only
}

void ChimeraSystem::EvaluateRhs(const Real *const system_state, const std::vector<Real> &k_prev,
                                std::vector<Real> &k_next, Real k_coef, Real dt) {
  EvalK(*this, system_state, k_prev, k_next, k_coef, dt);
}

// Ito form, ChimeraSystemPtr.cpp:131-144
void ChimeraSystem::EvaluateRhs(const Real *const system_state, std::vector<Real> &derivative,
                                std::vector<std::vector<Real>> &additional_derivative, Real dt) {
  (void)dt;
  EvalIto(*this, system_state, derivative, additional_derivative);
}

VortexArraysSystem::VortexArraysSystem(Real velocity, Real mu_plus, Real mu_minus, Real xi_a, Real xi_r, Real D_phi,
                                       Real kappa, Real rho, PeriodicBoundaryConditions &pbc_config, long n,
                                       Precision precision, uint64_t seed, NeighbourMode neighbour, Real verlet_skin)
    : DeviceSystem(VortexParams(velocity, mu_plus, mu_minus, xi_a, xi_r, D_phi, kappa, rho, pbc_config, n, precision,
                                seed, neighbour, verlet_skin)) {}

void VortexArraysSystem::EvaluateRhs(const Real *const system_state, const std::vector<Real> &k_prev,
                                     std::vector<Real> &k_next, Real k_coef, Real dt) {
  EvalK(*this, system_state, k_prev, k_next, k_coef, dt);
}

// Ito form, VortexArraysSystemPtr.cpp:122-135
void VortexArraysSystem::EvaluateRhs(const Real *const system_state, std::vector<Real> &derivative,
                                     std::vector<std::vector<Real>> &additional_derivative, Real dt) {
  (void)dt;
  EvalIto(*this, system_state, derivative, additional_derivative);
}

// ---------------------------------------------------------------------------
// Steppers: one device step of the host state, no wrap
// ---------------------------------------------------------------------------

namespace {

void HostStep(DeviceSystem &sys, Real *const state, int32_t method, Real dt, int64_t step, const Real *noise) {
  sys.Upload(state);
  Spc2Check(spc2_step_ex(sys.handle(), method, dt, step, 1, noise, SPC2_STEP_NO_WRAP), "spc2_step_ex");
  sys.Download(state);
}

// `count` draws of a fresh N(0, 1) from the generator, in order
void DrawNormals(std::mt19937 &gen, std::vector<Real> &out, size_t count) {
  out.resize(count);
  std::normal_distribution<Real> z(0.0, 1.0);
  for (Real &v : out) v = z(gen);
}

}  // namespace

void StochasticEulerStepper::Step(DeviceSystem &system, Real *const system_state, Real dt) {
  const Real *noise = nullptr;
  if (gen_) {
    DrawNormals(*gen_, noise_, (size_t)system.n());
    noise = noise_.data();
  }
  HostStep(system, system_state, SPC2_SE, dt, step_++, noise);
}

void StochasticEulerStepper::DoStep(ChimeraSystem &system, Real *const system_state, Real t, Real dt) {
  (void)t;
  Step(system, system_state, dt);
}

void StochasticEulerStepper::DoStep(VortexArraysSystem &system, Real *const system_state, Real t, Real dt) {
  (void)t;
  Step(system, system_state, dt);
}

StochasticRungeKuttaStepper::StochasticRungeKuttaStepper(std::mt19937 *generator, long n_frequencies)
    : gen_(generator) {
  if (gen_ && n_frequencies > 0) DrawNormals(*gen_, noise_, (size_t)n_frequencies);
}

void StochasticRungeKuttaStepper::Step(DeviceSystem &system, Real *const system_state, Real dt) {
  const Real *noise = nullptr;
  if (gen_) {
    DrawNormals(*gen_, noise_, 2 * (size_t)system.n());  // (U_1, U_2) per particle
    noise = noise_.data();
  }
  HostStep(system, system_state, SPC2_SRK, dt, step_++, noise);
}

void StochasticRungeKuttaStepper::DoStep(ChimeraSystem &system, Real *const system_state, Real t, Real dt) {
  (void)t;
  Step(system, system_state, dt);
}

void StochasticRungeKuttaStepper::DoStep(VortexArraysSystem &system, Real *const system_state, Real t, Real dt) {
  (void)t;
  Step(system, system_state, dt);
}

// The k_1..k_4 buffers live with the device system (one stepper per system, as
// the reference's engine uses it), so a phi entry with no partner keeps its
// value from the previous stage or step as in RungeKutta4StepperPtr.hpp:28-31.
void RungeKutta4Stepper::DoStep(ChimeraSystem &system, Real *const system_state, Real t, Real dt) {
  (void)t;
  HostStep(system, system_state, SPC2_RK4, dt, 0, nullptr);
}

void RungeKutta4Stepper::DoStep(VortexArraysSystem &system, Real *const system_state, Real t, Real dt) {
  (void)t;
  HostStep(system, system_state, SPC2_RK4, dt, 0, nullptr);
}

// ---------------------------------------------------------------------------
// BinaryObserver
// ---------------------------------------------------------------------------

namespace {
constexpr int kRingFrames = 4;

// calls from now until the one that writes, counting the next call as 1
long CallsUntilWrite(long counter, int every) {
  if (every <= 0) return std::numeric_limits<long>::max();
  return (long)((every - counter % every) % every) + 1;
}
}  // namespace

BinaryObserver::BinaryObserver(const std::string &state_path, const std::string &summary_path, int state_every,
                               int summary_every, bool float_summary)
    : state_every_(state_every), summary_every_(summary_every), float_summary_(float_summary) {
  // the output cadences are divisors (BinaryObserverPtr.cpp:153, :232)
  if (!state_path.empty() && state_every <= 0) throw Spc2Error("BinaryObserver: state_every must be > 0");
  if (!summary_path.empty() && summary_every <= 0) throw Spc2Error("BinaryObserver: summary_every must be > 0");
  if (!state_path.empty()) {
    state_file_.open(state_path, std::ios::binary | std::ios::out | std::ios::trunc);
    if (!state_file_.is_open()) throw Spc2Error("cannot open " + state_path);
  }
  if (!summary_path.empty()) {
    summary_file_.open(summary_path, std::ios::out | std::ios::trunc);
    if (!summary_file_.is_open()) throw Spc2Error("cannot open " + summary_path);
  }
  if (state_file_.is_open()) writer_ = std::thread(&BinaryObserver::WriterLoop, this);
}

BinaryObserver::~BinaryObserver() {
  try {
    Flush();
  } catch (const std::exception &e) {
    std::cerr << "BinaryObserver: " << e.what() << std::endl;
  }
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (writer_.joinable()) writer_.join();
  for (Frame &f : ring_) spc2_host_free(f.buf);
}

void BinaryObserver::WriterLoop() {
  for (;;) {
    int idx;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [this] { return stop_ || !queue_.empty(); });
      if (queue_.empty()) return;
      idx = queue_.front();
    }
    Frame &f = ring_[idx];
    state_file_.write(reinterpret_cast<const char *>(&f.t), sizeof(float));
    state_file_.write(reinterpret_cast<const char *>(f.buf), (std::streamsize)(frame_floats_ * sizeof(float)));
    {
      std::lock_guard<std::mutex> lk(mu_);
      queue_.pop_front();
      f.busy = false;
    }
    cv_.notify_all();
  }
}

// Hands a filled frame to the writer thread.
void BinaryObserver::Enqueue(int idx) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    queue_.push_back(idx);
  }
  cv_.notify_all();
}

// Next ring frame, once the writer has released it; (re)allocates the ring
// for `floats` floats per frame.
int BinaryObserver::AcquireFrame(size_t floats) {
  if (floats != frame_floats_) {
    Flush();
    for (Frame &f : ring_) spc2_host_free(f.buf);
    ring_.assign(kRingFrames, Frame{});
    for (Frame &f : ring_) {
      void *p = nullptr;
      Spc2Check(spc2_host_alloc(floats * sizeof(float), &p), "spc2_host_alloc");
      f.buf = static_cast<float *>(p);
    }
    frame_floats_ = floats;
    next_ = 0;
  }
  int idx = next_;
  next_ = (next_ + 1) % kRingFrames;
  std::unique_lock<std::mutex> lk(mu_);
  cv_.wait(lk, [&] { return !ring_[idx].busy; });
  ring_[idx].busy = true;
  return idx;
}

// Completes the asynchronous copy in flight, if any, and queues its frame.
void BinaryObserver::RetirePending() {
  if (pending_ < 0) return;
  Spc2Check(spc2_snapshot_wait(pending_handle_), "spc2_snapshot_wait");
  int idx = pending_;
  pending_ = -1;
  pending_handle_ = nullptr;
  Enqueue(idx);
}

// BinaryObserverPtr.cpp:151-171: float t followed by the state as 3N floats.
void BinaryObserver::SaveSystemState(Real *const system_state, long size, Real t) {
  if (state_file_.is_open() && !(counter_state_ % state_every_)) {
    RetirePending();
    int idx = AcquireFrame((size_t)size);
    Frame &f = ring_[idx];
    f.t = (float)t;
    for (long k = 0; k < size; ++k) f.buf[k] = (float)system_state[k];
    Enqueue(idx);
  }
  ++counter_state_;
}

void BinaryObserver::SaveSystemState(DeviceSystem &system, Real t) {
  if (state_file_.is_open() && !(counter_state_ % state_every_)) {
    RetirePending();
    int idx = AcquireFrame((size_t)kS * (size_t)system.n());
    ring_[idx].t = (float)t;
    Spc2Check(spc2_snapshot_async(system.handle(), ring_[idx].buf), "spc2_snapshot_async");
    pending_ = idx;
    pending_handle_ = system.handle();
  }
  ++counter_state_;
}

void BinaryObserver::WriteSummaryLine(Real t, float R, float psi) {
  float t_float = (float)t;
  summary_file_ << t_float << '\t' << R << '\t' << psi << '\n';
}

namespace {
// The reference's float order parameter (BinaryObserverPtr.cpp:237-247):
// complex<float> sum of (float cos, float sin) in index order, divided by N.
void FloatOrderParameter(const Real *state, long n, float &R, float &psi) {
  std::complex<float> z(0.0f, 0.0f);
  for (long i = 0; i < n; ++i)
    z += std::complex<float>((float)std::cos(state[kS * i + 2]), (float)std::sin(state[kS * i + 2]));
  z /= (float)n;
  R = std::abs(z);
  psi = std::arg(z);
}
}  // namespace

void BinaryObserver::SaveSummaryStatistics(const Real *const system_state, long size, Real t) {
  if (summary_file_.is_open() && !(counter_summary_ % summary_every_)) {
    float R, psi;
    FloatOrderParameter(system_state, size / kS, R, psi);
    WriteSummaryLine(t, R, psi);
  }
  ++counter_summary_;
}

void BinaryObserver::SaveSummaryStatistics(DeviceSystem &system, Real t) {
  if (summary_file_.is_open() && !(counter_summary_ % summary_every_)) {
    float R, psi;
    if (float_summary_) {
      std::vector<Real> st((size_t)kS * (size_t)system.n());
      system.Download(st.data());
      FloatOrderParameter(st.data(), system.n(), R, psi);
    } else {
      double Rd, psid;
      Spc2Check(spc2_order_parameter(system.handle(), &Rd, &psid), "spc2_order_parameter");
      R = (float)Rd;
      psi = (float)psid;
    }
    WriteSummaryLine(t, R, psi);
  }
  ++counter_summary_;
}

long BinaryObserver::CallsUntilOutput() const {
  long d = std::numeric_limits<long>::max();
  if (state_file_.is_open()) d = std::min(d, CallsUntilWrite(counter_state_, state_every_));
  if (summary_file_.is_open()) d = std::min(d, CallsUntilWrite(counter_summary_, summary_every_));
  return d;
}

void BinaryObserver::SkipCalls(long k) {
  counter_state_ += k;
  counter_summary_ += k;
}

void BinaryObserver::Flush() {
  RetirePending();
  {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [this] { return queue_.empty(); });
  }
  if (state_file_.is_open()) state_file_.flush();
  if (summary_file_.is_open()) summary_file_.flush();
}

// ---------------------------------------------------------------------------
// EngineConfig / SimulationEngine
// ---------------------------------------------------------------------------

void EngineConfig::Set(const std::string &key, const std::string &value) {
  auto d = [&] { return std::stod(value); };
  auto l = [&] { return std::stol(value); };
  if (key == "system") system = value;
  else if (key == "method") method = value;
  else if (key == "precision") precision = value;
  else if (key == "init") init = value;
  else if (key == "noise") noise = value;
  else if (key == "loop") loop = value;
  else if (key == "n") n = l();
  else if (key == "L") L = d();
  else if (key == "v0") v0 = d();
  else if (key == "sigma") sigma = d();
  else if (key == "rho") rho = d();
  else if (key == "alpha") alpha = d();
  else if (key == "D") D = d();
  else if (key == "rho0") rho0 = d();
  else if (key == "omega_std") omega_std = d();
  else if (key == "mu_plus") mu_plus = d();
  else if (key == "mu_minus") mu_minus = d();
  else if (key == "xi_a") xi_a = d();
  else if (key == "xi_r") xi_r = d();
  else if (key == "kappa") kappa = d();
  else if (key == "dt") dt = d();
  else if (key == "t0") t0 = d();
  else if (key == "t1") t1 = d();
  else if (key == "steps") steps = l();
  else if (key == "seed") seed = std::stoull(value);
  else if (key == "state") state_path = value;
  else if (key == "summary") summary_path = value;
  else if (key == "state_every") state_every = (int)l();
  else if (key == "summary_every") summary_every = (int)l();
  else if (key == "float_summary") float_summary = l() != 0;
  else if (key == "neighbour") neighbour = value;
  else if (key == "verlet_skin") verlet_skin = d();
  else if (key == "init_file") init_file = value;
  else if (key == "init_frame") init_frame = l();
  else if (key == "init_state_file") init_state_file = value;
  else if (key == "resume") resume = value;
  else if (key == "checkpoint") checkpoint = value;
  else if (key == "checkpoint_every") checkpoint_every = l();
  else throw Spc2Error("unknown configuration key '" + key + "'");
}

namespace {

Precision ParsePrecision(const std::string &s) {
  if (s == "fp64") return Precision::kFP64;
  if (s == "fp32k") return Precision::kFP32Kahan;
  if (s == "fp64f") return Precision::kFP64Factored;
  throw Spc2Error("precision must be fp64, fp32k or fp64f");
}

NeighbourMode ParseNeighbour(const std::string &s) {
  if (s == "auto") return NeighbourMode::kAuto;
  if (s == "cells") return NeighbourMode::kCells;
  if (s == "all_pairs") return NeighbourMode::kAllPairs;
  if (s == "verlet") return NeighbourMode::kVerlet;
  throw Spc2Error("neighbour must be auto, cells, all_pairs or verlet");
}

int32_t ParseMethod(const std::string &s) {
  if (s == "se") return SPC2_SE;
  if (s == "srk") return SPC2_SRK;
  if (s == "rk4") return SPC2_RK4;
  throw Spc2Error("method must be se, srk or rk4");
}

}  // namespace

SimulationEngine::SimulationEngine(const EngineConfig &cfg) : cfg_(cfg), pbc_(cfg.L, cfg.L) {
  if (cfg_.n <= 0) throw Spc2Error("n must be positive");
  if (cfg_.state_every <= 0 || cfg_.summary_every <= 0) throw Spc2Error("output intervals must be positive");
  if (cfg_.init != "philox" && cfg_.init != "mt19937") throw Spc2Error("init must be philox or mt19937");
  if (cfg_.noise != "philox" && cfg_.noise != "mt19937") throw Spc2Error("noise must be philox or mt19937");
  if (cfg_.loop != "device" && cfg_.loop != "host") throw Spc2Error("loop must be device or host");
  if (cfg_.checkpoint_every < 0) throw Spc2Error("checkpoint_every must be >= 0");
  if (cfg_.loop == "host" && (!cfg_.resume.empty() || !cfg_.checkpoint.empty()))
    throw Spc2Error("checkpoints need loop=device");
  ParseMethod(cfg_.method);
  const Precision prec = ParsePrecision(cfg_.precision);
  const NeighbourMode nb = ParseNeighbour(cfg_.neighbour);
  if (cfg_.system == "chimera") {
    system_ = new ChimeraSystem(cfg_.v0, cfg_.sigma, cfg_.rho, cfg_.alpha, cfg_.D, cfg_.rho0, pbc_, cfg_.n, prec,
                                cfg_.seed, nb, cfg_.verlet_skin);
  } else if (cfg_.system == "vortex") {
    system_ = new VortexArraysSystem(cfg_.v0, cfg_.mu_plus, cfg_.mu_minus, cfg_.xi_a, cfg_.xi_r, cfg_.D, cfg_.kappa,
                                     cfg_.rho, pbc_, cfg_.n, prec, cfg_.seed, nb, cfg_.verlet_skin);
  } else {
    throw Spc2Error("system must be chimera or vortex");
  }
  mt_.seed((std::mt19937::result_type)cfg_.seed);
}

SimulationEngine::~SimulationEngine() { delete system_; }

// Philox: x, y ~ U[0, L), phi ~ U[0, 2 pi), omega ~ N(0, omega_std) keyed by
// the seed.  mt19937: the reference's draws (SimulationEnginePtr.cpp:31-50)
// from a std::mt19937 seeded with the seed, then N normals for the natural
// frequencies when the method is SRK (StochasticRungeKuttaStepperPtr.hpp:25-35).
void SimulationEngine::InitializeRandomSystemState() {
  if (cfg_.init == "philox") {
    Spc2Check(spc2_init_random_state(system_->handle(), cfg_.omega_std), "spc2_init_random_state");
    return;
  }
  const long n = cfg_.n;
  std::vector<Real> st((size_t)kS * (size_t)n);
  std::uniform_real_distribution<Real> u_L(0.0, cfg_.L);
  std::uniform_real_distribution<Real> u_2pi(0.0, 2.0 * M_PI);
  for (long i = 0; i < n; ++i) {
    st[kS * i] = u_L(mt_);
    st[kS * i + 1] = u_L(mt_);
    st[kS * i + 2] = u_2pi(mt_);
  }
  system_->Upload(st.data());
  std::vector<Real> omega((size_t)n, 0.0);
  if (ParseMethod(cfg_.method) == SPC2_SRK) {
    std::normal_distribution<Real> z(0.0, 1.0);
    for (Real &w : omega) w = z(mt_) * cfg_.omega_std + 0.0;
  }
  system_->SetNaturalFrequencies(omega);
}

// Frame `frame` of a float trajectory (SimulationEnginePtr.cpp:52-73):
// skip frame * (1 + 3N) floats, read t and the 3N floats.
void SimulationEngine::InitializeSystemStateFromPreviousSolution(const std::string &file_name, long frame) {
  std::ifstream in(file_name, std::ios::binary | std::ios::in);
  if (!in.is_open()) throw Spc2Error("cannot open " + file_name);
  const size_t m = (size_t)kS * (size_t)cfg_.n;
  in.seekg((std::streamoff)((1 + m) * (size_t)frame * sizeof(float)), std::ios::beg);
  float t = 0.0f;
  std::vector<float> f(m);
  in.read(reinterpret_cast<char *>(&t), sizeof(float));
  in.read(reinterpret_cast<char *>(f.data()), (std::streamsize)(m * sizeof(float)));
  if (!in) throw Spc2Error("frame " + std::to_string(frame) + " is beyond the end of " + file_name);
  std::vector<Real> st(f.begin(), f.end());
  system_->Upload(st.data());
}

// SimulationEngine.cpp:52-73: 3N floats from the start of the file (the time
// stamp of a frame is not there), widened to the state.
void SimulationEngine::InitializeSystemStateFromFile(const std::string &file_name) {
  std::ifstream in(file_name, std::ios::binary | std::ios::in);
  if (!in.is_open()) throw Spc2Error("cannot open " + file_name);
  const size_t m = (size_t)kS * (size_t)cfg_.n;
  std::vector<float> f(m);
  in.read(reinterpret_cast<char *>(f.data()), (std::streamsize)(m * sizeof(float)));
  if (!in) throw Spc2Error(file_name + " holds fewer than 3N floats");
  std::vector<Real> st(f.begin(), f.end());
  system_->Upload(st.data());
}

namespace {
constexpr char kEngineCkptMagic[8] = {'S', 'P', 'C', '2', 'E', 'N', 'G', '1'};
}

// Engine checkpoint: magic, step index, t, the library blob, the mt19937
// stream as text (its standard serialisation).
void SimulationEngine::SaveCheckpoint(const std::string &file_name, Real t) const {
  uint64_t bytes = 0;
  Spc2Check(spc2_checkpoint_size(system_->handle(), &bytes), "spc2_checkpoint_size");
  std::vector<char> blob((size_t)bytes);
  Spc2Check(spc2_checkpoint_save(system_->handle(), blob.data(), bytes), "spc2_checkpoint_save");
  std::ostringstream mt;
  mt << mt_;
  const std::string mts = mt.str();
  const std::string tmp = file_name + ".tmp";
  {
    std::ofstream out(tmp, std::ios::binary | std::ios::out | std::ios::trunc);
    if (!out.is_open()) throw Spc2Error("cannot open " + tmp);
    const int64_t step = step_;
    const uint64_t mlen = mts.size();
    out.write(kEngineCkptMagic, sizeof(kEngineCkptMagic));
    out.write(reinterpret_cast<const char *>(&step), sizeof(step));
    out.write(reinterpret_cast<const char *>(&t), sizeof(t));
    out.write(reinterpret_cast<const char *>(&bytes), sizeof(bytes));
    out.write(blob.data(), (std::streamsize)bytes);
    out.write(reinterpret_cast<const char *>(&mlen), sizeof(mlen));
    out.write(mts.data(), (std::streamsize)mlen);
    if (!out) throw Spc2Error("writing " + tmp + " failed");
  }
  if (std::rename(tmp.c_str(), file_name.c_str()) != 0) throw Spc2Error("cannot rename " + tmp);
}

Real SimulationEngine::LoadCheckpoint(const std::string &file_name) {
  std::ifstream in(file_name, std::ios::binary | std::ios::in);
  if (!in.is_open()) throw Spc2Error("cannot open " + file_name);
  char magic[8];
  int64_t step = 0;
  Real t = 0.0;
  uint64_t bytes = 0, mlen = 0;
  in.read(magic, sizeof(magic));
  if (!in || std::memcmp(magic, kEngineCkptMagic, sizeof(magic)) != 0)
    throw Spc2Error(file_name + " is not an engine checkpoint");
  in.read(reinterpret_cast<char *>(&step), sizeof(step));
  in.read(reinterpret_cast<char *>(&t), sizeof(t));
  in.read(reinterpret_cast<char *>(&bytes), sizeof(bytes));
  if (!in || bytes > (uint64_t)1 << 40) throw Spc2Error(file_name + ": truncated checkpoint");
  std::vector<char> blob((size_t)bytes);
  in.read(blob.data(), (std::streamsize)bytes);
  in.read(reinterpret_cast<char *>(&mlen), sizeof(mlen));
  if (!in || mlen > (uint64_t)1 << 20) throw Spc2Error(file_name + ": truncated checkpoint");
  std::string mts((size_t)mlen, '\0');
  in.read(&mts[0], (std::streamsize)mlen);
  if (!in) throw Spc2Error(file_name + ": truncated checkpoint");
  Spc2Check(spc2_checkpoint_load(system_->handle(), blob.data(), bytes), "spc2_checkpoint_load");
  std::istringstream mt(mts);
  mt >> mt_;
  if (!mt) throw Spc2Error(file_name + ": bad mt19937 state");
  step_ = step;
  return t;
}

std::vector<Real> SimulationEngine::State() const {
  std::vector<Real> st((size_t)kS * (size_t)cfg_.n);
  system_->Download(st.data());
  return st;
}

// Number of loop iterations of `while (t <= t_1) t += dt` from t_start.
long SimulationEngine::StepCount(Real t_start) const {
  if (cfg_.t1 < 0.0) return cfg_.steps;
  long k = 0;
  for (Real t = t_start; t <= cfg_.t1; t += cfg_.dt) ++k;
  return k;
}

// The engine time loop (SimulationEnginePtr.cpp:100-115, :154-169): frame at
// t_0, then per step t += dt, DoStep, wrap, frame and summary on their own
// intervals.  The state stays on the device; steps between two outputs go to
// the device in one call.
void SimulationEngine::Run() {
  const int32_t method = ParseMethod(cfg_.method);
  Real t = cfg_.t0;
  if (!cfg_.resume.empty()) {
    t = LoadCheckpoint(cfg_.resume);
  } else if (!cfg_.init_file.empty()) {
    InitializeSystemStateFromPreviousSolution(cfg_.init_file, cfg_.init_frame);
  } else if (!cfg_.init_state_file.empty()) {
    InitializeSystemStateFromFile(cfg_.init_state_file);
  } else {
    InitializeRandomSystemState();
  }

  const long total = StepCount(t);
  if (cfg_.loop == "host") {
    RunHostLoop(total);
    return;
  }
  BinaryObserver obs(cfg_.state_path, cfg_.summary_path, cfg_.state_every, cfg_.summary_every, cfg_.float_summary);
  obs.SaveSystemState(*system_, t);
  const bool mt_noise = cfg_.noise == "mt19937" && method != SPC2_RK4;
  const long per_step = method == SPC2_SRK ? 2 * cfg_.n : cfg_.n;
  std::vector<Real> noise(mt_noise ? (size_t)per_step : 0);

  auto t_start = std::chrono::steady_clock::now();
  long done = 0;
  while (done < total) {
    long k = std::min(total - done, obs.CallsUntilOutput());
    if (cfg_.checkpoint_every > 0 && !cfg_.checkpoint.empty())
      k = std::min(k, cfg_.checkpoint_every - done % cfg_.checkpoint_every);
    if (mt_noise) {
      // the reference's per-step draws: a fresh N(0,1) per DoStep, particles
      // in index order, (U_1, U_2) per particle for SRK
      for (long j = 0; j < k; ++j) {
        std::normal_distribution<Real> z(0.0, 1.0);
        for (Real &v : noise) v = z(mt_);
        Spc2Check(spc2_step(system_->handle(), method, cfg_.dt, step_ + j, 1, noise.data()), "spc2_step");
      }
    } else {
      Spc2Check(spc2_step(system_->handle(), method, cfg_.dt, step_, k, nullptr), "spc2_step");
    }
    for (long j = 0; j < k; ++j) t += cfg_.dt;
    step_ += k;
    done += k;
    if (k > 1) obs.SkipCalls(k - 1);
    obs.SaveSystemState(*system_, t);
    obs.SaveSummaryStatistics(*system_, t);
    if (cfg_.checkpoint_every > 0 && !cfg_.checkpoint.empty() && done % cfg_.checkpoint_every == 0)
      SaveCheckpoint(cfg_.checkpoint, t);
  }
  obs.Flush();
  seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  if (!cfg_.checkpoint.empty()) SaveCheckpoint(cfg_.checkpoint, t);
}

// The reference's loop verbatim (SimulationEnginePtr.cpp:104-115) on a host
// state through the drop-in classes: DoStep, ApplyPeriodicBoundaryConditions,
// SaveSystemState, SaveSummaryStatistics (float form) every step.
void SimulationEngine::RunHostLoop(long total) {
  const int32_t method = ParseMethod(cfg_.method);
  std::vector<Real> state = State();
  const long size = (long)state.size();
  std::mt19937 *gen = cfg_.noise == "mt19937" ? &mt_ : nullptr;
  StochasticEulerStepper se(gen);
  StochasticRungeKuttaStepper srk(gen);
  RungeKutta4Stepper rk4;
  auto *chimera = dynamic_cast<ChimeraSystem *>(system_);
  auto *vortex = dynamic_cast<VortexArraysSystem *>(system_);
  BinaryObserver obs(cfg_.state_path, cfg_.summary_path, cfg_.state_every, cfg_.summary_every, true);
  Real t = cfg_.t0;
  obs.SaveSystemState(state.data(), size, t);
  auto t_start = std::chrono::steady_clock::now();
  for (long k = 0; k < total; ++k) {
    t += cfg_.dt;
    if (chimera) {
      if (method == SPC2_SE) se.DoStep(*chimera, state.data(), t, cfg_.dt);
      else if (method == SPC2_SRK) srk.DoStep(*chimera, state.data(), t, cfg_.dt);
      else rk4.DoStep(*chimera, state.data(), t, cfg_.dt);
    } else {
      if (method == SPC2_SE) se.DoStep(*vortex, state.data(), t, cfg_.dt);
      else if (method == SPC2_SRK) srk.DoStep(*vortex, state.data(), t, cfg_.dt);
      else rk4.DoStep(*vortex, state.data(), t, cfg_.dt);
    }
    pbc_.ApplyPeriodicBoundaryConditions(state.data(), size);
    obs.SaveSystemState(state.data(), size, t);
    obs.SaveSummaryStatistics(state.data(), size, t);
  }
  obs.Flush();
  seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  step_ += total;
  system_->Upload(state.data());
}

void SimulationEngine::RunChimeraSimulation() {
  if (cfg_.system != "chimera") throw Spc2Error("RunChimeraSimulation needs system=chimera");
  Run();
}

void SimulationEngine::RunVortexArraySimulation() {
  if (cfg_.system != "vortex") throw Spc2Error("RunVortexArraySimulation needs system=vortex");
  Run();
}
