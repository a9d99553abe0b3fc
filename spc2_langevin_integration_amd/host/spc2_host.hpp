// spc2 host layer: the reference's C++ class interfaces, forwarding to the
// MI355X library through its C ABI (include/spc2_hip.h).
//
// Class and method names, argument order and in/out semantics follow the
// reference so that a caller of it can switch:
//   PeriodicBoundaryConditions   BoundaryConfiguration/PeriodicBoundaryConditions.hpp:12-39
//   ChimeraSystem                DynamicalSystems/ChimeraSystemPtr.hpp:12-71
//   VortexArraysSystem           DynamicalSystems/VortexArraysSystemPtr.hpp:12-60
//   StochasticEulerStepper       Steppers/StochasticEulerStepperPtr.hpp:21-76
//   StochasticRungeKuttaStepper  Steppers/StochasticRungeKuttaStepperPtr.hpp:21-97
//   RungeKutta4Stepper           Steppers/RungeKutta4StepperPtr.hpp:15-33
//   BinaryObserver               Observers/BinaryObserverPtr.hpp:44-45
//   SimulationEngine             ControlEngines/SimulationEnginePtr.hpp:14-34
// What changes: N, L and the model parameters are run-time values (the
// reference fixes kN, kL at compile time and hard-codes the parameters in the
// engine), there is no MPI Thread argument, and the state lives on the GPU:
// the Real* state a stepper receives is a host mirror copied in and out around
// each DoStep, while SimulationEngine keeps the state resident and copies it
// out only for observer frames.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <fstream>
#include <limits>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/spc2_hip.h"

typedef double Real;
typedef float RealOutput;
constexpr int kS = 3;  // (x, y, phi) per particle, Definitions.hpp:21

// Raised for any failing spc2 call (the reference has no error channel).
class Spc2Error : public std::runtime_error {
 public:
  explicit Spc2Error(const std::string &what) : std::runtime_error(what) {}
};

void Spc2Check(spc2_status st, const char *what);

// Host geometry helpers with the reference arithmetic
// (PeriodicBoundaryConditions.cpp:7-151); the device path has its own copy.
class PeriodicBoundaryConditions {
 public:
  PeriodicBoundaryConditions(Real x_size, Real y_size);
  void ClassAEffectiveParticleDistance(Real x_i, Real y_i, Real x_j, Real y_j, Real &dx, Real &dy) const;
  void ApplyPeriodicBoundaryConditions(Real x, Real y, Real &x_pbc, Real &y_pbc) const;
  void ApplyPeriodicBoundaryConditions(std::vector<Real> &system_state) const;
  void ApplyPeriodicBoundaryConditions(Real *const system_state, long size) const;
  Real GetXSize() const { return x_size_; }
  Real GetYSize() const { return y_size_; }

 private:
  Real x_size_, y_size_, x_rsize_, y_rsize_;
};

// One device-resident simulation (owner of an spc2_sim_t).
class DeviceSystem {
 public:
  DeviceSystem(const spc2_params &p);
  virtual ~DeviceSystem();
  DeviceSystem(const DeviceSystem &) = delete;
  DeviceSystem &operator=(const DeviceSystem &) = delete;

  spc2_sim_t handle() const { return h_; }
  long n() const { return (long)p_.n; }
  const spc2_params &params() const { return p_; }
  void Upload(const Real *state);                 // host AoS -> device
  void Download(Real *state) const;               // device -> host AoS
  void SetNaturalFrequencies(const std::vector<Real> &omega);

 protected:
  spc2_params p_;
  spc2_sim_t h_ = nullptr;
};

enum class Precision { kFP64 = SPC2_FP64, kFP32Kahan = SPC2_FP32_KAHAN, kFP64Factored = SPC2_FP64_FACTORED };
// neighbour search: the reference's dispatch (linked cells when rho <= L/4, else
// all pairs), linked cells, all pairs, or a Verlet list of skin verlet_skin
// (0: 0.2 rho)
enum class NeighbourMode {
  kAuto = SPC2_NEIGH_AUTO,
  kCells = SPC2_NEIGH_CELLS,
  kAllPairs = SPC2_NEIGH_ALL_PAIRS,
  kVerlet = SPC2_NEIGH_VERLET
};

class ChimeraSystem : public DeviceSystem {
 public:
  // reference order (ChimeraSystemPtr.hpp:16-23) with N, precision and the noise
  // seed appended; L comes from the boundary object
  ChimeraSystem(Real v_0, Real sigma, Real rho, Real alpha, Real D_phi, Real rho_0,
                PeriodicBoundaryConditions &pbc_config, long n, Precision precision = Precision::kFP64,
                uint64_t seed = 42, NeighbourMode neighbour = NeighbourMode::kAuto, Real verlet_skin = 0.0);
  void EvaluateRhs(const Real *const system_state, const std::vector<Real> &k_prev, std::vector<Real> &k_next,
                   Real k_coef, Real dt);
  void EvaluateRhs(const Real *const system_state, std::vector<Real> &derivative,
                   std::vector<std::vector<Real>> &additional_derivative, Real dt);
  Real rho() const { return p_.rho; }
  Real D_phi() const { return p_.D; }
  Real alpha() const { return p_.alpha; }
  bool last_coefficient_ = true;
};

class VortexArraysSystem : public DeviceSystem {
 public:
  // reference order (VortexArraysSystemPtr.hpp:16-25) with N, precision and seed appended
  VortexArraysSystem(Real velocity, Real mu_plus, Real mu_minus, Real xi_a, Real xi_r, Real D_phi, Real kappa,
                     Real rho, PeriodicBoundaryConditions &pbc_config, long n, Precision precision = Precision::kFP64,
                     uint64_t seed = 42, NeighbourMode neighbour = NeighbourMode::kAuto,
                     Real verlet_skin = 0.0);
  void EvaluateRhs(const Real *const system_state, const std::vector<Real> &k_prev, std::vector<Real> &k_next,
                   Real k_coef, Real dt);
  void EvaluateRhs(const Real *const system_state, std::vector<Real> &derivative,
                   std::vector<std::vector<Real>> &additional_derivative, Real dt);
  Real D_phi() const { return p_.D; }
  bool last_coefficient_ = true;
};

// Steppers: DoStep(system, state, t, dt) advances the host state by one step on
// the device without the periodic wrap (the reference's engine wraps after
// DoStep, SimulationEnginePtr.cpp:109-111).  Noise: with a generator, the
// reference's draws (a fresh std::normal_distribution per DoStep, particles in
// index order, StochasticRungeKuttaStepperPtr.hpp:49-56); without one, Philox
// keyed by the stepper's own step counter.
class StochasticEulerStepper {
 public:
  explicit StochasticEulerStepper(std::mt19937 *generator = nullptr) : gen_(generator) {}
  void DoStep(ChimeraSystem &system, Real *const system_state, Real t, Real dt);
  void DoStep(VortexArraysSystem &system, Real *const system_state, Real t, Real dt);
  int64_t steps_taken() const { return step_; }

 private:
  void Step(DeviceSystem &system, Real *const system_state, Real dt);
  std::mt19937 *gen_;
  std::vector<Real> noise_;
  int64_t step_ = 0;
};

class StochasticRungeKuttaStepper {
 public:
  // n_frequencies > 0 draws that many natural frequencies N(0, 0) from the
  // generator first, as the reference's constructor does
  // (StochasticRungeKuttaStepperPtr.hpp:25-35); they are zero but consume draws.
  explicit StochasticRungeKuttaStepper(std::mt19937 *generator = nullptr, long n_frequencies = 0);
  void DoStep(ChimeraSystem &system, Real *const system_state, Real t, Real dt);
  void DoStep(VortexArraysSystem &system, Real *const system_state, Real t, Real dt);
  int64_t steps_taken() const { return step_; }

 private:
  void Step(DeviceSystem &system, Real *const system_state, Real dt);
  std::mt19937 *gen_;
  std::vector<Real> noise_;
  int64_t step_ = 0;
};

class RungeKutta4Stepper {
 public:
  void DoStep(ChimeraSystem &system, Real *const system_state, Real t, Real dt);
  void DoStep(VortexArraysSystem &system, Real *const system_state, Real t, Real dt);
};

// Trajectory frames (float t + 3N floats, BinaryObserverPtr.cpp:163-168) and
// summary lines "t \t |R| \t arg R" (:246-248), each every `*_every` calls
// starting with the first (:157, :235).  Frames of a device-resident system are
// copied asynchronously into a ring of pinned buffers and written by a
// background thread.
class BinaryObserver {
 public:
  // float_summary = true computes the order parameter on the host exactly as
  // the reference (complex<float> sum in index order, BinaryObserverPtr.cpp:238-247);
  // false uses the device FP64 reduction.
  BinaryObserver(const std::string &state_path, const std::string &summary_path, int state_every,
                 int summary_every, bool float_summary = false);
  ~BinaryObserver();
  // reference signatures, on a host state
  void SaveSystemState(Real *const system_state, long size, Real t);
  void SaveSummaryStatistics(const Real *const system_state, long size, Real t);
  // device-resident forms
  void SaveSystemState(DeviceSystem &system, Real t);
  void SaveSummaryStatistics(DeviceSystem &system, Real t);
  // calls (of each of the two Save functions) until the next one that writes
  long CallsUntilOutput() const;
  // advances both counters as k calls that write nothing would
  void SkipCalls(long k);
  // waits until every frame issued so far is on disk
  void Flush();

 private:
  struct Frame {
    float *buf = nullptr;
    float t = 0.0f;
    bool busy = false;
  };
  void WriterLoop();
  void Enqueue(int idx);
  int AcquireFrame(size_t floats);
  void RetirePending();
  void WriteSummaryLine(Real t, float R, float psi);
  std::ofstream state_file_, summary_file_;
  int state_every_, summary_every_;
  bool float_summary_;
  long counter_state_ = 0, counter_summary_ = 0;
  // ring of pinned frames; `pending_` is the frame whose copy is in flight
  std::vector<Frame> ring_;
  size_t frame_floats_ = 0;
  int next_ = 0, pending_ = -1;
  spc2_sim_t pending_handle_ = nullptr;
  std::deque<int> queue_;
  std::mutex mu_;
  std::condition_variable cv_;
  bool stop_ = false;
  std::thread writer_;
};

// Run-time configuration replacing the engine's hard-coded values
// (SimulationEnginePtr.cpp:82-91, :131-143) and Definitions.hpp:20-22.
struct EngineConfig {
  std::string system = "chimera";  // chimera | vortex
  std::string method = "srk";      // se | srk | rk4
  std::string precision = "fp64";  // fp64 | fp32k | fp64f
  std::string init = "philox";     // philox | mt19937 (the reference's draws from std::mt19937(seed))
  std::string noise = "philox";    // philox | mt19937
  // device: state resident, steps between outputs in one call; host: the
  // reference's loop through the drop-in classes on a host state
  std::string loop = "device";
  long n = 1000;
  Real L = 1.0;
  // chimera
  Real v0 = 0.01, sigma = 1.0, rho = 0.1, alpha = 1.45, D = 0.01, rho0 = 1.0, omega_std = 0.0;
  // vortex
  Real mu_plus = 100.0, mu_minus = 100.0, xi_a = 0.01, xi_r = 0.005, kappa = 20.0;
  Real dt = 0.01;
  long steps = 1000;
  // t1 >= 0: step while t <= t1 from t0 (the reference's loop), else `steps` steps
  Real t0 = 0.0, t1 = -1.0;
  uint64_t seed = 42;
  std::string state_path, summary_path;  // empty: no output
  int state_every = 10, summary_every = 10;
  bool float_summary = false;
  std::string neighbour = "auto";  // auto | cells | all_pairs | verlet
  Real verlet_skin = 0.0;          // Verlet list skin (0: 0.2 rho)
  std::string init_file;   // non-empty: start from a frame of a previous trajectory
  long init_frame = 0;
  std::string init_state_file;  // non-empty: start from a file of 3N floats (SimulationEngine.cpp:52-73)
  // exact restarts: `resume` loads an engine checkpoint (state, k buffers,
  // frequencies, Verlet list, t, step index and the mt19937 stream) and runs
  // `steps` more steps (or until t1); `checkpoint` writes one at the end of the
  // run and, with checkpoint_every > 0, every checkpoint_every steps
  std::string resume, checkpoint;
  long checkpoint_every = 0;
  // key=value parser used by the command-line runner
  void Set(const std::string &key, const std::string &value);
};

class SimulationEngine {
 public:
  explicit SimulationEngine(const EngineConfig &cfg);
  ~SimulationEngine();
  void InitializeRandomSystemState();
  // frame `frame` of a float trajectory file (SimulationEnginePtr.cpp:52-73 reads frame 2900)
  void InitializeSystemStateFromPreviousSolution(const std::string &file_name, long frame);
  // 3N floats from the start of a file, no time stamp (SimulationEngine.cpp:52-73)
  void InitializeSystemStateFromFile(const std::string &file_name);
  // exact checkpoint of the run so far (the library's checkpoint of the
  // handle, the time, the step index and the mt19937 stream), and its reverse
  void SaveCheckpoint(const std::string &file_name, Real t) const;
  Real LoadCheckpoint(const std::string &file_name);
  void RunChimeraSimulation();
  void RunVortexArraySimulation();
  // host copy of the current state (AoS, by particle id)
  std::vector<Real> State() const;
  DeviceSystem &system() { return *system_; }
  double seconds_in_loop() const { return seconds_; }
  int64_t steps_done() const { return step_; }

 private:
  void Run();
  void RunHostLoop(long total);
  long StepCount(Real t_start) const;
  EngineConfig cfg_;
  PeriodicBoundaryConditions pbc_;
  DeviceSystem *system_ = nullptr;
  std::mt19937 mt_;
  int64_t step_ = 0;
  double seconds_ = 0.0;
};
