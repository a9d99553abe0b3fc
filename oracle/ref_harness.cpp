// Golden-vector harness around the reference's own non-MPI classes.
//
// Test infrastructure only: this file is compiled together with the reference
// translation units (ChimeraSystem.cpp, LinkedList.cpp,
// PeriodicBoundaryConditions.cpp, Thread.cpp, RungeKutta4Stepper.cpp and the
// header-only steppers) where they lie, by oracle/Makefile, into
// oracle/_ref/ref_harness.  It never ships and nothing in the product links it.
//
// The reference fixes N = kN = 8192 and L = kL = 64 at compile time
// (Definitions.hpp:20-22); every case here therefore runs at that size.  The
// reference's main.cpp (which defines the global generator and starts MPI) is
// not compiled; this harness defines the generator instead
// (Definitions.hpp:24 declares it `extern`).
//
// Inputs are generated with splitmix64 so that the oracle (oracle/spc2_oracle.c)
// and the tests regenerate them bit-for-bit from the seed alone.
//
// Usage: ref_harness <case> <outdir> [args...]; every output is a raw
// little-endian array file <outdir>/<name>.bin.

#include "DynamicalSystems/ChimeraSystem.hpp"
#include "ForceSummation/LinkedList.hpp"
#include "BoundaryConfiguration/PeriodicBoundaryConditions.hpp"
#include "Parallelization/Thread.hpp"
#include "Steppers/StochasticEulerStepper.hpp"
#include "Steppers/StochasticRungeKuttaStepper.hpp"
#include "Steppers/RungeKutta4Stepper.hpp"

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

std::mt19937 mersenne_twister_generator;

namespace {

const double kTwoPi = 2.0 * M_PI;

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  // uniform in [0, 1) with 53 random bits
  double u01() { return double(next() >> 11) * 0x1.0p-53; }
};

std::string g_outdir;

template <class T>
void Write(const std::string &name, const std::vector<T> &v) {
  std::string path = g_outdir + "/" + name + ".bin";
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) { std::perror(path.c_str()); std::exit(2); }
  if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
  std::fclose(f);
}

// State layout kS = 3 AoS (x, y, phi), Definitions.hpp:17-21.
// mode 0: x, y ~ U[0, L), phi ~ U[0, 2pi)
// mode 1: x, y ~ U[-L/4, 5L/4)  (exercises the conditional wrap in the binning)
// mode 2: x, y on a lattice with spacing 1/4 cell (particles exactly on cell edges)
std::vector<Real> MakeState(uint64_t seed, int mode, double rho) {
  SplitMix g(seed);
  std::vector<Real> s(kS * kN);
  for (int i = 0; i < kN; ++i) {
    double ux = g.u01(), uy = g.u01(), up = g.u01();
    if (mode == 0) {
      s[kS * i] = ux * kL;
      s[kS * i + 1] = uy * kL;
    } else if (mode == 1) {
      s[kS * i] = (ux * 1.5 - 0.25) * kL;
      s[kS * i + 1] = (uy * 1.5 - 0.25) * kL;
    } else {
      int n = int(kL / rho);
      double h = kL / n;
      s[kS * i] = double(int(ux * 4 * n)) * 0.25 * h;
      s[kS * i + 1] = double(int(uy * 4 * n)) * 0.25 * h;
    }
    s[kS * i + 2] = up * kTwoPi;
  }
  return s;
}

std::vector<Real> MakeUniform(uint64_t seed, size_t n, double lo, double hi) {
  SplitMix g(seed);
  std::vector<Real> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = lo + (hi - lo) * g.u01();
  return v;
}

struct ChimeraArgs {
  double v0, sigma, rho, alpha, D;
};

ChimeraArgs ParseChimera(char **a) {
  return ChimeraArgs{std::atof(a[0]), std::atof(a[1]), std::atof(a[2]), std::atof(a[3]), std::atof(a[4])};
}

// Flatten the reference's vector<vector<int>> cell list to CSR.
void FlattenCells(const std::vector<std::vector<int>> &ll, std::vector<int32_t> &start,
                  std::vector<int32_t> &sorted, std::vector<int32_t> &cell_of) {
  start.assign(ll.size() + 1, 0);
  sorted.clear();
  cell_of.assign(kN, -1);
  for (size_t c = 0; c < ll.size(); ++c) {
    start[c] = int32_t(sorted.size());
    for (int i : ll[c]) {
      sorted.push_back(i);
      cell_of[i] = int32_t(c);
    }
  }
  start[ll.size()] = int32_t(sorted.size());
}

// Replays exactly the normals the reference's next DoStep will draw from the
// global generator: a copy of the engine state and a distribution in the same
// (empty-cache) state.  kN is even, so the static distribution inside
// ChimeraSystem::AddNoise (ChimeraSystem.cpp:926) has an empty cache at every
// call boundary, like a fresh one (StochasticRungeKuttaStepper.hpp:42).
std::vector<Real> PeekNormals(size_t count) {
  std::mt19937 copy = mersenne_twister_generator;
  std::normal_distribution<Real> d(0.0, 1.0);
  std::vector<Real> v(count);
  for (auto &x : v) x = d(copy);
  return v;
}

int CaseCells(char **a) {
  // cells <seed> <mode> <rho>
  uint64_t seed = std::strtoull(a[0], nullptr, 10);
  int mode = std::atoi(a[1]);
  double rho = std::atof(a[2]);
  int n = int(kL / rho);
  Thread thread(0, nullptr);
  PeriodicBoundaryConditions pbc(&thread, kL, kL);
  LinkedList ll(&thread, kL, kL, n, n, pbc);
  std::vector<Real> s = MakeState(seed, mode, rho);
  ll.ConstructLinkedList(s);
  std::vector<int32_t> start, sorted, cell_of;
  FlattenCells(ll.GetLinkedList(), start, sorted, cell_of);
  Write("cell_start", start);
  Write("sorted", sorted);
  Write("cell_of", cell_of);
  return 0;
}

int CasePbc(char **a) {
  // pbc <seed>: min-image displacements and wraps on random and crafted inputs
  uint64_t seed = std::strtoull(a[0], nullptr, 10);
  Thread thread(0, nullptr);
  PeriodicBoundaryConditions pbc(&thread, kL, kL);
  const int m = 4096;
  std::vector<Real> pts = MakeUniform(seed, 4 * m, -0.0, kL);
  // crafted: exact half-box, exact zero, exact L, just below L
  double crafted[][4] = {{0.0, 0.0, 32.0, 32.0}, {32.0, 32.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0},
                         {1.0, 2.0, 1.0 + 32.0, 2.0 - 32.0}, {0.0, 0.0, 63.999999999, 0.5},
                         {10.0, 10.0, 10.0 + 38.4, 10.0 - 38.4}, {63.0, 1.0, 0.5, 63.5}};
  for (int c = 0; c < 7; ++c)
    for (int k = 0; k < 4; ++k) pts[4 * c + k] = crafted[c][k];
  std::vector<Real> dxy(2 * m);
  for (int p = 0; p < m; ++p) {
    pbc.ClassAEffectiveParticleDistance(pts[4 * p], pts[4 * p + 1], pts[4 * p + 2], pts[4 * p + 3], dxy[2 * p],
                                        dxy[2 * p + 1]);
  }
  // wraps: scalar form on wide-range inputs and the array form
  std::vector<Real> w = MakeUniform(seed + 1, 2 * m, -2.0 * kL, 3.0 * kL);
  double wc[] = {-0.0, 0.0, 64.0, -64.0, 63.99999999999999, -1e-300, 128.0, -1e-17};
  for (int c = 0; c < 8; ++c) w[c] = wc[c];
  std::vector<Real> wrapped(2 * m);
  for (int p = 0; p < m; ++p) pbc.ApplyPeriodicBoundaryConditions(w[2 * p], w[2 * p + 1], wrapped[2 * p], wrapped[2 * p + 1]);
  Write("pts", pts);
  Write("dxy", dxy);
  Write("wrap_in", w);
  Write("wrap_out", wrapped);
  return 0;
}

int CaseRhsK(char **a) {
  // rhs_k <seed> <mode> <k_coef> <dt> <v0> <sigma> <rho> <alpha> <D>
  uint64_t seed = std::strtoull(a[0], nullptr, 10);
  int mode = std::atoi(a[1]);
  double k_coef = std::atof(a[2]), dt = std::atof(a[3]);
  ChimeraArgs p = ParseChimera(a + 4);
  Thread thread(0, nullptr);
  PeriodicBoundaryConditions pbc(&thread, kL, kL);
  ChimeraSystem sys(&thread, p.v0, p.sigma, p.rho, p.alpha, p.D, 1.0, pbc);
  std::vector<Real> s = MakeState(seed, mode, p.rho);
  // k_prev large enough that stage positions cross the box edges
  std::vector<Real> k_prev = MakeUniform(seed + 101, kS * kN, -50.0, 50.0);
  std::vector<Real> k_next = MakeUniform(seed + 202, kS * kN, 5.0, 6.0);  // sentinel: phi kept if n_i = 0
  sys.EvaluateRhs(s, k_prev, k_next, k_coef, dt);
  Write("k_next", k_next);
  return 0;
}

int CaseRhsIto(char **a) {
  // rhs_ito <seed> <mode> <init> <dt> <v0> <sigma> <rho> <alpha> <D>
  uint64_t seed = std::strtoull(a[0], nullptr, 10);
  int mode = std::atoi(a[1]);
  int init = std::atoi(a[2]);
  double dt = std::atof(a[3]);
  ChimeraArgs p = ParseChimera(a + 4);
  Thread thread(0, nullptr);
  PeriodicBoundaryConditions pbc(&thread, kL, kL);
  ChimeraSystem sys(&thread, p.v0, p.sigma, p.rho, p.alpha, p.D, 1.0, pbc);
  std::vector<Real> s = MakeState(seed, mode, p.rho);
  std::vector<Real> deriv(kS * kN, 0.0);
  std::vector<std::vector<Real>> add(kN, std::vector<Real>(2, 0.0));
  if (init) {
    deriv = MakeUniform(seed + 303, kS * kN, -1.0, 1.0);
    std::vector<Real> a2 = MakeUniform(seed + 404, 2 * kN, -1.0, 1.0);
    for (int i = 0; i < kN; ++i) add[i] = {a2[2 * i], a2[2 * i + 1]};
  }
  sys.EvaluateRhs(s, deriv, add, dt);
  std::vector<Real> add_flat(2 * kN);
  for (int i = 0; i < kN; ++i) {
    add_flat[2 * i] = add[i][0];
    add_flat[2 * i + 1] = add[i][1];
  }
  Write("derivative", deriv);
  Write("additional", add_flat);
  return 0;
}

// Runs `steps` steps of the reference time loop (SimulationEngine.cpp:115-121:
// DoStep, then the array wrap) from the seeded state, recording the normals each
// DoStep consumes, the state after the first step and the final state.
template <class Stepper>
int RunSteps(char **a, int normals_per_particle) {
  // <seed> <mtseed> <steps> <dt> <v0> <sigma> <rho> <alpha> <D>
  uint64_t seed = std::strtoull(a[0], nullptr, 10);
  uint32_t mtseed = uint32_t(std::strtoul(a[1], nullptr, 10));
  int steps = std::atoi(a[2]);
  double dt = std::atof(a[3]);
  ChimeraArgs p = ParseChimera(a + 4);
  mersenne_twister_generator.seed(mtseed);
  Thread thread(0, nullptr);
  PeriodicBoundaryConditions pbc(&thread, kL, kL);
  ChimeraSystem sys(&thread, p.v0, p.sigma, p.rho, p.alpha, p.D, 1.0, pbc);
  Stepper stepper(&thread);
  std::vector<Real> s = MakeState(seed, 0, p.rho);
  std::vector<Real> noise;
  std::vector<Real> first;
  double t = 0.0;
  for (int k = 0; k < steps; ++k) {
    if (normals_per_particle) {
      std::vector<Real> xi = PeekNormals(size_t(normals_per_particle) * kN);
      noise.insert(noise.end(), xi.begin(), xi.end());
    }
    t += dt;
    stepper.DoStep(sys, s, t, dt);
    pbc.ApplyPeriodicBoundaryConditions(s);
    if (k == 0) first = s;
  }
  Write("noise", noise);
  Write("state1", first);
  Write("state", s);
  return 0;
}

// Ensemble of independent runs of the reference SRK / SE stepper driven by its
// own mt19937 + std::normal_distribution noise: |R|(t) of the Kuramoto order
// parameter in FP64 every `every` steps.  Initial state from splitmix64(seed+r).
template <class Stepper>
int RunEnsemble(char **a) {
  // <seed> <members> <steps> <every> <dt> <v0> <sigma> <rho> <alpha> <D>
  uint64_t seed = std::strtoull(a[0], nullptr, 10);
  int members = std::atoi(a[1]);
  int steps = std::atoi(a[2]);
  int every = std::atoi(a[3]);
  double dt = std::atof(a[4]);
  ChimeraArgs p = ParseChimera(a + 5);
  Thread thread(0, nullptr);
  PeriodicBoundaryConditions pbc(&thread, kL, kL);
  std::vector<Real> out;
  for (int r = 0; r < members; ++r) {
    mersenne_twister_generator.seed(uint32_t(1000 + r));
    ChimeraSystem sys(&thread, p.v0, p.sigma, p.rho, p.alpha, p.D, 1.0, pbc);
    Stepper stepper(&thread);
    std::vector<Real> s = MakeState(seed + uint64_t(r), 0, p.rho);
    double t = 0.0;
    for (int k = 0; k <= steps; ++k) {
      if (k % every == 0) {
        double cr = 0.0, ci = 0.0;
        for (int i = 0; i < kN; ++i) {
          cr += std::cos(s[kS * i + 2]);
          ci += std::sin(s[kS * i + 2]);
        }
        out.push_back(std::sqrt(cr * cr + ci * ci) / kN);
      }
      if (k == steps) break;
      t += dt;
      stepper.DoStep(sys, s, t, dt);
      pbc.ApplyPeriodicBoundaryConditions(s);
    }
  }
  Write("R", out);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <case> <outdir> [args...]\n", argv[0]);
    return 1;
  }
  std::string c = argv[1];
  g_outdir = argv[2];
  char **a = argv + 3;
  int na = argc - 3;
  auto need = [&](int k) {
    if (na < k) {
      std::fprintf(stderr, "case %s needs %d args\n", c.c_str(), k);
      std::exit(1);
    }
  };
  if (c == "cells") { need(3); return CaseCells(a); }
  if (c == "pbc") { need(1); return CasePbc(a); }
  if (c == "rhs_k") { need(9); return CaseRhsK(a); }
  if (c == "rhs_ito") { need(9); return CaseRhsIto(a); }
  if (c == "se") { need(9); return RunSteps<StochasticEulerStepper>(a, 1); }
  if (c == "srk") { need(9); return RunSteps<StochasticRungeKuttaStepper>(a, 2); }
  if (c == "rk4") { need(9); return RunSteps<RungeKutta4Stepper>(a, 0); }
  if (c == "ens_srk") { need(10); return RunEnsemble<StochasticRungeKuttaStepper>(a); }
  if (c == "ens_se") { need(10); return RunEnsemble<StochasticEulerStepper>(a); }
  std::fprintf(stderr, "unknown case %s\n", c.c_str());
  return 1;
}
