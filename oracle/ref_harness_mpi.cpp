// Golden-vector and timing harness around the reference's MPI shared-memory
// classes: VortexArraysSystemPtr, ChimeraSystemPtr and the *StepperPtr
// steppers (the active build, Definitions.hpp:13).
//
// Test infrastructure only: compiled with MPICH's mpicxx together with the
// reference translation units (VortexArraysSystemPtr.cpp, ChimeraSystemPtr.cpp,
// RungeKutta4StepperPtr.cpp, ThreadSharedMemory.cpp, Thread.cpp,
// Parallelization.cpp, PeriodicBoundaryConditions.cpp) where they lie, by
// oracle/Makefile, into oracle/_ref/ref_harness_mpi, and run under mpiexec in
// the container that holds the reference.  It never ships.
//
// N = kN = 8192 and L = kL = 64 (Definitions.hpp:20-22).  Each rank owns the
// contiguous slice ThreadSharedMemory gives it (ThreadSharedMemory.cpp:12-22);
// outputs are gathered to rank 0, which writes them as raw little-endian
// arrays <outdir>/<name>.bin.  Inputs come from splitmix64 exactly as in
// ref_harness.cpp, so the oracle regenerates them from the seed.
//
// Usage: mpiexec -n P ref_harness_mpi <case> <outdir> [args...]

#include "DynamicalSystems/ChimeraSystemPtr.hpp"
#include "DynamicalSystems/VortexArraysSystemPtr.hpp"
#include "BoundaryConfiguration/PeriodicBoundaryConditions.hpp"
#include "Parallelization/Parallelization.hpp"
#include "Parallelization/ThreadSharedMemory.hpp"
#include "Steppers/RungeKutta4StepperPtr.hpp"
#include "Steppers/StochasticEulerStepperPtr.hpp"
#include "Steppers/StochasticRungeKuttaStepperPtr.hpp"

#include <mpi.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
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
  double u01() { return double(next() >> 11) * 0x1.0p-53; }
};

std::string g_outdir;
int g_rank = 0, g_size = 1;

template <class T>
void Write(const std::string &name, const std::vector<T> &v) {
  if (g_rank != 0) return;
  std::string path = g_outdir + "/" + name + ".bin";
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) { std::perror(path.c_str()); MPI_Abort(MPI_COMM_WORLD, 2); }
  if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
  std::fclose(f);
}

// The same inputs as ref_harness.cpp (mode 0: uniform in the box; mode 1:
// x, y ~ U[-L/4, 5L/4)).
std::vector<Real> MakeState(uint64_t seed, int mode) {
  SplitMix g(seed);
  std::vector<Real> s(kS * kN);
  for (int i = 0; i < kN; ++i) {
    double ux = g.u01(), uy = g.u01(), up = g.u01();
    if (mode == 0) {
      s[kS * i] = ux * kL;
      s[kS * i + 1] = uy * kL;
    } else {
      s[kS * i] = (ux * 1.5 - 0.25) * kL;
      s[kS * i + 1] = (uy * 1.5 - 0.25) * kL;
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

// Rank r's slice [r kN / P, (r + 1) kN / P) of a per-particle array of `comps`
// values, collected on rank 0 (the other entries of each rank's copy are
// whatever that rank holds, as in the reference).
std::vector<Real> Gather(const std::vector<Real> &v, int comps) {
  const int per = kN / g_size * comps;
  std::vector<Real> out(g_rank == 0 ? (size_t)kN * comps : 0);
  MPI_Gather(v.data() + (size_t)g_rank * per, per, MPI_DOUBLE, out.data(), per, MPI_DOUBLE, 0, MPI_COMM_WORLD);
  return out;
}

std::vector<Real> Flatten(const std::vector<std::vector<Real>> &a) {
  std::vector<Real> f(2 * a.size());
  for (size_t i = 0; i < a.size(); ++i) {
    f[2 * i] = a[i][0];
    f[2 * i + 1] = a[i][1];
  }
  return f;
}

// vortex arrays: <velocity> <mu+> <mu-> <xi_a> <xi_r> <D> <kappa> <rho>
struct VortexArgs {
  double v, mp, mm, xa, xr, D, kappa, rho;
};
VortexArgs ParseVortex(char **a) {
  return VortexArgs{std::atof(a[0]), std::atof(a[1]), std::atof(a[2]), std::atof(a[3]),
                    std::atof(a[4]), std::atof(a[5]), std::atof(a[6]), std::atof(a[7])};
}
// chimera: <v0> <sigma> <rho> <alpha> <D>
struct ChimeraArgs {
  double v0, sigma, rho, alpha, D;
};
ChimeraArgs ParseChimera(char **a) {
  return ChimeraArgs{std::atof(a[0]), std::atof(a[1]), std::atof(a[2]), std::atof(a[3]), std::atof(a[4])};
}

struct VortexSys {
  VortexArraysSystemPtr sys;
  VortexSys(ThreadSharedMemory &th, PeriodicBoundaryConditions &pbc, const VortexArgs &p)
      : sys(&th, p.v, p.mp, p.mm, p.xa, p.xr, p.D, p.kappa, p.rho, pbc) {}
};
struct ChimeraSys {
  ChimeraSystemPtr sys;
  ChimeraSys(ThreadSharedMemory &th, PeriodicBoundaryConditions &pbc, const ChimeraArgs &p)
      : sys(&th, p.v0, p.sigma, p.rho, p.alpha, p.D, 1.0, pbc) {}
};

// k-form RHS: <seed> <mode> <k_coef> <dt> <system args>
template <class Sys, class Args>
int RhsK(ThreadSharedMemory &th, char **a, Args p) {
  uint64_t seed = std::strtoull(a[0], nullptr, 10);
  int mode = std::atoi(a[1]);
  double k_coef = std::atof(a[2]), dt = std::atof(a[3]);
  PeriodicBoundaryConditions pbc(&th, kL, kL);
  Sys s(th, pbc, p);
  std::vector<Real> st = MakeState(seed, mode);
  std::vector<Real> k_prev = MakeUniform(seed + 101, kS * kN, -1.0, 1.0);
  std::vector<Real> k_next = MakeUniform(seed + 202, kS * kN, 5.0, 6.0);  // phi kept where no partner
  s.sys.EvaluateRhs(st.data(), k_prev, k_next, k_coef, dt);
  Write("k_next", Gather(k_next, kS));
  return 0;
}

// Ito form: <seed> <mode> <init> <dt> <system args>
template <class Sys, class Args>
int RhsIto(ThreadSharedMemory &th, char **a, Args p) {
  uint64_t seed = std::strtoull(a[0], nullptr, 10);
  int mode = std::atoi(a[1]);
  int init = std::atoi(a[2]);
  double dt = std::atof(a[3]);
  PeriodicBoundaryConditions pbc(&th, kL, kL);
  Sys s(th, pbc, p);
  std::vector<Real> st = MakeState(seed, mode);
  std::vector<Real> deriv(kS * kN, 0.0);
  std::vector<std::vector<Real>> add(kN, std::vector<Real>(2, 0.0));
  if (init) {
    deriv = MakeUniform(seed + 303, kS * kN, -1.0, 1.0);
    std::vector<Real> a2 = MakeUniform(seed + 404, 2 * kN, -1.0, 1.0);
    for (int i = 0; i < kN; ++i) add[i] = {a2[2 * i], a2[2 * i + 1]};
  }
  s.sys.EvaluateRhs(st.data(), deriv, add, dt);
  Write("derivative", Gather(deriv, kS));
  Write("additional", Gather(Flatten(add), 2));
  return 0;
}

// The normals the next DoStep draws on a single rank: a copy of the generator
// and a fresh distribution, as each DoStep constructs one.
std::vector<Real> PeekNormals(size_t count) {
  std::mt19937 copy = mersenne_twister_generator;
  std::normal_distribution<Real> d(0.0, 1.0);
  std::vector<Real> v(count);
  for (auto &x : v) x = d(copy);
  return v;
}

// `steps` steps of the engine loop (SimulationEnginePtr.cpp:104-115: DoStep,
// then the array wrap) from the seeded state; records the normals of each step
// (single rank), the state after the first step and the final state.
// <seed> <mtseed> <steps> <dt> <system args>
template <class Stepper, class Sys, class Args>
int Steps(ThreadSharedMemory &th, char **a, Args p, int normals_per_particle) {
  uint64_t seed = std::strtoull(a[0], nullptr, 10);
  uint32_t mtseed = uint32_t(std::strtoul(a[1], nullptr, 10));
  int steps = std::atoi(a[2]);
  double dt = std::atof(a[3]);
  mersenne_twister_generator.seed(mtseed);
  PeriodicBoundaryConditions pbc(&th, kL, kL);
  Sys s(th, pbc, p);
  Stepper stepper(&th);  // the SRK stepper draws its natural frequencies here
  std::vector<Real> st = MakeState(seed, 0);
  std::vector<Real> noise, first;
  double t = 0.0;
  for (int k = 0; k < steps; ++k) {
    if (normals_per_particle && g_size == 1) {
      std::vector<Real> xi = PeekNormals(size_t(normals_per_particle) * kN);
      noise.insert(noise.end(), xi.begin(), xi.end());
    }
    t += dt;
    stepper.DoStep(s.sys, st.data(), t, dt);
    pbc.ApplyPeriodicBoundaryConditions(st.data(), kS * kN);
    if (k == 0) first = Gather(st, kS);
  }
  Write("noise", noise);
  Write("state1", first);
  Write("state", Gather(st, kS));
  return 0;
}

// Ensemble of the reference's vortex SRK stepper driven by its own mt19937 +
// std::normal_distribution noise (generator seeded 2000 + member): |R|(t) of
// the Kuramoto order parameter in FP64 every `every` steps, from positions as
// MakeState(seed + member, 0) with every phase 1 (R = 1 at t = 0).  One rank.
// <seed> <members> <steps> <every> <dt> <vortex args>
int EnsembleVortexSrk(ThreadSharedMemory &th, char **a, VortexArgs p) {
  uint64_t seed = std::strtoull(a[0], nullptr, 10);
  int members = std::atoi(a[1]), steps = std::atoi(a[2]), every = std::atoi(a[3]);
  double dt = std::atof(a[4]);
  PeriodicBoundaryConditions pbc(&th, kL, kL);
  std::vector<Real> out;
  for (int r = 0; r < members; ++r) {
    mersenne_twister_generator.seed(uint32_t(2000 + r));
    VortexSys s(th, pbc, p);
    StochasticRungeKuttaStepperPtr stepper(&th);
    std::vector<Real> st = MakeState(seed + uint64_t(r), 0);
    for (int i = 0; i < kN; ++i) st[kS * i + 2] = 1.0;
    double t = 0.0;
    for (int k = 0; k <= steps; ++k) {
      if (k % every == 0) {
        double cr = 0.0, ci = 0.0;
        for (int i = 0; i < kN; ++i) {
          cr += std::cos(st[kS * i + 2]);
          ci += std::sin(st[kS * i + 2]);
        }
        out.push_back(std::sqrt(cr * cr + ci * ci) / kN);
      }
      if (k == steps) break;
      t += dt;
      stepper.DoStep(s.sys, st.data(), t, dt);
      pbc.ApplyPeriodicBoundaryConditions(st.data(), kS * kN);
    }
  }
  Write("R", out);
  return 0;
}

// Seconds per step of `steps` RK4 steps of the vortex system after `warm`
// untimed ones (rank 0 prints one JSON line).  <seed> <warm> <steps> <dt> <vortex args>
int TimeRk4(ThreadSharedMemory &th, char **a, VortexArgs p) {
  uint64_t seed = std::strtoull(a[0], nullptr, 10);
  int warm = std::atoi(a[1]), steps = std::atoi(a[2]);
  double dt = std::atof(a[3]);
  PeriodicBoundaryConditions pbc(&th, kL, kL);
  VortexSys s(th, pbc, p);
  RungeKutta4StepperPtr stepper(&th);
  std::vector<Real> st = MakeState(seed, 0);
  double t = 0.0;
  for (int k = 0; k < warm + steps; ++k) {
    if (k == warm) {
      MPI_Barrier(MPI_COMM_WORLD);
      t = MPI_Wtime();
    }
    stepper.DoStep(s.sys, st.data(), t, dt);
    pbc.ApplyPeriodicBoundaryConditions(st.data(), kS * kN);
  }
  MPI_Barrier(MPI_COMM_WORLD);
  double sec = (MPI_Wtime() - t) / steps;
  std::vector<Real> full = Gather(st, kS);
  if (g_rank == 0) {
    double R = 0.0, I = 0.0;
    for (int i = 0; i < kN; ++i) {
      R += std::cos(full[kS * i + 2]);
      I += std::sin(full[kS * i + 2]);
    }
    std::printf("{\"ranks\": %d, \"n\": %d, \"seconds_per_step\": %.9g, \"R\": %.17g}\n", g_size, kN, sec,
                std::sqrt(R * R + I * I) / kN);
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  LaunchParallelSession(argc, argv);
  int rc = 1;
  {
    ThreadSharedMemory th(argc, argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &g_rank);
    MPI_Comm_size(MPI_COMM_WORLD, &g_size);
    if (argc < 3) {
      if (g_rank == 0) std::fprintf(stderr, "usage: %s <case> <outdir> [args...]\n", argv[0]);
    } else {
      std::string c = argv[1];
      g_outdir = argv[2];
      char **a = argv + 3;
      const int na = argc - 3;
      const bool vort = c[0] == 'v', chim = c[0] == 'c';
      const int sys_args = vort ? 8 : 5;
      if (c == "ens_vsrk") {
        if (na < 13) {
          if (g_rank == 0) std::fprintf(stderr, "case %s: too few arguments\n", c.c_str());
        } else {
          rc = EnsembleVortexSrk(th, a, ParseVortex(a + 5));
        }
      } else if ((vort || chim || c == "time_rk4") && na < 4 + (c == "time_rk4" ? 8 : sys_args)) {
        if (g_rank == 0) std::fprintf(stderr, "case %s: too few arguments\n", c.c_str());
      } else if (c == "vrhs_k") {
        rc = RhsK<VortexSys>(th, a, ParseVortex(a + 4));
      } else if (c == "vrhs_ito") {
        rc = RhsIto<VortexSys>(th, a, ParseVortex(a + 4));
      } else if (c == "crhs_k") {
        rc = RhsK<ChimeraSys>(th, a, ParseChimera(a + 4));
      } else if (c == "crhs_ito") {
        rc = RhsIto<ChimeraSys>(th, a, ParseChimera(a + 4));
      } else if (c == "vse") {
        rc = Steps<StochasticEulerStepperPtr, VortexSys>(th, a, ParseVortex(a + 4), 1);
      } else if (c == "vsrk") {
        rc = Steps<StochasticRungeKuttaStepperPtr, VortexSys>(th, a, ParseVortex(a + 4), 2);
      } else if (c == "vrk4") {
        rc = Steps<RungeKutta4StepperPtr, VortexSys>(th, a, ParseVortex(a + 4), 0);
      } else if (c == "cse") {
        rc = Steps<StochasticEulerStepperPtr, ChimeraSys>(th, a, ParseChimera(a + 4), 1);
      } else if (c == "csrk") {
        rc = Steps<StochasticRungeKuttaStepperPtr, ChimeraSys>(th, a, ParseChimera(a + 4), 2);
      } else if (c == "crk4") {
        rc = Steps<RungeKutta4StepperPtr, ChimeraSys>(th, a, ParseChimera(a + 4), 0);
      } else if (c == "time_rk4") {
        rc = TimeRk4(th, a, ParseVortex(a + 4));
      } else if (g_rank == 0) {
        std::fprintf(stderr, "unknown case %s\n", c.c_str());
      }
    }
  }
  FinalizeParallelSession();
  return rc;
}
