// Command-line runner of the drop-in engine: the reference's main()
// (main.cpp:11-35) with the compile-time N, L and the engine's hard-coded
// parameters replaced by key=value arguments or a file of them.
//
//   spc2_run system=chimera method=srk n=262144 L=8.0954 rho=0.1 steps=1000
//            state=traj.bin summary=summary.txt   (one command line)
//   spc2_run config=run.cfg [key=value ...]
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "spc2_host.hpp"

namespace {

void ApplyArg(EngineConfig &cfg, const std::string &arg);

void ApplyFile(EngineConfig &cfg, const std::string &path) {
  std::ifstream in(path);
  if (!in.is_open()) throw Spc2Error("cannot open " + path);
  std::string line;
  while (std::getline(in, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream words(line);
    std::string w;
    while (words >> w) ApplyArg(cfg, w);
  }
}

void ApplyArg(EngineConfig &cfg, const std::string &arg) {
  const size_t eq = arg.find('=');
  if (eq == std::string::npos || eq == 0) throw Spc2Error("expected key=value, got '" + arg + "'");
  const std::string key = arg.substr(0, eq), value = arg.substr(eq + 1);
  if (key == "config")
    ApplyFile(cfg, value);
  else
    cfg.Set(key, value);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc > 1 && (std::string(argv[1]) == "-h" || std::string(argv[1]) == "--help")) {
    std::cout << "usage: spc2_run [config=FILE] [key=value ...]\n"
                 "keys: system method precision init noise loop n L v0 sigma rho alpha D rho0 omega_std\n"
                 "      mu_plus mu_minus xi_a xi_r kappa dt t0 t1 steps seed state summary\n"
                 "      state_every summary_every float_summary init_file init_frame init_state_file\n"
                 "      neighbour verlet_skin resume checkpoint checkpoint_every\n";
    return 0;
  }
  try {
    EngineConfig cfg;
    for (int a = 1; a < argc; ++a) ApplyArg(cfg, argv[a]);
    SimulationEngine engine(cfg);
    if (cfg.system == "chimera")
      engine.RunChimeraSimulation();
    else
      engine.RunVortexArraySimulation();
    double R = 0.0, psi = 0.0;
    Spc2Check(spc2_order_parameter(engine.system().handle(), &R, &psi), "spc2_order_parameter");
    std::cout.precision(17);
    std::cout << "{\"system\": \"" << cfg.system << "\", \"method\": \"" << cfg.method << "\", \"precision\": \""
              << cfg.precision << "\", \"n\": " << cfg.n << ", \"steps\": " << engine.steps_done()
              << ", \"seconds\": " << engine.seconds_in_loop()
              << ", \"R\": " << R << ", \"psi\": " << psi << "}" << std::endl;
  } catch (const std::exception &e) {
    std::cerr << "spc2_run: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
