"""Calibrates the CPU baseline against the reference itself.

Times the vortex-arrays RK4 step of the reference's MPI shared-memory build
(oracle/_ref/ref_harness_mpi_o3: VortexArraysSystemPtr + RungeKutta4StepperPtr
compiled -O3 where they lie) under `mpiexec -n P`, P = 1, 2, 4, 8, and the
oracle's restatement (liboracle.so, the CPU baseline of bench.py) on the same
workload with the same thread counts: N = 8192, L = 64 (the reference's
compile-time size) at rho = 2.236, xi_a = 0.4472, xi_r = 0.2236, about 31
partners per particle as in the N = 1,048,576 workload.  Median of 3 runs of
`--steps` steps after 2 warm-up steps each.  Prints a markdown table and a
JSON line with the oracle / reference time ratios.

Test infrastructure: runs only where the reference build exists.

    make -C oracle ref && python oracle/calibrate_cpu.py [--steps 10]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import numpy as np  # noqa: E402

import oracle as O  # noqa: E402

N, L = 8192, 64.0
VORTEX = dict(velocity=1.0, mu_plus=100.0, mu_minus=100.0, xi_a=0.4472, xi_r=0.2236, D=0.0, kappa=20.0, rho=2.236)
SEED, DT = 38, 0.01
MPIEXEC = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")
REF = os.path.join(HERE, "_ref", "ref_harness_mpi_o3")


def time_reference(ranks, steps):
    v = VORTEX
    args = [str(SEED), "2", str(steps), repr(DT)] + [repr(v[k]) for k in
                                                      ("velocity", "mu_plus", "mu_minus", "xi_a", "xi_r", "D",
                                                       "kappa", "rho")]
    out = subprocess.run([MPIEXEC, "-n", str(ranks), REF, "time_rk4", "/tmp"] + args, check=True,
                         capture_output=True, text=True).stdout
    return json.loads(out.strip().splitlines()[-1])["seconds_per_step"]


def time_oracle(threads, steps):
    v = VORTEX
    sysp = O.vortex(N, L, rho=v["rho"], velocity=v["velocity"], mu_plus=v["mu_plus"], mu_minus=v["mu_minus"],
                    xi_a=v["xi_a"], xi_r=v["xi_r"], kappa=v["kappa"], D=v["D"], threads=threads)
    st = O.make_state(SEED, 0, N, L, v["rho"])
    kbuf = np.zeros(12 * N)
    for _ in range(2):
        st = O.step_rk4(sysp, st, DT, kbuf)
    t0 = time.perf_counter()
    for _ in range(steps):
        st = O.step_rk4(sysp, st, DT, kbuf)
    return (time.perf_counter() - t0) / steps


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--ranks", default="1,2,4,8")
    args = ap.parse_args(argv)
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    rows = []
    for p in [int(x) for x in args.ranks.split(",")]:
        ref = statistics.median(time_reference(p, args.steps) for _ in range(3))
        orc = statistics.median(time_oracle(p, args.steps) for _ in range(3))
        rows.append({"ranks_or_threads": p, "reference_s_per_step": ref, "oracle_s_per_step": orc,
                     "reference_particle_steps_per_s": N / ref, "oracle_particle_steps_per_s": N / orc,
                     "oracle_over_reference_speed": ref / orc})
        print(f"# P={p}: reference {ref * 1e3:.2f} ms/step, oracle {orc * 1e3:.2f} ms/step", file=sys.stderr,
              flush=True)
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next(line.split(":", 1)[1].strip() for line in f if line.startswith("model name"))
    except (OSError, StopIteration):
        pass
    print(f"| P (MPI ranks / threads) | reference ms/step | oracle ms/step | reference particle-steps/s | "
          f"oracle particle-steps/s | oracle / reference speed |")
    print("|---:|---:|---:|---:|---:|---:|")
    for r in rows:
        print(f"| {r['ranks_or_threads']} | {r['reference_s_per_step'] * 1e3:.2f} | {r['oracle_s_per_step'] * 1e3:.2f}"
              f" | {r['reference_particle_steps_per_s']:.3e} | {r['oracle_particle_steps_per_s']:.3e} | "
              f"{r['oracle_over_reference_speed']:.2f} |")
    print(json.dumps({"cpu": cpu, "cores": os.cpu_count(), "n": N, "steps": args.steps, "workload": VORTEX,
                      "rows": rows}))


if __name__ == "__main__":
    main()
