"""Batched parameter-sweep ensemble: 512 independent chimera replicas of uneven
N stepped together (spc2_batch_*), against the same replicas run one after
another through single handles.

Sweep (one replica per grid point, 16 x 4 x 8 = 512):
  sigma 16 points linear in [0.2, 2], alpha 4 points linear in [1.2, 1.57],
  D 8 points log-spaced in [0.001, 0.1];
  N log-uniform in [500, 32768] (numpy default_rng(42)), density 4000
  (L = sqrt(N / 4000)), rho = 0.1, v0 = 0.01, omega ~ N(0, 0.1);
  Stochastic Euler (or, with --method srk, the stochastic Runge-Kutta step),
  dt = 0.01, FP32-compensated pair sums, seed 42 + replica.
Replicas with N < 640 have rho > L/4 and take the all-pairs path, as the
reference dispatches.

The sequential leg times `--sample` replicas spread over the N range alone for
the same number of steps and estimates the 512-replica sequential time by
interpolating time(N) in log-log; every sampled replica's final state must
equal the batch's bit for bit.

    python -m spc2_langevin_integration_amd.ensemble [--steps 5000] [--sample 24] [--method se|srk]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import _abi as spc2


def sweep(count=512, seed=42, precision="fp32k"):
    rng = np.random.default_rng(seed)
    sig = np.linspace(0.2, 2.0, 16)
    alp = np.linspace(1.2, 1.57, 4)
    dph = np.logspace(-3, -1, 8)
    grid = [(s, a, d) for s in sig for a in alp for d in dph]
    ns = np.exp(rng.uniform(math.log(500), math.log(32768), count)).astype(int)
    reps = []
    for r in range(count):
        s, a, d = grid[r % len(grid)]
        n = int(ns[r])
        L = math.sqrt(n / 4000.0)
        reps.append(spc2.params("chimera", precision, n=n, L=L, rho=0.1, v0=0.01, sigma=float(s), alpha=float(a),
                                D=float(d), seed=seed + r))
    return reps


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--replicas", type=int, default=512)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--sample", type=int, default=24)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--trial-steps", type=int, default=3)
    ap.add_argument("--method", default="se", choices=["se", "srk"])
    args = ap.parse_args(argv)
    reps = sweep(args.replicas)
    ns = np.array([p.n for p in reps])
    omega_std = 0.1

    with spc2.Batch(reps) as b:
        b.init_random_state(omega_std)
        tuned = None if args.no_tune else b.autotune(args.dt, args.trial_steps)
        targets = b.tile_targets()
        b.step(args.dt, 1, method=args.method)  # warm-up: the timed leg starts at step 1 for every replica
        t0 = time.perf_counter()
        b.step(args.dt, args.steps, method=args.method)
        batch_sec = time.perf_counter() - t0
        order = np.argsort(ns)
        pick = order[np.linspace(0, len(order) - 1, min(args.sample, len(order))).astype(int)]
        finals = {int(r): b.get_state(int(r)) for r in pick}
    print(f"# batch: {args.replicas} replicas x {args.steps} steps in {batch_sec:.2f} s", file=sys.stderr, flush=True)

    seq = []
    mismatches = 0
    for r in pick:
        p = reps[int(r)]
        with spc2.Sim(p) as s:
            s.init_random_state(omega_std)
            s.step(args.method, args.dt, 0, 1)
            s.synchronize()
            t0 = time.perf_counter()
            s.step(args.method, args.dt, 1, args.steps)
            s.synchronize()
            seq.append((int(p.n), time.perf_counter() - t0))
            mismatches += int(not np.array_equal(s.get_state(), finals[int(r)]))
        print(f"# alone: N={p.n} {seq[-1][1]:.3f} s", file=sys.stderr, flush=True)
    sn = np.log([q[0] for q in seq])
    st = np.log([q[1] for q in seq])
    est = np.exp(np.interp(np.log(ns), sn, st))
    seq_total = float(est.sum())
    done_times = np.cumsum(est[np.argsort(-ns)])  # sequential order: any; completion times for latency
    line = {
        "metric": "ensemble solves/s, batched vs sequential launches",
        "replicas": args.replicas, "steps": args.steps, "dt": args.dt, "method": args.method, "dtype": "fp32k",
        "particles_total": int(ns.sum()), "all_pairs_replicas": int((ns < 640).sum()),
        "batch_seconds": batch_sec, "batch_solves_per_s": args.replicas / batch_sec,
        "batch_particle_steps_per_s": float(ns.sum()) * args.steps / batch_sec,
        "batch_p99_latency_s": batch_sec,
        "sequential_seconds_estimated": seq_total, "sequential_solves_per_s": args.replicas / seq_total,
        "sequential_p99_latency_s": float(np.percentile(done_times, 99)),
        "speedup": seq_total / batch_sec,
        "sequential_sample": {"replicas": len(seq), "n": [q[0] for q in seq], "seconds": [q[1] for q in seq],
                              "bit_identical_to_batch": len(seq) - mismatches},
        "autotune": {"tuned_ms_per_step": tuned,
                     "tile_targets_by_log2n": {int(k): float(np.median(targets[np.floor(np.log2(ns)) == k]))
                                               for k in np.unique(np.floor(np.log2(ns)))}},
    }
    print(json.dumps(line), flush=True)
    return 0 if mismatches == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
