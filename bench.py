"""Benchmark: particle-steps/s of the vortex-arrays RK4 workload at N = 1,048,576
on one MI355X, FP32-compensated form (the headline value) next to the plain FP64
form, with the pair kernel's roofline position and the CPU oracle baseline.

Workload: vortex arrays, deterministic RK4 (D = 0), dt = 0.01, N = 1,048,576 in a
periodic box at density 4000 (L = sqrt(N / 4000) = 16.19086), rho = 0.05,
xi_a = 0.01, xi_r = 0.005, v = 0.025, mu+ = mu- = 100, kappa = 20; uniform
random positions and phases (Philox, seed 42).  Neighbours from a Verlet list
with skin 0.2 rho (the workload's stated skin; --skin), rebuilt before any
particle can have moved skin / 2 (--neighbour cells: the reference's linked
cells rebuilt at every RHS).  Inputs are generated on the device and stay
resident; a step is one full RK4 step (4 RHS evaluations + combine + wrap, plus
the list rebuilds that fall in it).  The FP32 form is also timed at the tuned
skin 0.125 rho (value_tuned_skin).

Besides the timings the line carries the FP32-vs-FP64 errors at the workload's
500-step horizon, framed by the chaos of the system: an ensemble of FP64 runs
from initial phases perturbed by 1e-7 gives the spread any rounding-size
difference grows to; the order parameter is also compared time-averaged, with
95% confidence intervals from batch means; and phase errors are split between
particles whose neighbour counts differed between the FP32 and FP64 runs at
some step (a cutoff crossing, not a rounding error) and the others.

Order: error analysis, pair-kernel profiling and the FP64 timings run first,
the headline FP32 timing last (the GPU busy right up to its timed region).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--neighbour verlet|cells]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp32k": 157.3, "fp64": 78.6, "fp64f": 78.6}  # TFLOP/s, MI355X vector FP32 / FP64
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the FP32 Verlet pair kernel (k_pass1_verlet_lds
# <VortAcc<P32K>>) on this workload (skin 0.2 rho), from the FETCH_SIZE (x2.000)
# and WRITE_SIZE (x1.000) counter passes calibrated against a 512 MiB device
# copy, stages 1-3 and 4 weighted 3:1; profiles/pair_kernel_vortex_rk4_n1m_fp32k_verlet.md.
TRAFFIC_VERLET_FP32K = 240.6e6

# algorithmic flop convention per RHS evaluation (T = sin/cos/sqrt/atan2 counted as 1)
FLOP_CANDIDATE = 14      # 2 sub, min-image 2x4, r^2 3, compare 1
FLOP_VORTEX_PAIR = 16 + 2  # sqrt, W(r), sin(phi_j - phi_i), compensated add
FLOP_COHESION = 12 + 2
FLOP_EPILOGUE = 4 + 2


def workload(n=1048576, name="vortex-rk4"):
    L = math.sqrt(n / 4000.0)
    if name == "chimera-srk":
        # chimera, SRK, dt = 0.005, rho = 0.1, v0 = 0.01, sigma = 1, alpha = 1.45,
        # D = 0.01, omega ~ N(0, 0.1) (N = 262,144: ~125 neighbours per particle)
        return dict(system="chimera", method="srk", n=n, L=L, rho=0.1, v0=0.01, sigma=1.0, alpha=1.45, D=0.01,
                    dt=0.005, omega_std=0.1, seed=42)
    return dict(system="vortex", method="rk4", n=n, L=L, rho=0.05, velocity=0.025, mu_plus=100.0, mu_minus=100.0,
                xi_a=0.01, xi_r=0.005, kappa=20.0, D=0.0, dt=0.01, omega_std=0.0, seed=42)


def make_sim(spc2, w, precision, neighbour="verlet"):
    skin = w.get("skin", 0.2) * w["rho"]
    if w["system"] == "chimera":
        p = spc2.params("chimera", precision, neighbour, n=w["n"], L=w["L"], rho=w["rho"], v0=w["v0"],
                        sigma=w["sigma"], alpha=w["alpha"], D=w["D"], seed=w["seed"], verlet_skin=skin)
    else:
        p = spc2.params("vortex", precision, neighbour, n=w["n"], L=w["L"], rho=w["rho"], velocity=w["velocity"],
                        mu_plus=w["mu_plus"], mu_minus=w["mu_minus"], xi_a=w["xi_a"], xi_r=w["xi_r"],
                        kappa=w["kappa"], D=w["D"], seed=w["seed"], verlet_skin=skin)
    s = spc2.Sim(p)
    s.init_random_state(w["omega_std"])
    return s


def time_precision(spc2, w, precision, steps, warmup, neighbour):
    s = make_sim(spc2, w, precision, neighbour)
    s.step(w["method"], w["dt"], 0, warmup)
    s.synchronize()
    b0 = s.verlet_stats()[0] if neighbour == "verlet" else 0
    t0 = time.perf_counter()
    s.step(w["method"], w["dt"], warmup, steps)
    s.synchronize()
    t1 = time.perf_counter()
    builds = (s.verlet_stats()[0] - b0) if neighbour == "verlet" else None
    R, _ = s.order_parameter()
    phases = s.get_state()[2::3]
    s.close()
    sec = t1 - t0
    return {"particle_steps_per_s": w["n"] * steps / sec, "ms_per_step": 1e3 * sec / steps, "R_final": R,
            "phases": phases, "list_builds": builds}


def fixed_bytes_per_particle(w, precision, neighbour):
    """Algorithmic HBM bytes per particle and pair-kernel launch, beside the
    list: what each launch must read and write once per particle (neighbour
    tiles are counted once, as the particle's own tile; their reuse comes
    from LDS and cache).

    Vortex RK4, every stage: own tile (FP32: x, y, cos, sin 16 B + FP64 stage
    phase 8 B; FP64: 24 B), list count 4 B, k written (linked cells: 24 B in
    stages 1-3, k_4's phi 8 B in the last; Verlet steps write only k's phi,
    8 B, since the combine rebuilds k_1..k_3's x / y from the stage phases).
    The last stage (a quarter of the launches) also reads and writes the state
    (48 B) and reads k1..k3 (72 B; Verlet: their phi components, 24 B).
    Verlet lists fuse the next stage's tiles into the epilogue: written (FP32:
    16 B tile + 16 B FP64 position + 8 B phase; FP64: 24 B + 8 B), with the
    build cell (4 B) and, in stages 1-3, the state (24 B) read.

    Chimera SRK, the mean of its two passes: pass 1 reads the own tile 24 B,
    count and id 8 B and writes the pass-2 inputs 28 B (60 B; the derivative
    itself is not written, nothing in the step reads it); pass 2 reads the
    tile 24 B, count and id 8 B, its pass-1 inputs 24 B and the staged pass-1
    derivative 4 B, omega 8 B, the state 24 B, and writes the state 24 B and
    (Verlet) the next step's tiles 40 B with the build cell 4 B read (160 B)."""
    fp32 = precision == "fp32k"
    if w["system"] == "chimera":
        return (60 + 160) / 2.0
    own = 24
    if neighbour == "verlet":
        return own + 4 + 8 + (24 + 48) / 4.0 + (40 if fp32 else 32) + 4 + 24 * 3 / 4.0
    return own + 4 + (24 * 3 + 8) / 4.0 + (72 + 48) / 4.0


def profile_pair_kernel(spc2, w, precision, warmup, steps, neighbour):
    """HIP-event timing of the pair kernels on the handle's stream over the
    timed region's steps (warmup .. warmup + steps of a fresh handle), and the
    pair counts of the same steps on another fresh handle."""
    s = make_sim(spc2, w, precision, neighbour)
    s.step(w["method"], w["dt"], 0, warmup)
    s.set_profiling(True)            # timing pass: events bound to the pair-kernel dispatches
    s.step(w["method"], w["dt"], warmup, steps)
    ms, launches = s.profile()
    s.close()
    s = make_sim(spc2, w, precision, neighbour)
    s.step(w["method"], w["dt"], 0, warmup)
    s.set_counting(True)             # counting pass over the same launches
    s.step(w["method"], w["dt"], warmup, steps)
    cnt = s.counts()
    lds = neighbour == "verlet" and s.verlet_layout()[0] == 1
    s.close()
    per = cnt / launches
    if w["system"] == "chimera":
        # pass 1 and pass 2 alternate and only pass 1 counts, so the counts per
        # RHS are twice those per launch. Both passes walk every candidate;
        # an interaction costs 7 + 1T in pass 1 and 8 + 1T in pass 2; the
        # figures are averaged over the two launches.
        per = per * 2
        flops = (FLOP_CANDIDATE * per[0] * 2 + (8 + 9) * per[1] + 2 * FLOP_EPILOGUE * w["n"]) / 2
    else:
        flops = (FLOP_CANDIDATE * per[0] + FLOP_VORTEX_PAIR * per[1] + FLOP_COHESION * per[2] +
                 FLOP_EPILOGUE * w["n"])
    avg_s = ms / launches / 1e3
    # list: 2 B per entry in the LDS layout (16-bit halo indices), 4 B in the
    # gather layout, once per launch (both SRK passes read it)
    entry_b = (2.0 if lds else 4.0) if neighbour == "verlet" else 0.0
    bytes_per = w["n"] * fixed_bytes_per_particle(w, precision, neighbour) + entry_b * per[0]
    return {"avg_ms": ms / launches, "launches": launches, "flops_per_launch": flops,
            "tflops": flops / avg_s / 1e12, "bytes_per_launch": bytes_per, "gbs": bytes_per / avg_s / 1e9,
            "candidates_per_particle": per[0] / w["n"], "neighbours_per_particle": per[1] / w["n"],
            "cohesion_per_particle": per[2] / w["n"], "lds_layout": lds}


def phase_diff(a, b):
    """|a - b| as angles, in [0, pi]."""
    import numpy as np
    return np.abs(np.mod(a - b + np.pi, 2 * np.pi) - np.pi)


def order_param(phases):
    import numpy as np
    return float(np.abs(np.exp(1j * phases).mean()))


def batch_means_ci(series, batches=5):
    """Mean of a correlated time series with a 95% confidence half-width from
    batch means (Student t, batches - 1 degrees of freedom)."""
    import numpy as np
    x = np.asarray(series, float)
    k = len(x) // batches
    m = x[:k * batches].reshape(batches, k).mean(1)
    t975 = {2: 12.706, 3: 4.303, 4: 3.182, 5: 2.776, 6: 2.571, 8: 2.365, 10: 2.262}.get(batches, 2.0)
    return float(x.mean()), float(t975 * m.std(ddof=1) / math.sqrt(batches))


def error_analysis(spc2, w, nb, horizon=500, members=8, flip_steps=25, every=10):
    """FP32-compensated against FP64 (the reference method) from the same
    initial state over `horizon` steps:
      * the order parameter R every `every` steps, its difference at the
        horizon against the spread of `members` FP64 runs whose initial phases
        are perturbed by U(-1e-7, 1e-7) (the growth of rounding-size
        differences in this chaotic system), and the time-averaged R of both
        runs with 95% confidence intervals;
      * phase errors after `flip_steps` and 100 steps, split between
        particles whose partner counts (within rho, within xi_r) differed
        between the two runs at one of the first `flip_steps` steps and the
        others;
      * FP64 against FP64 with the other neighbour structure (another
        summation order) as the floor at the same horizons."""
    import numpy as np
    res = {}
    series, phases = {}, {}
    checkpoints = (flip_steps, 100, horizon)
    other = "cells" if nb == "verlet" else "verlet"
    counts = {}
    for name, prec, neigh in (("fp64", "fp64", nb), ("fp32", "fp32k", nb), ("fp64_reordered", "fp64", other)):
        s = make_sim(spc2, w, prec, neigh)
        rs = [order_param(s.get_state()[2::3])]
        track = name in ("fp64", "fp32")
        cnt = []
        for k in range(horizon):
            s.step(w["method"], w["dt"], k, 1)
            if track and k < flip_steps:
                cnt.append(s.neighbour_counts())
            if (k + 1) % every == 0 or (k + 1) in checkpoints:
                ph = s.get_state()[2::3]
                if (k + 1) % every == 0:
                    rs.append(order_param(ph))
                if (k + 1) in checkpoints:
                    phases[(name, k + 1)] = ph
        counts[name] = cnt
        series[name] = rs
        s.close()
    flip = np.zeros(w["n"], bool)
    for (a32, b32), (a64, b64) in zip(counts["fp32"], counts["fp64"]):
        flip |= (a32 != a64) | (b32 != b64)
    for h in (flip_steps, 100):
        ref = phases[("fp64", h)]
        d = phase_diff(phases[("fp32", h)], ref)
        dr = phase_diff(phases[("fp64_reordered", h)], ref)
        entry = {"order_param_abs_err": abs(order_param(phases[("fp32", h)]) - order_param(ref)),
                 "phase_err_max": float(d.max()), "phase_err_rms": float(np.sqrt((d ** 2).mean())),
                 "fp64_reordered_phase_err_rms": float(np.sqrt((dr ** 2).mean())),
                 "fp64_reordered_order_param_abs_err": abs(order_param(phases[("fp64_reordered", h)]) -
                                                           order_param(ref))}
        for tag, m in (("with_count_flip", flip), ("without_count_flip", ~flip)):
            entry[tag] = {"particles": int(m.sum()),
                          "phase_err_max": float(d[m].max()) if m.any() else 0.0,
                          "phase_err_rms": float(np.sqrt((d[m] ** 2).mean())) if m.any() else 0.0}
        res[f"after_{h}_steps"] = entry
    # FP64 members from perturbed initial phases
    st0 = None
    R_members = []
    for m in range(members):
        s = make_sim(spc2, w, "fp64", nb)
        if st0 is None:
            st0 = s.get_state()
        st = st0.copy()
        st[2::3] += np.random.default_rng(1000 + m).uniform(-1e-7, 1e-7, w["n"])
        s.set_state(st)
        s.step(w["method"], w["dt"], 0, horizon)
        R_members.append(order_param(s.get_state()[2::3]))
        s.close()
    R64, R32 = series["fp64"][-1], series["fp32"][-1]
    spread = np.array(R_members) - R64
    m64, ci64 = batch_means_ci(series["fp64"][1:])
    m32, ci32 = batch_means_ci(series["fp32"][1:])
    res[f"after_{horizon}_steps"] = {
        "order_param_fp64": R64, "order_param_fp32": R32, "order_param_abs_err": abs(R32 - R64),
        "fp64_perturbed_members": members, "fp64_perturbation": "initial phases + U(-1e-7, 1e-7)",
        "fp64_perturbed_order_param_spread_rms": float(np.sqrt((spread ** 2).mean())),
        "fp64_perturbed_order_param_spread_max": float(np.abs(spread).max()),
        "fp32_err_within_fp64_spread": bool(abs(R32 - R64) <= max(np.abs(spread).max(), 1e-300)),
        "fp64_reordered_order_param_abs_err": abs(series["fp64_reordered"][-1] - R64),
        "time_avg_order_param_fp64": m64, "time_avg_order_param_fp64_ci95": ci64,
        "time_avg_order_param_fp32": m32, "time_avg_order_param_fp32_ci95": ci32,
        "time_avg_order_param_abs_err": abs(m32 - m64),
        "time_avg_within_ci95": bool(abs(m32 - m64) <= math.hypot(ci64, ci32)),
        "order_param_every": every,
    }
    return res


COPY_BYTES = 1 << 29  # calibration copy: reads and writes this many bytes


def pmc_run(spc2, w, nb, precision="fp32k"):
    """Workload steps and a calibration copy for the FETCH_SIZE / WRITE_SIZE
    counters: the copy reads and writes COPY_BYTES, so the counters' ratio to
    it converts their units to bytes."""
    import torch
    s = make_sim(spc2, w, precision, nb)
    s.step(w["method"], w["dt"], 0, 3)
    s.synchronize()
    s.close()
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"pmc_run": True, "copy_bytes": COPY_BYTES, "n": w["n"], "neighbour": nb,
                      "precision": precision}), flush=True)


def cpu_baseline(w, target_s=12.0, max_steps=40):
    """The oracle on the host cores: one step to time it, then as many more as
    fit about target_s seconds (the rate is over all steps taken)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import numpy as np
    threads = max(1, min(16, os.cpu_count() or 1))
    n = w["n"]
    st, om = O.init_state(w["seed"], n, w["L"], w["L"], w["omega_std"])
    if w["system"] == "chimera":
        sysp = O.chimera(n, w["L"], w["rho"], w["v0"], w["sigma"], w["alpha"], w["D"], threads=threads)

        def one(k, st):
            xi = O.philox_normals(w["seed"], k, 0, n, 2)
            t = time.perf_counter()
            st = O.step_srk(sysp, st, w["dt"], xi, omega=om)
            return st, time.perf_counter() - t
    else:
        sysp = O.vortex(n, w["L"], rho=w["rho"], velocity=w["velocity"], mu_plus=w["mu_plus"],
                        mu_minus=w["mu_minus"], xi_a=w["xi_a"], xi_r=w["xi_r"], kappa=w["kappa"], D=w["D"],
                        threads=threads)
        kbuf = np.zeros(12 * n)

        def one(k, st):
            t = time.perf_counter()
            st = O.step_rk4(sysp, st, w["dt"], kbuf)
            return st, time.perf_counter() - t
    st, sec = one(0, st)
    steps = max(1, min(max_steps, int(target_s / max(sec, 1e-9))))
    for k in range(1, steps):
        st, dt_s = one(k, st)
        sec += dt_s
    return {"value": n * steps / sec, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} {w['method'].upper()} steps of the same N={n} {w['system']} workload, FP64 "
                      f"reference method (oracle/spc2_oracle.c, -O3, bit-identical to the reference's own build), "
                      f"{threads} threads over contiguous particle slices; the port runs at 0.8-1.2x the speed of "
                      f"the reference's MPI build (-O3, mpiexec -n 1..8) on the same cores "
                      f"(profiles/cpu_reference_calibration.md)",
            "seconds": sec}


# Algorithmic bytes of the vortex RK4 + Verlet workload: 1,108 B per
# particle-step (4 x [24 + 12 + 12 + 45.2 x 4 list + 12 + 12] + 96 combine) over
# the step's 4 pair-kernel launches.
ALG_BYTES_PER_PARTICLE_STEP = 1108.0
# VALU issue of the same kernel from the SQ counter pass (SQ_INSTS_VALU x 4
# cycles over the 1024 SIMDs during the launch, stages weighted 3:1);
# profiles/pair_kernel_vortex_rk4_n1m_fp32k_verlet.md.
VALU_ISSUE_FRAC_VERLET_FP32K = 0.74


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--precisions", default="fp32k,fp64,fp64f",
                    help="forms to time; the first is the headline value")
    ap.add_argument("--neighbour", default="verlet", choices=["verlet", "cells"],
                    help="verlet (default) or cells: the reference's linked cells rebuilt at every RHS")
    ap.add_argument("--workload", default="vortex-rk4", choices=["vortex-rk4", "chimera-srk"],
                    help="vortex-rk4: the headline (N=1M); chimera-srk: chimera SRK at the given --n")
    ap.add_argument("--pmc-run", action="store_true",
                    help="for rocprofv3 --pmc: a few steps of the FP32 workload and a plain device copy of "
                         "COPY_BYTES bytes (the counter calibration), no timing")
    ap.add_argument("--pmc-precision", default="fp32k", choices=["fp32k", "fp64", "fp64f"],
                    help="the form --pmc-run steps")
    ap.add_argument("--skin", type=float, default=0.2,
                    help="Verlet skin as a fraction of rho (0.2: the workload's stated skin)")
    ap.add_argument("--tuned-skin", type=float, default=0.125,
                    help="skin of the extra FP32 timing (0.125: the fastest of 0.1-0.3 on this workload; 0: none)")
    ap.add_argument("--error-horizon", type=int, default=500,
                    help="steps of the FP32-vs-FP64 error analysis (the workload's 500; 0: none)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (a large system would shard by "
              "spatial slabs of cell rows, with one halo row exchanged per pass)", file=sys.stderr)
        sys.exit(2)

    import spc2_langevin_integration_amd as spc2
    w = workload(args.n, args.workload)
    w["skin"] = args.skin
    if args.pmc_run:
        pmc_run(spc2, w, args.neighbour, args.pmc_precision)
        return
    precs = args.precisions.split(",")
    head = precs[0]
    nb = args.neighbour
    verlet = nb == "verlet"
    headline = verlet and args.n == 1048576 and w["system"] == "vortex"

    # untimed analysis first
    err = None
    if "fp64" in precs and args.error_horizon > 0:
        err = error_analysis(spc2, w, nb, args.error_horizon)
        for k, v in err.items():
            print(f"# errors {k}: {json.dumps(v)}", file=sys.stderr, flush=True)
    prof = profile_pair_kernel(spc2, w, "fp32k", args.warmup, args.steps, nb)
    prof64 = profile_pair_kernel(spc2, w, "fp64", args.warmup, min(args.steps, 50), nb) if "fp64" in precs else None
    print(f"# pair kernel fp32k: {prof['avg_ms']:.4f} ms/launch, {prof['tflops']:.2f} TFLOP/s, "
          f"{prof['gbs']:.0f} GB/s" + (f"; fp64: {prof64['avg_ms']:.4f} ms/launch" if prof64 else ""),
          file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline(w)

    # timings: the other forms and the tuned skin, then the headline form last
    res = {}
    tuned = None
    for prec in precs[1:]:
        res[prec] = time_precision(spc2, w, prec, args.steps, args.warmup, nb)
        print(f"# {prec} ({nb}): {res[prec]['particle_steps_per_s']:.4e} particle-steps/s, "
              f"{res[prec]['ms_per_step']:.3f} ms/step, list builds {res[prec]['list_builds']}",
              file=sys.stderr, flush=True)
    if verlet and args.tuned_skin > 0 and args.tuned_skin != args.skin:
        wt = dict(w, skin=args.tuned_skin)
        tuned = time_precision(spc2, wt, head, args.steps, args.warmup, nb)
        print(f"# {head} at skin {args.tuned_skin} rho: {tuned['particle_steps_per_s']:.4e} particle-steps/s, "
              f"{tuned['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    res[head] = time_precision(spc2, w, head, args.steps, args.warmup, nb)
    v = res[head]
    print(f"# {head} ({nb}, skin {args.skin} rho): {v['particle_steps_per_s']:.4e} particle-steps/s, "
          f"{v['ms_per_step']:.3f} ms/step, list builds {v['list_builds']}", file=sys.stderr, flush=True)

    acc = "VortAcc" if w["system"] == "vortex" else "ChimAcc, Chim2Acc"
    kname = (f"k_pass1_verlet_lds<{acc}<P32K>>" if verlet else f"k_pass1_lds<{acc}<P32K>>") + \
        f" ({w['system']} pair sums, FP32 Kahan)"
    avg_s = prof["avg_ms"] / 1e3
    if headline:
        # 1,108 B per particle-step over 4 launches
        alg_bytes = ALG_BYTES_PER_PARTICLE_STEP / 4.0 * w["n"]
        roof = {"bound": "hbm", "kernel": kname, "achieved": alg_bytes / avg_s / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": alg_bytes / avg_s / 1e9 / HBM_PEAK_GBS,
                "traffic": TRAFFIC_VERLET_FP32K, "traffic_GBs": TRAFFIC_VERLET_FP32K / avg_s / 1e9,
                "algorithmic_bytes_per_launch": alg_bytes,
                "algorithmic_bytes_convention": "1,108 B per particle-step (4 launches), 45.2 list entries x 4 B",
                "limiter": "VALU issue", "valu_issue_frac": VALU_ISSUE_FRAC_VERLET_FP32K,
                "measured_list_bytes_model_per_launch": prof["bytes_per_launch"],
                "measured_list_bytes_model_GBs": prof["gbs"]}
    elif verlet:
        roof = {"bound": "hbm", "kernel": kname, "achieved": prof["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": prof["gbs"] / HBM_PEAK_GBS, "traffic": None,
                "algorithmic_bytes_per_launch": prof["bytes_per_launch"]}
    else:
        roof = {"bound": "valu", "kernel": kname, "achieved": prof["tflops"], "peak": PEAK["fp32k"],
                "unit": "TFLOP/s", "frac": prof["tflops"] / PEAK["fp32k"], "traffic": None}
    roof.update({"avg_launch_ms": prof["avg_ms"], "launches_timed": prof["launches"],
                 "flops_per_launch": prof["flops_per_launch"], "valu_tflops": prof["tflops"],
                 "valu_flop_frac": prof["tflops"] / PEAK["fp32k"],
                 "candidates_per_particle": prof["candidates_per_particle"],
                 "neighbours_per_particle": prof["neighbours_per_particle"],
                 "fp64_pair_kernel_avg_ms": prof64["avg_ms"] if prof64 else None})
    line = {
        "metric": ("particle-steps/s at N=1M (FP32-comp vs FP64, same GPU)" if headline else
                   f"particle-steps/s, {args.workload} at N={w['n']} (FP32-comp vs FP64, same GPU)"),
        "value": v["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": v["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair sums (Kahan) + f64 state" if head == "fp32k" else head,
        "data": "synthetic (Philox uniform positions and phases, seed 42)",
        "config": {"workload": (f"vortex-arrays RK4, N={w['n']}, density 4000 (L={w['L']:.5f}), rho=0.05, "
                                "xi_a=0.01, xi_r=0.005, v=0.025, mu+-=100, kappa=20, D=0, dt=0.01, "
                                if w["system"] == "vortex" else
                                f"chimera SRK, N={w['n']}, density 4000 (L={w['L']:.5f}), rho=0.1, v0=0.01, "
                                "sigma=1, alpha=1.45, D=0.01, omega~N(0,0.1), dt=0.005, ")
                               + (f"Verlet list, skin {args.skin:g} rho, rebuilt at skin/2 displacement" if verlet
                                  else "linked cells per RHS"),
                   "n": w["n"], "method": w["method"], "neighbour": nb, "skin_over_rho": args.skin if verlet else None,
                   "list_builds_in_timed_steps": v["list_builds"]},
        "value_tuned_skin": tuned["particle_steps_per_s"] if tuned else None,
        "ms_per_step_tuned_skin": tuned["ms_per_step"] if tuned else None,
        "tuned_skin_over_rho": args.tuned_skin if tuned else None,
        "fp64_value": res["fp64"]["particle_steps_per_s"] if "fp64" in res else None,
        "fp64_factored_value": res["fp64f"]["particle_steps_per_s"] if "fp64f" in res else None,
        "fp32_over_fp64": (v["particle_steps_per_s"] / res["fp64"]["particle_steps_per_s"]) if "fp64" in res else None,
        "errors_fp32_vs_fp64": err,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu_baseline"] = v["particle_steps_per_s"] / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
