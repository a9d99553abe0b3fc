"""Writes tests/golden/golden.json from the reference's own classes.

Runs oracle/_ref/ref_harness (the reference's non-MPI translation units compiled
where they lie, see oracle/Makefile) and oracle/_ref/ref_harness_mpi (its MPI
shared-memory translation units, under mpiexec with 1, 2 and 4 ranks) on
seeded inputs and records, per output
array, its SHA-256 over the raw little-endian bytes, a few sampled entries and
summary statistics.  Inputs are regenerated from the seed by splitmix64, so the
fixture holds no input data.  The reference fixes N = 8192, L = 64.

    make -C oracle ref && python tests/golden/make_golden.py [--only-vortex-ensemble]
"""
from __future__ import annotations

import hashlib
import json
import os
import platform
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

N, L = 8192, 64.0

# (name, case, args, outputs{name: dtype})
CHIM = lambda v0, sigma, rho, alpha, D: [repr(v0), repr(sigma), repr(rho), repr(alpha), repr(D)]
CASES = [
    ("cells_rho4", "cells", ["1", "0", "4.0"], {"cell_start": "i4", "sorted": "i4", "cell_of": "i4"}),
    ("cells_rho1_outside", "cells", ["2", "1", "1.0"], {"cell_start": "i4", "sorted": "i4", "cell_of": "i4"}),
    ("cells_rho3.7_lattice", "cells", ["3", "2", "3.7"], {"cell_start": "i4", "sorted": "i4", "cell_of": "i4"}),
    ("cells_rho16_n4", "cells", ["4", "1", "16.0"], {"cell_start": "i4", "sorted": "i4", "cell_of": "i4"}),
    ("cells_rho0.5", "cells", ["5", "0", "0.5"], {"cell_start": "i4", "sorted": "i4", "cell_of": "i4"}),
    ("pbc", "pbc", ["5"], {"dxy": "f8", "wrap_out": "f8"}),
    ("rhs_k_rho4", "rhs_k", ["7", "0", "0.5", "0.01"] + CHIM(0.01, 1.0, 4.0, 1.45, 0.01), {"k_next": "f8"}),
    ("rhs_k_rho1_outside", "rhs_k", ["8", "1", "1.0", "0.01"] + CHIM(0.5, 2.0, 1.0, 0.3, 0.1), {"k_next": "f8"}),
    ("rhs_k_rho4_c0", "rhs_k", ["10", "0", "0.0", "0.005"] + CHIM(1.0, 1.0, 4.0, 1.45, 0.01), {"k_next": "f8"}),
    ("rhs_k_allpairs", "rhs_k", ["9", "1", "0.5", "0.01"] + CHIM(0.01, 1.0, 20.0, 1.45, 0.01), {"k_next": "f8"}),
    ("rhs_ito_rho4", "rhs_ito", ["7", "0", "0", "0.01"] + CHIM(0.01, 1.0, 4.0, 1.45, 0.01),
     {"derivative": "f8", "additional": "f8"}),
    ("rhs_ito_rho1_inout", "rhs_ito", ["12", "1", "1", "0.01"] + CHIM(2.0, 1.0, 1.0, 0.0, 0.4),
     {"derivative": "f8", "additional": "f8"}),
    ("rhs_ito_allpairs", "rhs_ito", ["9", "0", "1", "0.01"] + CHIM(0.01, 1.0, 20.0, 1.45, 0.01),
     {"derivative": "f8", "additional": "f8"}),
    ("se_rho4", "se", ["11", "42", "3", "0.01"] + CHIM(0.01, 1.0, 4.0, 1.45, 0.01),
     {"noise": "f8", "state1": "f8", "state": "f8"}),
    ("se_rho1", "se", ["13", "7", "3", "0.01"] + CHIM(0.5, 1.0, 1.0, 1.45, 0.2),
     {"noise": "f8", "state1": "f8", "state": "f8"}),
    ("srk_rho4", "srk", ["11", "42", "3", "0.01"] + CHIM(0.01, 1.0, 4.0, 1.45, 0.01),
     {"noise": "f8", "state1": "f8", "state": "f8"}),
    ("srk_rho1", "srk", ["14", "9", "3", "0.005"] + CHIM(1.0, 2.0, 1.0, 1.2, 0.4),
     {"noise": "f8", "state1": "f8", "state": "f8"}),
    ("rk4_rho1", "rk4", ["15", "0", "3", "0.01"] + CHIM(0.5, 1.0, 1.0, 1.45, 0.0), {"state1": "f8", "state": "f8"}),
    ("rk4_rho4", "rk4", ["16", "0", "3", "0.01"] + CHIM(2.0, 1.0, 4.0, 1.45, 0.0), {"state1": "f8", "state": "f8"}),
    ("srk_traj20", "srk", ["17", "123", "20", "0.01"] + CHIM(1.0, 1.0, 2.0, 1.45, 0.01),
     {"noise": "f8", "state": "f8"}),
]

# The reference's MPI shared-memory classes (VortexArraysSystemPtr,
# ChimeraSystemPtr, the *StepperPtr steppers) through ref_harness_mpi under
# mpiexec.  Vortex arguments: velocity mu+ mu- xi_a xi_r D kappa rho; at
# density 2 (N = 8192, L = 64) rho = 2.236 gives the workload's ~31 partners,
# with xi_a = rho / 5 and xi_r = rho / 10 as in the reference's set.
VORT = lambda v, mp, mm, xa, xr, D, kappa, rho: [repr(x) for x in (v, mp, mm, xa, xr, D, kappa, rho)]
VW = VORT(1.0, 100.0, 100.0, 0.4472, 0.2236, 0.01, 20.0, 2.236)
VW_D0 = VORT(1.0, 100.0, 100.0, 0.4472, 0.2236, 0.0, 20.0, 2.236)
MPI_CASES = [
    ("vrhs_k", "vrhs_k", ["31", "0", "0.5", "0.01"] + VW, {"k_next": "f8"}),
    ("vrhs_k_outside", "vrhs_k", ["32", "1", "1.0", "0.02"] + VW, {"k_next": "f8"}),
    ("vrhs_k_allpairs", "vrhs_k", ["33", "0", "0.5", "0.01"] + VORT(1.0, 100.0, 80.0, 8.0, 4.0, 0.01, 20.0, 20.0),
     {"k_next": "f8"}),
    ("vrhs_ito", "vrhs_ito", ["34", "0", "0", "0.01"] + VW, {"derivative": "f8", "additional": "f8"}),
    ("vrhs_ito_inout", "vrhs_ito", ["35", "1", "1", "0.01"] + VW, {"derivative": "f8", "additional": "f8"}),
    ("vse", "vse", ["36", "42", "3", "0.01"] + VW, {"noise": "f8", "state1": "f8", "state": "f8"}),
    ("vsrk", "vsrk", ["37", "43", "3", "0.01"] + VW, {"noise": "f8", "state1": "f8", "state": "f8"}),
    ("vrk4", "vrk4", ["38", "0", "4", "0.01"] + VW_D0, {"state1": "f8", "state": "f8"}),
    ("crhs_k", "crhs_k", ["39", "1", "0.5", "0.01"] + CHIM(0.5, 1.0, 2.0, 1.45, 0.01), {"k_next": "f8"}),
    ("crhs_ito", "crhs_ito", ["40", "0", "1", "0.01"] + CHIM(0.5, 1.0, 2.0, 1.45, 0.05),
     {"derivative": "f8", "additional": "f8"}),
    ("cse_ptr", "cse", ["41", "44", "3", "0.01"] + CHIM(0.5, 1.0, 2.0, 1.45, 0.05),
     {"noise": "f8", "state1": "f8", "state": "f8"}),
    ("csrk_ptr", "csrk", ["42", "45", "3", "0.01"] + CHIM(0.5, 1.0, 2.0, 1.45, 0.05),
     {"noise": "f8", "state1": "f8", "state": "f8"}),
    ("crk4_ptr", "crk4", ["43", "0", "3", "0.01"] + CHIM(2.0, 1.0, 2.0, 1.45, 0.0), {"state1": "f8", "state": "f8"}),
]
# deterministic cases also run on several ranks (ThreadSharedMemory slices):
# their outputs must not depend on the rank count
MULTI_RANK = {"vrhs_k": ["k_next"], "vrhs_ito_inout": ["derivative", "additional"], "vrk4": ["state"],
              "crhs_ito": ["derivative", "additional"], "crk4_ptr": ["state"]}
RANKS = (2, 4)
MPIEXEC = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")
HARNESS_MPI = os.path.join(ROOT, "oracle", "_ref", "ref_harness_mpi")

# Ensemble of the reference's own SRK stepper with its own mt19937 noise
# (seeds 1000 + r): |R|(t) every 10 steps, for the 95 % CI comparison.
ENSEMBLE = ("ens_srk", ["21", "24", "200", "10", "0.01"] + CHIM(1.0, 1.0, 4.0, 1.45, 0.05))

# Ensemble of the reference's vortex SRK stepper (VortexArraysSystemPtr +
# StochasticRungeKuttaStepperPtr, its own mt19937 noise, generator seeds 2000 + r,
# one rank): |R|(t) every 10 steps from aligned phases (R = 1 at t = 0) as the
# state dephases, dt = 0.001, D = 0.5.
ENSEMBLE_VORTEX = ("ens_vsrk", ["51", "24", "200", "10", "0.001"] + VORT(1.0, 100.0, 100.0, 0.4472, 0.2236, 0.5,
                                                                          20.0, 2.236))

SAMPLE_IDX = [0, 1, 2, 3, 100, 1000, 4095, 4096, 8191]


def digest(a: np.ndarray) -> dict:
    a = np.ascontiguousarray(a)
    d = {
        "dtype": a.dtype.str,
        "size": int(a.size),
        "sha256": hashlib.sha256(a.tobytes()).hexdigest(),
        "sample_idx": [i for i in SAMPLE_IDX if i < a.size],
    }
    d["sample"] = [a[i].item() for i in d["sample_idx"]]
    if a.dtype.kind == "f" and a.size:
        d["sum"] = float(a.sum())
        d["absmax"] = float(np.abs(a).max())
    return d


def run_case(case, args, outputs, tmp, ranks=0):
    cmd = [HARNESS, case, tmp] + args if not ranks else [MPIEXEC, "-n", str(ranks), HARNESS_MPI, case, tmp] + args
    subprocess.run(cmd, check=True)
    return {k: digest(np.fromfile(os.path.join(tmp, k + ".bin"), np.dtype(dt))) for k, dt in outputs.items()}


def vortex_ensemble(tmp):
    case, args = ENSEMBLE_VORTEX
    subprocess.run([MPIEXEC, "-n", "1", HARNESS_MPI, case, tmp] + args, check=True)
    members = int(args[1])
    R = np.fromfile(os.path.join(tmp, "R.bin")).reshape(members, -1)
    return {"case": case, "args": args, "members": members, "every": int(args[3]), "R": R.tolist()}


def main():
    if not os.path.exists(HARNESS) or not os.path.exists(HARNESS_MPI):
        sys.exit("build the reference harnesses first: make -C oracle ref")
    if len(sys.argv) > 1 and sys.argv[1] == "--only-vortex-ensemble":
        # regenerates that entry alone, keeping the rest of the fixture
        path = os.path.join(HERE, "golden.json")
        with open(path) as f:
            out = json.load(f)
        with tempfile.TemporaryDirectory() as tmp:
            out["ensemble_vortex"] = vortex_ensemble(tmp)
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        print("wrote", path)
        return
    out = {
        "about": ("Golden vectors from the reference's own classes: the non-MPI ones (ref_harness.cpp) and "
                  "the MPI shared-memory ones under mpiexec (ref_harness_mpi.cpp, cases marked harness=mpi); "
                  "N=8192, L=64"),
        "N": N, "L": L,
        "toolchain": {
            "cxx": subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0],
            "libc": " ".join(platform.libc_ver()),
            "flags": "-std=c++17 -O2 -ffp-contract=off",
            "mpi": subprocess.run([MPIEXEC, "--version"], capture_output=True, text=True).stdout.split("\n")[1].strip(),
        },
        "cases": {},
    }
    with tempfile.TemporaryDirectory() as tmp:
        for name, case, args, outputs in CASES:
            out["cases"][name] = {"case": case, "args": args, "outputs": run_case(case, args, outputs, tmp)}
            print(name, flush=True)
        for name, case, args, outputs in MPI_CASES:
            res = run_case(case, args, outputs, tmp, ranks=1)
            entry = {"case": case, "args": args, "harness": "mpi", "outputs": res}
            if name in MULTI_RANK:
                entry["ranks"] = {}
                for r in RANKS:
                    rr = run_case(case, args, {k: outputs[k] for k in MULTI_RANK[name]}, tmp, ranks=r)
                    entry["ranks"][str(r)] = {k: v["sha256"] for k, v in rr.items()}
            out["cases"][name] = entry
            print(name, flush=True)
        case, args = ENSEMBLE
        subprocess.run([HARNESS, case, tmp] + args, check=True)
        R = np.fromfile(os.path.join(tmp, "R.bin"))
        members = int(args[1])
        R = R.reshape(members, -1)
        out["ensemble"] = {"case": case, "args": args, "members": members, "every": int(args[3]),
                           "R": R.tolist()}
        out["ensemble_vortex"] = vortex_ensemble(tmp)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
