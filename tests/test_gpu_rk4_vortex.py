"""The vortex-arrays RK4 step, RungeKutta4StepperPtr::DoStep for
VortexArraysSystemPtr (RungeKutta4StepperPtr.cpp:45-65 over the k-form RHS
VortexArraysSystemPtr.cpp:357-507), against the oracle's step_rk4 with its k
buffers carried from step to step.

Two workloads:
  * dense: the N = 1,048,576 workload's parameters (rho = 0.05, xi_a = 0.01,
    xi_r = 0.005, v = 0.025, mu+- = 100, kappa = 20, density 4000) at N = 40,000;
  * sparse: density 400 with the cohesion range xi_r = 0.05015 between rho = 0.05
    and the reference's cell side (L / int(L / rho) = 0.0501561): particles with
    a cohesion partner but none within rho add the cohesion mean onto the phase
    derivative that the k buffer kept from the previous step
    (VortexArraysSystemPtr.cpp:495-503), and particles without partners keep it
    as is.

Every neighbour path runs: linked cells, the Verlet list's gather passes, and
(FP32) its LDS-staged passes with lane slots by list length forced off and on.
The Verlet skin (0.004) makes the list rebuild every 8 steps.

Bars (tests/_parity.py): single steps from the oracle's state within 1e-12
(FP64 forms) and 1e-5 (FP32-compensated) of the increment scale; FP64
trajectories of 24 steps within that bar summed over the steps taken.
"""
import math
import os

import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle as O
import spc2_langevin_integration_amd as spc2
from _parity import TOL, TrajectoryBar, assert_state, periodic_diff

pytestmark = pytest.mark.gpu

DT, STEPS, SKIN = 0.01, 24, 0.004
BASE = dict(velocity=0.025, mu_plus=100.0, mu_minus=100.0, xi_a=0.01, kappa=20.0, D=0.0)
CASES = {
    "dense": dict(n=40000, L=math.sqrt(10.0), rho=0.05, xi_r=0.005, seed=42),
    "sparse": dict(n=99212, L=15.749, rho=0.05, xi_r=0.05015, seed=43),
}
# (neighbour mode, SPC2_VERLET_LDS, SPC2_VERLET_SORT); the FP64 forms always use
# the gather passes
PATHS = {
    "cells": ("cells", None, None),
    "verlet-gather": ("verlet", "0", None),
    "verlet-lds": ("verlet", "1", "0"),
    "verlet-lds-sorted": ("verlet", "1", "1"),
}
COMBOS = [("fp64", "cells"), ("fp64", "verlet-gather"), ("fp64", "verlet-lds"), ("fp64f", "cells"),
          ("fp64f", "verlet-gather"), ("fp64f", "verlet-lds"), ("fp32k", "cells"), ("fp32k", "verlet-gather"),
          ("fp32k", "verlet-lds"), ("fp32k", "verlet-lds-sorted")]


def threads():
    return max(1, min(16, os.cpu_count() or 1))


def sim_params(case, prec, neighbour):
    c = CASES[case]
    kw = dict(BASE, xi_r=c["xi_r"])
    return spc2.params("vortex", prec, neighbour, n=c["n"], L=c["L"], rho=c["rho"], seed=c["seed"],
                       verlet_skin=SKIN, **kw)


def oracle_sys(case):
    c = CASES[case]
    kw = dict(BASE, xi_r=c["xi_r"])
    return O.vortex(c["n"], c["L"], rho=c["rho"], threads=threads(), **kw)


_TRAJ = {}


def oracle_trajectory(case):
    """States before every step and after the last (24 RK4 steps), with the k
    buffers carried across steps as the stepper's members are."""
    if case not in _TRAJ:
        c = CASES[case]
        st, _ = O.init_state(c["seed"], c["n"], c["L"], c["L"], 0.0)
        sysp = oracle_sys(case)
        kbuf = np.zeros(12 * c["n"])
        traj = [st]
        for _ in range(STEPS):
            traj.append(O.step_rk4(sysp, traj[-1], DT, kbuf))
        _TRAJ[case] = traj
    return _TRAJ[case]


@pytest.fixture(scope="module")
def gpu():
    spc2.load()
    with spc2.Sim(spc2.params(n=16)) as s:  # fails loudly without a GPU
        s.init_random_state()


def open_sim(monkeypatch, case, prec, path):
    neighbour, lds, sort = PATHS[path]
    with monkeypatch.context() as m:
        for var, val in (("SPC2_VERLET_LDS", lds), ("SPC2_VERLET_SORT", sort)):
            if val is None:
                m.delenv(var, raising=False)
            else:
                m.setenv(var, val)
        return spc2.Sim(sim_params(case, prec, neighbour))


def check_layout(s, path):
    if PATHS[path][0] != "verlet":
        return
    lds, cap = s.verlet_layout()
    if path.startswith("verlet-lds"):
        assert lds == 1 and cap > 0, (path, lds, cap)
    else:
        assert (lds, cap) == (0, 0), (path, lds, cap)


def test_sparse_case_has_cohesion_only_particles():
    """The sparse workload exercises the cohesion-only update: some particles
    have a partner within xi_r but none within rho (periodic distances)."""
    c = CASES["sparse"]
    st, _ = O.init_state(c["seed"], c["n"], c["L"], c["L"], 0.0)
    pos = np.stack([st[0::3], st[1::3]], 1)
    tree = cKDTree(pos, boxsize=c["L"])
    d, _ = tree.query(pos, k=2)
    nearest = d[:, 1]
    cohesion_only = int(((nearest > c["rho"]) & (nearest <= c["xi_r"])).sum())
    isolated = int((nearest > c["xi_r"]).sum())
    print(f"sparse: {cohesion_only} cohesion-only and {isolated} isolated particles of {c['n']}")
    h = c["L"] / int(c["L"] / c["rho"])
    assert c["rho"] < c["xi_r"] <= h
    assert cohesion_only >= 10 and isolated >= 100


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec,path", COMBOS)
def test_vortex_rk4_single_steps_match_oracle(gpu, monkeypatch, case, prec, path):
    """Each step starts from the oracle's state; the handle's k buffers carry
    over from its previous step as the oracle's do."""
    traj = oracle_trajectory(case)
    L = CASES[case]["L"]
    worst = 0.0
    s = open_sim(monkeypatch, case, prec, path)
    with s:
        for k in range(STEPS):
            s.set_state(traj[k])
            if k == 0:
                check_layout(s, path)
            s.step("rk4", DT, k, 1)
            worst = max(worst, assert_state(s.get_state(), traj[k + 1], traj[k], L, TOL[prec],
                                            f"{case}/{prec}/{path} step {k}"))
    print(f"{case}/{prec}/{path}: largest error / bar over {STEPS} steps = {worst:.3e}")


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec,path", [c for c in COMBOS if c[0] != "fp32k"])
def test_vortex_rk4_fp64_trajectory_matches_oracle(gpu, monkeypatch, case, prec, path):
    """24 uninterrupted steps (three Verlet list builds): the FP64 forms follow
    the oracle's trajectory within the accumulated single-step bar."""
    traj = oracle_trajectory(case)
    L = CASES[case]["L"]
    worst = 0.0
    bar = TrajectoryBar(TOL[prec])
    s = open_sim(monkeypatch, case, prec, path)
    with s:
        s.set_state(traj[0])
        for k in range(STEPS):
            s.step("rk4", DT, k, 1)
            worst = max(worst, bar.check(s.get_state(), traj[k + 1], traj[k], L,
                                         f"{case}/{prec}/{path} trajectory step {k}"))
        if PATHS[path][0] == "verlet":
            assert s.verlet_stats()[0] >= 3
    print(f"{case}/{prec}/{path} trajectory: largest error / bar = {worst:.3e}")


@pytest.mark.parametrize("prec,path", [("fp32k", "verlet-lds"), ("fp64", "verlet-gather"), ("fp32k", "cells")])
def test_vortex_rk4_full_size_steps(gpu, monkeypatch, prec, path):
    """N = 1,048,576 at density 4000 (L = 16.19086): two steps from the
    workload's initial state against the oracle run on the host's cores."""
    n = 1048576
    L = math.sqrt(n / 4000.0)
    st, _ = O.init_state(42, n, L, L, 0.0)
    sysp = O.vortex(n, L, rho=0.05, threads=threads(), xi_r=0.005, **BASE)
    kbuf = np.zeros(12 * n)
    neighbour, lds, sort = PATHS[path]
    with monkeypatch.context() as m:
        if lds is None:
            m.delenv("SPC2_VERLET_LDS", raising=False)
        else:
            m.setenv("SPC2_VERLET_LDS", lds)
        m.delenv("SPC2_VERLET_SORT", raising=False)
        p = spc2.params("vortex", prec, neighbour, n=n, L=L, rho=0.05, xi_r=0.005, seed=42, **BASE)
        s = spc2.Sim(p)
    with s:
        s.init_random_state(0.0)
        assert np.array_equal(s.get_state(), st)
        check_layout(s, path)
        before = st
        for k in range(2):
            after = O.step_rk4(sysp, before, DT, kbuf)
            s.set_state(before)
            s.step("rk4", DT, k, 1)
            r = assert_state(s.get_state(), after, before, L, TOL[prec], f"N=1M {prec}/{path} step {k}")
            print(f"N=1M {prec}/{path} step {k}: error / bar = {r:.3e}")
            before = after
    # the stepper's phase increment is of order dt |a_phi|: a wrong term would
    # move it by far more than the bar
    assert np.abs(periodic_diff(after, st, L)[2::3]).max() > 1e-3
