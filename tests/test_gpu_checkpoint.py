"""Exact checkpoints (spc2_checkpoint_*) and the engine's restarts.

"Run K steps, save, load into a fresh handle, run M steps" must end bit for bit
where "run K + M steps" ends, on every neighbour path and method, with the save
falling between two Verlet list builds and the RK4 stepper's carried phase
derivatives in play (a sparse vortex state has particles without alignment
partners).  The engine's resume (spc2_run resume=...) must continue its run in
the same way, and its two initial-condition loaders must read the frame or the
floats they are given: InitializeSystemStateFromPreviousSolution
(SimulationEnginePtr.cpp:52-73, a frame of a trajectory) and
InitializeSystemStateFromFile (SimulationEngine.cpp:52-73, 3N floats).
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import spc2_langevin_integration_amd as spc2

pytestmark = pytest.mark.gpu

RUNNER = os.path.join(os.path.dirname(spc2.LIB_PATH), "spc2_run")
VORT = dict(velocity=0.025, mu_plus=100.0, mu_minus=100.0, xi_a=0.01, kappa=20.0)


@pytest.fixture(scope="module")
def gpu():
    spc2.load()
    with spc2.Sim(spc2.params(n=16)) as s:  # fails loudly without a GPU
        s.init_random_state()


def case_params(name, prec):
    if name == "vortex-sparse":  # density 400, cohesion range between rho and the cell side
        return spc2.params("vortex", prec, "verlet", n=99212, L=15.749, rho=0.05, xi_r=0.05015, D=0.0, seed=43,
                           verlet_skin=0.004, **VORT)
    if name == "vortex-dense":
        n = 40000
        return spc2.params("vortex", prec, "verlet", n=n, L=math.sqrt(n / 4000.0), rho=0.05, xi_r=0.005, D=0.01,
                           seed=44, verlet_skin=0.004, **VORT)
    if name == "chimera-cells":
        n = 30000
        return spc2.params("chimera", prec, "cells", n=n, L=math.sqrt(n / 4000.0), rho=0.1, v0=0.05, sigma=1.0,
                           alpha=1.45, D=0.02, seed=45)
    if name == "chimera-verlet":
        n = 30000
        return spc2.params("chimera", prec, "verlet", n=n, L=math.sqrt(n / 4000.0), rho=0.1, v0=0.2, sigma=1.0,
                           alpha=1.45, D=0.02, seed=46)
    n = 500  # all pairs (rho > L / 4)
    return spc2.params("chimera", prec, "auto", n=n, L=math.sqrt(n / 4000.0), rho=0.1, v0=0.01, sigma=1.0,
                       alpha=1.45, D=0.02, seed=47)


CASES = [("vortex-sparse", "rk4", "fp32k"), ("vortex-sparse", "rk4", "fp64"), ("vortex-dense", "srk", "fp32k"),
         ("vortex-dense", "se", "fp64f"), ("chimera-cells", "rk4", "fp64"), ("chimera-cells", "srk", "fp32k"),
         ("chimera-verlet", "se", "fp32k"), ("chimera-verlet", "rk4", "fp64"), ("chimera-allpairs", "srk", "fp64")]


@pytest.mark.parametrize("name,method,prec", CASES)
def test_checkpoint_continues_bit_for_bit(gpu, name, method, prec):
    p = case_params(name, prec)
    K, M, dt = 5, 14, 0.01
    with spc2.Sim(p) as a:
        a.init_random_state(0.3)
        a.step(method, dt, 0, K + M)
        straight = a.get_state()
        builds_straight = a.verlet_stats()[0] if p.neighbour == spc2.NEIGH_VERLET else 0
    with spc2.Sim(p) as b:
        b.init_random_state(0.3)
        b.step(method, dt, 0, K)
        blob = b.checkpoint()
        saved = b.get_state()
    with spc2.Sim(p) as c:
        c.restore(blob)
        assert np.array_equal(c.get_state(), saved)
        c.step(method, dt, K, M)
        resumed = c.get_state()
        if p.neighbour == spc2.NEIGH_VERLET:
            # the save fell between list builds, and the resumed run rebuilt again
            assert builds_straight >= 2 and c.verlet_stats()[0] >= 2
    assert np.array_equal(resumed, straight), np.abs(resumed - straight).max()


def test_checkpoint_refuses_other_parameters(gpu):
    p = case_params("chimera-cells", "fp64")
    with spc2.Sim(p) as a:
        a.init_random_state(0.1)
        blob = a.checkpoint()
    for change in (dict(seed=99), dict(sigma=1.1), dict(precision=spc2.FP32_KAHAN)):
        q = case_params("chimera-cells", "fp64")
        for k, v in change.items():
            setattr(q, k, v)
        with spc2.Sim(q) as b:
            with pytest.raises(spc2.Spc2Error, match="other parameters"):
                b.restore(blob)
    with spc2.Sim(p) as b:
        with pytest.raises(spc2.Spc2Error, match="not an spc2 checkpoint"):
            b.restore(b"x" * len(blob))


def run(tmp_path, tag, **cfg):
    args = [RUNNER] + [f"{k}={v}" for k, v in cfg.items()]
    print("spc2_run", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def frames(path, n):
    return np.fromfile(path, dtype=np.float32).reshape(-1, 1 + 3 * n)


ENGINE = dict(system="vortex", n=40000, L=math.sqrt(10.0), rho=0.05, v0=0.025, mu_plus=100.0, mu_minus=100.0,
              xi_a=0.01, xi_r=0.005, kappa=20.0, D=0.01, dt=0.01, seed=11, state_every=3, summary_every=3)


@pytest.mark.parametrize("method,prec,neighbour,noise", [("rk4", "fp32k", "verlet", "philox"),
                                                         ("srk", "fp64", "cells", "mt19937")])
def test_engine_resume_continues_the_run(gpu, tmp_path, method, prec, neighbour, noise):
    """spc2_run steps=K checkpoint=f, then resume=f steps=M, ends where one run
    of K + M steps ends (frames, time stamps, summary values), with the
    mt19937 stream restored where the noise comes from it."""
    cfg = dict(ENGINE, method=method, precision=prec, neighbour=neighbour, verlet_skin=0.004, noise=noise,
               init="mt19937" if noise == "mt19937" else "philox")
    n = cfg["n"]
    ck = tmp_path / "run.ckpt"
    full = run(tmp_path, "full", steps=21, state=tmp_path / "full.bin", summary=tmp_path / "full.txt", **cfg)
    run(tmp_path, "first", steps=9, checkpoint=ck, checkpoint_every=4, **cfg)
    second = run(tmp_path, "second", steps=12, resume=ck, state=tmp_path / "second.bin",
                 summary=tmp_path / "second.txt", **cfg)
    assert full["steps"] == 21 and second["steps"] == 21
    assert full["R"] == second["R"] and full["psi"] == second["psi"]
    ff, fs = frames(tmp_path / "full.bin", n), frames(tmp_path / "second.bin", n)
    # the full run's frames at steps 9, 12, ..., 21 are the resumed run's frames
    assert np.array_equal(ff[3:], fs), np.abs(ff[3:] - fs).max()
    lf = (tmp_path / "full.txt").read_text().splitlines()
    ls = (tmp_path / "second.txt").read_text().splitlines()
    assert lf[-len(ls):] == ls


def test_engine_checkpoint_every_writes_the_latest_state(gpu, tmp_path):
    """checkpoint_every=4 in a 9-step run: the file holds the state at the end
    of the run (the last write), and resuming it with steps=0 reproduces it."""
    cfg = dict(ENGINE, method="se", precision="fp64")
    ck = tmp_path / "c.ckpt"
    run(tmp_path, "a", checkpoint=ck, checkpoint_every=4, state=tmp_path / "a.bin", **dict(cfg, state_every=9, steps=9))
    run(tmp_path, "b", resume=ck, state=tmp_path / "b.bin", **dict(cfg, steps=0))
    fa, fb = frames(tmp_path / "a.bin", cfg["n"]), frames(tmp_path / "b.bin", cfg["n"])
    assert np.array_equal(fa[-1], fb[0])


def test_previous_solution_frame_seek_and_float_file(gpu, tmp_path):
    """init_file=F init_frame=k starts from frame k of F (time stamp skipped, as
    SimulationEnginePtr.cpp:58-66 seeks (1 + 3N) k floats); init_state_file=G
    starts from the 3N floats at the start of G (SimulationEngine.cpp:60-68)."""
    n, L = 3000, math.sqrt(3000 / 4000.0)
    cfg = dict(system="chimera", n=n, L=L, rho=0.1, v0=0.05, sigma=1.0, alpha=1.45, D=0.02, method="se",
               precision="fp64", dt=0.01, seed=5, state_every=4, summary_every=4)
    run(tmp_path, "traj", steps=12, state=tmp_path / "traj.bin", **cfg)
    tr = frames(tmp_path / "traj.bin", n)
    assert tr.shape[0] == 4
    for k in (0, 2, 3):
        run(tmp_path, f"seek{k}", steps=0, init_file=tmp_path / "traj.bin", init_frame=k,
            state=tmp_path / f"seek{k}.bin", **cfg)
        got = frames(tmp_path / f"seek{k}.bin", n)
        assert np.array_equal(got[0, 1:], tr[k, 1:]), k
    r = subprocess.run([RUNNER, "steps=0", f"init_file={tmp_path / 'traj.bin'}", "init_frame=4"] +
                       [f"{k}={v}" for k, v in cfg.items()], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "beyond the end" in r.stderr
    st = O.make_state(8, 0, n, L, 0.1).astype(np.float32)
    st.tofile(tmp_path / "ic.bin")
    run(tmp_path, "ic", steps=0, init_state_file=tmp_path / "ic.bin", state=tmp_path / "ic_out.bin", **cfg)
    got = frames(tmp_path / "ic_out.bin", n)
    assert np.array_equal(got[0, 1:], st)
