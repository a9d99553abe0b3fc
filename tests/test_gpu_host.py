"""GPU tests of the C++ drop-in host layer (spc2_run over libspc2_host.so) and of
the frame-snapshot and no-wrap entry points.

Bars:
  * trajectory frames (float t + 3N floats) against the oracle's FP64
    trajectory rounded to float: within 1 float ulp;
  * the reference's own loop through the drop-in classes on a host state
    (loop=host) against the device-resident loop (loop=device): the same bytes
    in the frame and summary files;
  * mt19937 mode: the reference's draws (std::mt19937(seed) through
    std::uniform_real_distribution and std::normal_distribution) replayed by the
    oracle, within 1 float ulp.
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


def run(tmp_path, tag, **cfg):
    state = tmp_path / f"{tag}.bin"
    summary = tmp_path / f"{tag}.txt"
    args = [RUNNER, f"state={state}", f"summary={summary}"] + [f"{k}={v}" for k, v in cfg.items()]
    print("spc2_run", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    n = int(cfg["n"])
    frames = np.fromfile(state, dtype=np.float32).reshape(-1, 1 + 3 * n)
    lines = [tuple(float(x) for x in ln.split("\t")) for ln in summary.read_text().splitlines()]
    return out, frames, lines, state.read_bytes(), summary.read_bytes()


def assert_frame(frame, ref_state, what):
    ref = ref_state.astype(np.float32)
    ulp = np.spacing(np.abs(ref))
    err = np.abs(frame.astype(np.float64) - ref.astype(np.float64))
    assert (err <= ulp).all(), f"{what}: {int((err > ulp).sum())} values beyond 1 float ulp (max {err.max():.3e})"


CHIM = dict(system="chimera", n=4000, L=1.0, rho=0.1, v0=0.01, sigma=1.0, alpha=1.45, D=0.01)


@pytest.fixture(scope="module")
def gpu():
    spc2.load()
    with spc2.Sim(spc2.params(n=16)) as s:  # fails loudly without a GPU
        s.init_random_state()


def test_device_loop_matches_oracle_trajectory(gpu, tmp_path):
    n, L, dt, steps = 4000, 1.0, 0.005, 20
    out, frames, lines, _, _ = run(tmp_path, "dev", method="srk", precision="fp64", dt=dt, steps=steps, seed=42,
                                   omega_std=0.1, state_every=10, summary_every=10, **CHIM)
    assert out["steps"] == steps and frames.shape[0] == 3 and len(lines) == 2
    sysp = O.chimera(n, L, 0.1, 0.01, 1.0, 1.45, 0.01, threads=8)
    st, om = O.init_state(42, n, L, L, 0.1)
    traj = [st]
    for k in range(steps):
        traj.append(O.step_srk(sysp, traj[-1], dt, O.philox_normals(42, k, 0, n, 2), omega=om))
    t = [0.0]
    for k in range(steps):
        t.append(t[-1] + dt)
    for f, j in enumerate((0, 10, 20)):  # frames at the first call and every 10th
        assert frames[f, 0] == np.float32(t[j])
        assert_frame(frames[f, 1:], traj[j], f"frame {f}")
    for line, j in zip(lines, (1, 11)):  # summary at the first post-step call and every 10th
        R, psi = O.order_param(traj[j])
        assert line[0] == pytest.approx(t[j], rel=1e-5)
        assert line[1] == pytest.approx(R, rel=5e-6) and line[2] == pytest.approx(psi, rel=5e-6, abs=1e-6)


def test_float_summary_matches_oracle_f32(gpu, tmp_path):
    """float_summary=1 writes the reference's own float order parameter
    (BinaryObserverPtr.cpp:237-247: complex<float> sum of float cos and sin in
    index order, divided by N) as text, t, |R| and arg R at the stream's default
    6 significant digits: the same text as the oracle's orc_order_param_f32 of
    the oracle's trajectory."""
    n, L, dt, steps = 4000, 1.0, 0.005, 20
    _, _, _, _, summary = run(tmp_path, "f32", method="srk", precision="fp64", dt=dt, steps=steps, seed=42,
                              omega_std=0.1, state_every=10, summary_every=5, float_summary=1, **CHIM)
    sysp = O.chimera(n, L, 0.1, 0.01, 1.0, 1.45, 0.01, threads=8)
    st, om = O.init_state(42, n, L, L, 0.1)
    traj, t = [st], [0.0]
    for k in range(steps):
        traj.append(O.step_srk(sysp, traj[-1], dt, O.philox_normals(42, k, 0, n, 2), omega=om))
        t.append(t[-1] + dt)
    lines = summary.decode().splitlines()
    assert len(lines) == 4
    for line, j in zip(lines, (1, 6, 11, 16)):
        R, psi = O.order_param_f32(traj[j])
        expect = "\t".join(f"{float(v):g}" for v in (np.float32(t[j]), np.float32(R), np.float32(psi)))
        assert line == expect, (j, line, expect)


VORT = dict(system="vortex", n=3000, L=1.0, rho=0.1, v0=0.1, mu_plus=100.0, mu_minus=100.0, xi_a=0.04,
            xi_r=0.02, kappa=20.0, D=0.01)


@pytest.mark.parametrize("cfg,method", [(CHIM, "se"), (CHIM, "rk4"), (VORT, "srk"), (VORT, "rk4"),
                                        (dict(CHIM, precision="fp32k"), "srk")])
def test_host_loop_bit_identical_to_device_loop(gpu, tmp_path, cfg, method):
    """The reference's loop (DoStep on a host state, host wrap, observer every
    step) through the drop-in classes gives the same bytes as the
    device-resident loop."""
    common = dict(method=method, dt=0.005, steps=12, seed=9, state_every=4, summary_every=3, float_summary=1)
    common.update(cfg)
    _, fd, _, bd, sd = run(tmp_path, "device", loop="device", **common)
    _, fh, _, bh, sh = run(tmp_path, "host", loop="host", **common)
    assert fd.shape[0] == 4
    assert bd == bh, f"frames differ: max {np.abs(fd - fh).max():.3e}"
    assert sd == sh


@pytest.mark.parametrize("loop", ["device", "host"])
@pytest.mark.parametrize("system,method", [("chimera", "srk"), ("vortex", "se")])
def test_mt19937_mode_replays_reference_draws(gpu, tmp_path, loop, system, method):
    n, L, dt, steps, seed = 1000, 1.0, 0.01, 6, 7
    cfg = dict(CHIM if system == "chimera" else VORT, n=n, L=L)
    _, frames, _, _, _ = run(tmp_path, "mt", loop=loop, init="mt19937", noise="mt19937", method=method, dt=dt,
                             steps=steps, seed=seed, state_every=3, summary_every=3, **cfg)
    m = O.MT19937Normals(seed)
    st = m.reference_initial_state(n, L)
    if system == "chimera":
        sysp = O.chimera(n, L, 0.1, 0.01, 1.0, 1.45, 0.01, threads=8)
    else:
        sysp = O.vortex(n, L, rho=0.1, velocity=0.1, mu_plus=100.0, mu_minus=100.0, xi_a=0.04, xi_r=0.02,
                        kappa=20.0, D=0.01, threads=8)
    if method == "srk":
        m.normals(n)  # the SRK stepper's natural-frequency draws, N(0, 0)
    traj = [st]
    for k in range(steps):
        if method == "srk":
            traj.append(O.step_srk(sysp, traj[-1], dt, m.normals(2 * n), omega=np.zeros(n)))
        else:
            traj.append(O.step_se(sysp, traj[-1], dt, m.normals(n), variant=0))
    for f, j in enumerate((0, 3, 6)):
        assert_frame(frames[f, 1:], traj[j], f"{loop} frame {f}")


def test_t1_loop_count(gpu, tmp_path):
    """t1 >= 0 runs the reference's `while (t <= t_1) t += dt` loop."""
    dt, t1 = 0.01, 0.05
    k, t = 0, 0.0
    while t <= t1:
        t += dt
        k += 1
    out, frames, _, _, _ = run(tmp_path, "t1", method="se", dt=dt, t1=t1, state_every=1, summary_every=1,
                               **dict(CHIM, n=500))
    assert out["steps"] == k and frames.shape[0] == k + 1
    assert frames[-1, 0] == np.float32(t)


def test_snapshot_and_no_wrap_step(gpu):
    n, L = 20000, math.sqrt(20000 / 4000.0)
    st = O.make_state(3, 0, n, L, 0.1)
    p = spc2.params("chimera", "fp64", n=n, L=L, rho=0.1, v0=0.5, sigma=1.0, alpha=1.45, D=0.05, seed=5)
    with spc2.Sim(p) as s:
        s.set_state(st)
        frame = s.snapshot()
        assert np.array_equal(frame, st.astype(np.float32))
        s.step("srk", 0.05, 0, 1, wrap=False)
        raw = s.get_state()
        s.set_state(st)
        s.step("srk", 0.05, 0, 1)
        wrapped = s.get_state()
        buf = spc2.PinnedFrame(3 * n)
        try:
            assert np.array_equal(s.snapshot(buf), wrapped.astype(np.float32))
        finally:
            buf.free()
    assert (raw[0::3] < 0).any() or (raw[0::3] >= L).any()  # some particle crossed the box edge
    expect = raw.copy()
    expect[0::3] -= np.floor(expect[0::3] * (1.0 / L)) * L
    expect[1::3] -= np.floor(expect[1::3] * (1.0 / L)) * L
    assert np.array_equal(expect, wrapped)
