"""Verlet neighbour lists (SPC2_NEIGH_VERLET) against the oracle and against the
linked-cell path.

The list must give the linked cells' neighbour sets at every evaluation, so:
  * pair counts (within rho, cohesion) equal the oracle's exactly, per call and
    summed over whole runs that cross several list rebuilds;
  * RHS values meet the same bars as the cell path (1e-12 FP64, 1e-5 FP32);
  * trajectories equal the cell path's up to the summation order (FP64).
"""
import math

import numpy as np
import pytest

import oracle as O
import spc2_langevin_integration_amd as spc2
from _parity import assert_rhs_terms

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp64f": 1e-12, "fp32k": 1e-5}
VORT = dict(velocity=0.025, mu_plus=100.0, mu_minus=100.0, xi_a=0.01, xi_r=0.005, kappa=20.0)


@pytest.fixture(scope="module")
def gpu():
    spc2.load()
    with spc2.Sim(spc2.params(n=16)) as s:  # fails loudly without a GPU
        s.init_random_state()


def periodic_diff(a, b, L):
    d = a - b
    d[0::3] -= np.round(d[0::3] / L) * L
    d[1::3] -= np.round(d[1::3] / L) * L
    return d


@pytest.mark.parametrize("prec", ["fp64", "fp64f", "fp32k"])
@pytest.mark.parametrize("kc", [0.5, 1.0])  # stage reach below / above skin / 2 (list from the stage)
def test_verlet_chimera_rhs_k(gpu, prec, kc):
    n, rho = 40000, 0.1
    L = math.sqrt(n / 4000.0)
    st = O.make_state(17, 1, n, L, rho)
    kp = O.uniform(170, 3 * n, -1.0, 1.0)
    kn = O.uniform(171, 3 * n, 5.0, 6.0)
    dt = 0.01  # reach |kc dt k| ~ 0.007 (kc = 0.5) or 0.014 (kc = 1) against skin / 2 = 0.01
    ref, rc, ts = O.rhs_k(O.chimera(n, L, rho, 0.3, 1.2, 1.4, 0.05, threads=8), st, kp, kc, dt, kn,
                          term_scale=True)
    p = spc2.params("chimera", prec, "verlet", n=n, L=L, rho=rho, v0=0.3, sigma=1.2, alpha=1.4, D=0.05)
    with spc2.Sim(p) as s:
        s.set_state(st)
        s.set_counting(True)
        out = s.eval_rhs(kp, kc, dt, kn)
        cnt = s.counts()
        builds, entries = s.verlet_stats()
    assert_rhs_terms(out, ref, ts, 0.3, TOL[prec], f"chimera verlet/{prec}", ref[2::3] == kn[2::3])
    assert np.array_equal(cnt[1:], rc[1:]), (cnt, rc)
    assert builds == (2 if kc == 1.0 else 1) and entries == cnt[0]


@pytest.mark.parametrize("prec", ["fp64", "fp64f", "fp32k"])
def test_verlet_vortex_rhs_both_forms(gpu, prec):
    n, rho, D = 40000, 0.05, 0.02
    L = math.sqrt(n / 4000.0)
    st = O.make_state(23, 1, n, L, rho)
    sysp = O.vortex(n, L, rho=rho, D=D, threads=8, **VORT)
    kp = O.uniform(230, 3 * n, -1.0, 1.0)
    kn = O.uniform(231, 3 * n, 5.0, 6.0)
    refk, rck, tsk = O.rhs_k(sysp, st, kp, 0.5, 0.01, kn, term_scale=True)
    rd, ra, rci = O.rhs_ito(sysp, st)
    _, _, tsi = O.rhs_k(sysp, st, np.zeros(3 * n), 0.0, 0.01, np.zeros(3 * n), term_scale=True)
    p = spc2.params("vortex", prec, "verlet", n=n, L=L, rho=rho, D=D, **VORT)
    with spc2.Sim(p) as s:
        s.set_state(st)
        s.set_counting(True)
        outk = s.eval_rhs(kp, 0.5, 0.01, kn)
        ck = s.counts()
        s.set_counting(True)
        gd, ga = s.eval_rhs_ito()
        ci = s.counts()
    tol = TOL[prec]
    assert_rhs_terms(outk, refk, tsk, 0.025, tol, f"vortex verlet k/{prec}", refk[2::3] == kn[2::3])
    assert_rhs_terms(gd, rd, tsi, 0.025, tol, f"vortex verlet ito/{prec}")
    for c in (0, 1):
        scale = np.abs(ra[c::2]).max()
        assert np.abs(ga[c::2] - ra[c::2]).max() <= tol * scale, f"vortex verlet ito/{prec} additional[{c}]"
    assert np.array_equal(ck[1:], rck[1:]) and np.array_equal(ci[1:], rci[1:])


def run(params, st, method, dt, steps, omega=None, chunks=(1,)):
    with spc2.Sim(params) as s:
        s.set_state(st)
        if omega is not None:
            s.set_omega(omega)
        s.set_counting(True)
        k = 0
        for c in chunks:
            s.step(method, dt, k, c)
            k += c
        s.step(method, dt, k, steps - k)
        out = s.get_state(), s.counts()
        stats = s.verlet_stats() if params.neighbour == spc2.NEIGH_VERLET else None
    return out[0], out[1], stats


@pytest.mark.parametrize("method", ["se", "rk4", "srk"])
def test_verlet_trajectory_matches_cells_across_rebuilds(gpu, method):
    """Fast particles (v0 dt = skin / 8) force a list rebuild every few steps;
    the run must see exactly the cell path's neighbour sets at every RHS."""
    n, rho, dt, steps = 30000, 0.1, 0.01, 40
    L = math.sqrt(n / 4000.0)
    v0 = 0.2 * rho / 8 / dt
    st = O.make_state(8, 0, n, L, rho)
    om = O.uniform(80, n, -0.5, 0.5)
    kw = dict(n=n, L=L, rho=rho, v0=v0, sigma=1.0, alpha=1.45, D=0.02, seed=11)
    a, ca, _ = run(spc2.params("chimera", "fp64", "cells", **kw), st, method, dt, steps, om, chunks=(3, 10))
    b, cb, (builds, _) = run(spc2.params("chimera", "fp64", "verlet", **kw), st, method, dt, steps, om,
                             chunks=(3, 10))
    assert np.array_equal(ca[1:], cb[1:]), (ca, cb)
    assert builds >= steps // 6
    assert np.abs(periodic_diff(a, b, L)).max() < 1e-9


def test_verlet_srk_noise_bound_keeps_neighbour_sets(gpu):
    """SRK with the built-in noise rebuilds the list on a bound of each step's
    displacement whose noise term (sqrt(2 D) v0 |dZ|, |dZ| bounded through
    Box-Muller's largest radius) here outweighs the drift v0 dt.  The
    neighbour sets must stay those of the linked cells at every RHS."""
    n, rho, dt, steps = 30000, 0.1, 0.01, 60
    L = math.sqrt(n / 4000.0)
    v0 = 0.2 * rho / 8 / dt
    st = O.make_state(14, 0, n, L, rho)
    om = O.uniform(81, n, -0.5, 0.5)
    kw = dict(n=n, L=L, rho=rho, v0=v0, sigma=1.0, alpha=1.45, D=5.0, seed=17)
    a, ca, _ = run(spc2.params("chimera", "fp64", "cells", **kw), st, "srk", dt, steps, om, chunks=(7, 20))
    b, cb, (builds, _) = run(spc2.params("chimera", "fp64", "verlet", **kw), st, "srk", dt, steps, om,
                             chunks=(7, 20))
    assert np.array_equal(ca[1:], cb[1:]), (ca, cb)
    # the bound per step is about 3.2 v0 dt here (noise term 2.2 v0 dt): a
    # rebuild at least every third step
    assert builds >= steps // 3
    assert np.abs(periodic_diff(a, b, L)).max() < 1e-9


def test_verlet_vortex_rk4_fp32_matches_fp64_cells(gpu):
    """The bench workload at a reduced N: vortex RK4 with Verlet lists, FP64
    against FP64 linked cells (the same neighbour sets at every RHS), and
    FP32-compensated against FP64 (phases and order parameter after 50 steps;
    the two trajectories drift apart by rounding, so their neighbour sets may
    differ at the cutoffs)."""
    n, rho, dt, steps = 65536, 0.05, 0.01, 50
    L = math.sqrt(n / 4000.0)
    st, _ = O.init_state(42, n, L, L, 0.0)
    kw = dict(n=n, L=L, rho=rho, D=0.0, **VORT)
    a, ca, _ = run(spc2.params("vortex", "fp64", "cells", **kw), st, "rk4", dt, steps)
    v, cv, _ = run(spc2.params("vortex", "fp64", "verlet", **kw), st, "rk4", dt, steps)
    assert np.array_equal(ca[1:], cv[1:]), (ca, cv)
    assert np.abs(periodic_diff(a, v, L)).max() < 1e-9
    b, cb, (builds, entries) = run(spc2.params("vortex", "fp32k", "verlet", **kw), st, "rk4", dt, steps)
    c, cc, _ = run(spc2.params("vortex", "fp32k", "cells", **kw), st, "rk4", dt, steps)
    assert builds >= 2 and entries > 0
    Ra = np.abs(np.exp(1j * a[2::3]).mean())
    rms = {}
    for name, g in (("verlet", b), ("cells", c)):
        dphi = np.abs(a[2::3] - g[2::3])
        rms[name] = np.sqrt((dphi ** 2).mean())
        dR = abs(np.abs(np.exp(1j * g[2::3]).mean()) - Ra)
        print(f"fp32k {name} vs fp64 cells after {steps} steps: max |dphi| {dphi.max():.3e}, "
              f"rms {rms[name]:.3e}, |dR| {dR:.3e}")
        assert dR < 1e-4
    # the list changes the FP32 summation order only: the same error level as the cells
    assert rms["verlet"] < 1e-3 and rms["verlet"] < 3 * rms["cells"] + 1e-6


def test_verlet_rebuild_schedule(gpu):
    """RK4 rebuilds exactly when the next step could carry a particle beyond
    skin / 2 of its build position: after m steps with (m + 1) v0 dt > skin / 2."""
    n, rho, dt = 20000, 0.1, 0.01
    L = math.sqrt(n / 4000.0)
    skin = 0.2 * rho
    v0 = skin / 2 / dt / 4.5  # 4 steps fit, the 5th would not
    p = spc2.params("chimera", "fp64", "verlet", n=n, L=L, rho=rho, v0=v0, sigma=1.0, alpha=1.45, D=0.0)
    st = O.make_state(2, 0, n, L, rho)
    with spc2.Sim(p) as s:
        s.set_state(st)  # build 1
        s.step("rk4", dt, 0, 4)
        assert s.verlet_stats()[0] == 1  # step m (0-based) fits while (m + 1) v0 dt <= skin / 2
        s.step("rk4", dt, 4, 1)
        assert s.verlet_stats()[0] == 2  # 5 v0 dt > skin / 2: rebuilt before step 4
        s.step("rk4", dt, 5, 7)
        assert s.verlet_stats()[0] == 3  # and again before step 8


def test_verlet_rejects_too_few_cells(gpu):
    p = spc2.params("chimera", "fp64", "verlet", n=100, L=1.0, rho=0.3)
    with pytest.raises(spc2.Spc2Error, match="Verlet"):
        spc2.Sim(p)


@pytest.mark.parametrize("prec", ["fp64", "fp32k"])
@pytest.mark.parametrize("method", ["rk4", "se", "srk"])
def test_verlet_fused_stage_tiles_equal_separate_stage_pass(gpu, prec, method):
    """Verlet SE and RK4 steps write the next stage's tiles in the pass-1
    epilogue, SRK steps in the pass-2 epilogue.  Interleaving a call that
    rebuilds the tiles from the state before every step (an Ito evaluation)
    must not change the trajectory by a single bit, across list rebuilds."""
    n, rho, dt, steps = 30000, 0.05, 0.01, 30
    L = math.sqrt(n / 4000.0)
    kw = dict(n=n, L=L, rho=rho, D=0.0 if method == "rk4" else 0.01, seed=4, verlet_skin=0.004, **VORT)
    st = O.make_state(12, 1, n, L, rho)
    p = spc2.params("vortex", prec, "verlet", **kw)
    with spc2.Sim(p) as s:
        s.set_state(st)
        s.step(method, dt, 0, steps)
        a, ba = s.get_state(), s.verlet_stats()[0]
    with spc2.Sim(p) as s:
        s.set_state(st)
        for k in range(steps):
            if k:  # between steps only: a call after the last one may rebuild the list once more
                s.eval_rhs_ito()
            s.step(method, dt, k, 1)
        b, bb = s.get_state(), s.verlet_stats()[0]
    assert ba == bb and ba >= 3
    assert np.array_equal(a, b), np.abs(a - b).max()


@pytest.mark.parametrize("system,prec", [("vortex", "fp32k"), ("vortex", "fp64"), ("chimera", "fp32k"),
                                         ("chimera", "fp64f")])
def test_verlet_rk4_combine_rebuilt_k_equals_stored_k(gpu, monkeypatch, system, prec):
    """Fused RK4 steps do not store k_1..k_3's x / y components; the combine
    rebuilds them as v0 (cos, sin) of the stage phases.  The trajectory must
    equal, bit for bit and across list rebuilds, the one whose combine reads
    them from memory (SPC2_RK4_STORE_KXY=1)."""
    n, rho, dt, steps = 30000, 0.05, 0.01, 30
    L = math.sqrt(n / 4000.0)
    if system == "vortex":
        kw = dict(n=n, L=L, rho=rho, D=0.0, seed=4, verlet_skin=0.004, **VORT)
    else:
        kw = dict(n=n, L=L, rho=rho, v0=0.02, sigma=1.0, alpha=1.45, D=0.0, seed=4, verlet_skin=0.004)
    st = O.make_state(12, 1, n, L, rho)
    om = O.uniform(81, n, -0.5, 0.5) if system == "chimera" else None
    p = spc2.params(system, prec, "verlet", **kw)
    out = []
    for store in ("0", "1"):
        with monkeypatch.context() as m:
            m.setenv("SPC2_RK4_STORE_KXY", store)
            with spc2.Sim(p) as s:
                s.set_state(st)
                if om is not None:
                    s.set_omega(om)
                s.step("rk4", dt, 0, 7)
                s.step("rk4", dt, 7, steps - 7)
                out.append((s.get_state(), s.verlet_stats()[0]))
    (a, ba), (b, bb) = out
    assert ba == bb and ba >= 3
    assert np.array_equal(a, b), np.abs(a - b).max()


@pytest.mark.parametrize("neighbour", ["cells", "verlet"])
def test_eval_rhs_between_rk4_steps_leaves_the_stepper_alone(gpu, neighbour):
    """A k-form evaluation with the caller's buffers between RK4 steps must not
    touch the stepper's k_1..k_4, whose phi components a particle without
    partners keeps from one step to the next (RungeKutta4StepperPtr.cpp:25-43
    holds them as stepper members).  Sparse system: about 45% of the particles
    have no partner within rho."""
    n, rho, dt = 4000, 0.05, 0.01
    L = math.sqrt(n / 100.0)
    st, om = O.init_state(21, n, L, L, 0.5)
    kw = dict(n=n, L=L, rho=rho, v0=0.02, sigma=1.0, alpha=1.45, D=0.0, seed=3)
    p = spc2.params("chimera", "fp64", neighbour, **kw)
    out = []
    for interleave in (False, True):
        with spc2.Sim(p) as s:
            s.set_state(st)
            s.set_omega(om)
            s.step("rk4", dt, 0, 3)
            if interleave:  # small stage reach: no list rebuild inside the call
                s.eval_rhs(O.uniform(22, 3 * n, -0.01, 0.01), 0.5, dt, O.uniform(23, 3 * n, 5.0, 6.0))
            s.step("rk4", dt, 3, 3)
            out.append(s.get_state())
    assert np.array_equal(out[0], out[1]), np.abs(out[0] - out[1]).max()


def test_verlet_ito_eval_with_a_rebuild_inside_keeps_in_out_arrays(gpu):
    """eval_rhs_ito's derivative and additional arrays are in/out.  When the call
    rebuilds the Verlet list (here: the previous k-form evaluation built it from a
    stage more than skin / 2 away), the particles are re-homed; the in/out values
    must still reach their own particles."""
    n, rho = 30000, 0.1
    L = math.sqrt(n / 4000.0)
    st = O.make_state(31, 1, n, L, rho)
    kw = dict(n=n, L=L, rho=rho, v0=0.3, sigma=1.2, alpha=1.4, D=0.05)
    p = spc2.params("chimera", "fp64", "verlet", **kw)
    d_in, a_in = O.uniform(32, 3 * n, 1.0, 2.0), O.uniform(33, 2 * n, -3.0, 3.0)
    with spc2.Sim(p) as s:
        s.set_state(st)
        s.eval_rhs(O.uniform(34, 3 * n, -1.0, 1.0), 1.0, 0.02, np.zeros(3 * n))  # reach above skin / 2
        b0 = s.verlet_stats()[0]
        d, a = s.eval_rhs_ito(d_in, a_in)
        assert s.verlet_stats()[0] == b0 + 1  # the call rebuilt the list
    with spc2.Sim(p) as s:
        s.set_state(st)
        d_ref, a_ref = s.eval_rhs_ito(d_in, a_in)
    assert np.abs(d - d_ref).max() <= 1e-12 * np.abs(d_ref).max()
    assert np.abs(a - a_ref).max() <= 1e-12 * np.abs(a_ref).max()


def run_layout(params, st, method, dt, steps, monkeypatch, lds, omega=None, sort=None):
    """run() with the FP32 pass kernels chosen: lds=False keeps the gather
    passes (SPC2_VERLET_LDS=0 at creation); sort = "0" / "1" forces the LDS
    layout's lane slots off / on (None: by list length)."""
    with monkeypatch.context() as m:
        if lds:
            m.delenv("SPC2_VERLET_LDS", raising=False)
        else:
            m.setenv("SPC2_VERLET_LDS", "0")
        if sort is None:
            m.delenv("SPC2_VERLET_SORT", raising=False)
        else:
            m.setenv("SPC2_VERLET_SORT", sort)
        with spc2.Sim(params) as s:
            s.set_state(st)
            if omega is not None:
                s.set_omega(omega)
            s.set_counting(True)
            s.step(method, dt, 0, 3)
            layout = s.verlet_layout()
            s.step(method, dt, 3, steps - 3)
            return s.get_state(), s.counts(), s.verlet_stats(), layout


@pytest.mark.parametrize("prec", ["fp32k", "fp64", "fp64f"])
@pytest.mark.parametrize("sort", ["0", "1"])
@pytest.mark.parametrize("system,method", [("chimera", "srk"), ("vortex", "rk4"), ("vortex", "se")])
def test_verlet_lds_passes_equal_gather_passes(gpu, monkeypatch, system, method, sort, prec):
    """The LDS-staged passes (FP32 tiles, or the FP64 forms' positions and
    phases) read the same tiles from each workgroup's staged halo, with the same
    arithmetic and order as the gather passes, whichever lane takes a particle:
    trajectories across several rebuilds are bit-identical."""
    n, dt, steps = 30000, 0.01, 24
    L = math.sqrt(n / 4000.0)
    if system == "chimera":
        # the factored form stages 40 B per halo particle: at rho = 0.1 (62
        # particles per build cell) a workgroup spanning two rows overflows its
        # LDS capacity and the build keeps the gather layout
        rho = 0.08 if prec == "fp64f" else 0.1
        kw = dict(n=n, L=L, rho=rho, v0=0.2 * rho / 8 / dt, sigma=1.0, alpha=1.45, D=0.02, seed=5)
        om = O.uniform(81, n, -0.5, 0.5)
    else:
        rho = 0.05
        kw = dict(n=n, L=L, rho=rho, D=0.01 if method == "se" else 0.0, seed=6, verlet_skin=0.004, **VORT)
        om = None
    st = O.make_state(9, 0 if system == "chimera" else 1, n, L, rho)
    p = spc2.params(system, prec, "verlet", **kw)
    a, ca, sa, la = run_layout(p, st, method, dt, steps, monkeypatch, True, om, sort)
    b, cb, sb, lb = run_layout(p, st, method, dt, steps, monkeypatch, False, om)
    assert la[0] == 1 and la[1] > 0, la
    assert lb == (0, 0), lb
    assert sa == sb and sa[0] >= 2
    assert np.array_equal(ca, cb), (ca, cb)
    assert np.array_equal(a, b), np.abs(a - b).max()


@pytest.mark.parametrize("prec", ["fp32k", "fp64", "fp64f"])
def test_verlet_lds_vortex_ito_equals_gather(gpu, monkeypatch, prec):
    """Vortex pass 2 (Ito form) from the staged halo equals the gather pass."""
    n, rho, D = 40000, 0.05, 0.02
    L = math.sqrt(n / 4000.0)
    st = O.make_state(23, 1, n, L, rho)
    p = spc2.params("vortex", prec, "verlet", n=n, L=L, rho=rho, D=D, **VORT)
    out = {}
    for lds in (True, False):
        with monkeypatch.context() as m:
            if lds:
                m.delenv("SPC2_VERLET_LDS", raising=False)
            else:
                m.setenv("SPC2_VERLET_LDS", "0")
            with spc2.Sim(p) as s:
                s.set_state(st)
                out[lds] = s.eval_rhs_ito() + (s.verlet_layout(),)
    assert out[True][2][0] == 1 and out[False][2] == (0, 0)
    assert np.array_equal(out[True][0], out[False][0]) and np.array_equal(out[True][1], out[False][1])


def test_verlet_lds_clustered_state_keeps_gather(gpu):
    """A dense cluster whose workgroup halo exceeds the LDS capacity keeps the
    gather passes (FP32 form), with the oracle's pair counts and RHS."""
    n, rho = 8000, 0.1
    L = math.sqrt(30000 / 4000.0)
    st = O.make_state(31, 0, n, L, rho)
    u = O.uniform(310, 2 * n, 0.0, 0.6)
    st[0::3] = 1.0 + u[0::2]
    st[1::3] = 1.0 + u[1::2]
    kp = O.uniform(311, 3 * n, -1.0, 1.0)
    kn = O.uniform(312, 3 * n, 5.0, 6.0)
    ref, rc, ts = O.rhs_k(O.chimera(n, L, rho, 0.3, 1.2, 1.4, 0.05, threads=8), st, kp, 0.5, 0.01, kn,
                          term_scale=True)
    p = spc2.params("chimera", "fp32k", "verlet", n=n, L=L, rho=rho, v0=0.3, sigma=1.2, alpha=1.4, D=0.05)
    with spc2.Sim(p) as s:
        s.set_state(st)
        assert s.verlet_layout() == (0, 0)
        s.set_counting(True)
        out = s.eval_rhs(kp, 0.5, 0.01, kn)
        cnt = s.counts()
    assert_rhs_terms(out, ref, ts, 0.3, TOL["fp32k"], "clustered chimera verlet fp32k", ref[2::3] == kn[2::3])
    assert np.array_equal(cnt[1:], rc[1:]), (cnt, rc)


def test_verlet_fresh_handles_build_sane_lists(gpu):
    """Handles created one after another in a process (each reusing the
    previous one's freed device memory) build their first list from a cleared
    cell histogram: the list length per particle stays at the density's
    expectation for every handle."""
    n, rho = 262144, 0.05
    L = math.sqrt(n / 4000.0)
    skin = 0.14 * rho
    expect = 4000.0 * math.pi * (rho + skin) ** 2
    for k in range(4):
        p = spc2.params("vortex", "fp32k", "verlet", n=n, L=L, rho=rho, D=0.0, verlet_skin=skin, seed=40 + k,
                        **VORT)
        with spc2.Sim(p) as s:
            s.init_random_state(0.0)
            builds, entries = s.verlet_stats()
            s.step("rk4", 0.01, 0, 2)
        assert builds == 1
        assert 0.9 * expect < entries / n < 1.1 * expect, (k, entries / n, expect)


def test_chimera_pass2_from_rho_bits_equals_recomputed(gpu, monkeypatch):
    """Chimera pass 2 on the LDS layout reads pass 1's within-rho decisions
    (a bit per list entry) instead of forming each separation again: SRK
    trajectories across rebuilds and the Ito evaluation are bit-identical to
    the pass 2 that recomputes them (SPC2_VERLET_RHO_BITS=0)."""
    n, rho, dt, steps = 30000, 0.1, 0.01, 24
    L = math.sqrt(n / 4000.0)
    kw = dict(n=n, L=L, rho=rho, v0=0.2 * rho / 8 / dt, sigma=1.0, alpha=1.45, D=0.02, seed=5)
    st = O.make_state(19, 0, n, L, rho)
    om = O.uniform(82, n, -0.5, 0.5)
    p = spc2.params("chimera", "fp32k", "verlet", **kw)
    out = {}
    for bits in ("1", "0"):
        with monkeypatch.context() as m:
            m.setenv("SPC2_VERLET_RHO_BITS", bits)
            m.delenv("SPC2_VERLET_LDS", raising=False)
            with spc2.Sim(p) as s:
                s.set_state(st)
                s.set_omega(om)
                ito = s.eval_rhs_ito()
                assert s.verlet_layout()[0] == 1
                s.step("srk", dt, 0, steps)
                out[bits] = (ito, s.get_state(), s.verlet_stats())
    (ia, sa, va), (ib, sb, vb) = out["1"], out["0"]
    assert va == vb and va[0] >= 2
    assert np.array_equal(ia[0], ib[0]) and np.array_equal(ia[1], ib[1])
    assert np.array_equal(sa, sb), np.abs(sa - sb).max()


@pytest.mark.parametrize("prec", ["fp64", "fp64f", "fp32k"])
@pytest.mark.parametrize("init", [0, 1])
def test_verlet_chimera_rhs_ito(gpu, prec, init):
    """The chimera Ito evaluation on a Verlet list (ChimeraSystemPtr.cpp:448-639):
    pass 1 and pass 2 from the staged halo (the FP32 pass 2 from pass 1's
    within-rho bits) against the oracle's rhs_ito: the derivative at the
    per-particle bars, each additional output against its largest reference
    value, and exact pair counts."""
    n, v0, D = 30000, 0.3, 0.05
    rho = 0.08 if prec == "fp64f" else 0.1  # the factored form's 40-byte pass-2 halo fits the LDS at 0.08
    L = math.sqrt(n / 4000.0)
    st = O.make_state(29, 0, n, L, rho)
    d0 = O.uniform(290, 3 * n, -1.0, 1.0) if init else np.zeros(3 * n)
    a0 = O.uniform(291, 2 * n, -1.0, 1.0) if init else np.zeros(2 * n)
    sysp = O.chimera(n, L, rho, v0, 1.2, 1.4, D, threads=8)
    rd, ra, rc = O.rhs_ito(sysp, st, d0, a0)
    _, _, ts = O.rhs_k(sysp, st, np.zeros(3 * n), 0.0, 0.0, d0, term_scale=True)
    p = spc2.params("chimera", prec, "verlet", n=n, L=L, rho=rho, v0=v0, sigma=1.2, alpha=1.4, D=D)
    with spc2.Sim(p) as s:
        s.set_state(st)
        s.set_counting(True)
        gd, ga = s.eval_rhs_ito(d0, a0)
        cnt = s.counts()
        assert s.verlet_layout()[0] == 1
    tol = TOL[prec]
    assert_rhs_terms(gd, rd, ts, v0, tol, f"chimera verlet ito/{prec}", rd[2::3] == d0[2::3])
    for c in (0, 1):
        scale = np.abs(ra[c::2]).max()
        err = np.abs(ga[c::2] - ra[c::2]).max()
        assert err <= tol * scale, f"chimera verlet ito/{prec} additional[{c}]: {err:.3e} > {tol} x {scale:.3e}"
    assert np.array_equal(cnt[1:], rc[1:]), (cnt, rc)


@pytest.mark.parametrize("prec", ["fp32k", "fp64"])
def test_verlet_chimera_srk_steps_match_oracle(gpu, prec):
    """The chimera SRK workload's method on a Verlet list (skin 0.2 rho, dt =
    0.005), each step from the oracle's state with the same Philox normals,
    against the oracle's step_srk at the single-step bars; the handle keeps its
    list across the steps, and pass 2 of the FP32 form reads pass 1's rho bits."""
    from _parity import assert_state
    n, rho, dt, steps = 30000, 0.1, 0.005, 6
    L = math.sqrt(n / 4000.0)
    st, om = O.init_state(7, n, L, L, 0.1)
    sysp = O.chimera(n, L, rho, 0.01, 1.0, 1.45, 0.01, threads=8)
    p = spc2.params("chimera", prec, "verlet", n=n, L=L, rho=rho, v0=0.01, sigma=1.0, alpha=1.45, D=0.01, seed=7)
    cur = st
    worst = 0.0
    with spc2.Sim(p) as s:
        s.set_omega(om)
        for k in range(steps):
            s.set_state(cur)
            s.step("srk", dt, k, 1)
            ref = O.step_srk(sysp, cur, dt, O.philox_normals(7, k, 0, n, 2), omega=om)
            worst = max(worst, assert_state(s.get_state(), ref, cur, L, TOL[prec], f"chimera verlet srk/{prec} step {k}"))
            cur = ref
        assert s.verlet_layout()[0] == 1
    print(f"chimera verlet srk/{prec}: largest error / bar {worst:.3e}")
