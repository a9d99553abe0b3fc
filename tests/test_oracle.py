"""Pins the CPU oracle (oracle/spc2_oracle.c) before it is trusted as the checker.

1. Bit-exact against golden vectors written by the reference's own classes
   (tests/golden/golden.json, oracle/ref_harness.cpp): cell lists, min-image,
   wraps, both chimera RHS forms on the linked-cell and all-pairs paths, SE / SRK
   / RK4 steps and the mt19937 + std::normal_distribution noise they consume.
2. When the reference build is present (oracle/_ref), the same comparison on
   fresh seeds.
   The MPI shared-memory classes (VortexArraysSystemPtr, ChimeraSystemPtr and
   the *StepperPtr steppers, oracle/ref_harness_mpi.cpp) the same way: the
   vortex-arrays RHS in both forms on the linked-cell and all-pairs paths, and
   the SE / SRK / RK4 Ptr steps of both systems, with their outputs on 2 and 4
   ranks equal to 1 rank's.
3. Closed forms, invariants and brute force on top: W(r) values, two-body sums,
   the cohesion atan2 identity, Philox KATs, the normal transform, the order
   parameter.
"""
import hashlib
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O

N, L = 8192, 64.0


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def chim_args(args):
    v0, sigma, rho, alpha, D = (float(x) for x in args)
    return O.chimera(N, L, rho, v0=v0, sigma=sigma, alpha=alpha, D=D)


def pbc_inputs(seed, m=4096):
    pts = O.uniform(seed, 4 * m, -0.0, L)
    crafted = [[0.0, 0.0, 32.0, 32.0], [32.0, 32.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0], [1.0, 2.0, 33.0, -30.0],
               [0.0, 0.0, 63.999999999, 0.5], [10.0, 10.0, 10.0 + 38.4, 10.0 - 38.4], [63.0, 1.0, 0.5, 63.5]]
    for c, row in enumerate(crafted):
        pts[4 * c:4 * c + 4] = row
    w = O.uniform(seed + 1, 2 * m, -2.0 * L, 3.0 * L)
    w[:8] = [-0.0, 0.0, 64.0, -64.0, 63.99999999999999, -1e-300, 128.0, -1e-17]
    return pts, w


def run_golden_case(name, case, args):
    """Recomputes a golden case with the oracle; returns {output: array}."""
    a = args
    if case == "cells":
        seed, mode, rho = int(a[0]), int(a[1]), float(a[2])
        s = O.chimera(N, L, rho)
        st = O.make_state(seed, mode, N, L, rho)
        co, cs, so, _ = O.cell_list(s, st)
        return {"cell_start": cs, "sorted": so, "cell_of": co}
    if case == "pbc":
        pts, w = pbc_inputs(int(a[0]))
        return {"dxy": O.min_image(L, pts), "wrap_out": O.wrap_points(L, w)}
    if case == "rhs_k":
        seed, mode, kc, dt = int(a[0]), int(a[1]), float(a[2]), float(a[3])
        s = chim_args(a[4:])
        st = O.make_state(seed, mode, N, L, s.rho)
        kp = O.uniform(seed + 101, 3 * N, -50.0, 50.0)
        kn = O.uniform(seed + 202, 3 * N, 5.0, 6.0)
        out, _ = O.rhs_k(s, st, kp, kc, dt, kn)
        return {"k_next": out}
    if case == "rhs_ito":
        seed, mode, init = int(a[0]), int(a[1]), int(a[2])
        s = chim_args(a[4:])
        st = O.make_state(seed, mode, N, L, s.rho)
        d0 = O.uniform(seed + 303, 3 * N, -1.0, 1.0) if init else None
        a0 = O.uniform(seed + 404, 2 * N, -1.0, 1.0) if init else None
        d, ad, _ = O.rhs_ito(s, st, d0, a0)
        return {"derivative": d, "additional": ad}
    if case in ("se", "srk", "rk4"):
        seed, mtseed, steps, dt = int(a[0]), int(a[1]), int(a[2]), float(a[3])
        s = chim_args(a[4:])
        mt = O.MT19937Normals(mtseed)
        st = O.make_state(seed, 0, N, L, s.rho)
        kbuf = np.zeros(12 * N)
        noise, first = [], None
        for k in range(steps):
            if case == "se":
                xi = mt.normals(N)
                noise.append(xi)
                st = O.step_se(s, st, dt, xi, variant=1)
            elif case == "srk":
                xi = mt.normals(2 * N)
                noise.append(xi)
                st = O.step_srk(s, st, dt, xi)
            else:
                st = O.step_rk4(s, st, dt, kbuf)
            if k == 0:
                first = st.copy()
        out = {"state": st, "state1": first}
        out["noise"] = np.concatenate(noise) if noise else np.zeros(0)
        return out
    if case in ("vrhs_k", "crhs_k"):
        seed, mode, kc, dt = int(a[0]), int(a[1]), float(a[2]), float(a[3])
        s = vort_args(a[4:]) if case[0] == "v" else chim_args(a[4:])
        st = O.make_state(seed, mode, N, L, s.rho)
        kp = O.uniform(seed + 101, 3 * N, -1.0, 1.0)
        kn = O.uniform(seed + 202, 3 * N, 5.0, 6.0)
        out, _ = O.rhs_k(s, st, kp, kc, dt, kn)
        return {"k_next": out}
    if case in ("vrhs_ito", "crhs_ito"):
        seed, mode, init = int(a[0]), int(a[1]), int(a[2])
        s = vort_args(a[4:]) if case[0] == "v" else chim_args(a[4:])
        st = O.make_state(seed, mode, N, L, s.rho)
        d0 = O.uniform(seed + 303, 3 * N, -1.0, 1.0) if init else None
        a0 = O.uniform(seed + 404, 2 * N, -1.0, 1.0) if init else None
        d, ad, _ = O.rhs_ito(s, st, d0, a0)
        return {"derivative": d, "additional": ad}
    if case in ("vse", "vsrk", "vrk4", "cse", "csrk", "crk4"):
        # the *StepperPtr steppers (StochasticEulerStepperPtr.hpp:32-76,
        # StochasticRungeKuttaStepperPtr.hpp:25-97, RungeKutta4StepperPtr.cpp:25-65)
        seed, mtseed, steps, dt = int(a[0]), int(a[1]), int(a[2]), float(a[3])
        s = vort_args(a[4:]) if case[0] == "v" else chim_args(a[4:])
        mt = O.MT19937Normals(mtseed)
        omega = None
        if case[1:] == "srk":  # the SRK stepper draws N(0, 0) frequencies at construction
            drawn = mt.normals(N) * 0.0 + 0.0
            if case == "csrk":  # only the chimera step adds them (:65 against :94)
                omega = drawn
        st = O.make_state(seed, 0, N, L, s.rho)
        kbuf = np.zeros(12 * N)
        noise, first = [], None
        for k in range(steps):
            if case[1:] == "se":
                xi = mt.normals(N)
                noise.append(xi)
                st = O.step_se(s, st, dt, xi, variant=0)
            elif case[1:] == "srk":
                xi = mt.normals(2 * N)
                noise.append(xi)
                st = O.step_srk(s, st, dt, xi, omega=omega)
            else:
                st = O.step_rk4(s, st, dt, kbuf)
            if k == 0:
                first = st.copy()
        out = {"state": st, "state1": first}
        out["noise"] = np.concatenate(noise) if noise else np.zeros(0)
        return out
    raise KeyError(case)


def vort_args(args):
    v, mp, mm, xa, xr, D, kappa, rho = (float(x) for x in args)
    return O.vortex(N, L, rho=rho, velocity=v, mu_plus=mp, mu_minus=mm, xi_a=xa, xi_r=xr, kappa=kappa, D=D,
                    threads=8)


def golden_case_names():
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")) as f:
        return list(json.load(f)["cases"])


@pytest.mark.parametrize("name", golden_case_names())
def test_oracle_matches_reference_golden(golden, name):
    case = golden["cases"][name]
    got = run_golden_case(name, case["case"], case["args"])
    for out, ref in case["outputs"].items():
        arr = got[out]
        assert arr.dtype.str == ref["dtype"] and arr.size == ref["size"], (name, out)
        samples = [arr[i].item() for i in ref["sample_idx"]]
        assert samples == ref["sample"], (name, out, samples[:4], ref["sample"][:4])
        assert sha(arr) == ref["sha256"], (name, out)


def test_reference_outputs_do_not_depend_on_the_rank_count(golden):
    """The reference's deterministic MPI cases gave the same bytes on 1, 2 and
    4 ranks (ThreadSharedMemory's contiguous slices, ThreadSharedMemory.cpp:12-22),
    so the oracle's single-process restatement stands for every rank count."""
    multi = [(n, c) for n, c in golden["cases"].items() if "ranks" in c]
    assert len(multi) >= 5
    for name, c in multi:
        for ranks, outs in c["ranks"].items():
            for out, h in outs.items():
                assert h == c["outputs"][out]["sha256"], (name, ranks, out)


HARNESS = O.REF_HARNESS
HARNESS_MPI = os.path.join(os.path.dirname(O.REF_HARNESS), "ref_harness_mpi")
MPIEXEC = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("case,args,outputs", [
    ("cells", ["77", "1", "2.5"], ("cell_start", "sorted", "cell_of")),
    ("rhs_k", ["78", "1", "0.5", "0.02", "0.3", "1.7", "2.0", "0.9", "0.05"], ("k_next",)),
    ("rhs_ito", ["79", "0", "1", "0.01", "0.7", "1.1", "3.0", "1.5", "0.2"], ("derivative", "additional")),
    ("srk", ["80", "5", "2", "0.02", "0.1", "1.0", "3.0", "1.45", "0.05"], ("noise", "state")),
    ("rk4", ["81", "0", "2", "0.02", "0.3", "1.0", "2.0", "1.45", "0.0"], ("state",)),
])
def test_oracle_matches_live_reference(case, args, outputs):
    with tempfile.TemporaryDirectory() as tmp:
        subprocess.run([HARNESS, case, tmp] + args, check=True)
        got = run_golden_case("live", case, args)
        for out in outputs:
            ref = np.fromfile(os.path.join(tmp, out + ".bin"), got[out].dtype)
            assert np.array_equal(got[out], ref), (case, out)


@pytest.mark.skipif(not (os.path.exists(HARNESS_MPI) and os.path.exists(MPIEXEC)),
                    reason="reference MPI build (oracle/_ref) or mpiexec not present")
@pytest.mark.parametrize("ranks,case,args,outputs", [
    (1, "vrhs_k", ["91", "1", "0.5", "0.02", "0.5", "90.0", "110.0", "0.5", "0.3", "0.02", "15.0", "2.5"],
     ("k_next",)),
    (2, "vrhs_ito", ["92", "0", "1", "0.01", "0.8", "100.0", "100.0", "0.4472", "0.2236", "0.05", "20.0", "2.236"],
     ("derivative", "additional")),
    (1, "vsrk", ["93", "6", "2", "0.02", "1.0", "100.0", "100.0", "0.4472", "0.2236", "0.05", "20.0", "2.236"],
     ("noise", "state")),
    (4, "vrk4", ["94", "0", "2", "0.01", "1.0", "100.0", "100.0", "0.4472", "0.2236", "0.0", "20.0", "2.236"],
     ("state",)),
    (1, "csrk", ["95", "7", "2", "0.01", "0.3", "1.0", "2.0", "1.45", "0.05"], ("noise", "state")),
])
def test_oracle_matches_live_mpi_reference(ranks, case, args, outputs):
    with tempfile.TemporaryDirectory() as tmp:
        subprocess.run([MPIEXEC, "-n", str(ranks), HARNESS_MPI, case, tmp] + args, check=True)
        got = run_golden_case("live", case, args)
        for out in outputs:
            ref = np.fromfile(os.path.join(tmp, out + ".bin"), got[out].dtype)
            assert np.array_equal(got[out], ref), (case, ranks, out)


# ---- closed forms and invariants ----------------------------------------------

def test_min_image_closed_forms():
    Lb = 10.0
    pts = np.array([0, 0, 3.0, -2.0,        # |dx| < L/2: unchanged
                    0, 0, 6.0, 0.0,         # dx = 0.6 L -> -0.4 L
                    0, 0, 5.0, 0.0,         # dx = L/2 -> -L/2 (truncation toward zero)
                    6.0, 0, 0.0, 0.0,       # dx = -0.6 L -> +0.4 L
                    0, 0, 4.999, -4.999], np.float64)
    d = O.min_image(Lb, pts).reshape(-1, 2)
    assert d[0].tolist() == [3.0, -2.0]
    assert d[1, 0] == pytest.approx(-4.0, abs=1e-15)
    assert d[2, 0] == -5.0
    assert d[3, 0] == pytest.approx(4.0, abs=1e-15)
    assert d[4].tolist() == [4.999, -4.999]


def test_wrap_closed_forms():
    w = O.wrap_points(10.0, np.array([-1.0, 10.0, 25.0, -0.0, 3.5, -20.0]))
    assert w.tolist() == [9.0, 0.0, 5.0, 0.0, 3.5, 0.0]


def test_cell_list_invariants():
    n, Lb, rho = 5000, 5.0, 0.4
    s = O.chimera(n, Lb, rho)
    st = O.make_state(3, 1, n, Lb, rho)
    co, cs, so, _ = O.cell_list(s, st)
    nc = O.num_cells(Lb, rho) ** 2
    assert cs[0] == 0 and cs[-1] == n and np.all(np.diff(cs) >= 0)
    assert np.array_equal(np.sort(so), np.arange(n))       # a permutation
    for c in range(nc):                                    # members ascend, belong to c
        mem = so[cs[c]:cs[c + 1]]
        assert np.all(np.diff(mem) > 0) and np.all(co[mem] == c)
    # brute force: the cell of each wrapped position
    x = np.mod(st[0::3], Lb)
    y = np.mod(st[1::3], Lb)
    ncx = O.num_cells(Lb, rho)
    assert np.array_equal(co, (y / Lb * ncx).astype(int) * ncx + (x / Lb * ncx).astype(int))


def brute_chimera(state, Lb, rho, v0, sigma, alpha):
    x, y, p = state[0::3], state[1::3], state[2::3]
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    dx -= np.trunc(dx * 2.0 / Lb) * Lb
    dy -= np.trunc(dy * 2.0 / Lb) * Lb
    m = (dx * dx + dy * dy <= rho * rho)
    np.fill_diagonal(m, False)
    n = m.sum(1)
    A = (np.sin(p[None, :] - p[:, None] - alpha) * m).sum(1)
    out = np.zeros_like(state)
    out[0::3] = v0 * np.cos(p)
    out[1::3] = v0 * np.sin(p)
    out[2::3] = np.where(n > 0, sigma * A / np.maximum(n, 1), 0.0)
    return out, n


@pytest.mark.parametrize("rho", [0.3, 0.45, 2.0])  # linked cells and all pairs (rho > L/4)
def test_chimera_rhs_brute_force(rho):
    n, Lb = 600, 3.0
    s = O.chimera(n, Lb, rho, v0=0.7, sigma=1.3, alpha=0.4)
    st = O.make_state(11, 0, n, Lb, rho)
    out, cnt = O.rhs_k(s, st, np.zeros(3 * n), 0.0, 0.01, np.zeros(3 * n))
    ref, nn = brute_chimera(st, Lb, rho, 0.7, 1.3, 0.4)
    assert cnt[1] == nn.sum()
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-13)


def test_chimera_special_cases():
    n, Lb, rho = 2000, 4.0, 0.25
    s = O.chimera(n, Lb, rho, v0=1.0, sigma=1.7, alpha=1.45)
    st = O.make_state(5, 0, n, Lb, rho)
    st[2::3] = 0.3                                     # all phases equal
    out, _ = O.rhs_k(s, st, np.zeros(3 * n), 0.0, 0.01, np.full(3 * n, 9.0))
    _, nn = brute_chimera(st, Lb, rho, 1.0, 1.7, 1.45)
    has = nn > 0
    np.testing.assert_allclose(out[2::3][has], 1.7 * math.sin(-1.45), rtol=1e-14)
    assert np.all(out[2::3][~has] == 9.0)              # no partner: entry left as is
    s0 = O.chimera(n, Lb, rho, sigma=0.0)
    out0, _ = O.rhs_k(s0, O.make_state(6, 0, n, Lb, rho), np.zeros(3 * n), 0.0, 0.01, np.zeros(3 * n))
    assert np.all(out0[2::3] == 0.0)
    # phase-shift invariance
    st2 = O.make_state(7, 0, n, Lb, rho)
    a1, _ = O.rhs_k(s, st2, np.zeros(3 * n), 0.0, 0.01, np.zeros(3 * n))
    st3 = st2.copy()
    st3[2::3] += 0.8
    a2, _ = O.rhs_k(s, st3, np.zeros(3 * n), 0.0, 0.01, np.zeros(3 * n))
    np.testing.assert_allclose(a1[2::3], a2[2::3], atol=1e-12)


def test_chimera_two_body_and_ito_identity():
    Lb = 4.0
    st = np.array([1.0, 1.0, 0.2, 1.05, 1.02, 1.1] + [3.0, 3.0, 0.0], np.float64)
    s = O.chimera(3, Lb, 0.25, v0=0.5, sigma=2.0, alpha=0.3, D=0.1)
    d, a, _ = O.rhs_ito(s, st)
    a01 = 2.0 * math.sin(1.1 - 0.2 - 0.3)
    a10 = 2.0 * math.sin(0.2 - 1.1 - 0.3)
    assert d[2] == pytest.approx(a01, rel=1e-15) and d[5] == pytest.approx(a10, rel=1e-15)
    assert d[8] == 0.0                                 # isolated
    c = math.cos(1.1 - 0.2 - 0.3)
    assert a[0] == pytest.approx(-math.sqrt(0.2) * c, rel=1e-14)
    assert a[1] == pytest.approx((a10 - a01) * math.cos(0.2 - 1.1 - 0.3) * 0 + (a10 - a01) * c - 0.2 * a01, rel=1e-13)


# ---- vortex arrays: the reference build does not cover them -------------------

def vortex_w(r, xa, rho, mp, mm):
    if r < xa:
        return mp * (1.0 - r * r / (xa * xa))
    if r < rho:
        return -mm * 4.0 * (r - xa) * (rho - r) / (rho - xa) / (rho - xa)
    return 0.0


def test_vortex_two_body_closed_forms():
    xa, xr, rho, mp, mm, kap, v = 0.01, 0.005, 0.05, 100.0, 80.0, 20.0, 0.025
    Lb = 1.0
    for r, W in [(0.0, mp), (xa, 0.0), (0.5 * (xa + rho), -mm), (0.003, mp * (1 - 0.09))]:
        th, pi, pj = 0.7, 0.4, 1.9
        st = np.array([0.5, 0.5, pi, 0.5 + r * math.cos(th), 0.5 + r * math.sin(th), pj], np.float64)
        s = O.vortex(2, Lb, rho=rho, velocity=v, mu_plus=mp, mu_minus=mm, xi_a=xa, xi_r=xr, kappa=kap)
        out, cnt = O.rhs_k(s, st, np.zeros(6), 0.0, 0.01, np.zeros(6))
        expect = W * math.sin(pj - pi)
        if r <= xr:
            ang = math.atan2(st[4] - 0.5, st[3] - 0.5) if r > 0 else 0.0
            expect += -kap * math.sin(ang - pi)
        assert out[2] == pytest.approx(expect, rel=1e-12, abs=1e-12), r
        assert out[0] == pytest.approx(v * math.cos(pi), rel=1e-15)


def brute_vortex(state, Lb, rho, xa, xr, mp, mm, kap, v):
    x, y, p = state[0::3], state[1::3], state[2::3]
    n = len(x)
    out = np.zeros_like(state)
    for i in range(n):
        dx = x - x[i]
        dy = y - y[i]
        dx -= np.trunc(dx * 2.0 / Lb) * Lb
        dy -= np.trunc(dy * 2.0 / Lb) * Lb
        dr = np.sqrt(dx * dx + dy * dy)
        dr[i] = np.inf
        w = np.array([vortex_w(r, xa, rho, mp, mm) for r in dr])
        al = dr <= rho
        co = dr <= xr
        out[3 * i] = v * math.cos(p[i])
        out[3 * i + 1] = v * math.sin(p[i])
        val = 0.0
        if al.sum():
            val = (w[al] * np.sin(p[al] - p[i])).sum() / al.sum()
        if co.sum():
            # sin(atan2(dy, dx) - phi_i) written as (dy cos phi_i - dx sin phi_i) / r
            val += (-kap * (dy[co] * math.cos(p[i]) - dx[co] * math.sin(p[i])) / dr[co]).sum() / co.sum()
        out[3 * i + 2] = val
    return out


@pytest.mark.parametrize("rho", [0.1, 0.3])  # linked cells and all pairs
def test_vortex_rhs_brute_force(rho):
    n, Lb = 700, 1.0
    xa, xr = 0.4 * rho, 0.2 * rho
    s = O.vortex(n, Lb, rho=rho, velocity=0.1, mu_plus=100, mu_minus=100, xi_a=xa, xi_r=xr, kappa=20)
    st = O.make_state(21, 0, n, Lb, rho)
    out, cnt = O.rhs_k(s, st, np.zeros(3 * n), 0.0, 0.01, np.zeros(3 * n))
    ref = brute_vortex(st, Lb, rho, xa, xr, 100, 100, 20, 0.1)
    np.testing.assert_allclose(out, ref, rtol=0, atol=2e-11)
    assert cnt[2] > 0  # cohesion exercised


def test_vortex_translation_invariance():
    n, Lb, rho = 1500, 1.0, 0.1
    s = O.vortex(n, Lb, rho=rho, xi_a=0.04, xi_r=0.02, velocity=0.1)
    st = O.make_state(4, 0, n, Lb, rho)
    a, _ = O.rhs_k(s, st, np.zeros(3 * n), 0.0, 0.01, np.zeros(3 * n))
    st2 = st.copy()
    st2[0::3] = np.mod(st2[0::3] + 0.25, Lb)       # shift by a whole number of cells
    st2[1::3] = np.mod(st2[1::3] + 0.5, Lb)
    b, _ = O.rhs_k(s, st2, np.zeros(3 * n), 0.0, 0.01, np.zeros(3 * n))
    np.testing.assert_allclose(a, b, atol=1e-10)


def _phi_drift(s, st):
    d, _, cnt = O.rhs_ito(s, st)
    return d[2::3], cnt


@pytest.mark.parametrize("rho", [0.2, 0.3])  # linked cells (rho <= L/4) and all pairs
def test_vortex_ito_terms_are_generator_derivatives(rho):
    """The vortex pass-2 terms are the Ito-Taylor terms of the pass-1 drift a_i
    with the neighbour counts held fixed:
        additional[i][0] = sqrt(2 D) da_i/dphi_i
        additional[i][1] = sum_j grad_j a_i . f_j + D sum_j d^2 a_i / dphi_j^2
    with f the pass-1 derivative (the drift of every coordinate).  Checked by
    central differences of the oracle's own pass 1, which the brute-force and
    closed-form tests above pin: a dropped term, a wrong factor (e.g. the 2 D
    that collects the diffusion of both phases of a pair) or a sign fails."""
    n, Lb, D = 90, 1.0, 0.05
    s = O.vortex(n, Lb, rho=rho, velocity=0.1, mu_plus=100.0, mu_minus=80.0, xi_a=0.45 * rho, xi_r=0.3 * rho,
                 kappa=20.0, D=D)
    st = O.make_state(5, 0, n, Lb, rho)
    d, add, cnt = O.rhs_ito(s, st)
    a, _ = _phi_drift(s, st)
    assert cnt[1] > 3 * n and cnt[2] > n // 4  # alignment and cohesion both exercised
    scale0 = np.abs(add[0::2]).max()
    scale1 = np.abs(add[1::2]).max()

    # L^1 a_i: first differences in phi_i
    h = 1e-6
    dphi = np.empty(n)
    for i in range(n):
        sp, sm = st.copy(), st.copy()
        sp[3 * i + 2] += h
        sm[3 * i + 2] -= h
        dphi[i] = (_phi_drift(s, sp)[0][i] - _phi_drift(s, sm)[0][i]) / (2 * h)
    np.testing.assert_allclose(add[0::2], math.sqrt(2 * D) * dphi, rtol=0, atol=1e-7 * scale0)

    # L^0 a_i: directional derivative along the drift + D x phase Laplacian
    eps = 1e-6
    ap, cp = _phi_drift(s, st + eps * d)
    am, cm = _phi_drift(s, st - eps * d)
    assert (cp == cnt).all() and (cm == cnt).all()  # neighbour sets unchanged
    directional = (ap - am) / (2 * eps)
    h2 = 1e-4
    lap = np.zeros(n)
    for j in range(n):
        sp, sm = st.copy(), st.copy()
        sp[3 * j + 2] += h2
        sm[3 * j + 2] -= h2
        lap += (_phi_drift(s, sp)[0] + _phi_drift(s, sm)[0] - 2 * a) / (h2 * h2)
    np.testing.assert_allclose(add[1::2], directional + D * lap, rtol=0, atol=1e-6 * scale1)


# ---- steppers: closed forms of the Ptr (active-build) operation order --------

def test_se_ptr_form_closed_form():
    n, Lb = 200, 5.0
    s = O.chimera(n, Lb, 0.5, v0=0.0, sigma=0.0, D=0.3)
    st = O.make_state(9, 0, n, Lb, 0.5)
    xi = O.uniform(3, n, -2.0, 2.0)
    om = O.uniform(4, n, -0.1, 0.1)
    dt = 0.01
    new = O.step_se(s, st, dt, xi, omega=om, variant=0)
    expect = st[2::3] + ((om + 0.0) * dt + math.sqrt(2 * 0.3) * (math.sqrt(dt) * xi))
    assert np.array_equal(new[2::3], expect)
    assert np.array_equal(new[0::3], st[0::3]) and np.array_equal(new[1::3], st[1::3])


def test_srk_deterministic_limit():
    # D = 0, sigma = 0: x += v0 cos(phi) dt - 0.5 (a_phi a_y) dt^2 with a_phi = 0
    n, Lb, dt = 100, 5.0, 0.05
    s = O.chimera(n, Lb, 0.5, v0=0.3, sigma=0.0, D=0.0)
    st = O.make_state(12, 0, n, Lb, 0.5)
    new = O.step_srk(s, st, dt, np.zeros(2 * n))
    x = np.mod(st[0::3] + 0.3 * np.cos(st[2::3]) * dt, Lb)
    np.testing.assert_allclose(new[0::3], x, atol=1e-14)
    assert np.array_equal(new[2::3], st[2::3])


def test_rk4_order_on_two_body_ode():
    # Two partners, v0 = 0, alpha = 0: the phase difference d = phi1 - phi0
    # obeys d' = -2 sigma sin d, solved by tan(d/2) = tan(d0/2) exp(-2 sigma t).
    # One RK4 step has local error O(dt^5): halving dt divides it by ~32.
    Lb, d0 = 4.0, 0.5
    s = O.chimera(2, Lb, 0.25, v0=0.0, sigma=1.0, alpha=0.0)
    st = np.array([1.0, 1.0, 0.0, 1.1, 1.0, d0], np.float64)
    errs = []
    for dt in (0.2, 0.1):
        out = O.step_rk4(s, st, dt, np.zeros(12 * 2))
        exact = 2 * math.atan(math.tan(d0 / 2) * math.exp(-2 * dt))
        errs.append(abs((out[5] - out[2]) - exact))
    assert 24.0 < errs[0] / errs[1] < 40.0, errs


# ---- order parameter -------------------------------------------------------------

def test_order_parameter_closed_forms():
    n = 1000
    st = np.zeros(3 * n)
    st[2::3] = 1.3
    R, psi = O.order_param(st)
    assert R == pytest.approx(1.0, abs=1e-13) and psi == pytest.approx(1.3, abs=1e-13)
    st[2::3] = 2 * np.pi * np.arange(n) / n
    R, _ = O.order_param(st)
    assert R < 1e-14
    Rf, _ = O.order_param_f32(st)
    assert Rf < 1e-5


# ---- noise -----------------------------------------------------------------------

def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 10 rounds
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_normal_transform_matches_libm_box_muller():
    n, seed, step = 4000, 123, 7
    z = O.philox_normals(seed, step, 0, n, 2).reshape(n, 2)
    ref = np.empty((n, 2))
    for i in range(n):
        w = O.philox([i, 0, step & 0xffffffff, step >> 32], [seed & 0xffffffff, seed >> 32])
        a = (w[1] << 32) | w[0]
        b = (w[3] << 32) | w[2]
        u1 = ((a >> 11) + 1) * 2.0 ** -53
        u2 = (b >> 11) * 2.0 ** -53
        r = math.sqrt(-2.0 * math.log(u1))
        ref[i] = (r * math.cos(2 * math.pi * u2), r * math.sin(2 * math.pi * u2))
    np.testing.assert_allclose(z, ref, rtol=0, atol=4e-15 * np.maximum(1, np.abs(ref)).max())


def test_mt19937_uniform_draws_known_answers():
    """std::uniform_real_distribution over std::mt19937, as the reference's
    InitializeRandomSystemState draws (SimulationEnginePtr.cpp:36-45):
    generate_canonical<double, 53> combines two 32-bit words as
    (w0 + w1 2^32) / 2^64.  The raw words come from numpy's legacy
    RandomState(seed), which seeds mt19937 as std::mt19937(seed) does; its
    10000th word at the default seed 5489 is the value the C++ standard
    requires of mt19937 (4123659995)."""
    w = np.random.RandomState(5489).randint(0, 2 ** 32, size=10000, dtype=np.uint32)
    assert int(w[-1]) == 4123659995
    for seed in (5489, 42, 2 ** 31 + 7):
        words = np.random.RandomState(seed).randint(0, 2 ** 32, size=600, dtype=np.uint32).astype(np.float64)
        canon = (words[0::2] + words[1::2] * 2.0 ** 32) / 2.0 ** 64
        lo, hi = -0.5, 3.25
        got = O.MT19937Normals(seed).uniform(300, lo, hi)
        assert np.array_equal(got, canon * (hi - lo) + lo)


def test_normal_transform_statistics():
    z = O.philox_normals(42, 0, 0, 200000, 2)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.01
    from scipy import stats
    assert stats.kstest(z, "norm").pvalue > 1e-3
