"""GPU parity: the gfx950 library (through its C ABI) against the oracle and the
reference's golden vectors.

Bars (written into each assertion; tests/_parity.py):
  * cell lists, neighbour sets and pair counts: bit-exact;
  * Philox noise and the Philox-keyed initial state: bit-exact against the host
    restatement;
  * FP64 k-form RHS: velocity components within 1e-12 x v0, each particle's
    phase derivative within 1e-12 x max(|a_ref|, its mean absolute term);
  * Ito additional terms: 1e-12 x the largest reference value of each output;
  * single steps: 1e-12 x the increment scale of each component (+ 2 ulp);
  * FP32-compensated form: the same with 1e-5;
  * stochastic runs: ensemble mean |R|(t) within the reference ensemble's
    confidence band (z <= 3 at every sampled time).
"""
import hashlib
import math

import numpy as np
import pytest

import oracle as O
import spc2_langevin_integration_amd as spc2
from _parity import assert_rhs_terms

pytestmark = pytest.mark.gpu

N8, L64 = 8192, 64.0
TOL = {"fp64": 1e-12, "fp64f": 1e-12, "fp32k": 1e-5}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def chim_params(n, L, rho, v0, sigma, alpha, D, precision="fp64", seed=42, neighbour="auto"):
    return spc2.params("chimera", precision, neighbour, n=n, L=L, rho=rho, v0=v0, sigma=sigma, alpha=alpha, D=D,
                       seed=seed)


def vort_params(n, L, rho=0.05, precision="fp64", D=0.0, neighbour="auto", **kw):
    d = dict(velocity=0.025, mu_plus=100.0, mu_minus=100.0, xi_a=0.01, xi_r=0.005, kappa=20.0)
    d.update(kw)
    return spc2.params("vortex", precision, neighbour, n=n, L=L, rho=rho, D=D, **d)


def orc_vortex(n, L, rho=0.05, D=0.0, threads=8, **kw):
    d = dict(velocity=0.025, mu_plus=100.0, mu_minus=100.0, xi_a=0.01, xi_r=0.005, kappa=20.0)
    d.update(kw)
    return O.vortex(n, L, rho=rho, D=D, threads=threads, **d)


def periodic_diff(a, b, L):
    d = a - b
    d[0::3] -= np.round(d[0::3] / L) * L
    d[1::3] -= np.round(d[1::3] / L) * L
    return d


def assert_state(gpu, ref, before, L, tol, what=""):
    """Step parity on the increments: |gpu - ref| <= tol x |increment| scale + 2 ulp of the state."""
    inc = periodic_diff(ref, before, L)
    d = np.abs(periodic_diff(gpu, ref, L))
    for c, name in ((0, "x"), (1, "y"), (2, "phi")):
        scale = max(np.abs(inc[c::3]).max(), 1e-300)
        ulp = 2 * np.spacing(np.abs(ref[c::3]).max())
        err = d[c::3].max()
        assert err <= tol * scale + ulp, f"{what}: {name} error {err:.3e} > {tol} x {scale:.3e} + {ulp:.1e}"


@pytest.fixture(scope="module")
def gpu():
    lib = spc2.load()
    with spc2.Sim(spc2.params(n=16)) as s:  # fails loudly without a GPU
        s.init_random_state()
    return lib


# ---- noise and initial state ---------------------------------------------------

@pytest.mark.parametrize("step", [0, 1, 12345, (1 << 33) + 5])
def test_noise_bit_identical_to_host(gpu, step):
    n = 5000
    with spc2.Sim(chim_params(n, 2.0, 0.1, 0.01, 1.0, 1.45, 0.01, seed=987654321123)) as s:
        assert np.array_equal(s.noise("se", step), O.philox_normals(987654321123, step, 0, n, 1))
        assert np.array_equal(s.noise("srk", step), O.philox_normals(987654321123, step, 0, n, 2))


def test_init_state_and_omega_bit_identical(gpu):
    n, L = 6000, 1.5
    with spc2.Sim(chim_params(n, L, 0.1, 0.0, 0.0, 1.45, 0.0, seed=7)) as s:
        s.init_random_state(0.1)
        st, om = O.init_state(7, n, L, L, 0.1)
        g = s.get_state()
        assert np.array_equal(g, st)
        # v0 = sigma = D = 0: one SE step adds exactly omega * dt to the phase
        s.step("se", 0.01)
        g2 = s.get_state()
        assert np.array_equal(g2[2::3], st[2::3] + ((om + 0.0) * 0.01 + 0.0))
        assert np.array_equal(g2[0::3], st[0::3])


# ---- cell lists ----------------------------------------------------------------

CELL_CASES = [
    # golden cases (reference hash) ...
    ("cells_rho4", 1, 0, 4.0), ("cells_rho1_outside", 2, 1, 1.0), ("cells_rho3.7_lattice", 3, 2, 3.7),
    ("cells_rho16_n4", 4, 1, 16.0), ("cells_rho0.5", 5, 0, 0.5),
]


@pytest.mark.parametrize("name,seed,mode,rho", CELL_CASES)
def test_cell_list_matches_reference_hash(gpu, golden, name, seed, mode, rho):
    st = O.make_state(seed, mode, N8, L64, rho)
    with spc2.Sim(chim_params(N8, L64, rho, 0.01, 1.0, 1.45, 0.01)) as s:
        s.set_state(st)
        co, cs, so = s.cell_list()
    ref = golden["cases"][name]["outputs"]
    assert sha(cs) == ref["cell_start"]["sha256"]
    assert sha(so) == ref["sorted"]["sha256"]
    assert sha(co) == ref["cell_of"]["sha256"]


@pytest.mark.parametrize("n,rho,mode", [
    (1000, 0.1, 0),          # the reference's CPU case: L = 1
    (262144, 0.1, 0),        # chimera SRK at density 4000: L = sqrt(N / 4000)
    (1048576, 0.05, 0),      # vortex arrays at N = 1M
    (1048576, 0.05, 1),      # the same size, positions outside the box
])
def test_cell_list_full_size_bit_exact(gpu, n, rho, mode):
    L = 1.0 if n == 1000 else math.sqrt(n / 4000.0)
    st = O.make_state(100 + n, mode, n, L, rho)
    with spc2.Sim(chim_params(n, L, rho, 0.01, 1.0, 1.45, 0.01)) as s:
        s.set_state(st)
        co, cs, so = s.cell_list()
    rco, rcs, rso, _ = O.cell_list(O.chimera(n, L, rho), st)
    assert np.array_equal(cs, rcs) and np.array_equal(so, rso) and np.array_equal(co, rco)


# ---- chimera RHS ----------------------------------------------------------------

RHS_K = ["rhs_k_rho4", "rhs_k_rho1_outside", "rhs_k_rho4_c0", "rhs_k_allpairs"]


@pytest.mark.parametrize("prec", ["fp64", "fp64f", "fp32k"])
@pytest.mark.parametrize("name", RHS_K)
def test_chimera_rhs_k(gpu, golden, name, prec):
    a = golden["cases"][name]["args"]
    seed, mode, kc, dt = int(a[0]), int(a[1]), float(a[2]), float(a[3])
    v0, sigma, rho, alpha, D = (float(x) for x in a[4:])
    st = O.make_state(seed, mode, N8, L64, rho)
    kp = O.uniform(seed + 101, 3 * N8, -50.0, 50.0)
    kn = O.uniform(seed + 202, 3 * N8, 5.0, 6.0)
    ref, rc, ts = O.rhs_k(O.chimera(N8, L64, rho, v0, sigma, alpha, D, threads=8), st, kp, kc, dt, kn,
                          term_scale=True)
    with spc2.Sim(chim_params(N8, L64, rho, v0, sigma, alpha, D, prec)) as s:
        s.set_state(st)
        s.set_counting(True)
        out = s.eval_rhs(kp, kc, dt, kn)
        cnt = s.counts()
    # the untouched phase entries (no partner) are exactly the sentinel inputs
    untouched = ref[2::3] == kn[2::3]
    r, plain = assert_rhs_terms(out, ref, ts, v0, TOL[prec], f"{name}/{prec}", untouched)
    print(f"{name}/{prec}: error / bar {r:.3e}, plain relative error {plain:.3e}")
    assert np.array_equal(cnt[:2], rc[:2]), (cnt, rc)   # identical neighbour sets, every precision


RHS_ITO = ["rhs_ito_rho4", "rhs_ito_rho1_inout", "rhs_ito_allpairs"]


@pytest.mark.parametrize("prec", ["fp64", "fp64f", "fp32k"])
@pytest.mark.parametrize("name", RHS_ITO)
def test_chimera_rhs_ito(gpu, golden, name, prec):
    a = golden["cases"][name]["args"]
    seed, mode, init = int(a[0]), int(a[1]), int(a[2])
    v0, sigma, rho, alpha, D = (float(x) for x in a[4:])
    st = O.make_state(seed, mode, N8, L64, rho)
    d0 = O.uniform(seed + 303, 3 * N8, -1.0, 1.0) if init else np.zeros(3 * N8)
    a0 = O.uniform(seed + 404, 2 * N8, -1.0, 1.0) if init else np.zeros(2 * N8)
    sysp = O.chimera(N8, L64, rho, v0, sigma, alpha, D, threads=8)
    rd, ra, _ = O.rhs_ito(sysp, st, d0, a0)
    _, _, ts = O.rhs_k(sysp, st, np.zeros(3 * N8), 0.0, 0.0, d0, term_scale=True)
    with spc2.Sim(chim_params(N8, L64, rho, v0, sigma, alpha, D, prec)) as s:
        s.set_state(st)
        gd, ga = s.eval_rhs_ito(d0, a0)
    tol = TOL[prec]
    assert_rhs_terms(gd, rd, ts, v0, tol, f"{name}/{prec} derivative", rd[2::3] == d0[2::3])
    # additional: each output against its largest reference value
    for c, what in ((0, "add0"), (1, "add1")):
        scale = np.abs(ra[c::2]).max()
        err = np.abs(ga[c::2] - ra[c::2]).max()
        assert err <= tol * scale, f"{name}/{prec} {what}: {err:.3e} > {tol} x {scale:.3e}"


# ---- vortex arrays RHS -----------------------------------------------------------

@pytest.mark.parametrize("prec", ["fp64", "fp64f", "fp32k"])
@pytest.mark.parametrize("n,mode,kc", [(40000, 0, 0.5), (40000, 1, 1.0), (300, 0, 0.5)])
def test_vortex_rhs_k(gpu, prec, n, mode, kc):
    rho = 0.05
    L = math.sqrt(n / 4000.0)   # n = 300: rho > L/4, all pairs
    st = O.make_state(31 + n, mode, n, L, rho)
    kp = O.uniform(77, 3 * n, -1.0, 1.0)
    kn = O.uniform(78, 3 * n, 5.0, 6.0)
    ref, rc, ts = O.rhs_k(orc_vortex(n, L), st, kp, kc, 0.01, kn, term_scale=True)
    with spc2.Sim(vort_params(n, L, precision=prec)) as s:
        s.set_state(st)
        s.set_counting(True)
        out = s.eval_rhs(kp, kc, 0.01, kn)
        cnt = s.counts()
    untouched = ref[2::3] == kn[2::3]
    r, plain = assert_rhs_terms(out, ref, ts, 0.025, TOL[prec], f"vortex/{prec}", untouched)
    print(f"vortex n={n}/{prec}: error / bar {r:.3e}, plain relative error {plain:.3e}")
    assert np.array_equal(cnt, rc), (cnt, rc)
    assert rc[2] > 0


@pytest.mark.parametrize("prec", ["fp64", "fp64f", "fp32k"])
@pytest.mark.parametrize("n,mode,init", [(40000, 0, 0), (40000, 1, 1), (300, 0, 1)])
def test_vortex_rhs_ito(gpu, prec, n, mode, init):
    """Both passes of the vortex Ito form; bars relative to the largest
    reference value of each output (the pass-2 sums cancel heavily)."""
    rho, D = 0.05, 0.02
    L = math.sqrt(n / 4000.0)   # n = 300: all pairs
    st = O.make_state(41 + n, mode, n, L, rho)
    d0 = O.uniform(90, 3 * n, -1.0, 1.0) if init else np.zeros(3 * n)
    a0 = O.uniform(91, 2 * n, -1.0, 1.0) if init else np.zeros(2 * n)
    rd, ra, rc = O.rhs_ito(orc_vortex(n, L, D=D), st, d0, a0)
    _, _, ts = O.rhs_k(orc_vortex(n, L, D=D), st, np.zeros(3 * n), 0.0, 0.01, d0, term_scale=True)
    with spc2.Sim(vort_params(n, L, precision=prec, D=D)) as s:
        s.set_state(st)
        s.set_counting(True)
        gd, ga = s.eval_rhs_ito(d0, a0)
        cnt = s.counts()
    tol = TOL[prec]
    assert_rhs_terms(gd, rd, ts, 0.025, tol, f"vortex ito/{prec} derivative", rd[2::3] == d0[2::3])
    for c, name in ((0, "add0"), (1, "add1")):
        scale = np.abs(ra[c::2]).max()
        err = np.abs(ga[c::2] - ra[c::2]).max()
        assert err <= tol * scale, f"vortex ito/{prec} {name}: {err:.3e} > {tol} x {scale:.3e}"
    assert np.array_equal(cnt[1:], rc[1:]) and rc[2] > 0, (cnt, rc)


@pytest.mark.parametrize("prec", ["fp64", "fp64f", "fp32k"])
def test_vortex_srk_steps_with_given_noise(gpu, prec):
    """StochasticRungeKuttaStepperPtr::DoStep on the vortex system
    (StochasticRungeKuttaStepperPtr.hpp:70-97): one step from the oracle's state
    each time, same normals."""
    n, rho, D, dt = 40000, 0.05, 0.01, 0.01
    L = math.sqrt(n / 4000.0)
    sysp = orc_vortex(n, L, D=D, threads=8)
    st = O.make_state(12, 0, n, L, rho)
    mt = O.MT19937Normals(2024)
    with spc2.Sim(vort_params(n, L, precision=prec, D=D)) as s:
        ref = st
        for k in range(3):
            before = ref
            xi = mt.normals(2 * n)
            ref = O.step_srk(sysp, before, dt, xi)
            s.set_state(before)
            s.step("srk", dt, step0=k, nsteps=1, noise=xi)
            assert_state(s.get_state(), ref, before, L, TOL[prec], f"vortex srk/{prec} step {k}")


def test_vortex_srk_philox_trajectory_matches_oracle(gpu):
    n, rho, D, dt, steps = 20000, 0.05, 0.01, 0.002, 5
    L = math.sqrt(n / 4000.0)
    sysp = orc_vortex(n, L, D=D, threads=8)
    st, _ = O.init_state(7, n, L, L, 0.0)
    ref = st
    for k in range(steps):
        ref = O.step_srk(sysp, ref, dt, O.philox_normals(7, k, 0, n, 2))
    with spc2.Sim(vort_params(n, L, D=D, seed=7)) as s:
        s.init_random_state(0.0)
        s.step("srk", dt, 0, steps)
        g = s.get_state()
    assert np.abs(periodic_diff(g, ref, L)).max() < 1e-9


@pytest.mark.parametrize("prec", ["fp64", "fp64f", "fp32k"])
@pytest.mark.parametrize("cap", ["0", "300"])
def test_lds_overflow_fallback_matches(gpu, prec, cap, monkeypatch):
    """Segments whose halo exceeds the LDS capacity read the same tiles from
    global memory; results must not depend on which segments were staged."""
    n, rho = 40000, 0.05
    L = math.sqrt(n / 4000.0)
    st = O.make_state(7, 1, n, L, rho)
    kp = O.uniform(70, 3 * n, -1.0, 1.0)
    kn = O.uniform(71, 3 * n, 5.0, 6.0)
    outs = []
    for c in (None, cap):
        if c is None:
            monkeypatch.delenv("SPC2_LDS_CAP", raising=False)
        else:
            monkeypatch.setenv("SPC2_LDS_CAP", c)
        with spc2.Sim(vort_params(n, L, precision=prec)) as s:
            s.set_state(st)
            outs.append(s.eval_rhs(kp, 0.5, 0.01, kn))
        with spc2.Sim(chim_params(n, L, 0.1, 0.3, 1.2, 1.4, 0.05, prec)) as s:
            s.set_state(st)
            outs.append(np.concatenate(s.eval_rhs_ito()))
    monkeypatch.delenv("SPC2_LDS_CAP", raising=False)
    assert np.array_equal(outs[0], outs[2]) and np.array_equal(outs[1], outs[3])


def clustered_rows_state(n, L, cell, seed=9):
    """Uniform particles except two cell rows: 80% of row 5's particles moved into
    row 10.  Row 10 then holds about 1.8x a row's mean, beyond the workgroups
    its row chunks are sized for (mean + 4 sigma), so the row's last workgroup
    takes the rest; row 5 keeps a few particles per cell, so 256 of its slots
    span more than 64 cells and its chunk reads its tiles from global memory."""
    st = O.make_state(seed, 0, n, L, 0.05)
    y = st[1::3].copy()
    row = np.floor(y / cell).astype(int)
    idx = np.flatnonzero(row == 5)
    move = idx[: int(0.8 * idx.size)]
    y[move] += 5 * cell
    st[1::3] = y
    return st


@pytest.mark.parametrize("prec", ["fp64", "fp32k"])
def test_cell_walk_overfull_and_sparse_rows(gpu, prec):
    """Cell-stencil sums with an overfull and a sparse cell row against the
    oracle: the row-chunk workgroups' overflow and global-memory paths."""
    n, rho = 160000, 0.05
    L = math.sqrt(n / 4000.0)
    nc = O.num_cells(L, rho)
    st = clustered_rows_state(n, L, L / nc)
    rows = np.bincount(np.floor(st[1::3] / (L / nc)).astype(int), minlength=nc)
    mean = n / nc
    assert rows[10] > mean + 4 * math.sqrt(mean) + 8 + 256 and rows[5] < 0.25 * mean, (rows[5], rows[10], mean)
    kp = O.uniform(72, 3 * n, -1.0, 1.0)
    kn = O.uniform(73, 3 * n, 5.0, 6.0)
    ref, rc, ts = O.rhs_k(orc_vortex(n, L), st, kp, 0.5, 0.01, kn, term_scale=True)
    with spc2.Sim(vort_params(n, L, precision=prec, neighbour="cells")) as s:
        s.set_state(st)
        s.set_counting(True)
        out = s.eval_rhs(kp, 0.5, 0.01, kn)
        cnt = s.counts()
    assert_rhs_terms(out, ref, ts, 0.025, TOL[prec], f"vortex clustered/{prec}", ref[2::3] == kn[2::3])
    assert np.array_equal(cnt, rc), (cnt, rc)
    # chimera (rho 0.1: 63 cells per row, so only the overfull row's path)
    v0 = 0.3
    sysc = O.chimera(n, L, 0.1, v0, 1.2, 1.4, 0.05, threads=8)
    d0, a0 = np.zeros(3 * n), np.zeros(2 * n)
    rd, ra, _ = O.rhs_ito(sysc, st, d0, a0)
    _, _, ts = O.rhs_k(sysc, st, np.zeros(3 * n), 0.0, 0.0, d0, term_scale=True)
    with spc2.Sim(chim_params(n, L, 0.1, v0, 1.2, 1.4, 0.05, prec, neighbour="cells")) as s:
        s.set_state(st)
        gd, ga = s.eval_rhs_ito(d0, a0)
    tol = TOL[prec]
    assert_rhs_terms(gd, rd, ts, v0, tol, f"chimera clustered/{prec} derivative", rd[2::3] == d0[2::3])
    for c, what in ((0, "add0"), (1, "add1")):
        scale = np.abs(ra[c::2]).max()
        err = np.abs(ga[c::2] - ra[c::2]).max()
        assert err <= tol * scale, f"chimera clustered/{prec} {what}: {err:.3e} > {tol} x {scale:.3e}"


# ---- steps with the reference's own noise ----------------------------------------

@pytest.mark.parametrize("prec", ["fp64", "fp64f", "fp32k"])
@pytest.mark.parametrize("method,name", [("se", "se_rho4"), ("se", "se_rho1"), ("srk", "srk_rho4"),
                                         ("srk", "srk_rho1"), ("rk4", "rk4_rho1"), ("rk4", "rk4_rho4")])
def test_chimera_steps_with_reference_noise(gpu, golden, method, name, prec):
    a = golden["cases"][name]["args"]
    seed, mtseed, steps, dt = int(a[0]), int(a[1]), int(a[2]), float(a[3])
    v0, sigma, rho, alpha, D = (float(x) for x in a[4:])
    sysp = O.chimera(N8, L64, rho, v0, sigma, alpha, D, threads=8)
    st = O.make_state(seed, 0, N8, L64, rho)
    mt = O.MT19937Normals(mtseed)
    kbuf = np.zeros(12 * N8)
    with spc2.Sim(chim_params(N8, L64, rho, v0, sigma, alpha, D, prec)) as s:
        s.set_state(st)
        ref = st
        for k in range(steps):
            before = ref
            if method == "se":
                xi = mt.normals(N8)
                ref = O.step_se(sysp, ref, dt, xi, variant=0)
            elif method == "srk":
                xi = mt.normals(2 * N8)
                ref = O.step_srk(sysp, ref, dt, xi)
            else:
                xi = None
                ref = O.step_rk4(sysp, ref, dt, kbuf)
            s.set_state(before)           # restart each step from the oracle's state
            s.step(method, dt, step0=k, nsteps=1, noise=xi)
            assert_state(s.get_state(), ref, before, L64, TOL[prec], f"{name}/{prec} step {k}")


@pytest.mark.parametrize("method", ["se", "srk", "rk4"])
@pytest.mark.parametrize("prec", ["fp64", "fp32k"])
def test_rehoming_is_invisible(gpu, method, prec, monkeypatch):
    """Particle arrays are moved to cell order every few steps; the moves must
    not change a single bit of the trajectory, the cell list or the k buffers."""
    n, rho, dt = 30000, 0.1, 0.01
    L = math.sqrt(n / 4000.0)
    st = O.make_state(9, 0, n, L, rho)
    outs = []
    for every in ("1", "1000000"):
        monkeypatch.setenv("SPC2_REHOME_EVERY", every)
        with spc2.Sim(chim_params(n, L, rho, 0.05, 1.0, 1.45, 0.05, prec, seed=3)) as s:
            s.set_state(st)
            s.step(method, dt, 0, 3)
            s.set_state(s.get_state())          # state round trip through the id order
            s.step(method, dt, 3, 3)
            co, cs, so = s.cell_list()
            outs.append((s.get_state(), co, cs, so, s.eval_rhs(np.zeros(3 * n), 0.0, dt, np.zeros(3 * n))))
    monkeypatch.delenv("SPC2_REHOME_EVERY", raising=False)
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    rco, rcs, rso, _ = O.cell_list(O.chimera(n, L, rho), outs[0][0])
    assert np.array_equal(outs[0][2], rcs) and np.array_equal(outs[0][3], rso) and np.array_equal(outs[0][1], rco)


@pytest.mark.parametrize("every", ["1", "16"])
def test_rk4_stale_phase_semantics(gpu, every, monkeypatch):
    # rho = 1 at density 2: isolated particles keep k_phi from earlier stages/steps
    monkeypatch.setenv("SPC2_REHOME_EVERY", every)
    n, L, rho, dt = N8, L64, 1.0, 0.05
    sysp = O.chimera(n, L, rho, 0.5, 1.0, 1.45, 0.0, threads=8)
    st = O.make_state(15, 0, n, L, rho)
    kbuf = np.zeros(12 * n)
    ref = st
    with spc2.Sim(chim_params(n, L, rho, 0.5, 1.0, 1.45, 0.0)) as s:
        s.set_state(st)
        for k in range(4):
            ref = O.step_rk4(sysp, ref, dt, kbuf)
        s.step("rk4", dt, 0, 4)
        g = s.get_state()
    assert np.abs(periodic_diff(g, ref, L)).max() < 1e-11


@pytest.mark.parametrize("method", ["se", "srk"])
def test_philox_step_equals_materialised_noise(gpu, method):
    n, L = 20000, math.sqrt(20000 / 4000.0)
    st = O.make_state(5, 0, n, L, 0.1)
    outs = []
    for use_noise in (False, True):
        with spc2.Sim(chim_params(n, L, 0.1, 0.01, 1.0, 1.45, 0.05, "fp32k", seed=99)) as s:
            s.set_state(st)
            for k in range(3):
                nz = s.noise(method, 10 + k) if use_noise else None
                s.step(method, 0.01, step0=10 + k, nsteps=1, noise=nz)
            outs.append(s.get_state())
    assert np.array_equal(outs[0], outs[1])


def test_srk_philox_trajectory_matches_oracle(gpu):
    n, L, rho, dt, steps = 30000, math.sqrt(30000 / 4000.0), 0.1, 0.005, 10
    sysp = O.chimera(n, L, rho, 0.01, 1.0, 1.45, 0.01, threads=8)
    st, om = O.init_state(42, n, L, L, 0.1)
    ref = st
    for k in range(steps):
        ref = O.step_srk(sysp, ref, dt, O.philox_normals(42, k, 0, n, 2), omega=om)
    with spc2.Sim(chim_params(n, L, rho, 0.01, 1.0, 1.45, 0.01, seed=42)) as s:
        s.init_random_state(0.1)
        s.step("srk", dt, 0, steps)
        g = s.get_state()
    assert np.abs(periodic_diff(g, ref, L)).max() < 1e-10


def test_order_parameter(gpu):
    n = 100000
    st = O.make_state(3, 0, n, 5.0, 0.1)
    st[2::3] = np.random.default_rng(0).vonmises(0.4, 2.0, n) + 12.0
    with spc2.Sim(chim_params(n, 5.0, 0.1, 0.01, 1.0, 1.45, 0.01)) as s:
        s.set_state(st)
        R, psi = s.order_parameter()
    Rr, pr = O.order_param(st)
    assert abs(R - Rr) < 1e-14 and abs(math.remainder(psi - pr, 2 * math.pi)) < 1e-13


# ---- stochastic parity: ensemble order parameter vs the reference's own SRK ------

def test_srk_ensemble_matches_reference_ci(gpu, golden):
    ens = golden["ensemble"]
    a = ens["args"]
    seed, members, steps, every, dt = int(a[0]), int(a[1]), int(a[2]), int(a[3]), float(a[4])
    v0, sigma, rho, alpha, D = (float(x) for x in a[5:])
    Rref = np.array(ens["R"])
    Rg = []
    for r in range(members):
        st = O.make_state(seed + r, 0, N8, L64, rho)
        with spc2.Sim(chim_params(N8, L64, rho, v0, sigma, alpha, D, "fp64", seed=5000 + r)) as s:
            s.set_state(st)
            row = [s.order_parameter()[0]]
            for k in range(steps // every):
                s.step("srk", dt, step0=k * every, nsteps=every)
                row.append(s.order_parameter()[0])
        Rg.append(row)
    Rg = np.array(Rg)
    assert np.allclose(Rg[:, 0], Rref[:, 0], rtol=0, atol=1e-13)
    mg, mr = Rg.mean(0)[1:], Rref.mean(0)[1:]
    se = np.sqrt(Rg.var(0, ddof=1)[1:] / members + Rref.var(0, ddof=1)[1:] / members)
    z = np.abs(mg - mr) / np.maximum(se, 1e-12)
    assert np.all(z <= 3.0), (z, mg, mr)


# ---- per-particle partner counts --------------------------------------------------

@pytest.mark.parametrize("prec", ["fp64", "fp32k"])
@pytest.mark.parametrize("system,neighbour,n", [("vortex", "cells", 40000), ("vortex", "verlet", 40000),
                                                ("vortex", "auto", 300), ("chimera", "cells", 30000),
                                                ("chimera", "verlet", 30000)])
def test_neighbour_counts_match_oracle(gpu, prec, system, neighbour, n):
    """spc2_neighbour_counts gives each particle's partners within rho and
    within xi_r, exactly the oracle's in every precision."""
    L = math.sqrt(n / 4000.0)
    rho = 0.05 if system == "vortex" else 0.1
    st = O.make_state(50 + n, 1, n, L, rho)
    if system == "vortex":
        p = vort_params(n, L, precision=prec, neighbour=neighbour)
        sysp = orc_vortex(n, L)
    else:
        p = chim_params(n, L, rho, 0.01, 1.0, 1.45, 0.01, prec, neighbour=neighbour)
        sysp = O.chimera(n, L, rho, threads=8)
    with spc2.Sim(p) as s:
        s.set_state(st)
        a, b = s.neighbour_counts()
    ra, rb = O.neighbour_counts(sysp, st)
    assert np.array_equal(a, ra) and np.array_equal(b, rb)
    assert ra.sum() > 0 and (rb.sum() > 0) == (system == "vortex")


@pytest.mark.parametrize("prec", ["fp64", "fp32k"])
def test_vortex_srk_ensemble_matches_reference_ci(gpu, golden, prec):
    """StochasticRungeKuttaStepperPtr on the vortex system
    (StochasticRungeKuttaStepperPtr.hpp:70-97): the ensemble mean |R|(t) of 24
    runs with Philox noise against the reference's own 24 runs with its mt19937
    noise (tests/golden, ensemble_vortex), from aligned phases as the state
    dephases.  Family-wise 95% over the 20 sampled times (Bonferroni:
    |z| <= 3.02 at each)."""
    ens = golden["ensemble_vortex"]
    a = ens["args"]
    seed, members, steps, every, dt = int(a[0]), int(a[1]), int(a[2]), int(a[3]), float(a[4])
    v, mp, mm, xa, xr, D, kappa, rho = (float(x) for x in a[5:])
    Rref = np.array(ens["R"])
    Rg = []
    for r in range(members):
        st = O.make_state(seed + r, 0, N8, L64, rho)
        st[2::3] = 1.0
        p = spc2.params("vortex", prec, n=N8, L=L64, rho=rho, velocity=v, mu_plus=mp, mu_minus=mm, xi_a=xa,
                        xi_r=xr, kappa=kappa, D=D, seed=7000 + r)
        with spc2.Sim(p) as s:
            s.set_state(st)
            row = [s.order_parameter()[0]]
            for k in range(steps // every):
                s.step("srk", dt, step0=k * every, nsteps=every)
                row.append(s.order_parameter()[0])
        Rg.append(row)
    Rg = np.array(Rg)
    assert np.allclose(Rg[:, 0], 1.0, rtol=0, atol=1e-12) and np.allclose(Rref[:, 0], 1.0, rtol=0, atol=1e-12)
    mg, mr = Rg.mean(0)[1:], Rref.mean(0)[1:]
    se = np.sqrt(Rg.var(0, ddof=1)[1:] / members + Rref.var(0, ddof=1)[1:] / members)
    z = np.abs(mg - mr) / np.maximum(se, 1e-12)
    print(f"vortex SRK ensemble {prec}: max |z| {z.max():.2f}; mean R at t = 0.1: gpu {mg[9]:.4f}, "
          f"reference {mr[9]:.4f}")
    assert np.all(z <= 3.02), (z, mg, mr)
