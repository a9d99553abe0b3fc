"""The continuous-batching request loop (spc2_server_*): requests that arrive
while others run join the batch as slots free up, and every result equals the
same request run alone through a single handle, bit for bit."""
import math
import time

import numpy as np
import pytest

import spc2_langevin_integration_amd as spc2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    spc2.load()
    with spc2.Sim(spc2.params(n=16)) as s:  # fails loudly without a GPU
        s.init_random_state()


def request(system, n, seed, sigma=1.0, alpha=1.45, D=0.02):
    L = math.sqrt(n / 4000.0)
    if system == "chimera":
        return spc2.params("chimera", "fp32k", n=n, L=L, rho=0.1, v0=0.01, sigma=sigma, alpha=alpha, D=D, seed=seed)
    return spc2.params("vortex", "fp32k", n=n, L=L, rho=0.1, velocity=0.05, mu_plus=100.0, mu_minus=100.0,
                       xi_a=0.02, xi_r=0.01, kappa=20.0, D=D, seed=seed)


WAVES = [
    [("chimera", 3000, 11, 60), ("vortex", 8000, 12, 35), ("chimera", 500, 13, 47)],
    [("vortex", 1200, 14, 100), ("chimera", 20000, 15, 20), ("chimera", 620, 16, 0)],
    [("chimera", 5000, 17, 73), ("vortex", 640, 18, 55), ("chimera", 1500, 19, 41)],
]


def alone(p, steps, dt, omega_std):
    with spc2.Sim(p) as s:
        s.init_random_state(omega_std)
        s.step("se", dt, 0, steps)
        return s.get_state()


def test_results_alone_and_in_the_mix(gpu):
    dt, omega_std = 0.01, 0.1
    with spc2.Server(max_replicas=3, super_steps=20, snapshot_every=20, keep_final=1, dt=dt,
                     omega_std=omega_std) as srv:
        ids = []
        for wave in WAVES:
            for system, n, seed, steps in wave:
                p = request(system, n, seed)
                ids.append((srv.submit(p, steps), p, steps))
            time.sleep(0.05)
        results = [(srv.wait(rid)[0], p, steps) for rid, p, steps in ids]
        st = srv.stats()
    for k, (got, p, steps) in enumerate(results):
        ref = alone(p, steps, dt, omega_std)
        assert np.array_equal(got, ref), f"request {k} (N={p.n}, {steps} steps) differs from the run alone"
    assert st["completed"] == len(ids) and st["queued"] == 0
    assert 2 <= st["max_in_flight"] <= 3
    assert st["repacks"] >= 4 and st["frames"] > 0
    assert 0 < st["latency_p50"] <= st["latency_p99"] <= st["latency_max"]


def test_budget_admission_and_drain(gpu):
    """A budget that holds about two requests keeps the others queued until
    slots free up; every request still completes."""
    reqs = [request("chimera", 4000, 30 + k) for k in range(6)]
    budget = int(2.5 * spc2.Batch.footprint(reqs[0]))
    with spc2.Server(hbm_budget=budget, super_steps=10, snapshot_every=0, keep_final=0) as srv:
        for p in reqs:
            srv.submit(p, 25)
        srv.drain(120.0)
        st = srv.stats()
    assert st["completed"] == 6 and st["max_in_flight"] == 2 and st["queue_wait_p99"] > 0
