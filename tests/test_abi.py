"""The C-ABI library loads and exports every symbol include/spc2_hip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import subprocess

import pytest

import spc2_langevin_integration_amd as spc2


def test_header_declares_entry_points():
    names = spc2.declared_symbols()
    for must in ("spc2_create", "spc2_destroy", "spc2_set_state", "spc2_get_state", "spc2_eval_rhs",
                 "spc2_eval_rhs_ito", "spc2_step", "spc2_noise", "spc2_order_parameter", "spc2_get_cell_list",
                 "spc2_get_counts", "spc2_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(spc2.LIB_PATH):
        pytest.fail(f"{spc2.LIB_PATH} not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(spc2.LIB_PATH)
    missing = [n for n in spc2.declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    # the offload bundle id of the embedded code object names the target
    out = subprocess.run(["strings", spc2.LIB_PATH], capture_output=True, text=True)
    assert "amdgcn-amd-amdhsa--gfx950" in out.stdout


def test_default_params_without_gpu():
    p = spc2.default_params()
    assert p.n == 1000 and p.rho == 0.1 and p.seed == 42


def test_create_rejects_bad_parameters_before_touching_the_gpu():
    for kw in ({"n": 0}, {"rho": -1.0}, {"L": 0.0}, {"D": -0.1}):
        with pytest.raises(spc2.Spc2Error):
            spc2.Sim(spc2.params(**kw))


def test_verlet_rejects_cohesion_range_beyond_the_reference_stencil():
    """With linked cells of side h = L / int(L / rho) the reference sees
    cohesion partners only within its 3 x 3 stencil; a Verlet list must not
    reach beyond h (rejected before any device call)."""
    kw = dict(n=1000, L=1.0, rho=0.05, velocity=0.025, mu_plus=100.0, mu_minus=100.0, xi_a=0.01, kappa=20.0)
    with pytest.raises(spc2.Spc2Error, match="xi_r"):
        spc2.Sim(spc2.params("vortex", "fp64", "verlet", xi_r=0.0501, **kw))
