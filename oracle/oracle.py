"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never by the product package.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

CHIMERA, VORTEX = 0, 1


class OrcSys(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("Lx", C.c_double), ("Ly", C.c_double), ("rho", C.c_double),
        ("system", C.c_int32), ("threads", C.c_int32),
        ("v0", C.c_double), ("sigma", C.c_double), ("alpha", C.c_double), ("D", C.c_double),
        ("mu_plus", C.c_double), ("mu_minus", C.c_double), ("xi_a", C.c_double), ("xi_r", C.c_double),
        ("kappa", C.c_double),
    ]


def build(force: bool = False) -> str:
    """Compiles liboracle.so with the committed Makefile (gcc, no GPU needed)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "-s", "oracle"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        sp = C.POINTER(OrcSys)
        L.orc_splitmix_uniform.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double, dp]
        L.orc_make_state.argtypes = [C.c_uint64, C.c_int32, C.c_int64, C.c_double, C.c_double, dp]
        L.orc_min_image.argtypes = [C.c_double, dp, C.c_int64, dp]
        L.orc_wrap_points.argtypes = [C.c_double, dp, C.c_int64, dp]
        L.orc_wrap_state.argtypes = [sp, dp]
        L.orc_num_cells.argtypes = [C.c_double, C.c_double]
        L.orc_num_cells.restype = C.c_int32
        L.orc_cell_list.argtypes = [sp, C.c_int32, C.c_int32, dp, ip, ip, ip, C.POINTER(C.c_int64)]
        L.orc_rhs_k.argtypes = [sp, dp, dp, C.c_double, C.c_double, dp, C.c_void_p]
        L.orc_rhs_k_scaled.argtypes = [sp, dp, dp, C.c_double, C.c_double, dp, C.c_void_p, C.c_void_p]
        L.orc_rhs_ito.argtypes = [sp, dp, dp, dp, C.c_void_p]
        L.orc_neighbour_counts.argtypes = [sp, dp, ip, ip]
        L.orc_step_se.argtypes = [sp, dp, C.c_double, C.c_void_p, dp, C.c_int32]
        L.orc_step_srk.argtypes = [sp, dp, C.c_double, C.c_void_p, dp]
        L.orc_step_rk4.argtypes = [sp, dp, C.c_double, dp]
        L.orc_order_param.argtypes = [dp, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_order_param_f32.argtypes = [dp, C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_philox_normals.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int64, C.c_int32, dp]
        L.orc_init_state.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_double, dp, dp]
        L.orc_mt_state_size.restype = C.c_size_t
        L.orc_mt_seed.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_mt_uniform.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.c_double, dp]
        L.orc_mt_normals.argtypes = [C.c_void_p, C.c_int64, dp]
        for f in ("orc_cell_list", "orc_rhs_k", "orc_rhs_k_scaled", "orc_rhs_ito", "orc_neighbour_counts", "orc_step_se", "orc_step_srk", "orc_step_rk4"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def chimera(n, L, rho, v0=0.01, sigma=1.0, alpha=1.45, D=0.01, threads=1, Ly=None):
    return OrcSys(n=n, Lx=L, Ly=L if Ly is None else Ly, rho=rho, system=CHIMERA, threads=threads,
                  v0=v0, sigma=sigma, alpha=alpha, D=D)


def vortex(n, L, rho=0.05, velocity=0.025, mu_plus=100.0, mu_minus=100.0, xi_a=0.01, xi_r=0.005,
           kappa=20.0, D=0.0, threads=1, Ly=None):
    return OrcSys(n=n, Lx=L, Ly=L if Ly is None else Ly, rho=rho, system=VORTEX, threads=threads,
                  v0=velocity, sigma=0.0, alpha=0.0, D=D, mu_plus=mu_plus, mu_minus=mu_minus,
                  xi_a=xi_a, xi_r=xi_r, kappa=kappa)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def uniform(seed, n, lo, hi):
    out = np.empty(n, np.float64)
    lib().orc_splitmix_uniform(seed, n, lo, hi, out)
    return out


def make_state(seed, mode, n, L, rho):
    out = np.empty(3 * n, np.float64)
    lib().orc_make_state(seed, mode, n, L, rho, out)
    return out


def min_image(L, pts4):
    pts4 = np.ascontiguousarray(pts4, np.float64)
    m = pts4.size // 4
    out = np.empty(2 * m, np.float64)
    lib().orc_min_image(L, pts4, m, out)
    return out


def wrap_points(L, pts2):
    pts2 = np.ascontiguousarray(pts2, np.float64)
    out = np.empty_like(pts2)
    lib().orc_wrap_points(L, pts2, pts2.size // 2, out)
    return out


def num_cells(L, rho):
    return lib().orc_num_cells(L, rho)


def cell_list(sysp, pos, ncx=None, ncy=None):
    ncx = num_cells(sysp.Lx, sysp.rho) if ncx is None else ncx
    ncy = num_cells(sysp.Ly, sysp.rho) if ncy is None else ncy
    n = sysp.n
    cell_of = np.empty(n, np.int32)
    start = np.empty(ncx * ncy + 1, np.int32)
    srt = np.empty(n, np.int32)
    clamped = C.c_int64(0)
    rc = lib().orc_cell_list(C.byref(sysp), ncx, ncy, np.ascontiguousarray(pos, np.float64), cell_of, start, srt,
                             C.byref(clamped))
    assert rc == 0
    return cell_of, start, srt, clamped.value


def rhs_k(sysp, state, k_prev, k_coef, dt, k_next, term_scale=False):
    """k-form RHS; with term_scale=True also returns each particle's mean
    absolute phase-derivative term (0 without partners)."""
    counts = np.zeros(3, np.int64)
    k_next = np.array(k_next, np.float64, copy=True)
    kp = k_next if k_prev is k_next else np.ascontiguousarray(k_prev, np.float64)
    ts = np.zeros(sysp.n, np.float64) if term_scale else None
    rc = lib().orc_rhs_k_scaled(C.byref(sysp), np.ascontiguousarray(state, np.float64), kp, k_coef, dt, k_next,
                                counts.ctypes.data_as(C.c_void_p), _ptr(ts))
    assert rc == 0
    return (k_next, counts, ts) if term_scale else (k_next, counts)


def neighbour_counts(sysp, state):
    """Per-particle partner counts at the state: (within rho, within xi_r)."""
    a = np.empty(sysp.n, np.int32)
    b = np.empty(sysp.n, np.int32)
    rc = lib().orc_neighbour_counts(C.byref(sysp), np.ascontiguousarray(state, np.float64), a, b)
    assert rc == 0
    return a, b


def rhs_ito(sysp, state, derivative=None, additional=None):
    n = sysp.n
    d = np.zeros(3 * n) if derivative is None else np.array(derivative, np.float64, copy=True)
    a = np.zeros(2 * n) if additional is None else np.array(additional, np.float64, copy=True)
    counts = np.zeros(3, np.int64)
    rc = lib().orc_rhs_ito(C.byref(sysp), np.ascontiguousarray(state, np.float64), d, a,
                           counts.ctypes.data_as(C.c_void_p))
    assert rc == 0
    return d, a, counts


def step_se(sysp, state, dt, xi, omega=None, variant=0):
    st = np.array(state, np.float64, copy=True)
    om = None if omega is None else np.ascontiguousarray(omega, np.float64)
    rc = lib().orc_step_se(C.byref(sysp), st, dt, _ptr(om), np.ascontiguousarray(xi, np.float64), variant)
    assert rc == 0
    return st


def step_srk(sysp, state, dt, xi2, omega=None):
    st = np.array(state, np.float64, copy=True)
    om = None if omega is None else np.ascontiguousarray(omega, np.float64)
    rc = lib().orc_step_srk(C.byref(sysp), st, dt, _ptr(om), np.ascontiguousarray(xi2, np.float64))
    assert rc == 0
    return st


def step_rk4(sysp, state, dt, kbuf):
    """Advances `state` (copied) one RK4 step; kbuf (4*3N) is updated in place."""
    st = np.array(state, np.float64, copy=True)
    assert kbuf.dtype == np.float64 and kbuf.flags.c_contiguous
    rc = lib().orc_step_rk4(C.byref(sysp), st, dt, kbuf)
    assert rc == 0
    return st


def wrap_state(sysp, state):
    st = np.array(state, np.float64, copy=True)
    lib().orc_wrap_state(C.byref(sysp), st)
    return st


def order_param(state):
    n = state.size // 3
    R, psi = C.c_double(), C.c_double()
    lib().orc_order_param(np.ascontiguousarray(state, np.float64), n, C.byref(R), C.byref(psi))
    return R.value, psi.value


def order_param_f32(state):
    n = state.size // 3
    R, psi = C.c_float(), C.c_float()
    lib().orc_order_param_f32(np.ascontiguousarray(state, np.float64), n, C.byref(R), C.byref(psi))
    return R.value, psi.value


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def philox_normals(seed, step, stream, n, per_particle):
    out = np.empty(n * per_particle, np.float64)
    lib().orc_philox_normals(seed, step, stream, n, per_particle, out)
    return out


def init_state(seed, n, Lx, Ly, omega_std):
    st = np.empty(3 * n, np.float64)
    om = np.empty(n, np.float64)
    lib().orc_init_state(seed, n, Lx, Ly, omega_std, st, om)
    return st, om


class MT19937Normals:
    """mt19937 + libstdc++ normal_distribution replay (fresh distribution per call)."""

    def __init__(self, seed):
        self._buf = C.create_string_buffer(lib().orc_mt_state_size())
        lib().orc_mt_seed(self._buf, seed)

    def normals(self, count):
        out = np.empty(count, np.float64)
        lib().orc_mt_normals(self._buf, count, out)
        return out

    def uniform(self, count, lo, hi):
        out = np.empty(count, np.float64)
        lib().orc_mt_uniform(self._buf, count, C.c_double(lo), C.c_double(hi), out)
        return out

    def reference_initial_state(self, n, L):
        """x, y ~ U[0, L), phi ~ U[0, 2 pi) drawn per particle in that order
        (SimulationEnginePtr.cpp:36-45)."""
        st = np.empty(3 * n, np.float64)
        for i in range(n):
            st[3 * i:3 * i + 2] = self.uniform(2, 0.0, L)
            st[3 * i + 2] = self.uniform(1, 0.0, 2.0 * math.pi)[0]
        return st
