"""ctypes binding of the C ABI declared in include/spc2_hip.h.

The product's compute path is libspc2_hip.so; this module only moves numpy
arrays across the ABI for tests, benchmarks and Python callers.  Loading fails
loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libspc2_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "spc2_hip.h")

CHIMERA, VORTEX = 0, 1
FP64, FP32_KAHAN, FP64_FACTORED = 0, 1, 2
SE, SRK, RK4 = 0, 1, 2
STEP_NO_WRAP = 1
NEIGH_AUTO, NEIGH_CELLS, NEIGH_ALL_PAIRS, NEIGH_VERLET = 0, 1, 2, 3

PRECISIONS = {"fp64": FP64, "fp32k": FP32_KAHAN, "fp64f": FP64_FACTORED}
METHODS = {"se": SE, "srk": SRK, "rk4": RK4}


class Params(C.Structure):
    _fields_ = [
        ("system", C.c_int32), ("precision", C.c_int32), ("neighbour", C.c_int32), ("reserved", C.c_int32),
        ("n", C.c_int64), ("Lx", C.c_double), ("Ly", C.c_double), ("rho", C.c_double),
        ("v0", C.c_double), ("sigma", C.c_double), ("alpha", C.c_double), ("D", C.c_double),
        ("mu_plus", C.c_double), ("mu_minus", C.c_double), ("xi_a", C.c_double), ("xi_r", C.c_double),
        ("kappa", C.c_double), ("seed", C.c_uint64), ("verlet_skin", C.c_double),
    ]


class Spc2Error(RuntimeError):
    pass


_lib = None


def load():
    """Loads libspc2_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Spc2Error(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                        "(make -C spc2_langevin_integration_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    h = C.c_void_p
    dp = C.POINTER(C.c_double)
    st = C.c_int
    sig = {
        "spc2_default_params": (None, [C.POINTER(Params)]),
        "spc2_create": (st, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
        "spc2_destroy": (st, [h]),
        "spc2_last_error": (C.c_char_p, []),
        "spc2_set_state": (st, [h, dp]),
        "spc2_get_state": (st, [h, dp]),
        "spc2_init_random_state": (st, [h, C.c_double]),
        "spc2_set_omega": (st, [h, dp]),
        "spc2_eval_rhs": (st, [h, dp, C.c_double, C.c_double, dp]),
        "spc2_eval_rhs_ito": (st, [h, dp, dp]),
        "spc2_step": (st, [h, C.c_int32, C.c_double, C.c_int64, C.c_int64, dp]),
        "spc2_step_ex": (st, [h, C.c_int32, C.c_double, C.c_int64, C.c_int64, dp, C.c_int32]),
        "spc2_snapshot_async": (st, [h, C.c_void_p]),
        "spc2_snapshot_wait": (st, [h]),
        "spc2_host_alloc": (st, [C.c_uint64, C.POINTER(C.c_void_p)]),
        "spc2_host_free": (st, [C.c_void_p]),
        "spc2_noise": (st, [h, C.c_int32, C.c_int64, dp]),
        "spc2_noise_noalloc": (st, [h, C.c_int32, C.c_int64, dp]),
        "spc2_neighbour_counts": (st, [h, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "spc2_checkpoint_size": (st, [h, C.POINTER(C.c_uint64)]),
        "spc2_checkpoint_save": (st, [h, C.c_void_p, C.c_uint64]),
        "spc2_checkpoint_load": (st, [h, C.c_void_p, C.c_uint64]),
        "spc2_order_parameter": (st, [h, dp, dp]),
        "spc2_get_cell_list": (st, [h, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "spc2_num_cells": (st, [h, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "spc2_set_counting": (st, [h, C.c_int32]),
        "spc2_get_counts": (st, [h, C.POINTER(C.c_int64)]),
        "spc2_set_profiling": (st, [h, C.c_int32]),
        "spc2_get_profile": (st, [h, dp, C.POINTER(C.c_int64)]),
        "spc2_synchronize": (st, [h]),
        "spc2_verlet_stats": (st, [h, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "spc2_verlet_layout": (st, [h, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "spc2_batch_footprint": (C.c_uint64, [C.POINTER(Params)]),
        "spc2_batch_create": (st, [C.POINTER(Params), C.c_int32, C.POINTER(C.c_void_p)]),
        "spc2_batch_create_from": (st, [C.POINTER(Params), C.c_int32, h, C.POINTER(C.c_int32), C.c_double,
                                        C.POINTER(C.c_void_p)]),
        "spc2_batch_destroy": (st, [h]),
        "spc2_batch_init_random_state": (st, [h, C.c_double]),
        "spc2_batch_set_state": (st, [h, C.c_int32, dp]),
        "spc2_batch_get_state": (st, [h, C.c_int32, dp]),
        "spc2_batch_set_omega": (st, [h, C.c_int32, dp]),
        "spc2_batch_step": (st, [h, C.c_int32, C.c_double, C.c_int64, C.POINTER(C.c_int64)]),
        "spc2_batch_steps_done": (st, [h, C.POINTER(C.c_int64)]),
        "spc2_batch_order_parameters": (st, [h, dp, dp]),
        "spc2_batch_snapshot_async": (st, [h, C.c_void_p]),
        "spc2_batch_snapshot_wait": (st, [h]),
        "spc2_batch_layout": (st, [h, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "spc2_batch_set_tile_targets": (st, [h, dp]),
        "spc2_batch_get_tile_targets": (st, [h, dp]),
        "spc2_batch_autotune": (st, [h, C.c_double, C.c_int32, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = load().spc2_last_error().decode()
        raise Spc2Error(f"spc2 status {rc}: {msg}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def default_params() -> Params:
    p = Params()
    load().spc2_default_params(C.byref(p))
    return p


def params(system="chimera", precision="fp64", neighbour="auto", **kw) -> Params:
    p = default_params()
    p.system = CHIMERA if system == "chimera" else VORTEX
    p.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    p.neighbour = {"auto": NEIGH_AUTO, "cells": NEIGH_CELLS, "all_pairs": NEIGH_ALL_PAIRS,
                   "verlet": NEIGH_VERLET}[neighbour]
    if "L" in kw:
        L = kw.pop("L")
        p.Lx = p.Ly = L
    if "velocity" in kw:
        kw["v0"] = kw.pop("velocity")
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Sim:
    """One device-resident simulation (an spc2_sim_t handle)."""

    def __init__(self, p: Params):
        self.p = p
        self.n = int(p.n)
        h = C.c_void_p()
        _check(load().spc2_create(C.byref(p), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            load().spc2_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, state):
        st = np.ascontiguousarray(state, np.float64)
        assert st.size == 3 * self.n
        _check(load().spc2_set_state(self.h, _dp(st)))

    def get_state(self):
        st = np.empty(3 * self.n, np.float64)
        _check(load().spc2_get_state(self.h, _dp(st)))
        return st

    def init_random_state(self, omega_std=0.0):
        _check(load().spc2_init_random_state(self.h, omega_std))

    def set_omega(self, omega):
        om = np.ascontiguousarray(omega, np.float64)
        assert om.size == self.n
        _check(load().spc2_set_omega(self.h, _dp(om)))

    def eval_rhs(self, k_prev, k_coef, dt, k_next):
        kp = np.ascontiguousarray(k_prev, np.float64)
        kn = np.array(k_next, np.float64, copy=True)
        _check(load().spc2_eval_rhs(self.h, _dp(kp), k_coef, dt, _dp(kn)))
        return kn

    def eval_rhs_ito(self, derivative=None, additional=None):
        d = np.zeros(3 * self.n) if derivative is None else np.array(derivative, np.float64, copy=True)
        a = np.zeros(2 * self.n) if additional is None else np.array(additional, np.float64, copy=True)
        _check(load().spc2_eval_rhs_ito(self.h, _dp(d), _dp(a)))
        return d, a

    def step(self, method, dt, step0=0, nsteps=1, noise=None, wrap=True):
        m = METHODS[method] if isinstance(method, str) else int(method)
        nz = None
        if noise is not None:
            nz = np.ascontiguousarray(noise, np.float64)
        flags = 0 if wrap else STEP_NO_WRAP
        _check(load().spc2_step_ex(self.h, m, dt, step0, nsteps, _dp(nz) if nz is not None else None, flags))

    def snapshot(self, buf=None):
        """Float frame [x, y, phi] by particle id through the asynchronous
        snapshot path; `buf` may be a PinnedFrame to reuse."""
        own = buf is None
        if own:
            buf = PinnedFrame(3 * self.n)
        try:
            _check(load().spc2_snapshot_async(self.h, buf.ptr))
            _check(load().spc2_snapshot_wait(self.h))
            return buf.array().copy()
        finally:
            if own:
                buf.free()

    def noise(self, method, step):
        m = METHODS[method] if isinstance(method, str) else int(method)
        out = np.empty(self.n * (2 if m == SRK else 1), np.float64)
        _check(load().spc2_noise(self.h, m, step, _dp(out)))
        return out

    def neighbour_counts(self):
        """Per-particle partner counts at the current state: (within rho, within xi_r)."""
        a = np.empty(self.n, np.int32)
        b = np.empty(self.n, np.int32)
        _check(load().spc2_neighbour_counts(self.h, _ip(a), _ip(b)))
        return a, b

    def checkpoint(self):
        """Exact checkpoint of the handle (bytes)."""
        size = C.c_uint64()
        _check(load().spc2_checkpoint_size(self.h, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        _check(load().spc2_checkpoint_save(self.h, buf, size.value))
        return buf.raw

    def restore(self, blob):
        """Loads a checkpoint written by a handle with the same parameters."""
        buf = C.create_string_buffer(bytes(blob), len(blob))
        _check(load().spc2_checkpoint_load(self.h, buf, len(blob)))

    def noise_noalloc(self, method, step):
        """noise() through the entry point that allocates no device memory."""
        m = METHODS[method] if isinstance(method, str) else int(method)
        out = np.empty(self.n * (2 if m == SRK else 1), np.float64)
        _check(load().spc2_noise_noalloc(self.h, m, step, _dp(out)))
        return out

    def order_parameter(self):
        R, psi = C.c_double(), C.c_double()
        _check(load().spc2_order_parameter(self.h, C.byref(R), C.byref(psi)))
        return R.value, psi.value

    def num_cells(self):
        nx, ny = C.c_int32(), C.c_int32()
        _check(load().spc2_num_cells(self.h, C.byref(nx), C.byref(ny)))
        return nx.value, ny.value

    def cell_list(self):
        nx, ny = self.num_cells()
        cell_of = np.empty(self.n, np.int32)
        start = np.empty(nx * ny + 1, np.int32)
        srt = np.empty(self.n, np.int32)
        _check(load().spc2_get_cell_list(self.h, _ip(cell_of), _ip(start), _ip(srt)))
        return cell_of, start, srt

    def set_counting(self, on=True):
        _check(load().spc2_set_counting(self.h, int(on)))

    def counts(self):
        c = np.zeros(3, np.int64)
        _check(load().spc2_get_counts(self.h, c.ctypes.data_as(C.POINTER(C.c_int64))))
        return c

    def set_profiling(self, on=True):
        _check(load().spc2_set_profiling(self.h, int(on)))

    def profile(self):
        ms, n = C.c_double(), C.c_int64()
        _check(load().spc2_get_profile(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def synchronize(self):
        _check(load().spc2_synchronize(self.h))

    def verlet_stats(self):
        """(list builds so far, entries of the current list)."""
        b, e = C.c_int64(), C.c_int64()
        _check(load().spc2_verlet_stats(self.h, C.byref(b), C.byref(e)))
        return b.value, e.value

    def verlet_layout(self):
        """(1 if the current list feeds the LDS-staged passes else 0, their LDS capacity)."""
        lds, cap = C.c_int32(), C.c_int32()
        _check(load().spc2_verlet_layout(self.h, C.byref(lds), C.byref(cap)))
        return lds.value, cap.value


class Batch:
    """Replicas stepped together (an spc2_batch_t handle)."""

    def __init__(self, replicas, prev=None, prev_index=None, omega_std=0.0):
        self.params = list(replicas)
        self.count = len(self.params)
        arr = (Params * self.count)(*self.params)
        h = C.c_void_p()
        if prev is None:
            _check(load().spc2_batch_create(arr, self.count, C.byref(h)))
        else:
            idx = np.ascontiguousarray(prev_index, np.int32)
            assert idx.size == self.count
            _check(load().spc2_batch_create_from(arr, self.count, prev.h, idx.ctypes.data_as(C.POINTER(C.c_int32)),
                                                 omega_std, C.byref(h)))
        self.h = h

    @staticmethod
    def footprint(p):
        return int(load().spc2_batch_footprint(C.byref(p)))

    def close(self):
        if self.h:
            load().spc2_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_random_state(self, omega_std=0.0):
        _check(load().spc2_batch_init_random_state(self.h, omega_std))

    def set_state(self, r, state):
        st = np.ascontiguousarray(state, np.float64)
        assert st.size == 3 * self.params[r].n
        _check(load().spc2_batch_set_state(self.h, r, _dp(st)))

    def get_state(self, r):
        st = np.empty(3 * self.params[r].n, np.float64)
        _check(load().spc2_batch_get_state(self.h, r, _dp(st)))
        return st

    def set_omega(self, r, omega):
        om = np.ascontiguousarray(omega, np.float64)
        _check(load().spc2_batch_set_omega(self.h, r, _dp(om)))

    def step(self, dt, nsteps, limit=None, method="se"):
        m = METHODS[method] if isinstance(method, str) else int(method)
        lim = None
        if limit is not None:
            lim = np.ascontiguousarray(limit, np.int64)
            assert lim.size == self.count
        _check(load().spc2_batch_step(self.h, m, dt, nsteps,
                                      lim.ctypes.data_as(C.POINTER(C.c_int64)) if lim is not None else None))

    def steps_done(self):
        out = np.zeros(self.count, np.int64)
        _check(load().spc2_batch_steps_done(self.h, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def order_parameters(self):
        R = np.empty(self.count)
        psi = np.empty(self.count)
        _check(load().spc2_batch_order_parameters(self.h, _dp(R), _dp(psi)))
        return R, psi

    def layout(self):
        first = np.empty(self.count, np.int32)
        n = np.empty(self.count, np.int32)
        _check(load().spc2_batch_layout(self.h, _ip(first), _ip(n)))
        return first, n

    def snapshot(self, buf=None):
        """Float frames of every replica (list of arrays, AoS [x, y, phi] by id)."""
        first, n = self.layout()
        total = int(n.sum())
        own = buf is None
        if own:
            buf = PinnedFrame(3 * total)
        try:
            _check(load().spc2_batch_snapshot_async(self.h, buf.ptr))
            _check(load().spc2_batch_snapshot_wait(self.h))
            a = buf.array().copy()
        finally:
            if own:
                buf.free()
        return [a[3 * f:3 * (f + k)] for f, k in zip(first, n)]

    def set_tile_targets(self, targets):
        t = np.ascontiguousarray(targets, np.float64)
        assert t.size == self.count
        _check(load().spc2_batch_set_tile_targets(self.h, _dp(t)))

    def autotune(self, dt, trial_steps=4):
        """Tunes the tile target per N bucket; returns the tuned ms per step."""
        ms = C.c_double()
        _check(load().spc2_batch_autotune(self.h, dt, trial_steps, C.byref(ms)))
        return ms.value

    def tile_targets(self):
        t = np.empty(self.count)
        _check(load().spc2_batch_get_tile_targets(self.h, _dp(t)))
        return t


class PinnedFrame:
    """Pinned host buffer of `count` floats from spc2_host_alloc."""

    def __init__(self, count):
        self.count = int(count)
        p = C.c_void_p()
        _check(load().spc2_host_alloc(4 * self.count, C.byref(p)))
        self.ptr = p

    def array(self):
        return np.ctypeslib.as_array(C.cast(self.ptr, C.POINTER(C.c_float)), shape=(self.count,))

    def free(self):
        if self.ptr:
            load().spc2_host_free(self.ptr)
            self.ptr = None


HOST_LIB_PATH = os.path.join(PKG_DIR, "libspc2_host.so")
SERVER_HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "spc2_server.h")


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_uint64), ("max_particles", C.c_int64), ("max_replicas", C.c_int32),
        ("super_steps", C.c_int32), ("snapshot_every", C.c_int32), ("keep_final", C.c_int32),
        ("dt", C.c_double), ("omega_std", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("admitted", C.c_int64), ("completed", C.c_int64), ("queued", C.c_int64),
        ("in_flight", C.c_int64), ("repacks", C.c_int64), ("frames", C.c_int64), ("frame_bytes", C.c_uint64),
        ("particle_steps", C.c_double), ("particle_steps_executed", C.c_double), ("wall_seconds", C.c_double),
        ("busy_seconds", C.c_double),
        ("latency_p50", C.c_double), ("latency_p99", C.c_double), ("latency_max", C.c_double),
        ("queue_wait_p50", C.c_double), ("queue_wait_p99", C.c_double),
        ("max_in_flight", C.c_int64), ("max_particles_in_flight", C.c_int64),
    ]


_host = None


def load_host():
    """Loads libspc2_host.so (the C++ host layer with the serving loop)."""
    global _host
    if _host is not None:
        return _host
    load()
    if not os.path.exists(HOST_LIB_PATH):
        raise Spc2Error(f"{HOST_LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(HOST_LIB_PATH)
    h = C.c_void_p
    st = C.c_int
    sig = {
        "spc2_server_default_config": (None, [C.POINTER(ServerConfig)]),
        "spc2_server_create": (st, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
        "spc2_server_destroy": (st, [h]),
        "spc2_server_submit": (st, [h, C.POINTER(Params), C.c_int64, C.POINTER(C.c_int64)]),
        "spc2_server_wait": (st, [h, C.c_int64, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "spc2_server_drain": (st, [h, C.c_double]),
        "spc2_server_stats_get": (st, [h, C.POINTER(ServerStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _host = L
    return L


class Server:
    """The continuous-batching request loop (spc2_server_t)."""

    def __init__(self, **cfg):
        c = ServerConfig()
        load_host().spc2_server_default_config(C.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.cfg = c
        self.n = {}
        h = C.c_void_p()
        _check(load_host().spc2_server_create(C.byref(c), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            load_host().spc2_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, p, steps):
        rid = C.c_int64()
        _check(load_host().spc2_server_submit(self.h, C.byref(p), steps, C.byref(rid)))
        self.n[rid.value] = int(p.n)
        return rid.value

    def wait(self, rid, timeout=120.0):
        st = np.empty(3 * self.n[rid], np.float64) if self.cfg.keep_final else None
        lat = C.c_double()
        _check(load_host().spc2_server_wait(self.h, rid, timeout, _dp(st) if st is not None else None,
                                            C.byref(lat)))
        return st, lat.value

    def drain(self, timeout=-1.0):
        _check(load_host().spc2_server_drain(self.h, timeout))

    def stats(self):
        s = ServerStats()
        _check(load_host().spc2_server_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}


def declared_symbols(header=HEADER_PATH):
    """Names of the functions include/spc2_hip.h declares."""
    import re
    txt = open(header).read()
    return sorted(set(re.findall(r"\b(spc2_[a-z0-9_]+)\s*\(", txt)))
