"""ctypes binding of the CPU oracle (oracle/sg_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module as the checker.  The product path
(spin-glass_amd/) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libsgoracle.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_philox.argtypes = [_u32p, _u32p, _u32p]
        L.orc_exp.argtypes = [ctypes.c_double]
        L.orc_exp.restype = ctypes.c_double
        L.orc_thresholds.argtypes = [ctypes.c_double, ctypes.c_double, _u64p]
        L.orc_create.argtypes = [ctypes.c_int] * 4 + [_dp, ctypes.c_double,
                                                      ctypes.c_uint64, ctypes.c_uint64,
                                                      ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_init.argtypes = [ctypes.c_void_p]
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.orc_pt_step.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.orc_sweep_count.argtypes = [ctypes.c_void_p]
        L.orc_sweep_count.restype = ctypes.c_int64
        L.orc_nfield.argtypes = [ctypes.c_void_p]
        L.orc_n_meas.argtypes = [ctypes.c_void_p]
        L.orc_n_meas.restype = ctypes.c_int64
        L.orc_get_meas.argtypes = [ctypes.c_void_p, ctypes.c_int64, _i32p]
        for name in ("orc_get_spins", "orc_set_spins", "orc_get_couplings", "orc_set_couplings"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _u64p]
        L.orc_energies.argtypes = [ctypes.c_void_p, _i32p, _i32p]
        L.orc_swap_counts.argtypes = [ctypes.c_void_p, _u64p, _u64p]
        L.orc_sweep_lattice.argtypes = [ctypes.c_int, _u64p, _u64p, _u64p, ctypes.c_uint64,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_int]
        L.orc_lattice_energy.argtypes = [ctypes.c_int, _u64p, _u64p, _i32p, _i32p]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(_ptr(c, _u32p), _ptr(k, _u32p), _ptr(out, _u32p))
    return out


def det_exp(x: float) -> float:
    return lib().orc_exp(float(x))


def thresholds(beta: float, H: float = 0.0) -> np.ndarray:
    T = np.zeros(14, dtype=np.uint64)
    lib().orc_thresholds(float(beta), float(H), _ptr(T, _u64p))
    return T.reshape(2, 7)


def sweep_lattice(L, S, J, T, spin_seed, g, rho, t, n, scalar=False):
    S = np.ascontiguousarray(S, dtype=np.uint64)
    J = np.ascontiguousarray(J, dtype=np.uint64)
    T = np.ascontiguousarray(T, dtype=np.uint64).reshape(14)
    lib().orc_sweep_lattice(L, _ptr(S, _u64p), _ptr(J, _u64p), _ptr(T, _u64p),
                            spin_seed, g, rho, t, n, 1 if scalar else 0)
    return S


def lattice_energy(L, S, J):
    S = np.ascontiguousarray(S, dtype=np.uint64)
    J = np.ascontiguousarray(J, dtype=np.uint64)
    U = np.zeros(64, dtype=np.int32)
    K = np.zeros(64, dtype=np.int32)
    lib().orc_lattice_energy(L, _ptr(S, _u64p), _ptr(J, _u64p), _ptr(U, _i32p), _ptr(K, _i32p))
    return U, K


def pt_step(sim, b: int) -> None:
    lib().orc_pt_step(sim._h, int(b))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


class OracleSim:
    """The oracle's whole-run object: couplings, spins, PT schedule, records."""

    def __init__(self, L, n_groups, n_replicas, betas, H=0.0, coupling_seed=1,
                 spin_seed=2, pt_every=1, meas_every=0, scalar=False):
        self.L, self.G, self.R = L, n_groups, n_replicas
        self.betas = np.ascontiguousarray(betas, dtype=np.float64)
        self.T = len(self.betas)
        self.N = L ** 3
        self.H = H
        self._h = lib().orc_create(L, n_groups, n_replicas, self.T, _ptr(self.betas, _dp),
                                   float(H), coupling_seed, spin_seed, pt_every,
                                   meas_every, 1 if scalar else 0)
        if not self._h:
            raise ValueError("invalid oracle parameters")
        self.nfield = lib().orc_nfield(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def init(self):
        lib().orc_init(self._h)

    def run(self, n_sweeps):
        lib().orc_run(self._h, int(n_sweeps))

    @property
    def sweep(self):
        return lib().orc_sweep_count(self._h)

    def spins(self):
        out = np.zeros((self.G, self.R, self.T, self.N), dtype=np.uint64)
        lib().orc_get_spins(self._h, _ptr(out, _u64p))
        return out

    def set_spins(self, S):
        S = np.ascontiguousarray(S, dtype=np.uint64)
        assert S.size == self.G * self.R * self.T * self.N
        lib().orc_set_spins(self._h, _ptr(S, _u64p))

    def couplings(self):
        out = np.zeros((self.G, 3, self.N), dtype=np.uint64)
        lib().orc_get_couplings(self._h, _ptr(out, _u64p))
        return out

    def set_couplings(self, J):
        J = np.ascontiguousarray(J, dtype=np.uint64)
        lib().orc_set_couplings(self._h, _ptr(J, _u64p))

    def energies(self):
        U = np.zeros((self.G, self.R, self.T, 64), dtype=np.int32)
        K = np.zeros_like(U)
        lib().orc_energies(self._h, _ptr(U, _i32p), _ptr(K, _i32p))
        return U, K

    def n_meas(self):
        return lib().orc_n_meas(self._h)

    def meas(self, i):
        out = np.zeros((self.G, self.T, 64, self.nfield), dtype=np.int32)
        lib().orc_get_meas(self._h, int(i), _ptr(out, _i32p))
        return out

    def swap_counts(self):
        n = max(self.T - 1, 1)
        a = np.zeros(n, dtype=np.uint64)
        b = np.zeros(n, dtype=np.uint64)
        lib().orc_swap_counts(self._h, _ptr(a, _u64p), _ptr(b, _u64p))
        return a[: self.T - 1], b[: self.T - 1]
