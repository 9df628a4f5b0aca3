"""Python binding of libsg (ctypes over include/sg.h) and a mirror of the
reference's host object.

`Sim` wraps one sg_handle.  `eamsim` keeps the reference's method names and
argument meaning (Program/src/eamsim.cuh:66-110): eamsim(H), init_lattice(),
load_checkpoint(file), run_simulation(file), save_checkpoint(file).  Errors
raise SGError (the reference throws `error`, Program/src/util.cuh:46-54).

There is no fallback: if libsg.so is missing or cannot load, every entry
point raises.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsg.so")

SG_FORM_PLAIN = 0
SG_FORM_BATCHED = 1
SG_KERNEL_AUTO, SG_KERNEL_RESIDENT, SG_KERNEL_STREAMED, SG_KERNEL_RESIDENT_SINGLE = 0, 1, 2, 3
SG_KERNEL_SPLIT = 4
RES_COLUMNS = ["beta", "e", "m", "q", "q2", "q4", "binder", "chi0", "chi1", "xi_L", "n_meas"]
NRES = len(RES_COLUMNS)

# the reference's temperature ladder, eamsim.cu:728-737
REF_BETAS = 2.0 * 2.0 ** (-4.0 * np.arange(24) / 23.0)


class SGError(RuntimeError):
    pass


class _Params(ctypes.Structure):
    _fields_ = [
        ("L", ctypes.c_int32), ("n_samples", ctypes.c_int32), ("n_temps", ctypes.c_int32),
        ("n_replicas", ctypes.c_int32), ("betas", ctypes.POINTER(ctypes.c_double)),
        ("T_min", ctypes.c_double), ("T_max", ctypes.c_double), ("H", ctypes.c_double),
        ("coupling_seed", ctypes.c_uint64), ("spin_seed", ctypes.c_uint64),
        ("pt_every", ctypes.c_int32), ("meas_every", ctypes.c_int32),
        ("device", ctypes.c_int32), ("form", ctypes.c_int32),
        ("kernel", ctypes.c_int32), ("tile", ctypes.c_int32),
    ]


SINK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32),
                        ctypes.c_int64, ctypes.c_int32)

# every symbol include/sg.h declares
EXPORTS = [
    "sg_params_default", "sg_create", "sg_destroy", "sg_last_error", "sg_init_lattice",
    "sg_run", "sg_save_checkpoint", "sg_load_checkpoint", "sg_nfield", "sg_n_groups",
    "sg_sweep_count", "sg_get_betas", "sg_get_spins", "sg_set_spins", "sg_get_couplings",
    "sg_set_couplings", "sg_get_energies", "sg_get_swap_counts", "sg_get_results",
    "sg_get_timing", "sg_set_profiling", "sg_get_kernel_stats", "sg_tune_kernel", "sg_tune_shapes",
    "sg_batch_params_default", "sg_batch_create", "sg_batch_destroy", "sg_batch_last_error",
    "sg_job_footprint", "sg_submit", "sg_step", "sg_drain", "sg_poll", "sg_result",
    "sg_job_times", "sg_batch_stats", "sg_batch_launch_stats", "sg_batch_alloc_stats", "sg_serve",
    "sg_probe_valu", "sg_probe_last_error", "sg_prepare", "sg_graph_count",
    "sg_tune_count", "sg_run_static", "sg_tune_tiles", "sg_get_shape",
]

JOB_QUEUED, JOB_RUNNING, JOB_DONE = 0, 1, 2


class _BatchParams(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32), ("pt_every", ctypes.c_int32),
        ("slice_pt_steps", ctypes.c_int32), ("form", ctypes.c_int32),
        ("arena_bytes", ctypes.c_int64), ("max_inflight", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]

_lib = None
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def load_library(path: str = LIB_PATH):
    """Load libsg.so and declare its signatures (no device call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SGError(f"{path} is missing: run `python spin-glass_amd/build_sg.py`")
    L = ctypes.CDLL(path)
    L.sg_params_default.argtypes = [ctypes.POINTER(_Params)]
    L.sg_create.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(_vp)]
    L.sg_destroy.argtypes = [_vp]
    L.sg_last_error.argtypes = [_vp]
    L.sg_last_error.restype = ctypes.c_char_p
    L.sg_init_lattice.argtypes = [_vp]
    L.sg_run.argtypes = [_vp, ctypes.c_int64, SINK, _vp]
    L.sg_save_checkpoint.argtypes = [_vp, _vp, ctypes.POINTER(ctypes.c_size_t)]
    L.sg_load_checkpoint.argtypes = [_vp, _vp, ctypes.c_size_t]
    L.sg_nfield.argtypes = [_vp]
    L.sg_n_groups.argtypes = [_vp]
    L.sg_sweep_count.argtypes = [_vp]
    L.sg_sweep_count.restype = ctypes.c_int64
    L.sg_get_betas.argtypes = [_vp, _dp]
    for n in ("sg_get_spins", "sg_set_spins", "sg_get_couplings", "sg_set_couplings"):
        getattr(L, n).argtypes = [_vp, _u64p]
    L.sg_get_energies.argtypes = [_vp, _i32p, _i32p]
    L.sg_get_swap_counts.argtypes = [_vp, _u64p, _u64p]
    L.sg_get_results.argtypes = [_vp, _dp]
    L.sg_get_timing.argtypes = [_vp, _dp, _dp]
    L.sg_set_profiling.argtypes = [_vp, ctypes.c_int32]
    L.sg_get_kernel_stats.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), _dp, _dp]
    L.sg_tune_kernel.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                 _i32p, _dp, _dp]
    L.sg_tune_shapes.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                 _i32p, _dp]
    L.sg_tune_tiles.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                _i32p, _i32p, _dp]
    L.sg_get_shape.argtypes = [_vp, _i32p, _i32p]
    _i64p = ctypes.POINTER(ctypes.c_int64)
    L.sg_batch_params_default.argtypes = [ctypes.POINTER(_BatchParams)]
    L.sg_batch_create.argtypes = [ctypes.POINTER(_BatchParams), ctypes.POINTER(_vp)]
    L.sg_batch_destroy.argtypes = [_vp]
    L.sg_batch_last_error.argtypes = [_vp]
    L.sg_batch_last_error.restype = ctypes.c_char_p
    L.sg_job_footprint.argtypes = [ctypes.POINTER(_Params)]
    L.sg_job_footprint.restype = ctypes.c_int64
    L.sg_submit.argtypes = [_vp, ctypes.POINTER(_Params), ctypes.c_int64, _i64p]
    L.sg_step.argtypes = [_vp, _i32p, _i32p]
    L.sg_drain.argtypes = [_vp]
    L.sg_poll.argtypes = [_vp, ctypes.c_int64, _i32p]
    L.sg_result.argtypes = [_vp, ctypes.c_int64, _dp, _u64p]
    L.sg_job_times.argtypes = [_vp, ctypes.c_int64, _dp, _dp, _dp]
    L.sg_batch_stats.argtypes = [_vp, _dp, _i64p, _i64p, _i64p]
    L.sg_batch_launch_stats.argtypes = [_vp, _i64p, _i64p, _i64p]
    L.sg_batch_alloc_stats.argtypes = [_vp, _i64p, _i64p]
    L.sg_serve.argtypes = [_vp, ctypes.POINTER(_Params), _i64p, _dp, ctypes.c_int32, _i64p, _dp]
    L.sg_run_static.argtypes = [ctypes.POINTER(_Params), _vp, _vp, ctypes.c_int64, ctypes.c_int64,
                                _vp, _vp, _vp]
    L.sg_tune_count.argtypes = []
    L.sg_tune_count.restype = ctypes.c_int64
    L.sg_prepare.argtypes = [_vp, ctypes.c_int64, ctypes.c_int64]
    L.sg_graph_count.argtypes = [_vp]
    L.sg_probe_valu.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp, _dp]
    L.sg_probe_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(status, h=None):
    if status != 0:
        msg = load_library().sg_last_error(h)
        raise SGError(f"libsg status {status}: {msg.decode() if msg else ''}")


def _fill_params(p, L, n_samples, n_replicas, betas, T_range, H, coupling_seed, spin_seed,
                 pt_every, meas_every, device=0, form=SG_FORM_PLAIN, kernel=SG_KERNEL_AUTO,
                 tile=0):
    """Fills an sg_params; returns the betas array that p.betas points into
    (keep it alive while p is used)."""
    p.L, p.n_samples, p.n_replicas = L, n_samples, n_replicas
    keep = None
    if betas is not None:
        keep = np.ascontiguousarray(betas, dtype=np.float64)
        p.n_temps = len(keep)
        p.betas = keep.ctypes.data_as(_dp)
    else:
        tmin, tmax, nt = T_range
        p.n_temps = nt
        p.T_min, p.T_max = tmin, tmax
        p.betas = None
    p.H = H
    p.coupling_seed, p.spin_seed = coupling_seed, spin_seed
    p.pt_every, p.meas_every, p.device, p.form = pt_every, meas_every, device, form
    p.kernel = kernel
    p.tile = tile
    return keep


class Sim:
    """One libsg handle: L, samples (64 per group), temperatures, replicas."""

    def __init__(self, L=16, n_samples=64, n_replicas=2, betas=None, T_range=None, H=0.0,
                 coupling_seed=0x5EED0001, spin_seed=0x5EED0002, pt_every=16,
                 meas_every=0, device=0, form=SG_FORM_PLAIN, kernel=SG_KERNEL_AUTO, tile=0):
        """tile: 0 = the tuner's pick (kernel AUTO) or the shape's first tile;
        t >= 1 = tile t - 1 of the shape (see tune_tiles)."""
        lib = load_library()
        p = _Params()
        self._betas = _fill_params(p, L, n_samples, n_replicas, betas, T_range, H,
                                   coupling_seed, spin_seed, pt_every, meas_every, device, form,
                                   kernel, tile)
        h = _vp()
        _check(lib.sg_create(ctypes.byref(p), ctypes.byref(h)), None)
        self._h = h
        self.L, self.N, self.R, self.T = L, L ** 3, n_replicas, p.n_temps
        self.n_samples = n_samples
        self.G = lib.sg_n_groups(h)
        self.nfield = lib.sg_nfield(h)
        self.H = H
        self.records = []
        self.keep_records = True

    def close(self):
        if getattr(self, "_h", None):
            load_library().sg_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _c(self, st):
        _check(st, self._h)

    @property
    def betas(self):
        b = np.zeros(self.T, dtype=np.float64)
        self._c(load_library().sg_get_betas(self._h, b.ctypes.data_as(_dp)))
        return b

    @property
    def sweep(self):
        return load_library().sg_sweep_count(self._h)

    def init_lattice(self):
        self._c(load_library().sg_init_lattice(self._h))

    def run(self, n_sweeps, sink=None):
        """Run n_sweeps; each measurement is appended to self.records as an
        int32 array [group][temp][lane][field] and passed to sink(sweep, rec)."""
        def _cb(user, sweep, rec, n, nfield):
            arr = np.ctypeslib.as_array(rec, shape=(n * nfield,)).copy()
            arr = arr.reshape(self.G, self.T, 64, nfield)
            if self.keep_records:
                self.records.append((sweep, arr))
            if sink is not None:
                sink(sweep, arr)
        cb = SINK(_cb) if (self.keep_records or sink is not None) else SINK()
        self._c(load_library().sg_run(self._h, int(n_sweeps), cb, None))

    def prepare(self, n_sweeps, from_sweep=-1):
        """Instantiate every captured launch sequence that a call run(n_sweeps)
        starting at from_sweep (default: the current sweep) would replay,
        running nothing."""
        self._c(load_library().sg_prepare(self._h, int(from_sweep), int(n_sweeps)))

    @property
    def graph_count(self):
        return load_library().sg_graph_count(self._h)

    def spins(self):
        out = np.zeros((self.G, self.R, self.T, self.N), dtype=np.uint64)
        self._c(load_library().sg_get_spins(self._h, out.ctypes.data_as(_u64p)))
        return out

    def set_spins(self, S):
        S = np.ascontiguousarray(S, dtype=np.uint64)
        if S.size != self.G * self.R * self.T * self.N:
            raise SGError("spin array has the wrong size")
        self._c(load_library().sg_set_spins(self._h, S.ctypes.data_as(_u64p)))

    def couplings(self):
        out = np.zeros((self.G, 3, self.N), dtype=np.uint64)
        self._c(load_library().sg_get_couplings(self._h, out.ctypes.data_as(_u64p)))
        return out

    def set_couplings(self, J):
        J = np.ascontiguousarray(J, dtype=np.uint64)
        if J.size != self.G * 3 * self.N:
            raise SGError("coupling array has the wrong size")
        self._c(load_library().sg_set_couplings(self._h, J.ctypes.data_as(_u64p)))

    def energies(self):
        U = np.zeros((self.G, self.R, self.T, 64), dtype=np.int32)
        K = np.zeros_like(U)
        self._c(load_library().sg_get_energies(self._h, U.ctypes.data_as(_i32p),
                                               K.ctypes.data_as(_i32p)))
        return U, K

    def swap_counts(self):
        n = max(self.T - 1, 1)
        a = np.zeros(n, dtype=np.uint64)
        b = np.zeros(n, dtype=np.uint64)
        self._c(load_library().sg_get_swap_counts(self._h, a.ctypes.data_as(_u64p),
                                                  b.ctypes.data_as(_u64p)))
        return a[: self.T - 1], b[: self.T - 1]

    def results(self):
        r = np.zeros((self.T, NRES), dtype=np.float64)
        self._c(load_library().sg_get_results(self._h, r.ctypes.data_as(_dp)))
        return r

    def timing(self):
        ms = ctypes.c_double()
        fl = ctypes.c_double()
        self._c(load_library().sg_get_timing(self._h, ctypes.byref(ms), ctypes.byref(fl)))
        return ms.value, fl.value

    def set_profiling(self, on=True):
        self._c(load_library().sg_set_profiling(self._h, 1 if on else 0))

    @property
    def shape(self):
        """(SG_KERNEL_* shape, 0-based tile) the handle sweeps with"""
        k, t = ctypes.c_int32(), ctypes.c_int32()
        self._c(load_library().sg_get_shape(self._h, ctypes.byref(k), ctypes.byref(t)))
        return k.value, t.value

    def kernel_stats(self, kind=0):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        fl = ctypes.c_double()
        self._c(load_library().sg_get_kernel_stats(self._h, kind, ctypes.byref(n),
                                                   ctypes.byref(ms), ctypes.byref(fl)))
        return n.value, ms.value, fl.value

    def save_checkpoint_bytes(self) -> bytes:
        lib = load_library()
        n = ctypes.c_size_t(0)
        self._c(lib.sg_save_checkpoint(self._h, None, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._c(lib.sg_save_checkpoint(self._h, buf, ctypes.byref(n)))
        return buf.raw[: n.value]

    def load_checkpoint_bytes(self, data: bytes):
        buf = ctypes.create_string_buffer(data, len(data))
        self._c(load_library().sg_load_checkpoint(self._h, buf, len(data)))


def tune_kernel(L, n_lattices, field=False, device=0):
    """The autotuner's choice for n_lattices lattices of size L:
    (SG_KERNEL_RESIDENT, _STREAMED, _RESIDENT_SINGLE or _SPLIT, lattice-pair
    ms, streamed ms)."""
    lib = load_library()
    k = ctypes.c_int32()
    r, st = ctypes.c_double(), ctypes.c_double()
    _check(lib.sg_tune_kernel(device, L, n_lattices, 1 if field else 0, ctypes.byref(k),
                              ctypes.byref(r), ctypes.byref(st)))
    return k.value, r.value, st.value


def tune_shapes(L, n_lattices, field=False, device=0):
    """The autotuner's choice and the measured ms of every shape:
    (kernel, {SG_KERNEL_RESIDENT: ms, SG_KERNEL_STREAMED: ms,
    SG_KERNEL_RESIDENT_SINGLE: ms, SG_KERNEL_SPLIT: ms}), 0 ms where the L has
    no such shape."""
    lib = load_library()
    k = ctypes.c_int32()
    ms = (ctypes.c_double * 5)()
    _check(lib.sg_tune_shapes(device, L, n_lattices, 1 if field else 0, ctypes.byref(k), ms))
    return k.value, {SG_KERNEL_RESIDENT: ms[1], SG_KERNEL_STREAMED: ms[2],
                     SG_KERNEL_RESIDENT_SINGLE: ms[3], SG_KERNEL_SPLIT: ms[4]}


def tune_tiles(L, n_lattices, field=False, device=0):
    """The autotuner's choice down to the tile: (kernel, tile, {(shape,
    tile): ms}) with every measured (shape, 0-based tile) pair."""
    lib = load_library()
    k, t = ctypes.c_int32(), ctypes.c_int32()
    ms = (ctypes.c_double * 20)()
    _check(lib.sg_tune_tiles(device, L, n_lattices, 1 if field else 0, ctypes.byref(k),
                             ctypes.byref(t), ms))
    return k.value, t.value, {(m, i): ms[m * 4 + i] for m in range(1, 5) for i in range(4)
                              if ms[m * 4 + i] > 0}


STATIC_SCRATCH_BYTES = 32 << 20   # SG_STATIC_SCRATCH_BYTES


def checkerboard_order(L):
    """perm with kernel_order[..., perm[s]] = natural[..., s]: the site order
    sg_run_static's spin buffers use (colour c = (x+y+z)&1 at c*N/2, then
    (x>>1) + (L/2)(y + L z))."""
    N = L ** 3
    s = np.arange(N)
    x, y, z = s % L, (s // L) % L, s // (L * L)
    return ((x + y + z) & 1) * (N // 2) + (x >> 1) + (L // 2) * (y + L * z)


def run_static(spins_ptr, couplings_ptr, energies_ptr, swaps_ptr, sweep0, n_sweeps, L=16,
               n_samples=64, n_replicas=2, betas=None, T_range=None, H=0.0, spin_seed=2,
               pt_every=1, device=0, stream=None):
    """sg_run_static over device buffers given by address (integers): spins
    [group][replica][temp][N] in checkerboard_order, couplings [group][3][N],
    energies [group][replica][temp][128] int32, swaps [T-1] uint64.  Queues
    the work on `stream` (an address, None = the null stream) and returns."""
    lib = load_library()
    p = _Params()
    keep = _fill_params(p, L, n_samples, n_replicas, betas, T_range, H, 0, spin_seed, pt_every,
                        0, device)
    st = lib.sg_run_static(ctypes.byref(p), spins_ptr, couplings_ptr, int(sweep0), int(n_sweeps),
                           energies_ptr, swaps_ptr, stream)
    del keep
    _check(st, None)


def tune_count():
    """Tuner measurements run in this process so far."""
    return load_library().sg_tune_count()


def probe_valu(kind=0, iters=4096, device=0):
    """VALU issue-rate probe: (lane-operations, ms, lane-ops per second) of
    one launch of 16 independent instruction chains per thread over every CU;
    kind 0 = v_bitop3_b32, 1 = v_mad_u64_u32."""
    lib = load_library()
    ops, ms = ctypes.c_double(), ctypes.c_double()
    st = lib.sg_probe_valu(device, kind, iters, ctypes.byref(ops), ctypes.byref(ms))
    if st != 0:
        raise SGError(f"libsg status {st}: {lib.sg_probe_last_error().decode()}")
    return ops.value, ms.value, ops.value / (ms.value * 1e-3)


# ------------------------------------------------------------------ engine
class Job:
    """Parameters of one engine job (the arguments of Sim, plus n_sweeps)."""

    def __init__(self, L, n_sweeps, n_samples=64, n_replicas=2, betas=None, T_range=None,
                 H=0.0, coupling_seed=0x5EED0001, spin_seed=0x5EED0002, meas_every=0,
                 kernel=SG_KERNEL_AUTO):
        self.L, self.n_sweeps, self.n_samples, self.n_replicas = L, n_sweeps, n_samples, n_replicas
        self.kernel = kernel
        self.betas, self.T_range, self.H = betas, T_range, H
        self.coupling_seed, self.spin_seed, self.meas_every = coupling_seed, spin_seed, meas_every

    @property
    def n_temps(self):
        return len(self.betas) if self.betas is not None else self.T_range[2]

    def params(self, pt_every=0):
        p = _Params()
        keep = _fill_params(p, self.L, self.n_samples, self.n_replicas, self.betas, self.T_range,
                            self.H, self.coupling_seed, self.spin_seed, pt_every, self.meas_every,
                            kernel=self.kernel)
        return p, keep

    def footprint(self):
        p, _keep = self.params()
        return load_library().sg_job_footprint(ctypes.byref(p))

    def flips(self):
        return float(self.L) ** 3 * 64 * ((self.n_samples + 63) // 64) * self.n_replicas * \
            self.n_temps * self.n_sweeps


class Batch:
    """One sg_batch engine: jobs of mixed L resident on one device, admitted
    by footprint, one sweep-kernel launch per (L, field) class and PT
    interval (form SG_FORM_BATCHED) or per job (SG_FORM_PLAIN)."""

    def __init__(self, pt_every=16, slice_pt_steps=4, form=SG_FORM_BATCHED,
                 arena_bytes=4 << 30, max_inflight=4, device=0):
        lib = load_library()
        bp = _BatchParams()
        lib.sg_batch_params_default(ctypes.byref(bp))
        bp.device, bp.pt_every, bp.slice_pt_steps, bp.form = device, pt_every, slice_pt_steps, form
        bp.arena_bytes, bp.max_inflight = arena_bytes, max_inflight
        b = _vp()
        st = lib.sg_batch_create(ctypes.byref(bp), ctypes.byref(b))
        if st != 0:
            msg = lib.sg_batch_last_error(None)
            raise SGError(f"libsg status {st}: {msg.decode() if msg else ''}")
        self._b = b
        self.pt_every = pt_every
        self._temps = {}

    def close(self):
        if getattr(self, "_b", None):
            load_library().sg_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        self.close()

    def _c(self, st):
        if st != 0:
            msg = load_library().sg_batch_last_error(self._b)
            raise SGError(f"libsg status {st}: {msg.decode() if msg else ''}")

    def submit(self, job: Job) -> int:
        p, _keep = job.params(self.pt_every)
        jid = ctypes.c_int64()
        self._c(load_library().sg_submit(self._b, ctypes.byref(p), int(job.n_sweeps),
                                         ctypes.byref(jid)))
        self._temps[jid.value] = job.n_temps
        return jid.value

    def step(self):
        nr, nq = ctypes.c_int32(), ctypes.c_int32()
        self._c(load_library().sg_step(self._b, ctypes.byref(nr), ctypes.byref(nq)))
        return nr.value, nq.value

    def drain(self):
        self._c(load_library().sg_drain(self._b))

    def poll(self, jid: int) -> int:
        s = ctypes.c_int32()
        self._c(load_library().sg_poll(self._b, int(jid), ctypes.byref(s)))
        return s.value

    def times(self, jid: int):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._c(load_library().sg_job_times(self._b, int(jid), ctypes.byref(a), ctypes.byref(b),
                                            ctypes.byref(c)))
        return a.value, b.value, c.value

    def result(self, jid: int):
        """(results [T][NRES], accepted swaps [T-1]); releases the job."""
        T = self._temps[jid]
        r = np.zeros((T, NRES), dtype=np.float64)
        sw = np.zeros(max(T - 1, 1), dtype=np.uint64)
        self._c(load_library().sg_result(self._b, int(jid), r.ctypes.data_as(_dp),
                                         sw.ctypes.data_as(_u64p)))
        del self._temps[jid]
        return r, sw[: T - 1]

    def stats(self):
        fl = ctypes.c_double()
        st, used, cap = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._c(load_library().sg_batch_stats(self._b, ctypes.byref(fl), ctypes.byref(st),
                                              ctypes.byref(used), ctypes.byref(cap)))
        return {"flips": fl.value, "steps": st.value, "arena_in_use": used.value,
                "arena_capacity": cap.value}

    def launch_stats(self):
        """Batched class intervals launched directly and through graph
        replays, and the graphs instantiated."""
        d, g, n = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._c(load_library().sg_batch_launch_stats(self._b, ctypes.byref(d), ctypes.byref(g),
                                                     ctypes.byref(n)))
        return {"intervals_direct": d.value, "intervals_graphed": g.value,
                "graphs_built": n.value}

    def alloc_stats(self):
        """Allocations beyond the arena: task-table growths and pinned
        blocks allocated for final copies (reused after retirement)."""
        g, b = ctypes.c_int64(), ctypes.c_int64()
        self._c(load_library().sg_batch_alloc_stats(self._b, ctypes.byref(g), ctypes.byref(b)))
        return {"table_growths": g.value, "pinned_blocks": b.value}

    def serve(self, jobs, arrivals):
        """Runs the native serving loop over jobs arriving at `arrivals`
        seconds; returns (ids, wall seconds)."""
        n = len(jobs)
        arr = (_Params * n)()
        keep = []
        for i, j in enumerate(jobs):
            p, k = j.params(self.pt_every)
            arr[i] = p
            keep.append(k)
        sweeps = np.ascontiguousarray([j.n_sweeps for j in jobs], dtype=np.int64)
        at = np.ascontiguousarray(arrivals, dtype=np.float64)
        ids = np.zeros(n, dtype=np.int64)
        wall = ctypes.c_double()
        _i64p = ctypes.POINTER(ctypes.c_int64)
        self._c(load_library().sg_serve(self._b, arr, sweeps.ctypes.data_as(_i64p),
                                         at.ctypes.data_as(_dp), n, ids.ctypes.data_as(_i64p),
                                         ctypes.byref(wall)))
        for i, j in zip(ids, jobs):
            self._temps[int(i)] = j.n_temps
        return ids, wall.value


# ------------------------------------------------------------------ drop-in
OBS_DTYPE = np.dtype([("e", "f4", (2,)), ("m", "f4", (2,)), ("q_0", "f4"),
                      ("q_1_r", "f4", (3,)), ("q_1_i", "f4", (3,))])


def obs_records(rec: np.ndarray, L: int, R: int, H: float, n_samples: int) -> np.ndarray:
    """Integer measurement records [group][temp][lane][field] -> the
    reference's float32 obs_s array [sample][temp] (Program/src/eamsim.cuh:
    55-62, eamsim.cu:408-415 and 504-509)."""
    G, T = rec.shape[0], rec.shape[1]
    N = L ** 3
    r = rec.transpose(0, 2, 1, 3).reshape(G * 64, T, -1)[:n_samples].astype(np.float64)
    out = np.zeros((n_samples, T), dtype=OBS_DTYPE)
    U, K = r[..., :R], r[..., R:2 * R]
    M = 2 * K - N
    E = (2 * U - 3 * N) - H * M
    ncp = min(R, 2)
    out["e"][..., :ncp] = (E / N)[..., :ncp]
    out["m"][..., :ncp] = (M / N)[..., :ncp]
    if R >= 2:
        W = r[..., 2 * R]
        out["q_0"] = (N - 2 * W) / N
        # sum_c cos/sin(2 pi c / L) W_d[c] in FP64, c in order (the csg
        # writer's arithmetic, so that both give the same float32 bits)
        cw = [math.cos(2.0 * math.pi * c / L) for c in range(L)]
        sw = [math.sin(2.0 * math.pi * c / L) for c in range(L)]
        for d in range(3):
            Wd = r[..., 2 * R + 1 + d * L: 2 * R + 1 + (d + 1) * L]
            re = np.zeros(Wd.shape[:-1])
            im = np.zeros(Wd.shape[:-1])
            for c in range(L):
                re = re + cw[c] * Wd[..., c]
                im = im + sw[c] * Wd[..., c]
            out["q_1_r"][..., d] = -2.0 / N * re
            out["q_1_i"][..., d] = -2.0 / N * im
    return out


class eamsim:
    """Drop-in for the reference's `eamsim` (Program/src/eamsim.cuh:66-110).

    eamsim(H) fixes the reference's configuration (L=16, 2048 samples, 2
    copies, the 24-temperature ladder, PT every 16 sweeps, a measurement
    every 1024, 262,144 sweeps per run_simulation call); keyword arguments
    override any of them."""

    def __init__(self, H: float, L=16, n_samples=2048, n_replicas=2, betas=REF_BETAS,
                 pt_every=16, meas_every=1024, sweeps_per_file=262144,
                 coupling_seed=0x5EED0001, spin_seed=0x5EED0002, form=SG_FORM_BATCHED):
        if not (0.0 <= H <= 4.0):
            raise SGError("H out of range")
        self.H = float(H)
        self.sweeps_per_file = sweeps_per_file
        self.sim = Sim(L=L, n_samples=n_samples, n_replicas=n_replicas, betas=betas, H=H,
                       coupling_seed=coupling_seed, spin_seed=spin_seed, pt_every=pt_every,
                       meas_every=meas_every, form=form)
        self.sim.keep_records = False

    def init_lattice(self):
        self.sim.init_lattice()

    def save_checkpoint(self, f):
        f.write(self.sim.save_checkpoint_bytes())

    def load_checkpoint(self, f):
        self.sim.load_checkpoint_bytes(f.read())

    def run_simulation(self, f):
        s = self.sim

        def sink(sweep, rec):
            f.write(obs_records(rec, s.L, s.R, s.H, s.n_samples).tobytes())

        s.run(self.sweeps_per_file, sink=sink)
