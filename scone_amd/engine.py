"""ctypes driver of libscone_amd.so (the HIP engine behind include/scone_gpu.h).

There is no CPU fallback: if the library or a GPU is missing, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import tables as T

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscone_amd.so")

# exported symbols of include/scone_gpu.h
EXPORTS = ["scn_create", "scn_destroy", "scn_last_error", "scn_device_info", "scn_eigen_cycle",
           "scn_norm_size_repr", "scn_eigen_run", "scn_eigen_start", "scn_eigen_step", "scn_eigen_finish",
           "scn_fission_source", "scn_get_stats", "scn_reset_stats",
           "scn_batch_create", "scn_batch_destroy", "scn_batch_last_error", "scn_batch_add", "scn_batch_cycle",
           "scn_batch_query", "scn_batch_states", "scn_batch_release", "scn_batch_stats",
           "scn_norm_size_repr_noalloc", "scn_batch_set_window", "scn_batch_bin",
           "scn_batch_secondary_bytes"]

SCN_ERR_BUSY = 11


class Stats(C.Structure):
    _fields_ = [("transport_kernel_ms", C.c_double), ("transport_launches", C.c_int64),
                ("n_histories", C.c_int64), ("n_segments", C.c_int64), ("n_collisions", C.c_int64),
                ("n_sites", C.c_int64), ("n_scores", C.c_int64), ("n_cycles", C.c_int64),
                ("cycle_ms", C.c_double)]


class EngineError(RuntimeError):
    pass


_lib = None


def load():
    """Load the HIP library; raises if it was not built (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER(T.Problem)
    PS = C.POINTER(T.ParticleState)
    d3 = C.c_double * 3
    L.scn_last_error.restype = C.c_char_p
    L.scn_last_error.argtypes = [C.c_void_p]
    L.scn_device_info.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_int32)]
    L.scn_create.argtypes = [P, C.c_int32, C.POINTER(C.c_void_p)]
    L.scn_destroy.argtypes = [C.c_void_p]
    L.scn_destroy.restype = None
    L.scn_eigen_cycle.argtypes = [C.c_void_p, PS, C.c_int32, C.c_int64, C.c_double, PS, C.c_int32,
                                  C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.scn_norm_size_repr.argtypes = [C.c_void_p, PS, C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int64]
    L.scn_eigen_run.argtypes = [C.c_void_p, C.POINTER(T.EigenSettings), PS, C.POINTER(T.EigenResult)]
    L.scn_eigen_start.argtypes = [C.c_void_p, C.POINTER(T.EigenSettings), PS]
    L.scn_eigen_step.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
    L.scn_eigen_finish.argtypes = [C.c_void_p, C.POINTER(T.EigenResult)]
    L.scn_fission_source.argtypes = [C.c_void_p, C.c_int32, C.c_int64, d3, d3, C.c_int32, C.c_int32, PS]
    L.scn_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
    L.scn_reset_stats.argtypes = [C.c_void_p]
    L.scn_batch_create.argtypes = [C.POINTER(T.BatchSettings), C.POINTER(C.c_void_p)]
    L.scn_batch_destroy.argtypes = [C.c_void_p]
    L.scn_batch_destroy.restype = None
    L.scn_batch_last_error.restype = C.c_char_p
    L.scn_batch_last_error.argtypes = [C.c_void_p]
    L.scn_batch_add.argtypes = [C.c_void_p, C.POINTER(T.FixedRequest), C.POINTER(C.c_int32)]
    L.scn_batch_cycle.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
    L.scn_batch_query.argtypes = [C.c_void_p, C.c_int32, C.POINTER(T.FixedResult), C.POINTER(C.c_double),
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_int32)]
    L.scn_batch_states.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int32]
    L.scn_batch_release.argtypes = [C.c_void_p, C.c_int32]
    L.scn_batch_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
    L.scn_batch_secondary_bytes.argtypes = [C.c_int32, C.c_int32]
    L.scn_batch_secondary_bytes.restype = C.c_int64
    L.scn_batch_set_window.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    L.scn_batch_bin.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    L.scn_norm_size_repr_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_int64,
                                             C.c_void_p, C.POINTER(C.c_int32)]
    _lib = L
    return L


def device_info():
    L = load()
    buf = C.create_string_buffer(256)
    ncu = C.c_int32()
    rc = L.scn_device_info(buf, 256, C.byref(ncu))
    if rc:
        raise EngineError(f"no GPU visible (code {rc})")
    return buf.value.decode(), ncu.value


def _s64(x: int) -> int:
    x &= (1 << 64) - 1
    return x - (1 << 64) if x >= (1 << 63) else x


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


class Engine:
    """One engine handle: device-resident tables, banks and tallies for one problem."""

    def __init__(self, comp: T.Compiled, capacity: int):
        self.L = load()
        self.comp = comp
        self.capacity = int(capacity)
        h = C.c_void_p()
        rc = self.L.scn_create(C.byref(comp.problem), self.capacity, C.byref(h))
        if rc:
            raise EngineError(f"scn_create failed ({rc}): {self.L.scn_last_error(None).decode()}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.scn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise EngineError(f"engine error {rc}: {self.L.scn_last_error(self.h).decode()}")

    def eigen_cycle(self, bank_in: np.ndarray, rng_state: int, k_eff: float, bins: bool = True):
        n_in = len(bank_in)
        cap = max(16, 4 * n_in)
        out = T.bank_array(cap)
        n_out = C.c_int32()
        nb = int(self.comp.problem.n_bins)
        b = np.zeros(max(1, nb)) if bins else None
        ks = np.zeros(4)
        bank_in = np.ascontiguousarray(bank_in)
        self._check(self.L.scn_eigen_cycle(self.h, T.bank_ptr(bank_in), n_in, _s64(rng_state), k_eff,
                                           T.bank_ptr(out), cap, C.byref(n_out),
                                           b.ctypes.data_as(C.POINTER(C.c_double)) if b is not None else None,
                                           ks.ctypes.data_as(C.POINTER(C.c_double))))
        return out[: n_out.value].copy(), (b[:nb] if b is not None else None), ks

    def norm_size_repr(self, bank: np.ndarray, tot_pop: int, rng_state: int):
        n = len(bank)
        cap = max(4 * tot_pop + 16, n)
        buf = T.bank_array(cap)
        buf[:n] = bank
        nn = C.c_int32(n)
        self._check(self.L.scn_norm_size_repr(self.h, T.bank_ptr(buf), C.byref(nn), cap, tot_pop, _s64(rng_state)))
        return buf[: nn.value].copy()

    def fission_source(self, n: int, rng_state: int, bottom=None, top=None, g_in: int = 1, attempts: int = 10000):
        bank = T.bank_array(n)
        bottom = self.comp.bounds[:3] if bottom is None else bottom
        top = self.comp.bounds[3:] if top is None else top
        self._check(self.L.scn_fission_source(self.h, n, _s64(rng_state), _d3(bottom), _d3(top), g_in, attempts,
                                              T.bank_ptr(bank)))
        return bank

    @staticmethod
    def _settings(settings: dict):
        S = T.EigenSettings()
        S.pop = settings["pop"]
        S.n_inactive = settings["inactive"]
        S.n_active = settings["active"]
        S.inactive_tally = settings.get("inactive_tally", 0)
        S.seed = _s64(settings["seed"])
        S.keff0 = settings.get("keff0", 1.0)
        S.events_per_launch = settings.get("events_per_launch", 0)
        S.block_size = settings.get("block_size", 0)
        S.cycle_graph = settings.get("cycle_graph", 0)
        S.refill = settings.get("refill", 0)
        S.select_window = settings.get("select_window", 0.0)
        return S

    def _result(self, S):
        R = T.EigenResult()
        nb = int(self.comp.problem.n_bins)
        n = S.n_inactive + S.n_active
        arrs = dict(kh=np.zeros(max(1, n)), ct=np.zeros(max(1, n)), ba=np.zeros(max(1, 2 * nb)),
                    bi=np.zeros(max(1, 2 * nb)))
        R.k_history = arrs["kh"].ctypes.data_as(C.POINTER(C.c_double))
        R.cycle_time = arrs["ct"].ctypes.data_as(C.POINTER(C.c_double))
        R.bins_active = arrs["ba"].ctypes.data_as(C.POINTER(C.c_double))
        R.bins_inactive = arrs["bi"].ctypes.data_as(C.POINTER(C.c_double))
        return R, arrs

    def eigen_run(self, settings: dict, initial_bank: np.ndarray):
        S = self._settings(settings)
        R, a = self._result(S)
        nb = int(self.comp.problem.n_bins)
        bank = np.ascontiguousarray(initial_bank)
        self._check(self.L.scn_eigen_run(self.h, C.byref(S), T.bank_ptr(bank), C.byref(R)))
        n = S.n_inactive + S.n_active
        self.last_cycle_time = a["ct"][:n].copy()
        return R, a["kh"][:n], a["ba"][: 2 * nb].reshape(-1, 2), a["bi"][: 2 * nb].reshape(-1, 2)

    def eigen_start(self, settings: dict, initial_bank: np.ndarray):
        self._S = self._settings(settings)
        bank = np.ascontiguousarray(initial_bank)
        self._check(self.L.scn_eigen_start(self.h, C.byref(self._S), T.bank_ptr(bank)))

    def eigen_step(self, n_cycles: int) -> int:
        done = C.c_int32()
        self._check(self.L.scn_eigen_step(self.h, n_cycles, C.byref(done)))
        return done.value

    def eigen_finish(self):
        S = self._S
        R, a = self._result(S)
        nb = int(self.comp.problem.n_bins)
        self._check(self.L.scn_eigen_finish(self.h, C.byref(R)))
        n = S.n_inactive + S.n_active
        return R, a["kh"][:n], a["ct"][:n], a["ba"][: 2 * nb].reshape(-1, 2), a["bi"][: 2 * nb].reshape(-1, 2)

    def stats(self) -> dict:
        s = Stats()
        self._check(self.L.scn_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def reset_stats(self):
        self._check(self.L.scn_reset_stats(self.h))


class Busy(EngineError):
    """The batch engine has no free slot or budget for a request right now."""


def secondary_bytes(transport=T.TRANSPORT_DT, buffer=0) -> int:
    """HBM-budget bytes of the per-lane buffers of secondaries that a batch engine reserves once a
    solve with fission or weight windows is resident (`buffer` entries per lane, 0 = 50)."""
    n = load().scn_batch_secondary_bytes(transport, buffer)
    if n < 0:
        raise EngineError("scn_batch_secondary_bytes: no GPU visible to HIP")
    return int(n)


class Batch:
    """Fixed-source engine with resident problem slots (scn_batch_* of include/scone_gpu.h)."""

    def __init__(self, transport=T.TRANSPORT_DT, max_slots=512, hbm_budget=8 << 30, tile=0, k0=0, n_launch=0,
                 use_graph=1, refill=0):
        L = load()
        S = T.BatchSettings()
        S.transport = transport
        S.max_slots = max_slots
        S.hbm_budget = int(hbm_budget)
        S.tile = tile
        S.k0 = k0
        S.n_launch = n_launch
        S.use_graph = use_graph
        S.refill = refill
        self.settings = S
        h = C.c_void_p()
        rc = L.scn_batch_create(C.byref(S), C.byref(h))
        if rc:
            raise EngineError(f"scn_batch_create failed ({rc}): {L.scn_batch_last_error(None).decode()}")
        self.h = h
        self.L = L
        self._held = {}

    def close(self):
        if self.h:
            self.L.scn_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            msg = self.L.scn_batch_last_error(self.h).decode()
            if rc == SCN_ERR_BUSY:
                raise Busy(msg)
            raise EngineError(f"batch engine error {rc}: {msg}")

    def add(self, comp, src, pop, n_batches, seed, stop_bin=-1, stop_rel=0.0, record=False, buffer=0,
            window=1, adapt=0) -> int:
        """Admit a solve; raises Busy if it does not fit now.  Returns its slot.  `window` batches of
        the solve are transported together in each cycle (closed in order: same results); with
        `adapt` (1-100) and a stop bin, each cycle transports only the batches the stop bin's trend
        still asks for (at most `window`)."""
        R = T.FixedRequest()
        R.problem = C.pointer(comp.problem)
        R.source = C.pointer(src.source)
        R.pop = pop
        R.n_batches = n_batches
        R.seed = _s64(seed)
        R.stop_bin = stop_bin
        R.stop_rel = stop_rel
        R.record = 1 if record else 0
        R.buffer = buffer
        R.window = window
        R.adapt = adapt
        slot = C.c_int32(-1)
        self._check(self.L.scn_batch_add(self.h, C.byref(R), C.byref(slot)))
        self._held[slot.value] = (comp, src, pop, record)
        return slot.value

    def cycle(self, n=1) -> int:
        """Run up to n cycles; returns the number of slots still running."""
        nr = C.c_int32()
        self._check(self.L.scn_batch_cycle(self.h, n, C.byref(nr)))
        return nr.value

    def query(self, slot):
        """(FixedResult, bins (n_bins, 2) or None, hist_rng or None, hist_coll or None)."""
        r = T.FixedResult()
        held = self._held.get(slot)
        if held is not None and not held[3]:
            # one call: the bins come back with the state (dropped while the solve runs)
            nb = int(held[0].problem.n_bins)
            bins = np.zeros(max(1, 2 * nb))
            self._check(self.L.scn_batch_query(self.h, slot, C.byref(r), bins.ctypes.data_as(C.POINTER(C.c_double)),
                                               None, None))
            if r.state != T.SLOT_DONE:
                return r, None, None, None
            return r, bins[: 2 * int(r.n_bins)].reshape(-1, 2), None, None
        self._check(self.L.scn_batch_query(self.h, slot, C.byref(r), None, None, None))
        if r.state != T.SLOT_DONE:
            return r, None, None, None
        nb = int(r.n_bins)
        bins = np.zeros(max(1, 2 * nb))
        comp, src, pop, record = self._held[slot]
        nh = r.n_batches * pop if record else 0
        hr = np.zeros(max(1, nh), dtype=np.uint64) if record else None
        hc = np.zeros(max(1, nh), dtype=np.int32) if record else None
        self._check(self.L.scn_batch_query(
            self.h, slot, C.byref(r), bins.ctypes.data_as(C.POINTER(C.c_double)),
            hr.ctypes.data_as(C.POINTER(C.c_uint64)) if hr is not None else None,
            hc.ctypes.data_as(C.POINTER(C.c_int32)) if hc is not None else None))
        return r, bins[: 2 * nb].reshape(-1, 2), (hr[:nh] if hr is not None else None), \
            (hc[:nh] if hc is not None else None)

    def states(self) -> np.ndarray:
        """SLOT_* of every slot."""
        n = self.settings.max_slots
        out = np.zeros(n, dtype=np.int32)
        self._check(self.L.scn_batch_states(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), n))
        return out

    def set_window(self, slot, window):
        """Batches of a running solve transported per cycle from now on (1 .. its admitted window)."""
        self._check(self.L.scn_batch_set_window(self.h, slot, window))

    def bin_result(self, slot, b):
        """(mean, sd, batches closed) of bin b of a running or done solve."""
        out = (C.c_double * 2)()
        n = C.c_int32()
        self._check(self.L.scn_batch_bin(self.h, slot, b, out, C.byref(n)))
        return out[0], out[1], n.value

    def release(self, slot):
        self._check(self.L.scn_batch_release(self.h, slot))
        self._held.pop(slot, None)

    def stats(self) -> Stats:
        s = Stats()
        self._check(self.L.scn_batch_stats(self.h, C.byref(s)))
        return s

    def run_all(self, max_cycles=1 << 30):
        """Cycle until no slot is running."""
        n = 0
        while n < max_cycles:
            if self.cycle(1) == 0:
                break
            n += 1
        return n
