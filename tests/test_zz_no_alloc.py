"""scn_norm_size_repr_noalloc (population control with its scratch in a static device array)
against scn_norm_size_repr through the C ABI: bit-identical banks on consecutive calls, the second
with fewer sites, in every excess regime (drop, duplicate with and without whole copies, none).
The banks' device memory comes from the HIP runtime the library itself links (hipMalloc through
ctypes, in the test); the library call allocates nothing."""
import ctypes as C

import numpy as np
import pytest

import binding as O
from helpers import assert_banks_identical, problem
from scone_amd import tables as T
from scone_amd.engine import Engine, load

pytestmark = pytest.mark.gpu

REC = C.sizeof(T.ParticleState)
H2D, D2H = 1, 2


def _hip():
    load()   # the library's HIP runtime is then the one already mapped into the process
    H = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [C.c_void_p]
    return H


class DevBuf:
    def __init__(self, H, nbytes):
        self.H = H
        self.p = C.c_void_p()
        assert H.hipMalloc(C.byref(self.p), max(1, nbytes)) == 0
        self.n = nbytes

    def put(self, arr):
        raw = np.ascontiguousarray(arr)
        assert raw.nbytes <= self.n
        assert self.H.hipMemcpy(self.p, raw.ctypes.data, raw.nbytes, H2D) == 0

    def get(self, nbytes):
        out = np.zeros(nbytes, dtype=np.uint8)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, nbytes, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.p)


def _noalloc(bank: np.ndarray, tot: int, rng: int):
    L = load()
    H = _hip()
    n = len(bank)
    cap = max(tot, 1)
    d_in = DevBuf(H, n * REC)
    d_out = DevBuf(H, cap * REC)
    d_in.put(bank)
    n_out = C.c_int32(-1)
    print(f"noalloc call n={n} tot={tot}", flush=True)
    rc = L.scn_norm_size_repr_noalloc(d_in.p, n, d_out.p, cap, tot, rng, None, C.byref(n_out))
    assert rc == 0, L.scn_last_error(None)
    out = d_out.get(n_out.value * REC).view(bank.dtype)
    d_in.free()
    d_out.free()
    return out


@pytest.mark.parametrize("excess", [900, -700, -2500, 0])
def test_noalloc_matches_engine_on_two_calls(excess):
    inp, mg, comp = problem("inf_medium_2g", pop=2000)
    eng = Engine(comp, 4000)
    bank = O.fission_source(comp, 2000, 99)
    sites, _, _, _ = O.eigen_cycle(comp, bank, 1234, 1.0)
    # first call: every site; second: a shorter brood-sorted prefix, same excess sign
    for n in (len(sites), len(sites) * 3 // 5):
        part = sites[:n]
        tot = max(1, n - excess if excess >= 0 else n - excess * n // len(sites))
        ref = eng.norm_size_repr(part, tot, 777 + n)
        got = _noalloc(part, tot, 777 + n)
        assert len(ref) == tot
        assert_banks_identical(ref, got)


def test_noalloc_rejects_sizes_beyond_the_scratch():
    L = load()
    n_out = C.c_int32(0)
    # the size check returns before any device access, so the (unused) pointers need not be valid
    rc = L.scn_norm_size_repr_noalloc(C.c_void_p(16), (1 << 20) + 1, C.c_void_p(16), 10, 10, 1, None,
                                      C.byref(n_out))
    assert rc == 1
    assert b"SCN_NOALLOC_MAX_SITES" in L.scn_last_error(None)


def test_noalloc_rejects_unsorted_broods():
    L = load()
    H = _hip()
    inp, mg, comp = problem("inf_medium_2g", pop=500)
    bank = O.fission_source(comp, 500, 5)
    sites, _, _, _ = O.eigen_cycle(comp, bank, 77, 1.0)
    shuffled = sites[::-1].copy()
    d_in = DevBuf(H, len(sites) * REC)
    d_out = DevBuf(H, len(sites) * REC)
    d_in.put(shuffled)
    n_out = C.c_int32(0)
    rc = L.scn_norm_size_repr_noalloc(d_in.p, len(sites), d_out.p, len(sites), len(sites) - 10, 3, None,
                                      C.byref(n_out))
    d_in.free()
    d_out.free()
    assert rc == 1
    assert b"sortByBroodID" in L.scn_last_error(None)
