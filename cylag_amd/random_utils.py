"""Module-level PCG32 stream (cylag/lsm/random_utils.pyx:9-262).

As in the reference, one global stream is shared by every solver of the
process; a Solver loads it onto the device before a step and stores the
advanced state back afterwards, so consecutive runs continue one stream.
The arithmetic runs in libcylag_hip.so's host functions.
"""
import ctypes as C

from . import _hip

_STATE = C.c_uint64(0)
_INC = C.c_uint64(1)
_SEEDED = [False]


def _lib():
    lib = _hip.load()
    if not _SEEDED[0]:
        _SEEDED[0] = True
        lib.cyl_pcg32_seed(C.byref(_STATE), C.byref(_INC), 42, 0xda3e06d11bb2b203)
    return lib


def pcg32_seed(seed, stream):
    _lib().cyl_pcg32_seed(C.byref(_STATE), C.byref(_INC), seed & (2**64 - 1), stream & (2**64 - 1))


def pcg32_advance(delta):
    _lib().cyl_pcg32_advance(C.byref(_STATE), _INC.value, delta & (2**64 - 1))


def get_state():
    """(state, inc) of the global stream (no counterpart in the reference)."""
    _lib()
    return int(_STATE.value), int(_INC.value)


def set_state(state, inc):
    _lib()
    _STATE.value = state & (2**64 - 1)
    _INC.value = inc & (2**64 - 1)


def pcg32_next():
    return _lib().cyl_pcg32_next(C.byref(_STATE), _INC.value)


def random_uniform():
    return pcg32_next() * (1.0 / 4294967296.0)


def random_gaussian():
    return _lib().cyl_random_normal(C.byref(_STATE), _INC.value, 3)


def random_gaussian_pair():
    a, b = C.c_double(), C.c_double()
    _lib().cyl_random_pair(C.byref(_STATE), _INC.value, 3, C.byref(a), C.byref(b))
    return a.value, b.value


def random_gaussian_ziggurat():
    return _lib().cyl_random_normal(C.byref(_STATE), _INC.value, 2)


def random_gaussian_ziggurat_pair():
    return random_gaussian_ziggurat(), random_gaussian_ziggurat()


def generate_random_pair(method):
    if method not in (1, 2, 3):
        raise ValueError("Unknown RNG method")
    a, b = C.c_double(), C.c_double()
    _lib().cyl_random_pair(C.byref(_STATE), _INC.value, method, C.byref(a), C.byref(b))
    return a.value, b.value


def generate_random_single(method):
    if method not in (1, 2, 3):
        raise ValueError("Unknown RNG method")
    return _lib().cyl_random_normal(C.byref(_STATE), _INC.value, method)


