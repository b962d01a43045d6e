"""The reference's STATES files (include/oa_state.h) from Python: one state
at a time, in the stored k-vector order of the step library.

Encodings: 0 text, 1 binary, 2 text gzip, 3 binary gzip (".gz" is appended
to the path for 2 and 3, as the reference does)."""
import ctypes

import numpy as np

from . import _lib
from ._lib import OaError, ptr

TEXT, BINARY, TEXT_GZ, BINARY_GZ = range(4)


def _k(kx, ky, kz):
    return [np.ascontiguousarray(a, np.int32) for a in (kx, ky, kz)]


def write_state(path, psi, kx, ky, kz, grid, fmt=BINARY):
    """Write one state psi[npts] (complex, stored order kx/ky/kz)."""
    L = _lib.load()
    k = _k(kx, ky, kz)
    a = np.ascontiguousarray(psi, np.complex128)
    g = np.ascontiguousarray(grid, np.int32)
    st = L.oa_state_write(str(path).encode(), int(fmt), int(a.size), ptr(k[0]), ptr(k[1]),
                          ptr(k[2]), ptr(a), ptr(g))
    if st != _lib.OA_OK:
        raise OaError(st, f"oa_state_write({path})")


def read_state(path, kx, ky, kz, fmt=BINARY):
    """(psi[npts] in the stored order kx/ky/kz, header grid)."""
    L = _lib.load()
    k = _k(kx, ky, kz)
    psi = np.zeros(k[0].size, np.complex128)
    g = np.zeros(3, np.int32)
    st = L.oa_state_read(str(path).encode(), int(fmt), int(psi.size), ptr(k[0]), ptr(k[1]),
                         ptr(k[2]), ptr(psi), ptr(g))
    if st != _lib.OA_OK:
        raise OaError(st, f"oa_state_read({path})")
    return psi, tuple(int(v) for v in g)
