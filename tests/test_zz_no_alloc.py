"""oa_ortho_transform_noalloc: the allocation-free T = (S/2)^{-1/2} against
oa_ortho_transform (same kernels, so the same bits) and the closed form, on
two calls in a row with the second smaller, and its size limit."""
import numpy as np
import pytest

from openatom_amd import _lib

pytestmark = pytest.mark.gpu


def overlap_like(n, seed, eps=3e-3):
    """A symmetric S near 2I (reference scaling) with |S/2 - I| ~ eps."""
    rng = np.random.default_rng(seed)
    E = rng.standard_normal((n, n))
    E = (E + E.T) * (eps / np.sqrt(n))
    return 2.0 * (np.eye(n) + E)


def test_noalloc_matches_allocating_entry_point(gpu):
    for n in (512, 200, 37):  # largest allowed first, then smaller inputs
        S = overlap_like(n, seed=100 + n)
        T, st = _lib.ortho_transform_noalloc(S)
        Tr, str_ = _lib.ortho_transform(S)
        w, U = np.linalg.eigh(S / 2.0)
        Tc = (U * (1.0 / np.sqrt(w))) @ U.T
        e_cf = np.abs(T - Tc).max() / np.abs(Tc).max()
        print(f"[noalloc n={n}] status {st} vs allocating: max|dT| "
              f"{np.abs(T - Tr).max():.1e}, vs eigh {e_cf:.2e}", flush=True)
        assert st == _lib.OA_OK and str_ == _lib.OA_OK
        assert np.array_equal(T, Tr)
        assert e_cf < 1e-10


def test_noalloc_reports_nonconvergence_like_allocating(gpu):
    rng = np.random.default_rng(3)
    U, _ = np.linalg.qr(rng.standard_normal((64, 64)))
    S = 2.0 * (U * np.linspace(1e-3, 1.9, 64)) @ U.T
    T, st = _lib.ortho_transform_noalloc(S)
    Tr, str_ = _lib.ortho_transform(S)
    assert st == str_ == _lib.OA_ENOCONV
    assert np.array_equal(T, Tr)


def test_noalloc_rejects_sizes_beyond_its_array(gpu):
    n = _lib.ORTHO_NOALLOC_MAX_N + 1
    S = 2.0 * np.eye(n)
    T = np.zeros_like(S)
    st = _lib.load().oa_ortho_transform_noalloc(_lib.ptr(T), _lib.ptr(S), n)
    assert st == _lib.OA_EINVAL
    assert not T.any()
