"""oa_ortho_transform: the stateless CPORTHOG::CP_orthoTransform (T of a
caller's S, no context) against the oracle's Newton-Schulz and the closed
form (S/2)^{-1/2} by eigendecomposition (the reference's get_diag_Tmat)."""
import numpy as np
import pytest

import oracle
from openatom_amd import _lib

pytestmark = pytest.mark.gpu

TOL = 1e-10


def overlap_like(n, seed, eps=3e-3):
    """A symmetric S near 2I (reference scaling) with |S/2 - I| ~ eps."""
    rng = np.random.default_rng(seed)
    E = rng.standard_normal((n, n))
    E = (E + E.T) * (eps / np.sqrt(n))
    return 2.0 * (np.eye(n) + E)


@pytest.mark.parametrize("n", [1, 16, 37, 128, 200, 512])
def test_ortho_transform_matches_oracle(gpu, n):
    S = overlap_like(n, seed=n)
    T, st = _lib.ortho_transform(S)
    To, it, err = oracle.ns(S, tol=1e-15, max_iter=10)
    w, U = np.linalg.eigh(S / 2.0)
    Tc = (U * (1.0 / np.sqrt(w))) @ U.T
    e_or = np.abs(T - To).max() / np.abs(To).max()
    e_cf = np.abs(T - Tc).max() / np.abs(Tc).max()
    ortho = np.abs(T @ (S / 2.0) @ T - np.eye(n)).max()
    print(f"[ortho_transform n={n}] it={it} vs oracle {e_or:.2e} vs eigh {e_cf:.2e} "
          f"|T S/2 T - I| {ortho:.2e}", flush=True)
    assert st == _lib.OA_OK and it > 0
    assert e_or < TOL and e_cf < TOL and ortho < 1e-12


def test_ortho_transform_reports_nonconvergence(gpu):
    """With an eigenvalue of S/2 at 1e-3, ten iterations miss 1e-15: OA_ENOCONV where the
    reference EXITs (cp_lowdin.C:247-253), as the oracle reports too."""
    rng = np.random.default_rng(3)
    U, _ = np.linalg.qr(rng.standard_normal((64, 64)))
    S = 2.0 * (U * np.linspace(1e-3, 1.9, 64)) @ U.T
    T, st = _lib.ortho_transform(S)
    To, it, err = oracle.ns(S, tol=1e-15, max_iter=10)
    assert it < 0
    assert st == _lib.OA_ENOCONV
