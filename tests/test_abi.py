"""CPU checks of the C ABI boundary: the library builds, loads and exports
every symbol include/*.h declares; its host-side k-vector enumeration
(no GPU involved) matches the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from openatom_amd import _lib, system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            src = open(os.path.join(ROOT, "include", h)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            syms |= set(re.findall(r"\b(oa_[a-z0-9_]+)\s*\(", src))
    return sorted(syms)


def test_header_declares_the_binding_table():
    assert set(declared_symbols()) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_default_options():
    o = _lib.default_options()
    assert o.precision == _lib.OA_FP64 and o.ns_max_iter == 30 and o.ns_tol == 1e-15
    assert o.use_lambda == 1 and o.dt == 5.0


@pytest.mark.parametrize("name", ["water8", "water32", "water128"])
def test_library_kvecs_match_oracle(name):
    w = system.WORKLOADS[name]
    h = system.cubic_hmat(w["box_ang"] * system.BOHR_PER_ANG)
    a = _lib.kvecs(w["ecut_ry"], h)
    b = oracle.stored_kvecs(w["ecut_ry"], h)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_library_kvecs_orthorhombic():
    h = np.diag([b * system.BOHR_PER_ANG for b in (5.0, 5.4, 5.8)]).reshape(9)
    for x, y in zip(_lib.kvecs(14.0, h), oracle.stored_kvecs(14.0, h)):
        assert np.array_equal(x, y)
