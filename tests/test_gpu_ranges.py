"""State-range entry points (oa_fft_g2r_states, oa_vpsi_r2g_states): the
per-state-plane granularity of the reference's doStpFFTGtoR/RtoG calls.

Covering the states in uneven ranges gives the density of the all-state
call (up to the order of the chunk sums, 1e-13), and each range's forces
equal those of the all-state call bit for bit (the per-state arithmetic is
the same); the EKE shares add up to the total.  Both are also checked
against the oracle at 1e-10.  Bad ranges and engines without ranges are
rejected."""
import numpy as np
import pytest

import oracle
from openatom_amd import _lib, system
from openatom_amd.step import CPContext
from tests.helpers import rel, small_system

pytestmark = pytest.mark.gpu

_cache = {}


def get_system(name):
    if name not in _cache:
        _cache[name] = {
            "water8": lambda: system.build("water8", seed=21, kvec_source=oracle.stored_kvecs),
            "square27": lambda: small_system(box=(5.0, 5.2, 5.8), grid=(27, 27, 30), nstates=19,
                                             seed=6),
        }[name]()
    return _cache[name]


def ctx(s, precision="fp64", **kw):
    return CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"],
                     precision=precision, **kw)


@pytest.mark.parametrize("name,precision", [("water8", "fp64"), ("square27", "fp64"),
                                            ("water8", "mixed")])
def test_ranges_equal_all_states(gpu, name, precision):
    s = get_system(name)
    n = s["nstates"]
    ranges = [(0, 5), (5, n - 12), (n - 7, 7)]
    assert sum(r[1] for r in ranges) == n
    with ctx(s, precision) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        c.fft_g2r()
        rho_all = c.get_density()
        eke_all = c.vpsi_r2g()
        f_all = c.get_force()
        for i, (s0, ns) in enumerate(ranges):
            c.fft_g2r_states(s0, ns, first=(i == 0))
        rho_r = c.get_density()
        eke_r = 0.0
        for s0, ns in ranges[::-1]:
            eke_r += c.vpsi_r2g_states(s0, ns)
        f_r = c.get_force()
    o = oracle.System(s["grid"], s["hmat"], s["ecut_ry"], n, s["kx"], s["ky"], s["kz"])
    pr = o.g2r_all(s["psi"])
    rho_o = o.density(pr)
    f_o = o.vpsi_r2g_all(pr, s["vks"])
    eke_o = o.eke_force(s["psi"], f_o)
    tol = 1e-10 if precision == "fp64" else 1e-5
    print(f"[{name} {precision}] rho ranges vs all {rel(rho_r, rho_all):.1e}, vs oracle "
          f"{rel(rho_r, rho_o):.1e}; F bitwise {np.array_equal(f_r, f_all)}, vs oracle "
          f"{rel(f_r, f_o):.1e}; EKE {eke_r:.12f} all {eke_all:.12f} oracle {eke_o:.12f}",
          flush=True)
    assert rel(rho_r, rho_all) < 1e-13
    assert np.array_equal(f_r, f_all)
    assert abs(eke_r - eke_all) <= 1e-13 * abs(eke_all)
    assert rel(rho_r, rho_o) < tol and rel(f_r, f_o) < tol
    assert abs(eke_r - eke_o) <= tol * abs(eke_o)


def test_bad_ranges_rejected(gpu):
    s = get_system("water8")
    n = s["nstates"]
    with ctx(s) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        for s0, ns in ((-1, 2), (0, 0), (n - 1, 2), (n, 1)):
            with pytest.raises(_lib.OaError):
                c.fft_g2r_states(s0, ns)
            with pytest.raises(_lib.OaError):
                c.vpsi_r2g_states(s0, ns)


def test_ranges_need_the_plane_engine(gpu):
    s = get_system("water8")
    with ctx(s, fft_mode=1) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        with pytest.raises(_lib.OaError):
            c.fft_g2r_states(0, 4)
