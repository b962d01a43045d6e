"""State FFT passes pruned to the sphere (fft_full = 0, the default) against
the full transforms (fft_full = 1).

Along each axis where every stored |k| is <= (N - 1) / 4, the two-factor z
passes and the fused plane passes skip the radix-pass inputs that are zero
outside the sphere and the outputs that are not kept.  Only additions of
exact zeros and unused values go, so the density, forces and kinetic energy
agree with the full transforms to rounding (1e-13 relative in FP64, 1e-6
mixed), and both agree with the oracle (1e-10 FP64, 1e-5 mixed).  The
cases: water8 (48^3, |k| <= 10), a serving shape (5 H2O at 70 Ry, 54^3 =
9 x 6), water8's sphere on a 40^3 grid, which does not satisfy the bound
(both settings then run the full transforms), and a 108^3 serving shape
whose FP64 plane passes run in 512-thread workgroups."""
import numpy as np
import pytest

import oracle
from openatom_amd import system
from openatom_amd.step import CPContext
from tests.helpers import rel

pytestmark = pytest.mark.gpu

_cache = {}


def get_system(name):
    if name not in _cache:
        if name == "water8":
            s = system.build("water8", seed=21, kvec_source=oracle.stored_kvecs)
        elif name == "serve5":
            s = system.build(system.serving_workload(5, 70.0), seed=8,
                             kvec_source=oracle.stored_kvecs)
        elif name == "wide108":  # 40 H2O at 70 Ry: 108^3, FP64 planes in 512-thread workgroups
            s = system.build(system.serving_workload(40, 70.0), seed=9, nstates=24,
                             kvec_source=oracle.stored_kvecs)
        else:  # water8's sphere on a 40^3 grid (4 kmax + 1 = 41 > 40)
            s = system.build(dict(system.WORKLOADS["water8"], grid=40), seed=21,
                             kvec_source=oracle.stored_kvecs)
        _cache[name] = s
    return _cache[name]


def run(s, precision, fft_full):
    with CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"],
                   precision=precision, fft_full=fft_full) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        c.fft_g2r()
        rho = c.get_density()
        eke = c.vpsi_r2g()
        f = c.get_force()
    return rho, f, eke


@pytest.mark.parametrize("name", ["water8", "serve5", "water8_grid40", "wide108"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pruned_equals_full(gpu, name, precision):
    s = get_system(name)
    kabs = max(int(np.max(np.abs(s[k]))) for k in ("kx", "ky", "kz"))
    pruned = 4 * kabs + 1 <= min(s["grid"][:2])
    assert pruned == (name != "water8_grid40")
    rho_p, f_p, eke_p = run(s, precision, 0)
    rho_f, f_f, eke_f = run(s, precision, 1)
    o = oracle.System(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"])
    pr = o.g2r_all(s["psi"])
    rho_o = o.density(pr)
    f_o = o.vpsi_r2g_all(pr, s["vks"])
    eke_o = o.eke_force(s["psi"], f_o)
    same = 1e-13 if precision == "fp64" else 1e-6
    tol = 1e-10 if precision == "fp64" else 1e-5
    print(f"[{name} {precision} pruned={pruned}] pruned vs full: rho {rel(rho_p, rho_f):.1e} "
          f"F {rel(f_p, f_f):.1e}; vs oracle: rho {rel(rho_p, rho_o):.1e} F {rel(f_p, f_o):.1e} "
          f"EKE {abs(eke_p - eke_o) / abs(eke_o):.1e}", flush=True)
    assert rel(rho_p, rho_f) < same and rel(f_p, f_f) < same
    assert abs(eke_p - eke_f) <= same * abs(eke_f)
    for rho, f, eke in ((rho_p, f_p, eke_p), (rho_f, f_f, eke_f)):
        assert rel(rho, rho_o) < tol and rel(f, f_o) < tol
        assert abs(eke - eke_o) <= tol * abs(eke_o)
