"""Split-BF16 products of mixed-precision contexts (opt.gemm_split) through
the C ABI, each against the FP64 oracle on the same inputs, with the error
stated per result.

The split rounds each operand to FP32 (2^-24), splits it exactly into three
BF16 terms and sums products in FP32 per K slice (gram) or over the states
(rotate), so:
- Lambda = Re Psi^H F (bit 0, default on in mixed precision; Lambda only
  steers the search direction) and the projected F (bit 3, on request):
  <= 1e-6 relative;
- S = Psi^H Psi: <= 1e-6 relative (on request; the FP32 sums put S some 1e-7
  off, above what ||Psi^H Psi/2 - I|| <= 1e-10 needs);
- Psi T in the update form Psi + Psi (T - I): <= 1e-8 relative (on request);
- whole mixed steps: |dE| <= 1e-6 Ha/atom against the oracle.
"""
import numpy as np
import pytest

import oracle
from openatom_amd import system
from openatom_amd.step import CPContext
from tests.helpers import rel, small_system

pytestmark = pytest.mark.gpu

_cache = {}


def get_system(name):
    if name not in _cache:
        _cache[name] = {
            "water8": lambda: system.build("water8", seed=11, kvec_source=oracle.stored_kvecs),
            "ragged37": lambda: small_system(nstates=37, ecut=18.0, seed=8),
            # more than 128 states: split gram, FP64 projection and rotation
            "states150": lambda: small_system(box=(5.0, 5.2, 5.6), ecut=36.0,
                                              grid=(40, 40, 45), nstates=150, seed=13),
        }[name]()
    return _cache[name]


def make_ctx(s, split, precision="mixed", **kw):
    return CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                     s["kz"], precision=precision, gemm_split=split, **kw)


def make_oracle(s):
    return oracle.System(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                         s["kz"])


def report(tag, **errs):
    print(tag, " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)


def forces(s, seed=5):
    """Force-like coefficients: the oracle's F of the system's V, scaled."""
    o = make_oracle(s)
    F = o.vpsi_r2g_all(o.g2r_all(s["psi"]), s["vks"])
    o.eke_force(s["psi"], F)
    return F


@pytest.mark.parametrize("name", ["water8", "ragged37", "states150"])
def test_split_lambda(gpu, name):
    s = get_system(name)
    o = make_oracle(s)
    F = forces(s)
    lam_o = o.lambda_(s["psi"], F)
    Fo = F.copy()
    o.lambda_apply(s["psi"], lam_o, Fo)
    with make_ctx(s, 9) as c:  # the gram and the projection
        c.set_psi(s["psi"])
        c.set_force(F)
        lam = c.lambda_()
        Fp = c.get_force()
    e_lam, e_f = rel(lam, lam_o), rel(Fp, Fo)
    report(f"[{name} split Lambda]", lam=e_lam, projected_F=e_f)
    assert e_lam < 1e-6 and e_f < 1e-6
    # the FP64 context is untouched by gemm_split
    with make_ctx(s, 15, precision="fp64") as c:
        c.set_psi(s["psi"])
        c.set_force(F)
        assert rel(c.lambda_(), lam_o) < 1e-10


@pytest.mark.parametrize("name", ["water8", "states150"])
def test_split_overlap(gpu, name):
    s = get_system(name)
    o = make_oracle(s)
    So = o.overlap(s["psi"])
    with make_ctx(s, 2) as c:
        c.set_psi(s["psi"])
        S = c.overlap()
    e = rel(S, So)
    report(f"[{name} split S]", S=e, asym=float(np.abs(S - S.T).max()))
    assert e < 1e-6
    assert np.array_equal(S, S.T)  # mirrored from one triangle


@pytest.mark.parametrize("name", ["water8", "ragged37"])
def test_split_rotation_update_form(gpu, name):
    s = get_system(name)
    o = make_oracle(s)
    # a T of the size one step produces: (S/2)^{-1/2} of an SD-updated psi
    psi1 = s["psi"].copy()
    F = forces(s)
    o.lambda_apply(psi1, o.lambda_(psi1, F), F)
    o.sd_update(o.cmass(25.0, 2.0), 5.0, F, psi1)
    o.red_sync(psi1)
    T, it, err = oracle.ns(o.overlap(psi1), tol=1e-15, max_iter=30)
    out_o = o.rotate(psi1, T)
    with make_ctx(s, 4) as c:
        c.set_psi(psi1)
        c.set_tmat(T)
        c.rotate()
        out = c.get_psi()
    e = rel(out, out_o)
    Sg = system.gamma_overlap(out, s["kx"])
    resid = float(np.abs(Sg / 2 - np.eye(s["nstates"])).max())
    report(f"[{name} split Psi T]", psi=e, T_minus_I=float(np.abs(T - np.eye(len(T))).max()),
           resid=resid)
    assert e < 1e-8


@pytest.mark.parametrize("split", [1, 9, 15])
def test_split_steps_energy(gpu, split):
    """Three mixed steps against the oracle's FP64 steps: dE per atom and
    ||Psi^H Psi/2 - I|| stated; Lambda alone keeps the 1e-10 residual."""
    s = get_system("water8")
    o = make_oracle(s)
    psi_o = s["psi"].copy()
    with make_ctx(s, split, dt=5.0) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        for k in range(3):
            r = c.step()
            ro = o.step(s["vks"], psi_o, dt=5.0, ns_tol=1e-15, ns_max_iter=30)
            de = abs(r["eelec"] - ro["eelec"]) / s["natoms"]
            report(f"[water8 gemm_split={split} step {k}]", dE_atom=de, resid=r["resid_max"],
                   psi=rel(c.get_psi(), psi_o))
            assert de < 1e-6
            if split in (1, 9):
                assert r["resid_max"] < 1e-10
            else:
                assert r["resid_max"] < 1e-5
