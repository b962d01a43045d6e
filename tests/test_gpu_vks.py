"""V_KS from rho on the device (vks_from_rho contexts) against the oracle:
PZ-LDA exchange-correlation, Hartree and the local pseudopotential of fixed
ions, built from each step's own density.  FP64 bar: 1e-10 relative on
V(r), on E_H, E_xc, E_ext and on whole steps."""
import numpy as np
import pytest

import oracle
from openatom_amd import OaError, system
from openatom_amd.step import CPContext
from tests.helpers import oracle_pseudo, rel, small_system

pytestmark = pytest.mark.gpu

TOL = 1e-10

_cache = {}


def get_system(name):
    if name not in _cache:
        _cache[name] = {
            # orthorhombic, unequal grid (odd Ny): two-factor per-pass engine
            "small": lambda: small_system(),
            # square x-y planes: the fused plane engine
            "square27": lambda: small_system(box=(5.0, 5.2, 5.8), grid=(27, 27, 30),
                                             nstates=19, seed=5),
            "water8": lambda: system.build("water8", seed=11, kvec_source=oracle.stored_kvecs),
        }[name]()
    return _cache[name]


def make_ctx(s, precision="fp64", **kw):
    c = CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"],
                  precision=precision, vks_from_rho=1, **kw)
    t = s["pseudo"]
    c.set_local_pseudo(s["pos"], s["species"], t["gmin"], t["dg"], t["vps"], t["vps_g0"])
    return c


def oracle_vloc(s):
    o = oracle.System(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                      s["kz"])
    lp, keep = oracle_pseudo(s)
    return o, o.vloc_g(lp)


def report(tag, **errs):
    print(tag, " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)


@pytest.mark.parametrize("mode", [0, 1, 3])
@pytest.mark.parametrize("name", ["small", "square27", "water8"])
def test_vks_from_density(gpu, name, mode):
    """oa_fft_g2r of a vks_from_rho context leaves V_KS[rho] and its
    energies; the oracle builds them from the oracle's own rho."""
    s = get_system(name)
    o, vl = oracle_vloc(s)
    rho = o.density(o.g2r_all(s["psi"]))
    v_ref, e_ref = o.vks(rho, vl)
    with make_ctx(s, fft_mode=mode) as c:
        c.set_psi(s["psi"])
        c.fft_g2r()
        v = c.get_vks()
        r = c.result()
        c.vks_build()  # rebuilding from the same rho gives the same V
        v2 = c.get_vks()
    errs = dict(v=rel(v, v_ref), ehart=rel(r["ehart"], e_ref["ehart"]),
                exc=rel(r["exc"], e_ref["exc"]), eext=rel(r["eloc"], e_ref["eext"]))
    report(f"[{name} mode{mode}] vks", **errs)
    assert all(e < TOL for e in errs.values())
    assert np.array_equal(v, v2)


@pytest.mark.parametrize("name", ["small", "water8"])
def test_vks_steps_match_oracle(gpu, name):
    """Three self-consistent steepest-descent steps (V rebuilt from each
    step's rho) through the C ABI against the oracle's."""
    s = get_system(name)
    o, vl = oracle_vloc(s)
    psi_o = s["psi"].copy()
    with make_ctx(s) as c:
        c.set_psi(s["psi"])
        for k in range(3):
            g = c.step()
            r = o.step(None, psi_o, dt=5.0, ns_tol=1e-15, ns_max_iter=30, vloc=vl)
            errs = {f: rel(g[f], r[f]) for f in ("eelec", "eke", "eloc", "ehart", "exc")}
            report(f"[{name} step {k}]", **errs)
            assert all(e < TOL for e in errs.values())
        e_psi = rel(c.get_psi(), psi_o)
    report(f"[{name}] psi after 3 steps", psi=e_psi)
    assert e_psi < TOL


@pytest.mark.parametrize("name", ["square27", "water8"])
def test_vks_autotuned_tiling_matches_oracle(gpu, name):
    """opt.autotune also times the V_KS passes' row pairs, columns and z
    lines per workgroup (2, 4, 8, 16 where the LDS fits); whatever it picks,
    V_KS, its energies and three steps match the oracle at 1e-10."""
    s = get_system(name)
    o, vl = oracle_vloc(s)
    rho = o.density(o.g2r_all(s["psi"]))
    v_ref, e_ref = o.vks(rho, vl)
    psi_o = s["psi"].copy()
    with make_ctx(s, autotune=1) as c:
        log = c.tuning_log()
        print(f"[{name} autotune]\n{log}", flush=True)
        for knob in ("V_KS row pairs", "V_KS columns", "V_KS z lines"):
            assert knob in log
        c.set_psi(s["psi"])
        c.fft_g2r()
        v = c.get_vks()
        r = c.result()
        errs = dict(v=rel(v, v_ref), ehart=rel(r["ehart"], e_ref["ehart"]),
                    exc=rel(r["exc"], e_ref["exc"]))
        c.set_psi(s["psi"])
        for k in range(3):
            g = c.step()
            ro = o.step(None, psi_o, dt=5.0, ns_tol=1e-15, ns_max_iter=30, vloc=vl)
            errs[f"eelec{k}"] = rel(g["eelec"], ro["eelec"])
        errs["psi"] = rel(c.get_psi(), psi_o)
    report(f"[{name} autotuned V_KS]", **errs)
    assert all(e < TOL for e in errs.values())


def test_vks_mixed_precision_error(gpu):
    """FP32 FFT chain, V_KS built in FP64 from the FP32-derived rho: the
    energies after three steps stay within 1e-6 Ha/atom of the oracle."""
    s = get_system("water8")
    o, vl = oracle_vloc(s)
    psi_o = s["psi"].copy()
    with make_ctx(s, precision="mixed") as c:
        c.set_psi(s["psi"])
        for k in range(3):
            g = c.step()
            r = o.step(None, psi_o, dt=5.0, ns_tol=1e-15, ns_max_iter=30, vloc=vl)
    de = abs(g["eelec"] - r["eelec"]) / s["natoms"]
    report("[water8 mixed] vks steps", dE_per_atom=de, ehart=rel(g["ehart"], r["ehart"]),
           exc=rel(g["exc"], r["exc"]))
    assert de < 1e-6


def test_vks_graph_replay_equals_eager(gpu):
    s = get_system("water8")
    out = {}
    for graph in (0, 1):
        with make_ctx(s, use_graph=graph) as c:
            c.set_psi(s["psi"])
            out[graph] = [c.step()["eelec"] for _ in range(4)] + [c.get_psi()]
    assert out[0][:4] == out[1][:4]
    assert np.array_equal(out[0][4], out[1][4])


def test_vks_mode_errors(gpu):
    s = get_system("small")
    with make_ctx(s) as c:
        with pytest.raises(OaError):
            c.set_vks(s["vks"])  # V comes from rho in this context
    with CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                   s["kz"]) as c:
        with pytest.raises(OaError):
            c.vks_build()  # not a vks_from_rho context
    # a grid that cannot hold the density sphere |G|^2 <= 4 ecut
    with pytest.raises(OaError):
        CPContext((18, 27, 30), s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"],
                  vks_from_rho=1)
