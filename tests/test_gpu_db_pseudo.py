"""The 8-water workload with the reference's own pseudopotential database
(data/DATABASE O_BLYP_PP30 and H_BLYP_PP30 through pi_md.vps: hydrogen local,
oxygen local p channel plus the Kleinman-Bylander s channel), FP64 on the
GPU against the oracle at 1e-10: one step stage by stage (rho, V_KS, S, T,
coefficients and every energy term, E_NL included) and three whole steps."""
import numpy as np
import pytest

import oracle
from openatom_amd import system
from openatom_amd.step import CPContext
from tests.helpers import oracle_nonlocal, oracle_pseudo, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10
_cache = {}


def water8_db():
    if "w" not in _cache:
        _cache["w"] = system.build("water8", seed=11, pseudo="db",
                                   kvec_source=oracle.stored_kvecs)
    return _cache["w"]


def make_ctx(s, **kw):
    c = CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"],
                  vks_from_rho=1, dt=5.0, ns_tol=1e-15, ns_max_iter=30, **kw)
    t = s["pseudo"]
    c.set_local_pseudo(s["pos"], s["species"], t["gmin"], t["dg"], t["vps"], t["vps_g0"])
    k = s["kb"]
    c.set_nonlocal(k["pos"], k["species"], k["gmin"], k["dg"], k["proj"], k["proj_g0"],
                   k["vnorm"])
    return c


def report(tag, **errs):
    print(tag, " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)


def test_water8_database_pseudopotentials_match_oracle(gpu):
    s = water8_db()
    assert s["pseudo"]["vps"].shape[2] == 4000 and s["kb"]["order"] == ("O",)
    o = oracle.System(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                      s["kz"])
    lp, keep = oracle_pseudo(s)
    vl = o.vloc_g(lp)
    nl, keep2 = oracle_nonlocal(s)
    psi_o = s["psi"].copy()
    r1 = o.step(None, psi_o, dt=5.0, ns_tol=1e-15, ns_max_iter=30, vloc=vl, nl=nl,
                want=("rho", "S", "T", "v"))
    with make_ctx(s) as c:
        c.set_psi(s["psi"])
        c.fft_g2r()
        rho, v = c.get_density(), c.get_vks()
        c.vpsi_r2g()
        c.nonlocal_force()
        c.lambda_()
        c.sd_update()
        S = c.overlap()
        T, it, err, ok = c.ortho_ns(1e-15, 30)
        c.rotate()
        g1 = c.result()
        _, enl = c.nonlocal_get()
        errs = dict(rho=rel(rho, r1["rho"]), vks=rel(v, r1["v"]), S=rel(S, r1["S"]),
                    T=rel(T, r1["T"]), psi=rel(c.get_psi(), psi_o))
        for k in ("eke", "eloc", "ehart", "exc"):
            errs[k] = abs(g1[k] - r1[k]) / abs(r1[k])
        report("[water8 db step 1]", **errs)
        assert ok and all(e < TOL for e in errs.values()), errs
        for k in range(3):
            g = c.step()
            r = o.step(None, psi_o, dt=5.0, ns_tol=1e-15, ns_max_iter=30, vloc=vl, nl=nl)
            e = dict(eelec=abs(g["eelec"] - r["eelec"]) / abs(r["eelec"]),
                     enl=abs(g["enl"] - r["enl"]) / abs(r["enl"]), psi=rel(c.get_psi(), psi_o))
            report(f"[water8 db step {k + 2}] E={g['eelec']:.10f}", **e)
            assert all(x < TOL for x in e.values()), e
    assert abs(enl) > 0.0
