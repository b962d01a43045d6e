"""Non-local KB pseudopotential (l = 0, the reference's G-space method) on
the device against the oracle: projections Z, E_NL and ion forces
(oa_nonlocal_get), the coefficient force (oa_nonlocal_force after
oa_vpsi_r2g), and whole steps with E_NL in eelec, in fixed-V and V_KS
contexts, steepest descent and conjugate gradient.  FP64 bar: 1e-10
relative."""
import numpy as np
import pytest

import oracle
from openatom_amd import OaError, system
from openatom_amd.step import CPContext
from tests.helpers import oracle_nonlocal, oracle_pseudo, rel, small_system

pytestmark = pytest.mark.gpu

TOL = 1e-10
_cache = {}


def get_system(name):
    if name not in _cache:
        _cache[name] = {
            "small": lambda: small_system(nmol=3, nstates=6, seed=7),
            "square27": lambda: small_system(nmol=3, box=(5.0, 5.2, 5.8), grid=(27, 27, 30),
                                             nstates=19, seed=5),
            "water8": lambda: system.build("water8", seed=11, kvec_source=oracle.stored_kvecs),
        }[name]()
    return _cache[name]


def osys(s):
    return oracle.System(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                         s["kz"])


def make_ctx(s, vks=False, **kw):
    c = CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"],
                  vks_from_rho=int(vks), **kw)
    t = s["kb"]
    c.set_nonlocal(t["pos"], t["species"], t["gmin"], t["dg"], t["proj"], t["proj_g0"],
                   t["vnorm"])
    if vks:
        p = s["pseudo"]
        c.set_local_pseudo(s["pos"], s["species"], p["gmin"], p["dg"], p["vps"], p["vps_g0"])
    return c


def report(tag, **errs):
    print(tag, " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)


@pytest.mark.parametrize("name", ["small", "square27", "water8"])
def test_projections_energy_ion_forces(gpu, name):
    s = get_system(name)
    o = osys(s)
    nl, _ = oracle_nonlocal(s)
    z_ref = o.nl_zmat(nl, s["psi"])
    fa_ref = np.zeros((nl.natoms, 3))
    e_ref = o.nl_apply(nl, s["psi"], None, fa_ref)
    with make_ctx(s) as c:
        c.set_psi(s["psi"])
        z, e, fa = c.nonlocal_get(atom_forces=True)
        z2, e2 = c.nonlocal_get()
    errs = dict(z=rel(z, z_ref), enl=rel(e, e_ref), fatm=rel(fa, fa_ref))
    report(f"[{name}] nonlocal", **errs)
    assert all(v < TOL for v in errs.values())
    assert np.array_equal(z, z2) and e == e2


@pytest.mark.parametrize("name", ["small", "square27", "water8"])
def test_coefficient_force(gpu, name):
    """F after oa_vpsi_r2g plus oa_nonlocal_force against the oracle's
    V psi + kinetic + non-local forces."""
    s = get_system(name)
    o = osys(s)
    nl, _ = oracle_nonlocal(s)
    f_ref = o.vpsi_r2g_all(o.g2r_all(s["psi"]), s["vks"])
    o.eke_force(s["psi"], f_ref)
    f_loc = f_ref.copy()
    o.nl_apply(nl, s["psi"], f_ref)
    with make_ctx(s) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        c.fft_g2r()
        c.vpsi_r2g()
        c.nonlocal_force()
        f = c.get_force()
    e = rel(f - f_loc, f_ref - f_loc)
    report(f"[{name}] force", f=rel(f, f_ref), f_nl=e)
    assert rel(f, f_ref) < TOL and e < 1e-9


@pytest.mark.parametrize("vks", [False, True])
@pytest.mark.parametrize("name", ["small", "water8"])
def test_steps_match_oracle(gpu, name, vks):
    s = get_system(name)
    o = osys(s)
    nl, _ = oracle_nonlocal(s)
    vl = o.vloc_g(oracle_pseudo(s)[0]) if vks else None
    psi_o = s["psi"].copy()
    with make_ctx(s, vks=vks) as c:
        c.set_psi(s["psi"])
        if not vks:
            c.set_vks(s["vks"])
        for k in range(3):
            g = c.step()
            r = o.step(None if vks else s["vks"], psi_o, dt=5.0, ns_tol=1e-15, ns_max_iter=30,
                       vloc=vl, nl=nl)
            errs = {f: rel(g[f], r[f]) for f in ("eelec", "eke", "eloc", "enl")}
            report(f"[{name} vks={vks} step {k}]", **errs)
            assert all(v < TOL for v in errs.values())
        e_psi = rel(c.get_psi(), psi_o)
    report(f"[{name} vks={vks}] psi after 3 steps", psi=e_psi)
    assert e_psi < TOL


def test_cg_steps_match_oracle(gpu):
    s = get_system("water8")
    o = osys(s)
    nl, _ = oracle_nonlocal(s)
    cg = oracle.CgMinimizer(s["nstates"], len(s["kx"]))
    psi_o = s["psi"].copy()
    with make_ctx(s, use_cg=1) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        for k in range(4):
            g = c.step()
            r = o.step(s["vks"], psi_o, dt=5.0, ns_tol=1e-15, ns_max_iter=30, cg=cg, nl=nl)
            errs = {f: rel(g[f], r[f]) for f in ("eelec", "enl", "fnorm")}
            report(f"[water8 cg step {k}]", **errs)
            assert all(v < TOL for v in errs.values())
        assert rel(c.get_psi(), psi_o) < TOL


def test_mixed_precision_with_nonlocal(gpu):
    """The non-local part stays FP64 in mixed contexts: after three steps
    the energy is within 1e-6 Ha/atom of the oracle's FP64 steps."""
    s = get_system("water8")
    o = osys(s)
    nl, _ = oracle_nonlocal(s)
    psi_o = s["psi"].copy()
    with make_ctx(s, precision="mixed") as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        for k in range(3):
            g = c.step()
            r = o.step(s["vks"], psi_o, dt=5.0, ns_tol=1e-15, ns_max_iter=30, nl=nl)
    de = abs(g["eelec"] - r["eelec"]) / s["natoms"]
    report("[water8 mixed] nonlocal steps", dE_per_atom=de, enl=rel(g["enl"], r["enl"]))
    assert de < 1e-6


def test_graph_replay_and_reset(gpu):
    """Graph replay equals eager with the non-local stage; replacing the
    atoms after captured steps recaptures (the new table is used)."""
    s = get_system("water8")
    out = {}
    moved = s["kb"]["pos"] + np.array([0.3, -0.2, 0.1])
    t = s["kb"]
    for graph in (0, 1):
        with make_ctx(s, use_graph=graph) as c:
            c.set_psi(s["psi"])
            c.set_vks(s["vks"])
            e = [c.step()["eelec"] for _ in range(3)]
            c.set_nonlocal(moved, t["species"], t["gmin"], t["dg"], t["proj"], t["proj_g0"],
                           t["vnorm"])
            e += [c.step()["enl"] for _ in range(2)]
            out[graph] = e + [c.get_psi()]
    assert out[0][:5] == out[1][:5]
    assert np.array_equal(out[0][5], out[1][5])


def test_errors(gpu):
    s = get_system("small")
    t = s["kb"]
    with CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                   s["kz"]) as c:
        with pytest.raises(OaError):
            c._nl_natoms = 1
            c.nonlocal_get()  # nothing set
        with pytest.raises(OaError):  # species out of range
            c.set_nonlocal(t["pos"], t["species"] + 5, t["gmin"], t["dg"], t["proj"],
                           t["proj_g0"], t["vnorm"])
        with pytest.raises(OaError):  # table ends below the cutoff
            c.set_nonlocal(t["pos"], t["species"], t["gmin"], t["dg"], t["proj"][:, :, :100],
                           t["proj_g0"], t["vnorm"])
