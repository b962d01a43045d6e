"""Parity of the HIP path (through the C ABI) with the CPU oracle, stage by
stage and for whole steps.  FP64 bar: 1e-10 relative (max-abs error over
max-abs value) on psi(r), rho, F, EKE, Lambda, S, T, rotated psi and the
electronic energy.  Mixed precision: FP32 FFTs, errors stated below."""
import numpy as np
import pytest

import oracle
from openatom_amd import system
from openatom_amd.step import CPContext
from tests.helpers import rel, small_system

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _systems():
    return {
        "small": lambda: small_system(),
        "ragged37": lambda: small_system(nstates=37, ecut=18.0, seed=8),
        "water8": lambda: system.build("water8", seed=11, kvec_source=oracle.stored_kvecs),
        # square x-y planes for the fused plane engine: odd Ny (27 = 3 x 9),
        # and 128 = 8 x 16 with 17 kx columns (two y batches of 16) and four
        # x rounds of pair-rows
        "square27": lambda: small_system(box=(5.0, 5.2, 5.8), grid=(27, 27, 30), nstates=19,
                                         seed=5),
        "dense128": lambda: small_system(box=(6.0, 6.0, 5.0), ecut=80.0, grid=(128, 128, 30),
                                         nstates=5, seed=9),
        # more than 128 states: two 128-row gram blocks, the non-persistent
        # rotate and the separate |F|^2 reduction
        "states150": lambda: small_system(box=(5.0, 5.2, 5.6), ecut=36.0, grid=(40, 40, 45),
                                          nstates=150, seed=13),
    }


_cache = {}


def get_system(name):
    if name not in _cache:
        _cache[name] = _systems()[name]()
    return _cache[name]


def make_ctx(s, precision="fp64", **kw):
    if kw.get("fft_mode") == 4 and s["grid"][0] != s["grid"][1]:
        pytest.skip("fused two-factor planes need nx == ny")
    return CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                     s["kz"], precision=precision, **kw)


def make_oracle(s):
    return oracle.System(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                         s["kz"])


def report(tag, **errs):
    print(tag, " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)


# fft_mode: 0 auto, 1 Stockham per pass, 2 Stockham fused planes,
# 3 two-factor per pass, 4 two-factor fused planes (square planes only)
MODES = [0, 1, 2, 3, 4]
FFT_SYSTEMS = ["small", "ragged37", "water8", "square27", "dense128"]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", FFT_SYSTEMS)
def test_g2r_density(gpu, name, mode):
    s = get_system(name)
    o = make_oracle(s)
    with make_ctx(s, fft_mode=mode) as c:
        c.set_psi(s["psi"])
        c.fft_g2r()
        pr = o.g2r_all(s["psi"])
        e_psir = max(rel(c.get_psir(st), pr[st]) for st in range(s["nstates"]))
        e_rho = rel(c.get_density(), o.density(pr))
    report(f"[{name} mode{mode}] g2r", psir=e_psir, rho=e_rho)
    assert e_psir < TOL and e_rho < TOL


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", FFT_SYSTEMS)
def test_r2g_force_eke(gpu, name, mode):
    s = get_system(name)
    o = make_oracle(s)
    with make_ctx(s, fft_mode=mode) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        c.fft_g2r()
        eke = c.vpsi_r2g()
        F = c.get_force()
    pr = o.g2r_all(s["psi"])
    Fo = o.vpsi_r2g_all(pr, s["vks"])
    eko = o.eke_force(s["psi"], Fo)
    report(f"[{name} mode{mode}] r2g", force=rel(F, Fo), eke=abs(eke - eko) / abs(eko))
    assert rel(F, Fo) < TOL
    assert abs(eke - eko) < TOL * abs(eko)


# Newton-Schulz forms: 0 autotuned, 1 launches per iteration, 2 cooperative
# 32 x 32 tiles, 3/4/5 cooperative 16-row blocks split over 1/2/4 workgroups,
# 6 one workgroup in a plain launch (n <= 64; launches per iteration above)
@pytest.mark.parametrize("ns_mode", [0, 1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("name", ["small", "ragged37", "water8", "states150"])
def test_lambda_sd_overlap_ns_rotate(gpu, name, ns_mode):
    s = get_system(name)
    split = {4: 2, 5: 4}.get(ns_mode, 1)
    if split > (s["nstates"] + 15) // 16:
        pytest.skip(f"{split} column groups need at least {split} column tiles")
    o = make_oracle(s)
    dt = 5.0
    with make_ctx(s, dt=dt, ns_mode=ns_mode) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        c.fft_g2r()
        c.vpsi_r2g()
        lam = c.lambda_()
        F = c.get_force()
        fn = c.sd_update()
        psi_sd = c.get_psi()
        S = c.overlap()
        T, it, err, ok = c.ortho_ns(1e-15, 30)
        c.rotate()
        psi_out = c.get_psi()
    # oracle chain on the same inputs
    pr = o.g2r_all(s["psi"])
    Fo = o.vpsi_r2g_all(pr, s["vks"])
    o.eke_force(s["psi"], Fo)
    lamo = o.lambda_(s["psi"], Fo)
    o.lambda_apply(s["psi"], lamo, Fo)
    psio = s["psi"].copy()
    o.sd_update(o.cmass(25.0, 2.0), dt, Fo, psio)
    o.red_sync(psio)
    So = o.overlap(psio)
    To, ito, erro = oracle.ns(So, tol=1e-15, max_iter=30)
    outo = o.rotate(psio, To)
    errs = dict(lam=rel(lam, lamo), force=rel(F, Fo), fnorm=abs(fn - np.linalg.norm(Fo)) / np.linalg.norm(Fo),
                psi_sd=rel(psi_sd, psio), S=rel(S, So), T=rel(T, To), psi_out=rel(psi_out, outo))
    report(f"[{name} ns{ns_mode}] chain it={it}/{abs(ito)} err={err:.2e}", **errs)
    assert ok and it >= 1
    for k, v in errs.items():
        assert v < TOL, (k, v)
    # orthonormal output (norm-2 convention)
    Sout = system.gamma_overlap(psi_out, s["kx"])
    assert np.abs(Sout / 2 - np.eye(s["nstates"])).max() < 1e-12


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", ["small", "water8", "square27"])
def test_full_steps_match_oracle(gpu, name, mode):
    s = get_system(name)
    o = make_oracle(s)
    psio = s["psi"].copy()
    with make_ctx(s, dt=5.0, ns_tol=1e-15, ns_max_iter=30, fft_mode=mode) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        for k in range(3):
            r = c.step()
            ro = o.step(s["vks"], psio, dt=5.0, ns_tol=1e-15, ns_max_iter=30, use_lambda=True)
            psi = c.get_psi()
            de = abs(r["eelec"] - ro["eelec"]) / abs(ro["eelec"])
            report(f"[{name} mode{mode}] step{k}", dE=de, eke=abs(r["eke"] - ro["eke"]) / abs(ro["eke"]),
                   psi=rel(psi, psio), resid=r["resid_max"])
            assert de < TOL
            assert rel(psi, psio) < TOL
            assert r["resid_max"] < 1e-12
            assert r["ns_iters"] >= 1


@pytest.mark.parametrize("mode", [0, 3])
@pytest.mark.parametrize("name", ["water8", "dense128"])
def test_mixed_precision_errors_stated(gpu, name, mode):
    """Mixed mode (FP32 FFT chain, FP64 master coefficients and FP64 GEMMs):
    errors against the FP64 oracle for one step.  Bars: |dE| <= 1e-6 Ha per
    atom, rho within 1e-5 relative, ||Psi^H Psi/2 - I|| <= 1e-10."""
    s = get_system(name)
    o = make_oracle(s)
    psio = s["psi"].copy()
    ro = o.step(s["vks"], psio, dt=5.0, ns_tol=1e-15, ns_max_iter=30, use_lambda=True, want=("rho",))
    with make_ctx(s, precision="mixed", dt=5.0, fft_mode=mode) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        r = c.step()
        rho = c.get_density()
        psi = c.get_psi()
    natoms = s["natoms"]
    dE_atom = abs(r["eelec"] - ro["eelec"]) / natoms
    report(f"[{name} mixed mode{mode}]", dE_per_atom=dE_atom, rho=rel(rho, ro["rho"]),
           psi=rel(psi, psio), resid=r["resid_max"])
    assert dE_atom < 1e-6
    assert rel(rho, ro["rho"]) < 1e-5
    assert r["resid_max"] < 1e-10


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("name", ["small", "water8", "states150"])
def test_cg_steps_match_oracle(gpu, name, graph):
    """Conjugate-gradient minimisation (use_cg) against the oracle's CG over
    six steps: coefficients and energies at 1e-10, the same |F| sequence."""
    s = get_system(name)
    o = make_oracle(s)
    psio = s["psi"].copy()
    cgo = oracle.CgMinimizer(s["nstates"], len(s["kx"]))
    with make_ctx(s, dt=20.0, use_cg=1, use_graph=graph) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        for k in range(6):
            r = c.step()
            ro = o.step(s["vks"], psio, dt=20.0, ns_tol=1e-15, ns_max_iter=30, use_lambda=True,
                        cg=cgo)
            psi = c.get_psi()
            report(f"[{name} cg graph{graph}] step{k}", dE=abs(r["eelec"] - ro["eelec"]),
                   psi=rel(psi, psio), f=abs(r["fnorm"] - ro["fnorm"]) / ro["fnorm"])
            assert abs(r["eelec"] - ro["eelec"]) < TOL * abs(ro["eelec"])
            assert abs(r["fnorm"] - ro["fnorm"]) < TOL * ro["fnorm"]
            assert rel(psi, psio) < TOL
        # oa_psi_set restarts the minimisation: the first step repeats
        c.set_psi(s["psi"])
        r0 = c.step()
    o2 = make_oracle(s)
    p2 = s["psi"].copy()
    ro0 = o2.step(s["vks"], p2, dt=20.0, ns_tol=1e-15, ns_max_iter=30, use_lambda=True,
                  cg=oracle.CgMinimizer(s["nstates"], len(s["kx"])))
    assert abs(r0["eelec"] - ro0["eelec"]) < TOL * abs(ro0["eelec"])


def test_cg_needs_lambda(gpu):
    s = get_system("small")
    from openatom_amd import OaError
    with pytest.raises(OaError):
        make_ctx(s, use_cg=1, use_lambda=0)


@pytest.mark.parametrize("name", ["states150"])
def test_many_states_steps_match_oracle(gpu, name):
    s = get_system(name)
    o = make_oracle(s)
    psio = s["psi"].copy()
    with make_ctx(s, dt=5.0) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        for k in range(2):
            r = c.step()
            ro = o.step(s["vks"], psio, dt=5.0, ns_tol=1e-15, ns_max_iter=30, use_lambda=True)
            report(f"[{name}] step{k}", dE=abs(r["eelec"] - ro["eelec"]),
                   psi=rel(c.get_psi(), psio), resid=r["resid_max"])
            assert abs(r["eelec"] - ro["eelec"]) < TOL * abs(ro["eelec"])
            assert rel(c.get_psi(), psio) < TOL
            assert r["resid_max"] < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_equals_eager(gpu, precision):
    """use_graph replays the captured kernel chain (both psi-buffer parities,
    several replays with host synchronisation between them): every step must
    equal the eager step bit for bit, and the FP64 chain the oracle."""
    s = get_system("water8")
    o = make_oracle(s)
    psio = s["psi"].copy()
    with make_ctx(s, precision=precision, dt=5.0) as ce, \
            make_ctx(s, precision=precision, dt=5.0, use_graph=1) as cg:
        for c in (ce, cg):
            c.set_psi(s["psi"])
            c.set_vks(s["vks"])
        for k in range(5):
            re_ = ce.step()
            rg = cg.step()
            pe, pg = ce.get_psi(), cg.get_psi()
            report(f"[water8 {precision} graph] step{k}", dpsi=float(np.abs(pe - pg).max()),
                   dE=abs(re_["eelec"] - rg["eelec"]))
            assert np.array_equal(pe, pg)
            assert re_["eelec"] == rg["eelec"] and re_["fnorm"] == rg["fnorm"]
            if precision == "fp64":
                ro = o.step(s["vks"], psio, dt=5.0, ns_tol=1e-15, ns_max_iter=30,
                            use_lambda=True)
                assert abs(rg["eelec"] - ro["eelec"]) < TOL * abs(ro["eelec"])
                assert rel(pg, psio) < TOL


def test_enumerated_layout_equals_explicit(gpu):
    """npts = 0 (library enumerates) gives the same k-vectors as passing them."""
    s = get_system("water8")
    c = CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"])
    kx, ky, kz = c.kvecs()
    c.close()
    assert np.array_equal(kx, s["kx"]) and np.array_equal(ky, s["ky"]) and np.array_equal(kz, s["kz"])


def test_bad_layout_rejected(gpu):
    s = get_system("small")
    kx = s["kx"].copy()
    ky = s["ky"].copy()
    kz = s["kz"].copy()
    # break kz contiguity inside a line
    i = int(np.where((kx > 0))[0][3])
    kz[i] += 5
    from openatom_amd import OaError
    with pytest.raises(OaError):
        CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], kx, ky, kz)


@pytest.mark.parametrize("name,precision", [("water8", "fp64"), ("square27", "fp64"),
                                            ("water8", "mixed")])
def test_autotuned_context_matches_oracle(gpu, name, precision):
    """opt.autotune times candidate chunkings, K slices and Newton-Schulz
    modes at creation; whatever it picks, three steps still match the
    oracle (FP64 at 1e-10; mixed at the FP32 FFT bar) and the log names
    every knob it measured."""
    s = get_system(name)
    o = make_oracle(s)
    psio = s["psi"].copy()
    with make_ctx(s, precision=precision, autotune=1) as c:
        log = c.tuning_log()
        print(f"[{name} {precision} autotune]\n{log}", flush=True)
        assert "K slices" in log
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        for k in range(3):
            r = c.step()
            ro = o.step(s["vks"], psio, dt=5.0, ns_tol=1e-15, ns_max_iter=30)
        e_psi = rel(c.get_psi(), psio)
    report(f"[{name} {precision} autotune]", dE=abs(r["eelec"] - ro["eelec"]), psi=e_psi)
    if precision == "fp64":
        assert abs(r["eelec"] - ro["eelec"]) < TOL * abs(ro["eelec"]) and e_psi < TOL
    else:
        assert abs(r["eelec"] - ro["eelec"]) / s["natoms"] < 1e-6 and e_psi < 1e-5
