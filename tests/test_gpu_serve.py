"""The request loop (include/oa_serve.h): independent minimisations of
different sizes admitted by HBM budget and run concurrently on one GPU.
Each request's result must be the same alone and in the mix, bit for bit,
and the same as driving its context step by step."""
import numpy as np
import pytest

import oracle
from openatom_amd import _lib, serve, system
from openatom_amd.step import CPContext
from tests.helpers import small_system

pytestmark = pytest.mark.gpu

# conjugate gradient; Newton-Schulz by launches (no spinning kernels while
# other requests share the device)
OPTS = dict(dt=20.0, use_cg=1, ns_mode=1)
F_TOL = 1e-4
MAX_STEPS = 400

_cache = {}


def systems():
    if not _cache:
        # each converges to |F| <= 1e-4 within 150..360 CG steps on the oracle
        _cache["list"] = [
            system.build(system.serving_workload(4, 50.0), seed=5,
                         kvec_source=oracle.stored_kvecs),                   # 45^3, 16 states
            small_system(box=(5.0, 5.2, 5.8), grid=(27, 27, 30), nstates=8, seed=5),
            system.build("water8", seed=11, kvec_source=oracle.stored_kvecs),  # 48^3, 32 states
            small_system(box=(6.0, 6.0, 6.0), ecut=20.0, grid=(36, 36, 36), nstates=12, seed=21),
        ]
    return _cache["list"]


def run(srv, reqs):
    for tag, s in reqs:
        srv.submit(s, tag, f_tol=F_TOL, max_steps=MAX_STEPS, want_psi=True, **OPTS)
    return {r["tag"]: r for r in srv.drain()}


def test_alone_equals_mix(gpu):
    ss = systems()
    alone = {}
    for t, s in enumerate(ss):
        with serve.Server(max_active=1) as srv:
            alone.update(run(srv, [(t, s)]))
    with serve.Server(max_active=4) as srv:
        mix = run(srv, list(enumerate(ss)))
        st = srv.stats()
    print("[serve mix]", st, flush=True)
    assert st["peak_active"] >= 2 and st["finished"] == len(ss)
    for t in alone:
        a, m = alone[t], mix[t]
        print(f"[serve] req {t}: steps {a['steps']} E {a['eelec']:.10f} |F| {a['fnorm']:.2e} "
              f"alone {a['run_ms']:.1f} ms, in mix {m['run_ms']:.1f} ms", flush=True)
        assert a["status"] == 0 and m["status"] == 0
        assert a["fnorm"] <= F_TOL
        assert a["steps"] == m["steps"] and a["eelec"] == m["eelec"]
        assert np.array_equal(a["psi"], m["psi"])


def test_admission_respects_budget(gpu):
    ss = systems()
    need = [serve.ctx_bytes(s, **OPTS) for s in ss]
    budget = sorted(need)[-1] + sorted(need)[-2]  # the two largest fit, not three
    with serve.Server(hbm_budget=budget, max_active=4) as srv:
        res = run(srv, list(enumerate(ss)))
        st = srv.stats()
    print("[serve budget]", budget, need, st, flush=True)
    assert st["peak_hbm_bytes"] <= budget
    assert st["peak_active"] <= 3
    assert all(r["status"] == 0 for r in res.values())
    assert all(r["hbm_bytes"] == need[t] for t, r in res.items())


def test_request_equals_direct_steps(gpu):
    """The server's result is the one of stepping the context directly."""
    s = systems()[3]
    with serve.Server(max_active=2) as srv:
        r = run(srv, [(7, s)])[7]
    with CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"],
                   **OPTS) as c:
        c.set_psi(s["psi"])
        c.set_vks(s["vks"])
        for k in range(MAX_STEPS):
            o = c.step()
            if o["fnorm"] <= F_TOL:
                break
        psi = c.get_psi()
    assert r["steps"] == k + 1 and r["eelec"] == o["eelec"]
    assert np.array_equal(r["psi"], psi)


def test_oversized_request_rejected(gpu):
    s = systems()[2]
    from openatom_amd import OaError
    with serve.Server(hbm_budget=1 << 20, max_active=1) as srv:
        with pytest.raises(OaError):
            srv.submit(s, 1, **OPTS)


def test_mixed_request_finishes_in_fp64(gpu):
    """A mixed-precision request whose tolerance lies below the FP32 FFT
    floor continues in FP64 (fp64_below) and reaches the FP64 solve's
    minimum; admission charged the larger of the two footprints."""
    s = systems()[2]  # water8
    f_tol = 1e-6
    with serve.Server(max_active=2) as srv:
        srv.submit(s, 1, f_tol=f_tol, max_steps=800, **OPTS)
        srv.submit(s, 2, f_tol=f_tol, max_steps=800, precision="mixed", **OPTS)
        res = {r["tag"]: r for r in srv.drain()}
    r64, rmx = res[1], res[2]
    print(f"[serve mixed->fp64] fp64: {r64['steps']} steps E {r64['eelec']:.10f}; mixed: "
          f"{rmx['steps']} steps ({rmx['fp64_steps']} in FP64) E {rmx['eelec']:.10f}", flush=True)
    assert r64["status"] == 0 and rmx["status"] == 0
    assert rmx["fnorm"] <= f_tol and rmx["fp64_steps"] > 0
    assert abs(rmx["eelec"] - r64["eelec"]) / s["natoms"] < 1e-8
    need = max(serve.ctx_bytes(s, **OPTS), serve.ctx_bytes(s, precision="mixed", **OPTS))
    assert rmx["hbm_bytes"] == need


def test_scf_request_equals_direct_steps(gpu):
    """A self-consistent request (ions instead of a fixed V: V_KS rebuilt
    from each step's rho) gives the result of stepping a vks_from_rho
    context directly, alone and in a mix with fixed-V requests."""
    ss = systems()
    s = ss[0]
    t = s["pseudo"]
    with serve.Server(max_active=1) as srv:
        srv.submit(s, "scf", f_tol=F_TOL, max_steps=MAX_STEPS, want_psi=True, scf=True, **OPTS)
        alone = srv.drain()[0]
    with serve.Server(max_active=4) as srv:
        srv.submit(s, "scf", f_tol=F_TOL, max_steps=MAX_STEPS, want_psi=True, scf=True, **OPTS)
        for k, o in enumerate(ss[1:]):
            srv.submit(o, k, f_tol=F_TOL, max_steps=MAX_STEPS, **OPTS)
        mix = {r["tag"]: r for r in srv.drain()}["scf"]
    with CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"],
                   vks_from_rho=1, **OPTS) as c:
        c.set_local_pseudo(s["pos"], s["species"], t["gmin"], t["dg"], t["vps"], t["vps_g0"])
        c.set_psi(s["psi"])
        for k in range(MAX_STEPS):
            o = c.step()
            if o["fnorm"] <= F_TOL:
                break
        psi = c.get_psi()
    print(f"[serve scf] steps {alone['steps']} E {alone['eelec']:.10f} (direct {k + 1}, "
          f"{o['eelec']:.10f})", flush=True)
    # converged, or the same step cap reached everywhere
    assert alone["status"] in (0, 4) and mix["status"] == alone["status"]
    assert alone["steps"] == k + 1 == mix["steps"]
    assert alone["eelec"] == o["eelec"] == mix["eelec"]
    assert np.array_equal(alone["psi"], psi) and np.array_equal(mix["psi"], psi)


def test_repeated_tag_results_stay_paired(gpu):
    """Two requests under one tag: each result carries its own coefficients."""
    ss = systems()
    with serve.Server(max_active=2) as srv:
        srv.submit(ss[0], "t", f_tol=F_TOL, max_steps=MAX_STEPS, want_psi=True, **OPTS)
        srv.submit(ss[3], "t", f_tol=F_TOL, max_steps=MAX_STEPS, want_psi=True, **OPTS)
        res = srv.drain()
    assert sorted(r["psi"].shape for r in res) == sorted([ss[0]["psi"].shape, ss[3]["psi"].shape])
    assert all(r["tag"] == "t" and r["status"] == 0 for r in res)


def test_concurrent_sizing_and_creation(gpu):
    """oa_ctx_bytes from several host threads while another thread creates
    and steps contexts (the submit path beside the scheduler): every sizing
    agrees and the contexts still step correctly."""
    import threading
    ss = systems()
    want = [serve.ctx_bytes(s, **OPTS) for s in ss]
    errors = []

    def sizer():
        for _ in range(20):
            for s, w in zip(ss, want):
                b = serve.ctx_bytes(s, **OPTS)
                if b != w:
                    errors.append((b, w))

    def creator():
        for s in ss:
            with CPContext(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                           s["kz"], **OPTS) as c:
                c.set_psi(s["psi"])
                c.set_vks(s["vks"])
                c.step()

    th = [threading.Thread(target=sizer) for _ in range(6)] + [threading.Thread(target=creator)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors


def test_pooled_context_equals_fresh(gpu):
    """A request that takes over the pooled context of an earlier request of
    the same shape (other coefficients, left in the pool when it finished)
    gives the result of a fresh context, bit for bit; a request of another
    shape does not take it."""
    ss = systems()
    s = ss[3]
    warm = dict(s, psi=np.ascontiguousarray(s["psi"][::-1]))  # same shape, states reversed
    with serve.Server(max_active=1) as srv:
        fresh = run(srv, [("x", s)])["x"]
    with serve.Server(max_active=1) as srv:
        res = run(srv, [("warm", warm), ("other", ss[1]), ("x", s)])
    print(f"[serve pool] fresh: {fresh['steps']} steps E {fresh['eelec']:.10f}; pooled: "
          f"{res['x']['steps']} steps E {res['x']['eelec']:.10f}, reused "
          f"{[res[k]['reused'] for k in ('warm', 'other', 'x')]}", flush=True)
    assert fresh["reused"] == 0 and res["warm"]["reused"] == 0 and res["other"]["reused"] == 0
    assert res["x"]["reused"] == 1
    assert res["x"]["status"] == 0 and res["x"]["steps"] == fresh["steps"]
    assert res["x"]["eelec"] == fresh["eelec"]
    assert np.array_equal(res["x"]["psi"], fresh["psi"])


def test_mixed_switch_pools_both_contexts(gpu):
    """At the FP64 switch the request's mixed context goes to the pool: a
    second mixed request of the same shape takes it over at admission and
    the FP64 context at its own switch, and gives the result of a fresh
    server bit for bit."""
    s = systems()[0]  # 45^3, 16 states
    warm = dict(s, psi=np.ascontiguousarray(s["psi"][::-1]))  # same shape, states reversed
    kw = dict(f_tol=1e-5, max_steps=800, precision="mixed", want_psi=True, **OPTS)
    with serve.Server(max_active=1) as srv:
        srv.submit(s, "x", **kw)
        fresh = {r["tag"]: r for r in srv.drain()}["x"]
    with serve.Server(max_active=1) as srv:
        srv.submit(warm, "warm", **kw)
        srv.submit(s, "x", **kw)
        res = {r["tag"]: r for r in srv.drain()}
    print(f"[serve mixed pool] fresh: {fresh['steps']} steps ({fresh['fp64_steps']} FP64) "
          f"E {fresh['eelec']:.10f}; pooled: {res['x']['steps']} steps, reused "
          f"{[res[k]['reused'] for k in ('warm', 'x')]}", flush=True)
    assert fresh["fp64_steps"] > 0 and res["warm"]["fp64_steps"] > 0
    assert res["warm"]["reused"] == 0 and res["x"]["reused"] == 1
    assert res["x"]["status"] == 0 and res["x"]["steps"] == fresh["steps"]
    assert res["x"]["fp64_steps"] == fresh["fp64_steps"]
    assert res["x"]["eelec"] == fresh["eelec"]
    assert np.array_equal(res["x"]["psi"], fresh["psi"])


def test_cancel_withdraws_unfinished_requests(gpu):
    """oa_server_cancel: queued requests finish at once and the admitted one
    after its step in flight, all with OA_ECANCELED; a request submitted
    afterwards runs normally and equals one run on a fresh server."""
    ss = systems()
    with serve.Server(max_active=1) as srv:
        for t in range(4):
            srv.submit(ss[t], t, f_tol=F_TOL, max_steps=MAX_STEPS, **OPTS)
        srv.cancel()
        srv.submit(ss[0], "after", f_tol=F_TOL, max_steps=MAX_STEPS, want_psi=True, **OPTS)
        res = {r["tag"]: r for r in srv.drain()}
    with serve.Server(max_active=1) as srv:
        fresh = run(srv, [("after", ss[0])])["after"]
    print(f"[serve cancel] {[(t, res[t]['status'], res[t]['steps']) for t in range(4)]}; after: "
          f"{res['after']['status']} {res['after']['steps']} steps", flush=True)
    assert all(res[t]["status"] == _lib.OA_ECANCELED for t in range(4))
    assert all(res[t]["steps"] < MAX_STEPS for t in range(4))
    assert res["after"]["status"] == 0 and res["after"]["steps"] == fresh["steps"]
    assert np.array_equal(res["after"]["psi"], fresh["psi"])
