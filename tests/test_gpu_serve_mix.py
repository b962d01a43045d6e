"""Two dozen requests drawn from the serving workload's distribution: water
clusters of 4-64 molecules (log-uniform), four states per molecule, Ecut in
{50, 60, 70} Ry, ions carried by the request (V_KS rebuilt from each step's
density), conjugate gradient at the serving time step, half of them in
mixed precision.  Each runs a fixed number of steps (the step cap, so the
test time stays bounded) alone on its own server and in one mix of all 24
with up to 8 on the GPU at once: step counts, energies, |F| and the final
coefficients must agree bit for bit."""
import numpy as np
import pytest

from openatom_amd import serve, system

pytestmark = pytest.mark.gpu

NREQ = 24
STEPS = 25
OPTS = dict(dt=160.0, use_cg=1, ns_mode=6)


def draw():
    rng = np.random.default_rng(2024)
    reqs = []
    for i in range(NREQ):
        nmol = int(round(np.exp(rng.uniform(np.log(4), np.log(64)))))
        ecut = float(rng.choice([50.0, 60.0, 70.0]))
        s = system.build(system.serving_workload(nmol, ecut), seed=100 + i)
        reqs.append((i, s, "mixed" if i % 2 else "fp64"))
    return reqs


def submit(srv, tag, s, precision):
    srv.submit(s, tag, f_tol=1e-5, max_steps=STEPS, precision=precision, want_psi=True,
               scf=True, **OPTS)


def test_drawn_requests_alone_equal_mix(gpu):
    reqs = draw()
    print("[serve mix24] sizes", [(s["nstates"] // 4, s["ecut_ry"], p) for _, s, p in reqs],
          flush=True)
    alone = {}
    for tag, s, p in reqs:
        with serve.Server(max_active=1) as srv:
            submit(srv, tag, s, p)
            alone[tag] = srv.drain()[0]
    with serve.Server(max_active=8) as srv:
        for tag, s, p in reqs:
            submit(srv, tag, s, p)
        mix = {r["tag"]: r for r in srv.drain()}
        st = srv.stats()
    print("[serve mix24]", st, flush=True)
    assert st["peak_active"] >= 2 and st["finished"] == NREQ
    for tag, s, p in reqs:
        a, m = alone[tag], mix[tag]
        assert a["status"] == m["status"] and a["steps"] == m["steps"] == STEPS
        assert a["eelec"] == m["eelec"] and a["fnorm"] == m["fnorm"]
        assert np.isfinite(a["eelec"]) and np.all(np.isfinite(a["psi"]))
        assert np.array_equal(a["psi"], m["psi"])
