#!/usr/bin/env python3
"""Throughput of the CP electronic step on one MI355X.

One step = one minimisation iteration of the reference's GSpaceDriver order:
G->R state FFTs + rho, V psi + R->G FFTs + kinetic force/EKE, Lambda
projection, steepest-descent update + kx=0 sync, S = Psi^H Psi,
Newton-Schulz (S/2)^{-1/2}, Psi T, and the ||Psi^H Psi/2 - I|| check.

Workload (default water32): 32 H2O in a 9.86 A box, 128 doubly occupied
states at Gamma, 70 Ry (32,961 stored plane waves per state), 100^3 grid,
seeded coefficients; V(r) = V_xc + V_H + V_loc (PZ-LDA, Hartree, seeded
Gaussian local pseudopotential) rebuilt from each step's density on the
device (synthetic data).
`value` is the mixed-precision (FP32 FFT chain, FP64 coefficients,
GEMMs and Newton-Schulz) steps/s with inputs resident in HBM; the plain
FP64 path is timed in the same run and reported beside it.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("CP steps/s (128 states, 70 Ry) mixed vs FP64 at ΔE≤1e-6 Ha/atom; "
          "p99 solve ms")
HBM_PEAK_GBS = 8000.0
FP64_MFMA_TFLOPS = 78.6

PROF_STEPS = 20  # steps with a per-stage event breakdown after the timed steps

STAGES = ("k_zfwd", "k_y_fwd", "k_xc2r", "k_xr2c", "k_y_bwd", "k_zbwd",
          "k_plane_fwd", "k_plane_bwd", "k_plane_g2r", "k_plane_r2g", "k_plane_r2g_rho")
PLANE3 = ("k_plane_g2r", "k_plane_r2g", "k_plane_r2g_rho")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(stage, geo, csz, rsz):
    """Algorithmic bytes per state of one FFT pass in the reference's 3-pass
    pruned model: B_z = 2 n_lines Nz c, B_y = 2 nplane_x Nz Ny c,
    B_x = Ny Nz (Nx/2+1) c + Nx Ny Nz r (c, r: complex and real sizes).
    A fused plane kernel does the y and the x pass: B_y + B_x."""
    nx, ny, nz, nl, npx = geo["nx"], geo["ny"], geo["nz"], geo["nlines"], geo["nplane_x"]
    bz = 2 * nl * nz * csz
    by = 2 * npx * nz * ny * csz
    bx = ny * nz * (nx // 2 + 1) * csz + nx * ny * nz * rsz
    if stage in ("k_zfwd", "k_zbwd"):
        return bz
    if stage in ("k_y_fwd", "k_y_bwd"):
        return by
    if stage in ("k_plane_fwd", "k_plane_bwd") + PLANE3:
        return by + bx
    return bx


def plane_bytes(stage, geo, csz, nst, launch):
    """HBM bytes a two-factor fused plane pass must move (psi(r) stays on
    chip): G->R reads the populated z-lines of every state (n_lines Nz c)
    and writes one rho partial (N doubles) per state chunk; R->G reads and
    writes the z-lines and reads V (N doubles) once per chunk, and with rho
    fused into it also writes one rho partial per chunk."""
    n = geo["nx"] * geo["ny"] * geo["nz"]
    lines = geo["nlines"] * geo["nz"] * csz
    if stage == "k_plane_g2r":
        return nst * lines + launch["nchunk"] * n * 8
    if stage == "k_plane_r2g_rho":
        return 2 * nst * lines + 2 * launch["pcp"] * n * 8
    return 2 * nst * lines + launch["pcp"] * n * 8


# options every benchmark context takes (set from the command line)
RUN_OPTS = {}

# CG time step of the serving requests (a.u.).  The fixed-step CG of the
# V_KS minimisations converges faster with the step up to at least this
# value across the mix's sizes (4-64 H2O): measured steps to |F| <= 1e-5
# per size are recorded where the serving runs are documented.
DEFAULT_SERVE_DT = 160.0


def run_precision(sysd, precision, steps, warmup, device=0, use_graph=0, vks=True, autotune=1,
                  nonlocal_kb=False, prof_steps=None, gemm_split=None):
    """vks: V(r) = V_xc + V_H + V_loc rebuilt from each step's rho on the
    device (the density chares' work inside the step); else the seeded
    local potential as a fixed input.  nonlocal_kb: add the l = 0 KB
    non-local energy and force of the oxygen atoms to every step."""
    from openatom_amd.step import CPContext
    # The reference step has no ||Psi^H Psi - I|| evaluation (Ortho checks the
    # T iterates); it is measured once after the timed steps instead.
    extra = dict(RUN_OPTS)
    if gemm_split is not None:
        extra["gemm_split"] = int(gemm_split)
    c = CPContext(sysd["grid"], sysd["hmat"], sysd["ecut_ry"], sysd["nstates"], sysd["kx"],
                  sysd["ky"], sysd["kz"], precision=precision, device=device, dt=5.0,
                  ns_tol=1e-15 if precision == "fp64" else 1e-14, ns_max_iter=30,
                  want_resid=0, use_graph=use_graph, vks_from_rho=int(vks),
                  autotune=int(autotune), **extra)
    c.set_psi(sysd["psi"])
    if vks:
        t = sysd["pseudo"]
        c.set_local_pseudo(sysd["pos"], sysd["species"], t["gmin"], t["dg"], t["vps"],
                           t["vps_g0"])
    else:
        c.set_vks(sysd["vks"])
    if nonlocal_kb:
        t = sysd["kb"]
        c.set_nonlocal(t["pos"], t["species"], t["gmin"], t["dg"], t["proj"], t["proj_g0"],
                       t["vnorm"])
    first = c.step()
    for _ in range(max(0, warmup - 1)):
        c.step(fetch=False)
    c.sync()
    # the timed steps record events only around the step and the R->G plane
    # pass (the roofline kernel); a per-stage breakdown follows in steps of
    # its own, since an event at every stage boundary slows the step
    c.set_timing(True, level=2)
    t0 = time.perf_counter()
    for _ in range(steps):
        c.step(fetch=False)
    c.sync()
    t1 = time.perf_counter()
    tim_timed = {k: v for k, v in c.timings().items() if v > 0.0}
    c.set_timing(True, level=1)
    for _ in range(PROF_STEPS if prof_steps is None else prof_steps):
        c.step(fetch=False)
    c.sync()
    last = c.result()
    tim = c.timings()
    last["resid_max"], last["resid_rms"] = c.residual()
    geo = dict(nx=c.nx, ny=c.ny, nz=c.nz, nlines=c.nlines, nplane_x=c.nplane_x, npts=c.npts)
    psi = c.get_psi()
    cfg = dict(c.config)
    tuning = c.tuning_log()
    c.close()
    return dict(precision=precision, ms=(t1 - t0) * 1e3 / steps, steps_per_s=steps / (t1 - t0),
                first=first, last=last, timings=tim, timings_timed=tim_timed, geo=geo, psi=psi,
                launch=cfg,
                tuning=[t for t in tuning.split("\n") if t])


def quality_trace(sysd, nsteps, device=0):
    """Mixed precision against the GPU FP64 path from the same start, in
    lockstep, every step: dE per atom, max|d rho| (absolute and relative to
    max rho) and ||Psi^H Psi/2 - I|| of the mixed coefficients (FP64)."""
    import numpy as np
    from openatom_amd.step import CPContext
    ctx = {}
    for prec in ("mixed", "fp64"):
        c = CPContext(sysd["grid"], sysd["hmat"], sysd["ecut_ry"], sysd["nstates"], sysd["kx"],
                      sysd["ky"], sysd["kz"], precision=prec, device=device, dt=5.0,
                      ns_tol=1e-15 if prec == "fp64" else 1e-14, ns_max_iter=30, want_resid=0,
                      vks_from_rho=1, autotune=1)
        t = sysd["pseudo"]
        c.set_local_pseudo(sysd["pos"], sysd["species"], t["gmin"], t["dg"], t["vps"],
                           t["vps_g0"])
        c.set_psi(sysd["psi"])
        ctx[prec] = c
    natoms = sysd["natoms"]
    tr = dict(dE_per_atom=[], max_drho=[], drho_rel=[], resid_max=[])
    first_mixed = None
    for k in range(nsteps):
        gm = ctx["mixed"].step()
        gd = ctx["fp64"].step()
        if first_mixed is None:
            first_mixed = gm
        rm, rd = ctx["mixed"].get_density(), ctx["fp64"].get_density()
        d = float(np.abs(rm - rd).max())
        tr["dE_per_atom"].append(abs(gm["eelec"] - gd["eelec"]) / natoms)
        tr["max_drho"].append(d)
        tr["drho_rel"].append(d / float(np.abs(rd).max()))
        tr["resid_max"].append(ctx["mixed"].residual()[0])
    for c in ctx.values():
        c.close()
    worst = {k: max(v) for k, v in tr.items()}
    return dict(steps=nsteps, worst=worst, last={k: v[-1] for k, v in tr.items()},
                meets_bars=bool(worst["dE_per_atom"] <= 1e-6 and worst["resid_max"] <= 1e-10),
                first_step_mixed=first_mixed)


def roofline(res, nst, traffic=None):
    """Dominant kernel (from the per-stage breakdown) and its roofline, with
    its duration measured over the timed steps when they recorded it."""
    tim = res["timings"]
    fp32 = res["precision"] == "mixed"
    csz, rsz = (8, 4) if fp32 else (16, 8)
    geo = res["geo"]
    K = 2 * geo["npts"]
    cands = {k: tim[k] for k in STAGES + ("overlap", "rotate") if k in tim}
    dom = max(cands, key=cands.get)
    ms = res.get("timings_timed", {}).get(dom) or cands[dom]
    if dom in PLANE3:
        # fused pass: its own compulsory bytes; the 3-pass model's bytes for the
        # same work are reported beside it
        b = plane_bytes(dom, geo, csz, nst, res["launch"])
        ach = b / (ms * 1e-3) / 1e9
        bm = stage_bytes(dom, geo, csz, rsz) * nst
        r = dict(kernel=dom, bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                 frac=ach / HBM_PEAK_GBS, traffic=traffic, algorithmic_bytes=b, avg_ms=ms,
                 three_pass_model_bytes=bm, three_pass_model_GBs=bm / (ms * 1e-3) / 1e9)
    elif dom in STAGES:
        b = stage_bytes(dom, geo, csz, rsz) * nst
        ach = b / (ms * 1e-3) / 1e9
        r = dict(kernel=dom, bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                 frac=ach / HBM_PEAK_GBS, traffic=traffic, algorithmic_bytes=b, avg_ms=ms)
    else:
        fl = 2.0 * nst * nst * K
        ach = fl / (ms * 1e-3) / 1e12
        r = dict(kernel=dom, bound="mfma", achieved=ach, peak=FP64_MFMA_TFLOPS, unit="TFLOP/s",
                 frac=ach / FP64_MFMA_TFLOPS, traffic=traffic, algorithmic_flops=fl, avg_ms=ms)
    return r


def _serving_system(args):
    from openatom_amd import system
    nmol, ecut, seed = args
    return system.build(system.serving_workload(nmol, ecut), seed=seed)


def _noop(i):
    return i


def _one_blas_thread():
    """Worker initializer: one BLAS thread per builder process (the workers
    run side by side on the host cores)."""
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(1)
    except ImportError:
        pass


class ServingMix:
    """The request mix of the serving workload: water clusters of 4..max_mol
    molecules (log-uniform), 4 states per molecule, Ecut uniform in
    {50, 60, 70} Ry, Poisson arrivals at `rate` per second.  Each request's
    system is built when it is needed, in worker processes forked before any
    GPU work, and dropped once submitted (the server copies its inputs): a
    2,000-request mix would hold tens of GB of coefficients at once."""

    def __init__(self, nreq, rate, seed, max_mol, workers=12):
        import numpy as np
        from concurrent.futures import ProcessPoolExecutor
        import multiprocessing as mp
        rng = np.random.default_rng(seed)
        nmol = np.floor(np.exp(rng.uniform(np.log(4), np.log(max_mol + 1), nreq))).astype(int)
        self.nmol = [int(n) for n in np.clip(nmol, 4, max_mol)]
        ecut = rng.choice([50.0, 60.0, 70.0], nreq)
        self.arrive = np.cumsum(rng.exponential(1.0 / rate, nreq))
        self.jobs = [(n, float(e), int(seed * 1000 + i))
                     for i, (n, e) in enumerate(zip(self.nmol, ecut))]
        self.ex = ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("fork"),
                                      initializer=_one_blas_thread)
        list(self.ex.map(_noop, range(4 * workers)))  # every worker forked now, before the GPU

    def __len__(self):
        return len(self.jobs)

    def build(self, i):
        return self.ex.submit(_serving_system, self.jobs[i])

    def close(self):
        self.ex.shutdown()


def serving_run(mix, arrive, precision, f_tol, max_steps, budget, max_active, opts, ahead=64,
                deadline=0.0):
    """Submit each request at its arrival time (or as soon as its system is
    built, when that is later), collect results as they finish.  Latency =
    finish - arrival (queueing and any build delay included).  deadline > 0:
    at that many seconds every request not finished yet is withdrawn (the
    server's cancel; requests not yet submitted are not submitted), and the
    summary covers the requests that finished within it."""
    import numpy as np
    from openatom_amd import serve
    n = len(mix)
    res = {}
    lag = np.zeros(n)
    futs = {}
    with serve.Server(hbm_budget=budget, max_active=max_active) as srv:
        t0 = time.perf_counter()
        i = 0
        t_log = 0.0
        cancelled_at = None
        while len(res) < i or (cancelled_at is None and i < n):
            now = time.perf_counter() - t0
            if deadline > 0 and cancelled_at is None and now >= deadline:
                srv.cancel()
                cancelled_at = now
                for f in futs.values():
                    f.cancel()
                log(f"[serve] {precision} deadline {deadline:g} s: {len(res)} finished, "
                    f"{i} submitted, the rest withdrawn")
            if cancelled_at is None:
                for k in range(i, min(n, i + ahead)):
                    if k not in futs:
                        futs[k] = mix.build(k)
            while cancelled_at is None and i < n and arrive[i] <= now and futs[i].done():
                srv.submit(futs.pop(i).result(), i, f_tol=f_tol, max_steps=max_steps,
                           precision=precision, **opts)
                lag[i] = time.perf_counter() - t0 - arrive[i]
                i += 1
            for r in srv.poll():
                res[r["tag"]] = r
            if now - t_log >= 30.0:  # progress of a long run
                t_log = now
                st = srv.stats()
                log(f"[serve] {precision} t={now:.0f}s submitted {i}/{n} done {len(res)} "
                    f"active {int(st['active'])} queued {int(st['queued'])} "
                    f"device {st['hbm_bytes'] / 2**30:.1f} GiB")
            time.sleep(5e-4)
        wall = time.perf_counter() - t0 if cancelled_at is None else cancelled_at
        st = srv.stats()
    ECANCELED = 6
    done = [k for k in res if res[k]["status"] != ECANCELED]  # finished, not withdrawn
    lat = np.array([res[k]["total_ms"] + lag[k] * 1e3 for k in done])
    steps = np.array([res[k]["steps"] for k in done])
    ok = [k for k in done if res[k]["status"] == 0]
    for r in res.values():
        r.pop("psi", None)
    summary = {
        "reused_contexts": int(sum(res[k]["reused"] for k in done)),
        "solves_per_s": round(len(ok) / wall, 3),
        "p50_ms": round(float(np.percentile(lat, 50)), 2),
        "p99_ms": round(float(np.percentile(lat, 99)), 2),
        "max_ms": round(float(lat.max()), 2),
        "mean_queue_ms": round(float(np.mean([res[k]["queue_ms"] for k in done])), 2),
        "converged": len(ok), "requests": n,
        "steps_mean": round(float(steps.mean()), 1), "steps_max": int(steps.max()),
        "wall_s": round(wall, 3), "peak_active": int(st["peak_active"]),
        "peak_hbm_gib": round(st["peak_hbm_bytes"] / 2**30, 3)}
    if cancelled_at is not None:
        summary.update(deadline_s=deadline, finished_in_deadline=len(done),
                       submitted_by_deadline=i, withdrawn=len(res) - len(done),
                       not_submitted=n - i)
    return dict(res=res, summary=summary)


def serving_leg(args, mix, precisions):
    """The serving workload in each precision, with the largest energy
    difference per atom between precisions over requests both converged."""
    import numpy as np
    arrive = mix.arrive
    # fixed ions: V_KS = V_xc + V_H + V_loc rebuilt from every step's rho
    sopts = dict(dt=args.serve_dt, use_cg=1, ns_mode=6, scf=True)
    serving = {"workload": f"{len(mix)} independent fixed-ion minimisations of water clusters "
                           f"(4-{args.serve_max_mol} H2O log-uniform, 4 states/molecule, "
                           f"Ecut in {{50,60,70}} Ry, V_KS from each step's rho), Poisson "
                           f"arrivals at {args.serve_rate:g}/s, conjugate gradient (time step "
                           f"{args.serve_dt:g} a.u.) to |F| <= {args.serve_ftol:g}",
               "nmol_hist": {str(k): int(v) for k, v in zip(*np.unique(
                   mix.nmol, return_counts=True))},
               "max_active": args.serve_active, "hbm_budget_gib": args.serve_budget_gib}
    runs_s = {}
    for prec in precisions:
        log(f"[bench] serving ({prec}): {len(mix)} requests")
        r = serving_run(mix, arrive, prec, args.serve_ftol, args.serve_max_steps,
                        int(args.serve_budget_gib * 2**30), args.serve_active, sopts,
                        deadline=args.serve_deadline)
        log(f"[bench] serving {prec}: {r['summary']}")
        serving[prec] = r["summary"]
        runs_s[prec] = r["res"]
    if "mixed" in runs_s and "fp64" in runs_s:
        rm, r64 = runs_s["mixed"], runs_s["fp64"]
        de = [abs(rm[k]["eelec"] - r64[k]["eelec"]) / (3 * mix.nmol[k])
              for k in range(len(mix))
              if k in rm and k in r64 and rm[k]["status"] == 0 and r64[k]["status"] == 0]
        serving["max_dE_per_atom_mixed_vs_fp64"] = max(de) if de else None
    return serving


def cpu_baseline(sysd, nsteps, warmup=3):
    """The C oracle's FP64 steps (V_KS rebuilt from rho each step, as on
    the GPU) on the host cores; V_loc(G) is set up before the clock."""
    import numpy as np
    import oracle
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    nthr = max(1, min(16, ncores))
    oracle.set_threads(nthr)
    o = oracle.System(sysd["grid"], sysd["hmat"], sysd["ecut_ry"], sysd["nstates"], sysd["kx"],
                      sysd["ky"], sysd["kz"])
    t = sysd["pseudo"]
    lp, keep = oracle.local_pseudo(sysd["pos"], sysd["species"], t["gmin"], t["dg"], t["vps"],
                                   t["vps_g0"])
    vloc = o.vloc_g(lp)
    psi = np.array(sysd["psi"], copy=True)
    outs, times = [], []
    for k in range(warmup + nsteps):
        t0 = time.perf_counter()
        outs.append(o.step(None, psi, dt=5.0, ns_tol=1e-15, ns_max_iter=30, vloc=vloc))
        times.append(time.perf_counter() - t0)
    tt = np.array(times[warmup:])
    return dict(value=nsteps / tt.sum(), unit="steps/s", cores=nthr, kind="port",
                sample=f"{warmup} warm-up + {nsteps} timed full FP64 CP steps (V_KS from rho "
                       f"included) of the same workload on the C oracle (OpenMP, {nthr} threads "
                       f"of {ncores} visible cores); value = timed steps / their total time",
                seconds=float(tt.sum()), median_s=round(float(np.median(tt)), 4),
                p99_s=round(float(np.percentile(tt, 99)), 4),
                eelec=[x["eelec"] for x in outs], psi=psi)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="water32")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--cpu-warmup", type=int, default=3)
    ap.add_argument("--quality-steps", type=int, default=100,
                    help="steps of the lockstep mixed-vs-FP64 error trace (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="eager launches only")
    ap.add_argument("--no-autotune", action="store_true",
                    help="heuristic launch shapes instead of the autotuner's measured choice")
    ap.add_argument("--fixed-v", action="store_true",
                    help="time the step with the seeded local potential as a fixed V(r) "
                         "instead of V_KS rebuilt from rho")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="before the runs, copy a buffer of this many MiB three times "
                         "(a plain copy kernel that scales the HBM byte counters)")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the serving sample (0: skip it)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-max-mol", type=int, default=64)
    ap.add_argument("--serve-ftol", type=float, default=1e-5)
    ap.add_argument("--serve-deadline", type=float, default=0.0,
                    help="seconds after which unfinished serving requests are withdrawn "
                         "(0: run every request to the end)")
    ap.add_argument("--serve-dt", type=float, default=DEFAULT_SERVE_DT,
                    help="fictitious time step of the requests' CG (a.u.)")
    ap.add_argument("--serve-max-steps", type=int, default=1500)
    ap.add_argument("--serve-active", type=int, default=8, help="requests on the GPU at once")
    ap.add_argument("--serve-budget-gib", type=float, default=64.0)
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving workload (--serve-requests requests, each "
                         "precision of --serve-precisions) and print its JSON line")
    ap.add_argument("--serve-precisions", default="fp64,mixed")
    ap.add_argument("--ns-mode", type=int, default=None,
                    help="force the Newton-Schulz kernel (ns_mode option; default: autotuned). "
                         "1 = one launch per product, no cooperative launch")
    ap.add_argument("--psir-keep", type=int, default=None,
                    help="psir_keep option: 1 keeps psi(r) between the plane passes, 0 "
                    "recomputes it in the R->G pass (default: the library's, 1)")
    ap.add_argument("--traffic-from", nargs="+", default=None,
                    help="rocprofv3 --pmc output directories (request-size counters): write "
                         "profiles/traffic_<workload>.json from them and exit (no GPU work)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the step on a single "
              "MI355X (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    if args.ns_mode is not None:
        RUN_OPTS["ns_mode"] = args.ns_mode
    if args.psir_keep is not None:
        RUN_OPTS["psir_keep"] = args.psir_keep
    if args.traffic_from:
        def prec(name):
            return "mixed" if "<float" in name else ("fp64" if "<double" in name else None)
        t = traffic_from_counters(args.traffic_from, os.path.join(
            ROOT, "profiles", f"traffic_{args.workload}.json"), prec)
        print(json.dumps({k: t[k] for k in ("fp64", "mixed")}), flush=True)
        return

    import numpy as np
    from openatom_amd import system

    mix = None
    if args.serve_requests > 0:
        log(f"[bench] building the serving mix: {args.serve_requests} requests")
        mix = ServingMix(args.serve_requests, args.serve_rate, args.seed, args.serve_max_mol)
    if args.serve_only:
        if mix is None:
            sys.exit("--serve-only needs --serve-requests > 0")
        out = serving_leg(args, mix, args.serve_precisions.split(","))
        mix.close()
        print(json.dumps({"serving": out}), flush=True)
        return

    if args.copy_calib_mb > 0:
        copy_calibration(args.copy_calib_mb)
    log(f"[bench] building workload {args.workload} (seed {args.seed})")
    sysd = system.build(args.workload, seed=args.seed)
    nst = sysd["nstates"]
    natoms = sysd["natoms"]

    # per precision: an eager run (the throughput reported, with per-stage
    # HIP-event timings for the roofline) and, unless --no-graph, a run that
    # replays the captured step graph, reported beside it
    runs = {}
    vks = not args.fixed_v
    for prec in ("fp64", "mixed"):
        for g in ((0, 1) if not args.no_graph else (0,)):
            log(f"[bench] {prec} run ({'graph' if g else 'eager'})")
            r = run_precision(sysd, prec, args.steps, args.warmup, use_graph=g, vks=vks,
                              autotune=not args.no_autotune)
            log(f"[bench] {prec} {'graph' if g else 'eager'} {r['ms']:.3f} ms/step")
            runs[(prec, g)] = r
    # the same step with V(r) a fixed input (no density-chare work), beside it
    fixed = None
    if vks and not args.no_graph:
        fixed = {}
        for prec in ("fp64", "mixed"):
            r = run_precision(sysd, prec, args.steps, args.warmup, vks=False,
                              autotune=not args.no_autotune)
            log(f"[bench] {prec} fixed V(r) {r['ms']:.3f} ms/step")
            fixed[prec] = {"steps_per_s": round(r["steps_per_s"], 3),
                           "ms_per_step": round(r["ms"], 4)}
    # the same V_KS step with the non-local KB part of the oxygen atoms added
    with_nl = None
    if vks and not args.no_graph:
        with_nl = {}
        for prec in ("fp64", "mixed"):
            r = run_precision(sysd, prec, args.steps, args.warmup, vks=True,
                              autotune=not args.no_autotune, nonlocal_kb=True)
            log(f"[bench] {prec} with non-local KB {r['ms']:.3f} ms/step")
            with_nl[prec] = {"steps_per_s": round(r["steps_per_s"], 3),
                             "ms_per_step": round(r["ms"], 4),
                             "enl_last": r["last"]["enl"],
                             "nonlocal_plus_lambda_ms": round(
                                 r["timings"].get("nonlocal+reduce_lambda", 0.0), 4),
                             "lambda_only_ms": round(
                                 runs[(prec, 0)]["timings"].get("reduce_lambda", 0.0), 4)}
    # the mixed step with the Lambda products in FP64 (gemm_split 0), the
    # default split-BF16 gram (1), and gram and projection split (9)
    lam_forms = None
    if vks and not args.no_graph:
        lam_forms = {}
        for gs in (0, 1, 9):
            r = run_precision(sysd, "mixed", args.steps, args.warmup, vks=True,
                              autotune=not args.no_autotune, gemm_split=gs)
            log(f"[bench] mixed gemm_split={gs} {r['ms']:.3f} ms/step")
            lam_forms[str(gs)] = {"ms_per_step": round(r["ms"], 4),
                                  "lambda_ms": round(r["timings"].get("reduce_lambda", 0.0), 4),
                                  "sd_ms": round(r["timings"].get("sd_sync", 0.0), 4)}
    qt = None
    if vks and args.quality_steps > 0:
        log(f"[bench] quality trace: mixed vs FP64 in lockstep over {args.quality_steps} steps")
        qt = quality_trace(sysd, args.quality_steps)
        log(f"[bench] quality worst: {qt['worst']}")
    r64, rmx = runs[("fp64", 0)], runs[("mixed", 0)]
    r64v, rmxv = r64, rmx
    graph = {p: round(runs[(p, 1)]["steps_per_s"], 3) for p in ("fp64", "mixed")
             if (p, 1) in runs}

    # quality: same start, same number of steps
    de_atom = abs(rmx["last"]["eelec"] - r64["last"]["eelec"]) / natoms
    de_first = abs(rmx["first"]["eelec"] - r64["first"]["eelec"]) / natoms

    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    rl = roofline(rmx, nst)
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        traffic = tr.get("mixed", {}).get(rl["kernel"])
        rl["traffic"] = traffic
    rl64 = roofline(r64, nst)

    serving = None
    if mix is not None:
        serving = serving_leg(args, mix, ("fp64", "mixed"))
        mix.close()
        del mix

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline: {args.cpu_steps} oracle steps")
        cb = cpu_baseline(sysd, args.cpu_steps, args.cpu_warmup)
        # the GPU FP64 path after the same number of steps from the same start
        cpu = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        cpu["seconds"] = round(cb["seconds"], 3)
        cpu["gpu_fp64_first_step_dE_rel_vs_cpu"] = abs(
            r64["first"]["eelec"] - cb["eelec"][0]) / abs(cb["eelec"][0])
        cpu["gpu_mixed_first_step_dE_per_atom_vs_cpu"] = abs(
            rmx["first"]["eelec"] - cb["eelec"][0]) / natoms
        cpu["median_step_s"] = cb["median_s"]
        cpu["p99_step_s"] = cb["p99_s"]

    out = {
        "metric": METRIC,
        "value": round(rmxv["steps_per_s"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(rmxv["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 FFT + f64 (coefficients, GEMMs, Newton-Schulz, energies)",
        "data": ("synthetic (seeded water box, Gaussian coefficients; V(r) = PZ-LDA V_xc + "
                 "Hartree + seeded-Gaussian local pseudopotential rebuilt from each step's rho)"
                 if vks else
                 "synthetic (seeded water box, Gaussian coefficients, fixed Gaussian local "
                 "potential)"),
        "config": {"workload": f"{args.workload}: {sysd['nmol']} H2O, {nst} states at Gamma, "
                               f"{sysd['ecut_ry']:.0f} Ry, {sysd['grid'][0]}^3 grid, "
                               f"{r64['geo']['npts']} stored plane waves/state",
                   "global_batch": nst, "seq_len": r64["geo"]["npts"], "parallelism": "none (1 GPU)"},
        "step_launch": "eager launches on one stream",
        "graph_replay_steps_per_s": graph,
        "fixed_v_steps": fixed,
        "with_nonlocal_kb": with_nl,
        "mixed_lambda_forms": lam_forms,
        "fp64": {"steps_per_s": round(r64v["steps_per_s"], 3), "ms_per_step": round(r64v["ms"], 4),
                 "ns_iters": r64["last"]["ns_iters"], "resid_max": r64["last"]["resid_max"],
                 "roofline": rl64, "timings_ms": r64["timings"],
                 "timed_steps_events_ms": r64["timings_timed"], "launch": r64["launch"],
                 "autotune": r64["tuning"]},
        "speedup_mixed_vs_fp64": round(rmxv["steps_per_s"] / r64v["steps_per_s"], 3),
        "quality": {"dE_per_atom_mixed_vs_fp64_after_steps": de_atom,
                    "dE_per_atom_first_step": de_first,
                    "energies_fp64_last": {k: r64["last"][k] for k in
                                           ("eelec", "eke", "eloc", "ehart", "exc")},
                    "resid_max_mixed": rmx["last"]["resid_max"],
                    "ns_iters_mixed": rmx["last"]["ns_iters"],
                    "steps_compared": args.warmup + args.steps + PROF_STEPS,
                    "per_step_trace": None if qt is None else
                    {k: qt[k] for k in ("steps", "worst", "last", "meets_bars")}},
        "timings_ms": rmx["timings"],
        "timed_steps_events_ms": rmx["timings_timed"],
        "timings_note": ("timed steps record HIP events around the step and the R->G plane pass "
                         "only (timed_steps_events_ms, the roofline's duration); timings_ms is a "
                         "per-stage breakdown from 20 further steps with an event at every stage "
                         "boundary, which add a few us each"),
        "launch": rmx["launch"],
        "autotune": rmx["tuning"],
        "roofline": rl,
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


def traffic_from_counters(dirs, out_path, precision_of):
    """HBM bytes per launch of each kernel from rocprofv3 request-size
    counters, no calibration: reads = 32 RDREQ_32B + 64 RDREQ_64B + 128
    RDREQ_128B, writes = 64 WRREQ_64B + 32 (WRREQ - WRREQ_64B) (TCC/EA
    interface, summed over the TCC instances).  Per kernel: the median of its
    last three dispatches (the timed steps, after the autotuner's trials).
    The runtime's copy kernel of a known size checks the byte counts.
    precision_of(kernel name) -> "fp64" | "mixed" | None."""
    import csv
    import numpy as np
    per = {}
    for d in dirs:
        with open(os.path.join(d, "run_counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                k = (r["Kernel_Name"], int(r["Dispatch_Id"]))
                per.setdefault(k, {})[r["Counter_Name"]] = float(r["Counter_Value"])
    # reads and writes come from separate passes: match dispatches of a kernel by order
    by_kernel = {}
    for (name, did), c in sorted(per.items(), key=lambda x: x[0][1]):
        by_kernel.setdefault(name, []).append(c)
    out = {"unit": "bytes per launch (HBM, TCC/EA request sizes: reads 32/64/128 B, "
                   "writes 32/64 B; no calibration factor)",
           "commands": ["cd /tmp && export TMPDIR=/tmp"] + [
               f"timeout -s KILL 120 rocprofv3 --pmc {c} --output-format csv -d <out>/{d} -o run "
               "-- python3 <repo>/bench.py --steps 2 --warmup 1 --no-cpu --no-graph "
               "--serve-requests 0 --copy-calib-mb 256"
               for c, d in (("TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_64B_sum "
                             "TCC_EA0_RDREQ_128B_sum", "pmc_rd"),
                            ("TCC_EA0_WRREQ_sum TCC_EA0_WRREQ_64B_sum", "pmc_wr"))] + [
               "python3 <repo>/bench.py --traffic-from <out>/pmc_rd <out>/pmc_wr"],
           "per_launch": "median of each kernel's last three dispatches (the timed steps, "
                         "after the autotuner's trials)",
           "fp64": {}, "mixed": {}, "detail": {}}
    for name, cs in by_kernel.items():
        rd = [c for c in cs if "TCC_EA0_RDREQ_sum" in c]
        wr = [c for c in cs if "TCC_EA0_WRREQ_sum" in c]
        rb = [32 * c["TCC_EA0_RDREQ_32B_sum"] + 64 * c["TCC_EA0_RDREQ_64B_sum"] +
              128 * c["TCC_EA0_RDREQ_128B_sum"] for c in rd[-3:]]
        other = [c["TCC_EA0_RDREQ_sum"] - c["TCC_EA0_RDREQ_32B_sum"] - c["TCC_EA0_RDREQ_64B_sum"]
                 - c["TCC_EA0_RDREQ_128B_sum"] for c in rd[-3:]]
        wb = [64 * c["TCC_EA0_WRREQ_64B_sum"] + 32 * (c["TCC_EA0_WRREQ_sum"] -
                                                       c["TCC_EA0_WRREQ_64B_sum"])
              for c in wr[-3:]]
        if not rb or not wb:
            continue
        det = {"read_bytes": float(np.median(rb)), "write_bytes": float(np.median(wb)),
               "reads_of_other_sizes": float(np.median(other)), "launches": len(rd)}
        out["detail"][name] = det
        prec = precision_of(name)
        key = STAGE_OF_KERNEL(name)
        if prec and key:
            out[prec][key] = int(det["read_bytes"] + det["write_bytes"])
        if "copyBuffer" in name:  # the known-size check: the largest runtime copy
            big = max(range(len(rd)), key=lambda i: rd[i]["TCC_EA0_RDREQ_sum"])
            c, w = rd[big], wr[big] if big < len(wr) else wr[-1]
            out["copy_check"] = {
                "read_bytes": 32 * c["TCC_EA0_RDREQ_32B_sum"] + 64 * c["TCC_EA0_RDREQ_64B_sum"]
                + 128 * c["TCC_EA0_RDREQ_128B_sum"],
                "write_bytes": 64 * w["TCC_EA0_WRREQ_64B_sum"] + 32 * (
                    w["TCC_EA0_WRREQ_sum"] - w["TCC_EA0_WRREQ_64B_sum"]),
                "copy_bytes": 256 << 20}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    return out


def STAGE_OF_KERNEL(name):
    """The bench stage a kernel's bytes belong to (the roofline's names).
    k3_r2g's eighth template argument is RHO: true in the fixed-V step's
    R->G pass, which also accumulates the density."""
    if "k2_zfwd" in name:
        return "k_zfwd"
    if "k2_zbwd" in name:
        return "k_zbwd"
    if "k3_g2r" in name:
        return "k_plane_g2r"
    if "k3_r2g" in name:
        targs = name[name.index("<") + 1:].split(", ")
        return "k_plane_r2g_rho" if targs[7] == "true" else "k_plane_r2g"
    return None


def copy_calibration(mib):
    """Three device-to-device copies of `mib` MiB through the HIP runtime (its
    __amd_rocclr_copyBuffer kernel): the known-size dispatches that scale the
    FETCH_SIZE / WRITE_SIZE counters of a --pmc run."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nb = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), nb) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, nb) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, nb, 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copy calibration: 3 x {mib} MiB device-to-device")


if __name__ == "__main__":
    main()
