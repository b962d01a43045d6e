#!/usr/bin/env python3
"""Throughput of the CP electronic step on one MI355X.

One step = one minimisation iteration of the reference's GSpaceDriver order:
G->R state FFTs + rho, V psi + R->G FFTs + kinetic force/EKE, Lambda
projection, steepest-descent update + kx=0 sync, S = Psi^H Psi,
Newton-Schulz (S/2)^{-1/2}, Psi T, and the ||Psi^H Psi/2 - I|| check.

Workload (default water32): 32 H2O in a 9.86 A box, 128 doubly occupied
states at Gamma, 70 Ry (32,961 stored plane waves per state), 100^3 grid,
seeded coefficients and a seeded Gaussian local potential (synthetic data).
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

STAGES = ("k_zfwd", "k_y_fwd", "k_xc2r", "k_xr2c", "k_y_bwd", "k_zbwd",
          "k_plane_fwd", "k_plane_bwd", "k_plane_g2r", "k_plane_r2g")


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
    if stage in ("k_plane_fwd", "k_plane_bwd", "k_plane_g2r", "k_plane_r2g"):
        return by + bx
    return bx


def plane_bytes(stage, geo, csz, nst, launch):
    """HBM bytes a two-factor fused plane pass must move (psi(r) stays on
    chip): G->R reads the populated z-lines of every state (n_lines Nz c)
    and writes one rho partial (N doubles) per state chunk; R->G reads and
    writes the z-lines and reads V (N doubles) once per chunk."""
    n = geo["nx"] * geo["ny"] * geo["nz"]
    lines = geo["nlines"] * geo["nz"] * csz
    if stage == "k_plane_g2r":
        return nst * lines + launch["nchunk"] * n * 8
    return 2 * nst * lines + launch["pcp"] * n * 8


def run_precision(sysd, precision, steps, warmup, device=0, use_graph=0):
    from openatom_amd.step import CPContext
    # The reference step has no ||Psi^H Psi - I|| evaluation (Ortho checks the
    # T iterates); it is measured once after the timed steps instead.
    c = CPContext(sysd["grid"], sysd["hmat"], sysd["ecut_ry"], sysd["nstates"], sysd["kx"],
                  sysd["ky"], sysd["kz"], precision=precision, device=device, dt=5.0,
                  ns_tol=1e-15 if precision == "fp64" else 1e-14, ns_max_iter=30,
                  want_resid=0, use_graph=use_graph)
    c.set_psi(sysd["psi"])
    c.set_vks(sysd["vks"])
    first = c.step()
    for _ in range(max(0, warmup - 1)):
        c.step(fetch=False)
    c.sync()
    c.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        c.step(fetch=False)
    c.sync()
    t1 = time.perf_counter()
    last = c.result()
    tim = c.timings()
    last["resid_max"], last["resid_rms"] = c.residual()
    geo = dict(nx=c.nx, ny=c.ny, nz=c.nz, nlines=c.nlines, nplane_x=c.nplane_x, npts=c.npts)
    psi = c.get_psi()
    cfg = dict(c.config)
    c.close()
    return dict(precision=precision, ms=(t1 - t0) * 1e3 / steps, steps_per_s=steps / (t1 - t0),
                first=first, last=last, timings=tim, geo=geo, psi=psi, launch=cfg)


def roofline(res, nst, traffic=None):
    """Dominant kernel of the timed region and its roofline."""
    tim = res["timings"]
    fp32 = res["precision"] == "mixed"
    csz, rsz = (8, 4) if fp32 else (16, 8)
    geo = res["geo"]
    K = 2 * geo["npts"]
    cands = {k: tim[k] for k in STAGES + ("overlap", "rotate") if k in tim}
    dom = max(cands, key=cands.get)
    ms = cands[dom]
    if dom in ("k_plane_g2r", "k_plane_r2g"):
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


def cpu_baseline(sysd, nsteps):
    import numpy as np
    import oracle
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    nthr = max(1, min(16, ncores))
    oracle.set_threads(nthr)
    o = oracle.System(sysd["grid"], sysd["hmat"], sysd["ecut_ry"], sysd["nstates"], sysd["kx"],
                      sysd["ky"], sysd["kz"])
    psi = np.array(sysd["psi"], copy=True)
    t0 = time.perf_counter()
    outs = []
    for _ in range(nsteps):
        outs.append(o.step(sysd["vks"], psi, dt=5.0, ns_tol=1e-15, ns_max_iter=30))
    dt = time.perf_counter() - t0
    return dict(value=nsteps / dt, unit="steps/s", cores=nthr, kind="port",
                sample=f"{nsteps} full FP64 CP steps of the same workload on the C oracle "
                       f"(OpenMP, {nthr} threads of {ncores} visible cores)",
                seconds=dt, eelec=[x["eelec"] for x in outs], psi=psi)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="water32")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="eager launches only")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="before the runs, copy a buffer of this many MiB three times "
                         "(a plain copy kernel that scales the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the step on a single "
              "MI355X (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import oracle
    from openatom_amd import system

    if args.copy_calib_mb > 0:
        copy_calibration(args.copy_calib_mb)
    log(f"[bench] building workload {args.workload} (seed {args.seed})")
    sysd = system.build(args.workload, seed=args.seed, kvec_source=oracle.stored_kvecs)
    nst = sysd["nstates"]
    natoms = sysd["natoms"]

    # per precision: an eager run (per-stage HIP-event timings for the
    # roofline) and, unless --no-graph, a run that replays the captured step
    # graph (the throughput reported)
    runs = {}
    for prec in ("fp64", "mixed"):
        for g in ((0, 1) if not args.no_graph else (0,)):
            log(f"[bench] {prec} run ({'graph' if g else 'eager'})")
            r = run_precision(sysd, prec, args.steps, args.warmup, use_graph=g)
            log(f"[bench] {prec} {'graph' if g else 'eager'} {r['ms']:.3f} ms/step")
            runs[(prec, g)] = r
    r64, rmx = runs[("fp64", 0)], runs[("mixed", 0)]
    r64v = runs.get(("fp64", 1), r64)
    rmxv = runs.get(("mixed", 1), rmx)

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

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline: {args.cpu_steps} oracle steps")
        cb = cpu_baseline(sysd, args.cpu_steps)
        # the GPU FP64 path after the same number of steps from the same start
        cpu = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        cpu["seconds"] = round(cb["seconds"], 3)
        cpu["gpu_fp64_first_step_dE_rel_vs_cpu"] = abs(
            r64["first"]["eelec"] - cb["eelec"][0]) / abs(cb["eelec"][0])

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
        "data": "synthetic (seeded water box, Gaussian coefficients, Gaussian local potential)",
        "config": {"workload": f"{args.workload}: {sysd['nmol']} H2O, {nst} states at Gamma, "
                               f"{sysd['ecut_ry']:.0f} Ry, {sysd['grid'][0]}^3 grid, "
                               f"{r64['geo']['npts']} stored plane waves/state",
                   "global_batch": nst, "seq_len": r64["geo"]["npts"], "parallelism": "none (1 GPU)"},
        "step_launch": "HIP graph replay" if rmxv is not rmx else "eager launches",
        "eager": {"mixed_steps_per_s": round(rmx["steps_per_s"], 3),
                  "fp64_steps_per_s": round(r64["steps_per_s"], 3)},
        "fp64": {"steps_per_s": round(r64v["steps_per_s"], 3), "ms_per_step": round(r64v["ms"], 4),
                 "ns_iters": r64["last"]["ns_iters"], "resid_max": r64["last"]["resid_max"],
                 "roofline": rl64, "timings_ms": r64["timings"], "launch": r64["launch"]},
        "speedup_mixed_vs_fp64": round(rmxv["steps_per_s"] / r64v["steps_per_s"], 3),
        "quality": {"dE_per_atom_mixed_vs_fp64_after_steps": de_atom,
                    "dE_per_atom_first_step": de_first,
                    "resid_max_mixed": rmx["last"]["resid_max"],
                    "ns_iters_mixed": rmx["last"]["ns_iters"],
                    "steps_compared": args.warmup + args.steps},
        "timings_ms": rmx["timings"],
        "launch": rmx["launch"],
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


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
