"""The C++ shim a maintainer adds to the reference tree (tests/cpp/
OaStepBackend.h, the class INTEGRATION.md describes), compiled as an
ordinary C++ caller of the C ABI with a CkAbort stub.  CPU: it compiles and
links against liboa_step.so.  GPU: one staged step and one whole step
through it against the oracle at 1e-10."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from openatom_amd import _lib
from tests.helpers import rel, small_system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "openatom_amd")


def build_shim(out):
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "tests", "cpp"),
                    os.path.join(ROOT, "tests", "cpp", "shim_step.cpp"), "-L" + LIBDIR,
                    "-loa_step", "-Wl,-rpath," + LIBDIR, "-L/opt/rocm/lib", "-lamdhip64", "-o",
                    str(out)], check=True)


def test_shim_compiles_and_links(tmp_path):
    _lib.load()  # the library is built
    exe = tmp_path / "shim_step"
    build_shim(exe)
    assert os.access(exe, os.X_OK)


@pytest.mark.gpu
def test_shim_step_matches_oracle(gpu, tmp_path):
    s = small_system(nstates=11, seed=17)
    d = str(tmp_path)
    nst, npts = s["nstates"], len(s["kx"])
    with open(os.path.join(d, "sys.txt"), "w") as f:
        f.write(f"{s['grid'][0]} {s['grid'][1]} {s['grid'][2]} {nst} {npts} "
                f"{s['ecut_ry']!r} 5.0\n")
    np.asarray(s["hmat"], np.float64).tofile(os.path.join(d, "hmat.f64"))
    for a in ("kx", "ky", "kz"):
        np.asarray(s[a], np.int32).tofile(os.path.join(d, f"{a}.i32"))
    np.ascontiguousarray(s["psi"], np.complex128).tofile(os.path.join(d, "psi.f64"))
    np.ascontiguousarray(s["vks"], np.float64).tofile(os.path.join(d, "vks.f64"))
    exe = tmp_path / "shim_step"
    build_shim(exe)
    r = subprocess.run([str(exe), d], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    nz, ny, nx = s["grid"][2], s["grid"][1], s["grid"][0]
    rho = np.fromfile(os.path.join(d, "rho.f64")).reshape(nz, ny, nx)
    lam = np.fromfile(os.path.join(d, "lam.f64")).reshape(nst, nst)
    S = np.fromfile(os.path.join(d, "S.f64")).reshape(nst, nst)
    T = np.fromfile(os.path.join(d, "T.f64")).reshape(nst, nst)
    psi1 = np.fromfile(os.path.join(d, "psi1.f64"), np.complex128).reshape(nst, npts)
    psi2 = np.fromfile(os.path.join(d, "psi2.f64"), np.complex128).reshape(nst, npts)
    eke = float(open(os.path.join(d, "eke.txt")).read())
    eelec, fnorm, _ = open(os.path.join(d, "step.txt")).read().split()
    o = oracle.System(s["grid"], s["hmat"], s["ecut_ry"], nst, s["kx"], s["ky"], s["kz"])
    pr = o.g2r_all(s["psi"])
    Fo = o.vpsi_r2g_all(pr, s["vks"])
    eko = o.eke_force(s["psi"], Fo)
    lamo = o.lambda_(s["psi"], Fo)
    psio = s["psi"].copy()
    ro = o.step(s["vks"], psio, dt=5.0, ns_tol=1e-15, ns_max_iter=30, want=("rho", "S", "T"))
    errs = dict(rho=rel(rho, o.density(pr)), eke=abs(eke - eko) / abs(eko),
                lam=rel(lam, lamo), S=rel(S, ro["S"]), T=rel(T, ro["T"]), psi1=rel(psi1, psio))
    ro2 = o.step(s["vks"], psio, dt=5.0, ns_tol=1e-15, ns_max_iter=30)
    errs.update(eelec=abs(float(eelec) - ro2["eelec"]) / abs(ro2["eelec"]),
                fnorm=abs(float(fnorm) - ro2["fnorm"]) / ro2["fnorm"], psi2=rel(psi2, psio))
    print("[shim]", " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    assert all(v < 1e-10 for v in errs.values()), errs
