"""Host-side logic of the request loop, on the CPU: how the scheduler reads
the event after a request's step (compiled with the host compiler from the
library's own header), and how the Python side pairs native results with
the caller's tags and psi arrays."""
import ctypes
import os
import subprocess

import numpy as np

from openatom_amd import serve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_event_status_classification(tmp_path):
    """hipSuccess retires a step, hipErrorNotReady keeps polling, and any
    other status fails the request (it must not stay in flight for ever)."""
    exe = tmp_path / "serve_poll_check"
    subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(ROOT, "openatom_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "serve_poll_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "serve poll: ok" in out.stdout


class _FakeLib:
    """Stands in for the native server: records submissions and returns
    results for the native tags it was given, in reverse order."""

    def __init__(self):
        self.tags = []

    def oa_server_submit(self, h, req):
        r = req._obj
        self.tags.append(int(r.tag))
        # the scheduler writes psi_out of each request
        n = 2 * r.sys.nstates * r.sys.npts
        out = (ctypes.c_double * n).from_address(r.psi_out)
        for i in range(n):
            out[i] = float(r.tag)
        return 0

    def oa_server_poll(self, h, buf, nmax):
        n = 0
        while self.tags and n < nmax:
            buf[n].tag = self.tags.pop()
            buf[n].status = 0
            n += 1
        return n


def test_duplicate_tags_keep_their_own_arrays():
    s = dict(hmat=np.eye(3).reshape(9), grid=(4, 4, 4), ecut_ry=1.0, nstates=2,
             kx=np.array([0, 0, 1], np.int32), ky=np.array([0, 1, 0], np.int32),
             kz=np.array([0, 0, 0], np.int32), psi=np.ones((2, 3), np.complex128),
             vks=np.zeros((4, 4, 4)))
    srv = serve.Server.__new__(serve.Server)
    srv._L = _FakeLib()
    srv.h = None
    import itertools
    srv._ids = itertools.count(1)
    srv._pending = {}
    srv.submit(s, "same", want_psi=True)
    srv.submit(s, "same", want_psi=True)
    res = srv.poll()
    assert [r["tag"] for r in res] == ["same", "same"]
    # each result carries the array its own native request wrote
    vals = sorted(float(r["psi"].real.flat[0]) for r in res)
    assert vals == [1.0, 2.0]
    assert not srv._pending


def test_serving_grids_hold_the_density_sphere():
    """Every shape of the serving mix (4-64 molecules, 50/60/70 Ry) gets a
    grid with N >= 2 kr + 1, kr the largest component of the density sphere
    |G|^2 <= 4 Ecut found by enumerating it (the library refuses smaller
    grids), and the smallest compiled size that does."""
    from openatom_amd import _lib, system
    for nmol in range(4, 65):
        for e in (50.0, 60.0, 70.0):
            w = system.serving_workload(nmol, e)
            L = w["box_ang"] * system.BOHR_PER_ANG
            kx, ky, kz = _lib.kvecs(4.0 * e, np.diag([L, L, L]))
            need = 2 * int(max(kx.max(), np.abs(ky).max(), np.abs(kz).max())) + 1
            assert w["grid"] >= need, (nmol, e, w["grid"], need)
            smaller = [n for n in system.FFT_SIZES if need <= n < w["grid"]]
            assert not smaller, (nmol, e, w["grid"], smaller)
