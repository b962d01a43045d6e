"""Python side of the request loop (include/oa_serve.h): independent
fixed-ion electronic minimisations of different sizes on one GPU, admitted
by HBM budget and advanced concurrently by the library's scheduler thread."""
import ctypes
import itertools

import numpy as np

from . import _lib
from ._lib import OaError, OaLocalPseudo, OaOptions, OaSystem, ptr


class OaRequest(ctypes.Structure):
    _fields_ = [("sys", OaSystem), ("opt", OaOptions), ("psi0", ctypes.c_void_p),
                ("vks", ctypes.c_void_p), ("f_tol", ctypes.c_double),
                ("max_steps", ctypes.c_int), ("psi_out", ctypes.c_void_p),
                ("tag", ctypes.c_longlong), ("fp64_below", ctypes.c_double),
                ("pseudo", ctypes.POINTER(OaLocalPseudo))]


class OaResult(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_longlong), ("status", ctypes.c_int), ("steps", ctypes.c_int),
                ("eelec", ctypes.c_double), ("fnorm", ctypes.c_double),
                ("queue_ms", ctypes.c_double), ("run_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("hbm_bytes", ctypes.c_size_t),
                ("fp64_steps", ctypes.c_int), ("reused", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _bind(L):
    if getattr(L, "_serve_bound", False):
        return L
    vp = ctypes.c_void_p
    L.oa_ctx_bytes.argtypes = [ctypes.POINTER(OaSystem), ctypes.POINTER(OaOptions)]
    L.oa_ctx_bytes.restype = ctypes.c_size_t
    L.oa_server_create.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                   ctypes.POINTER(vp)]
    L.oa_server_create.restype = ctypes.c_int
    L.oa_server_submit.argtypes = [vp, ctypes.POINTER(OaRequest)]
    L.oa_server_submit.restype = ctypes.c_int
    L.oa_server_poll.argtypes = [vp, ctypes.POINTER(OaResult), ctypes.c_int]
    L.oa_server_poll.restype = ctypes.c_int
    L.oa_server_drain.argtypes = [vp]
    L.oa_server_drain.restype = ctypes.c_int
    L.oa_server_cancel.argtypes = [vp]
    L.oa_server_cancel.restype = ctypes.c_int
    L.oa_server_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.oa_server_stats.restype = ctypes.c_int
    L.oa_server_destroy.argtypes = [vp]
    L.oa_server_destroy.restype = ctypes.c_int
    L._serve_bound = True
    return L


def _system_struct(s, keep):
    sysd = OaSystem()
    for i, v in enumerate(np.asarray(s["hmat"], np.float64).reshape(9)):
        sysd.hmat[i] = float(v)
    for i, v in enumerate(s["grid"]):
        sysd.grid[i] = int(v)
    sysd.ecut_ry = float(s["ecut_ry"])
    sysd.nstates = int(s["nstates"])
    k = [np.ascontiguousarray(s[a], np.int32) for a in ("kx", "ky", "kz")]
    keep.extend(k)
    sysd.npts = int(k[0].size)
    sysd.kx, sysd.ky, sysd.kz = (ptr(a) for a in k)
    sysd.occ = None
    return sysd


def _options(precision="fp64", **opts):
    o = _lib.default_options()
    o.precision = _lib.OA_MIXED if precision == "mixed" else _lib.OA_FP64
    for key, v in opts.items():
        if not hasattr(o, key):
            raise TypeError(f"unknown option {key}")
        setattr(o, key, v)
    return o


def ctx_bytes(system, precision="fp64", **opts):
    """Device bytes one context of this system would hold."""
    L = _bind(_lib.load())
    keep = []
    sysd = _system_struct(system, keep)
    o = _options(precision, **opts)
    return int(L.oa_ctx_bytes(ctypes.byref(sysd), ctypes.byref(o)))


class Server:
    """Request loop on one GPU.  submit() returns immediately; results come
    back from poll()/drain() in completion order, keyed by tag."""

    def __init__(self, device=0, hbm_budget=32 << 30, max_active=4):
        self._L = _bind(_lib.load())
        h = ctypes.c_void_p()
        st = self._L.oa_server_create(int(device), int(hbm_budget), int(max_active),
                                      ctypes.byref(h))
        if st != _lib.OA_OK:
            raise OaError(st, "oa_server_create failed")
        self.h = h
        # The native tag of a request is an id unique within this server; it
        # maps back to the caller's tag and to the arrays the scheduler may
        # still write (psi_out), which stay alive until the result returns.
        self._ids = itertools.count(1)
        self._pending = {}  # native id -> (caller's tag, arrays kept alive, psi_out)

    def submit(self, system, tag, f_tol=1e-5, max_steps=1000, precision="fp64", want_psi=False,
               fp64_below=None, scf=False, **opts):
        """fp64_below: mixed precision continues in FP64 once |F| reaches
        it or stalls (default 30 f_tol for mixed; 0 never switches).
        scf: the request carries its ions (system["pos"], ["species"],
        ["pseudo"]) and every step rebuilds V_KS from its own density;
        otherwise system["vks"] is a fixed V(r).  Tags need not be unique."""
        keep = []
        req = OaRequest()
        req.sys = _system_struct(system, keep)
        req.opt = _options(precision, **opts)
        psi0 = np.ascontiguousarray(system["psi"], np.complex128)
        keep.append(psi0)
        req.psi0 = ptr(psi0)
        if scf:
            t = system["pseudo"]
            lp, arrs = _lib.local_pseudo_struct(system["pos"], system["species"], t["gmin"],
                                                t["dg"], t["vps"], t["vps_g0"])
            keep += [lp, arrs]
            req.vks = None
            req.pseudo = ctypes.pointer(lp)
        else:
            vks = np.ascontiguousarray(system["vks"], np.float64)
            keep.append(vks)
            req.vks = ptr(vks)
            req.pseudo = None
        req.f_tol = float(f_tol)
        req.max_steps = int(max_steps)
        out = np.zeros_like(psi0) if want_psi else None
        req.psi_out = ptr(out)
        nid = next(self._ids)
        req.tag = nid
        if fp64_below is None:
            fp64_below = 30.0 * f_tol if precision == "mixed" else 0.0
        req.fp64_below = float(fp64_below)
        st = self._L.oa_server_submit(self.h, ctypes.byref(req))
        if st != _lib.OA_OK:
            raise OaError(st, "oa_server_submit failed")
        self._pending[nid] = (tag, keep, out)

    def poll(self, max_results=64):
        buf = (OaResult * max_results)()
        n = self._L.oa_server_poll(self.h, buf, max_results)
        out = []
        for i in range(n):
            d = buf[i].as_dict()
            tag, _, psi = self._pending.pop(d["tag"])
            d["tag"] = tag
            d["psi"] = psi
            out.append(d)
        return out

    def drain(self):
        st = self._L.oa_server_drain(self.h)
        if st != _lib.OA_OK:
            raise OaError(st, "oa_server_drain failed")
        res = []
        while True:
            r = self.poll()
            if not r:
                return res
            res += r

    def cancel(self):
        """Withdraw every request not finished yet (results with status
        OA_ECANCELED: queued ones at once, admitted ones after their step in
        flight)."""
        st = self._L.oa_server_cancel(self.h)
        if st != _lib.OA_OK:
            raise OaError(st, "oa_server_cancel failed")

    def stats(self):
        a = (ctypes.c_double * 8)()
        self._L.oa_server_stats(self.h, a)
        keys = ("active", "queued", "finished", "hbm_bytes", "peak_active", "peak_hbm_bytes",
                "steps_run", "submitted")
        return dict(zip(keys, list(a)))

    def close(self):
        if getattr(self, "h", None):
            self._L.oa_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
