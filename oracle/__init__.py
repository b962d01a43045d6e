"""ctypes view of the CPU oracle (oracle/liboa_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the host baseline.
The product library never loads it.  Build with ``make -C oracle``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboa_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(_LIB)
        c_int, c_double, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        L.oo_kmax.argtypes = [c_double, _dp, _ip]
        L.oo_kvec_half.argtypes = [c_double, _dp, _ip, vp, vp, vp]
        L.oo_kvec_half.restype = c_int
        L.oo_flip.argtypes = [c_int, _ip, _ip, _ip, vp, vp, vp]
        L.oo_flip.restype = c_int
        L.oo_stored_kvecs.argtypes = [c_double, _dp, vp, vp, vp]
        L.oo_stored_kvecs.restype = c_int
        L.oo_inv3.argtypes = [_dp, _dp]
        L.oo_inv3.restype = c_double
        L.oo_fft1d.argtypes = [c_int, c_int, _dp]
        L.oo_dft1d.argtypes = [c_int, c_int, _dp, _dp]
        L.oo_create.argtypes = [c_int, c_int, c_int, _dp, c_double, c_int, vp,
                                c_int, _ip, _ip, _ip]
        L.oo_create.restype = vp
        L.oo_destroy.argtypes = [vp]
        for name in ("oo_nlines", "oo_nplane_x", "oo_nkx0", "oo_nkx0_red"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = c_int
        L.oo_volume.argtypes = [vp]
        L.oo_volume.restype = c_double
        L.oo_g2.argtypes = [vp, _dp]
        L.oo_g2r.argtypes = [vp, _dp, _dp]
        L.oo_g2r_all.argtypes = [vp, _dp, _dp]
        L.oo_density.argtypes = [vp, _dp, _dp]
        L.oo_vpsi_r2g.argtypes = [vp, c_int, _dp, _dp, _dp]
        L.oo_vpsi_r2g_all.argtypes = [vp, _dp, _dp, _dp]
        L.oo_eke_force.argtypes = [vp, _dp, _dp]
        L.oo_eke_force.restype = c_double
        L.oo_lambda.argtypes = [vp, _dp, _dp, _dp]
        L.oo_lambda_apply.argtypes = [vp, _dp, _dp, _dp]
        L.oo_cmass.argtypes = [vp, c_double, c_double, _dp]
        L.oo_sd_update.argtypes = [vp, _dp, c_double, _dp, _dp]
        L.oo_red_sync.argtypes = [vp, _dp]
        L.oo_overlap.argtypes = [vp, _dp, _dp]
        L.oo_ns.argtypes = [c_int, _dp, c_double, c_int, _dp,
                            ctypes.POINTER(c_double)]
        L.oo_ns.restype = c_int
        L.oo_rotate.argtypes = [vp, _dp, _dp, _dp]
        L.oo_eloc.argtypes = [vp, _dp, _dp]
        L.oo_eloc.restype = c_double
        L.oo_step.argtypes = [vp, ctypes.POINTER(StepOpts), _dp, _dp, vp, vp,
                              vp, vp, ctypes.POINTER(StepOut)]
        L.oo_step.restype = c_int
        L.oo_cg_gamma.argtypes = [ctypes.POINTER(CgState), c_double]
        L.oo_cg_gamma.restype = c_double
        L.oo_cg_update.argtypes = [vp, _dp, c_double, c_double, _dp, _dp, _dp]
        L.oo_vps_eval.argtypes = [ctypes.POINTER(LocalPseudo), c_int, c_double]
        L.oo_vps_eval.restype = c_double
        L.oo_vloc_g.argtypes = [vp, ctypes.POINTER(LocalPseudo), _dp]
        L.oo_pz_lda.argtypes = [ctypes.c_size_t, _dp, _dp, _dp]
        L.oo_vks.argtypes = [vp, _dp, vp, _dp, ctypes.POINTER(VksEnergies)]
        L.oo_nl_proj.argtypes = [ctypes.POINTER(NonLocal), c_int, c_double]
        L.oo_nl_proj.restype = c_double
        L.oo_nl_zmat.argtypes = [vp, ctypes.POINTER(NonLocal), _dp, _dp, vp]
        L.oo_nl_apply.argtypes = [vp, ctypes.POINTER(NonLocal), _dp, vp, vp]
        L.oo_nl_apply.restype = c_double
        L.oo_set_threads.argtypes = [c_int]
        L.oo_get_threads.restype = c_int
        _lib = L
    return _lib


class CgState(ctypes.Structure):
    """oo_cg_state: CG iteration counter, resets, |F|^2 and |F| now and before."""
    _fields_ = [("iter", ctypes.c_int), ("num_reset", ctypes.c_int),
                ("fovlap", ctypes.c_double), ("fovlap_old", ctypes.c_double),
                ("fmag", ctypes.c_double), ("fmag_old", ctypes.c_double)]


class StepOpts(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("tau_fs", ctypes.c_double),
                ("ecut_mass", ctypes.c_double), ("ns_tol", ctypes.c_double),
                ("ns_max_iter", ctypes.c_int), ("use_lambda", ctypes.c_int),
                ("use_cg", ctypes.c_int), ("cg_state", ctypes.POINTER(CgState)),
                ("cg", ctypes.c_void_p), ("vks_from_rho", ctypes.c_int),
                ("vloc", ctypes.c_void_p), ("vks_out", ctypes.c_void_p),
                ("nl", ctypes.c_void_p)]


class StepOut(ctypes.Structure):
    _fields_ = [("eke", ctypes.c_double), ("eloc", ctypes.c_double),
                ("eelec", ctypes.c_double), ("fnorm", ctypes.c_double),
                ("ns_iters", ctypes.c_int), ("ns_err", ctypes.c_double),
                ("ehart", ctypes.c_double), ("exc", ctypes.c_double),
                ("enl", ctypes.c_double)]


class LocalPseudo(ctypes.Structure):
    """oo_local_pseudo: atoms, species and the spline tables of the local
    pseudopotential form factors."""
    _fields_ = [("natoms", ctypes.c_int), ("nspecies", ctypes.c_int),
                ("nspline", ctypes.c_int), ("pos", ctypes.c_void_p),
                ("species", ctypes.c_void_p), ("gmin", ctypes.c_double),
                ("dg", ctypes.c_double), ("vps", ctypes.c_void_p),
                ("vps_g0", ctypes.c_void_p)]


class NonLocal(ctypes.Structure):
    """oo_nonlocal: non-local atoms, species, the spline tables of the l = 0
    KB projectors and their normalisations."""
    _fields_ = [("natoms", ctypes.c_int), ("nspecies", ctypes.c_int),
                ("nspline", ctypes.c_int), ("pos", ctypes.c_void_p),
                ("species", ctypes.c_void_p), ("gmin", ctypes.c_double),
                ("dg", ctypes.c_double), ("proj", ctypes.c_void_p),
                ("proj_g0", ctypes.c_void_p), ("vnorm", ctypes.c_void_p)]


class VksEnergies(ctypes.Structure):
    _fields_ = [("ehart", ctypes.c_double), ("exc", ctypes.c_double),
                ("eext", ctypes.c_double), ("muxc", ctypes.c_double)]


def local_pseudo(pos, species, gmin, dg, vps, vps_g0):
    """LocalPseudo over numpy arrays; returns (struct, arrays kept alive).
    vps: [nspecies][4][nspline] spline coefficients c0..c3."""
    keep = [np.ascontiguousarray(pos, np.float64).reshape(-1, 3),
            np.ascontiguousarray(species, np.int32),
            np.ascontiguousarray(vps, np.float64),
            np.ascontiguousarray(vps_g0, np.float64)]
    lp = LocalPseudo(keep[0].shape[0], keep[2].shape[0], keep[2].shape[2],
                     _ptr(keep[0]), _ptr(keep[1]), float(gmin), float(dg),
                     _ptr(keep[2]), _ptr(keep[3]))
    return lp, keep


def nonlocal_(pos, species, gmin, dg, proj, proj_g0, vnorm):
    """NonLocal over numpy arrays; returns (struct, arrays kept alive).
    proj: [nspecies][4][nspline] spline coefficients c0..c3."""
    keep = [np.ascontiguousarray(pos, np.float64).reshape(-1, 3),
            np.ascontiguousarray(species, np.int32),
            np.ascontiguousarray(proj, np.float64),
            np.ascontiguousarray(proj_g0, np.float64),
            np.ascontiguousarray(vnorm, np.float64)]
    nl = NonLocal(keep[0].shape[0], keep[2].shape[0], keep[2].shape[2],
                  _ptr(keep[0]), _ptr(keep[1]), float(gmin), float(dg),
                  _ptr(keep[2]), _ptr(keep[3]), _ptr(keep[4]))
    return nl, keep


def nl_proj(nl, species, g):
    return lib().oo_nl_proj(ctypes.byref(nl), int(species), float(g))


def pz_lda(rho):
    """(vxc, sums) of CP_exc_calc on a flat density array."""
    r = np.ascontiguousarray(rho, np.float64).reshape(-1)
    v = np.zeros_like(r)
    sums = np.zeros(3)
    lib().oo_pz_lda(r.size, r, v, sums)
    return v.reshape(np.shape(rho)), sums


def vps_eval(lp, species, g):
    return lib().oo_vps_eval(ctypes.byref(lp), int(species), float(g))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def stored_kvecs(ecut_ry, hmat):
    """Stored (kx, ky, kz) of the reference layout for a cell/cutoff."""
    h = np.ascontiguousarray(hmat, dtype=np.float64).reshape(9)
    n = lib().oo_stored_kvecs(ecut_ry, h, None, None, None)
    kx = np.zeros(n, np.int32); ky = np.zeros(n, np.int32); kz = np.zeros(n, np.int32)
    lib().oo_stored_kvecs(ecut_ry, h, _ptr(kx), _ptr(ky), _ptr(kz))
    return kx, ky, kz


def fft1d(x, sign):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().oo_fft1d(a.size, sign, a.view(np.float64))
    return a


def dft1d(x, sign):
    a = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(a)
    lib().oo_dft1d(a.size, sign, a.view(np.float64), out.view(np.float64))
    return out


class System:
    """One Gamma-point system on the oracle (reference order of points)."""

    def __init__(self, grid, hmat, ecut_ry, nstates, kx, ky, kz, occ=None):
        self.nx, self.ny, self.nz = (int(g) for g in grid)
        self.nst = int(nstates)
        self.kx = np.ascontiguousarray(kx, np.int32)
        self.ky = np.ascontiguousarray(ky, np.int32)
        self.kz = np.ascontiguousarray(kz, np.int32)
        self.npts = self.kx.size
        self._occ = None if occ is None else np.ascontiguousarray(occ, np.float64)
        h = np.ascontiguousarray(hmat, np.float64).reshape(9)
        self.h = lib().oo_create(self.nx, self.ny, self.nz, h, float(ecut_ry),
                                 self.nst, _ptr(self._occ), self.npts,
                                 self.kx, self.ky, self.kz)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oo_destroy(self.h)
            self.h = None

    @property
    def ngrid(self):
        return self.nx * self.ny * self.nz

    @property
    def nlines(self):
        return lib().oo_nlines(self.h)

    @property
    def nplane_x(self):
        return lib().oo_nplane_x(self.h)

    @property
    def nkx0(self):
        return lib().oo_nkx0(self.h)

    @property
    def volume(self):
        return lib().oo_volume(self.h)

    def g2(self):
        out = np.zeros(self.npts)
        lib().oo_g2(self.h, out)
        return out

    @staticmethod
    def _c(psi):
        return np.ascontiguousarray(psi, np.complex128).view(np.float64)

    def g2r(self, c):
        out = np.zeros((self.nz, self.ny, self.nx))
        lib().oo_g2r(self.h, self._c(c).reshape(-1), out.reshape(-1))
        return out

    def g2r_all(self, psi):
        out = np.zeros((self.nst, self.nz, self.ny, self.nx))
        lib().oo_g2r_all(self.h, self._c(psi).reshape(-1), out.reshape(-1))
        return out

    def density(self, psir):
        out = np.zeros((self.nz, self.ny, self.nx))
        lib().oo_density(self.h, np.ascontiguousarray(psir).reshape(-1), out.reshape(-1))
        return out

    def vpsi_r2g(self, istate, psir_s, v):
        f = np.zeros(self.npts, np.complex128)
        lib().oo_vpsi_r2g(self.h, istate, np.ascontiguousarray(psir_s).reshape(-1),
                          np.ascontiguousarray(v).reshape(-1), f.view(np.float64))
        return f

    def vpsi_r2g_all(self, psir, v):
        f = np.zeros((self.nst, self.npts), np.complex128)
        lib().oo_vpsi_r2g_all(self.h, np.ascontiguousarray(psir).reshape(-1),
                              np.ascontiguousarray(v).reshape(-1),
                              f.view(np.float64).reshape(-1))
        return f

    def eke_force(self, psi, f):
        """Applies addForces + CP_eke_calc to f in place; returns EKE."""
        return lib().oo_eke_force(self.h, self._c(psi).reshape(-1),
                                  f.view(np.float64).reshape(-1))

    def lambda_(self, psi, f):
        lam = np.zeros((self.nst, self.nst))
        lib().oo_lambda(self.h, self._c(psi).reshape(-1), self._c(f).reshape(-1),
                        lam.reshape(-1))
        return lam

    def lambda_apply(self, psi, lam, f):
        lib().oo_lambda_apply(self.h, self._c(psi).reshape(-1),
                              np.ascontiguousarray(lam).reshape(-1),
                              f.view(np.float64).reshape(-1))

    def cmass(self, tau_fs=25.0, ecut_mass=2.0):
        out = np.zeros(self.npts)
        lib().oo_cmass(self.h, tau_fs, ecut_mass, out)
        return out

    def sd_update(self, cmass, dt, f, psi):
        lib().oo_sd_update(self.h, cmass, dt, self._c(f).reshape(-1),
                           psi.view(np.float64).reshape(-1))

    def red_sync(self, psi):
        lib().oo_red_sync(self.h, psi.view(np.float64).reshape(-1))

    def overlap(self, psi):
        S = np.zeros((self.nst, self.nst))
        lib().oo_overlap(self.h, self._c(psi).reshape(-1), S.reshape(-1))
        return S

    def rotate(self, psi, T):
        out = np.zeros((self.nst, self.npts), np.complex128)
        lib().oo_rotate(self.h, self._c(psi).reshape(-1),
                        np.ascontiguousarray(T, np.float64).reshape(-1),
                        out.view(np.float64).reshape(-1))
        return out

    def eloc(self, v, rho):
        return lib().oo_eloc(self.h, np.ascontiguousarray(v).reshape(-1),
                             np.ascontiguousarray(rho).reshape(-1))

    def vloc_g(self, lp):
        """V_loc(G) on the half grid [Nz][Ny][Nx/2+1] (complex)."""
        out = np.zeros((self.nz, self.ny, self.nx // 2 + 1), np.complex128)
        lib().oo_vloc_g(self.h, ctypes.byref(lp), out.view(np.float64).reshape(-1))
        return out

    def vks(self, rho, vloc=None):
        """(V_KS(r), energies dict) from rho(r) and V_loc(G) (or None)."""
        r = np.ascontiguousarray(rho, np.float64).reshape(-1)
        v = np.zeros_like(r)
        e = VksEnergies()
        vl = None if vloc is None else np.ascontiguousarray(vloc, np.complex128)
        lib().oo_vks(self.h, r, _ptr(None if vl is None else vl.view(np.float64)), v,
                     ctypes.byref(e))
        return v.reshape(self.nz, self.ny, self.nx), {k: getattr(e, k) for k, _ in e._fields_}

    def nl_zmat(self, nl, psi, derivs=False):
        """Z [nst][natoms] of the non-local projection (and dZ/dR as
        [nst][3][natoms] if derivs)."""
        p = np.ascontiguousarray(psi, np.complex128)
        z = np.zeros((self.nst, nl.natoms))
        dz = np.zeros((self.nst, 3, nl.natoms)) if derivs else None
        lib().oo_nl_zmat(self.h, ctypes.byref(nl), p.view(np.float64).reshape(-1), z,
                         _ptr(dz))
        return (z, dz) if derivs else z

    def nl_apply(self, nl, psi, f=None, fatm=None):
        """E_NL; the non-local forces are added to f (complex [nst][npts])
        and fatm ([natoms][3]) when given."""
        p = np.ascontiguousarray(psi, np.complex128)
        return lib().oo_nl_apply(self.h, ctypes.byref(nl), p.view(np.float64).reshape(-1),
                                 _ptr(None if f is None else f.view(np.float64)), _ptr(fatm))

    def cg_update(self, cmass, dt, gamma, f, cg, psi):
        """cg = gamma cg + F (|G|^2 <= ecut), psi += dt/m cg, in place."""
        lib().oo_cg_update(self.h, np.ascontiguousarray(cmass, np.float64), float(dt),
                           float(gamma), f.view(np.float64).reshape(-1),
                           cg.view(np.float64).reshape(-1), psi.view(np.float64).reshape(-1))

    def step(self, v, psi, dt=5.0, tau_fs=25.0, ecut_mass=2.0, ns_tol=1e-15,
             ns_max_iter=100, use_lambda=True, want=(), cg=None, vloc=None, nl=None):
        """One minimisation step; psi (complex [nst][npts]) updated in place.
        cg: None for steepest descent, or a Minimizer from new_cg() whose
        direction and state the step updates (conjugate gradient).
        vloc: None for a fixed V(r) = v, or the half-grid V_loc(G) of
        vloc_g(): V_KS is then built from this step's rho (v is ignored and
        may be None; want "v" returns the V_KS used).  nl: None, or the
        NonLocal of nonlocal_() whose energy and forces join the step's."""
        o = StepOpts(dt, tau_fs, ecut_mass, ns_tol, ns_max_iter, int(use_lambda))
        if nl is not None:
            o.nl = ctypes.cast(ctypes.pointer(nl), ctypes.c_void_p)
        vk = None
        if vloc is not None:
            vloc = np.ascontiguousarray(vloc, np.complex128)
            vk = np.zeros((self.nz, self.ny, self.nx))
            o.vks_from_rho = 1
            o.vloc = vloc.ctypes.data_as(ctypes.c_void_p)
            o.vks_out = vk.ctypes.data_as(ctypes.c_void_p)
            v = vk
        if cg is not None:
            o.use_cg = 1
            o.cg_state = ctypes.pointer(cg.state)
            o.cg = cg.direction.ctypes.data_as(ctypes.c_void_p)
        out = StepOut()
        extra = {}
        rho = np.zeros((self.nz, self.ny, self.nx)) if "rho" in want else None
        S = np.zeros((self.nst, self.nst)) if "S" in want else None
        T = np.zeros((self.nst, self.nst)) if "T" in want else None
        f = np.zeros((self.nst, self.npts), np.complex128) if "f" in want else None
        rc = lib().oo_step(self.h, ctypes.byref(o), np.ascontiguousarray(v).reshape(-1),
                           psi.view(np.float64).reshape(-1), _ptr(rho), _ptr(S),
                           _ptr(T), _ptr(None if f is None else f.view(np.float64)),
                           ctypes.byref(out))
        for k, a in (("rho", rho), ("S", S), ("T", T), ("f", f), ("v", vk)):
            if a is not None:
                extra[k] = a
        res = dict(eke=out.eke, eloc=out.eloc, eelec=out.eelec, fnorm=out.fnorm,
                   ns_iters=out.ns_iters, ns_err=out.ns_err, ehart=out.ehart, exc=out.exc,
                   enl=out.enl, rc=rc)
        res.update(extra)
        return res


class CgMinimizer:
    """Persistent conjugate-gradient direction and state of one minimisation."""

    def __init__(self, nst, npts):
        self.state = CgState()
        self.direction = np.zeros((nst, npts), np.complex128)


def cg_gamma(state, fovlap):
    return lib().oo_cg_gamma(ctypes.byref(state), float(fovlap))


def ns(S, tol=1e-15, max_iter=100):
    S = np.ascontiguousarray(S, np.float64)
    n = S.shape[0]
    T = np.zeros_like(S)
    err = ctypes.c_double()
    it = lib().oo_ns(n, S.reshape(-1), tol, max_iter, T.reshape(-1), ctypes.byref(err))
    return T, it, err.value


def set_threads(n):
    lib().oo_set_threads(int(n))


def get_threads():
    return lib().oo_get_threads()
