"""Host-side view of one device context of the CP electronic-step library.

Method names follow the reference's chare-facing calls where one exists
(paths relative to the OpenAtom tree):

  fft_g2r      FFTcache::doStpFFTGtoR_Gchare/_Rchare + density reduction
               (fftCache.C:1133-1243, CP_State_RealSpacePlane.C:349-517)
  vpsi_r2g     unpackVks + doStpFFTRtoG_* + doIFFT + addForces + CP_eke_calc
               (CP_State_RealSpacePlane.C:530-597, fftCache.C:1099-1319,
                CP_State_GSpacePlane.C:1582-1778)
  lambda_      asymmetric PairCalculator forward/backward (ckPairCalculator.C)
  sd_update    CP_integrate_min_STD + redundant-psi sync
  overlap      symmetric PairCalculator::multiplyForward (S ~ 2I)
  ortho_ns     Ortho Newton-Schulz (ortho.C:862-985)
  rotate       PairCalculator::bwMultiplyHelper (Psi T)
  step         GSpaceDriver minimisation iteration (gspace.ci:106-190)

All calls go through liboa_step.so; nothing here computes physics.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import OaError, ptr


class CPContext:
    # engine codes reported by oa_get_config (fft_mode option: 0 auto, 1-4 force
    # Stockham per pass, Stockham fused, two-factor per pass, two-factor fused)
    ENGINES = {0: "two-factor register FFT, one kernel per pass", 1: "Stockham per pass",
               2: "Stockham fused planes", 3: "two-factor fused y+x planes",
               4: "two-factor fused y+x planes, rho from the R->G pass (fixed V)"}

    def __init__(self, grid, hmat, ecut_ry, nstates, kx=None, ky=None, kz=None,
                 occ=None, precision="fp64", device=0, **opts):
        L = _lib.load()
        self._L = L
        self.grid = tuple(int(g) for g in grid)
        self.nx, self.ny, self.nz = self.grid
        self.nst = int(nstates)
        sysd = _lib.OaSystem()
        for i, v in enumerate(np.asarray(hmat, np.float64).reshape(9)):
            sysd.hmat[i] = float(v)
        for i, v in enumerate(self.grid):
            sysd.grid[i] = v
        sysd.ecut_ry = float(ecut_ry)
        sysd.nstates = self.nst
        self._occ = None if occ is None else np.ascontiguousarray(occ, np.float64)
        sysd.occ = ptr(self._occ)
        if kx is not None:
            self._k = [np.ascontiguousarray(a, np.int32) for a in (kx, ky, kz)]
            sysd.npts = int(self._k[0].size)
            sysd.kx, sysd.ky, sysd.kz = (ptr(a) for a in self._k)
        else:
            sysd.npts = 0
        o = _lib.default_options()
        o.precision = _lib.OA_MIXED if precision == "mixed" else _lib.OA_FP64
        for k, v in opts.items():
            if not hasattr(o, k):
                raise TypeError(f"unknown option {k}")
            setattr(o, k, v)
        self.opts = o
        self.precision = precision
        h = ctypes.c_void_p()
        st = L.oa_create(int(device), ctypes.byref(sysd), ctypes.byref(o), ctypes.byref(h))
        if st != _lib.OA_OK:
            raise OaError(st, "oa_create failed (see stderr)")
        self.h = h
        info = np.zeros(4, np.int32)
        self._chk(L.oa_layout_info(self.h, ptr(info)))
        self.npts, self.nlines, self.nplane_x, self.nkx0 = (int(v) for v in info)
        cfg = np.zeros(8, np.int32)
        self._chk(L.oa_get_config(self.h, ptr(cfg)))
        self.config = dict(zip(("engine", "zlines", "ycols", "xpairs", "kcp", "pcp", "cs",
                                "nchunk"), (int(v) for v in cfg)))
        self.config["engine_name"] = self.ENGINES.get(self.config["engine"], "?")

    # ------------------------------------------------------------------ plumbing
    def _chk(self, st):
        if st != _lib.OA_OK:
            msg = self._L.oa_last_error(self.h).decode() if getattr(self, "h", None) else ""
            raise OaError(st, msg)

    def close(self):
        if getattr(self, "h", None):
            self._L.oa_destroy(self.h)
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

    @property
    def ngrid(self):
        return self.nx * self.ny * self.nz

    def kvecs(self):
        kx = np.zeros(self.npts, np.int32); ky = np.zeros_like(kx); kz = np.zeros_like(kx)
        self._chk(self._L.oa_get_kvecs(self.h, ptr(kx), ptr(ky), ptr(kz)))
        return kx, ky, kz

    # ------------------------------------------------------------------ data
    def set_psi(self, psi):
        a = np.ascontiguousarray(psi, np.complex128)
        assert a.shape == (self.nst, self.npts), a.shape
        self._chk(self._L.oa_psi_set(self.h, ptr(a)))

    def get_psi(self):
        a = np.zeros((self.nst, self.npts), np.complex128)
        self._chk(self._L.oa_psi_get(self.h, ptr(a)))
        return a

    def set_vks(self, v):
        a = np.ascontiguousarray(v, np.float64)
        assert a.size == self.ngrid
        self._chk(self._L.oa_vks_set(self.h, ptr(a)))

    def get_force(self):
        a = np.zeros((self.nst, self.npts), np.complex128)
        self._chk(self._L.oa_force_get(self.h, ptr(a)))
        return a

    def set_local_pseudo(self, pos, species, gmin, dg, vps, vps_g0):
        """Fixed ions of a vks_from_rho context: positions (bohr), species
        indices and per-species spline tables vps[nspecies][4][nspline]."""
        lp, keep = _lib.local_pseudo_struct(pos, species, gmin, dg, vps, vps_g0)
        self._chk(self._L.oa_local_pseudo_set(self.h, ctypes.byref(lp)))

    def set_nonlocal(self, pos, species, gmin, dg, proj, proj_g0, vnorm):
        """Non-local KB atoms (l = 0 channel): positions (bohr), species
        indices, per-species projector splines proj[nspecies][4][nspline],
        their G = 0 values and normalisations.  Every later step adds E_NL
        and the non-local force."""
        keep = [np.ascontiguousarray(pos, np.float64).reshape(-1, 3),
                np.ascontiguousarray(species, np.int32),
                np.ascontiguousarray(proj, np.float64),
                np.ascontiguousarray(proj_g0, np.float64),
                np.ascontiguousarray(vnorm, np.float64)]
        nl = _lib.OaNonlocal(keep[0].shape[0], keep[2].shape[0], keep[2].shape[2],
                             ptr(keep[0]), ptr(keep[1]), float(gmin), float(dg),
                             ptr(keep[2]), ptr(keep[3]), ptr(keep[4]))
        self._chk(self._L.oa_nonlocal_set(self.h, ctypes.byref(nl)))
        self._nl_natoms = keep[0].shape[0]

    def nonlocal_get(self, atom_forces=False):
        """(Z [nstates][natoms], E_NL[, ion forces [natoms][3]]) of the
        current coefficients."""
        z = np.zeros((self.nst, self._nl_natoms))
        e = ctypes.c_double(0.0)
        fa = np.zeros((self._nl_natoms, 3)) if atom_forces else None
        self._chk(self._L.oa_nonlocal_get(self.h, ptr(z), ctypes.byref(e), ptr(fa)))
        return (z, e.value, fa) if atom_forces else (z, e.value)

    def nonlocal_force(self):
        """F += the non-local force of the current coefficients."""
        self._chk(self._L.oa_nonlocal_force(self.h))

    def read_states(self, directory, fmt=1, spin=0, kpt=0, bead=0, temper=0):
        """Load every state from the reference's STATES layout
        (directory/Spin.s_Kpt.k_Bead.b_Temper.t/state<i>.out)."""
        self._chk(self._L.oa_states_read(self.h, str(directory).encode(), int(fmt), spin, kpt,
                                         bead, temper))

    def write_states(self, directory, fmt=1, spin=0, kpt=0, bead=0, temper=0, iteration=0):
        """Write every state (and TimeStamp) in the reference's STATES layout;
        the Spin/Kpt/Bead/Temper directory must exist."""
        self._chk(self._L.oa_states_write(self.h, str(directory).encode(), int(fmt), spin, kpt,
                                          bead, temper, int(iteration)))

    def tuning_log(self):
        """What the autotuner measured and chose (empty unless autotune=1)."""
        return self._L.oa_tuning_log(self.h).decode()

    def vks_build(self):
        """V_KS from the context's current rho (after fft_g2r)."""
        self._chk(self._L.oa_vks_build(self.h))

    def get_vks(self):
        a = np.zeros((self.nz, self.ny, self.nx))
        self._chk(self._L.oa_vks_get(self.h, ptr(a)))
        return a

    def set_force(self, f):
        a = np.ascontiguousarray(f, np.complex128)
        self._chk(self._L.oa_force_set(self.h, ptr(a)))

    def get_psir(self, state):
        a = np.zeros((self.nz, self.ny, self.nx))
        self._chk(self._L.oa_psir_get(self.h, int(state), ptr(a)))
        return a

    def get_density(self):
        a = np.zeros((self.nz, self.ny, self.nx))
        self._chk(self._L.oa_density_get(self.h, ptr(a)))
        return a

    # ------------------------------------------------------------------ stages
    def fft_g2r(self):
        self._chk(self._L.oa_fft_g2r(self.h))

    def fft_g2r_states(self, s0, ns, first=True):
        """G->R and the density share of the states [s0, s0 + ns); rho is
        replaced (first) or added to."""
        self._chk(self._L.oa_fft_g2r_states(self.h, int(s0), int(ns), 1 if first else 0))

    def vpsi_r2g_states(self, s0, ns):
        """V psi + R->G for the states [s0, s0 + ns); returns their EKE."""
        e = ctypes.c_double(0.0)
        self._chk(self._L.oa_vpsi_r2g_states(self.h, int(s0), int(ns), ctypes.byref(e)))
        return e.value

    def vpsi_r2g(self):
        e = ctypes.c_double()
        self._chk(self._L.oa_vpsi_r2g(self.h, ctypes.byref(e)))
        return e.value

    def lambda_(self):
        lam = np.zeros((self.nst, self.nst))
        self._chk(self._L.oa_lambda(self.h, ptr(lam)))
        return lam

    def sd_update(self):
        f = ctypes.c_double()
        self._chk(self._L.oa_sd_update(self.h, ctypes.byref(f)))
        return f.value

    def overlap(self):
        S = np.zeros((self.nst, self.nst))
        self._chk(self._L.oa_overlap(self.h, ptr(S)))
        return S

    def ortho_ns(self, tol=1e-15, max_iter=30):
        T = np.zeros((self.nst, self.nst))
        it = ctypes.c_int()
        err = ctypes.c_double()
        st = self._L.oa_ortho_ns(self.h, float(tol), int(max_iter), ctypes.byref(it),
                                 ctypes.byref(err), ptr(T))
        if st not in (_lib.OA_OK, _lib.OA_ENOCONV):
            self._chk(st)
        return T, it.value, err.value, st == _lib.OA_OK

    def set_tmat(self, T):
        a = np.ascontiguousarray(T, np.float64)
        self._chk(self._L.oa_tmat_set(self.h, ptr(a)))

    def rotate(self):
        self._chk(self._L.oa_rotate(self.h))

    def step(self, fetch=True):
        if fetch:
            out = _lib.OaStepOut()
            self._chk(self._L.oa_step(self.h, ctypes.byref(out)))
            return out.as_dict()
        self._chk(self._L.oa_step(self.h, None))
        return None

    def result(self):
        out = _lib.OaStepOut()
        self._chk(self._L.oa_step_result(self.h, ctypes.byref(out)))
        return out.as_dict()

    def residual(self):
        """(max, rms) of |Psi^H Psi / 2 - I| for the current coefficients."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._chk(self._L.oa_residual(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def sync(self):
        self._chk(self._L.oa_sync(self.h))

    def set_timing(self, on=True, level=1):
        """level 1: every stage; level 2: only the step and the R->G plane
        pass (fewer events in the stream, so the step runs at full speed)."""
        self._chk(self._L.oa_set_timing(self.h, int(level) if on else 0))

    TIMING_NAMES = ("k_zfwd", "k_y_fwd", "k_xc2r", "k_xr2c", "k_y_bwd", "k_zbwd",
                    "reduce_lambda", "sd_sync", "overlap", "ns", "rotate", "resid", "step")
    TIMING_NAMES_PLANE = ("k_zfwd", "k_plane_g2r", "k_rho_reduce", "k_plane_r2g", "none",
                          "k_zbwd", "reduce_lambda", "sd_sync", "overlap", "ns", "rotate",
                          "resid", "step")
    TIMING_NAMES_PLANE_RHO = ("k_zfwd", "none", "none", "k_plane_r2g_rho", "k_rho_reduce",
                              "k_zbwd", "reduce_lambda", "sd_sync", "overlap", "ns", "rotate",
                              "resid", "step")
    TIMING_NAMES_FUSED = ("k_zfwd", "k_plane_fwd", "k_rho_reduce", "k_plane_bwd", "none",
                          "k_zbwd", "reduce_lambda", "sd_sync", "overlap", "ns", "rotate",
                          "resid", "step")

    def timings(self):
        """Per-stage device ms averaged over the timed steps (HIP events on the
        context's stream); names follow the FFT mode in use."""
        names = {2: self.TIMING_NAMES_FUSED, 3: self.TIMING_NAMES_PLANE,
                 4: self.TIMING_NAMES_PLANE_RHO}.get(self.config["engine"], self.TIMING_NAMES)
        if self.opts.vks_from_rho:  # stage 2 also builds V_KS from rho
            names = names[:2] + (names[2] + "+vks",) + names[3:]
        if getattr(self, "_nl_natoms", 0):  # the non-local part runs before Lambda
            names = tuple("nonlocal+" + n if n == "reduce_lambda" else n for n in names)
        ms = np.zeros(len(names))
        self._chk(self._L.oa_timings(self.h, ptr(ms), len(ms)))
        out = dict(zip(names, ms.tolist()))
        out.pop("none", None)
        return out
