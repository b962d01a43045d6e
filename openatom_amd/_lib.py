"""ctypes binding of liboa_step.so (the C ABI declared in include/oa_step.h).

The library is the product: there is no fallback.  If it has not been built
the import of this module still succeeds, but every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboa_step.so")

OA_OK, OA_EINVAL, OA_EHIP, OA_ENOMEM, OA_ENOCONV, OA_ESTATE, OA_ECANCELED = range(7)
OA_FP64, OA_MIXED = 0, 1
STATUS_NAMES = {0: "OA_OK", 1: "OA_EINVAL", 2: "OA_EHIP", 3: "OA_ENOMEM",
                4: "OA_ENOCONV", 5: "OA_ESTATE", 6: "OA_ECANCELED"}

# every symbol include/oa_step.h and include/oa_serve.h declare
EXPORTS = (
    "oa_default_options", "oa_kvecs", "oa_create", "oa_destroy", "oa_last_error",
    "oa_layout_info", "oa_ctx_device_bytes", "oa_get_config", "oa_get_kvecs", "oa_psi_set",
    "oa_psi_get", "oa_vks_set",
    "oa_fft_g2r", "oa_fft_g2r_states", "oa_psir_get", "oa_density_get", "oa_vpsi_r2g",
    "oa_vpsi_r2g_states", "oa_force_get",
    "oa_force_set", "oa_lambda", "oa_sd_update", "oa_overlap", "oa_ortho_ns",
    "oa_tmat_set", "oa_ortho_transform", "oa_ortho_transform_noalloc", "oa_rotate", "oa_step", "oa_residual", "oa_sync", "oa_step_result", "oa_timings",
    "oa_set_timing", "oa_local_pseudo_set", "oa_vks_build", "oa_vks_get", "oa_tuning_log",
    "oa_nonlocal_set", "oa_nonlocal_get", "oa_nonlocal_force",
    # state files (oa_state.h)
    "oa_state_read", "oa_state_write", "oa_states_read", "oa_states_write",
    # request loop (oa_serve.h)
    "oa_ctx_bytes", "oa_server_create", "oa_server_submit", "oa_server_poll", "oa_server_drain",
    "oa_server_stats", "oa_server_destroy", "oa_server_cancel",
)


class OaSystem(ctypes.Structure):
    _fields_ = [("hmat", ctypes.c_double * 9), ("grid", ctypes.c_int * 3),
                ("ecut_ry", ctypes.c_double), ("nstates", ctypes.c_int),
                ("occ", ctypes.c_void_p), ("npts", ctypes.c_int),
                ("kx", ctypes.c_void_p), ("ky", ctypes.c_void_p), ("kz", ctypes.c_void_p)]


class OaOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("ns_max_iter", ctypes.c_int),
                ("ns_tol", ctypes.c_double), ("dt", ctypes.c_double),
                ("tau_mass", ctypes.c_double), ("ecut_mass", ctypes.c_double),
                ("use_lambda", ctypes.c_int), ("use_cg", ctypes.c_int),
                ("want_resid", ctypes.c_int),
                ("fft_mode", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("ns_mode", ctypes.c_int), ("vks_from_rho", ctypes.c_int),
                ("autotune", ctypes.c_int), ("gemm_split", ctypes.c_int),
                ("ns_fp32", ctypes.c_int), ("fft_full", ctypes.c_int),
                ("psir_keep", ctypes.c_int)]


class OaStepOut(ctypes.Structure):
    _fields_ = [("eke", ctypes.c_double), ("eloc", ctypes.c_double),
                ("eelec", ctypes.c_double), ("fnorm", ctypes.c_double),
                ("ns_iters", ctypes.c_int), ("ns_err", ctypes.c_double),
                ("resid_max", ctypes.c_double), ("resid_rms", ctypes.c_double),
                ("ms", ctypes.c_double), ("ehart", ctypes.c_double), ("exc", ctypes.c_double),
                ("enl", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OaLocalPseudo(ctypes.Structure):
    _fields_ = [("natoms", ctypes.c_int), ("nspecies", ctypes.c_int), ("nspline", ctypes.c_int),
                ("pos", ctypes.c_void_p), ("species", ctypes.c_void_p),
                ("gmin", ctypes.c_double), ("dg", ctypes.c_double), ("vps", ctypes.c_void_p),
                ("vps_g0", ctypes.c_void_p)]


class OaNonlocal(ctypes.Structure):
    _fields_ = [("natoms", ctypes.c_int), ("nspecies", ctypes.c_int), ("nspline", ctypes.c_int),
                ("pos", ctypes.c_void_p), ("species", ctypes.c_void_p),
                ("gmin", ctypes.c_double), ("dg", ctypes.c_double), ("proj", ctypes.c_void_p),
                ("proj_g0", ctypes.c_void_p), ("vnorm", ctypes.c_void_p)]


class OaError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_lib = None


def load():
    """Load liboa_step.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                           "`make -C openatom_amd/csrc` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    pd = ctypes.POINTER(ctypes.c_double)
    L.oa_default_options.argtypes = [ctypes.POINTER(OaOptions)]
    L.oa_default_options.restype = None
    L.oa_kvecs.argtypes = [cd, pd, vp, vp, vp]
    L.oa_kvecs.restype = ci
    L.oa_create.argtypes = [ci, ctypes.POINTER(OaSystem), ctypes.POINTER(OaOptions),
                            ctypes.POINTER(vp)]
    L.oa_last_error.argtypes = [vp]
    L.oa_last_error.restype = ctypes.c_char_p
    L.oa_tuning_log.argtypes = [vp]
    L.oa_tuning_log.restype = ctypes.c_char_p
    for name, args in (("oa_destroy", [vp]), ("oa_layout_info", [vp, vp]), ("oa_get_config", [vp, vp]),
                       ("oa_get_kvecs", [vp, vp, vp, vp]), ("oa_psi_set", [vp, vp]),
                       ("oa_psi_get", [vp, vp]), ("oa_vks_set", [vp, vp]),
                       ("oa_fft_g2r", [vp]), ("oa_psir_get", [vp, ci, vp]),
                       ("oa_density_get", [vp, vp]), ("oa_vpsi_r2g", [vp, vp]),
                       ("oa_force_get", [vp, vp]), ("oa_force_set", [vp, vp]),
                       ("oa_lambda", [vp, vp]), ("oa_sd_update", [vp, vp]),
                       ("oa_overlap", [vp, vp]),
                       ("oa_ortho_ns", [vp, cd, ci, vp, vp, vp]),
                       ("oa_tmat_set", [vp, vp]), ("oa_rotate", [vp]),
                       ("oa_fft_g2r_states", [vp, ci, ci, ci]), ("oa_vpsi_r2g_states", [vp, ci, ci, vp]),
                       ("oa_ortho_transform", [vp, vp, ci]),
                       ("oa_ortho_transform_noalloc", [vp, vp, ci]),
                       ("oa_step", [vp, vp]), ("oa_residual", [vp, vp, vp]), ("oa_sync", [vp]),
                       ("oa_step_result", [vp, vp]), ("oa_timings", [vp, vp, ci]),
                       ("oa_set_timing", [vp, ci]),
                       ("oa_local_pseudo_set", [vp, ctypes.POINTER(OaLocalPseudo)]),
                       ("oa_vks_build", [vp]), ("oa_vks_get", [vp, vp]),
                       ("oa_state_read", [ctypes.c_char_p, ci, ci, vp, vp, vp, vp, vp]),
                       ("oa_state_write", [ctypes.c_char_p, ci, ci, vp, vp, vp, vp, vp]),
                       ("oa_states_read", [vp, ctypes.c_char_p, ci, ci, ci, ci, ci]),
                       ("oa_states_write", [vp, ctypes.c_char_p, ci, ci, ci, ci, ci, ci])):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ci
    _lib = L
    return L


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def local_pseudo_struct(pos, species, gmin, dg, vps, vps_g0):
    """(OaLocalPseudo, arrays it points into): fixed ions (bohr), species
    indices and per-species spline tables vps[nspecies][4][nspline]."""
    keep = [np.ascontiguousarray(pos, np.float64).reshape(-1, 3),
            np.ascontiguousarray(species, np.int32),
            np.ascontiguousarray(vps, np.float64),
            np.ascontiguousarray(vps_g0, np.float64)]
    lp = OaLocalPseudo(keep[0].shape[0], keep[2].shape[0], keep[2].shape[2],
                       ptr(keep[0]), ptr(keep[1]), float(gmin), float(dg),
                       ptr(keep[2]), ptr(keep[3]))
    return lp, keep


def default_options():
    o = OaOptions()
    load().oa_default_options(ctypes.byref(o))
    return o


def ortho_transform(S):
    """(T, status) with T = (S/2)^{-1/2} of a host S (reference scaling) on
    the current device: oa_ortho_transform, the stateless
    CPORTHOG::CP_orthoTransform."""
    S = np.ascontiguousarray(S, np.float64)
    n = S.shape[0]
    T = np.zeros_like(S)
    st = load().oa_ortho_transform(ptr(T), ptr(S), n)
    if st not in (OA_OK, OA_ENOCONV):
        raise OaError(st, "oa_ortho_transform failed")
    return T, st


ORTHO_NOALLOC_MAX_N = 512


def ortho_transform_noalloc(S):
    """(T, status) as ortho_transform, through oa_ortho_transform_noalloc:
    no device allocation, n <= ORTHO_NOALLOC_MAX_N."""
    S = np.ascontiguousarray(S, np.float64)
    n = S.shape[0]
    T = np.zeros_like(S)
    st = load().oa_ortho_transform_noalloc(ptr(T), ptr(S), n)
    if st not in (OA_OK, OA_ENOCONV):
        raise OaError(st, "oa_ortho_transform_noalloc failed")
    return T, st


def kvecs(ecut_ry, hmat):
    """Stored k-vectors (reference order) computed by the library."""
    L = load()
    h = (ctypes.c_double * 9)(*np.asarray(hmat, dtype=np.float64).reshape(9))
    n = L.oa_kvecs(float(ecut_ry), h, None, None, None)
    kx = np.zeros(n, np.int32); ky = np.zeros(n, np.int32); kz = np.zeros(n, np.int32)
    L.oa_kvecs(float(ecut_ry), h, ptr(kx), ptr(ky), ptr(kz))
    return kx, ky, kz
