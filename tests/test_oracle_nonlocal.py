"""The oracle's non-local KB part (l = 0, the reference's G-space method)
pinned by what the mathematics fixes, independently of its own
formulation: the projections against a full-sphere sum with the analytic
projector (no weights, explicit -G partners), the energy's closed form,
the coefficient forces as the gradient of E_NL in the reference's
convention (kx = 0 halved), the ion forces as -dE/dR, translation
invariance, and the occupation acting on the energy only."""
import numpy as np
import pytest

import oracle
from openatom_amd import system
from tests.helpers import oracle_nonlocal, small_system

Y00 = 1.0 / np.sqrt(4.0 * np.pi)


@pytest.fixture(scope="module")
def case():
    s = small_system(nmol=3, nstates=5, seed=7)
    assert len(s["kb"]["atoms"]) == 3  # three O: the KB atoms
    return s


def osys(s, occ=None):
    return oracle.System(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                         s["kz"], occ=occ)


def gvecs(s):
    hi = np.linalg.inv(np.asarray(s["hmat"]).reshape(3, 3))
    k = np.stack([s["kx"], s["ky"], s["kz"]], 1).astype(np.float64)
    return 2.0 * np.pi * k @ hi  # rows: (kx, ky, kz) . hmati rows


def full_sphere_z(s, psi, pos=None):
    """Z_sI = sum over the FULL sphere of P(|G|) y00 c(G) exp(i G.R_I) with
    the analytic Gaussian projector: stored points as they are (the kx = 0
    plane is stored whole) plus the kx > 0 points mirrored to -G with
    conj(c).  Returns the complex sums (their imaginary parts vanish)."""
    G = gvecs(s)
    hx = s["kx"] > 0
    Gf = np.concatenate([G, -G[hx]])
    cf = np.concatenate([psi, np.conj(psi[:, hx])], axis=1)
    B, sig, _ = s["kbparams"]["O"]
    P, _ = system.gaussian_form_factor(B, sig, np.linalg.norm(Gf, axis=1))
    R = s["kb"]["pos"] if pos is None else pos
    ph = np.exp(1j * Gf @ np.asarray(R).T)  # [points][atoms]
    return (cf * (P * Y00)[None, :]) @ ph


def test_projections_match_full_sphere_sum(case):
    s = case
    z = osys(s).nl_zmat(oracle_nonlocal(s)[0], s["psi"])
    zf = full_sphere_z(s, s["psi"])
    assert np.abs(zf.imag).max() < 1e-12 * np.abs(zf).max()
    assert np.abs(z - zf.real).max() < 1e-10 * np.abs(zf).max()


def test_energy_closed_form_with_occupations(case):
    s = case
    occ = np.array([1.0, 0.5, 2.0, 1.0, 0.25])
    nl, _ = oracle_nonlocal(s)
    zf = full_sphere_z(s, s["psi"]).real
    vn = s["kb"]["vnorm"][s["kb"]["species"]] / abs(np.linalg.det(
        np.asarray(s["hmat"]).reshape(3, 3)))
    e_ref = float(np.sum(occ[:, None] * vn[None, :] * zf ** 2))
    e = osys(s, occ).nl_apply(nl, s["psi"])
    assert abs(e - e_ref) < 1e-10 * abs(e_ref)


def test_coefficient_force_is_the_energy_gradient(case):
    """E_NL is quadratic in the coefficients, so central differences are
    exact up to rounding.  Reference convention: F = -(dE/dRe c + i dE/dIm c)
    for kx > 0; on the kx = 0 plane (c and its stored partner move together)
    half of that; at G = 0 (real) the whole derivative."""
    s = case
    o = osys(s)
    nl, _ = oracle_nonlocal(s)
    f = np.zeros_like(s["psi"])
    o.nl_apply(nl, s["psi"], f)
    kx, ky, kz = s["kx"], s["ky"], s["kz"]
    key = {(a, b, c): i for i, (a, b, c) in enumerate(zip(kx, ky, kz))}
    g0 = key[(0, 0, 0)]
    uniq0 = [i for i in range(len(kx)) if kx[i] == 0 and (ky[i] > 0 or (ky[i] == 0 and kz[i] > 0))]
    pos_kx = [i for i in range(len(kx)) if kx[i] > 0]
    rng = np.random.default_rng(0)
    h = 1e-4
    for st in (0, 3):
        for p in [g0] + list(rng.choice(uniq0, 3)) + list(rng.choice(pos_kx, 3)):
            partner = key[(0, -ky[p], -kz[p])] if kx[p] == 0 else None
            grad = []
            for d in ((1.0, 0.0), (0.0, 1.0)) if p != g0 else ((1.0, 0.0),):
                e = []
                for sgn in (1.0, -1.0):
                    psi = s["psi"].copy()
                    dz = sgn * h * complex(*d)
                    psi[st, p] += dz
                    if partner is not None and partner != p:
                        psi[st, partner] += np.conj(dz)
                    e.append(o.nl_apply(nl, psi))
                grad.append((e[0] - e[1]) / (2 * h))
            g = complex(grad[0], grad[1] if len(grad) > 1 else 0.0)
            scale = 1.0 if (kx[p] > 0 or p == g0) else 0.5
            expect = -scale * g
            assert abs(f[st, p] - expect) < 1e-7 * max(abs(expect), np.abs(f).max() * 1e-3), \
                (st, p, f[st, p], expect)


def test_ion_forces_are_minus_energy_gradient(case):
    s = case
    o = osys(s)
    nl, _ = oracle_nonlocal(s)
    fa = np.zeros((len(s["kb"]["atoms"]), 3))
    o.nl_apply(nl, s["psi"], None, fa)
    h = 1e-5
    for ia in range(fa.shape[0]):
        for q in range(3):
            e = []
            for sgn in (1.0, -1.0):
                pos = s["kb"]["pos"].copy()
                pos[ia, q] += sgn * h
                e.append(o.nl_apply(oracle_nonlocal(s, pos)[0], s["psi"]))
            fd = -(e[0] - e[1]) / (2 * h)
            assert abs(fa[ia, q] - fd) < 1e-6 * np.abs(fa).max(), (ia, q, fa[ia, q], fd)


def test_translation_invariance(case):
    """R -> R + t with c(G) -> c(G) exp(-i G.t) leaves every Z unchanged:
    pins the sign of the structure-factor phase."""
    s = case
    t = np.array([0.37, -1.21, 0.83])
    o = osys(s)
    z0 = o.nl_zmat(oracle_nonlocal(s)[0], s["psi"])
    psi_t = s["psi"] * np.exp(-1j * (gvecs(s) @ t))[None, :]
    z1 = o.nl_zmat(oracle_nonlocal(s, s["kb"]["pos"] + t)[0], psi_t)
    assert np.abs(z1 - z0).max() < 1e-11 * np.abs(z0).max()


def test_occupation_weights_energy_not_forces(case):
    s = case
    nl, _ = oracle_nonlocal(s)
    out = {}
    for occ in (None, np.full(s["nstates"], 2.0)):
        f = np.zeros_like(s["psi"])
        fa = np.zeros((len(s["kb"]["atoms"]), 3))
        e = osys(s, occ).nl_apply(nl, s["psi"], f, fa)
        out[occ is None] = (e, f, fa)
    assert abs(out[False][0] - 2.0 * out[True][0]) < 1e-12 * abs(out[True][0])
    assert np.array_equal(out[False][1], out[True][1])
    assert np.array_equal(out[False][2], out[True][2])


def test_step_adds_nonlocal_energy(case):
    s = case
    o = osys(s)
    nl, _ = oracle_nonlocal(s)
    e_nl = o.nl_apply(nl, s["psi"])
    psi = s["psi"].copy()
    r = o.step(s["vks"], psi, nl=nl)
    assert r["enl"] == e_nl
    assert abs(r["eelec"] - (r["eke"] + r["eloc"] + r["enl"])) < 1e-13 * abs(r["eelec"])
