"""Pins of the CPU oracle (oracle/oa_oracle.c) against numpy, brute force,
closed forms and invariants.  CPU only.  Each test names the oracle function
it pins and a mistake it would catch."""
import itertools

import numpy as np
import pytest

import oracle
from openatom_amd import system
from tests.helpers import (dense_grid, oracle_pseudo, periodic_gaussians, psi_r_numpy, rel,
                           small_system)


@pytest.fixture(scope="module")
def small():
    s = small_system()
    o = oracle.System(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"])
    return s, o


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [2, 3, 4, 5, 7, 12, 45, 48, 54, 100, 128, 160])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_vs_numpy_and_bruteforce(n, sign):
    """oo_fft1d: sign convention (FFTW_FORWARD = e^{-i}) and no normalisation."""
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    assert rel(oracle.fft1d(x, sign), ref) < 1e-13
    assert rel(oracle.dft1d(x, sign), ref) < 1e-13


# ---------------------------------------------------------------- k-vectors

@pytest.mark.parametrize("name,count", [("water8", 2395), ("water32", 32961), ("water128", 130496)])
def test_stored_counts(name, count):
    """oo_stored_kvecs: stored points incl. the redundant kx=0 half."""
    w = system.WORKLOADS[name]
    h = system.cubic_hmat(w["box_ang"] * system.BOHR_PER_ANG)
    kx, ky, kz = oracle.stored_kvecs(w["ecut_ry"], h)
    assert kx.size == count


def test_kvec_set_bruteforce():
    """oo_kvec_half + oo_flip: the stored set is exactly {|G|^2 <= Ecut, kx >= 0}
    (a cubic and an orthorhombic cell), G=0 once, the kx=0 plane closed under
    G -> -G, and the layout rules of stateSlab.C:285-380."""
    for box, ecut in (((6.21, 6.21, 6.21), 30.0), ((5.0, 5.4, 5.8), 14.0)):
        hmat = np.diag([b * system.BOHR_PER_ANG for b in box]).reshape(9)
        kx, ky, kz = oracle.stored_kvecs(ecut, hmat)
        got = set(zip(kx.tolist(), ky.tolist(), kz.tolist()))
        assert len(got) == kx.size
        hi = np.linalg.inv(hmat.reshape(3, 3))
        kmax = [int(np.sqrt(ecut) / (2 * np.pi * np.linalg.norm(hi[:, d]))) + 2 for d in range(3)]
        want = set()
        for a in range(0, kmax[0] + 1):
            for b in range(-kmax[1], kmax[1] + 1):
                for c in range(-kmax[2], kmax[2] + 1):
                    G = 2 * np.pi * hi @ np.array([a, b, c], float)
                    if G @ G <= ecut:
                        want.add((a, b, c))
        assert got == want
        n0 = int((kx == 0).sum())
        assert (kx[:n0] == 0).all() and (kx[n0:] > 0).all()
        red = (kx == 0) & ((ky < 0) | ((ky == 0) & (kz < 0)))
        nred = int(red.sum())
        assert red[:nred].all() and not red[nred:].any()
        assert (kx[nred], ky[nred], kz[nred]) == (0, 0, 0)
        # lines contiguous, kz ascending by one
        for i in range(1, kx.size):
            if kx[i] == kx[i - 1] and ky[i] == ky[i - 1]:
                assert kz[i] == kz[i - 1] + 1


# ---------------------------------------------------------------- G <-> R

def test_g2r_vs_numpy(small):
    """oo_g2r: psi(r) = sum c e^{-iG.r} on an unequal grid and orthorhombic cell
    (catches a swapped axis, a wrong sign, a missing conjugate-mirror term)."""
    s, o = small
    for st in range(3):
        c = s["psi"][st]
        ref = psi_r_numpy(c, s["kx"], s["ky"], s["kz"], s["grid"])
        assert rel(o.g2r(c), ref) < 1e-13


def test_g2r_parseval_and_real(small):
    s, o = small
    c = s["psi"][0]
    pr = o.g2r(c)
    full = np.fft.fftn(dense_grid(c, s["kx"], s["ky"], s["kz"], s["grid"]))
    assert np.abs(full.imag).max() < 1e-12 * np.abs(full.real).max()
    N = np.prod(s["grid"])
    assert abs((pr ** 2).sum() / N - 2.0) < 1e-12  # norm-2 states


def test_g2r_nonhermitian_kx0_takes_real_part(small):
    """FFTW c2r ignores Im of the kx=0 column: a non-Hermitian kx=0 plane gives
    psi = Re(sum_kx0 c e^{-iGr}) + 2 Re(sum_kx>0 c e^{-iGr})."""
    s, o = small
    rng = np.random.default_rng(9)
    c = rng.normal(size=len(s["kx"])) + 1j * rng.normal(size=len(s["kx"]))
    nx, ny, nz = s["grid"]
    kx, ky, kz = s["kx"], s["ky"], s["kz"]
    x = np.arange(nx); y = np.arange(ny); z = np.arange(nz)
    ph = lambda a, b, cc: np.exp(-2j * np.pi * (cc * z[:, None, None] / nz + b * y[None, :, None] / ny
                                                + a * x[None, None, :] / nx))
    ref = np.zeros((nz, ny, nx))
    for v, a, b, cc in zip(c, kx, ky, kz):
        term = v * ph(a, b, cc)
        ref += term.real if a == 0 else 2.0 * term.real
    assert rel(o.g2r(c), ref) < 1e-12


def test_density_sum_rule(small):
    """oo_density: rho = (1/Omega) sum occ psi^2; (Omega/N) sum rho = 2 nstates."""
    s, o = small
    pr = o.g2r_all(s["psi"])
    rho = o.density(pr)
    assert rel(rho, (pr ** 2).sum(axis=0) / o.volume) < 1e-14
    N = np.prod(s["grid"])
    assert abs(rho.sum() * o.volume / N - 2 * s["nstates"]) < 1e-10
    assert rho.min() >= 0.0


def test_vpsi_r2g_vs_numpy(small):
    """oo_vpsi_r2g: F(G) = -(2/N) sum_r V psi e^{+iG.r} (doIFFT, before addForces)."""
    s, o = small
    pr = o.g2r_all(s["psi"])
    N = np.prod(s["grid"])
    for st in range(2):
        f = o.vpsi_r2g(st, pr[st], s["vks"])
        full = np.fft.ifftn(s["vks"] * pr[st]) * N
        nx, ny, nz = s["grid"]
        ref = full[s["kz"] % nz, s["ky"] % ny, s["kx"] % nx] * (-2.0 / N)
        assert rel(f, ref) < 1e-13


def test_eke_force_closed_form(small):
    """oo_eke_force: addForces doubles kx>0, F -= w g^2 psi, EKE = 1/2 sum_full g^2|c|^2."""
    s, o = small
    rng = np.random.default_rng(2)
    f0 = rng.normal(size=s["psi"].shape) + 1j * rng.normal(size=s["psi"].shape)
    f = f0.copy()
    eke = o.eke_force(s["psi"], f)
    g2 = system.gvec_sq(s["kx"], s["ky"], s["kz"], s["hmat"])
    w = system.gamma_weights(s["kx"])
    ref_f = f0 * w[None, :] - s["psi"] * (w * g2)[None, :]
    assert rel(f, ref_f) < 1e-14
    ref_e = 0.5 * (w[None, :] * g2[None, :] * np.abs(s["psi"]) ** 2).sum()
    assert abs(eke - ref_e) < 1e-12 * abs(ref_e)


def test_kinetic_force_is_minus_the_energy_gradient(small):
    """oo_eke_force: EKE is quadratic in the coefficients, so central
    differences give its gradient up to rounding.  Reference convention (the
    same as the non-local force's): F = -(dE/dRe c + i dE/dIm c) for kx > 0;
    on the kx = 0 plane, where c and its stored partner -G move together,
    half of that; at G = 0 (real) the whole derivative.  A dropped kx > 0
    doubling, a w on the wrong plane or a missing 1/2 fails it."""
    s, o = small
    kx, ky, kz = s["kx"], s["ky"], s["kz"]
    f = np.zeros_like(s["psi"])
    o.eke_force(s["psi"], f)
    key = {(a, b, c): i for i, (a, b, c) in enumerate(zip(kx, ky, kz))}
    g0 = key[(0, 0, 0)]
    uniq0 = [i for i in range(len(kx)) if kx[i] == 0 and (ky[i] > 0 or (ky[i] == 0 and kz[i] > 0))]
    pos_kx = [i for i in range(len(kx)) if kx[i] > 0]
    rng = np.random.default_rng(5)
    h = 1e-3

    def eke(psi):
        return o.eke_force(psi, np.zeros_like(psi))
    for st in (0, s["nstates"] - 1):
        for p in [g0] + list(rng.choice(uniq0, 4)) + list(rng.choice(pos_kx, 4)):
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
                    e.append(eke(psi))
                grad.append((e[0] - e[1]) / (2 * h))
            g = complex(grad[0], grad[1] if len(grad) > 1 else 0.0)
            scale = 1.0 if (kx[p] > 0 or p == g0) else 0.5
            expect = -scale * g
            assert abs(f[st, p] - expect) < 1e-8 * max(abs(expect), 1.0), (st, p, f[st, p], expect)


def test_fictitious_mass_preconditions_the_kinetic_force():
    """oo_cmass (CP_create_mass): the two branches meet at g^2 = ecut_mass, and
    above it an SD step with the kinetic force alone scales every coefficient
    by the same factor 1 - dt ecut_mass / cmass0 (the mass cancels w g^2);
    cmass0 = 2 tau^2 CP_EMAGIC with tau in atomic units of time, checked
    against scipy's value of that unit."""
    from scipy import constants
    s = small_system(ecut=40.0)
    o = oracle.System(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"], s["kz"])
    tau_fs, ecut_mass, dt = 25.0, 2.0, 5.0
    cm = o.cmass(tau_fs, ecut_mass)
    g2 = system.gvec_sq(s["kx"], s["ky"], s["kz"], s["hmat"])
    w = system.gamma_weights(s["kx"])
    au_fs = constants.hbar / constants.physical_constants["Hartree energy"][0] * 1e15
    cmass0 = 2.0 * (tau_fs / au_fs) ** 2 * 0.0005
    low = g2 <= ecut_mass
    assert low.any() and (~low).any()
    assert np.allclose(cm[low], w[low] * cmass0, rtol=1e-5, atol=0.0)
    # continuity: the branch above, evaluated at the cutoff, is the branch below
    hi_at_cut = w * ecut_mass * (cm[~low] / (w[~low] * g2[~low])).mean()
    assert np.allclose(hi_at_cut, w * cmass0, rtol=1e-5)
    # the preconditioned kinetic step is uniform above ecut_mass
    psi = s["psi"].copy()
    f = np.zeros_like(psi)
    o.eke_force(psi, f)
    o.sd_update(cm, dt, f, psi)
    ratio = psi[:, ~low] / s["psi"][:, ~low]
    factor = 1.0 - dt * ecut_mass / cmass0
    assert np.allclose(ratio, factor, rtol=1e-5, atol=0.0)


def test_overlap_and_lambda(small):
    """oo_overlap = full-sphere overlap (2I for norm-2 orthonormal input);
    oo_lambda = Re sum_stored conj(psi_i) F_j; oo_lambda_apply weights 1/2, 1."""
    s, o = small
    S = o.overlap(s["psi"])
    assert np.abs(S - 2 * np.eye(s["nstates"])).max() < 1e-12
    rng = np.random.default_rng(4)
    psi = s["psi"] + 0.01 * (rng.normal(size=s["psi"].shape) + 1j * rng.normal(size=s["psi"].shape))
    assert rel(o.overlap(psi), system.gamma_overlap(psi, s["kx"])) < 1e-13
    F = rng.normal(size=psi.shape) + 1j * rng.normal(size=psi.shape)
    lam = o.lambda_(psi, F)
    ref = (np.conj(psi) @ F.T).real
    assert rel(lam, ref) < 1e-13
    F2 = F.copy()
    o.lambda_apply(psi, lam, F2)
    w = np.where(s["kx"] == 0, 0.5, 1.0)
    assert rel(F2, F - w[None, :] * (lam.T @ psi)) < 1e-13


def test_lambda_projection_is_tangent(small):
    """After F -= w Psi Lambda the true gradient is orthogonal to every state
    (full-sphere inner product with the kx>0 force halved)."""
    s, o = small
    rng = np.random.default_rng(5)
    F = rng.normal(size=s["psi"].shape) + 1j * rng.normal(size=s["psi"].shape)
    red, par = system.redundant_pairs(s["kx"], s["ky"], s["kz"])
    F[:, red] = np.conj(F[:, par])
    lam = o.lambda_(s["psi"], F)
    o.lambda_apply(s["psi"], lam, F)
    w = np.where(s["kx"] == 0, 1.0, 0.5)  # per-coefficient gradient
    Ft = F * w[None, :]
    ov = (system.gamma_weights(s["kx"])[None, :] * np.conj(s["psi"]) @ Ft.T).real
    assert np.abs(ov).max() < 1e-12 * np.abs(F).max()


def test_ns_vs_eigh():
    """oo_ns converges to (S/2)^{-1/2} (get_diag_Tmat's closed form,
    cp_lowdin.C:136-153) and reports rms changes; T (S/2) T = I."""
    rng = np.random.default_rng(1)
    n = 24
    A = rng.normal(size=(n, n)) * 0.05
    S = 2.0 * (np.eye(n) + 0.5 * (A + A.T))
    T, it, err = oracle.ns(S, tol=1e-14, max_iter=50)
    lam, U = np.linalg.eigh(S / 2)
    Tref = (U / np.sqrt(lam)) @ U.T
    assert it > 0 and err <= 1e-14
    assert rel(T, Tref) < 1e-13
    assert np.abs(T @ (S / 2) @ T - np.eye(n)).max() < 1e-13


def test_ns_nonconvergence_reported():
    """max_iter reached with err > tol -> negative iteration count (the
    reference aborts, ortho.C:129-133)."""
    rng = np.random.default_rng(2)
    n = 16
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    S = 2.0 * (Q * np.linspace(0.3, 2.5, n)) @ Q.T
    T, it, err = oracle.ns(S, tol=1e-15, max_iter=3)
    assert it == -3 and err > 1e-15


def test_rotate_cmass_sd_redsync(small):
    s, o = small
    rng = np.random.default_rng(6)
    n = s["nstates"]
    T = np.eye(n) + 0.01 * rng.normal(size=(n, n))
    assert rel(o.rotate(s["psi"], T), T.T @ s["psi"]) < 1e-14
    cm = o.cmass(25.0, 2.0)
    g2 = system.gvec_sq(s["kx"], s["ky"], s["kz"], s["hmat"])
    w = system.gamma_weights(s["kx"])
    tau = 25.0 / 0.0241888
    m0 = 2 * tau * tau * 0.0005
    ref = np.where(g2 > 2.0, w * g2 * m0 / 2.0, w * m0)
    assert rel(cm, ref) < 1e-14
    psi = s["psi"].copy()
    F = rng.normal(size=psi.shape) + 1j * rng.normal(size=psi.shape)
    o.sd_update(cm, 5.0, F, psi)
    assert rel(psi, s["psi"] + F * (5.0 / cm)[None, :]) < 1e-14
    o.red_sync(psi)
    red, par = system.redundant_pairs(s["kx"], s["ky"], s["kz"])
    assert np.array_equal(psi[:, red], np.conj(psi[:, par]))


def test_step_orthonormal_and_descends(small):
    """oo_step: output is orthonormal (S = 2I) and, at a small dt, E_elec
    decreases along successive steps with fixed V."""
    s, o = small
    psi = s["psi"].copy()
    es = []
    for _ in range(4):
        r = o.step(s["vks"], psi, dt=2.0, ns_tol=1e-14, ns_max_iter=50, use_lambda=True)
        assert r["rc"] == 0
        es.append(r["eelec"])
    assert all(b < a for a, b in zip(es, es[1:]))
    S = o.overlap(psi)
    assert np.abs(S - 2 * np.eye(s["nstates"])).max() < 1e-12


def test_cg_gamma_reset_rule():
    """oo_cg_gamma restates CP_State_GSpacePlane.C:2651-2671: gamma = 0 on the
    first iteration and on a reset (|F| > 2|F|_old, or |F| > 1.05|F|_old once
    five steps ran without a reset), else fovlap/fovlap_old."""
    st = oracle.CgState()
    assert oracle.cg_gamma(st, 4.0) == 0.0 and st.iter == 1 and st.num_reset == 1
    assert oracle.cg_gamma(st, 1.0) == pytest.approx(0.25)       # decreasing |F|
    assert oracle.cg_gamma(st, 1.0) == pytest.approx(1.0)        # flat: no reset
    assert oracle.cg_gamma(st, 1.1) == pytest.approx(1.1)        # +5%, only 3 steps since reset
    assert oracle.cg_gamma(st, 5.0) == 0.0                       # |F| more than doubles
    assert st.num_reset == 1
    for _ in range(4):
        oracle.cg_gamma(st, 5.0)
    assert st.num_reset == 5
    assert oracle.cg_gamma(st, 5.0 * 1.2) == 0.0                 # > 1.05 |F|_old after 5 steps
    st2 = oracle.CgState()
    oracle.cg_gamma(st2, 1.0)
    for _ in range(5):
        oracle.cg_gamma(st2, 1.0)
    assert oracle.cg_gamma(st2, 1.0 * 1.05 ** 2 * 0.99) == pytest.approx(1.05 ** 2 * 0.99)


def test_cg_update_closed_form_and_sd_limit(small):
    """cg = gamma cg + F (|G|^2 <= ecut), psi += dt/m cg (cp_min_CG.C:55-105);
    with gamma = 0 and a zero direction it is the SD update."""
    s, o = small
    rng = np.random.default_rng(12)
    cm = o.cmass(25.0, 2.0)
    F = rng.normal(size=s["psi"].shape) + 1j * rng.normal(size=s["psi"].shape)
    cg0 = rng.normal(size=s["psi"].shape) + 1j * rng.normal(size=s["psi"].shape)
    psi, cg = s["psi"].copy(), cg0.copy()
    o.cg_update(cm, 7.0, 0.3, F, cg, psi)
    cg_ref = 0.3 * cg0 + F   # every stored point has |G|^2 <= Ecut
    assert rel(cg, cg_ref) < 1e-15
    assert rel(psi, s["psi"] + cg_ref * (7.0 / cm)[None, :]) < 1e-14
    psi_cg, psi_sd = s["psi"].copy(), s["psi"].copy()
    o.cg_update(cm, 7.0, 0.0, F, np.zeros_like(F), psi_cg)
    o.sd_update(cm, 7.0, F, psi_sd)
    assert np.array_equal(psi_cg, psi_sd)


def test_cg_minimisation_beats_sd(small):
    """At the same dt, conjugate gradient lowers E_elec faster than steepest
    descent and the force norm falls by orders of magnitude."""
    s, o = small
    psi_sd, psi_cg = s["psi"].copy(), s["psi"].copy()
    cg = oracle.CgMinimizer(s["nstates"], len(s["kx"]))
    f0 = None
    for _ in range(50):
        rs = o.step(s["vks"], psi_sd, dt=20.0, ns_tol=1e-14, ns_max_iter=50, use_lambda=True)
        rc = o.step(s["vks"], psi_cg, dt=20.0, ns_tol=1e-14, ns_max_iter=50, use_lambda=True,
                    cg=cg)
        f0 = f0 or rc["fnorm"]
    assert rc["eelec"] < rs["eelec"]
    assert rc["fnorm"] < 0.1 * f0 < rs["fnorm"]
    S = o.overlap(psi_cg)
    assert np.abs(S - 2 * np.eye(s["nstates"])).max() < 1e-12


# ---------------------------------------------------------------- V_KS from rho

def test_pz_lda_potential_is_energy_derivative():
    """V_xc = d(rho eps_xc)/d rho: the potential branch (mufact) agrees with
    the energy branch (cfact, xfact) of CP_exc_calc by central differences."""
    rho = np.geomspace(1e-4, 10.0, 25)
    vxc, _ = oracle.pz_lda(rho)

    def e_density(r):
        return 0.75 * oracle.pz_lda(np.array([r]))[1][0] + oracle.pz_lda(np.array([r]))[1][1]

    for r, v in zip(rho, vxc):
        h = r * 1e-5
        fd = (e_density(r + h) - e_density(r - h)) / (2 * h)
        assert abs(v - fd) < 1e-9 * max(1.0, abs(v))


def test_pz_lda_dirac_exchange_and_rs1_continuity():
    """Exchange of the uniform gas: eps_x rs = -(3/4)(9/(4 pi^2))^(1/3)
    = -0.458165293...; the Perdew-Zunger correlation branches meet at rs = 1
    (eps_c and V_c continuous to the fit's 1e-4)."""
    c = -0.75 * (9.0 / (4.0 * np.pi ** 2)) ** (1.0 / 3.0)
    assert abs(c + 0.4581652932831429) < 1e-15
    for rs in (0.3, 1.0, 2.0, 7.5):
        r = 3.0 / (4.0 * np.pi * rs ** 3)
        _, s = oracle.pz_lda(np.array([r]))
        assert abs(0.75 * s[0] / r * rs - c) < 1e-13
    r1 = 3.0 / (4.0 * np.pi)
    lo = [oracle.pz_lda(np.array([r1 * (1 + d)])) for d in (-1e-12, 1e-12)]
    assert abs(lo[0][1][1] - lo[1][1][1]) / r1 < 1e-4
    assert abs(lo[0][0][0] - lo[1][0][0]) < 1e-4


def test_vps_spline_interpolates_form_factor():
    """The spline row and offset convention of CP_get_vpsnow reproduces the
    analytic Gaussian form factor between the knots."""
    params = {"O": (-6.0, 1.0), "H": (-1.2, 0.6)}
    t = system.local_pseudo_tables(params, 30.0)
    lp, keep = oracle.local_pseudo(np.zeros((1, 3)), [0], t["gmin"], t["dg"], t["vps"],
                                   t["vps_g0"])
    rng = np.random.default_rng(3)
    for s, name in enumerate(t["order"]):
        for g in rng.uniform(0.0, np.sqrt(120.0), 50):
            f, _ = system.gaussian_form_factor(*params[name], g)
            assert abs(oracle.vps_eval(lp, s, g) - f) < 1e-10 * abs(t["vps_g0"][s])


def _hartree_case():
    s = small_system(ecut=20.0, grid=(24, 27, 30))
    o = oracle.System(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                      s["kz"])
    return s, o


@pytest.mark.parametrize("k", [(0, 2, 1), (1, -1, 2), (2, 0, 0), (3, 1, -1)])
def test_hartree_cosine_density_closed_form(k):
    """rho = rho0 + a cos(G0.r): V_H = 4 pi a cos(G0.r)/G0^2 and
    E_H = pi Omega a^2 / G0^2 (kx = 0 and kx > 0 wave vectors, orthorhombic
    cell, unequal odd/even grid)."""
    s, o = _hartree_case()
    nx, ny, nz = s["grid"]
    L = np.diag(np.asarray(s["hmat"]).reshape(3, 3))
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx,
                          indexing="ij")
    a, rho0 = 0.004, 0.02
    ph = 2 * np.pi * (k[0] * x + k[1] * y + k[2] * z)
    G2 = sum((2 * np.pi * k[d] / L[d]) ** 2 for d in range(3))
    rho = rho0 + a * np.cos(ph)
    v, e = o.vks(rho, None)
    vxc, sums = oracle.pz_lda(rho)
    assert rel(e["ehart"], np.pi * np.prod(L) * a * a / G2) < 1e-12
    assert np.abs(v - vxc - 4 * np.pi * a * np.cos(ph) / G2).max() < 1e-12
    dv = np.prod(L) / (nx * ny * nz)
    assert rel(e["exc"], (0.75 * sums[0] + sums[1]) * dv) < 1e-14
    assert e["eext"] == 0.0


def test_local_pseudo_is_periodic_gaussian_sum():
    """At uniform rho (no Hartree term, uniform V_xc) V_KS - V_xc is the
    periodic sum of the atoms' Gaussians (atoms at asymmetric positions, so
    a conjugation or sign error of the structure factor moves every
    Gaussian to -R), and E_ext = rho0 sum_I v_I(0)."""
    s = small_system(ecut=60.0, grid=(36, 40, 40), box=(5.0, 5.4, 5.8))
    o = oracle.System(s["grid"], s["hmat"], s["ecut_ry"], s["nstates"], s["kx"], s["ky"],
                      s["kz"])
    lp, keep = oracle_pseudo(s)
    vl = o.vloc_g(lp)
    nx, ny, nz = s["grid"]
    rho0 = 0.02
    rho = np.full((nz, ny, nx), rho0)
    v, e = o.vks(rho, vl)
    vxc, _ = oracle.pz_lda(rho)
    Vp = periodic_gaussians(s["pos"], s["spec"], s["vparams"], s["hmat"], s["grid"])
    assert np.abs(v - vxc - Vp).max() < 1e-9 * np.abs(Vp).max()
    assert rel(e["eext"], rho0 * s["pseudo"]["vps_g0"][s["species"]].sum()) < 1e-13
    assert e["ehart"] < 1e-25


def test_scf_steepest_descent_lowers_energy(small):
    """With V_KS rebuilt from each step's rho, E = EKE + E_H + E_xc + E_ext
    decreases along small steepest-descent steps, and the step reports the
    V_KS of its own rho."""
    s, o = small
    lp, keep = oracle_pseudo(s)
    vl = o.vloc_g(lp)
    psi = s["psi"].copy()
    psir = o.g2r_all(psi)
    rho = o.density(psir)
    v_expect, e_expect = o.vks(rho, vl)
    es = []
    for k in range(6):
        r = o.step(None, psi, dt=5.0, vloc=vl, want=("v", "rho"))
        if k == 0:
            assert np.array_equal(r["rho"], rho)
            assert np.array_equal(r["v"], v_expect)
            assert r["ehart"] == e_expect["ehart"] and r["exc"] == e_expect["exc"]
            assert r["eloc"] == e_expect["eext"]
        es.append(r["eelec"])
        assert abs(r["eelec"] - (r["eke"] + r["eloc"] + r["ehart"] + r["exc"])) < 1e-12
    assert all(b < a for a, b in zip(es, es[1:]))
