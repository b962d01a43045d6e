"""The pseudopotential-database reader (openatom_amd/pseudo_db.py), on the
CPU, pinned by closed forms on synthetic files and by an independent
spline library:

- the L = 0 radial quadrature of slow_bess_vps against the analytic
  Fourier transform of a Gaussian;
- the local form factor of a table whose short-range part is a Gaussian and
  whose tail is the charge z1 (the file's convention: the long-range
  -z erf(alpha r)/r is added analytically): v(g) and v(0) in closed form;
- a Kleinman-Bylander s channel with Gaussian v_s - v_loc and r*phi:
  projector form factor and normalisation in closed form;
- spline_fit against scipy's clamped cubic spline with the same end slopes;
- the spline range against the density sphere of a small cell by brute force;
- the water entries of the reference's own database parse with the charges
  and channel counts pi_md.vps names.
"""
import os

import numpy as np
from scipy.interpolate import CubicSpline

from openatom_amd import pseudo_db as P
from openatom_amd import system


def write_pseud(path, nr, rmax, z1, a1, channels, zpol=0.0, gamma=1.0):
    """A database file: header, then each channel's (v, r*phi) columns."""
    with open(path, "w") as f:
        f.write(f"{nr} {rmax!r} {len(channels) - 1}\n{z1!r} {a1!r} 0.0 1.0\n{zpol!r} {gamma!r}\n")
        for v, rphi in channels:
            for a, b in zip(v, rphi):
                f.write(f"{float(a)!r} {float(b)!r}\n")


def gauss_ft(A, s, g):
    """Fourier transform of A exp(-r^2 / (2 s^2))."""
    return A * (2.0 * np.pi * s * s) ** 1.5 * np.exp(-0.5 * s * s * g * g)


def test_radial_quadrature_matches_gaussian_transform():
    nr, rmax = 4000, 16.0
    dr = rmax / nr
    r = np.arange(nr) * dr
    g = np.linspace(0.3, 12.0, 60)
    f = np.exp(-r * r / (2 * 0.7 ** 2))
    got = P.radial_l0(f * r, r, dr, g, 1)
    assert np.abs(got - gauss_ft(1.0, 0.7, g)).max() < 1e-9


def test_local_form_factor_closed_form(tmp_path):
    nr, rmax, A, s, z, al = 5000, 18.0, -3.0, 0.6, 4.0, 1.3
    r = np.arange(nr) * (rmax / nr)
    vshort = A * np.exp(-r * r / (2 * s * s))
    path = os.path.join(tmp_path, "X.pseud")
    write_pseud(path, nr, rmax, z, al, [(vshort, np.zeros(nr))])
    p = P.parse_pseud(path)
    assert (p["nr"], p["n_ang"], p["z1"], p["alpha1"]) == (nr, 0, z, al)
    g = np.linspace(0.4, 15.0, 80)
    f, f0 = P.local_form_factor(p, 0, g)
    want = gauss_ft(A, s, g) - 4.0 * np.pi * z * np.exp(-g * g / (4 * al * al)) / (g * g)
    assert np.abs(f - want).max() < 1e-8
    assert abs(f0 - (gauss_ft(A, s, 0.0) + np.pi * z / (al * al))) < 1e-8


def test_kb_projector_closed_form(tmp_path):
    nr, rmax = 5000, 18.0
    r = np.arange(nr) * (rmax / nr)
    B, sb, t = 2.0, 0.8, 0.9
    vloc = -np.exp(-r * r / 2.0)            # channel 1 (local)
    vs = vloc + B * np.exp(-r * r / (2 * sb * sb))   # channel 0
    rphi = r * np.exp(-r * r / (2 * t * t))
    path = os.path.join(tmp_path, "Y.pseud")
    write_pseud(path, nr, rmax, 6.0, 1.0, [(vs, rphi), (vloc, rphi)])
    p = P.parse_pseud(path)
    g = np.linspace(0.4, 12.0, 50)
    pf, p0, vn = P.kb_projector(p, 0, 1, g)
    # v_rphi = B r exp(-r^2/(2u^2)) with 1/u^2 = 1/sb^2 + 1/t^2
    u = 1.0 / np.sqrt(1.0 / sb ** 2 + 1.0 / t ** 2)
    assert np.abs(pf - gauss_ft(B, u, g)).max() < 1e-8
    assert abs(p0 - gauss_ft(B, u, 0.0)) < 1e-8
    # amat = int r^2 B exp(-r^2 (1/(2 sb^2) + 1/t^2)) dr = B sqrt(pi)/(4 a^1.5)
    a = 1.0 / (2 * sb ** 2) + 1.0 / t ** 2
    assert abs(1.0 / vn - B * np.sqrt(np.pi) / (4.0 * a ** 1.5)) < 1e-9


def test_spline_fit_is_the_clamped_cubic_spline():
    x = np.sort(np.random.default_rng(5).uniform(0.2, 9.0, 300))
    y = np.exp(-x) * np.sin(3 * x) + 0.1 * x
    c1, c2, c3 = P.spline_fit(y, x)
    s0 = (y[1] - y[0]) / (x[1] - x[0])
    sn = (y[-1] - y[-2]) / (x[-1] - x[-2])
    cs = CubicSpline(x, y, bc_type=((1, s0), (1, sn)))
    assert np.abs(c1[:-1] - cs.c[2]).max() < 1e-12
    assert np.abs(c2[:-1] - cs.c[1]).max() < 1e-9
    assert np.abs(c3[:-1] - cs.c[0]).max() < 1e-7
    assert c2[-1] == 0.0 and c3[-1] == 0.0


def test_spline_range_brute_force():
    h = np.diag([7.0, 7.5, 8.0]).reshape(9)
    ecut = 12.0
    gmin, gmax, gtrue = P.spline_range(h, ecut)
    gs = []
    for a in range(-20, 21):
        for b in range(-20, 21):
            for c in range(-20, 21):
                G = 2 * np.pi * np.array([a / 7.0, b / 7.5, c / 8.0])
                g2 = G @ G
                if 0 < g2 <= 4 * ecut:
                    gs.append(np.sqrt(g2))
    assert abs(gtrue - min(gs)) < 1e-12
    assert abs(gmin - 0.75 * min(gs)) < 1e-12 and abs(gmax - 4 / 3 * max(gs)) < 1e-12


def test_water_database_entries():
    o = P.parse_pseud(os.path.join(P.DB_DIR, "O_BLYP_PP30.pseud"))
    h = P.parse_pseud(os.path.join(P.DB_DIR, "H_BLYP_PP30.pseud"))
    assert (o["nr"], o["n_ang"], o["z1"] + o["z2"]) == (1000, 1, 6.0)
    assert (h["nr"], h["n_ang"], h["z1"] + h["z2"]) == (10000, 0, 1.0)
    vps = open(os.path.join(P.DB_DIR, "pi_md.vps")).read()
    for name, e in P.WATER_SPECIES.items():
        assert f"\\ATOM1{{{name}}} \\VPS_TYP{{{e['typ']}}} \\N_ANG{{{e['n_ang']}}} " \
               f"\\LOC_OPT{{{e['loc_opt']}}}" in vps
    w = system.WORKLOADS["water8"]
    hm = system.cubic_hmat(w["box_ang"] * system.BOHR_PER_ANG)
    loc, kb = P.water_tables(hm, w["ecut_ry"])
    assert loc["vps"].shape == (2, 4, 4000) and kb["proj"].shape == (1, 4, 4000)
    assert kb["order"] == ("O",) and np.allclose(loc["q_pseud"], [6.0, 1.0])
    # the potential is -z/r far from the ion: g^2 v(g) -> -4 pi z as g -> 0
    for name, z in (("O", 6.0), ("H", 1.0)):
        e = P.WATER_SPECIES[name]
        p = P.parse_pseud(os.path.join(P.DB_DIR, e["file"]))
        g = np.array([0.02, 0.04])
        f, _ = P.local_form_factor(p, e["loc_opt"], g)
        assert np.all(np.abs(g * g * f + 4.0 * np.pi * z) < 0.01 * 4.0 * np.pi * z)
