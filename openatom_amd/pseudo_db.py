"""Pseudopotential database files into the spline tables of the step
library (host-side set-up, done once per cell).

A database file (``*.pseud``) holds, whitespace-separated: ``nr rmax n_ang``,
``z1 alpha1 z2 alpha2``, ``zpol gamma`` and then ``n_ang + 1`` angular
channels of ``nr`` pairs ``v_l(r) r*phi_l(r)`` on r_i = (i-1) rmax/nr
(the reads of make_vps_splin, control_vps_params.C:769-784, 920-941).

- Local part (channel ``loc_opt``): the form factor of the short-range table
  plus the analytic long-range Coulomb tail of the charges z1, z2
  (slow_bess_vps, control_vps_params.C:1266-1525, L = 0 with ilong = 1):
  v(g) = sum_{i>=2} 4 pi dr j0(g r_i) r_i^2 v_loc(r_i)
         - sum_c 4 pi z_c exp(-g^2 / (4 alpha_c^2)) / g^2 (+ polarisation),
  v(0) = sum_{i>=2} 4 pi dr r_i^2 v_loc(r_i) + sum_c pi z_c / alpha_c^2.
- Kleinman-Bylander channels l != loc_opt (label KB, l = 0 only here): the
  projector v_rphi = (v_l - v_loc) r*phi_l, its normalisation
  vnorm = 1 / sum_i r*phi_l v_rphi dr and its form factor
  P(g) = sum_{i>=2} 4 pi dr j0(g r_i) r_i v_rphi(r_i), P(0) likewise with
  j0 = 1 (control_vps_params.C:1010-1068).
- Splines: PINY's grid g_i = gmin + (i-1) dg, i = 1..nsplin, over
  [0.75 |G|_min, 4/3 |G|_max] of the density sphere |G|^2 <= 4 ecut_ry
  (setkvec3d_sm, set_cp_ewald.C:474-505), dg = (gmax - gmin)/nsplin, fitted
  by spline_fit (friend_lib.C:134-200): a cubic spline whose end slopes are
  the first and last divided differences and whose interior slopes solve
  the standard tridiagonal continuity system.  The tables are the
  vps0..vps3 rows that CP_get_vpsnow (cp_hart_ext.C:433-475) and
  get_rad_proj (cp_struct_fact.C:268-300) evaluate.
"""
import os

import numpy as np

DB_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests",
                      "golden", "pseudo")

# species entries of the water inputs (pi_md.vps): label, n_ang, loc_opt, file
WATER_SPECIES = {
    "O": dict(typ="KB", n_ang=1, loc_opt=1, file="O_BLYP_PP30.pseud"),
    "H": dict(typ="LOC", n_ang=0, loc_opt=0, file="H_BLYP_PP30.pseud"),
}


def parse_pseud(path):
    """The tables of one database file (reads them as make_vps_splin does:
    whitespace-separated, channel after channel)."""
    with open(path) as f:
        tok = f.read().split()
    nr, rmax, n_ang = int(tok[0]), float(tok[1]), int(tok[2])
    z1, a1, z2, a2, zpol, gamma = (float(t) for t in tok[3:9])
    data = np.array(tok[9:9 + 2 * nr * (n_ang + 1)], np.float64)
    if data.size != 2 * nr * (n_ang + 1):
        raise ValueError(f"{path}: {data.size // 2} table lines, expected "
                         f"{nr * (n_ang + 1)}")
    data = data.reshape(n_ang + 1, nr, 2)
    dr = rmax / nr
    return dict(nr=nr, rmax=rmax, n_ang=n_ang, z1=z1, alpha1=a1, z2=z2, alpha2=a2, zpol=zpol,
                gamma=gamma, dr=dr, r=np.arange(nr) * dr, v=data[:, :, 0], rphi=data[:, :, 1])


def spline_range(hmat, ecut_ry):
    """(gmin_spl, gmax_spl, gmin_true) of the density sphere
    |G|^2 <= 4 ecut_ry, G != 0 (setkvec3d_sm, set_cp_ewald.C:474-505)."""
    h = np.asarray(hmat, np.float64).reshape(3, 3)
    hi = np.linalg.inv(h)
    gcut2 = 4.0 * ecut_ry
    # the largest index per axis that can lie inside the sphere
    kmax = [int(np.ceil(np.sqrt(gcut2) / (2.0 * np.pi) * np.linalg.norm(h[:, d]))) + 1
            for d in range(3)]
    a = np.arange(-kmax[0], kmax[0] + 1)
    b = np.arange(-kmax[1], kmax[1] + 1)
    c = np.arange(-kmax[2], kmax[2] + 1)
    K = np.stack(np.meshgrid(a, b, c, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    G = 2.0 * np.pi * K @ hi
    g2 = (G * G).sum(1)
    g = np.sqrt(g2[(g2 <= gcut2) & (g2 > 0.0)])
    gmin, gmax = g.min(), g.max()
    return 0.75 * gmin, 4.0 / 3.0 * gmax, gmin


def spline_grid(hmat, ecut_ry, nsplin=4000):
    gmin, gmax, _ = spline_range(hmat, ecut_ry)
    dg = (gmax - gmin) / nsplin
    return gmin, dg, gmin + dg * np.arange(nsplin)


def spline_fit(c0, x):
    """Rows (c1, c2, c3) of spline_fit (friend_lib.C:134-200): end slopes from
    the end divided differences, interior slopes from the tridiagonal
    system 2 (d_m + d_{m+1}) s_m + d_{m+1} s_{m-1} + d_m s_{m+1} =
    3 (d_m q_{m+1} + d_{m+1} q_m) (q the divided differences), solved by
    forward elimination and back substitution; then per interval
    c2 = (q - s_i - e)/dx, c3 = e/dx^2 with e = s_i + s_{i+1} - 2 q."""
    c0 = np.asarray(c0, np.float64)
    x = np.asarray(x, np.float64)
    n = c0.size
    d = np.zeros(n)
    q = np.zeros(n)
    d[1:] = x[1:] - x[:-1]
    q[1:] = (c0[1:] - c0[:-1]) / d[1:]
    s = np.zeros(n)
    s[0] = q[1]
    s[-1] = q[-1]
    diag = np.zeros(n)
    diag[0] = 1.0
    rhs = s.copy()
    for m in range(1, n - 1):
        rhs[m] = 3.0 * (d[m] * q[m + 1] + d[m + 1] * q[m])
        diag[m] = 2.0 * (d[m] + d[m + 1])
    # elimination: the first row is the fixed end slope (its off-diagonal
    # in row 1 acts through d[2] s[0])
    for m in range(1, n - 1):
        gfac = -d[m + 1] / diag[m - 1]
        diag[m] += gfac * (d[m - 1] if m > 1 else 0.0)
        rhs[m] += gfac * rhs[m - 1]
    for m in range(n - 2, 0, -1):
        rhs[m] = (rhs[m] - d[m] * rhs[m + 1]) / diag[m]
    s[1:-1] = rhs[1:-1]
    c1 = s
    c2 = np.zeros(n)
    c3 = np.zeros(n)
    dx = d[1:]
    e = s[:-1] + s[1:] - 2.0 * q[1:]
    c2[:-1] = (q[1:] - s[:-1] - e) / dx
    c3[:-1] = e / (dx * dx)
    return c1, c2, c3


def _j0(x):
    return np.sin(x) / x


def radial_l0(f, r, dr, g, power):
    """sum_{i>=2} 4 pi dr j0(g r_i) r_i^power f_i for every g (the L = 0
    quadrature of slow_bess_vps; the r = 0 point is skipped)."""
    rr = r[1:]
    w = 4.0 * np.pi * dr * rr ** power * f[1:]
    out = np.empty(g.size)
    for i0 in range(0, g.size, 256):  # blocks of g keep the temporaries small
        gb = g[i0:i0 + 256]
        out[i0:i0 + 256] = (_j0(gb[:, None] * rr[None, :]) * w[None, :]).sum(1)
    return out


def local_form_factor(p, loc_opt, g):
    """(v(g) on the grid, v(0)) of the local channel with its long-range
    tail (slow_bess_vps with ilong = 1, L = 0)."""
    vloc = p["v"][loc_opt]
    f = radial_l0(vloc * p["r"], p["r"], p["dr"], g, 1)
    g0 = 4.0 * np.pi * p["dr"] * float(np.sum(p["r"][1:] ** 2 * vloc[1:]))
    g2 = g * g
    for z, al in ((p["z1"], p["alpha1"]), (p["z2"], p["alpha2"])):
        if z != 0.0:
            f -= z * 4.0 * np.pi * np.exp(-0.25 * g2 / (al * al)) / g2
            g0 += 4.0 * np.pi * z / (4.0 * al * al)
    if p["zpol"] != 0.0:
        cpol = (1.0, -6.0, 15.0, -20.0, 15.0, -6.0, 1.0)
        for i, ci in enumerate(cpol):
            ag = i * p["gamma"]
            pre = 0.5 * p["zpol"] * 2.0 * np.pi * ci
            f -= pre * ((ag * ag - g2) * np.arctan2(g, ag) / g + ag * np.log(ag * ag + g2))
            if i >= 1:
                g0 -= 0.5 * p["zpol"] * 4.0 * np.pi * ci * ag * np.log(ag)
    return f, g0


def kb_projector(p, l, loc_opt, g):
    """(P(g), P(0), vnorm) of the Kleinman-Bylander channel l (l = 0)."""
    if l != 0:
        raise ValueError("only the l = 0 channel (the reference's G-space method)")
    v_rphi = (p["v"][l] - p["v"][loc_opt]) * p["rphi"][l]
    amat = float(np.sum(p["rphi"][l] * v_rphi) * p["dr"])
    f = radial_l0(v_rphi, p["r"], p["dr"], g, 1)
    g0 = 4.0 * np.pi * p["dr"] * float(np.sum(p["r"][1:] * v_rphi[1:]))
    return f, g0, 1.0 / amat


def _rows(f, g):
    c1, c2, c3 = spline_fit(f, g)
    return np.stack([f, c1, c2, c3])


def water_tables(hmat, ecut_ry, db_dir=DB_DIR, nsplin=4000, species_order=("O", "H")):
    """Local tables (dict of oa_local_pseudo fields) and the KB tables of the
    oxygen s channel (dict of oa_nonlocal fields) for a water cell."""
    gmin, dg, g = spline_grid(hmat, ecut_ry, nsplin)
    vps = np.zeros((len(species_order), 4, nsplin))
    vps_g0 = np.zeros(len(species_order))
    q = np.zeros(len(species_order))
    kb_order, proj, proj_g0, vnorm = [], [], [], []
    for s, name in enumerate(species_order):
        e = WATER_SPECIES[name]
        p = parse_pseud(os.path.join(db_dir, e["file"]))
        f, f0 = local_form_factor(p, e["loc_opt"], g)
        vps[s] = _rows(f, g)
        vps_g0[s] = f0
        q[s] = p["z1"] + p["z2"]
        if e["typ"] == "KB":
            for l in range(e["n_ang"] + 1):
                if l == e["loc_opt"]:
                    continue
                pf, p0, vn = kb_projector(p, l, e["loc_opt"], g)
                kb_order.append(name)
                proj.append(_rows(pf, g))
                proj_g0.append(p0)
                vnorm.append(vn)
    local = dict(gmin=gmin, dg=dg, vps=vps, vps_g0=vps_g0, order=tuple(species_order),
                 q_pseud=q)
    kb = dict(gmin=gmin, dg=dg, proj=np.array(proj), proj_g0=np.array(proj_g0),
              vnorm=np.array(vnorm), order=tuple(kb_order))
    return local, kb
