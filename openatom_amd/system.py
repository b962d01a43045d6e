"""Seeded synthetic inputs for the CP electronic step: liquid-water boxes,
Gaussian-seeded orthonormal coefficients and a Gaussian local potential.

Workloads (sizes follow the measurement plan of the build):
  water8   :   8 H2O, L = 6.21 A,   32 states, 30 Ry, 48^3
  water32  :  32 H2O, L = 9.86 A,  128 states, 70 Ry, 100^3
  water128 : 128 H2O, L = 15.65 A, 512 states, 70 Ry, 160^3

Coefficients follow c_s(G) = (xi1 + i xi2) exp(-|G|^2/4) on the unique half
sphere, Hermitian on the kx=0 plane, c(0) real, orthonormalised under the
Gamma inner product <a,b> = sum_stored w_p Re(a* b) (w = 1 on kx=0, 2
elsewhere) and scaled to the reference's norm 2
(src_piny_physics_v1.0/interface/coords_cp/gen_wave.C:699-710).
"""
import numpy as np

from . import _lib

BOHR_PER_ANG = 1.8897261246257702
OH_ANG = 0.9572
HOH_DEG = 104.52

WORKLOADS = {
    "water8": dict(nmol=8, box_ang=6.21, nstates=32, ecut_ry=30.0, grid=48),
    "water32": dict(nmol=32, box_ang=9.86, nstates=128, ecut_ry=70.0, grid=100),
    "water128": dict(nmol=128, box_ang=15.65, nstates=512, ecut_ry=70.0, grid=160),
}

# FFT lengths with compiled two-factor kernels (openatom_amd/csrc/oa_fft2_sizes.h)
FFT_SIZES = (24, 27, 30, 32, 36, 40, 45, 48, 50, 54, 60, 64, 72, 75, 80, 81, 90, 96, 100, 108,
             120, 128, 144, 150, 160)


def serving_workload(nmol, ecut_ry):
    """Water box of the request mix: nmol molecules at the density of the
    32-molecule box (side 9.86 (nmol/32)^(1/3) A), four doubly occupied states
    per molecule, and the smallest compiled cubic grid that holds the density
    sphere |G|^2 <= 4 Ecut: N >= 2 kr + 1, kr the largest integer component
    inside it, floor(2 sqrt(Ecut) L / 2 pi) in a cube of side L (twice the
    wavefunction sphere's kmax can fall one short of it)."""
    box = 9.86 * (nmol / 32.0) ** (1.0 / 3.0)
    L = box * BOHR_PER_ANG
    kr = int(np.floor(2.0 * np.sqrt(ecut_ry) * L / (2.0 * np.pi)))
    need = 2 * kr + 1
    grid = next((n for n in FFT_SIZES if n >= need), FFT_SIZES[-1])
    return dict(nmol=int(nmol), box_ang=box, nstates=4 * int(nmol), ecut_ry=float(ecut_ry),
                grid=grid)


def _rand_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def place_water(nmol, box_ang, seed, min_oo_ang=2.5, max_tries=200000):
    """O uniform in a cubic box, min-image O-O >= min_oo_ang; H at 0.9572 A and
    104.52 deg with a uniform random orientation.  Returns (pos_bohr, species)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = box_ang
    oxy = []
    tries = 0
    while len(oxy) < nmol:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place water molecules")
        p = rng.uniform(0.0, L, size=3)
        ok = True
        for q in oxy:
            d = p - q
            d -= L * np.round(d / L)
            if np.dot(d, d) < min_oo_ang ** 2:
                ok = False
                break
        if ok:
            oxy.append(p)
    half = np.deg2rad(HOH_DEG) / 2.0
    h1 = OH_ANG * np.array([np.sin(half), np.cos(half), 0.0])
    h2 = OH_ANG * np.array([-np.sin(half), np.cos(half), 0.0])
    pos, spec = [], []
    for o in oxy:
        R = _rand_rotation(rng)
        pos += [o, o + R @ h1, o + R @ h2]
        spec += ["O", "H", "H"]
    pos = np.mod(np.array(pos), L) * BOHR_PER_ANG
    return pos, spec


def cubic_hmat(box_bohr):
    return np.diag([box_bohr] * 3).reshape(9).astype(np.float64)


def gvec_sq(kx, ky, kz, hmat):
    hi = np.linalg.inv(np.asarray(hmat, np.float64).reshape(3, 3))
    k = np.stack([kx, ky, kz]).astype(np.float64)
    G = 2.0 * np.pi * (hi @ k)
    return (G * G).sum(axis=0)


def redundant_pairs(kx, ky, kz):
    """(redundant index, unique partner index) on the kx=0 plane."""
    idx = {(int(b), int(c)): i for i, (a, b, c) in enumerate(zip(kx, ky, kz)) if a == 0}
    red, par = [], []
    for i, (a, b, c) in enumerate(zip(kx, ky, kz)):
        if a == 0 and (b < 0 or (b == 0 and c < 0)):
            red.append(i)
            par.append(idx[(-int(b), -int(c))])
    return np.array(red, np.int64), np.array(par, np.int64)


def gamma_weights(kx):
    return np.where(np.asarray(kx) == 0, 1.0, 2.0)


def gamma_overlap(psi, kx):
    """S_ij = sum_p w_p Re(conj psi_i psi_j): the full-sphere overlap."""
    w = np.sqrt(gamma_weights(kx))[None, :]
    Ar, Ai = psi.real * w, psi.imag * w
    return Ar @ Ar.T + Ai @ Ai.T


def initial_coeffs(kx, ky, kz, hmat, nstates, seed, norm=2.0):
    """Seeded Gaussian coefficients, Gram-Schmidt orthonormalised (Cholesky-QR
    twice, which reproduces Gram-Schmidt's triangular factor), norm `norm`."""
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    npts = len(kx)
    g2 = gvec_sq(kx, ky, kz, hmat)
    amp = np.exp(-g2 / 4.0)
    xi = rng.normal(size=(nstates, npts)) + 1j * rng.normal(size=(nstates, npts))
    psi = xi * amp[None, :]
    red, par = redundant_pairs(kx, ky, kz)
    psi[:, red] = np.conj(psi[:, par])
    g0 = np.where((np.asarray(kx) == 0) & (np.asarray(ky) == 0) & (np.asarray(kz) == 0))[0]
    psi[:, g0] = psi[:, g0].real
    from scipy.linalg import solve_triangular
    for _ in range(2):
        S = gamma_overlap(psi, kx)
        Lc = np.linalg.cholesky(S)
        # Lc is real: two real triangular solves instead of a complex LU solve
        re = solve_triangular(Lc, psi.real, lower=True, check_finite=False)
        im = solve_triangular(Lc, psi.imag, lower=True, check_finite=False)
        psi = np.ascontiguousarray(re + 1j * im)  # the solves return Fortran order
    return psi * np.sqrt(norm)


def gaussian_vloc(pos, spec, hmat, grid, seed):
    """V(r) = sum_a A_a exp(-|r - R_a|^2 / (2 sigma_a^2)) with per-species
    (A, sigma) drawn from the seed; orthorhombic minimum image.  The ranges
    give each molecule four bound levels below a gap of about 0.7 Ha (for
    8 H2O at 30 Ry: 32 levels spread over 1.8 Ha, then the gap), like the
    four doubly occupied valence levels of water, so minimisations converge
    as they do for a real insulator.  Returns (V[Nz][Ny][Nx], params)."""
    rng = np.random.Generator(np.random.PCG64(seed + 104729))
    params = {"O": (float(rng.uniform(-6.5, -5.5)), float(rng.uniform(0.95, 1.05))),
              "H": (float(rng.uniform(-1.3, -1.1)), float(rng.uniform(0.55, 0.65)))}
    h = np.asarray(hmat, np.float64).reshape(3, 3)
    Lx, Ly, Lz = h[0, 0], h[1, 1], h[2, 2]
    nx, ny, nz = grid
    xs = np.arange(nx) * (Lx / nx)
    ys = np.arange(ny) * (Ly / ny)
    zs = np.arange(nz) * (Lz / nz)
    V = np.zeros((nz, ny, nx))
    for p, s in zip(pos, spec):
        A, sig = params[s]
        dx = xs - p[0]; dx -= Lx * np.round(dx / Lx)
        dy = ys - p[1]; dy -= Ly * np.round(dy / Ly)
        dz = zs - p[2]; dz -= Lz * np.round(dz / Lz)
        c = -1.0 / (2.0 * sig * sig)
        ex = np.exp(c * dx * dx)
        ey = np.exp(c * dy * dy)
        ez = np.exp(c * dz * dz)
        V += A * ez[:, None, None] * ey[None, :, None] * ex[None, None, :]
    return V, params


def gaussian_form_factor(A, sigma, g):
    """Fourier transform of A exp(-r^2 / (2 sigma^2)): A (2 pi sigma^2)^(3/2)
    exp(-sigma^2 g^2 / 2), and its derivative in g."""
    f = A * (2.0 * np.pi * sigma * sigma) ** 1.5 * np.exp(-0.5 * sigma * sigma * g * g)
    return f, -sigma * sigma * g * f


def _spline_grid(ecut_ry, dg):
    """g_i = i dg up to the density cutoff sqrt(4 ecut_ry), with margin."""
    gmax = np.sqrt(4.0 * ecut_ry) + 2 * dg
    nspl = int(np.ceil(gmax / dg)) + 2
    return np.arange(nspl + 1) * dg, nspl


def _hermite_rows(f, d, dg):
    """Rows c0..c3 of the cubic Hermite interpolant of values f and slopes d
    on a uniform grid, evaluated as ((c3 h + c2) h + c1) h + c0."""
    df = (f[1:] - f[:-1]) / dg
    return np.stack([f[:-1], d[:-1], (3.0 * df - 2.0 * d[:-1] - d[1:]) / dg,
                     (d[:-1] + d[1:] - 2.0 * df) / (dg * dg)])


def local_pseudo_tables(params, ecut_ry, dg=0.002, species_order=("O", "H")):
    """Spline tables of the Gaussian local potential in the layout of PINY's
    vps0..vps3 (cubic Hermite on g_i = i dg up to the density cutoff
    sqrt(4 ecut_ry)): returns dict(gmin, dg, vps[nspecies][4][nspline],
    vps_g0[nspecies], order)."""
    g, nspl = _spline_grid(ecut_ry, dg)
    vps = np.zeros((len(species_order), 4, nspl))
    g0 = np.zeros(len(species_order))
    for s, name in enumerate(species_order):
        A, sig = params[name]
        f, d = gaussian_form_factor(A, sig, g)
        vps[s] = _hermite_rows(f, d, dg)
        g0[s] = f[0]
    return dict(gmin=0.0, dg=dg, vps=vps, vps_g0=g0, order=tuple(species_order))


def kb_params(seed):
    """Seeded l = 0 Kleinman-Bylander channel for oxygen (the KB species of
    the water inputs; hydrogen is local only): a Gaussian projector
    beta(r) = B exp(-r^2 / (2 s^2)) and its normalisation vnorm.  Positive
    vnorm makes the channel repulsive, as the s channel of an oxygen
    pseudopotential with a p local channel is.  Returns {"O": (B, s, vnorm)}."""
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    s = float(rng.uniform(0.50, 0.60))
    return {"O": (float(rng.uniform(0.9, 1.1)) / (2.0 * np.pi * s * s) ** 1.5, s,
                  float(rng.uniform(15.0, 25.0)))}


def nonlocal_tables(params, ecut_ry, dg=0.002, species_order=("O",)):
    """Spline tables of the KB projectors in the layout of PINY's l = 0
    vps rows (as local_pseudo_tables): dict(gmin, dg, proj[nspecies][4]
    [nspline], proj_g0[nspecies], vnorm[nspecies], order)."""
    g, nspl = _spline_grid(ecut_ry, dg)
    proj = np.zeros((len(species_order), 4, nspl))
    g0 = np.zeros(len(species_order))
    vn = np.zeros(len(species_order))
    for s, name in enumerate(species_order):
        B, sig, vnorm = params[name]
        f, d = gaussian_form_factor(B, sig, g)
        proj[s] = _hermite_rows(f, d, dg)
        g0[s] = f[0]
        vn[s] = vnorm
    return dict(gmin=0.0, dg=dg, proj=proj, proj_g0=g0, vnorm=vn, order=tuple(species_order))


def kb_setup(pos, spec, ecut_ry, seed):
    """(KB parameters, dict of nonlocal_tables plus the non-local atoms:
    indices, positions and species indices into the table order)."""
    kb = kb_params(seed)
    nlt = nonlocal_tables(kb, ecut_ry)
    nl_atoms = [i for i, s in enumerate(spec) if s in nlt["order"]]
    nlt.update(atoms=np.array(nl_atoms, np.int32), pos=np.asarray(pos)[nl_atoms],
               species=np.array([nlt["order"].index(spec[i]) for i in nl_atoms], np.int32))
    return kb, nlt


def build(workload="water32", seed=1234, nstates=None, kvec_source=None, pseudo="gaussian"):
    """Build one seeded system.  kvec_source(ecut_ry, hmat) -> (kx, ky, kz)
    defaults to the library's own enumeration (oa_kvecs).  pseudo: "gaussian"
    for the seeded Gaussian local potential and KB projector, "db" for the
    local and oxygen s-channel KB tables of the reference's pseudopotential
    database (openatom_amd/pseudo_db.py; the fixed V(r) of the system, used
    only by fixed-potential steps, stays the seeded Gaussian one)."""
    w = dict(WORKLOADS[workload]) if isinstance(workload, str) else dict(workload)
    nst = int(nstates or w["nstates"])
    L = w["box_ang"] * BOHR_PER_ANG
    hmat = cubic_hmat(L)
    grid = (w["grid"],) * 3 if np.isscalar(w["grid"]) else tuple(w["grid"])
    pos, spec = place_water(w["nmol"], w["box_ang"], seed)
    src = kvec_source or _lib.kvecs
    kx, ky, kz = src(w["ecut_ry"], hmat)
    psi = initial_coeffs(kx, ky, kz, hmat, nst, seed)
    V, params = gaussian_vloc(pos, spec, hmat, grid, seed)
    if pseudo == "db":
        from . import pseudo_db
        tab, nlt = pseudo_db.water_tables(hmat, w["ecut_ry"])
        kb = None
        nl_atoms = [i for i, sp in enumerate(spec) if sp in nlt["order"]]
        nlt.update(atoms=np.array(nl_atoms, np.int32), pos=np.asarray(pos)[nl_atoms],
                   species=np.array([nlt["order"].index(spec[i]) for i in nl_atoms], np.int32))
    elif pseudo == "gaussian":
        tab = local_pseudo_tables(params, w["ecut_ry"])
        kb, nlt = kb_setup(pos, spec, w["ecut_ry"], seed)
    else:
        raise ValueError(f"pseudo must be 'gaussian' or 'db', not {pseudo!r}")
    species = np.array([tab["order"].index(s) for s in spec], np.int32)
    return dict(name=workload if isinstance(workload, str) else "custom", seed=seed,
                hmat=hmat, grid=grid, ecut_ry=float(w["ecut_ry"]), nstates=nst,
                natoms=len(spec), nmol=w["nmol"], pos=pos, spec=spec, kx=kx, ky=ky, kz=kz,
                psi=psi, vks=V, vparams=params, species=species, pseudo=tab,
                kbparams=kb, kb=nlt, pseudo_source=pseudo)
