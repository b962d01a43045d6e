"""Shared test helpers: small seeded systems, independent numpy statements of
the reference formulas (for pinning the oracle), error metrics."""
import numpy as np

from openatom_amd import system


def small_system(nmol=2, box=(5.0, 5.4, 5.8), ecut=14.0, grid=(24, 27, 30), nstates=6,
                 seed=3, kvec_source=None):
    """Orthorhombic, unequal-grid system: catches axis and transpose mistakes."""
    w = dict(nmol=nmol, box_ang=box[0], nstates=nstates, ecut_ry=ecut, grid=grid)
    import oracle
    hmat = np.diag([b * system.BOHR_PER_ANG for b in box]).reshape(9)
    src = kvec_source or oracle.stored_kvecs
    kx, ky, kz = src(ecut, hmat)
    pos, spec = system.place_water(nmol, min(box), seed, min_oo_ang=2.0)
    psi = system.initial_coeffs(kx, ky, kz, hmat, nstates, seed)
    V, params = system.gaussian_vloc(pos, spec, hmat, grid, seed)
    tab = system.local_pseudo_tables(params, ecut)
    species = np.array([tab["order"].index(s) for s in spec], np.int32)
    kb, nlt = system.kb_setup(pos, spec, ecut, seed)
    return dict(name="small", seed=seed, hmat=hmat, grid=tuple(grid), ecut_ry=ecut,
                nstates=nstates, natoms=len(spec), kx=kx, ky=ky, kz=kz, psi=psi, vks=V,
                pos=pos, spec=spec, vparams=params, species=species, pseudo=tab,
                kbparams=kb, kb=nlt)


def oracle_pseudo(s):
    """(oracle LocalPseudo struct, arrays kept alive) of a built system."""
    import oracle
    t = s["pseudo"]
    return oracle.local_pseudo(s["pos"], s["species"], t["gmin"], t["dg"], t["vps"],
                               t["vps_g0"])


def oracle_nonlocal(s, pos=None):
    """(oracle NonLocal struct, arrays kept alive) of a built system's KB
    atoms (optionally at other positions)."""
    import oracle
    t = s["kb"]
    return oracle.nonlocal_(t["pos"] if pos is None else pos, t["species"], t["gmin"], t["dg"],
                            t["proj"], t["proj_g0"], t["vnorm"])


def periodic_gaussians(pos, spec, params, hmat, grid, images=2):
    """sum_I sum_n A exp(-|r - R_I - L n|^2 / (2 sigma^2)) over image shifts
    n in [-images, images]^3 of an orthorhombic cell: the exact periodic
    Gaussian potential the local-pseudopotential form factors stand for."""
    h = np.asarray(hmat, np.float64).reshape(3, 3)
    L = np.array([h[0, 0], h[1, 1], h[2, 2]])
    nx, ny, nz = grid
    ax = [np.arange(n) * (L[d] / n) for d, n in enumerate((nx, ny, nz))]
    V = np.zeros((nz, ny, nx))
    sh = np.arange(-images, images + 1)
    for p, s in zip(pos, spec):
        A, sig = params[s]
        c = -1.0 / (2.0 * sig * sig)
        e = []
        for d in range(3):
            dd = ax[d][:, None] - p[d] - L[d] * sh[None, :]
            e.append(np.exp(c * dd * dd).sum(axis=1))
        V += A * e[2][:, None, None] * e[1][None, :, None] * e[0][None, None, :]
    return V


def dense_grid(c, kx, ky, kz, grid):
    """Full complex grid [z][y][x] holding c(G) and, off the kx=0 plane, c(-G)=c*."""
    nx, ny, nz = grid
    C = np.zeros((nz, ny, nx), complex)
    kx = np.asarray(kx); ky = np.asarray(ky); kz = np.asarray(kz)
    np.add.at(C, (kz % nz, ky % ny, kx % nx), c)
    m = kx > 0
    np.add.at(C, ((-kz[m]) % nz, (-ky[m]) % ny, (-kx[m]) % nx), np.conj(c[m]))
    return C


def psi_r_numpy(c, kx, ky, kz, grid):
    """psi(r) = sum_G c(G) exp(-i G.r): numpy fftn uses the same sign."""
    return np.fft.fftn(dense_grid(c, kx, ky, kz, grid)).real


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    d = np.abs(a - b).max()
    s = np.abs(b).max()
    return d / s if s > 0 else d
