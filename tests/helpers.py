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
    V, _ = system.gaussian_vloc(pos, spec, hmat, grid, seed)
    return dict(name="small", seed=seed, hmat=hmat, grid=tuple(grid), ecut_ry=ecut,
                nstates=nstates, natoms=len(spec), kx=kx, ky=ky, kz=kz, psi=psi, vks=V)


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
