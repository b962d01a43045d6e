"""STATES files (include/oa_state.h) on the CPU: the library's host-side
readers and writers against independent Python statements of the
reference's encodings (header, record layout, point order, G = 0 last,
"%g" text), round trips in all four encodings, files in any point order,
and the error cases."""
import gzip
import struct

import numpy as np
import pytest

import oracle
from openatom_amd import OaError, states, system

GRID = (24, 27, 30)


def case():
    h = np.diag([b * system.BOHR_PER_ANG for b in (5.0, 5.4, 5.8)]).reshape(9)
    kx, ky, kz = oracle.stored_kvecs(14.0, h)
    psi = system.initial_coeffs(kx, ky, kz, h, 2, 5)[1]  # Hermitian on kx = 0
    return kx, ky, kz, psi


def unique_order(kx, ky, kz):
    """The reference writer's order: unique points by (kx, ky, kz), G = 0 last."""
    idx = [i for i in range(len(kx))
           if not (kx[i] == 0 and (ky[i] < 0 or (ky[i] == 0 and kz[i] < 0)))]
    idx.sort(key=lambda i: (kx[i], ky[i], kz[i]))
    assert (kx[idx[0]], ky[idx[0]], kz[idx[0]]) == (0, 0, 0)
    return idx[1:] + idx[:1]


def test_text_writer_matches_reference_encoding(tmp_path):
    kx, ky, kz, psi = case()
    p = tmp_path / "state1.out"
    states.write_state(p, psi, kx, ky, kz, GRID, fmt=states.TEXT)
    order = unique_order(kx, ky, kz)
    expect = [f"{len(order)} {GRID[0]} {GRID[1]} {GRID[2]}\n"]
    expect += ["%g %g %d %d %d \n" % (psi[i].real, psi[i].imag, kx[i], ky[i], kz[i])
               for i in order]
    assert p.read_text() == "".join(expect)


@pytest.mark.parametrize("fmt", [states.BINARY, states.BINARY_GZ])
def test_binary_layout_and_exact_round_trip(tmp_path, fmt):
    kx, ky, kz, psi = case()
    p = tmp_path / "state1.out"
    states.write_state(p, psi, kx, ky, kz, GRID, fmt=fmt)
    raw = (gzip.decompress((tmp_path / "state1.out.gz").read_bytes())
           if fmt == states.BINARY_GZ else p.read_bytes())
    order = unique_order(kx, ky, kz)
    expect = struct.pack("=4i", len(order), *GRID) + b"".join(
        struct.pack("=ddiii", psi[i].real, psi[i].imag, kx[i], ky[i], kz[i]) for i in order)
    assert raw == expect
    back, grid = states.read_state(p, kx, ky, kz, fmt=fmt)
    assert grid == GRID
    assert np.array_equal(back, psi)  # redundant half rebuilt as conj(partner), bit for bit


@pytest.mark.parametrize("fmt", [states.TEXT, states.TEXT_GZ])
def test_text_round_trip_six_digits(tmp_path, fmt):
    kx, ky, kz, psi = case()
    p = tmp_path / "state1.out"
    states.write_state(p, psi, kx, ky, kz, GRID, fmt=fmt)
    back, _ = states.read_state(p, kx, ky, kz, fmt=fmt)
    assert np.abs(back - psi).max() <= 5e-6 * np.abs(psi).max()


def test_reader_accepts_any_point_order(tmp_path):
    """Files written by other generators (e.g. PINY's own half-sphere order)
    are matched point by point."""
    kx, ky, kz, psi = case()
    order = unique_order(kx, ky, kz)
    rng = np.random.default_rng(1)
    shuffled = list(rng.permutation(order))
    p = tmp_path / "s.out"
    p.write_bytes(struct.pack("=4i", len(order), *GRID) + b"".join(
        struct.pack("=ddiii", psi[i].real, psi[i].imag, kx[i], ky[i], kz[i])
        for i in shuffled))
    back, _ = states.read_state(p, kx, ky, kz, fmt=states.BINARY)
    assert np.array_equal(back, psi)


def test_reader_rejects_missing_and_foreign_points(tmp_path):
    kx, ky, kz, psi = case()
    order = unique_order(kx, ky, kz)
    hdr = lambda n: struct.pack("=4i", n, *GRID)  # noqa: E731
    rec = lambda i: struct.pack("=ddiii", psi[i].real, psi[i].imag, kx[i], ky[i], kz[i])  # noqa

    missing = tmp_path / "missing.out"
    missing.write_bytes(hdr(len(order) - 1) + b"".join(rec(i) for i in order[1:]))
    with pytest.raises(OaError):
        states.read_state(missing, kx, ky, kz)
    foreign = tmp_path / "foreign.out"
    foreign.write_bytes(hdr(len(order) + 1) + b"".join(rec(i) for i in order) +
                        struct.pack("=ddiii", 1.0, 0.0, 99, 0, 0))
    with pytest.raises(OaError):
        states.read_state(foreign, kx, ky, kz)
    short = tmp_path / "short.out"
    short.write_bytes(hdr(len(order)) + b"".join(rec(i) for i in order[:-3]))
    with pytest.raises(OaError):
        states.read_state(short, kx, ky, kz)
    with pytest.raises(OaError):
        states.read_state(tmp_path / "absent.out", kx, ky, kz)
