"""Error model of the split-BF16 products of the mixed-precision step, on the
CPU (numpy emulation of the device arithmetic):

- the three-term split x = h + m + l (each a BF16 value, the residuals exact
  in FP64) represents x to about 2^-26 relative, and keeping the six term
  products with i + j <= 2 loses about 2^-26 of a product;
- the matrix cores sum those products in FP32 within a K slice, so a gram
  entry carries the FP32 rounding of its slice sums.  Emulated on seeded
  water8 coefficients: S = Psi^H Psi in split BF16 lands at about 1e-8 to
  1e-7 of its entries, four orders above the 1e-10 orthonormality bar, while
  Lambda's relative error (which only perturbs the search direction) is of
  the same size.  This is why S and Psi T are split only on request.
"""
import numpy as np

from openatom_amd import system


def bf16(x32):
    """Round float32 to the nearest BF16 (ties to even), returned as float32."""
    u = np.asarray(x32, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def split3(x):
    x = np.asarray(x, np.float64)
    h = bf16(x.astype(np.float32)).astype(np.float64)
    r1 = x - h
    m = bf16(r1.astype(np.float32)).astype(np.float64)
    r2 = r1 - m
    lo = bf16(r2.astype(np.float32)).astype(np.float64)
    return h, m, lo


def split_gram(A, B, slice_len=256, chunk=16):
    """C = A B^T as the device forms it: split operands, six term products,
    exact 16-term chunk sums rounded into an FP32 accumulator per slice,
    slices summed in FP64."""
    Ah, Am, Al = split3(A)
    Bh, Bm, Bl = split3(B)
    terms = ((Al, Bh), (Am, Bm), (Ah, Bl), (Am, Bh), (Ah, Bm), (Ah, Bh))
    K = A.shape[1]
    C = np.zeros((A.shape[0], B.shape[0]))
    for s0 in range(0, K, slice_len):
        acc = np.zeros((A.shape[0], B.shape[0]), np.float32)
        for c0 in range(s0, min(K, s0 + slice_len), chunk):
            c1 = min(K, c0 + chunk)
            for a, b in terms:
                part = a[:, c0:c1] @ b[:, c0:c1].T  # exact BF16 x BF16 products
                acc = (acc.astype(np.float64) + part).astype(np.float32)
        C += acc.astype(np.float64)
    return C


def test_split_represents_operands_to_2_pow_minus_26():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(200000) * np.exp(rng.uniform(-20, 20, 200000))
    h, m, lo = split3(x)
    rel = np.abs((h + m + lo) - x) / np.abs(x)
    assert rel.max() < 2.0 ** -25
    # each term is a BF16 value: 8 significant bits
    for t in (h, m, lo):
        t32 = t.astype(np.float32)
        assert np.array_equal(bf16(t32), t32)


def test_six_products_keep_a_product_to_2_pow_minus_24():
    rng = np.random.default_rng(2)
    a = rng.standard_normal(100000)
    b = rng.standard_normal(100000)
    ah, am, al = split3(a)
    bh, bm, bl = split3(b)
    p = al * bh + am * bm + ah * bl + am * bh + ah * bm + ah * bh
    rel = np.abs(p - a * b) / np.abs(a * b)
    assert rel.max() < 2.0 ** -24
    # dropping the three smallest products is what bounds it, not the split
    full = sum(x * y for x in (ah, am, al) for y in (bh, bm, bl))
    assert np.abs(full - a * b).max() <= np.abs(p - a * b).max()


def _gamma_rows(psi, kx):
    """Real rows whose plain dot products give the Gamma-point overlap in the
    reference scaling: kx = 0 points weight 1, the others weight 2."""
    w = np.where(np.asarray(kx) == 0, 1.0, np.sqrt(2.0))
    z = psi * w[None, :]
    return np.concatenate([z.real, z.imag], axis=1)


def test_split_overlap_error_is_far_above_the_orthonormality_bar():
    s = system.build("water8", seed=11, nstates=16)
    A = _gamma_rows(s["psi"], s["kx"])
    S64 = A @ A.T
    S = split_gram(A, A)
    err = np.abs(S - S64).max()
    # about FP32 accumulation error of unit-sized entries ...
    assert 1e-10 < err < 1e-5
    # ... so ||Psi^H Psi/2 - I|| <= 1e-10 cannot rest on a split-BF16 S
    assert err / 2 > 1e-10


def test_split_lambda_error_is_fp32_sized():
    rng = np.random.default_rng(4)
    s = system.build("water8", seed=11, nstates=16)
    A = _gamma_rows(s["psi"], s["kx"])
    F = A * rng.standard_normal(A.shape) * 0.1
    L64 = A @ F.T
    L = split_gram(A, F)
    rel = np.abs(L - L64).max() / np.abs(L64).max()
    assert rel < 1e-6
