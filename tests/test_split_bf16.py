"""Error model of the split-BF16 products of the mixed-precision step, on the
CPU (numpy emulation of the device arithmetic):

- each operand is rounded to FP32, f = fl32(x) (2^-24 relative), and f is
  split exactly into three BF16 terms f = h + m + l (round to nearest at
  each level, the residuals exact in FP32); keeping the six term products
  with i + j <= 2 loses less than 2^-24 of a product;
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
    """The device split: f = fl32(x), h = bf16(f), m = bf16(f - h),
    l = bf16(f - h - m); every subtraction exact in FP32."""
    f = np.asarray(x, np.float64).astype(np.float32)
    h = bf16(f)
    r1 = f - h
    m = bf16(r1)
    r2 = r1 - m
    lo = bf16(r2)
    return h.astype(np.float64), m.astype(np.float64), lo.astype(np.float64)


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


def test_split_is_exact_for_the_fp32_rounding():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(200000) * np.exp(rng.uniform(-20, 20, 200000))
    h, m, lo = split3(x)
    f = x.astype(np.float32).astype(np.float64)
    assert np.array_equal(h + m + lo, f)  # three BF16 terms hold an FP32 value exactly
    rel = np.abs((h + m + lo) - x) / np.abs(x)
    assert rel.max() <= 2.0 ** -24
    assert np.all(np.abs(m) <= 2.0 ** -8 * np.abs(f)) and np.all(np.abs(lo) <= 2.0 ** -16 * np.abs(f))
    # each term is a BF16 value: 8 significant bits
    for t in (h, m, lo):
        t32 = t.astype(np.float32)
        assert np.array_equal(bf16(t32), t32)


def test_six_products_keep_the_product_of_the_fp32_roundings():
    rng = np.random.default_rng(2)
    a = rng.standard_normal(100000)
    b = rng.standard_normal(100000)
    ah, am, al = split3(a)
    bh, bm, bl = split3(b)
    p = al * bh + am * bm + ah * bl + am * bh + ah * bm + ah * bh
    fa, fb = a.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)
    # the three dropped products (m l, l m, l l) are below 2^-24 of the product
    # of the roundings
    assert (np.abs(p - fa * fb) / np.abs(fa * fb)).max() < 2.0 ** -24
    # against the FP64 product the FP32 rounding of the operands dominates
    assert (np.abs(p - a * b) / np.abs(a * b)).max() < 2.0 ** -22


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
