"""The three-piece bf16 split behind the FP32 GEMM's SNU_GEMM_SPLIT form,
checked on the CPU against exact arithmetic (no GPU).

k_gemm_split writes every FP32 operand x as bf16 pieces h = bf16(x),
m = bf16(x - h), l = bf16(x - h - m) (round to nearest even, the rounding of
gfx950's v_cvt_pk_bf16_f32) and sums six products hh + hm + mh + hl + mm + lh
in FP32.  The claims its comments rest on:
  * x - h and x - h - m are exact in FP32, and h + m + l == x exactly;
  * |m| <= 2^-8 |x| and |l| <= 2^-16 |x|;
  * each bf16 x bf16 product is exact in FP32;
  * the terms left out (ml, lm, ll) total at most about 2^-23 |a| |b|,
    while leaving out any of the six kept terms can exceed that.
Products of float32 values are exact in float64, so float64 is the reference.
"""
import numpy as np


def bf16_rne(x32: np.ndarray) -> np.ndarray:
    """float32 -> bf16 (round to nearest, ties to even), returned as float32."""
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split3(x32: np.ndarray):
    h = bf16_rne(x32)
    r = (x32 - h).astype(np.float32)
    m = bf16_rne(r)
    r2 = (r - m).astype(np.float32)
    l = bf16_rne(r2)
    return h, m, l, r, r2


def _samples(rng, n):
    """FP32 values over a wide exponent range, both signs, plus values whose
    low bits force ties and carries in the rounding."""
    x = (rng.standard_normal(n) * 2.0 ** rng.integers(-60, 61, n)).astype(np.float32)
    bits = rng.integers(0, 1 << 23, n).astype(np.uint32) | np.uint32(0x3F800000)
    bits[: n // 4] |= np.uint32(0x8000)       # exact ties at the first split
    bits[n // 4: n // 2] |= np.uint32(0xFFFF)  # carries into h
    y = bits.view(np.float32) * np.float32(1.5)
    return np.concatenate([x, y, -y])


def test_pieces_reproduce_every_operand_exactly():
    rng = np.random.default_rng(11)
    x = _samples(rng, 200000)
    h, m, l, r, r2 = split3(x)
    x64 = x.astype(np.float64)
    # the residuals are exact in FP32
    assert np.array_equal(r.astype(np.float64), x64 - h.astype(np.float64))
    assert np.array_equal(r2.astype(np.float64), r.astype(np.float64) - m.astype(np.float64))
    # the last piece needs no rounding, so the pieces sum to x exactly
    assert np.array_equal(l, r2)
    assert np.array_equal(h.astype(np.float64) + m.astype(np.float64) + l.astype(np.float64), x64)
    ax = np.abs(x64)
    assert np.all(np.abs(m.astype(np.float64)) <= 2.0 ** -8 * ax)
    assert np.all(np.abs(l.astype(np.float64)) <= 2.0 ** -16 * ax)
    # every piece is a bf16 (low 16 bits clear)
    for p in (h, m, l):
        assert not np.any(p.view(np.uint32) & 0xFFFF)


def test_bf16_products_are_exact_in_fp32():
    rng = np.random.default_rng(12)
    a = bf16_rne(_samples(rng, 50000))
    b = bf16_rne(_samples(rng, 50000)[::-1].copy())
    keep = (np.abs(a.astype(np.float64) * b.astype(np.float64)) < 1e30) & \
           (np.abs(a.astype(np.float64) * b.astype(np.float64)) > 1e-30)
    p32 = (a[keep] * b[keep]).astype(np.float64)
    assert np.array_equal(p32, a[keep].astype(np.float64) * b[keep].astype(np.float64))


def test_six_products_bound_and_each_is_needed():
    rng = np.random.default_rng(13)
    n = 200000
    a = (rng.standard_normal(n) * 2.0 ** rng.integers(-20, 21, n)).astype(np.float32)
    b = (rng.standard_normal(n) * 2.0 ** rng.integers(-20, 21, n)).astype(np.float32)
    ha, ma, la, _, _ = (p.astype(np.float64) for p in split3(a))
    hb, mb, lb, _, _ = (p.astype(np.float64) for p in split3(b))
    exact = a.astype(np.float64) * b.astype(np.float64)
    terms = {"hh": ha * hb, "hm": ha * mb, "mh": ma * hb, "hl": ha * lb, "mm": ma * mb, "lh": la * hb}
    six = sum(terms.values())
    scale = np.abs(a.astype(np.float64)) * np.abs(b.astype(np.float64))
    err = np.abs(six - exact) / scale
    # the terms left out: at most about 2^-23 |a||b|, the size of two FP32
    # roundings of one product
    assert err.max() <= 2.0 ** -23 * (1 + 2.0 ** -7)
    # the dropped terms are exactly ml + lm + ll
    assert np.allclose(exact - six, ma * lb + la * mb + la * lb, rtol=0, atol=1e-300)
    # leaving out any kept term breaks FP32 accuracy for some operands
    for k, t in terms.items():
        worst = (np.abs(six - t - exact) / scale).max()
        assert worst > 2.0 ** -23, (k, worst)
