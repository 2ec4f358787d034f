"""The matrix-core matcher's packed candidate flags, checked over every case
of their arithmetic (no GPU): for every threshold t in 0..63 and every
triple of prefix dot products (even values in [-128, 128]), the fields of
one accumulator (csrc/pdq_match_mfma.hip: flags_c0 with two products,
flags_c0_tri with three) are exact integers below 2^24 in every order of
accumulation, and a product at the bound (dot >= 128 - 2t) always sets a
flag bit of the float's bit pattern: the prefix filter cannot lose a
candidate.  It may add some: only where the fields' own slack or borrows
(distant pairs) say so."""
import numpy as np

DOTS = np.arange(-128, 129, 2, dtype=np.int64)


def flags(v, mask):
    f = v.astype(np.float32)
    assert np.all(f.astype(np.int64) == v), "not exact in f32"
    return (f.view(np.uint32).astype(np.int64) & mask) != 0


def test_two_products():
    a, b = (x.ravel() for x in np.meshgrid(DOTS, DOTS, indexing="ij"))
    for t in range(64):
        bound = 128 - 2 * t
        k = 256 - bound
        c0 = 2 ** 23 + k + 512 * k
        lo, hi = c0 - 128 - 512 * 128, c0 + 128 + 512 * 128  # any partial sum
        assert 2 ** 23 <= lo and hi < 2 ** 24
        v = c0 + a + 512 * b
        got = flags(v, (1 << 8) | (1 << 17))
        assert np.array_equal(got, (a >= bound) | (b >= bound)), t  # exact, both ways


def test_three_products():
    a, b, c = (x.ravel() for x in np.meshgrid(DOTS, DOTS, DOTS, indexing="ij"))
    mask = (1 << 7) | (1 << 15) | (1 << 22) | (1 << 23)
    for t in range(64):
        bound = 128 - 2 * t
        c0 = 2 ** 23 + (128 - bound) + 2 ** 8 * (129 - bound // 2) + 2 ** 16 * (64 - bound // 2)
        lo = c0 - 128 - 2 ** 7 * 128 - 2 ** 15 * 128
        hi = c0 + 128 + 2 ** 7 * 128 + 2 ** 15 * 128
        assert 2 ** 22 < lo and hi < 2 ** 24  # integers, exact in any order
        v = c0 + a + 2 ** 7 * b + 2 ** 15 * c
        got = flags(v, mask)
        at_bound = (a >= bound) | (b >= bound) | (c >= bound)
        assert not (at_bound & ~got).any(), t  # no candidate lost
        extra = got & ~at_bound
        allowed = (b == bound - 2) | (a < bound - 128) | (c < bound - 128)
        assert not (extra & ~allowed).any(), t
