"""CPU: the constant tables.  The product's (mpmath, yume-pdq_amd/csrc) and
the oracle's (GNU MPFR, oracle/) generators agree bit for bit, and both match
the closed forms of build.rs:316-387."""
import os
import re
import struct
from fractions import Fraction

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def product_tables():
    text = open(os.path.join(REPO, "yume-pdq_amd", "csrc", "pdq_tables.h")).read()

    def grab(name):
        body = re.search(r"#define " + name + r" \\\n(.*?)\n\n", text + "\n\n", re.S).group(1)
        vals = [v.strip() for v in body.replace("\\", "").split(",") if v.strip()]
        return np.array([float.fromhex(v[:-1]) if v != "0.0f" else 0.0 for v in vals], np.float64)

    return grab("PDQ_DCT_F32_INIT").astype(np.float32), grab("PDQ_TENT_X8_INIT").astype(np.float32)


def test_product_tables_equal_oracle_tables(oracle):
    pd, pt = product_tables()
    od, _, ot = oracle.tables()
    assert pd.shape == od.shape == (16 * 127 + 16,)
    assert np.array_equal(pd.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(pt.view(np.uint32), ot.view(np.uint32))


def test_dct_closed_form(oracle):
    """d(i,j) = sqrt(2/127) cos(pi (i+1)(2j+1) / 254) rounded to f32, except the
    eight entries whose exact value is 0 (odd frequency, j = 63), which hold
    the 128-bit residue (|v| < 1e-37)."""
    import mpmath
    d32, d64, _ = oracle.tables()
    with mpmath.workdps(60):
        for k in range(16):
            for j in range(127):
                exact = mpmath.sqrt(mpmath.mpf(2) / 127) * mpmath.cos(mpmath.pi * (k + 1) * (2 * j + 1) / 254)
                v = d32[k * 127 + j]
                if (k + 1) % 2 == 1 and j == 63:
                    assert abs(v) < 1e-37 and abs(d64[k * 127 + j]) < 1e-37
                else:
                    assert d64[k * 127 + j] == float(exact), (k, j)
                    assert v == np.float32(float(exact)), (k, j)
    assert np.all(d32[16 * 127:] == 0)


def test_dct_orthonormal_rows(oracle):
    _, d64, _ = oracle.tables()
    D = d64[:16 * 127].reshape(16, 127)
    assert np.allclose(D @ D.T, np.eye(16), atol=1e-13)


def test_tent_weights_closed_form(oracle):
    """Interior weights are round_f32(a_i a_j), a = (0,1,2,3,2,1,0)/9; the
    128-bit box filter leaves positive residues only in row 0 (cols 4, 5)."""
    _, _, t = oracle.tables()
    t = t.reshape(7, 8)
    a = [0, 1, 2, 3, 2, 1, 0]
    for i in range(7):
        for j in range(7):
            w = Fraction(a[i] * a[j], 81)
            if w:
                assert t[i, j] == np.float32(float(w)), (i, j)
                # exactness of the f32 rounding of the rational
                assert abs(Fraction(float(t[i, j])) - w) <= Fraction(1, 2 ** 24) * w
            elif i == 0 and j in (4, 5):
                assert 0 < t[i, j] < 1e-38  # subnormal residue kept by build.rs:373-375
            else:
                assert t[i, j] == 0
    assert np.all(t[:, 7] == 0)
    # the residues, bit for bit
    assert struct.unpack("<I", struct.pack("<f", t[0, 4]))[0] == 0x00015555
    assert struct.unpack("<I", struct.pack("<f", t[0, 5]))[0] == 0x0000AAAB
