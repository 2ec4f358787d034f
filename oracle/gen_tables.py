#!/usr/bin/env python3
"""Oracle copy of the PDQ constant tables, computed with GNU MPFR at 128 bits.

TEST INFRASTRUCTURE ONLY: the tables written here feed the CPU oracle
(`oracle/pdq_oracle.c`), which checks the HIP product; nothing shipped reads them.

The reference generates its tables at build time with MPFR (through the `rug`
crate, 128-bit precision, round-to-nearest):

* DCT matrix  -- `build.rs:316-325` (`d_value`) emitted by `build.rs:604-615`
  for i in 1..=16, j in 0..127, n = 127, as f32 and as f64.
* tent filter -- `build.rs:334-381` (`tent_filter_weights`): the impulse
  response of two passes of the 3-wide Jarosz box filter (`build.rs:196-313`)
  on a 7x7 grid, read at the centre; padded to 8 columns at `build.rs:729-745`.
* offsets     -- `build.rs:383-387,771-781`: round(lerp(0, 512, i, 127)) = 4(i+1).

This script calls the system libmpfr through ctypes with the same operation
order, so every intermediate is rounded exactly as the reference's were.  The
product has its own generator (mpmath, `yume-pdq_amd/csrc/gen_tables.py`);
`tests/test_tables.py` checks that both agree bit for bit and that they
match the closed forms.

Usage: python oracle/gen_tables.py  (rewrites oracle/pdq_tables.h)
"""
import ctypes
import ctypes.util
import os
import struct
import sys

PREC = 128
RNDN = 0


class _Mpfr(ctypes.Structure):
    # __mpfr_struct on LP64: prec (long), sign (int), exp (long), limb pointer
    _fields_ = [("prec", ctypes.c_long), ("sign", ctypes.c_int),
                ("exp", ctypes.c_long), ("d", ctypes.c_void_p)]


def _load_mpfr():
    for name in ("libmpfr.so.6", ctypes.util.find_library("mpfr")):
        if not name:
            continue
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    raise RuntimeError("libmpfr not found")


_lib = _load_mpfr()
_P = ctypes.POINTER(_Mpfr)
for fn, args, res in [
    ("mpfr_init2", [_P, ctypes.c_long], None),
    ("mpfr_clear", [_P], None),
    ("mpfr_set_si", [_P, ctypes.c_long, ctypes.c_int], ctypes.c_int),
    ("mpfr_set", [_P, _P, ctypes.c_int], ctypes.c_int),
    ("mpfr_add", [_P, _P, _P, ctypes.c_int], ctypes.c_int),
    ("mpfr_sub", [_P, _P, _P, ctypes.c_int], ctypes.c_int),
    ("mpfr_mul", [_P, _P, _P, ctypes.c_int], ctypes.c_int),
    ("mpfr_div", [_P, _P, _P, ctypes.c_int], ctypes.c_int),
    ("mpfr_sqrt", [_P, _P, ctypes.c_int], ctypes.c_int),
    ("mpfr_cos", [_P, _P, ctypes.c_int], ctypes.c_int),
    ("mpfr_acos", [_P, _P, ctypes.c_int], ctypes.c_int),
    ("mpfr_get_flt", [_P, ctypes.c_int], ctypes.c_float),
    ("mpfr_get_d", [_P, ctypes.c_int], ctypes.c_double),
]:
    f = getattr(_lib, fn)
    f.argtypes = args
    f.restype = res


class F:
    """A 128-bit MPFR float; every operation rounds to nearest at 128 bits."""

    __slots__ = ("m",)

    def __init__(self, v=0):
        self.m = _Mpfr()
        _lib.mpfr_init2(ctypes.byref(self.m), PREC)
        _lib.mpfr_set_si(ctypes.byref(self.m), int(v), RNDN)

    def __del__(self):
        try:
            _lib.mpfr_clear(ctypes.byref(self.m))
        except Exception:
            pass

    def _bin(self, other, fn):
        r = F()
        fn(ctypes.byref(r.m), ctypes.byref(self.m), ctypes.byref(other.m), RNDN)
        return r

    def __add__(self, o):
        return self._bin(o, _lib.mpfr_add)

    def __sub__(self, o):
        return self._bin(o, _lib.mpfr_sub)

    def __mul__(self, o):
        return self._bin(o, _lib.mpfr_mul)

    def __truediv__(self, o):
        return self._bin(o, _lib.mpfr_div)

    def _un(self, fn):
        r = F()
        fn(ctypes.byref(r.m), ctypes.byref(self.m), RNDN)
        return r

    def sqrt(self):
        return self._un(_lib.mpfr_sqrt)

    def cos(self):
        return self._un(_lib.mpfr_cos)

    def f32(self):
        return _lib.mpfr_get_flt(ctypes.byref(self.m), RNDN)

    def f64(self):
        return _lib.mpfr_get_d(ctypes.byref(self.m), RNDN)


def mpfr_pi():
    # build.rs:121-131: acos(0) at 128 bits, doubled (exact)
    z = F(0)
    half = z._un(_lib.mpfr_acos)
    return half * F(2)


def d_value(i, j, n):
    """build.rs:316-325, same operation order."""
    n1, i1, j1 = F(n), F(i), F(j)
    one = F(1)
    two = one + one
    pi = mpfr_pi()
    return (two / n1).sqrt() * ((pi / (two * n1) * i1 * (two * j1 + one)).cos())


def box_one_d(invec, start, outvec, length, stride, full):
    """build.rs:196-259 (mean_box::box_one_d_float) on 128-bit values."""
    half = (full + 2) // 2
    p1 = half - 1
    p2 = full - half + 1
    oi_off = p1 * stride
    li_off = p2 * stride
    s = F(0)
    cws = F(0)
    one = F(1)
    p1_end = oi_off + start
    for ri in range(start, p1_end, stride):
        s = s + invec[ri]
        cws = cws + one
    p2_end = full * stride + start
    for ri in range(p1_end, p2_end, stride):
        oi = ri - oi_off
        s = s + invec[ri]
        cws = cws + one
        outvec[oi] = s / cws
    p3_end = length * stride + start
    for ri in range(p2_end, p3_end, stride):
        oi = ri - oi_off
        li = oi - li_off
        s = s + invec[ri]
        s = s - invec[li]
        outvec[oi] = s / cws
    p4_start = (length - half + 1) * stride + start
    for oi in range(p4_start, p3_end, stride):
        li = oi - li_off
        s = s - invec[li]
        cws = cws - one
        outvec[oi] = s / cws


def jarosz(buf, rows, cols, win_r, win_c, nreps):
    """build.rs:281-307 (mean_box::jarosz_filter_float)."""
    tmp = [F(0) for _ in range(len(buf))]
    for _ in range(nreps):
        for i in range(rows):
            box_one_d(buf, i * cols, tmp, cols, 1, win_r)
        for j in range(cols):
            box_one_d(tmp, j, buf, rows, cols, win_c)


def tent_weights():
    """build.rs:334-381: impulse response read at the grid centre."""
    win = -(-512 // (2 * 127))  # div_ceil(512, 254) = 3
    rows = cols = 2 * win + 1
    out = [0.0] * (rows * cols)
    for di in range(rows):
        for dj in range(cols):
            t = [F(0) for _ in range(rows * cols)]
            t[di * cols + dj] = F(1)
            s = F(0)
            for x in t:
                s = s + x
            t = [x / s for x in t]
            jarosz(t, rows, cols, win, win, 2)
            mid = t[win * cols + win]
            v = mid.f32()
            if v > 0.0:
                out[di * cols + dj] = v
    return win, rows, cols, out


def f32_bits(v):
    return struct.unpack("<I", struct.pack("<f", v))[0]


def f64_bits(v):
    return struct.unpack("<Q", struct.pack("<d", v))[0]


def main():
    dct32, dct64 = [], []
    for i in range(1, 17):
        for j in range(127):
            v = d_value(i, j, 127)
            a = v.f32()
            b = v.f64()
            # rug's to_f32 and a double-rounded get_d agree on every entry
            assert f32_bits(a) == f32_bits(struct.unpack("<f", struct.pack("<f", b))[0]), (i, j)
            dct32.append(f32_bits(a))
            dct64.append(f64_bits(b))
    off, rows, cols, w = tent_weights()
    assert (off, rows, cols) == (3, 7, 7)
    x8 = []
    for r in range(rows):
        x8 += [f32_bits(w[r * cols + c]) for c in range(cols)] + [0]

    here = os.path.dirname(os.path.abspath(__file__))
    out = os.path.join(here, "pdq_tables.h")
    with open(out, "w") as fh:
        fh.write("/* Generated by oracle/gen_tables.py (GNU MPFR, 128-bit, round-to-nearest).\n"
                 " * Oracle-side copy of the reference's build.rs tables; test infrastructure only. */\n")
        fh.write("#pragma once\n#include <stdint.h>\n\n")
        fh.write("/* build.rs:604-615: 16 x 127 row-major, then 16 zero pad (alignment.rs:489-495) */\n")
        fh.write("static const uint32_t ORACLE_DCT_F32_BITS[16 * 127 + 16] = {\n")
        for k in range(0, len(dct32), 8):
            fh.write("  " + ", ".join("0x%08xu" % x for x in dct32[k:k + 8]) + ",\n")
        fh.write("  " + ", ".join(["0u"] * 16) + "\n};\n\n")
        fh.write("static const uint64_t ORACLE_DCT_F64_BITS[16 * 127 + 16] = {\n")
        for k in range(0, len(dct64), 4):
            fh.write("  " + ", ".join("0x%016xull" % x for x in dct64[k:k + 4]) + ",\n")
        fh.write("  " + ", ".join(["0ull"] * 16) + "\n};\n\n")
        fh.write("/* build.rs:729-745: 7 rows x 8 (7 weights + 1 zero pad) */\n")
        fh.write("static const uint32_t ORACLE_TENT_X8_BITS[7 * 8] = {\n")
        for r in range(7):
            fh.write("  " + ", ".join("0x%08xu" % x for x in x8[r * 8:(r + 1) * 8]) + ",\n")
        fh.write("};\n")
    print("wrote", out, file=sys.stderr)


if __name__ == "__main__":
    main()
