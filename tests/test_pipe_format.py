"""CPU: the pipe output formats (host-only formatter in the library, no GPU
call), pinned on the reference's own documented outputs:

* `pipe -f q+hex+lf` of test-data/aaa-orig.jpg prints
  "100.000:58f8f0cee0f4a84f06370a32038f67f0b36e2ed596623e1d33e6b39c4e9c9b22"
  (src/main.rs:148-154), and `hex+lf` prints the d8f8f0ce... hash of
  test-data/aaa-orig.jpg.expect:1 per line (src/main.rs:163-169);
* fnr-test/compliance.sh:44-52 compares `pipe -f bin` with the reference hash
  through `xxd -r -p | xxd -b -c 1`: the 32 bytes in order, each most
  significant bit first;
* the separator is the last '+' component (src/main.rs:1106-1133).
"""
import os
import struct

import numpy as np
import pytest

import yume_pdq

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DOC_Q_HEX = "100.000:58f8f0cee0f4a84f06370a32038f67f0b36e2ed596623e1d33e6b39c4e9c9b22"


def expect_hashes():
    return [line.strip() for line in open(os.path.join(GOLDEN, "aaa_orig_expect.txt")) if line.strip()]


def frames_of(hex_hashes, quality=1.0):
    n = len(hex_hashes)
    h = np.zeros((n, 8, 32), np.uint8)
    for i, x in enumerate(hex_hashes):
        h[i, :] = np.frombuffer(bytes.fromhex(x), np.uint8)
    q = np.full((n, 8), quality, np.float32)
    t = np.zeros((n, 8), np.float32)
    return h, q, t


def test_documented_q_hex_line():
    qual, hx = DOC_Q_HEX.split(":")
    h, q, t = frames_of([hx], quality=float(qual) / 100.0)
    assert yume_pdq.format_frames("q+hex+lf", h, q, t) == (DOC_Q_HEX + "\n").encode()


def test_documented_hex_lines():
    exp = expect_hashes()
    h, q, t = frames_of([exp[0]] * 5)
    assert yume_pdq.format_frames("hex+lf", h, q, t) == ((exp[0] + "\n") * 5).encode()
    assert yume_pdq.format_frames("HEX", h, q, t) == (exp[0].upper() * 5).encode()


def test_bin_matches_compliance_script_bit_order():
    exp = expect_hashes()
    h, q, t = frames_of(exp)
    got = yume_pdq.format_frames("bin", h, q, t).decode()
    # xxd -b -c 1 of the hash bytes, concatenated
    want = "".join("".join(format(b, "08b") for b in bytes.fromhex(x)) for x in exp)
    assert got == want
    assert yume_pdq.format_frames("BIN+nul", h, q, t).decode() == "".join(
        "".join(format(b, "08b") for b in bytes.fromhex(x)) + "\0" for x in exp)


def test_raw_and_quality_prefixed_raw():
    exp = expect_hashes()
    h, q, t = frames_of(exp, quality=0.75)
    assert yume_pdq.format_frames("raw", h, q, t) == b"".join(bytes.fromhex(x) for x in exp)
    rec = struct.pack("<f", 0.75)
    assert yume_pdq.format_frames("q+RAW+crlf", h, q, t) == b"".join(
        rec + bytes.fromhex(x) + b"\r\n" for x in exp)


@pytest.mark.parametrize("sep,bytes_", [("nl", b"\n"), ("NL", b"\n"), ("lf", b"\n"), ("LF", b"\n"),
                                        ("crlf", b"\r\n"), ("CRLF", b"\r\n"), ("cr", b"\r"), ("CR", b"\r"),
                                        ("nul", b"\0"), ("NUL", b"\0"), ("null", b"\0"), ("NULL", b"\0")])
def test_separators(sep, bytes_):
    exp = expect_hashes()
    h, q, t = frames_of(exp[:2])
    assert yume_pdq.format_frames("hex+" + sep, h, q, t) == b"".join(x.encode() + bytes_ for x in exp[:2])


@pytest.mark.parametrize("bad", ["", "lf", "hex+", "hex+foo", "q+", "q+q+hex", "Hex", "q+Bin+lf+lf", "diag"])
def test_invalid_formats(bad):
    h, q, t = frames_of(expect_hashes()[:1])
    with pytest.raises(yume_pdq.YumePdqError):
        yume_pdq.format_frames(bad, h, q, t)


def test_dihedral_records_use_each_orientation():
    """With dihedrals, eight records per frame in the stored orientation
    order, each with its own quality (visit_dihedrals passes the new quality,
    src/main.rs:806-812)."""
    rng = np.random.default_rng(3)
    h = rng.integers(0, 256, size=(3, 8, 32), dtype=np.uint8)
    q = rng.uniform(0, 1, size=(3, 8)).astype(np.float32)
    t = rng.normal(size=(3, 8)).astype(np.float32)
    got = yume_pdq.format_frames("q+hex+lf", h, q, t, dihedrals=True).decode().splitlines()
    assert len(got) == 24
    for i in range(3):
        for d in range(8):
            qs, hx = got[8 * i + d].split(":")
            assert hx == bytes(h[i, d]).hex()
            assert abs(float(qs) - 100.0 * float(q[i, d])) <= 5e-4 + 1e-6
    one = yume_pdq.format_frames("hex", h, q, t).decode()
    assert one == "".join(bytes(h[i, 0]).hex() for i in range(3))


def test_quality_prefix_rounding():
    """{0:02.3} of the f32 product quality * 100: three decimals of the exact
    value (ties to even), no padding."""
    h, _, t = frames_of(expect_hashes()[:1])
    for qv, want in [(1.0, "100.000"), (0.5, "50.000"), (0.0, "0.000"), (0.001, "0.100"),
                     (0.123456, "12.346"), (float("nan"), "NaN")]:
        q = np.full((1, 8), qv, np.float32)
        line = yume_pdq.format_frames("q+bin", h, q, t).decode()
        assert line.split(":")[0] == want


def test_diagnostic_layout():
    """src/main.rs:738-770: name, index, quality, threshold ({:.7}), the hash
    as 256 bits, the PDQF with '*' on values above the original threshold."""
    rng = np.random.default_rng(5)
    h = rng.integers(0, 256, size=(1, 8, 32), dtype=np.uint8)
    q = np.full((1, 8), 0.25, np.float32)
    t = np.array([[1.5, -2.0, 0, 0, 0, 0, 0, 0.125]], np.float32)
    p = rng.normal(size=(1, 8, 16, 16)).astype(np.float32)
    text = yume_pdq.format_frames("diagnostic", h, q, t, dihedrals=True, dihedral_pdqf=p).decode()
    blocks = text.split("[")[1:]
    assert len(blocks) == 8
    first = ("[" + blocks[0]).splitlines()
    assert first[0] == '["original"]'
    assert first[1] == "Index: 0"
    assert first[2] == "Quality: 25.000"
    assert first[3] == "Threshold: 1.5000000"
    assert first[4] == "Hash:"
    assert first[5] == "".join(format(b, "08b") for b in h[0, 0])
    assert first[6] == "PDQF:"
    row0 = first[7].split("\t")
    assert len(row0) == 16
    for j, cell in enumerate(row0):
        v = float(p[0, 0, 0, j])
        assert cell == ("*" if v > 1.5 else "") + f"{v:.5f}"
    last = ("[" + blocks[7]).splitlines()
    assert last[0] == "[Dihedrals { packed: %d }]" % 0x0001FF00  # ROTATED_90
    assert last[1] == "Index: 7"
    assert last[3] == "Threshold: 0.1250000"
    # marks use the original orientation's threshold, not this one's
    cells = [c for line in last[7:23] for c in line.split("\t")]
    assert sum(c.startswith("*") for c in cells) == int((p[0, 7] > 1.5).sum())
    with pytest.raises(yume_pdq.YumePdqError):
        yume_pdq.format_frames("diagnostic", h, q, t)  # the PDQFs are required


def test_bound_covers_output():
    rng = np.random.default_rng(9)
    h = rng.integers(0, 256, size=(4, 8, 32), dtype=np.uint8)
    q = rng.uniform(0, 1, size=(4, 8)).astype(np.float32)
    t = (rng.normal(size=(4, 8)) * 1e30).astype(np.float32)
    p = rng.uniform(-3e38, 3e38, size=(4, 8, 256)).astype(np.float32)
    lib = yume_pdq.lib()
    for fmt in ["raw", "q+raw", "hex+crlf", "q+HEX", "bin+lf", "q+bin+nul", "diagnostic+lf"]:
        for dih in (False, True):
            out = yume_pdq.format_frames(fmt, h, q, t, dihedrals=dih, dihedral_pdqf=p)
            assert len(out) <= lib.yume_pdq_amd_format_bound(fmt.encode(), 4, int(dih))
