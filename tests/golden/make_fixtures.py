#!/usr/bin/env python3
"""Writes the reference-held golden vectors for the hash path as small fixtures.

Run in a container where the reference checkout is mounted (path below); the
fixtures are committed so nothing at test time reads the reference.

* aaa_orig_512_u8.npy -- test-data/aaa-orig.jpg preprocessed exactly as the
  reference's FFI harness does (integration/hash.py:65-69: PIL convert('L'),
  resize((512, 512), BILINEAR)); stored as the u8 plane.
* aaa_orig_expect.txt -- the CI whitelist test-data/aaa-orig.jpg.expect:1-3
  (hex of the 32 output bytes; build.yml:96-128 accepts any of them).
* readme_fm.json      -- README.md:305-452 doctest: the 32 expected bytes
  (README.md:432-436, tolerance <= 2 bits, :447-451) and the PDQF printed to
  0.1 precision (README.md:397-412).
"""
import json
import os
import re
import sys

import numpy as np
from PIL import Image

REF = os.environ.get("YUME_PDQ_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    img = Image.open(os.path.join(REF, "test-data/aaa-orig.jpg")).convert("L")
    img = img.resize((512, 512), Image.Resampling.BILINEAR)
    arr = np.array(img, dtype=np.uint8)
    np.save(os.path.join(HERE, "aaa_orig_512_u8.npy"), arr)

    with open(os.path.join(REF, "test-data/aaa-orig.jpg.expect")) as fh:
        expect = [ln.strip() for ln in fh if ln.strip()]
    with open(os.path.join(HERE, "aaa_orig_expect.txt"), "w") as fh:
        fh.write("\n".join(expect) + "\n")

    readme = open(os.path.join(REF, "README.md")).read()
    m = re.search(r"let expected = \[(.*?)\];", readme, re.S)
    expected = [int(x, 16) for x in re.findall(r"0x([0-9a-fA-F]{2})", m.group(1))]
    assert len(expected) == 32
    m = re.search(r"You should see:\n(.*?)\*/", readme, re.S)
    rows = [[float(v) for v in ln.split()] for ln in m.group(1).strip().splitlines()]
    assert len(rows) == 16 and all(len(r) == 16 for r in rows)
    with open(os.path.join(HERE, "readme_fm.json"), "w") as fh:
        json.dump({"source": "README.md:305-452", "expected_bytes": expected,
                   "max_bits_different": 2, "pdqf_printed_1dp": rows}, fh, indent=1)
    print("fixtures written", file=sys.stderr)


if __name__ == "__main__":
    main()
