"""Cuts the real-signal fixtures from the reference's own 100 ms captures
(sample/readme.txt): the first 30 ms of L1_20211125_004000_12MHz_I.bin
(12 MHz, I-sampled int8, IF 3 MHz: GPS L1 C/A, L1C-P and Galileo E1-B) and
the first 14 ms of L6_20211125_004000_12MHz_IQ.bin (12 MHz, IQ-sampled int8
pairs, IF 0: QZSS L6D).  Data only: the bytes are copied as they are, and
slices.json records their source, offset, length and SHA-256.

Run in a tree where the reference is mounted:
    python3 tests/golden/make_sample_slices.py
"""
import hashlib
import json
import os

REF = "/root/reference/sample"
HERE = os.path.dirname(os.path.abspath(__file__))
SLICES = [("L1_20211125_004000_12MHz_I.bin", 0, 360000, "L1_20211125_004000_12MHz_I_30ms.bin",
           dict(fs=12e6, fi=3e6, IQ=1)),
          ("L6_20211125_004000_12MHz_IQ.bin", 0, 336000, "L6_20211125_004000_12MHz_IQ_14ms.bin",
           dict(fs=12e6, fi=0.0, IQ=2))]


def main():
    meta = {}
    for src, off, n, dst, info in SLICES:
        with open(os.path.join(REF, src), "rb") as f:
            f.seek(off)
            data = f.read(n)
        assert len(data) == n
        with open(os.path.join(HERE, dst), "wb") as f:
            f.write(data)
        meta[dst] = dict(source="sample/" + src, offset=off, bytes=n,
                         sha256=hashlib.sha256(data).hexdigest(), **info)
    with open(os.path.join(HERE, "slices.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
