"""Writes tests/golden/philox_kat.txt: Philox4x32-10 input/output vectors
produced by the reference's own vendored Random123 philox.h (compiled from
the reference tree by `make -C oracle ref` into oracle/_ref/philox_kat).

Inputs: the published Random123 known-answer inputs, the counter layouts
GO_Fish uses (reference 3P/_internal/shared.cuh:67-82 with keys
{0xbeeff00d, 0xdecafbad}), and seeded random keys/counters.
Run: python tests/golden/make_philox_kat.py   (needs /root/reference)
"""
import os
import subprocess
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(REPO, "oracle", "_ref", "philox_kat")


def inputs():
    rows = [
        (0, 0, 0, 0, 0, 0),
        (0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF),
        (0xa4093822, 0x299f31d0, 0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
        (0xbeeff00d, 0xdecafbad, 2, 1, 0, 0),
        (0xbeeff00d, 0xdecafbad, 1, 7, 0, 0),
        (0xbeeff00d, 0xdecafbad, 0, 123, 2, 0),
    ]
    rng = np.random.default_rng(20240601)
    for _ in range(200):
        k = rng.integers(0, 2**32, size=2, dtype=np.uint64)
        c = rng.integers(0, 2**32, size=4, dtype=np.uint64)
        rows.append(tuple(int(v) for v in (*k, *c)))
    for _ in range(56):  # GO_Fish-shaped counters: small ints
        k = rng.integers(0, 2**32, size=2, dtype=np.uint64)
        c = rng.integers(0, 100000, size=4, dtype=np.uint64)
        rows.append(tuple(int(v) for v in (*k, *c)))
    return rows


def main():
    if not os.path.exists(BIN):
        subprocess.check_call(["make", "-C", os.path.join(REPO, "oracle"), "ref"])
    text = "\n".join(" ".join(str(v) for v in r) for r in inputs()) + "\n"
    out = subprocess.run([BIN], input=text, capture_output=True, text=True, check=True).stdout
    with open(os.path.join(HERE, "philox_kat.txt"), "w") as f:
        f.write("# key0 key1 c0 c1 c2 c3 -> out0 out1 out2 out3 (hex); from reference Random123 philox.h\n")
        f.write(out)


if __name__ == "__main__":
    main()
