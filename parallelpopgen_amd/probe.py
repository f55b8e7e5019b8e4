"""Counter-calibration probe: runs the library's plain copy kernel
(gf_device_math op COPY) over a buffer of known size, so that rocprofv3
FETCH_SIZE / WRITE_SIZE readings of other kernels can be scaled to bytes.

    python -m parallelpopgen_amd.probe --mib 512
reads and writes MiB mebibytes once (one kernel dispatch)."""
import argparse

import numpy as np

from . import capi


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=512)
    a = ap.parse_args()
    n = a.mib * (1 << 20) // 4
    x = np.arange(n, dtype=np.float32)
    y = capi.device_math("copy", x)
    assert np.array_equal(x, y)
    print("copy probe: %d bytes read, %d bytes written" % (4 * n, 4 * n))


if __name__ == "__main__":
    main()
