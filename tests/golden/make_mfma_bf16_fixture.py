"""Convert the dump of tests/golden/mfma_bf16_probe.hip into the fixture
mfma_bf16_cases.npz: for a fixed sample of outputs (every 8th column of every
batch), the accumulator c, the 16 BF16 operand pairs (as bit patterns) and the
hardware result d.

    python tests/golden/make_mfma_bf16_fixture.py mfma_bf16_cases.bin
"""
import os
import struct
import sys

import numpy as np


def main(path):
    b = open(path, "rb").read()
    nb = struct.unpack("<i", b[:4])[0]
    o = 4
    A = np.frombuffer(b, np.uint16, nb * 512, o).reshape(nb, 32, 16)
    o += nb * 512 * 2
    B = np.frombuffer(b, np.uint16, nb * 512, o).reshape(nb, 16, 32)
    o += nb * 512 * 2
    C = np.frombuffer(b, np.float32, nb * 1024, o).reshape(nb, 32, 32)
    o += nb * 1024 * 4
    D = np.frombuffer(b, np.float32, nb * 1024, o).reshape(nb, 32, 32)
    t, m, n = np.meshgrid(np.arange(nb), np.arange(32), np.arange(0, 32, 8), indexing="ij")
    t, m, n = t.ravel(), m.ravel(), n.ravel()
    a = A[t, m, :]                      # [cases, 16]
    bb = B[t, :, n]                     # [cases, 16]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mfma_bf16_cases.npz")
    np.savez_compressed(out, c=C[t, m, n], a_bits=a, b_bits=bb, d=D[t, m, n], batch_type=(t % 5).astype(np.int8))
    print(out, len(t), "cases")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "mfma_bf16_cases.bin")
