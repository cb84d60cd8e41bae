"""Generate tests/golden/split_field_cases.npz on an MI355X: states along the
oracle's split-form trajectories of one SK instance (N = 77, 40 replicas,
dt = 0.5), and the split field the GPU computes on each (sb_field through the
v_mfma_f32_32x32x16_bf16 tiles); one replica of each state is kept.  The
states are those where an earlier form of the oracle's model (the 32-bit
window of a sum taken relative to max(E_p, e(acc)) instead of the sum's own
exponent) first disagreed with the hardware: fields whose MFMA sums cross a
binade while the low products are added.  They came from that earlier
model's trajectories, so rerunning this script with the current oracle
reaches other states.

    python tests/golden/make_split_field_cases.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from annealing_amd import sb  # noqa: E402
from oracle import oracle as O  # noqa: E402

# (variant, schedule length, state index after that many steps, replica kept)
PICKS = [("bsb", 64, 4, 29), ("bsb", 64, 5, 24), ("bsb", 64, 18, 2), ("asb_m1", 64, 53, 20),
         ("asb_m1", 64, 54, 29), ("asb_m1", 120, 30, 16)]


def main():
    N, R = 77, 40
    J = sb.generate(sb.GEN_SK, N, 1.0, N + R)
    X, F = [], []
    with sb.Context(0) as ctx:
        prob = ctx.problem(J)
        for variant, steps, k, r in PICKS:
            v = sb.VARIANTS[variant]
            x, y = O.init_state("uniform", N, R, seed=3)
            O.run_f32(v, J, x, y, 0, k, steps, 0.5, 1.0, O.c0_auto(v, J), nthreads=8, order=O.ORDER_SPLIT)
            xs = np.ascontiguousarray(x[:, :N])
            X.append(xs[r])
            F.append(ctx.field(prob, xs, precision="split", variant="bsb")[r])
        prob.close()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "split_field_cases.npz")
    np.savez_compressed(out, J=J, X=np.stack(X), F=np.stack(F))
    print(out, len(PICKS), "states")


if __name__ == "__main__":
    main()
