"""Every compiled tile shape of every kernel family gives the same results
(forced with the SB_SHAPE / SB_BATCH_SHAPE debug overrides).  Sizes are chosen
so that row and replica tiles are partial."""
import os

import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu

# Compiled shapes per family (csrc/sb_kernels.hip: SB_*_SHAPES, SB_BATCH_*).
SHAPES = {0: 13, 1: 11, 2: 11, 3: 12}
BATCH_SHAPES = {0: 5, 1: 3, 2: 3, 3: 3}
N, R, STEPS = 300, 40, 30


def solve_forced(ctx, prob, family, shape, **kw):
    os.environ["SB_SHAPE"] = f"{family}:{shape}"
    try:
        return ctx.solve(prob, state=True, **kw)
    finally:
        del os.environ["SB_SHAPE"]


@pytest.fixture(scope="module")
def ctx():
    c = sb.Context(0)
    yield c
    c.close()


def test_dsb_every_shape_matches_oracle(ctx, orc):
    J = sb.generate(sb.GEN_KN, N, 1.0, 31)
    prob = ctx.problem(J)
    xo, yo, _ = orc.solve_f32(orc.DSB, J, R, STEPS, init="uniform", seed=2, dt=0.5)
    xo, yo = xo[:, :N], yo[:, :N]
    for fam, prec in ((0, "fp32"), (1, "split"), (2, "split"), (3, "split")):
        for s in range(SHAPES[fam]):
            sol = solve_forced(ctx, prob, fam, s, variant="dsb", precision=prec, steps=STEPS, replicas=R,
                               init="uniform", seed=2, dt=0.5)
            bad = np.where(~(np.all(sol.x == xo, axis=1) & np.all(sol.y == yo, axis=1)))[0]
            assert bad.size == 0, f"family {fam} shape {s}: replicas {bad.tolist()}"


@pytest.mark.parametrize("variant", ["bsb", "asb"])
def test_every_shape_same_result(ctx, variant):
    J = sb.generate(sb.GEN_KN, N, 1.0, 32)
    prob = ctx.problem(J)
    dt = 1.0 if variant == "asb" else 0.5
    for prec, fams in (("fp32", (0,)), ("split", (1, 2))):
        ref = None
        for fam in fams:
            for s in range(SHAPES[fam]):
                sol = solve_forced(ctx, prob, fam, s, variant=variant, precision=prec, steps=STEPS, replicas=R,
                                   init="uniform", seed=3, dt=dt)
                if ref is None:
                    ref = sol
                    continue
                assert np.array_equal(sol.x, ref.x, equal_nan=True), f"{prec} family {fam} shape {s}"
                assert np.array_equal(sol.y, ref.y, equal_nan=True), f"{prec} family {fam} shape {s}"


@pytest.mark.parametrize("precision,variant,kind", [("fp32", "bsb", sb.GEN_KN), ("split", "bsb", sb.GEN_KN),
                                                    ("split", "dsb", sb.GEN_KN), ("split", "bsb", sb.GEN_SK),
                                                    ("split", "dsb", sb.GEN_SK)])
def test_every_batch_shape_equals_solo(ctx, precision, variant, kind):
    # K_N (integer J) runs the FP8-J families, SK (real J) the BF16-J family.
    Js = [sb.generate(kind, n, 1.0, 40 + n) for n in (N, 77)]
    probs = [ctx.problem(J) for J in Js]
    params = [sb.make_params(variant=variant, precision=precision, steps=STEPS, replicas=r, init="uniform", seed=5,
                             dt=0.5) for r in (R, 9)]
    solos = [ctx.solve(p, state=True, variant=variant, precision=precision, steps=STEPS, replicas=r,
                       init="uniform", seed=5, dt=0.5) for p, r in zip(probs, (R, 9))]
    if precision == "fp32":
        fam = 0
    elif kind == sb.GEN_SK:
        fam = 1
    else:
        fam = 3 if variant == "dsb" else 2
    for b in range(BATCH_SHAPES[fam]):
        os.environ["SB_BATCH_SHAPE"] = str(b)
        try:
            sols, _ = ctx.batch_solve(probs, params, state=True)
        finally:
            del os.environ["SB_BATCH_SHAPE"]
        for i in range(2):
            assert np.array_equal(sols[i].x, solos[i].x), f"batch shape {b} instance {i}"
            assert np.array_equal(sols[i].y, solos[i].y), f"batch shape {b} instance {i}"
