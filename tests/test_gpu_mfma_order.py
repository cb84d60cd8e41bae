"""The FP32 MFMA tile path against the oracle's MFMA order, bit for bit.

v_mfma_f32_32x32x2_f32 applies its two products to the accumulator as two
fmaf, the lane-half-0 product first (established on MI355X with crafted
operands), and the FP32 tiles feed MFMA c of each 8-wide k block with k = c and
k = 4 + c.  Every field element is therefore one fmaf chain over the positions
0, 4, 1, 5, 2, 6, 3, 7 of each 8-block (orc_field_f32_mfma): a fixed
permutation of the reference's sequential sum (Bifurcation_m2.cu.c:93-97), the
same for every tile shape, batch shape and replica half."""
import os

import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def ctx():
    c = sb.Context(0)
    yield c
    c.close()


def instance(variant, N, seed):
    # real couplings make the summation order visible (integer J x sign is exact)
    if variant == "dsb":
        return sb.generate(sb.GEN_SK, N, 1.0, seed)
    return sb.generate(sb.GEN_SK_DILUTED, N, 0.5, seed)


@pytest.mark.parametrize("variant", ["bsb", "dsb", "asb", "asb_m1"])
@pytest.mark.parametrize("N,R,steps", [(77, 5, 200), (300, 64, 150), (1000, 130, 60)])
def test_fp32_tiles_bit_exact_vs_mfma_order(ctx, orc, variant, N, R, steps):
    J = instance(variant, N, N + R)
    v = sb.VARIANTS[variant]
    dt = 1.0 if variant == "asb" else 0.5
    prob = ctx.problem(J)
    sol = ctx.solve(prob, state=True, variant=v, precision="fp32", steps=steps, replicas=R, init="uniform", seed=4,
                    dt=dt)
    xo, yo, _ = orc.solve_f32(v, J, R, steps, init="uniform", seed=4, dt=dt, order=orc.ORDER_MFMA, nthreads=THREADS)
    assert np.array_equal(sol.x, xo[:, :N], equal_nan=True)
    assert np.array_equal(sol.y, yo[:, :N], equal_nan=True)
    prob.close()


def test_fp32_field_operator_equals_mfma_order(ctx, orc):
    N, R = 1111, 9
    J = sb.generate(sb.GEN_SK, N, 1.0, 12)
    x, _ = orc.init_state("uniform", N, R, seed=3)
    prob = ctx.problem(J)
    for variant, sign in (("bsb", False), ("dsb", True)):
        F = ctx.field(prob, x[:, :N].copy(), "fp32", variant)
        assert np.array_equal(F, orc.field_f32(J, x, use_sign=sign, order=orc.ORDER_MFMA)[:, :N]), variant
    prob.close()


def test_replica_halves_and_batch_equal_mfma_order(ctx, orc):
    # R = 1024 runs as two concurrent replica halves; the first replicas of
    # each half follow the oracle.  A batched FP32 instance does too.
    N, steps = 2000, 80
    J = sb.generate(sb.GEN_SK, N, 1.0, 2)
    prob = ctx.problem(J)
    sol = ctx.solve(prob, state=True, variant="bsb", precision="fp32", steps=steps, replicas=1024, init="uniform",
                    seed=2, dt=1.0)
    pick = [0, 1, 511, 512, 513, 1023]
    x, y = orc.init_state("uniform", N, 1024, seed=2)
    xs, ys = np.ascontiguousarray(x[pick]), np.ascontiguousarray(y[pick])
    c0 = orc.c0_auto(orc.BSB, J)
    # the keyed init depends on (replica, spin) only, so the picked rows
    # advance like replicas 0, 1, 511, ... of the full solve
    orc.run_f32(orc.BSB, J, xs, ys, 0, steps, steps, 1.0, 1.0, c0, nthreads=THREADS, order=orc.ORDER_MFMA)
    assert np.array_equal(sol.x[pick], xs[:, :N]) and np.array_equal(sol.y[pick], ys[:, :N])
    b = ctx.problem(sb.generate(sb.GEN_SK, 700, 1.0, 7))
    params = [sb.make_params(variant="bsb", precision="fp32", steps=steps, replicas=R, init="uniform", seed=1, dt=0.5)
              for R in (96, 40)]
    sols, _ = ctx.batch_solve([prob, b], params, state=True)
    Jb = b.read_j()
    xo, yo, _ = orc.solve_f32(orc.BSB, Jb, 40, steps, init="uniform", seed=1, dt=0.5, order=orc.ORDER_MFMA,
                              nthreads=THREADS)
    assert np.array_equal(sols[1].x, xo[:, :700]) and np.array_equal(sols[1].y, yo[:, :700])
    prob.close()
    b.close()
