"""Sparse family (CSR J, spin-major state): the field is the sequential fmaf
chain over the nonzeros in ascending column order, so a sparse solve equals the
oracle's sequential (SEQ) order bit for bit, like SB_PREC_ORDERED."""
import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sb.Context(0)
    yield c
    c.close()


def sparse_problem(ctx, J, h=None):
    p = ctx.problem(J, h)
    p.set_sparse(True)
    return p


@pytest.mark.parametrize("variant", ["bsb", "dsb", "asb", "asb_m1"])
@pytest.mark.parametrize("N,R", [(300, 1), (517, 40), (700, 130)])
def test_sparse_bit_exact_vs_sequential_oracle(ctx, orc, variant, N, R):
    J = sb.generate(sb.GEN_ER, N, 0.04, 70 + N) if variant != "bsb" else sb.generate(sb.GEN_SK_DILUTED, N, 0.04,
                                                                                         70 + N)
    v = sb.VARIANTS[variant]
    dt = 0.5  # bounded trajectories: skipped zero terms differ from 0 * NaN only after a blow-up
    steps = 60
    prob = sparse_problem(ctx, J)
    assert prob.nnz == int(np.count_nonzero(J))
    for prec in ("fp32", "split"):
        sol = ctx.solve(prob, state=True, variant=v, precision=prec, steps=steps, replicas=R, init="uniform",
                        seed=9, dt=dt)
        xo, yo, _ = orc.solve_f32(v, J, R, steps, init="uniform", seed=9, dt=dt)
        assert np.array_equal(sol.x, xo[:, :N], equal_nan=True), (prec, np.nanmax(np.abs(sol.x - xo[:, :N])))
        assert np.array_equal(sol.y, yo[:, :N], equal_nan=True), prec
    prob.close()


def test_sparse_with_field_and_energies(ctx, orc):
    N, R, steps = 640, 24, 50
    J = sb.generate(sb.GEN_ER, N, 0.03, 81)
    h = np.random.default_rng(82).integers(-2, 3, N).astype(np.float32)
    prob = sparse_problem(ctx, J, h)
    sol = ctx.solve(prob, state=True, variant="bsb", precision="fp32", steps=steps, replicas=R, init="uniform",
                    seed=3, dt=0.5)
    xo, yo, _ = orc.solve_f32(sb.BSB, J, R, steps, init="uniform", seed=3, dt=0.5, h=h)
    assert np.array_equal(sol.x, xo[:, :N]) and np.array_equal(sol.y, yo[:, :N])
    s = orc.spins(np.ascontiguousarray(xo), N)
    assert sol.exact and np.array_equal(sol.energy_i64, orc.energy_i64(J, s, h))
    # the sparse form equals the ordered dense path; turning it off goes back to MFMA tiles
    o = ctx.solve(prob, state=True, variant="bsb", precision="ordered", steps=steps, replicas=R, init="uniform",
                  seed=3, dt=0.5)
    assert np.array_equal(o.x, sol.x)
    prob.set_sparse(False)
    assert prob.nnz == 0
    d = ctx.solve(prob, variant="bsb", precision="fp32", steps=steps, replicas=R, init="uniform", seed=3, dt=0.5)
    assert d.exact and abs(d.energy.mean() - sol.energy.mean()) <= 0.05 * abs(sol.energy.mean())
    prob.close()


def test_sparse_field_operator(ctx, orc):
    N, R = 900, 7
    J = sb.generate(sb.GEN_SK_DILUTED, N, 0.02, 83)
    x, _ = orc.init_state("uniform", N, R, seed=5)
    prob = sparse_problem(ctx, J)
    for variant, sign in (("bsb", False), ("dsb", True)):
        F = ctx.field(prob, x[:, :N].copy(), "fp32", variant)
        assert np.array_equal(F, orc.field_f32(J, x, use_sign=sign)[:, :N]), variant
    prob.close()


def test_sparse_problem_rejected_by_batch(ctx):
    prob = sparse_problem(ctx, sb.generate(sb.GEN_ER, 200, 0.05, 84))
    params = [sb.make_params(variant="bsb", precision="fp32", steps=10, replicas=16)]
    with pytest.raises(sb.SBError):
        ctx.batch_solve([prob], params)
    prob.close()
