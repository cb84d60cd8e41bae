"""Energies and cut values read back by sb_solve (the reference's readout and
energy loop, SB/CUDA_GPU/m2_opt.cu:189-202, H = -sum_{i<j} J_ij s_i s_j -
sum_i h_i s_i per README.md:11).  Integer J with |J| <= 127 runs the exact
integer reduction (int8 MFMA, int32 accumulators, int64 sums); other J the
FP64 form.  The cut value of the max-cut mapping W = -J is checked against
its definition, the sum of W_ij over the pairs split by the spins."""
import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sb.Context(0)
    yield c
    c.close()


def sym_int(N, lo, hi, seed, density=1.0):
    rng = np.random.default_rng(seed)
    J = rng.integers(lo, hi + 1, (N, N)).astype(np.float32)
    if density < 1.0:
        J *= rng.random((N, N)) < density
    J = np.triu(J, 1)
    return (J + J.T).astype(np.float32)


def brute_cut(J, s):
    """Sum of W_ij = -J_ij over i < j with s_i != s_j (int64 or float64)."""
    W = -np.triu(J.astype(np.float64), 1)
    out = np.zeros(s.shape[0])
    for r in range(s.shape[0]):
        split = s[r][:, None] != s[r][None, :]
        out[r] = W[split].sum()
    return out


@pytest.mark.parametrize("N,R,lo,hi", [(77, 5, -127, 127), (640, 130, -1, 1), (1000, 64, -100, 100),
                                       (129, 257, -16, 16)])
def test_integer_energy_exact_and_cut(ctx, orc, N, R, lo, hi):
    J = sym_int(N, lo, hi, N + R)
    prob = ctx.problem(J)
    for prec in ("fp32", "split"):
        sol = ctx.solve(prob, variant="bsb", precision=prec, steps=40, replicas=R, init="uniform", seed=3, dt=0.5)
        assert sol.exact
        H = orc.energy_i64(J, sol.spins)
        assert np.array_equal(sol.energy_i64, H)
        assert np.array_equal(sol.energy, H.astype(np.float64))
        assert np.array_equal(sol.cut, brute_cut(J, sol.spins))
    prob.close()


def test_integer_beyond_int8_uses_fp64_and_stays_exact(ctx, orc):
    J = sym_int(300, -1000, 1000, 7)
    prob = ctx.problem(J)
    sol = ctx.solve(prob, variant="dsb", precision="fp32", steps=30, replicas=40, init="uniform", seed=1, dt=0.5)
    assert sol.exact
    assert np.array_equal(sol.energy_i64, orc.energy_i64(J, sol.spins))
    assert np.array_equal(sol.cut, brute_cut(J, sol.spins))
    prob.close()


def test_energy_with_integer_and_real_field(ctx, orc):
    N, R = 333, 70
    J = sym_int(N, -3, 3, 11, density=0.3)
    rng = np.random.default_rng(12)
    for h, exact in ((rng.integers(-5, 6, N).astype(np.float32), True),
                     (rng.normal(0, 1, N).astype(np.float32), False)):
        prob = ctx.problem(J, h=h)
        sol = ctx.solve(prob, variant="bsb", precision="fp32", steps=50, replicas=R, init="uniform", seed=2,
                        dt=0.5)
        assert sol.exact == exact
        if exact:
            assert np.array_equal(sol.energy_i64, orc.energy_i64(J, sol.spins, h=h))
        else:
            assert np.allclose(sol.energy, orc.energy_f64(J, sol.spins, h=h), rtol=1e-12, atol=1e-9)
        # the cut ignores the field
        assert np.array_equal(sol.cut, brute_cut(J, sol.spins))
        prob.close()


def test_real_j_energy_and_cut(ctx, orc):
    J = sb.generate(sb.GEN_SK, 500, 1.0, 4)
    prob = ctx.problem(J)
    sol = ctx.solve(prob, variant="bsb", precision="split", steps=40, replicas=66, init="uniform", seed=1, dt=0.5)
    assert not sol.exact
    assert np.allclose(sol.energy, orc.energy_f64(J, sol.spins), rtol=1e-12, atol=1e-10)
    assert np.allclose(sol.cut, brute_cut(J, sol.spins), rtol=1e-12, atol=1e-9)
    prob.close()


def test_device_generated_instance_cut(ctx):
    # sb_problem_generate sums J right of the diagonal on the device.
    prob = ctx.generate(sb.GEN_ER, 900, 0.06, 5)
    J = prob.read_j()
    sol = ctx.solve(prob, variant="bsb", precision="fp32", steps=60, replicas=9, init="uniform", seed=1, dt=1.25)
    assert np.array_equal(sol.cut, brute_cut(J, sol.spins))
    prob.close()
