"""The full configuration sizes (K2000 +-1 at R=4096, SK N=2000 at R=1024)
through properties that do not depend on size.  The oracle runs only a few
replicas at this N; a replica's trajectory does not depend on R (keyed init,
a summation order fixed per kernel family), so the first replicas of a full run
are compared with an oracle run of those replicas alone, and every replica is
checked by invariants: |x| <= 1 after bSB/dSB steps, spins = sign(x) with
sign(0) = -1, and exact energies recomputed from the returned spins."""
import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu

N = 2000


@pytest.fixture(scope="module")
def ctx():
    c = sb.Context(0)
    yield c
    c.close()


def check_readout(orc, J, sol, sample, exact):
    x = sol.x
    assert np.array_equal(sol.spins, np.where(x > 0, 1, -1).astype(np.int8))
    s = np.ascontiguousarray(sol.spins[sample])
    if exact:
        assert sol.exact and np.array_equal(sol.energy_i64[sample], orc.energy_i64(J, s))
    else:
        ref = orc.energy_f64(J, s)
        assert np.allclose(sol.energy[sample], ref, rtol=1e-12, atol=1e-9)
    assert sol.energy[sol.best] == sol.energy.min()


def test_k2000_r4096_dsb_every_path_equals_oracle(ctx, orc):
    # Integer J: every partial sum of the field is an exact integer, so the FP8
    # dSB kernel, the FP32 kernel and the oracle agree bit for bit at full size.
    J = sb.generate(sb.GEN_KN, N, 1.0, 3)
    prob = ctx.problem(J)
    R, steps = 4096, 40
    kw = dict(variant="dsb", steps=steps, replicas=R, init="uniform", seed=3, dt=1.0)
    split = ctx.solve(prob, state=True, precision="split", **kw)
    fp32 = ctx.solve(prob, state=True, precision="fp32", **kw)
    assert np.array_equal(split.x, fp32.x) and np.array_equal(split.y, fp32.y)
    r = 6
    xo, yo, _ = orc.solve_f32(sb.DSB, J, r, steps, init="uniform", seed=3, dt=1.0, nthreads=8)
    assert np.array_equal(split.x[:r], xo[:, :N]) and np.array_equal(split.y[:r], yo[:, :N])
    assert np.all(np.abs(split.x) <= 1.0)
    sample = np.random.default_rng(0).choice(R, 24, replace=False)
    check_readout(orc, J, split, sample, exact=True)
    prob.close()


def test_k2000_r4096_bsb_split_invariants(ctx, orc):
    J = sb.generate(sb.GEN_KN, N, 1.0, 3)
    prob = ctx.problem(J)
    R, steps = 4096, 120
    sol = ctx.solve(prob, state=True, variant="bsb", precision="split", steps=steps, replicas=R, init="uniform",
                    seed=3, dt=1.0)
    assert np.all(np.isfinite(sol.x)) and np.all(np.isfinite(sol.y))
    assert np.all(np.abs(sol.x) <= 1.0)
    sample = np.random.default_rng(1).choice(R, 24, replace=False)
    check_readout(orc, J, sol, sample, exact=True)
    # the same replicas in a smaller solve follow the same trajectories
    small = ctx.solve(prob, state=True, variant="bsb", precision="split", steps=steps, replicas=96,
                      init="uniform", seed=3, dt=1.0)
    assert np.array_equal(small.x, sol.x[:96]) and np.array_equal(small.y, sol.y[:96])
    prob.close()


def test_sk2000_r1024_fp32_one_step_bound(ctx, orc):
    # The headline workload's kernel, one step at full size: the FP32 MFMA
    # field differs from the oracle's sequential sum by at most the FP32
    # summation bound, which the step scales by beta.
    J = sb.generate(sb.GEN_SK, N, 1.0, 2)
    prob = ctx.problem(J)
    R, dt = 1024, 1.0
    sol = ctx.solve(prob, state=True, variant="bsb", precision="fp32", steps=1, replicas=R, init="uniform", seed=2,
                    dt=dt)
    xo, yo, c0 = orc.solve_f32(sb.BSB, J, R, 1, init="uniform", seed=2, dt=dt, nthreads=8)
    x0, _ = orc.init_state("uniform", N, R, seed=2)
    absF = np.abs(x0[:, :N]).astype(np.float64) @ np.abs(J).T.astype(np.float64)
    beta = np.float32(dt * c0)
    tol = float(beta) * 2 * N * 2.0**-24 * absF + 4 * np.spacing(np.abs(yo[:, :N])) + 1e-30
    assert np.all(np.abs(sol.y.astype(np.float64) - yo[:, :N]) <= tol)
    prob.close()


def test_sk2000_r1024_fp32_readout(ctx, orc):
    J = sb.generate(sb.GEN_SK, N, 1.0, 2)
    prob = ctx.problem(J)
    sol = ctx.solve(prob, state=True, variant="bsb", precision="fp32", steps=200, replicas=1024, init="uniform",
                    seed=2, dt=1.0)
    assert np.all(np.abs(sol.x) <= 1.0)
    sample = np.random.default_rng(2).choice(1024, 24, replace=False)
    check_readout(orc, J, sol, sample, exact=False)
    prob.close()


@pytest.mark.parametrize("R", [1536, 2048, 2049, 3000])
def test_dsb_replica_halves_match_single_stream_paths(ctx, orc, R):
    # dSB on FP8 J from R = 1536 up runs as two replica halves on two streams
    # (here 768 + 768, 1024 + 1024, 1024 + 1025, 1408 + 1592); integer J makes every path
    # exact, so it must equal the FP32 path bit for bit and the oracle on the
    # replicas at both ends.
    n = 300
    J = sb.generate(sb.GEN_KN, n, 1.0, 41)
    prob = ctx.problem(J)
    kw = dict(variant="dsb", steps=30, replicas=R, init="uniform", seed=5, dt=0.5)
    split = ctx.solve(prob, state=True, precision="split", **kw)
    fp32 = ctx.solve(prob, state=True, precision="fp32", **kw)
    assert np.array_equal(split.x, fp32.x) and np.array_equal(split.y, fp32.y)
    assert np.array_equal(split.energy_i64, fp32.energy_i64)
    xo, yo, _ = orc.solve_f32(sb.DSB, J, R, 30, init="uniform", seed=5, dt=0.5, nthreads=8)
    for r in (0, R // 2 - 1, R // 2, R - 1):
        assert np.array_equal(split.x[r], xo[r, :n]) and np.array_equal(split.y[r], yo[r, :n]), r
    prob.close()


@pytest.mark.parametrize("variant,precision,J_kind,R", [("bsb", "fp32", "sk", 768), ("dsb", "fp32", "sk", 1024),
                                                        ("asb", "fp32", "sk", 1024), ("dsb", "split", "kn", 1536)])
def test_replica_halves_equal_single_launch(ctx, variant, precision, J_kind, R):
    # From R = 768 (FP32 family) and R = 1536 (dSB on FP8 J) a solve runs as two
    # replica halves on two streams; one replica fewer runs as a single launch.
    # The first R - 1 replicas must be bit-identical (a replica's trajectory
    # does not depend on R or on the tile shape).
    n = 700
    J = sb.generate(sb.GEN_SK if J_kind == "sk" else sb.GEN_KN, n, 1.0, 43)
    prob = ctx.problem(J)
    kw = dict(variant=variant, precision=precision, steps=40, init="uniform", seed=6, dt=0.5 if variant != "asb" else 1.0)
    halves = ctx.solve(prob, state=True, replicas=R, **kw)
    single = ctx.solve(prob, state=True, replicas=R - 1, **kw)
    assert np.array_equal(halves.x[:R - 1], single.x, equal_nan=True)
    assert np.array_equal(halves.y[:R - 1], single.y, equal_nan=True)
    assert np.array_equal(halves.energy[:R - 1], single.energy)
    prob.close()
