"""External field h (edge-list "i i w" lines): the force is (J.x)_i + h_i,
added after the coupling sum in every kernel family, and energies include
-sum_i h_i s_i (exact int64 for integer J and h)."""
import os
import tempfile

import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sb.Context(0)
    yield c
    c.close()


def int_field(N, seed, lo=-3, hi=3):
    return np.random.default_rng(seed).integers(lo, hi + 1, N).astype(np.float32)


@pytest.mark.parametrize("variant", ["bsb", "dsb", "asb", "asb_m1"])
def test_ordered_with_field_bit_exact(ctx, orc, variant):
    N, R, steps = 150, 6, 80
    J = sb.generate(sb.GEN_KN, N, 1.0, 51)
    h = int_field(N, 52)
    dt = 1.0 if variant == "asb" else 0.5
    v = sb.VARIANTS[variant]
    prob = ctx.problem(J, h)
    sol = ctx.solve(prob, state=True, variant=v, precision="ordered", steps=steps, replicas=R, init="uniform",
                    seed=4, dt=dt)
    xo, yo, _ = orc.solve_f32(v, J, R, steps, init="uniform", seed=4, dt=dt, h=h)
    assert np.array_equal(sol.x, xo[:, :N], equal_nan=True)
    assert np.array_equal(sol.y, yo[:, :N], equal_nan=True)
    s = orc.spins(np.ascontiguousarray(xo), N)
    assert sol.exact and np.array_equal(sol.energy_i64, orc.energy_i64(J, s, h))
    prob.close()


@pytest.mark.parametrize("variant", ["bsb", "dsb", "asb_m1"])
def test_thin_with_field_bit_exact(ctx, orc, variant):
    N, R, steps = 1100, 3, 40
    J = sb.generate(sb.GEN_SK if variant == "bsb" else sb.GEN_KN, N, 1.0, 53)
    h = (0.2 * np.random.default_rng(54).standard_normal(N)).astype(np.float32)
    v = sb.VARIANTS[variant]
    prob = ctx.problem(J, h)
    sol = ctx.solve(prob, state=True, variant=v, precision="fp32", steps=steps, replicas=R, init="uniform",
                    seed=5, dt=0.5)
    xo, yo, _ = orc.solve_f32(v, J, R, steps, init="uniform", seed=5, dt=0.5, order=orc.ORDER_LANES, h=h)
    assert np.array_equal(sol.x, xo[:, :N]) and np.array_equal(sol.y, yo[:, :N])
    assert np.allclose(sol.energy, orc.energy_f64(J, orc.spins(np.ascontiguousarray(xo), N), h), rtol=1e-12,
                       atol=1e-9)
    prob.close()


def test_dsb_integer_field_all_paths_identical(ctx, orc):
    # Integer J and h: every partial sum is an exact integer, so the FP32 MFMA,
    # split and FP8 dSB paths and the oracle agree bit for bit.
    N, R, steps = 333, 40, 60
    J = sb.generate(sb.GEN_KN, N, 1.0, 55)
    h = int_field(N, 56, -5, 5)
    prob = ctx.problem(J, h)
    xo, yo, _ = orc.solve_f32(sb.DSB, J, R, steps, init="uniform", seed=6, dt=0.5, h=h)
    for prec in ("fp32", "split"):
        sol = ctx.solve(prob, state=True, variant="dsb", precision=prec, steps=steps, replicas=R, init="uniform",
                        seed=6, dt=0.5)
        assert np.array_equal(sol.x, xo[:, :N]) and np.array_equal(sol.y, yo[:, :N]), prec
    prob.close()


def test_fp32_field_one_step_bound(ctx, orc):
    N, R = 300, 40
    J = sb.generate(sb.GEN_SK, N, 1.0, 57)
    h = (0.3 * np.random.default_rng(58).standard_normal(N)).astype(np.float32)
    prob = ctx.problem(J, h)
    c0 = 0.37
    sol = ctx.solve(prob, state=True, variant="bsb", precision="fp32", steps=1, replicas=R, init="uniform",
                    seed=7, dt=0.9, c0=c0)
    xo, yo, _ = orc.solve_f32(sb.BSB, J, R, 1, init="uniform", seed=7, dt=0.9, c0=c0, h=h)
    x0, _ = orc.init_state("uniform", N, R, seed=7)
    absF = np.abs(x0[:, :N]).astype(np.float64) @ np.abs(J).T.astype(np.float64) + np.abs(h)
    beta = np.float32(0.9 * c0)
    tol = float(beta) * 2 * N * 2.0**-24 * absF + 4 * np.spacing(np.abs(yo[:, :N])) + 1e-30
    assert np.all(np.abs(sol.y.astype(np.float64) - yo[:, :N]) <= tol)
    prob.close()


def test_field_from_edgelist_file_and_batch(ctx):
    N = 120
    J = sb.generate(sb.GEN_ER, N, 0.3, 59)
    h = int_field(N, 60)
    lines = [f"{N}"] + [f"{i} {j} {int(J[i, j])}" for i in range(N) for j in range(i + 1, N) if J[i, j] != 0]
    lines += [f"{i} {i} {int(h[i])}" for i in range(N) if h[i] != 0]
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        f.write("\n".join(lines) + "\n")
        path = f.name
    try:
        pf = ctx.problem_from_file(path)
    finally:
        os.unlink(path)
    pd = ctx.problem(J, h)
    kw = dict(variant="bsb", precision="split", steps=50, replicas=48, init="uniform", seed=8, dt=0.5)
    a = ctx.solve(pf, state=True, **kw)
    b = ctx.solve(pd, state=True, **kw)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.energy_i64, b.energy_i64) and a.exact
    # batched with a field-free instance: each equals its solo solve
    p0 = ctx.problem(sb.generate(sb.GEN_KN, 200, 1.0, 61))
    params = [sb.make_params(**kw), sb.make_params(**kw)]
    sols, _ = ctx.batch_solve([pd, p0], params, state=True)
    c = ctx.solve(p0, state=True, **kw)
    assert np.array_equal(sols[0].x, b.x) and np.array_equal(sols[1].x, c.x)
    for p in (pf, pd, p0):
        p.close()


def test_sb_field_includes_h(ctx, orc):
    N, R = 257, 5
    J = sb.generate(sb.GEN_KN, N, 1.0, 62)
    h = int_field(N, 63)
    x, _ = orc.init_state("uniform", N, R, seed=3)
    xs = x[:, :N].copy()
    prob = ctx.problem(J, h)
    F = ctx.field(prob, xs, "ordered", "bsb")
    assert np.array_equal(F, (orc.field_f32(J, x)[:, :N] + h).astype(np.float32))
    prob.set_field(None)
    assert np.array_equal(ctx.field(prob, xs, "ordered", "bsb"), orc.field_f32(J, x)[:, :N])
    prob.close()
