"""Sparse family (CSR J, spin-major state): the field is the sequential fmaf
chain over the nonzeros in ascending column order, so a sparse solve equals the
oracle's sequential (SEQ) order bit for bit, like SB_PREC_ORDERED."""
import os

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


@pytest.mark.parametrize("N", [1, 2, 5, 40])
@pytest.mark.parametrize("R", [1, 64, 70])
def test_sparse_empty_rows_and_tiny_instances(ctx, orc, N, R):
    # Spins without a nonzero (every third spin isolated), an all-zero J
    # (N <= 2: no nonzeros at all) and fewer spins than a workgroup's four.
    J = np.zeros((N, N), np.float32)
    if N >= 5:
        J = sb.generate(sb.GEN_ER, N, 0.3, 90 + N)
        J[:, ::3] = 0.0
        J[::3, :] = 0.0
    prob = sparse_problem(ctx, J)
    assert prob.nnz == int(np.count_nonzero(J))
    for variant in ("bsb", "dsb"):
        v = sb.VARIANTS[variant]
        sol = ctx.solve(prob, state=True, variant=v, precision="fp32", steps=30, replicas=R, init="uniform",
                        seed=N, dt=0.5, c0=0.3)
        xo, yo, _ = orc.solve_f32(v, J, R, 30, init="uniform", seed=N, dt=0.5, c0=0.3)
        assert np.array_equal(sol.x, xo[:, :N]) and np.array_equal(sol.y, yo[:, :N]), variant
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


@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_sparse_problems_in_a_batch(ctx, orc, precision):
    # Sparse and dense problems batched together: the sparse instances run the
    # gather tiles inside the persistent launch and equal both their solo
    # sparse solve and the oracle's sequential order, bit for bit.
    specs = [("sparse", sb.GEN_ER, 600, 0.02, "bsb", 48, 80), ("dense", sb.GEN_KN, 300, 1.0, "bsb", 64, 60),
             ("sparse", sb.GEN_SK_DILUTED, 1500, 0.01, "dsb", 130, 50), ("sparse", sb.GEN_ER, 257, 0.05, "asb", 3, 70),
             ("dense", sb.GEN_SK, 400, 1.0, "dsb", 2, 40)]
    probs, params, Js = [], [], []
    for i, (form, kind, N, p, v, R, steps) in enumerate(specs):
        J = sb.generate(kind, N, p, 90 + i)
        prob = ctx.problem(J)
        if form == "sparse":
            prob.set_sparse(True)
            assert prob.nnz == int(np.count_nonzero(J))
        probs.append(prob)
        Js.append(J)
        params.append(sb.make_params(variant=v, precision=precision, steps=steps, replicas=R, init="uniform",
                                     seed=i, dt=1.0 if v == "asb" else 0.5))
    sols, st = ctx.batch_solve(probs, params, spins=True, state=True)
    assert st.graph_used == 1
    for i, (form, kind, N, p, v, R, steps) in enumerate(specs):
        kw = dict(state=True, variant=v, precision=precision, steps=steps, replicas=R, init="uniform", seed=i,
                  dt=1.0 if v == "asb" else 0.5)
        solo = ctx.solve(probs[i], **kw)
        assert np.array_equal(sols[i].x, solo.x) and np.array_equal(sols[i].y, solo.y), i
        assert np.array_equal(sols[i].energy, solo.energy), i
        if form == "sparse":
            xo, yo, _ = orc.solve_f32(sb.VARIANTS[v], Js[i], R, steps, init="uniform", seed=i, dt=kw["dt"])
            assert np.array_equal(sols[i].x, xo[:, :N]) and np.array_equal(sols[i].y, yo[:, :N]), i
    for prob in probs:
        prob.close()


def test_server_runs_low_density_requests_sparse(ctx, orc):
    # sparse_density=0.02: a request with at most 2 % nonzeros gets its CSR form
    # built at admission and follows the oracle's sequential order (bSB on real
    # J included).
    GB = 1 << 30
    specs = [(sb.GEN_SK_DILUTED, 2000, 0.01, 31, "bsb", 64, 120), (sb.GEN_ER, 700, 0.015, 32, "dsb", 33, 90),
             (sb.GEN_ER, 500, 0.5, 33, "bsb", 32, 60)]
    with sb.Server(ctx, 8 * GB, chunk_steps=16, precision="split", sparse_density=0.02) as srv:
        tickets = [srv.submit_generated(k, N, p, g, variant=v, precision="split", steps=s, replicas=R,
                                        init="uniform", seed=g, dt=0.5) for k, N, p, g, v, R, s in specs]
        res = [srv.wait(t) for t in tickets]
    for (k, N, p, g, v, R, s), r in zip(specs, res):
        prob = ctx.generate(k, N, p, g)
        J = prob.read_j()
        if p <= 0.02:
            prob.set_sparse(True)
            xo, _, _ = orc.solve_f32(sb.VARIANTS[v], J, R, s, init="uniform", seed=g, dt=0.5,
                                     nthreads=min(16, os.cpu_count() or 1))
            assert np.array_equal(r.spins, orc.spins(np.ascontiguousarray(xo), N)), N
        solo = ctx.solve(prob, variant=v, precision="split", steps=s, replicas=R, init="uniform", seed=g, dt=0.5)
        assert np.array_equal(r.spins, solo.spins) and np.array_equal(r.energy, solo.energy), N
        prob.close()
