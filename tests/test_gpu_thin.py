"""Few-replica (R <= 4) matrix-vector families: the FP32 path is bit-exact
against the oracle's lane-order field (the thin kernels' documented summation
order), and the field does not depend on the thin shape or on whether J is
read as FP32, BF16 or FP8 when those hold the same values."""
import ctypes as C
import os

import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu

THIN_F32, THIN_BF16, THIN_FP8 = 4, 5, 6
THIN_SHAPES = 4
THIN_BATCH_SHAPES = 3


@pytest.fixture(scope="module")
def ctx():
    c = sb.Context(0)
    yield c
    c.close()


def solve_forced(ctx, prob, family, shape, **kw):
    os.environ["SB_SHAPE"] = f"{family}:{shape}"
    try:
        return ctx.solve(prob, state=True, **kw)
    finally:
        del os.environ["SB_SHAPE"]


@pytest.mark.parametrize("variant", ["bsb", "asb", "dsb", "asb_m1"])
# R = 1, 2 and 3-4 run the kernels compiled for 1, 2 and 4 replicas; N covers
# rows of one partial 16-element group (5), one group per lane (64, 300), two
# (1025, 1100) and three (2100, 3000).
@pytest.mark.parametrize("N,R", [(300, 1), (1100, 3), (2100, 4), (5, 2), (64, 2), (1025, 2), (3000, 1)])
def test_thin_fp32_bit_exact_vs_lane_order_oracle(ctx, orc, variant, N, R):
    J = sb.generate(sb.GEN_SK if variant == "bsb" else sb.GEN_KN, N, 1.0, 21)
    v = sb.VARIANTS[variant]
    dt = 1.0 if variant == "asb" else 0.5
    steps = 40
    prob = ctx.problem(J)
    sol = ctx.solve(prob, state=True, variant=v, precision="fp32", steps=steps, replicas=R, init="uniform",
                    seed=6, dt=dt)
    xo, yo, _ = orc.solve_f32(v, J, R, steps, init="uniform", seed=6, dt=dt, order=orc.ORDER_LANES)
    assert np.array_equal(sol.x, xo[:, :N], equal_nan=True), np.nanmax(np.abs(sol.x - xo[:, :N]))
    assert np.array_equal(sol.y, yo[:, :N], equal_nan=True)
    s = orc.spins(np.ascontiguousarray(xo), N)
    assert np.array_equal(sol.spins, s)
    prob.close()


@pytest.mark.parametrize("variant", ["bsb", "dsb"])
@pytest.mark.parametrize("N", [1, 2, 3, 16, 17])
@pytest.mark.parametrize("R", [1, 2, 4])
def test_thin_tiny_instances_bit_exact(ctx, orc, variant, N, R):
    # Fewer spins than one 16-element group, and one wave's rows mostly past N.
    J = sb.generate(sb.GEN_KN, N, 1.0, 30 + N) if N > 1 else np.zeros((1, 1), np.float32)
    v = sb.VARIANTS[variant]
    prob = ctx.problem(J)
    sol = ctx.solve(prob, state=True, variant=v, precision="fp32", steps=25, replicas=R, init="uniform", seed=N,
                    dt=0.5, c0=0.3)
    xo, yo, _ = orc.solve_f32(v, J, R, 25, init="uniform", seed=N, dt=0.5, c0=0.3, order=orc.ORDER_LANES)
    assert np.array_equal(sol.x, xo[:, :N]) and np.array_equal(sol.y, yo[:, :N])
    prob.close()


@pytest.mark.parametrize("variant", ["bsb", "dsb"])
def test_thin_every_shape_and_j_format_identical(ctx, variant):
    # Integer J: FP32, BF16 and FP8 hold the same values, so every thin family
    # and shape must give the same trajectory.
    N, R = 777, 3
    J = sb.generate(sb.GEN_KN, N, 1.0, 22)
    prob = ctx.problem(J)
    ref = None
    for prec, fams in (("fp32", (THIN_F32,)), ("split", (THIN_FP8, THIN_BF16))):
        for f in fams:
            for s in range(THIN_SHAPES):
                sol = solve_forced(ctx, prob, f, s, variant=variant, precision=prec, steps=50, replicas=R,
                                   init="uniform", seed=8, dt=0.5)
                if ref is None:
                    ref = sol
                    continue
                assert np.array_equal(sol.x, ref.x) and np.array_equal(sol.y, ref.y), (prec, f, s)
    prob.close()


def test_thin_split_real_j_field_bound(ctx, orc):
    # Real J in the split precision form: the thin path reads BF16 J, so the
    # field is that of bf16(J) in FP32.
    N, R = 500, 2
    J = sb.generate(sb.GEN_SK, N, 1.0, 23)
    Jb = (J.view(np.uint32) + 0x7FFF + ((J.view(np.uint32) >> 16) & 1)) & 0xFFFF0000
    Jb = Jb.view(np.float32)
    x, _ = orc.init_state("uniform", N, R, seed=4)
    with sb.Context(0) as c:
        prob = c.problem(J)
        F = c.field(prob, x[:, :N].copy(), "split", "bsb")
        prob.close()
    F64 = orc.field_f64(Jb, x.astype(np.float64))[:, :N]
    bound = N * 2.0**-24 * (np.abs(x[:, :N]).astype(np.float64) @ np.abs(Jb).T.astype(np.float64))
    assert np.all(np.abs(F - F64) <= bound)
    Fl = orc.field_f32(Jb, x, order=orc.ORDER_LANES)[:, :N]
    assert np.array_equal(F, Fl)


@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_thin_batch_equals_solo(ctx, precision):
    Js = [sb.generate(sb.GEN_ER, 400, 0.2, 31), sb.generate(sb.GEN_KN, 1300, 1.0, 32),
          sb.generate(sb.GEN_SK, 90, 1.0, 33)]
    Rs, variants = [1, 4, 40], ["bsb", "dsb", "bsb"]
    probs = [ctx.problem(J) for J in Js]
    params = [sb.make_params(variant=v, precision=precision, steps=35, replicas=r, init="uniform", seed=9 + i,
                             dt=0.5) for i, (v, r) in enumerate(zip(variants, Rs))]
    solos = [ctx.solve(p, state=True, variant=v, precision=precision, steps=35, replicas=r, init="uniform",
                       seed=9 + i, dt=0.5) for i, (p, v, r) in enumerate(zip(probs, variants, Rs))]
    for b in range(THIN_BATCH_SHAPES):
        os.environ["SB_BATCH_SHAPE"] = str(b)
        try:
            sols, _ = ctx.batch_solve(probs, params, state=True)
        finally:
            del os.environ["SB_BATCH_SHAPE"]
        for i in range(3):  # two thin instances and one on MFMA tiles
            assert np.array_equal(sols[i].x, solos[i].x), (b, i)
            assert np.array_equal(sols[i].y, solos[i].y), (b, i)
    for p in probs:
        p.close()


def test_thin_sb_step_matches_solve(orc):
    import torch
    N, R, steps = 600, 2, 30
    J = sb.generate(sb.GEN_KN, N, 1.0, 24)
    ld = orc.pad_to(N)
    lib = sb.load()
    params = sb.make_params(variant=sb.BSB, precision=sb.PREC_SPLIT, steps=steps, replicas=R, init=sb.INIT_UNIFORM,
                            seed=4, dt=0.5)
    c0 = sb.c0_auto(sb.BSB, N, 1.0, sb.instance_stats(J)[0])
    with sb.Context(0) as ctx:
        prob = ctx.problem(J)
        sol = ctx.solve(prob, state=True, variant="bsb", precision="fp32", steps=steps, replicas=R, init="uniform",
                        seed=4, dt=0.5)
        Jp = np.zeros((N, ld), np.float32)
        Jp[:, :N] = J
        forms = {sb.J_FP32: torch.from_numpy(Jp).cuda(),
                 sb.J_BF16: torch.from_numpy((Jp.view(np.uint32) >> 16).astype(np.uint16).view(np.int16)).cuda()}
        for jfmt, Jd in forms.items():
            X = torch.zeros((R, ld), device="cuda")
            Xn = torch.zeros((R, ld), device="cuda")
            Y = torch.zeros((R, ld), device="cuda")
            stream = torch.cuda.current_stream().cuda_stream
            assert lib.sb_init_state(ctx.handle, sb.INIT_UNIFORM, 4, N, R, ld, X.data_ptr(), Y.data_ptr(),
                                     stream) == 0
            prec = sb.PREC_FP32 if jfmt == sb.J_FP32 else sb.PREC_SPLIT
            for k in range(steps):
                cf = sb.step_coeffs(params, c0, k)
                rc = lib.sb_step(ctx.handle, prec, Jd.data_ptr(), jfmt, N, ld, X.data_ptr(), Xn.data_ptr(),
                                 Y.data_ptr(), R, ld, C.byref(cf), stream)
                assert rc == 0, lib.sb_last_error_message()
                X, Xn = Xn, X
            torch.cuda.synchronize()
            assert np.array_equal(X.cpu().numpy()[:, :N], sol.x), jfmt
            assert np.array_equal(Y.cpu().numpy()[:, :N], sol.y), jfmt
        prob.close()


def _ctx_with(env):
    import os
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return sb.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def test_solo_graph_replays_equal_eager_launches(orc):
    # Launch-bound solves (R <= 4, or small MFMA steps) replay their step loop
    # from a cached graph.  Every replay, with other eager solves and
    # synchronisations in between, must equal the step-by-step launches.
    cases = [(sb.GEN_ER, 800, 0.06, "bsb", "fp32", 1, 300), (sb.GEN_KN, 2000, 1.0, "dsb", "split", 2, 150),
             (sb.GEN_SK, 300, 1.0, "bsb", "fp32", 40, 120), (sb.GEN_SK, 500, 1.0, "asb", "split", 3, 90),
             (sb.GEN_SK, 3000, 1.0, "asb_m1", "fp32", 4, 33)]
    eager = _ctx_with({"SB_SOLO_GRAPH": "0"})
    with sb.Context(0) as ctx:
        for kind, N, p, v, prec, R, steps in cases:
            J = sb.generate(kind, N, p, 5)
            h = np.random.default_rng(N).standard_normal(N).astype(np.float32) if N in (800, 3000) else None
            kw = dict(state=True, variant=v, precision=prec, steps=steps, replicas=R, init="uniform", seed=3,
                      dt=1.0 if v == "asb" else 0.5)
            pe = eager.problem(J, h)
            ref = eager.solve(pe, **kw)
            pg = ctx.problem(J, h)
            other = ctx.problem(sb.generate(sb.GEN_KN, 100, 1.0, 9))
            for rep in range(3):
                got = ctx.solve(pg, **kw)
                assert np.array_equal(got.x, ref.x, equal_nan=True), (N, R, rep)
                assert np.array_equal(got.y, ref.y, equal_nan=True), (N, R, rep)
                assert np.array_equal(got.energy, ref.energy, equal_nan=True)
                ctx.solve(other, variant="bsb", precision="fp32", steps=7, replicas=9, init="uniform", seed=rep)
            pg.close()
            other.close()
            pe.close()
    eager.close()
