"""GPU parity of the HIP path (through the C ABI) against the CPU oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def ctx():
    c = sb.Context(0)
    yield c
    c.close()


def sym(J):
    J = np.triu(np.asarray(J, np.float32), 1)
    return (J + J.T).astype(np.float32)


def kn_pm1(N, seed):
    return sb.generate(sb.GEN_KN, N, 1.0, seed)


def sk(N, seed):
    return sb.generate(sb.GEN_SK, N, 1.0, seed)


def oracle_run(orc, variant, J, R, steps, init, seed, dt, a0=1.0, c0=None, K=1.0, M=2):
    x, y, c0 = orc.solve_f32(variant, J, R, steps, init=init, seed=seed, dt=dt, a0=a0, c0=c0, K=K, M=M)
    N = J.shape[0]
    return x[:, :N], y[:, :N], c0


# ---- init (R3) ---------------------------------------------------------------

@pytest.mark.parametrize("init", ["reference", "uniform"])
def test_init_bit_exact(orc, init):
    import torch
    lib = sb.load()
    c = sb.Context(0)
    N, R, ld = 77, 9, 128
    X = torch.full((R, ld), 7.0, device="cuda")
    Y = torch.full((R, ld), 7.0, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.sb_init_state(c.handle, sb.INITS[init], 5, N, R, ld, X.data_ptr(), Y.data_ptr(), stream)
    assert rc == 0
    torch.cuda.synchronize()
    xo, yo = orc.init_state(init, N, R, seed=5, ld=ld)
    assert np.array_equal(X.cpu().numpy(), xo) and np.array_equal(Y.cpu().numpy(), yo)
    c.close()


# ---- ordered-sum mode: bit-exact trajectories ---------------------------------

@pytest.mark.parametrize("variant", ["asb", "bsb", "dsb", "asb_m1"])
@pytest.mark.parametrize("N,R,steps,init", [(67, 5, 120, "uniform"), (200, 12, 150, "reference")])
def test_ordered_matches_oracle_bit_exact(ctx, orc, variant, N, R, steps, init):
    J = sk(N, 11) if variant == "bsb" else kn_pm1(N, 12)
    v = sb.VARIANTS[variant]
    dt = 1.0 if variant == "asb" else 0.5
    prob = ctx.problem(J)
    sol = ctx.solve(prob, state=True, variant=v, precision="ordered", steps=steps, replicas=R,
                    init=init, seed=3, dt=dt)
    xo, yo, c0 = oracle_run(orc, v, J, R, steps, init, 3, dt)
    assert sol.c0 == pytest.approx(c0, rel=1e-12)
    # (aSB can run away to inf/NaN for some replicas; NaN == NaN here)
    assert np.array_equal(sol.x, xo, equal_nan=True), np.nanmax(np.abs(sol.x - xo))
    assert np.array_equal(sol.y, yo, equal_nan=True)
    s = orc.spins(np.ascontiguousarray(xo), N)
    assert np.array_equal(sol.spins, s)
    if prob.integer:
        assert np.array_equal(sol.energy_i64, orc.energy_i64(J, s))
    else:
        assert np.allclose(sol.energy, orc.energy_f64(J, s), rtol=1e-12, atol=1e-9)


# ---- FP32 MFMA path ------------------------------------------------------------

@pytest.mark.parametrize("N", [1, 2, 3, 33, 64, 65, 130, 300])
@pytest.mark.parametrize("R", [1, 5, 40])
def test_fp32_one_step_within_field_bound(ctx, orc, N, R):
    J = sk(N, N) if N > 1 else np.zeros((1, 1), np.float32)
    prob = ctx.problem(J)
    c0 = 0.37
    sol = ctx.solve(prob, state=True, variant="bsb", precision="fp32", steps=1, replicas=R,
                    init="uniform", seed=N, dt=0.9, c0=c0)
    xo, yo, _ = oracle_run(orc, sb.BSB, J, R, 1, "uniform", N, 0.9, c0=c0)
    x0, _ = orc.init_state("uniform", N, R, seed=N)
    absF = np.abs(x0[:, :N]).astype(np.float64) @ np.abs(J).T.astype(np.float64)
    beta = np.float32(0.9 * c0)
    tol = float(beta) * 2 * N * 2.0**-24 * absF + 4 * np.spacing(np.abs(yo)) + 1e-30
    assert np.all(np.abs(sol.y.astype(np.float64) - yo) <= tol)


def test_fp32_trajectory_statistical_parity(ctx, orc):
    # Against the sequential order (the reference's naive kernel) the MFMA
    # tiles sum in a different order: trajectories may split where a spin sits
    # near x = 0; most replicas end with identical spins and the energy
    # distribution is unchanged.
    N, R, steps = 256, 64, 400
    J = kn_pm1(N, 21)
    prob = ctx.problem(J)
    sol = ctx.solve(prob, state=True, variant="bsb", precision="fp32", steps=steps, replicas=R,
                    init="uniform", seed=9, dt=0.5)
    xo, yo, c0 = oracle_run(orc, sb.BSB, J, R, steps, "uniform", 9, 0.5)
    # bit for bit in the FP32 tiles' own summation order
    xm, ym, _ = orc.solve_f32(sb.BSB, J, R, steps, init="uniform", seed=9, dt=0.5, order=orc.ORDER_MFMA)
    assert np.array_equal(sol.x, xm[:, :N]) and np.array_equal(sol.y, ym[:, :N])
    so = orc.spins(np.ascontiguousarray(xo), N)
    same = np.all(sol.spins == so, axis=1)
    Ho = orc.energy_i64(J, so)
    # exact energies for the GPU's own spins
    assert np.array_equal(sol.energy_i64, orc.energy_i64(J, sol.spins))
    assert np.array_equal(sol.energy_i64[same], Ho[same])
    print(f"N={N} R={R} {steps} steps: identical final spins in {same.sum()}/{R} replicas, "
          f"mean energy gap {abs(sol.energy.mean() - Ho.mean()) / abs(Ho.mean()):.2e}")
    # Measured on MI355X: 64 of 64 replicas end with identical spins.
    assert same.sum() >= 63, same.sum()
    assert abs(sol.energy.mean() - Ho.mean()) <= 1e-3 * abs(Ho.mean())


def test_dsb_integer_all_paths_identical(ctx, orc):
    # Integer J: every partial sum of J.sign(x) is an exact integer in FP32, so
    # FP32 MFMA, FP8 x FP8 MFMA, BF16 split and the oracle agree bit for bit.
    N, R, steps = 300, 40, 200
    J = kn_pm1(N, 31)
    prob = ctx.problem(J)
    assert prob.split_jfmt == sb.J_FP8
    sols = [ctx.solve(prob, state=True, variant="dsb", precision=p, steps=steps, replicas=R,
                      init="uniform", seed=2, dt=0.5) for p in ("fp32", "split", "ordered")]
    xo, yo, _ = oracle_run(orc, sb.DSB, J, R, steps, "uniform", 2, 0.5)
    for name, s in zip(("fp32", "split", "ordered"), sols):
        bad = np.where(~(np.all(s.x == xo, axis=1) & np.all(s.y == yo, axis=1)))[0]
        assert bad.size == 0, f"{name}: replicas {bad.tolist()} differ"


def test_stale_workspace_planes_do_not_leak(orc):
    # Regression: the second B-operand plane buffer was not initialised, so
    # its padding columns (N..ld) kept bytes of an earlier, larger solve; J is
    # zero there but 0 * NaN is NaN.  Leave NaNs behind with a diverging aSB
    # solve at N=300, then solve dSB at N=260 (padding 260..319) on the same
    # context and compare with the oracle.
    with sb.Context(0) as c:
        big = c.problem(kn_pm1(300, 90))
        blown = c.solve(big, state=True, variant="asb", precision="split", steps=40, replicas=48,
                        init="uniform", seed=1, dt=1.0, c0=1.0)
        assert np.isnan(blown.x).any()
        J = kn_pm1(260, 91)
        small = c.problem(J)
        xo, yo, _ = oracle_run(orc, sb.DSB, J, 48, 60, "uniform", 2, 0.5)
        for prec in ("split", "fp32"):
            s = c.solve(small, state=True, variant="dsb", precision=prec, steps=60, replicas=48, init="uniform",
                        seed=2, dt=0.5)
            assert np.array_equal(s.x, xo) and np.array_equal(s.y, yo), prec


def test_results_independent_of_tile_and_batch_shape(ctx):
    # Keyed init + per-element k order independent of the tile: replica r's
    # trajectory does not depend on R (which changes the tile grid).
    J = sk(190, 5)
    prob = ctx.problem(J)
    a = ctx.solve(prob, state=True, variant="bsb", precision="fp32", steps=80, replicas=37, init="uniform",
                  seed=4, dt=0.5)
    b = ctx.solve(prob, state=True, variant="bsb", precision="fp32", steps=80, replicas=300, init="uniform",
                  seed=4, dt=0.5)
    assert np.array_equal(a.x, b.x[:37]) and np.array_equal(a.y, b.y[:37])
    c = ctx.solve(prob, state=True, variant="bsb", precision="split", steps=80, replicas=37, init="uniform",
                  seed=4, dt=0.5)
    d = ctx.solve(prob, state=True, variant="bsb", precision="split", steps=80, replicas=300, init="uniform",
                  seed=4, dt=0.5)
    assert np.array_equal(c.x, d.x[:37])


# ---- precision form --------------------------------------------------------------

def test_split_one_step_field_error(ctx, orc):
    # X = hi + lo (BF16 pair, |x - hi - lo| <= 2^-17 |x|) and integer J exact in
    # FP8/BF16: the field error is bounded by 2^-16 sum|J||x| plus accumulation.
    N, R = 500, 32
    J = kn_pm1(N, 41)
    prob = ctx.problem(J)
    c0 = 0.05
    sol = ctx.solve(prob, state=True, variant="bsb", precision="split", steps=1, replicas=R, init="uniform",
                    seed=6, dt=1.0, c0=c0)
    x0, y0 = orc.init_state("uniform", N, R, seed=6)
    F64 = orc.field_f64(J, x0.astype(np.float64))[:, :N]
    absF = np.abs(x0[:, :N]).astype(np.float64) @ np.abs(J).T.astype(np.float64)
    a = orc.coeffs(sb.BSB, 0, 1, 1.0, 1.0, c0)
    y_exact = y0[:, :N].astype(np.float64) - float(a[0]) * x0[:, :N] + float(a[1]) * F64
    err = np.abs(sol.y - y_exact)
    tol = float(a[1]) * (2.0**-16 + 2 * N * 2.0**-24) * absF + 4 * np.spacing(np.abs(y_exact).astype(np.float32))
    assert np.all(err <= tol)


def test_split_bsb_energy_close_to_fp32(ctx):
    N, R, steps = 512, 128, 500
    J = kn_pm1(N, 51)
    prob = ctx.problem(J)
    f = ctx.solve(prob, variant="bsb", precision="fp32", steps=steps, replicas=R, init="uniform", seed=1, dt=0.5)
    s = ctx.solve(prob, variant="bsb", precision="split", steps=steps, replicas=R, init="uniform", seed=1, dt=0.5)
    assert s.exact and f.exact
    assert abs(s.energy.mean() - f.energy.mean()) <= 0.01 * abs(f.energy.mean())


def test_split_real_j_uses_bf16(ctx):
    J = sk(300, 61)
    prob = ctx.problem(J)
    assert prob.split_jfmt == sb.J_BF16 and not prob.integer
    f = ctx.solve(prob, variant="bsb", precision="fp32", steps=300, replicas=64, init="uniform", seed=1, dt=0.5)
    s = ctx.solve(prob, variant="bsb", precision="split", steps=300, replicas=64, init="uniform", seed=1, dt=0.5)
    assert abs(s.energy.mean() - f.energy.mean()) <= 0.02 * abs(f.energy.mean())


# ---- energy (R7) ------------------------------------------------------------------

def test_energy_kernel_exact(ctx, orc):
    for J in (kn_pm1(257, 3), sym(np.random.default_rng(0).integers(-9, 10, (100, 100)))):
        prob = ctx.problem(J)
        sol = ctx.solve(prob, variant="bsb", steps=30, replicas=70, init="uniform", seed=8, dt=0.5)
        assert sol.exact
        assert np.array_equal(sol.energy_i64, orc.energy_i64(J, sol.spins))
        assert sol.best == int(np.argmin(sol.energy_i64))
    J = sk(129, 2)
    sol = ctx.solve(ctx.problem(J), variant="bsb", steps=30, replicas=33, init="uniform", seed=8, dt=0.5)
    assert not sol.exact
    assert np.allclose(sol.energy, orc.energy_f64(J, sol.spins), rtol=1e-12, atol=1e-10)


def test_replicas_beyond_the_grid_y_limit(ctx, orc):
    # R above 65535, the y extent of a launch grid: the init, aSB prologue,
    # ordered step and readout kernels stride over replicas.
    N, R, steps = 20, 70000, 5
    J = kn_pm1(N, 5)
    prob = ctx.problem(J)
    for variant, precision, order in (("asb", "ordered", orc.ORDER_SEQ), ("bsb", "fp32", orc.ORDER_MFMA)):
        v = sb.VARIANTS[variant]
        sol = ctx.solve(prob, state=True, variant=v, precision=precision, steps=steps, replicas=R, init="uniform",
                        seed=2, dt=0.5)
        xo, yo, _ = orc.solve_f32(v, J, R, steps, init="uniform", seed=2, dt=0.5, order=order)
        assert np.array_equal(sol.x, xo[:, :N]) and np.array_equal(sol.y, yo[:, :N]), variant
        so = orc.spins(np.ascontiguousarray(xo[:, :N]), N)
        assert np.array_equal(sol.spins, so), variant
        assert np.array_equal(sol.energy_i64, orc.energy_i64(J, so)), variant
    prob.close()


# ---- errors and the CLI boundary --------------------------------------------------

def test_invalid_params_rejected(ctx):
    prob = ctx.problem(kn_pm1(10, 1))
    for kw in (dict(steps=0), dict(replicas=0), dict(variant=5), dict(precision=9), dict(M=0)):
        with pytest.raises(sb.SBError) as e:
            ctx.solve(prob, **kw)
        assert e.value.status == 1


@pytest.mark.parametrize("mode", [["--precision", "ordered"], ["--precision", "fp32", "--sparse"]])
def test_cli_writes_output_txt(ctx, orc, mode):
    J = kn_pm1(40, 77)
    lines = ["40"] + [f"{i} {j} {int(J[i, j])}" for i in range(40) for j in range(i + 1, 40)]
    exe = os.path.join(ROOT, "annealing_amd", "bin", "sb_solve")
    with tempfile.TemporaryDirectory() as d:
        inst = os.path.join(d, "k40.txt")
        open(inst, "w").write("\n".join(lines) + "\n")
        r = subprocess.run([exe, inst, "--variant", "bsb", *mode, "--steps", "200",
                            "--replicas", "6", "--init", "uniform", "--seed", "3", "--dt", "0.5",
                            "--spins", "spins.txt"], cwd=d, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        H = np.loadtxt(os.path.join(d, "output.txt"), dtype=np.int64)
        S = np.loadtxt(os.path.join(d, "spins.txt"), dtype=np.int8)
    xo, _, _ = oracle_run(orc, sb.BSB, J, 6, 200, "uniform", 3, 0.5)
    so = orc.spins(np.ascontiguousarray(xo), 40)
    assert np.array_equal(S, so)
    assert np.array_equal(H, orc.energy_i64(J, so))


@pytest.mark.parametrize("name", ["split", "fp8bf16"])
def test_cli_split_matches_oracle_split_order(ctx, orc, name):
    # --precision split through the CLI: integer J stored as FP8 (exact), X as
    # BF16 hi + lo on the BF16 MFMA; the oracle's model of that MFMA gives the
    # same spins, hence the same int64 energies.
    N, R = 40, 6
    J = kn_pm1(N, 79)
    lines = [f"{N}"] + [f"{i} {j} {int(J[i, j])}" for i in range(N) for j in range(i + 1, N)]
    exe = os.path.join(ROOT, "annealing_amd", "bin", "sb_solve")
    with tempfile.TemporaryDirectory() as d:
        inst = os.path.join(d, "k40.txt")
        open(inst, "w").write("\n".join(lines) + "\n")
        r = subprocess.run([exe, inst, "--variant", "bsb", "--precision", name, "--steps", "200",
                            "--replicas", str(R), "--init", "uniform", "--seed", "3", "--dt", "0.5",
                            "--spins", "spins.txt"], cwd=d, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        H = np.loadtxt(os.path.join(d, "output.txt"), dtype=np.int64)
        S = np.loadtxt(os.path.join(d, "spins.txt"), dtype=np.int8)
    xo, _, _ = orc.solve_f32(orc.BSB, J, R, 200, init="uniform", seed=3, dt=0.5, order=orc.ORDER_SPLIT)
    so = orc.spins(np.ascontiguousarray(xo[:, :N]), N)
    assert np.array_equal(S, so)
    assert np.array_equal(H, orc.energy_i64(J, so))


def test_cli_fractional_field_writes_fp64_energies(ctx, orc):
    # Integer J with a fractional field line: energies are not integers, so
    # output.txt carries the FP64 values (not truncated integers).
    N = 30
    J = kn_pm1(N, 78)
    h = np.zeros(N, np.float32)
    h[3], h[17] = 0.5, -1.25
    lines = [f"{N}"] + [f"{i} {j} {int(J[i, j])}" for i in range(N) for j in range(i + 1, N)]
    lines += ["3 3 0.5", "17 17 -1.25"]
    exe = os.path.join(ROOT, "annealing_amd", "bin", "sb_solve")
    with tempfile.TemporaryDirectory() as d:
        inst = os.path.join(d, "k30h.txt")
        open(inst, "w").write("\n".join(lines) + "\n")
        r = subprocess.run([exe, inst, "--variant", "bsb", "--precision", "ordered", "--steps", "100",
                            "--replicas", "4", "--init", "uniform", "--seed", "2", "--dt", "0.5",
                            "--spins", "spins.txt", "--cut", "cut.txt"], cwd=d, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        H = np.loadtxt(os.path.join(d, "output.txt"), dtype=np.float64)
        S = np.loadtxt(os.path.join(d, "spins.txt"), dtype=np.int8)
        cut = np.loadtxt(os.path.join(d, "cut.txt"), dtype=np.int64)
    x, y = orc.init_state("uniform", N, 4, seed=2)
    c0 = orc.c0_auto(orc.BSB, J)
    orc.run_f32(orc.BSB, J, x, y, 0, 100, 100, 0.5, 1.0, c0, h=h)
    so = orc.spins(x, N)
    assert np.array_equal(S, so)
    Ho = orc.energy_f64(J, so, h=h)
    assert np.array_equal(H, Ho)
    assert np.any(H != np.round(H))
    W = -np.triu(J.astype(np.int64), 1)
    assert np.array_equal(cut, [W[so[r][:, None] != so[r][None, :]].sum() for r in range(4)])


def test_split_choice_not_shared_between_fp8_and_bf16_problems():
    # The autotuner caches its (family, shape) choice per size bucket.  An
    # FP8-exact problem chooses among the FP8- and BF16-J split families; a
    # real-valued problem of the same size has no FP8 copy of J, so it must
    # make its own choice (the cache key holds the candidate set) and give the
    # result of a fresh context.
    N, R, steps = 600, 96, 40
    kw = dict(variant="bsb", precision="split", steps=steps, replicas=R, init="uniform", seed=5, dt=0.5)
    with sb.Context(0) as ctx:
        p_int = ctx.problem(sb.generate(sb.GEN_KN, N, 1.0, seed=5))
        p_real = ctx.problem(sb.generate(sb.GEN_SK, N, 1.0, seed=5))
        assert p_int.split_jfmt == sb.J_FP8 and p_real.split_jfmt == sb.J_BF16
        ctx.solve(p_int, **kw)
        after = ctx.solve(p_real, **kw)
        p_int.close()
        p_real.close()
    with sb.Context(0) as ctx:
        p_real = ctx.problem(sb.generate(sb.GEN_SK, N, 1.0, seed=5))
        fresh = ctx.solve(p_real, **kw)
        p_real.close()
    assert np.array_equal(after.spins, fresh.spins) and np.array_equal(after.energy, fresh.energy)
