"""GPU tests of the low-level operator boundary (sb_step, sb_init_state,
sb_asb_prologue, sb_energy) on caller-owned device buffers, the replacement of
the reference's UpdateTwice + cublasSgemv pair (SB/CUDA_GPU/m2_opt.cu:168-169)."""
import ctypes as C

import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu


def pad(N, m=64):
    return (N + m - 1) // m * m


def fp8_e4m3_bits(v: np.ndarray) -> np.ndarray:
    """E4M3 encoding of integers |v| <= 16 (sign, 4-bit exponent bias 7, 3-bit mantissa)."""
    out = np.zeros(v.shape, np.uint8)
    nz = v != 0
    a = np.abs(v[nz]).astype(np.int64)
    e = np.floor(np.log2(a)).astype(np.int64)
    m = ((a << 3) >> e) & 7
    out[nz] = ((v[nz] < 0).astype(np.uint8) << 7) | ((e + 7) << 3).astype(np.uint8) | m.astype(np.uint8)
    return out


def bf16_bits(v: np.ndarray) -> np.ndarray:
    return (v.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)  # exact for small integers


class Dev:
    """Device buffers of one instance for sb_step (torch only allocates)."""

    def __init__(self, J, R, seed, init="uniform"):
        import torch
        self.torch = torch
        self.lib = sb.load()
        self.ctx = sb.Context(0)
        N = J.shape[0]
        self.N, self.R, self.ld = N, R, pad(N)
        Jp = np.zeros((N, self.ld), np.float32)
        Jp[:, :N] = J
        self.J32 = torch.from_numpy(Jp).cuda()
        self.Jbf = torch.from_numpy(bf16_bits(Jp).view(np.int16)).cuda()
        self.J8 = torch.from_numpy(fp8_e4m3_bits(Jp).view(np.int8)).cuda()
        self.X = torch.zeros((R, self.ld), device="cuda")
        self.Xn = torch.zeros((R, self.ld), device="cuda")
        self.Y = torch.zeros((R, self.ld), device="cuda")
        self.stream = torch.cuda.current_stream().cuda_stream
        rc = self.lib.sb_init_state(self.ctx.handle, sb.INITS[init], seed, N, R, self.ld, self.X.data_ptr(),
                                    self.Y.data_ptr(), self.stream)
        assert rc == 0

    def step(self, precision, jfmt, coeffs):
        J = {sb.J_FP32: self.J32, sb.J_BF16: self.Jbf, sb.J_FP8: self.J8}[jfmt]
        rc = self.lib.sb_step(self.ctx.handle, precision, J.data_ptr(), jfmt, self.N, self.ld, self.X.data_ptr(),
                              self.Xn.data_ptr(), self.Y.data_ptr(), self.R, self.ld, C.byref(coeffs), self.stream)
        assert rc == 0, self.lib.sb_last_error_message()
        self.X, self.Xn = self.Xn, self.X

    def state(self):
        self.torch.cuda.synchronize()
        return self.X.cpu().numpy()[:, :self.N], self.Y.cpu().numpy()[:, :self.N]

    def close(self):
        self.torch.cuda.synchronize()
        self.ctx.close()


def run_steps(dev, params, c0, precision, jfmt):
    for k in range(params.steps):
        dev.step(precision, jfmt, sb.step_coeffs(params, c0, k))
    return dev.state()


@pytest.mark.parametrize("variant", ["bsb", "dsb"])
def test_sb_step_matches_solve(variant):
    N, R, steps = 150, 40, 60
    J = sb.generate(sb.GEN_KN, N, 1.0, 5)
    v = sb.VARIANTS[variant]
    params = sb.make_params(variant=v, precision=sb.PREC_SPLIT, steps=steps, replicas=R, init=sb.INIT_UNIFORM,
                            seed=4, dt=0.5)
    c0 = sb.c0_auto(v, N, 1.0, sb.instance_stats(J)[0])
    with sb.Context(0) as ctx:
        prob = ctx.problem(J)
        for prec in ("fp32", "split"):
            sol = ctx.solve(prob, state=True, variant=v, precision=prec, steps=steps, replicas=R, init="uniform",
                            seed=4, dt=0.5)
            for jfmt in ((sb.J_FP32,) if prec == "fp32" else (sb.J_BF16, sb.J_FP8)):
                dev = Dev(J, R, 4)
                x, y = run_steps(dev, params, c0, sb.PRECISIONS[prec], jfmt)
                dev.close()
                assert np.array_equal(x, sol.x) and np.array_equal(y, sol.y), (prec, jfmt)
        prob.close()


def test_split_fp8_and_bf16_j_identical_for_integer_j():
    # FP8 E4M3 J widened by v_cvt_scalef32_pk_bf16_fp8 must equal BF16 J exactly.
    N, R = 333, 24
    rng = np.random.default_rng(3)
    J = rng.integers(-16, 17, (N, N)).astype(np.float32)
    J = np.triu(J, 1)
    J = J + J.T
    params = sb.make_params(variant=sb.BSB, precision=sb.PREC_SPLIT, steps=25, replicas=R, dt=0.5)
    a = Dev(J, R, 9)
    xa, ya = run_steps(a, params, 0.01, sb.PREC_SPLIT, sb.J_BF16)
    a.close()
    b = Dev(J, R, 9)
    xb, yb = run_steps(b, params, 0.01, sb.PREC_SPLIT, sb.J_FP8)
    b.close()
    assert np.array_equal(xa, xb) and np.array_equal(ya, yb)


def test_asb_prologue_and_steps_match_oracle(orc):
    N, R, steps = 96, 7, 40
    J = sb.generate(sb.GEN_KN, N, 1.0, 8)
    params = sb.make_params(variant=sb.ASB, steps=steps, replicas=R, init=sb.INIT_REFERENCE, seed=0)
    c0 = sb.c0_auto(sb.ASB, N, 1.0, sb.instance_stats(J)[0])
    dev = Dev(J, R, 0, init="reference")
    rc = dev.lib.sb_asb_prologue(dev.ctx.handle, C.byref(params), N, R, dev.ld, dev.X.data_ptr(), dev.Y.data_ptr(),
                                 dev.stream)
    assert rc == 0
    x, y = run_steps(dev, params, c0, sb.PREC_ORDERED, sb.J_FP32)
    dev.close()
    xo, yo, _ = orc.solve_f32(orc.ASB, J, R, steps, init="reference", seed=0, dt=1.0)
    assert np.array_equal(x, xo[:, :N], equal_nan=True) and np.array_equal(y, yo[:, :N], equal_nan=True)


def test_sb_energy_on_device_buffers(orc):
    import torch
    N, R = 210, 19
    J = sb.generate(sb.GEN_ER, N, 0.3, 2)
    dev = Dev(J, R, 6)
    H = torch.zeros(R, dtype=torch.float64, device="cuda")
    rc = dev.lib.sb_energy(dev.ctx.handle, dev.J32.data_ptr(), N, dev.ld, dev.X.data_ptr(), R, dev.ld,
                           H.data_ptr(), dev.stream)
    assert rc == 0
    x, _ = dev.state()
    dev.close()
    s = orc.spins(np.ascontiguousarray(x), N)
    assert np.array_equal(H.cpu().numpy().astype(np.int64), orc.energy_i64(J, s))


def test_sb_step_rejects_bad_layouts():
    import torch
    lib = sb.load()
    with sb.Context(0) as ctx:
        buf = torch.zeros(64 * 64, device="cuda")
        c = sb.StepCoeffs()
        c.variant = sb.BSB
        p = buf.data_ptr()
        # ld not a multiple of 64 / smaller than N / FP32 path with an FP8 J
        assert lib.sb_step(ctx.handle, 0, p, 0, 60, 60, p, p, p, 1, 64, C.byref(c), None) == 1
        assert lib.sb_step(ctx.handle, 0, p, 0, 70, 64, p, p, p, 1, 64, C.byref(c), None) == 1
        assert lib.sb_step(ctx.handle, 0, p, 2, 60, 64, p, p, p, 1, 64, C.byref(c), None) == 1


def test_sb_field_paths(orc):
    N, R = 257, 6
    J = sb.generate(sb.GEN_KN, N, 1.0, 12)
    x, _ = orc.init_state("uniform", N, R, seed=3)
    xs = x[:, :N].copy()
    F32o = orc.field_f32(J, x)[:, :N]
    F64 = orc.field_f64(J, x.astype(np.float64))[:, :N]
    bound = N * 2.0**-24 * (np.abs(xs).astype(np.float64) @ np.abs(J).T.astype(np.float64))
    with sb.Context(0) as ctx:
        prob = ctx.problem(J)
        assert np.array_equal(ctx.field(prob, xs, "ordered", "bsb"), F32o)
        assert np.all(np.abs(ctx.field(prob, xs, "fp32", "bsb") - F64) <= 2 * bound)
        assert np.all(np.abs(ctx.field(prob, xs, "split", "bsb") - F64) <= 2 * bound + 2.0**-16 * bound / (N * 2.0**-24))
        exact = (np.where(xs > 0, 1, -1).astype(np.int64) @ J.astype(np.int64).T).astype(np.float32)
        for prec in ("fp32", "split", "ordered"):
            assert np.array_equal(ctx.field(prob, xs, prec, "dsb"), exact), prec
        # Amplitudes beyond the wall: no variant clamps or zeroes the field.
        big = 3.0 * xs
        for var in ("asb", "bsb", "asb_m1"):
            assert np.array_equal(ctx.field(prob, big, "ordered", var),
                                  orc.field_f32(J, np.ascontiguousarray(3.0 * x))[:, :N]), var
            assert np.allclose(ctx.field(prob, big, "fp32", var), 3.0 * F64, rtol=1e-5, atol=1e-4), var
            assert np.allclose(ctx.field(prob, big, "split", var), 3.0 * F64, rtol=1e-4, atol=1e-3), var
        for prec in ("fp32", "split", "ordered"):
            assert np.array_equal(ctx.field(prob, big, prec, "dsb"), exact), prec
        prob.close()
