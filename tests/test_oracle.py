"""Pins of the CPU restatement (oracle/sb_oracle.c) against what fixes it
independently: known-answer values, the reference's recorded trajectory data,
closed forms, invariants and brute force.  CPU only."""
import numpy as np
import pytest

RNG = np.random.default_rng(12345)


# ---- R2 xorshift32 (SB/CUDA_GPU/m2_opt.cu:38-46) ----------------------------

def test_xorshift32_known_answers(orc):
    # Marsaglia (2003), "Xorshift RNGs", xor32 with y = 2463534242.
    assert orc.xorshift32(2463534242) == 723471715
    # The reference's spin-1 stream (seed st_1 = 1), cf. the recorded pos.txt.
    assert orc.xorshift32(1) == 270369
    assert orc.xorshift32(270369) == 67634689
    # State 0 is a fixed point: the reference's spin 0 always starts at 0.
    assert orc.xorshift32(0) == 0


# ---- R3 init (m2_opt.cu:48-56) ------------------------------------------------

def test_reference_init_matches_recorded_trajectory(orc, golden_dir):
    pos = np.loadtxt(f"{golden_dir}/bifurcation_pos.txt")
    x, y = orc.init_state("reference", N=4, R=1)
    assert x[0, 0] == 0.0 and y[0, 0] == 0.0  # spin 0: state 0
    # pos[0] = x0 + DeltaT * y0 for spin 1 with DeltaT = 0.05 (x-first update);
    # printed with 6 decimals.
    assert abs(float(x[0, 1]) + 0.05 * float(y[0, 1]) - pos[0]) <= 5e-7
    assert x[0, 1] == np.float32(270369 / 2**32 / 10.0)


def test_reference_init_trials_continue_the_stream(orc):
    # Trial r takes draws 2r+1, 2r+2 of each spin's stream (state never reset).
    x, y = orc.init_state("reference", N=3, R=3)
    st = 2
    for r in range(3):
        st = orc.xorshift32(st)
        assert x[r, 2] == np.float32(orc.ref_draw_value(st))
        st = orc.xorshift32(st)
        assert y[r, 2] == np.float32(orc.ref_draw_value(st))
    assert np.all((x[:, :3] >= 0) & (x[:, :3] < 0.1))


def test_uniform_init_range_and_keying(orc):
    x, y = orc.init_state("uniform", N=100, R=8, seed=7)
    v = np.concatenate([x[:, :100].ravel(), y[:, :100].ravel()])
    assert v.min() >= -0.1 and v.max() < 0.1
    assert abs(v.mean()) < 0.01
    assert np.all(x[:, 100:] == 0)
    # value depends only on (seed, replica, spin, component)
    x2, _ = orc.init_state("uniform", N=120, R=3, seed=7)
    assert np.array_equal(x2[:3, :100], x[:3, :100])


# ---- R4 aSB local update (m2_opt.cu:58-75) --------------------------------------

def test_asb_x_first_recurrence_matches_recorded(orc, golden_dir):
    pos = np.loadtxt(f"{golden_dir}/bifurcation_pos.txt")
    spd = np.loadtxt(f"{golden_dir}/bifurcation_speed.txt")
    # The recorded run (Bifurcation_m1, DeltaT=0.05) prints x after the x
    # update and y after the y update: x_{k+1} = x_k + a0 dt y_k.
    worst = 0.0
    for k in range(len(pos) - 1):
        xn, _ = orc.asb_local_f32(pos[k], spd[k], 0.0, 1.0, 1.0, 0.05, 1)
        worst = max(worst, abs(xn - pos[k + 1]))
    assert worst <= 1.2e-6  # 6-decimal printing


def test_asb_kerr_fixed_point(orc):
    # Beyond the bifurcation (p > a0) the uncoupled oscillator has fixed points
    # x* = +-sqrt((p - a0)/K), y = 0; a sign error in the Kerr or detuning
    # term moves the state away.
    a0, p, K = 0.5, 0.9, 1.0
    xs = np.float32(np.sqrt((p - a0) / K))
    x, y = float(xs), 0.0
    for _ in range(200):
        x, y = orc.asb_local_f32(x, y, p, a0, K, 0.25, 2)
    assert abs(x - xs) < 1e-5 and abs(y) < 1e-5
    # and x = 0 is unstable there: a tiny push swings out towards the
    # turning point x^2 = 2 (p - a0) / K of the conserved energy
    x, y, peak = 1e-4, 0.0, 0.0
    for _ in range(400):
        x, y = orc.asb_local_f32(x, y, p, a0, K, 0.25, 2)
        peak = max(peak, abs(x))
    assert 0.8 < peak < 0.95


# ---- R9 bSB update (Goto et al. 2021) ------------------------------------------

def test_bsb_symplectic_invariant(orc):
    # With constant coefficients and F = 0, y-first symplectic Euler conserves
    # Q = w x^2 + a0 y^2 - h a0 w x y (w = a0 - a, h = dt).  An x-first order
    # or a sign error in the restoring term breaks it at O(h).
    a0, w, h = 1.0, 0.7, 0.3
    alpha, gamma = np.float32(h * w), np.float32(h * a0)
    x, y = 0.05, 0.02
    Q0 = w * x * x + a0 * y * y - h * a0 * w * x * y
    for _ in range(2000):
        x, y = orc.bsb_update_f32(0.0, alpha, 0.0, gamma, x, y)
        Q = w * x * x + a0 * y * y - h * a0 * w * x * y
        assert abs(Q - Q0) <= 1e-4 * Q0


def test_bsb_two_spin_normal_modes(orc):
    # N = 2, J12 = J: modes u = x1 + x2 and v = x1 - x2 decouple with restoring
    # constants w -+ c0 J.  Check each mode's invariant (FP64 run; a(t) moves
    # by 5e-5 over the 50 steps, which bounds the drift).
    J = np.array([[0.0, 1.0], [1.0, 0.0]], np.float32)
    a0, c0, dt, steps = 1.0, 0.2, 0.1, 10**6  # a ~ 0 over the steps run
    x = np.zeros((1, 64), np.float64)
    y = np.zeros((1, 64), np.float64)
    x[0, :2] = [0.01, 0.004]
    y[0, :2] = [0.0, 0.002]
    def inv(x, y, s):
        u = x[0, 0] + s * x[0, 1]
        q = y[0, 0] + s * y[0, 1]
        w = a0 - s * c0
        return w * u * u + a0 * q * q - dt * a0 * w * u * q
    Iu, Iv = inv(x, y, 1), inv(x, y, -1)
    orc.run_f64(orc.BSB, J, x, y, 0, 50, steps, dt, a0, c0)
    assert abs(inv(x, y, 1) - Iu) < 2e-4 * abs(Iu)
    assert abs(inv(x, y, -1) - Iv) < 2e-4 * abs(Iv)
    # with the modes' constants swapped (a sign error on c0 J) it is not kept
    x2 = np.zeros((1, 64)); y2 = np.zeros((1, 64))
    x2[0, :2] = [0.01, 0.004]; y2[0, :2] = [0.0, 0.002]
    orc.run_f64(orc.BSB, -J, x2, y2, 0, 50, steps, dt, a0, c0)
    assert abs(inv(x2, y2, 1) - Iu) > 1e-2 * abs(Iu)


def test_bsb_wall_and_odd_symmetry(orc):
    N = 48
    J = RNG.standard_normal((N, N)).astype(np.float32)
    J = np.triu(J, 1)
    J = (J + J.T).astype(np.float32)
    x, y = orc.init_state("uniform", N, 4, seed=3)
    xm, ym = -x, -y
    c0 = orc.c0_auto(orc.BSB, J)
    for k in range(0, 300, 30):
        orc.run_f32(orc.BSB, J, x, y, k, k + 30, 300, 0.5, 1.0, c0)
        orc.run_f32(orc.BSB, J, xm, ym, k, k + 30, 300, 0.5, 1.0, c0)
        assert np.all(np.abs(x) <= 1.0)
        # clamped spins have y = 0 after the step that hit the wall
        assert np.array_equal(xm, -x) and np.array_equal(ym, -y)
    assert np.mean(np.abs(x[:, :N]) == 1.0) > 0.5  # bifurcated to the walls


def test_dsb_field_is_exact_integer(orc):
    N = 70
    J = RNG.integers(-16, 17, (N, N)).astype(np.float32)
    J = np.triu(J, 1)
    J = J + J.T
    x, _ = orc.init_state("uniform", N, 5, seed=1)
    F = orc.field_f32(J, x, use_sign=True)
    s = np.where(x[:, :N] > 0, 1, -1).astype(np.int64)
    exact = s @ J.astype(np.int64).T
    assert np.array_equal(F[:, :N].astype(np.int64), exact)


def test_field_f32_error_bound(orc):
    N = 300
    J = (RNG.standard_normal((N, N)) / np.sqrt(N)).astype(np.float32)
    J = np.triu(J, 1)
    J = J + J.T
    x, _ = orc.init_state("uniform", N, 3, seed=2)
    F32 = orc.field_f32(J, x)[:, :N].astype(np.float64)
    F64 = orc.field_f64(J, x.astype(np.float64))[:, :N]
    ref = x[:, :N].astype(np.float64) @ J.astype(np.float64).T
    assert np.allclose(F64, ref, rtol=0, atol=1e-12)
    bound = N * 2.0**-24 * (np.abs(x[:, :N]).astype(np.float64) @ np.abs(J).T.astype(np.float64))
    assert np.all(np.abs(F32 - F64) <= bound)


@pytest.mark.parametrize("order", ["lanes", "mfma"])
def test_reordered_field_pins(orc, order):
    # Integer couplings times integer-valued state are exact in any order, so
    # the lane-order and MFMA-order fields must equal the integer product: a
    # dropped, repeated or misplaced group, a wrong butterfly step or a
    # permutation that misses a position fails.  The sizes cover one partial
    # group, whole lane sweeps, partial last sweeps and partial 8- and 32-blocks.
    ORD = orc.ORDER_LANES if order == "lanes" else orc.ORDER_MFMA
    for N in (5, 16, 33, 70, 1023, 1025, 2100):
        J = RNG.integers(-16, 17, (N, N)).astype(np.float32)
        J = np.triu(J, 1)
        J = J + J.T
        ld = orc.pad_to(N)
        xi = np.zeros((3, ld), np.int64)
        xi[:, :N] = RNG.integers(-3, 4, (3, N))
        x = xi.astype(np.float32)
        exact = xi[:, :N] @ J.astype(np.int64).T
        F = orc.field_f32(J, x, order=ORD)
        assert np.array_equal(F[:, :N].astype(np.int64), exact), N
        s = np.where(x[:, :N] > 0, 1, -1).astype(np.int64)
        Fs = orc.field_f32(J, x, use_sign=True, order=ORD)
        assert np.array_equal(Fs[:, :N].astype(np.int64), s @ J.astype(np.int64).T), N
    # Real couplings: within the FP32 bound of the FP64 field, and a different
    # rounding sequence from the sequential chain.
    N = 1500
    J = (RNG.standard_normal((N, N)) / np.sqrt(N)).astype(np.float32)
    J = np.triu(J, 1)
    J = J + J.T
    x, _ = orc.init_state("uniform", N, 3, seed=4)
    Fl = orc.field_f32(J, x, order=ORD)[:, :N]
    Fq = orc.field_f32(J, x)[:, :N]
    F64 = orc.field_f64(J, x.astype(np.float64))[:, :N]
    bound = N * 2.0**-24 * (np.abs(x[:, :N]).astype(np.float64) @ np.abs(J).T.astype(np.float64))
    assert np.all(np.abs(Fl - F64) <= bound)
    assert np.any(Fl != Fq)


@pytest.mark.parametrize("order", ["lanes", "mfma"])
def test_reordered_run_equals_sequential_for_exact_fields(orc, order):
    # dSB with integer J: every field is an exact integer, so the orders give
    # the same trajectory bit for bit.
    ORD = orc.ORDER_LANES if order == "lanes" else orc.ORDER_MFMA
    N, R, steps = 300, 4, 60
    J = RNG.integers(-2, 3, (N, N)).astype(np.float32)
    J = np.triu(J, 1)
    J = J + J.T
    xa, ya, _ = orc.solve_f32(orc.DSB, J, R, steps, seed=5, dt=0.5)
    xb, yb, _ = orc.solve_f32(orc.DSB, J, R, steps, seed=5, dt=0.5, order=ORD)
    assert np.array_equal(xa, xb) and np.array_equal(ya, yb)


def test_mfma_order_chain_on_crafted_rows(orc):
    # The MFMA order is one fmaf chain over the positions 0,4,1,5,2,6,3,7 of
    # each 8-block.  A row whose only nonzero products sit at k = 1 and k = 4
    # with values that round differently in the two orders: 1 + 2^-24 + 2^-24
    # is 1 in either fmaf chain that adds them one at a time, while a row with
    # a large product at k = 4 first and a cancelling one at k = 1 keeps the
    # rounding error of the k = 4 product only if k = 4 comes first.
    N = 8
    J = np.zeros((N, N), np.float32)
    a = np.float32(1.0 + 2.0**-23)
    J[0, 4] = a
    J[0, 1] = -a
    J[4, 0] = a
    J[1, 0] = -a
    x = np.zeros((1, 64), np.float32)
    x[0, 4] = a
    x[0, 1] = a
    F = orc.field_f32(J, x, order=orc.ORDER_MFMA)
    # k = 4 first: fmaf(a, a, 0) = fl(a^2); then fmaf(a, -a, fl(a^2)) = fl(a^2) - a^2 = -2^-46
    assert F[0, 0] == np.float32(-(2.0**-46))
    Fs = orc.field_f32(J, x, order=orc.ORDER_SEQ)
    # sequential: k = 1 first gives +2^-46
    assert Fs[0, 0] == np.float32(2.0**-46)


def _bf16_from_bits(bits):
    return (np.asarray(bits, np.uint32) << 16).view(np.float32)


def test_mfma_bf16_model_matches_hardware_vectors(orc, golden_dir):
    # 10,240 outputs of v_mfma_f32_32x32x16_bf16 measured on an MI355X with
    # random operands (dense, wide exponent spread, two products only, one
    # product per half, cancelling pairs; tests/golden/mfma_bf16_probe.hip):
    # the oracle's model reproduces every one bit for bit; exact rounding of
    # c + sum a b misses more than a thousand of them.
    z = np.load(f"{golden_dir}/mfma_bf16_cases.npz")
    a, b, c, d = _bf16_from_bits(z["a_bits"]), _bf16_from_bits(z["b_bits"]), z["c"], z["d"]
    model = np.array([orc.mfma_bf16(c[q], a[q], b[q]) for q in range(len(c))], np.float32)
    assert np.array_equal(model.view(np.uint32), d.view(np.uint32))
    exact = np.array([np.float32(np.sum(a[q].astype(np.float64) * b[q].astype(np.float64)) + np.float64(c[q]))
                      for q in range(len(c))], np.float32)
    assert np.count_nonzero(exact != d) > 1000
    for t in range(5):
        assert np.any(z["batch_type"] == t)


def test_mfma_bf16_model_crafted_cases(orc):
    # The measured rules, one at a time (each case fails if the rule is dropped):
    def mf(c, prods):
        a = np.zeros(16, np.float32)
        b = np.zeros(16, np.float32)
        for k, v in prods:
            a[k], b[k] = v, 1.0
        return float(orc.mfma_bf16(np.float32(c), a, b))
    # products within one step are added exactly: 1 + 2^-24 + 2^-24 = 1 + 2^-23
    assert mf(1.0, [(0, 2.0**-24), (1, 2.0**-24)]) == 1.0 + 2.0**-23
    # the two steps round separately: k = 0 and k = 8 give 1 (each a tie to even)
    assert mf(1.0, [(0, 2.0**-24), (8, 2.0**-24)]) == 1.0
    # products 25 binades below the largest product are truncated away
    assert mf(0.0, [(0, 2.0**24), (1, 1.0), (2, -2.0**24)]) == 1.0
    assert mf(0.0, [(0, 2.0**25), (1, 1.0), (2, -2.0**25)]) == 0.0
    # the accumulator is floored at 2^(E_p - 24) as well
    assert mf(2.0**-24, [(0, 1.0), (1, -1.0)]) == 2.0**-24
    assert mf(2.0**-25, [(0, 1.0), (1, -1.0)]) == 0.0
    # with a larger accumulator the sum keeps 31 binades below it
    assert mf(1.0, [(0, 2.0**-24), (1, 2.0**-31)]) == 1.0 + 2.0**-23
    assert mf(1.0, [(0, 2.0**-24), (1, 2.0**-32)]) == 1.0
    # ... floored: a tiny negative product still pulls a tie down
    assert mf(1.0 + 2.0**-23, [(0, 2.0**-24), (1, -2.0**-40)]) == 1.0 + 2.0**-23
    assert mf(1.0 + 2.0**-23, [(0, 2.0**-24)]) == 1.0 + 2.0**-22  # tie to even
    # no products: the accumulator is returned as is
    assert mf(0.1, []) == np.float32(0.1)


def test_split_field_matches_hardware_states(orc, golden_dir):
    # Six states on which the MI355X split tiles were run (sb_field): every
    # field value equals the oracle's.  Each state holds a field whose MFMA sum
    # crosses a binade as the low products are added, which pins the 32-bit
    # window to the sum's own exponent (a window relative to the larger of the
    # accumulator and the products misses one value in each).
    z = np.load(f"{golden_dir}/split_field_cases.npz")
    J, X, F = z["J"], z["X"], z["F"]
    N = J.shape[0]
    Xp = np.zeros((X.shape[0], orc.pad_to(N)), np.float32)
    Xp[:, :N] = X
    Fo = orc.field_f32(J, Xp, order=orc.ORDER_SPLIT)[:, :N]
    assert np.array_equal(Fo.view(np.uint32), F.view(np.uint32))


def test_split_field_within_bound_of_fp64(orc):
    # The split form's field on real J: BF16-rounded J times x = hi + lo, so it
    # is within the BF16 rounding of J (2^-9 relative per coupling) plus the
    # hi/lo residue (2^-17 |x|) and the MFMA's truncation of the FP64 field.
    N = 700
    J = (RNG.standard_normal((N, N)) / np.sqrt(N)).astype(np.float32)
    J = np.triu(J, 1)
    J = J + J.T
    x, _ = orc.init_state("uniform", N, 3, seed=9)
    F = orc.field_f32(J, x, order=orc.ORDER_SPLIT)[:, :N].astype(np.float64)
    F64 = orc.field_f64(J, x.astype(np.float64))[:, :N]
    ax = np.abs(x[:, :N]).astype(np.float64) @ np.abs(J).T.astype(np.float64)
    assert np.all(np.abs(F - F64) <= 2.0**-8 * ax + 1e-6)
    # with J exact in BF16 (integers) the error is the split residue and the truncation only
    Ji = RNG.integers(-16, 17, (N, N)).astype(np.float32)
    Ji = np.triu(Ji, 1)
    Ji = Ji + Ji.T
    Fi = orc.field_f32(Ji, x, order=orc.ORDER_SPLIT)[:, :N].astype(np.float64)
    Fi64 = orc.field_f64(Ji, x.astype(np.float64))[:, :N]
    axi = np.abs(x[:, :N]).astype(np.float64) @ np.abs(Ji).T.astype(np.float64)
    assert np.all(np.abs(Fi - Fi64) <= 2.0**-15 * axi)
    # dSB on integer J: sign products are integers, the split field is exact
    s = np.where(x[:, :N] > 0, 1, -1).astype(np.int64)
    Fs = orc.field_f32(Ji, x, use_sign=True, order=orc.ORDER_SPLIT)[:, :N]
    assert np.array_equal(Fs.astype(np.int64), s @ Ji.astype(np.int64).T)


# ---- R12 aSB-m1 (Bifurcation_m1.cu.c:55-92) ------------------------------------

def _sk(N, seed):
    rng = np.random.default_rng(seed)
    J = (rng.standard_normal((N, N)) / np.sqrt(N)).astype(np.float32)
    J = np.triu(J, 1)
    return (J + J.T).astype(np.float32)


def test_m1_fp64_equals_asb_one_substep(orc):
    # x += a0 y dt, then y += dt (-K x^3 + p x - a0 x) + xi dt F(x) is aSB with
    # one local sub-step and the kick on the same x: in FP64 the two code paths
    # (aSB's is pinned by the recorded trajectory) must agree to rounding.
    N, R, steps = 64, 3, 120
    J = _sk(N, 5)
    x0, y0 = orc.init_state("uniform", N, R, seed=3)
    xa, ya = x0.astype(np.float64), y0.astype(np.float64)
    xb, yb = xa.copy(), ya.copy()
    c0 = orc.c0_auto(orc.ASB, J)
    orc.run_f64(orc.ASB_M1, J, xa, ya, 0, steps, steps, 0.5, 1.0, c0, 1.0, 1)
    orc.run_f64(orc.ASB, J, xb, yb, 0, steps, steps, 0.5, 1.0, c0, 1.0, 1)
    assert np.max(np.abs(xa - xb)) < 1e-11 and np.max(np.abs(ya - yb)) < 1e-11
    assert np.max(np.abs(xa)) > 0.05  # the trajectory moved


def test_m1_fp32_follows_fp64(orc):
    # The FP32 restatement stays within a few ulp-scaled units of its FP64
    # trajectory; a dropped term or wrong sign would not.
    N, R, steps = 64, 3, 40
    J = _sk(N, 6)
    c0 = orc.c0_auto(orc.ASB_M1, J)
    x32, y32 = orc.init_state("uniform", N, R, seed=4)
    x64, y64 = x32.astype(np.float64), y32.astype(np.float64)
    orc.run_f32(orc.ASB_M1, J, x32, y32, 0, steps, steps, 0.5, 1.0, c0, 1.0, 1)
    orc.run_f64(orc.ASB_M1, J, x64, y64, 0, steps, steps, 0.5, 1.0, c0, 1.0, 1)
    assert np.max(np.abs(x32 - x64)) < 1e-5 and np.max(np.abs(y32 - y64)) < 1e-5
    # and it is its own rounding sequence, not aSB's
    xa, ya = orc.init_state("uniform", N, R, seed=4)
    orc.run_f32(orc.ASB, J, xa, ya, 0, steps, steps, 0.5, 1.0, c0, 1.0, 1)
    assert not np.array_equal(ya, y32)


def test_m1_single_step_known_values(orc):
    # One spin, J = 0, DeltaT = 0.5 (every product with DeltaT is exact):
    # x' = 0.05 + 0.02 * 0.5 = 0.06; y' = fl32(-x'^3 / 2 + p x' / 2 - x' / 2)
    # + fl32(F xi / 2) + y, accumulated in FP32.
    x1 = orc.m1_x_f32(np.float32(0.05), np.float32(0.02), 1.0, 0.5)
    assert x1 == np.float32(np.float32(0.05) + np.float32(0.01))
    x = np.float32(x1)
    ny = np.float32(-(np.float64(x * x * x)) * 0.5 + np.float64(np.float32(0.25) * x) * 0.5 - np.float64(x) * 0.5)
    ny = np.float32(ny + np.float32(np.float64(np.float32(0.3)) * 0.01565 * 0.5))
    ny = np.float32(ny + np.float32(0.02))
    assert orc.m1_y_f32(0.3, x, 0.02, 0.25, 1.0, 1.0, 0.5, 0.01565) == ny


# ---- external field h (next-step row: "i i w" lines) -----------------------------

def _augment(J, h):
    """The field-free model on N + 1 spins with J[i][N] = J[N][i] = h_i."""
    N = J.shape[0]
    Ja = np.zeros((N + 1, N + 1), np.float32)
    Ja[:N, :N] = J
    Ja[:N, N] = h
    Ja[N, :N] = h
    return Ja


def test_field_energy_equals_auxiliary_spin(orc):
    # H(s; J, h) = -sum_{i<j} J s s - sum h s is the field-free energy of the
    # augmented model with the extra spin at +1; its ground state (a global
    # minimum up to the all-flip symmetry) is the field model's.
    rng = np.random.default_rng(21)
    N = 11
    J = rng.integers(-3, 4, (N, N)).astype(np.float32)
    J = np.triu(J, 1)
    J = (J + J.T).astype(np.float32)
    h = rng.integers(-4, 5, N).astype(np.float32)
    Ja = _augment(J, h)
    S = np.where(rng.random((40, N)) < 0.5, 1, -1).astype(np.int8)
    Sa = np.concatenate([S, np.ones((40, 1), np.int8)], axis=1)
    assert np.array_equal(orc.energy_i64(J, S, h), orc.energy_i64(Ja, Sa))
    assert np.allclose(orc.energy_f64(J, S, h), orc.energy_f64(Ja, Sa), rtol=0, atol=1e-9)
    Hg, sg = orc.ground_state(Ja)
    sg = sg if sg[N] == 1 else -sg
    assert orc.energy_i64(J, sg[None, :N], h)[0] == Hg
    # J = 0: the ground energy is -sum |h|
    assert orc.ground_state(_augment(np.zeros((N, N), np.float32), h))[0] == -np.abs(h).sum()


def test_field_step_equals_auxiliary_spin_step(orc):
    # One step: the field F_i = fl32((J.x)_i + h_i) equals the augmented
    # sequential sum whose last term is h_i * 1, bit for bit (SEQ order).  bSB
    # and dSB take the field on the incoming state, where the extra spin is
    # still +1 (aSB moves x before its field).
    N, R = 90, 3
    rng = np.random.default_rng(22)
    J = (rng.standard_normal((N, N)) / np.sqrt(N)).astype(np.float32)
    J = np.triu(J, 1)
    J = (J + J.T).astype(np.float32)
    h = (0.3 * rng.standard_normal(N)).astype(np.float32)
    for variant in (orc.BSB, orc.DSB):
        x, y = orc.init_state("uniform", N, R, seed=2)
        xa, ya = orc.init_state("uniform", N + 1, R, seed=2)
        xa[:, :N], ya[:, :N] = x[:, :N], y[:, :N]
        xa[:, N] = 1.0
        ya[:, N] = 0.0
        orc.run_f32(variant, J, x, y, 0, 1, 50, 0.5, 1.0, 0.2, 1.0, 2, h=h)
        orc.run_f32(variant, _augment(J, h), xa, ya, 0, 1, 50, 0.5, 1.0, 0.2, 1.0, 2)
        assert np.array_equal(x[:, :N], xa[:, :N]) and np.array_equal(y[:, :N], ya[:, :N]), variant


def test_edgelist_field_lines(orc):
    text = "4\n0 1 1\n2 2 -3\n1 3 -1\n0 0 2\n"
    rc, J, h = orc.parse_edgelist_field(text)
    assert rc == 0 and list(h) == [2.0, 0.0, -3.0, 0.0]
    assert J[0, 1] == 1 and J[1, 3] == -1 and J[2, 2] == 0
    rc, _ = orc.parse_edgelist(text)
    assert rc != 0  # without a field output, "i i w" stays an error


# ---- R7 energy (m2_opt.cu:192-202) ----------------------------------------------

def _sym_int(N, lo, hi, rng):
    J = rng.integers(lo, hi + 1, (N, N)).astype(np.float32)
    J = np.triu(J, 1)
    return (J + J.T).astype(np.float32)


def test_energy_matches_quadratic_form(orc):
    J = _sym_int(11, -5, 5, RNG)
    s = RNG.choice([-1, 1], size=(20, 11)).astype(np.int8)
    H = orc.energy_i64(J, s)
    q = -0.5 * np.einsum("ri,ij,rj->r", s.astype(np.int64), J.astype(np.int64), s.astype(np.int64))
    assert np.array_equal(H, q.astype(np.int64))
    Hf = orc.energy_f64(J, s)
    assert np.array_equal(Hf, H.astype(np.float64))


def test_energy_invariants(orc):
    # all +1: H = -sum_{i<j} J_ij ; K_N +-1: H = P (mod 2), P = N(N-1)/2 pairs
    for N in (9, 10, 13):
        J = np.where(RNG.random((N, N)) < 0.5, -1.0, 1.0).astype(np.float32)
        J = np.triu(J, 1)
        J = (J + J.T).astype(np.float32)
        s = np.ones((1, N), np.int8)
        assert orc.energy_i64(J, s)[0] == -int(np.triu(J, 1).sum())
        P = N * (N - 1) // 2
        sr = RNG.choice([-1, 1], size=(16, N)).astype(np.int8)
        assert np.all(orc.energy_i64(J, sr) % 2 == P % 2)
    with pytest.raises(ValueError):
        orc.energy_i64(np.array([[0, 0.5], [0.5, 0]], np.float32), np.ones((1, 2), np.int8))


def test_ground_state_closed_forms(orc):
    # ferromagnet: all aligned, H = -P
    N = 10
    J = np.ones((N, N), np.float32) - np.eye(N, dtype=np.float32)
    H, s = orc.ground_state(J)
    assert H == -N * (N - 1) / 2 and abs(int(s.sum())) == N
    # even antiferromagnetic ring: every bond satisfied, H = -N
    N = 12
    J = np.zeros((N, N), np.float32)
    for i in range(N):
        J[i, (i + 1) % N] = J[(i + 1) % N, i] = -1
    H, s = orc.ground_state(J)
    assert H == -N
    # complete bipartite K_{a,b} with J = -1 across: max cut a*b
    a, b = 5, 7
    J = np.zeros((a + b, a + b), np.float32)
    J[:a, a:] = -1
    J[a:, :a] = -1
    H, s = orc.ground_state(J)
    assert H == -a * b


def test_bsb_reaches_easy_ground_states(orc):
    # bSB finds the max cut of an even ring and of K_{a,b}.
    for J in (
        np.roll(np.eye(16, dtype=np.float32), 1, 1) * -1 + np.roll(np.eye(16, dtype=np.float32), -1, 1) * -1,
        np.block([[np.zeros((6, 6)), -np.ones((6, 9))], [-np.ones((9, 6)), np.zeros((9, 9))]]).astype(np.float32),
    ):
        Hgs, _ = orc.ground_state(J)
        x, y, _ = orc.solve_f32(orc.BSB, J, R=8, steps=400, seed=5, dt=0.5)
        H = orc.energy_i64(J, orc.spins(x, J.shape[0]))
        assert H.min() == Hgs


# ---- R1 loader (m2_opt.cu:111-123) ------------------------------------------------

def test_edgelist_oracle(orc):
    rc, J = orc.parse_edgelist("4\n0 1 1\n1 2 -2\n2 3 3\n0 1 5\n")
    assert rc == 0
    assert J[0, 1] == 5 and J[1, 0] == 5 and J[1, 2] == -2 and J[3, 2] == 3
    assert np.array_equal(J, J.T) and np.all(np.diag(J) == 0)
    assert orc.parse_edgelist("3\n0 0 1\n")[0] == 1  # i == j
    assert orc.parse_edgelist("3\n0 3 1\n")[0] == 4  # out of range
    assert orc.parse_edgelist("3\n0 1\n")[0] == 3    # truncated
    rc, J = orc.parse_edgelist("3\n1 2 1\n", one_based=True)
    assert rc == 0 and J[0, 1] == 1


# ---- parameters -------------------------------------------------------------------

def test_reference_default_xi(orc):
    # m2_opt.cu:18: xi = 0.01565 "0.7*detune/(rho*sqrt(N))" for K2000 +-1.
    N = 2000
    J = np.ones((N, N), np.float32) - np.eye(N, dtype=np.float32)
    assert abs(orc.c0_auto(orc.ASB, J) - 0.01565) < 5e-6
    # coeff1 = DeltaT * xi as float (m2_opt.cu:148)
    assert orc.coeffs(orc.ASB, 0, 100, 1.0, 1.0, 0.01565)[1] == np.float32(0.01565)
    # amplitude = s / (float)STEP
    assert orc.coeffs(orc.ASB, 7, 20, 1.0, 1.0, 0.01565)[3] == np.float32(7) / np.float32(20)
