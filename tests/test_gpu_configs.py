"""The benchmark configurations run at their stated shapes, through the C ABI,
against the CPU oracle (the reference's solver loop, SB/CUDA_GPU/m2_opt.cu:152-202,
and its trial loop over replicas, :152-154):

  C1  ER N=800 p=0.06 +-1 (J = -W), R=1, 1000 bSB steps, dt=1.25: the thin
      R = 1 path is bit-exact with the oracle's 64-lane order (LANES);
  C2  SK N=2000, FP32 MFMA path, 2000 steps: bit-exact with the oracle's
      MFMA order, and against the oracle's sequential-sum run (the reference's
      naive order) within 0.1 % in mean energy;
  C4  512 ER instances N ~ U[64, 4096], R=64, one batched launch per step:
      every instance equals its solo solve, sampled energies are exact against
      the oracle and dSB instances (integer J: exact in any order) equal the
      oracle bit for bit;
  C5  server requests at the size extremes (N=8000, R=256, BF16 and FP8 J)
      mixed with small ones: the same result alone and in the mix, exact or
      FP64 energies against the oracle, dSB and thin FP32 requests equal the
      oracle's trajectories.
"""
import os

import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu

GB = 1 << 30
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def ctx():
    c = sb.Context(0)
    yield c
    c.close()


def brute_cut(J, s):
    W = -np.triu(J.astype(np.float64), 1)
    return np.array([W[s[r][:, None] != s[r][None, :]].sum() for r in range(s.shape[0])])


# ---- C1 --------------------------------------------------------------------------

@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_c1_er800_r1_thin_path_bit_exact(ctx, orc, precision):
    N, steps, dt = 800, 1000, 1.25
    J = sb.generate(sb.GEN_ER, N, 0.06, 1)
    prob = ctx.problem(J)
    sol = ctx.solve(prob, state=True, variant="bsb", precision=precision, steps=steps, replicas=1,
                    init="uniform", seed=1, dt=dt)
    c0 = orc.c0_auto(orc.BSB, J)
    assert sol.c0 == c0 and abs(c0 - 0.0722) < 1e-3  # 0.5 / (sqrt(N) sigma_J), sigma_J ~ sqrt(0.06)
    x, y = orc.init_state("uniform", N, 1, seed=1)
    orc.run_f32(orc.BSB, J, x, y, 0, steps, steps, dt, 1.0, c0, order=orc.ORDER_LANES)
    assert np.array_equal(sol.x, x[:, :N]) and np.array_equal(sol.y, y[:, :N])
    s = orc.spins(x, N)
    assert np.array_equal(sol.spins, s)
    assert np.array_equal(sol.energy_i64, orc.energy_i64(J, s))
    assert np.array_equal(sol.cut, brute_cut(J, s))
    prob.close()


# ---- C2 --------------------------------------------------------------------------

def test_c2_sk2000_fp32_against_oracle(ctx, orc):
    # Bit for bit against the oracle's MFMA order (the FP32 tiles' own
    # summation order), and statistically against the reference's sequential
    # order (Bifurcation_m2.cu.c:93-97).
    N, R, steps = 2000, 32, 2000
    J = sb.generate(sb.GEN_SK, N, 1.0, 2)
    prob = ctx.problem(J)
    sol = ctx.solve(prob, state=True, variant="bsb", precision="fp32", steps=steps, replicas=R, init="uniform",
                    seed=2, dt=1.0)
    c0 = orc.c0_auto(orc.BSB, J)
    assert sol.c0 == c0
    x, y = orc.init_state("uniform", N, R, seed=2)
    orc.run_f32(orc.BSB, J, x, y, 0, steps, steps, 1.0, 1.0, c0, nthreads=THREADS, order=orc.ORDER_MFMA)
    assert np.array_equal(sol.x, x[:, :N]) and np.array_equal(sol.y, y[:, :N])
    assert np.allclose(sol.energy, orc.energy_f64(J, sol.spins), rtol=1e-12, atol=1e-10)
    xq, yq = orc.init_state("uniform", N, R, seed=2)
    orc.run_f32(orc.BSB, J, xq, yq, 0, steps, steps, 1.0, 1.0, c0, nthreads=THREADS)
    so = orc.spins(xq, N)
    Ho = orc.energy_f64(J, so)
    same = np.all(sol.spins == so, axis=1)
    gap = abs(sol.energy.mean() - Ho.mean()) / abs(Ho.mean())
    print(f"C2 SK N={N} R={R} {steps} steps: bit-exact with the MFMA order; against the sequential order "
          f"identical final spins in {same.sum()}/{R} replicas, mean energy gap {gap:.2e}")
    assert np.allclose(sol.energy[same], Ho[same], rtol=1e-12, atol=1e-10)
    assert gap <= 1e-3
    # Measured on MI355X: 29 of 32 replicas end with identical spins.
    assert same.sum() >= 28, same.sum()
    prob.close()


# ---- C4 --------------------------------------------------------------------------

def test_c4_batched_512_instances(orc):
    n, R, steps = 512, 64, 20
    rng = np.random.default_rng(4)
    Ns = rng.integers(64, 4097, n)
    # every 64th instance runs dSB: its integer field is exact in any order, so
    # the batched trajectory must equal the oracle's bit for bit
    variants = ["dsb" if i % 64 == 0 else "bsb" for i in range(n)]
    with sb.Context(0) as ctx:
        probs = [ctx.generate(sb.GEN_ER, int(N), 0.05, 1000 + i) for i, N in enumerate(Ns)]
        params = [sb.make_params(variant=v, precision="split", steps=steps, replicas=R, init="uniform", seed=i,
                                 dt=1.0) for i, v in enumerate(variants)]
        sols, st = ctx.batch_solve(probs, params, spins=True, state=True)
        assert st.graph_used == 1 and st.workgroups >= 256
        for i, p in enumerate(probs):
            solo = ctx.solve(p, state=True, variant=variants[i], precision="split", steps=steps, replicas=R,
                             init="uniform", seed=i, dt=1.0)
            assert np.array_equal(sols[i].x, solo.x) and np.array_equal(sols[i].y, solo.y), i
            assert np.array_equal(sols[i].energy_i64, solo.energy_i64), i
        order = np.argsort(Ns)
        sample = sorted(set(order[[0, n // 4, n // 2, 3 * n // 4, n - 1]].tolist()) | {0, 64, 128})
        for i in sample:
            J = probs[i].read_j()
            s = sols[i]
            assert s.exact
            assert np.array_equal(s.energy_i64, orc.energy_i64(J, s.spins)), i
            assert np.array_equal(s.cut, brute_cut(J, s.spins)), i
            N = int(Ns[i])
            if variants[i] == "dsb":
                x, y = orc.init_state("uniform", N, 8, seed=i)
                orc.run_f32(orc.DSB, J, x, y, 0, steps, steps, 1.0, 1.0, s.c0, nthreads=THREADS)
                assert np.array_equal(s.x[:8], x[:, :N]) and np.array_equal(s.y[:8], y[:, :N]), i
            else:
                # bSB split instances follow the oracle's BF16-MFMA model bit
                # for bit (two sampled replicas).
                x0, y0 = orc.init_state("uniform", N, R, seed=i)
                pk = [0, R - 1]
                x, y = np.ascontiguousarray(x0[pk]), np.ascontiguousarray(y0[pk])
                orc.run_f32(orc.BSB, J, x, y, 0, steps, steps, 1.0, 1.0, s.c0, nthreads=THREADS,
                            order=orc.ORDER_SPLIT)
                assert np.array_equal(s.x[pk], x[:, :N]) and np.array_equal(s.y[pk], y[:, :N]), i
        for p in probs:
            p.close()


# ---- C5 --------------------------------------------------------------------------

# (kind, N, density, generator seed, variant, R, steps)
C5_REQS = [
    (sb.GEN_SK_DILUTED, 8000, 0.6, 11, "bsb", 256, 30),   # BF16 J
    (sb.GEN_ER, 8000, 0.4, 12, "bsb", 256, 30),           # FP8 J
    (sb.GEN_ER, 100, 0.9, 13, "dsb", 32, 60),
    (sb.GEN_SK_DILUTED, 150, 0.01, 14, "bsb", 64, 45),
    (sb.GEN_ER, 700, 0.02, 15, "dsb", 128, 40),
    (sb.GEN_SK_DILUTED, 2500, 1.0, 16, "bsb", 32, 35),
]


def c5_submit(srv, spec):
    kind, N, p, gseed, v, R, steps = spec
    return srv.submit_generated(kind, N, p, gseed, variant=v, precision="split", steps=steps, replicas=R,
                                init="uniform", seed=gseed, dt=1.0)


def test_c5_server_size_extremes(ctx, orc):
    with sb.Server(ctx, 64 * GB, chunk_steps=10, precision="split") as srv:
        alone = [srv.wait(c5_submit(srv, spec)) for spec in C5_REQS]
        tickets = [c5_submit(srv, spec) for spec in C5_REQS]
        mixed = [srv.wait(t) for t in tickets]
        assert srv.stats().max_active_seen >= 2
    for i, (a, m) in enumerate(zip(alone, mixed)):
        assert np.array_equal(a.spins, m.spins), i
        assert np.array_equal(a.energy, m.energy), i
        assert np.array_equal(a.cut, m.cut), i
    for i, spec in enumerate(C5_REQS):
        kind, N, p, gseed, v, R, steps = spec
        prob = ctx.generate(kind, N, p, gseed)
        J = prob.read_j()
        prob.close()
        r = alone[i]
        pick = [0, 1, R // 2, R - 1]
        if kind == sb.GEN_ER:
            assert r.exact
            assert np.array_equal(r.energy_i64[pick], orc.energy_i64(J, r.spins[pick])), i
        else:
            assert not r.exact
            assert np.allclose(r.energy[pick], orc.energy_f64(J, r.spins[pick]), rtol=1e-12, atol=1e-9), i
        if v == "dsb":
            x, y = orc.init_state("uniform", N, R, seed=gseed)
            orc.run_f32(orc.DSB, J, x, y, 0, steps, steps, 1.0, 1.0, r.c0, nthreads=THREADS)
            assert np.array_equal(r.spins, orc.spins(x, N)), i
        else:
            # bSB in the split form: two sampled replicas follow the oracle's
            # model of the BF16 MFMA bit for bit (a replica depends only on itself).
            x0, y0 = orc.init_state("uniform", N, R, seed=gseed)
            pk = [0, R - 1]
            x, y = np.ascontiguousarray(x0[pk]), np.ascontiguousarray(y0[pk])
            orc.run_f32(orc.BSB, J, x, y, 0, steps, steps, 1.0, 1.0, r.c0, nthreads=THREADS, order=orc.ORDER_SPLIT)
            assert np.array_equal(r.spins[pk], orc.spins(x, N)), i


def test_server_fp32_requests_equal_oracle(ctx, orc):
    # R <= 4 requests run the thin kernels inside the batched launch, larger R
    # the FP32 MFMA tiles; their trajectories are the oracle's 64-lane and MFMA
    # orders bit for bit, in any mix.
    specs = [(sb.GEN_ER, 800, 0.06, 21, "bsb", 1, 300), (sb.GEN_SK, 1500, 1.0, 22, "bsb", 3, 200),
             (sb.GEN_ER, 600, 0.3, 23, "bsb", 64, 150), (sb.GEN_KN, 333, 1.0, 24, "dsb", 2, 250)]
    with sb.Server(ctx, 16 * GB, chunk_steps=20, precision="fp32") as srv:
        tickets = [srv.submit_generated(k, N, p, g, variant=v, precision="fp32", steps=s, replicas=R,
                                        init="uniform", seed=g, dt=1.0) for k, N, p, g, v, R, s in specs]
        res = [srv.wait(t) for t in tickets]
    for (k, N, p, g, v, R, s), r in zip(specs, res):
        prob = ctx.generate(k, N, p, g)
        J = prob.read_j()
        prob.close()
        x, y = orc.init_state("uniform", N, R, seed=g)
        var = orc.BSB if v == "bsb" else orc.DSB
        order = orc.ORDER_LANES if R <= 4 else orc.ORDER_MFMA
        orc.run_f32(var, J, x, y, 0, s, s, 1.0, 1.0, r.c0, nthreads=THREADS, order=order)
        assert np.array_equal(r.spins, orc.spins(x, N)), (N, R)
