"""Several host threads sharing one context: every call takes the context's
lock, so concurrent solves (including first-time autotuning and graph capture)
return exactly what the same solves return one after another."""
import threading

import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu

# (generator, N, density, variant, precision, replicas, steps)
JOBS = [
    [(sb.GEN_SK, 300, 1.0, "bsb", "fp32", 40, 60), (sb.GEN_KN, 200, 1.0, "dsb", "split", 3, 80)],
    [(sb.GEN_KN, 500, 1.0, "bsb", "split", 96, 50), (sb.GEN_ER, 800, 0.06, "bsb", "fp32", 1, 100)],
    [(sb.GEN_SK_DILUTED, 700, 0.3, "asb", "fp32", 8, 40), (sb.GEN_SK, 150, 1.0, "bsb", "ordered", 5, 30)],
    [(sb.GEN_ER, 1000, 0.05, "dsb", "fp32", 130, 40), (sb.GEN_KN, 64, 1.0, "asb_m1", "split", 2, 70)],
]


def run_jobs(ctx, jobs, seed0):
    out = []
    for k, (kind, N, p, variant, precision, R, steps) in enumerate(jobs):
        prob = ctx.generate(kind, N, p, seed0 + k)
        sol = ctx.solve(prob, state=True, variant=variant, precision=precision, steps=steps, replicas=R,
                        init="uniform", seed=seed0 + k, dt=0.5)
        out.append((sol.x.copy(), sol.y.copy(), sol.energy.copy()))
        prob.close()
    return out


def test_concurrent_solves_on_one_context():
    with sb.Context(0) as ctx:
        results = [None] * len(JOBS)
        errors = []

        def worker(t):
            try:
                results[t] = run_jobs(ctx, JOBS[t], 100 * (t + 1))
            except Exception as e:  # reported below, in the main thread
                errors.append((t, repr(e)))

        threads = [threading.Thread(target=worker, args=(t,), daemon=True) for t in range(len(JOBS))]
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=240)
        assert not any(th.is_alive() for th in threads)
        assert not errors, errors
    with sb.Context(0) as ctx:
        for t, jobs in enumerate(JOBS):
            serial = run_jobs(ctx, jobs, 100 * (t + 1))
            for k, (a, b) in enumerate(zip(results[t], serial)):
                assert all(np.array_equal(u, v) for u, v in zip(a, b)), (t, k)


def test_server_and_direct_solves_share_a_context():
    # A server's worker thread and direct solves from this thread interleave on
    # one context; each returns what it returns alone.
    reqs = [(sb.GEN_ER, 600, 0.1, 31, "bsb", 64, 120), (sb.GEN_SK_DILUTED, 900, 0.5, 32, "bsb", 32, 90),
            (sb.GEN_KN, 300, 1.0, 33, "dsb", 128, 150), (sb.GEN_ER, 1500, 0.02, 34, "bsb", 8, 60)]

    def submit(srv, spec):
        kind, N, p, g, v, R, steps = spec
        return srv.submit_generated(kind, N, p, g, variant=v, precision="split", steps=steps, replicas=R,
                                    init="uniform", seed=g, dt=0.5)

    with sb.Context(0) as ctx:
        with sb.Server(ctx, 16 * (1 << 30), chunk_steps=10, precision="split") as srv:
            tickets = [submit(srv, s) for s in reqs]
            direct = run_jobs(ctx, JOBS[1] + JOBS[3], 500)
            mixed = [srv.wait(t, timeout_ms=120000) for t in tickets]
        with sb.Server(ctx, 16 * (1 << 30), chunk_steps=10, precision="split") as srv:
            alone = [srv.wait(submit(srv, s), timeout_ms=120000) for s in reqs]
        serial = run_jobs(ctx, JOBS[1] + JOBS[3], 500)
    for i, (a, m) in enumerate(zip(alone, mixed)):
        assert np.array_equal(a.spins, m.spins) and np.array_equal(a.energy, m.energy), i
    for k, (a, b) in enumerate(zip(direct, serial)):
        assert all(np.array_equal(u, v) for u, v in zip(a, b)), k
