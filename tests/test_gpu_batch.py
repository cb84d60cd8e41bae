"""Batched solve (many instances, one persistent launch per step, graph
replay): each instance's result equals its solo sb_solve bit for bit."""
import numpy as np
import pytest

from annealing_amd import sb

pytestmark = pytest.mark.gpu


def make_mix(kinds):
    out = []
    for i, (kind, N) in enumerate(kinds):
        if kind == "er":
            J = sb.generate(sb.GEN_ER, N, 0.2, 100 + i)
        elif kind == "sk":
            J = sb.generate(sb.GEN_SK, N, 1.0, 100 + i)
        else:
            J = sb.generate(sb.GEN_KN, N, 1.0, 100 + i)
        out.append(J)
    return out


MIX = [("er", 37), ("sk", 64), ("kn", 130), ("er", 257), ("sk", 600), ("kn", 96), ("er", 1000)]


@pytest.mark.parametrize("precision", ["fp32", "split"])
def test_batch_equals_solo(precision):
    Js = make_mix(MIX)
    variants = ["bsb", "dsb", "asb", "bsb", "asb_m1", "dsb", "bsb"]
    Rs = [5, 64, 33, 100, 16, 1, 64]
    steps = [40, 100, 73, 60, 50, 90, 30]
    with sb.Context(0) as ctx:
        probs = [ctx.problem(J) for J in Js]
        params = [sb.make_params(variant=v, precision=precision, steps=s, replicas=R, init="uniform", seed=7 + i,
                                 dt=1.0 if v == "asb" else 0.5)
                  for i, (v, R, s) in enumerate(zip(variants, Rs, steps))]
        sols, st = ctx.batch_solve(probs, params, spins=True, state=True)
        assert st.graph_used == 1 and st.work_items > 0 and st.workgroups >= 256
        for i, p in enumerate(probs):
            solo = ctx.solve(p, state=True, variant=variants[i], precision=precision, steps=steps[i],
                             replicas=Rs[i], init="uniform", seed=7 + i, dt=1.0 if variants[i] == "asb" else 0.5)
            assert np.array_equal(sols[i].x, solo.x, equal_nan=True), i
            assert np.array_equal(sols[i].y, solo.y, equal_nan=True), i
            assert np.array_equal(sols[i].spins, solo.spins), i
            assert np.array_equal(sols[i].energy, solo.energy), i


def test_batch_energies_exact(orc):
    Js = make_mix([("er", 300), ("kn", 200)])
    with sb.Context(0) as ctx:
        probs = [ctx.problem(J) for J in Js]
        params = [sb.make_params(variant="bsb", precision="split", steps=100, replicas=40, init="uniform", seed=i,
                                 dt=0.5) for i in range(2)]
        sols, _ = ctx.batch_solve(probs, params)
        for J, s in zip(Js, sols):
            assert s.exact
            assert np.array_equal(s.energy_i64, orc.energy_i64(J, s.spins))


def test_batch_rejects_mixed_or_ordered_precision():
    Js = make_mix([("er", 50), ("er", 60)])
    with sb.Context(0) as ctx:
        probs = [ctx.problem(J) for J in Js]
        mixed = [sb.make_params(precision="fp32"), sb.make_params(precision="split")]
        with pytest.raises(sb.SBError):
            ctx.batch_solve(probs, mixed)
        ordered = [sb.make_params(precision="ordered"), sb.make_params(precision="ordered")]
        with pytest.raises(sb.SBError):
            ctx.batch_solve(probs, ordered)
