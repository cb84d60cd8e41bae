"""CPU tests of libsbamd.so's host side: it loads, exports every symbol
include/sb.h declares, and its host-only helpers agree with the oracle.
No GPU compute is called."""
import os
import re
import subprocess

import numpy as np
import pytest

from annealing_amd import sb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RNG = np.random.default_rng(7)


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sb.h")).read()
    return sorted(set(re.findall(r"SB_API\s+[^;(]*?\b(sb_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = sb.load()
    names = declared_symbols()
    assert len(names) >= 25
    assert set(names) == set(sb.EXPORTS)
    for n in names:
        assert hasattr(lib, n), n
    nm = subprocess.run(["nm", "-D", "--defined-only", sb.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (sb_\w+)", nm))
    assert set(names) <= exported


def test_version_and_status_strings():
    lib = sb.load()
    assert b"gfx950" in lib.sb_version()
    assert lib.sb_status_string(0) == b"ok"
    assert lib.sb_status_string(8) == b"request exceeds the HBM budget"


def test_default_params_are_reference_defines():
    p = sb.default_params()  # SB/CUDA_GPU/m2_opt.cu:10-22
    assert (p.variant, p.steps, p.M, p.replicas) == (sb.ASB, 100, 2, 1)
    assert (p.dt, p.a0, p.K) == (1.0, 1.0, 1.0)
    assert p.init == sb.INIT_REFERENCE


def _random_edgelist(N, m, real=False):
    lines = [str(N)]
    for _ in range(m):
        i, j = RNG.integers(0, N, 2)
        if i == j:
            continue
        w = f"{RNG.standard_normal():.6f}" if real else str(int(RNG.integers(-3, 4)))
        lines.append(f"{i} {j} {w}")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("real", [False, True])
def test_edgelist_parse_matches_oracle(orc, real):
    for N, m in ((1, 0), (5, 8), (60, 400)):
        text = _random_edgelist(N, m, real)
        J = sb.parse_edgelist(text)
        rc, Jo = orc.parse_edgelist(text)
        assert rc == 0 and np.array_equal(J, Jo)


def test_edgelist_errors_match_oracle(orc):
    cases = {
        "3\n0 0 1\n": 1,      # i == j (m2_opt.cu:119 asserts)
        "3\n0 3 1\n": 4,      # index out of range
        "3\n-1 2 1\n": 4,
        "3\n0 1\n": 3,        # truncated triple
        "x\n": 3,             # no N
        "3\n0 1 abc\n": 3,
    }
    for text, code in cases.items():
        with pytest.raises(sb.SBError) as e:
            sb.parse_edgelist(text)
        assert e.value.status == code
        assert orc.parse_edgelist(text)[0] == code
    J = sb.parse_edgelist("3\n1 3 2\n", one_based=True)
    assert J[0, 2] == 2 and J[2, 0] == 2


def test_generators():
    N = 300
    er = sb.generate(sb.GEN_ER, N, 0.06, seed=1)
    sk = sb.generate(sb.GEN_SK, N, 1.0, seed=2)
    kn = sb.generate(sb.GEN_KN, N, 1.0, seed=3)
    for J in (er, sk, kn):
        assert np.array_equal(J, J.T) and np.all(np.diag(J) == 0)
    off = ~np.eye(N, dtype=bool)
    assert set(np.unique(er)) <= {-1.0, 0.0, 1.0}
    assert abs(np.mean(er[off] != 0) - 0.06) < 0.01
    assert abs(sk[off].var() * N - 1.0) < 0.05 and abs(sk[off].mean()) < 0.01
    assert set(np.unique(kn[off])) == {-1.0, 1.0}
    assert np.array_equal(sb.generate(sb.GEN_ER, N, 0.06, seed=1), er)
    assert not np.array_equal(sb.generate(sb.GEN_ER, N, 0.06, seed=9), er)


def test_stats_and_c0_match_oracle(orc):
    for J in (sb.generate(sb.GEN_ER, 200, 0.1, 4), sb.generate(sb.GEN_SK, 150, 1.0, 5)):
        sigma, integer, mx = sb.instance_stats(J)
        assert abs(sigma - orc.sigma(J)) <= 1e-12 * sigma
        for v in (sb.ASB, sb.BSB, sb.DSB):
            assert abs(sb.c0_auto(v, J.shape[0], 1.0, sigma) - orc.c0_auto(v, J)) <= 1e-12
    assert sb.instance_stats(sb.generate(sb.GEN_KN, 50, 1.0, 1))[1:] == (True, 1.0)


def test_step_coeffs_match_oracle(orc):
    for v in (sb.ASB, sb.BSB, sb.DSB):
        p = sb.make_params(variant=v, steps=37, dt=0.7, a0=1.3)
        for k in (0, 1, 17, 36):
            c = sb.step_coeffs(p, 0.0123, k)
            o = orc.coeffs(v, k, 37, 0.7, 1.3, 0.0123)
            if v == sb.ASB:
                assert np.float32(c.beta) == o[1]
                # the fused step applies the next step's local sub-steps
                if k + 1 < 37:
                    assert np.float32(c.p_next) == orc.coeffs(v, k + 1, 37, 0.7, 1.3, 0.0123)[3]
                assert c.do_local == (1 if k + 1 < 37 else 0)
                assert c.dts == 0.7 / 2
            else:
                assert (np.float32(c.alpha), np.float32(c.beta), np.float32(c.gamma)) == tuple(o[:3])


def test_invalid_arguments_are_status_codes():
    lib = sb.load()
    assert lib.sb_default_params(None) == 1
    assert lib.sb_generate(7, 10, 0.5, 0, None) == 1
    p = sb.make_params(steps=10)
    c = sb.StepCoeffs()
    import ctypes
    assert lib.sb_step_coeffs_for(ctypes.byref(p), 0.1, 10, ctypes.byref(c)) == 1


def test_cli_usage_exits_zero():
    exe = os.path.join(ROOT, "annealing_amd", "bin", "sb_solve")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Usage" in r.stdout


def test_missing_library_raises():
    # No CPU fallback: without the built HIP library the binding refuses to load.
    code = "from annealing_amd import sb; sb.load('/nonexistent/libsbamd.so')"
    r = subprocess.run(["python3", "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "not built" in r.stderr


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_context_without_gpu_is_an_error():
    # Without a GPU the solver reports an error status; it never computes on the CPU.
    with pytest.raises(sb.SBError):
        sb.Context(0)


def kernel_scratch(text):
    """{kernel: scratch bytes per lane} from hipcc's kernel-resource-usage remarks."""
    out, name = {}, None
    for line in text.splitlines():
        m = re.search(r"remark: Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"remark:\s+ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            out[name] = int(m.group(1))
    return out


def test_scratch_parser_finds_a_spill():
    text = ("a.hip:1:1: remark: Function Name: k_one [-Rpass-analysis=kernel-resource-usage]\n"
            "a.hip:1:1: remark:     ScratchSize [bytes/lane]: 0 [-Rpass-analysis=kernel-resource-usage]\n"
            "a.hip:9:1: remark: Function Name: k_two [-Rpass-analysis=kernel-resource-usage]\n"
            "a.hip:9:1: remark:     ScratchSize [bytes/lane]: 264 [-Rpass-analysis=kernel-resource-usage]\n")
    assert kernel_scratch(text) == {"k_one": 0, "k_two": 264}


def test_no_kernel_spills_to_scratch():
    # The Makefile keeps hipcc's resource-usage remarks per translation unit.
    # A spill to scratch (e.g. the batched kernel, at two waves per SIMD and
    # 256 VGPRs, holding a tile that needs more) slows every launch.
    import glob
    files = sorted(glob.glob(os.path.join(ROOT, "annealing_amd", "_build", "*.usage")))
    if not files:
        pytest.skip("kernels not built by annealing_amd/csrc/Makefile in this tree")
    usage = {}
    for f in files:
        usage.update(kernel_scratch(open(f).read()))
    assert len(usage) >= 150 and any("k_batch_step" in k for k in usage)
    assert {k: v for k, v in usage.items() if v} == {}
