#!/usr/bin/env python3
"""Benchmark: replica spin-updates/s of bSB on SK N=2000, R=1024 (FP32 plain path).

One bench step = one complete solve of the workload (keyed init, 2000 fused SB
steps over 1024 replicas, exact energies and best-replica readout), with J
already resident in HBM.  value = N * R * sb_steps * K / wall seconds of the K
timed solves.  Extra legs (same JSON line): the precision form on K2000 +-1 at
R=4096 (BF16 hi/lo split and FP8, vs FP32 at the same R), the roofline of the
fused step kernel from HIP events, and the CPU oracle on a bounded sample.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP8_TFLOPS = 5000.0
PEAK_HBM_GBS = 8000.0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def solve_timed(ctx, prob, K, W, **kw):
    for _ in range(W):
        ctx.solve(prob, spins=False, **kw)
    t0 = time.perf_counter()
    last = None
    for _ in range(K):
        last = ctx.solve(prob, spins=False, **kw)
    wall = time.perf_counter() - t0
    return wall, last


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed solves")
    ap.add_argument("--warmup", type=int, default=1, help="untimed solves")
    ap.add_argument("--sb-steps", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-precision", action="store_true", help="skip the precision-form leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 28  # 2^28 floats = 1 GiB read + 1 GiB written per copy
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "copies": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(replicas shard across GPUs with no communication; see DESIGN.md)", file=sys.stderr)
        return 2

    from annealing_amd import sb

    N, R, S = 2000, 1024, args.sb_steps
    ctx = sb.Context(0)
    J = sb.generate(sb.GEN_SK, N, 1.0, seed=2)
    prob = ctx.problem(J)
    kw = dict(variant="bsb", precision="fp32", steps=S, replicas=R, init="uniform", seed=2, dt=1.0)
    print(f"# bench: SK N={N} R={R} steps={S} fp32 bSB, warmup={args.warmup} timed={args.steps}",
          file=sys.stderr, flush=True)
    wall, last = solve_timed(ctx, prob, args.steps, args.warmup, **kw)
    value = N * R * S * args.steps / wall
    avg_launch_s = last.step_ms / 1e3 / S
    flops = 2.0 * N * N * R
    achieved = flops / avg_launch_s / 1e12
    out = {
        "metric": "replica spin-updates/s, bSB N=2000 R=1024; p50/p99 solve latency",
        "value": value,
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded SK J_ij ~ N(0,1)/sqrt(N), keyed uniform init",
        "config": {"workload": "bSB SK N=2000 R=1024, 2000 steps, dt=1.0, c0 auto (0.5/(sqrt(N) sigma_J)), "
                               "FP32 plain path, one solve per step",
                   "N": N, "replicas": R, "sb_steps": S, "variant": "bsb", "precision": "fp32"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "fused step (FP32 MFMA 32x32x2)", "avg_launch_us": avg_launch_s * 1e6,
                     "flops_per_launch": flops},
        "best_energy": float(last.energy.min()),
        "mean_energy": float(last.energy.mean()),
    }
    print(f"# fp32 leg: {value:.4g} spin-updates/s, {avg_launch_s*1e6:.1f} us/launch", file=sys.stderr, flush=True)

    if not args.no_precision:
        # K2000 +-1, R=4096: FP32 vs split (BF16 hi/lo, FP8 J) at the same seeds.
        N3, R3 = 2000, 4096
        J3 = sb.generate(sb.GEN_KN, N3, 1.0, seed=3)
        p3 = ctx.problem(J3)
        leg = {}
        for name, kw3 in (("fp32_bsb", dict(variant="bsb", precision="fp32")),
                          ("split_bsb", dict(variant="bsb", precision="split")),
                          ("split_dsb", dict(variant="dsb", precision="split")),
                          ("fp32_dsb", dict(variant="dsb", precision="fp32"))):
            common = dict(steps=S, replicas=R3, init="uniform", seed=3, dt=1.0)
            w3, r3 = solve_timed(ctx, p3, 1, 1, **kw3, **common)
            leg[name] = {"spin_updates_per_s": N3 * R3 * S / w3, "solve_s": w3,
                         "avg_launch_us": r3.step_ms * 1e3 / S,
                         "mean_energy": float(r3.energy.mean()), "best_energy": float(r3.energy.min())}
            print(f"# {name}: {leg[name]['spin_updates_per_s']:.4g} spin-updates/s", file=sys.stderr, flush=True)
        leg["speedup_split_bsb_vs_fp32"] = leg["split_bsb"]["spin_updates_per_s"] / leg["fp32_bsb"]["spin_updates_per_s"]
        leg["speedup_split_dsb_vs_fp32"] = leg["split_dsb"]["spin_updates_per_s"] / leg["fp32_dsb"]["spin_updates_per_s"]
        leg["mean_energy_gap_split_vs_fp32_bsb"] = (
            (leg["split_bsb"]["mean_energy"] - leg["fp32_bsb"]["mean_energy"]) / abs(leg["fp32_bsb"]["mean_energy"]))
        leg["workload"] = "K2000 +-1 (FP8-exact J), R=4096, 2000 steps, dt=1.0, seed 3"
        out["precision_form"] = leg
        p3.close()

    if not args.no_cpu:
        from oracle import oracle as O
        Rc, Sc = 2, S
        x, y = O.init_state("uniform", N, Rc, seed=2)
        c0 = O.c0_auto(O.BSB, J)
        t0 = time.perf_counter()
        O.run_f32(O.BSB, J, x, y, 0, Sc, Sc, 1.0, 1.0, c0, 1.0, 2, 1)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": N * Rc * Sc / tc, "unit": "spin-updates/s", "cores": 1, "kind": "port",
                               "sample": f"same SK N={N} instance, {Rc} replicas x {Sc} steps, FP32 sequential-fmaf "
                                         f"oracle, 1 thread, {tc:.1f} s on {cpu_model()} (nproc={os.cpu_count()})"}
    prob.close()
    ctx.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
