#!/usr/bin/env python3
"""Benchmark: replica spin-updates/s of bSB on SK N=2000, R=1024 (FP32 plain path).

One bench step = one complete solve of that workload (keyed init, 2000 fused SB
steps over 1024 replicas, exact energies and best-replica readout), with J
already resident in HBM.  value = N * R * sb_steps * K / wall seconds of the K
timed solves.

Further legs on the same JSON line (each can be skipped):
  precision_form  K2000 +-1, R=4096: BF16 hi/lo split and FP8 (dSB) vs FP32
  batched         512 ER instances, N ~ U[64,4096], R=64, 1500 steps, one
                  persistent launch per step replayed from a hipGraph
  serving         Poisson arrivals at 200 req/s for 60 s (--serve-seconds), N
                  log-uniform [100,8000], SK/+-1 halves, density U[0.01,1],
                  R in {32,64,128,256}, steps U[1000,5000]: p50/p99 latency
                  (the offered load is several times one GPU's capacity, so
                  the queue drains for a few minutes after the window)
  accuracy        field error of the FP32 and split paths against FP64
                  trajectories (SK N=2000 8 replicas, K2000 16 replicas) and
                  the final-energy gaps per replica
  c1, cpu_baseline  ER N=800 p=0.06, R=1, 1000 steps, dt=1.25: the GPU thin
                  path, and the CPU oracle on one thread and on all threads
"""
import argparse
import json
import math
import os
import platform
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP8_TFLOPS = 5000.0
PEAK_HBM_GBS = 8000.0
GB = 1 << 30


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


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


def measured_traffic(workload):
    """HBM bytes per launch of the kernel this workload runs, from the committed
    rocprofv3 counter summary (profiles/pmc_traffic.json), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        w, cal = t["workloads"][workload], t["calibration"]
    except (OSError, KeyError, ValueError):
        return None
    return (w["fetch_kib"] * cal["fetch_scale"] + w["write_kib"] * cal["write_scale"]) * 1024.0


def leg_precision(sb, ctx, S):
    N3, R3 = 2000, 4096
    p3 = ctx.problem(sb.generate(sb.GEN_KN, N3, 1.0, seed=3))
    leg = {}
    for name, kw3 in (("fp32_bsb", dict(variant="bsb", precision="fp32")),
                      ("split_bsb", dict(variant="bsb", precision="split")),
                      ("fp32_dsb", dict(variant="dsb", precision="fp32")),
                      ("split_dsb", dict(variant="dsb", precision="split"))):
        common = dict(steps=S, replicas=R3, init="uniform", seed=3, dt=1.0)
        w3, r3 = solve_timed(ctx, p3, 1, 1, **kw3, **common)
        us = r3.step_ms * 1e3 / S
        flops = 2.0 * N3 * N3 * R3
        leg[name] = {"spin_updates_per_s": N3 * R3 * S / w3, "solve_s": w3, "avg_step_us": us,
                     "alg_tflops": flops / us / 1e6, "mean_energy": float(r3.energy.mean()),
                     "best_energy": float(r3.energy.min())}
        log(f"{name}: {leg[name]['spin_updates_per_s']:.4g} spin-updates/s, {us:.1f} us/step")
    f, s = leg["fp32_bsb"], leg["split_bsb"]
    leg["speedup_split_bsb_vs_fp32"] = s["spin_updates_per_s"] / f["spin_updates_per_s"]
    leg["speedup_split_dsb_vs_fp32"] = leg["split_dsb"]["spin_updates_per_s"] / leg["fp32_dsb"]["spin_updates_per_s"]
    leg["mean_energy_gap_split_vs_fp32_bsb"] = (s["mean_energy"] - f["mean_energy"]) / abs(f["mean_energy"])
    # split bSB issues two BF16 MFMA passes: 2 x the algorithmic flops
    leg["split_bsb_roofline"] = {"bound": "mfma", "achieved_issued": 2 * s["alg_tflops"], "peak": PEAK_BF16_TFLOPS,
                                 "unit": "TFLOP/s", "frac": 2 * s["alg_tflops"] / PEAK_BF16_TFLOPS,
                                 "traffic": measured_traffic("bSB K2000 R=4096 split")}
    # dSB FP8: state traffic per step, 16 B/spin-replica (x, y read+write) + the 1 B sign plane read
    # and the next one written, + J (FP8) once
    st_bytes = (16 + 2) * N3 * R3 + N3 * N3
    gbs = st_bytes / (leg["split_dsb"]["avg_step_us"] * 1e-6) / 1e9
    leg["split_dsb_roofline"] = {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                 "frac": gbs / PEAK_HBM_GBS, "bytes_per_launch": st_bytes,
                                 "traffic": measured_traffic("dSB K2000 R=4096 split")}
    leg["workload"] = "K2000 +-1 (FP8-exact J), R=4096, 2000 steps, dt=1.0, seed 3"
    p3.close()
    return leg


def leg_matvec(sb, ctx, S):
    """R = 1: the reference's own shape (one replica, one SGEMV per step).  The
    thin kernels stream J once per step; achieved GB/s counts J bytes only."""
    leg = {"workload": f"R=1 bSB, {S} steps, dt=1.0; SK (FP32 J, split form reads BF16 J) and K_N +-1 "
                       "(split form reads FP8 J) generated on the device, seed 2"}
    for N, kind, kname in ((16384, sb.GEN_SK, "sk"), (16384, sb.GEN_KN, "kn"), (2000, sb.GEN_SK, "sk")):
        prob = ctx.generate(kind, N, 1.0, seed=2)
        for prec in ("fp32", "split"):
            if kname == "sk" and prec == "split" and N != 16384:
                continue
            if kname == "kn" and prec == "fp32":
                continue
            jb = 4 if prec == "fp32" else (1 if prob.split_jfmt == sb.J_FP8 else 2)
            w, r = solve_timed(ctx, prob, 1, 1, variant="bsb", precision=prec, steps=S, replicas=1,
                               init="uniform", seed=2, dt=1.0)
            us = r.step_ms * 1e3 / S
            gbs = N * N * jb / (us * 1e-6) / 1e9
            name = f"{kname}{N}_{prec}_r1"
            leg[name] = {"avg_launch_us": us, "j_bytes_per_launch": N * N * jb, "j_format": {4: "fp32", 2: "bf16",
                         1: "fp8"}[jb], "achieved_gbs": gbs, "frac_of_hbm_peak": gbs / PEAK_HBM_GBS,
                         "spin_updates_per_s": N * S / w}
            log(f"matvec {name}: {us:.1f} us/launch, {gbs:.0f} GB/s of J")
        prob.close()
    return leg


def leg_sparse(sb, ctx, S):
    """Low-density instances: the CSR (sparse) family against the dense paths on
    the same J and seeds.  The sparse sums are the sequential order; the dense
    paths are MFMA tiles (FP32, and the split form on FP8 J)."""
    leg = {"workload": f"ER +-1 (J = -W) generated on the device, bSB, {S} steps, dt=0.5, seed 5"}
    for N, p, R in ((16384, 0.01, 64), (4096, 0.05, 64), (4096, 0.01, 256)):
        prob = ctx.generate(sb.GEN_ER, N, p, seed=5)
        kw = dict(spins=False, variant="bsb", steps=S, replicas=R, init="uniform", seed=1, dt=0.5)
        prob.set_sparse(True)
        nnz = prob.nnz
        ctx.solve(prob, precision="fp32", **kw)
        sp = ctx.solve(prob, precision="fp32", **kw)
        prob.set_sparse(False)
        out = {"nnz": nnz, "sparse_fp32_us": sp.step_ms * 1e3 / S,
               "sparse_mean_energy": float(sp.energy.mean())}
        for prec in ("fp32", "split"):
            ctx.solve(prob, precision=prec, **kw)
            d = ctx.solve(prob, precision=prec, **kw)
            out[f"dense_{prec}_us"] = d.step_ms * 1e3 / S
            out[f"dense_{prec}_mean_energy"] = float(d.energy.mean())
        us = out["sparse_fp32_us"]
        # CSR (4 B index + 4 B value) and the state (16 B per spin-replica) from HBM;
        # one 256-byte gather of X^T per nonzero and 64-replica block from L2
        out["hbm_gbs_alg"] = (8.0 * nnz + 16.0 * N * R) / (us * 1e-6) / 1e9
        out["gather_gbs"] = nnz * 4.0 * ((R + 63) // 64 * 64) / (us * 1e-6) / 1e9
        out["speedup_vs_dense_split"] = out["dense_split_us"] / us
        out["speedup_vs_dense_fp32"] = out["dense_fp32_us"] / us
        leg[f"er{N}_p{p}_r{R}"] = out
        log(f"sparse N={N} p={p} R={R}: {us:.1f} us/step vs dense split {out['dense_split_us']:.1f}, "
            f"fp32 {out['dense_fp32_us']:.1f}")
        prob.close()
    return leg


def leg_batched(sb, ctx, n_inst, steps):
    rng = np.random.default_rng(4)
    Ns = rng.integers(64, 4097, n_inst)
    t0 = time.perf_counter()
    probs = [ctx.generate(sb.GEN_ER, int(n), 0.05, 1000 + i) for i, n in enumerate(Ns)]
    gen_s = time.perf_counter() - t0
    leg = {"workload": f"{n_inst} ER instances N~U[64,4096] p=0.05 +-1, R=64, {steps} steps, bSB, seed 4",
           "instances": n_inst, "sum_N": int(Ns.sum()), "generate_s": gen_s}
    work = float(Ns.sum()) * 64 * steps
    means = {}
    for prec in ("fp32", "split"):
        params = [sb.make_params(variant="bsb", precision=prec, steps=steps, replicas=64, init="uniform",
                                 seed=i, dt=1.0) for i in range(n_inst)]
        # warm-up (autotune of the (N, R) buckets) on a short run
        warm = [sb.make_params(variant="bsb", precision=prec, steps=2, replicas=64, init="uniform", seed=i)
                for i in range(n_inst)]
        ctx.batch_solve(probs, warm, spins=False)
        t0 = time.perf_counter()
        sols, st = ctx.batch_solve(probs, params, spins=False)
        wall = time.perf_counter() - t0
        means[prec] = float(np.mean([s.energy.mean() for s in sols]))
        leg[prec] = {"spin_updates_per_s": work / wall, "wall_s": wall, "device_step_ms": st.step_ms,
                     "avg_launch_us": st.step_ms * 1e3 / steps, "work_items": st.work_items,
                     "workgroups": st.workgroups, "graph_used": bool(st.graph_used)}
        log(f"batched {prec}: {work / wall:.4g} spin-updates/s ({wall:.2f} s, {st.work_items} tiles/step)")
    leg["speedup_split_vs_fp32"] = leg["split"]["spin_updates_per_s"] / leg["fp32"]["spin_updates_per_s"]
    leg["mean_energy_gap_split_vs_fp32"] = (means["split"] - means["fp32"]) / abs(means["fp32"])
    # Algorithmic bytes per step: J once (FP8 split, FP32 plain) + the state of
    # every spin-replica (X, Y read and written: 16 B; the split form reads
    # FP32 X and splits it into BF16 hi + lo while staging).
    jb = float((Ns.astype(np.float64) ** 2).sum())
    sr = float(Ns.sum()) * 64
    for prec, bj, bs in (("split", 1.0, 16.0), ("fp32", 4.0, 16.0)):
        t = leg[prec]["avg_launch_us"] * 1e-6
        leg[prec]["alg_bytes_per_step"] = bj * jb + bs * sr
        leg[prec]["hbm_gbs_alg"] = (bj * jb + bs * sr) / t / 1e9
        leg[prec]["hbm_gbs_j_only"] = bj * jb / t / 1e9
        # HBM bytes per step from the committed counter runs (same workload)
        tr = measured_traffic(f"batched ER512 R=64 {prec}") if n_inst == 512 else None
        leg[prec]["traffic"] = tr
        leg[prec]["hbm_gbs_measured"] = tr / t / 1e9 if tr else None
    for p in probs:
        p.close()
    return leg


def leg_serving(sb, ctx, seconds, rate, drain_cap_s, chunk=50, sparse_density=0.0):
    rng = np.random.default_rng(5)
    n_req = int(rng.poisson(rate * seconds))
    arrivals = np.sort(rng.uniform(0.0, seconds, n_req))
    Ns = np.exp(rng.uniform(math.log(100), math.log(8000), n_req)).astype(int)
    kinds = np.where(rng.random(n_req) < 0.5, sb.GEN_SK_DILUTED, sb.GEN_ER)
    dens = rng.uniform(0.01, 1.0, n_req)
    Rs = rng.choice([32, 64, 128, 256], n_req)
    steps = rng.integers(1000, 5001, n_req)
    srv = sb.Server(ctx, hbm_budget_bytes=128 * GB, chunk_steps=chunk, precision="split",
                    sparse_density=sparse_density)
    tickets = [None] * n_req
    lat = {}
    done_evt = threading.Event()
    lock = threading.Lock()

    # Requests retire out of submission order (shortest work first), so several
    # threads collect them, each blocking on its own share of the tickets in
    # turn; the latency is stamped by the server when a request retires, not
    # when it is collected here.
    n_wait = 32
    progress = {"retired": 0, "last": time.perf_counter()}

    def waiter(w):
        for i in range(w, n_req, n_wait):
            while tickets[i] is None:
                time.sleep(0.001)
            try:
                r = srv.wait(tickets[i], spins=False)
                with lock:
                    lat[i] = r.total_ms
            except sb.SBError:
                pass
            with lock:
                progress["retired"] += 1
                retired = progress["retired"]
                report = time.perf_counter() - progress["last"] > 30.0
                if report:
                    progress["last"] = time.perf_counter()
                if retired == n_req:
                    done_evt.set()
            if report:  # long windows: report progress as requests retire
                log(f"serving: {retired}/{n_req} requests retired")

    for w in range(n_wait):
        threading.Thread(target=waiter, args=(w,), daemon=True).start()
    t0 = time.perf_counter()
    for i in range(n_req):
        delay = arrivals[i] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        tickets[i] = srv.submit_generated(int(kinds[i]), int(Ns[i]), float(dens[i]), 5000 + i, variant="bsb",
                                          precision="split", steps=int(steps[i]), replicas=int(Rs[i]),
                                          init="uniform", seed=i, dt=1.0)
    submit_s = time.perf_counter() - t0
    finished = done_evt.wait(timeout=drain_cap_s)
    if not finished:
        for i in range(n_req):
            if i not in lat:
                try:
                    srv.cancel(tickets[i])
                except sb.SBError:
                    pass
        done_evt.wait(timeout=120)
    total_s = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    done_idx = sorted(lat)
    L = np.array([lat[i] for i in done_idx]) if done_idx else np.zeros(1)
    work_all = (Ns.astype(np.float64) ** 2 * 2 * Rs * steps)
    work_done = float(work_all[done_idx].sum()) if done_idx else 0.0
    spin_upd = float((Ns * Rs * steps)[done_idx].sum()) if done_idx else 0.0
    leg = {
        "workload": f"Poisson {rate} req/s for {seconds} s (seed 5): N log-uniform [100,8000], half diluted SK "
                    "(BF16 J) half ER +-1 (FP8 J), density U[0.01,1], R in {32,64,128,256}, steps U[1000,5000], "
                    f"bSB split precision, chunk {chunk} steps, HBM budget 128 GiB, sparse family at density "
                    f"<= {sparse_density}",
        "requests": n_req, "completed": len(done_idx), "cancelled_at_drain_cap": n_req - len(done_idx),
        "p50_ms": float(np.percentile(L, 50)), "p99_ms": float(np.percentile(L, 99)),
        "mean_ms": float(L.mean()), "submit_window_s": submit_s, "total_s": total_s,
        "throughput_req_per_s": len(done_idx) / total_s,
        "spin_updates_per_s": spin_upd / total_s,
        "offered_tflops_alg": float(work_all.sum()) / seconds / 1e12,
        "served_tflops_alg": work_done / max(st.busy_ms * 1e-3, 1e-9) / 1e12,
        "device_busy_s": st.busy_ms / 1e3, "chunks": int(st.chunks), "max_active": int(st.max_active_seen),
    }
    leg["offered_over_capacity"] = leg["offered_tflops_alg"] / max(leg["served_tflops_alg"], 1e-9)
    log(f"serving: {len(done_idx)}/{n_req} done, p50 {leg['p50_ms']:.0f} ms, p99 {leg['p99_ms']:.0f} ms, "
        f"offered/capacity {leg['offered_over_capacity']:.2f}")
    return leg


def accuracy_run(sb, ctx, O, kind, N, R, seed, S, nthreads, snaps_per_run=5):
    """FP64 oracle trajectory of R replicas against the FP32 and split solves
    at the same seed: the field error of both GPU products on the FP64 state
    (rounded to FP32) at a few steps, and per-replica final energies."""
    J = sb.generate(kind, N, 1.0, seed=seed)
    prob = ctx.problem(J)
    c0 = O.c0_auto(O.BSB, J)
    x32, y32 = O.init_state("uniform", N, R, seed=seed)
    x, y = x32.astype(np.float64), y32.astype(np.float64)
    snaps = sorted({0, S // 4, S // 2, 3 * S // 4, S - 1}) if snaps_per_run else []
    err = {"fp32": [], "split": []}
    k = 0
    t0 = time.perf_counter()
    for sn in snaps:
        O.run_f64(O.BSB, J, x, y, k, sn, S, 1.0, 1.0, c0, 1.0, 2, nthreads)
        k = sn
        xs = x[:, :N].astype(np.float32)
        F64 = O.field_f64(J, np.ascontiguousarray(np.pad(xs.astype(np.float64), ((0, 0), (0, x.shape[1] - N)))))[:, :N]
        for prec in ("fp32", "split"):
            F = ctx.field(prob, xs, precision=prec, variant="bsb").astype(np.float64)
            d = np.abs(F - F64)
            err[prec].append({"step": sn, "max_abs": float(d.max()), "rel_max": float(d.max() / np.abs(F64).max()),
                              "rel_rms": float(np.sqrt((d ** 2).mean() / (F64 ** 2).mean()))})
    O.run_f64(O.BSB, J, x, y, k, S, S, 1.0, 1.0, c0, 1.0, 2, nthreads)
    oracle_s = time.perf_counter() - t0
    s64 = np.where(x[:, :N] > 0, 1, -1).astype(np.int8)
    H64 = O.energy_i64(J, s64) if kind == sb.GEN_KN else O.energy_f64(J, s64)
    H64 = np.asarray(H64, dtype=np.float64)
    out = {"fp64_energy": H64.tolist(), "fp64_mean_energy": float(H64.mean()), "fp64_best_energy": float(H64.min()),
           "oracle_s": oracle_s}
    if snaps:
        out["field_error"] = err
    for prec in ("fp32", "split"):
        sol = ctx.solve(prob, variant="bsb", precision=prec, steps=S, replicas=R, init="uniform", seed=seed, dt=1.0,
                        c0=c0)
        E = sol.energy
        out[f"{prec}_energy"] = E.tolist()
        out[f"{prec}_mean_gap_vs_fp64"] = (float(E.mean()) - out["fp64_mean_energy"]) / abs(out["fp64_mean_energy"])
        out[f"{prec}_best_gap_vs_fp64"] = (float(E.min()) - out["fp64_best_energy"]) / abs(out["fp64_best_energy"])
        rel = (E - H64) / np.abs(H64)
        diff = np.nonzero(E != H64)[0]
        out[f"{prec}_replicas_differing"] = int(diff.size)
        out[f"{prec}_per_replica_gap"] = {int(r): float(rel[r]) for r in diff}
    prob.close()
    return out


def leg_accuracy(sb, ctx, S, nthreads):
    """FP64 oracle trajectories against the FP32 and split paths.  SK N=2000
    (BF16-rounded real J in the split form) with 8 replicas, and K2000 +-1
    (FP8-exact J) with 16 replicas: field error per checkpoint step, final
    energies per replica and the mean/best gaps."""
    from oracle import oracle as O
    out = {"workload": f"bSB, {S} steps, dt=1.0, c0 auto, keyed uniform init; FP64 oracle trajectory on "
                       f"{nthreads} threads; sk2000: SK N=2000 seed 2, 8 replicas; k2000: K2000 +-1 seed 3, "
                       "16 replicas",
           "sk2000": accuracy_run(sb, ctx, O, sb.GEN_SK, 2000, 8, 2, S, nthreads),
           "k2000": accuracy_run(sb, ctx, O, sb.GEN_KN, 2000, 16, 3, S, nthreads)}
    for name in ("sk2000", "k2000"):
        a = out[name]
        log(f"accuracy {name}: split mean gap {a['split_mean_gap_vs_fp64']:.2e}, fp32 {a['fp32_mean_gap_vs_fp64']:.2e}, "
            f"field rel_max split {max(e['rel_max'] for e in a['field_error']['split']):.2e}")
    return out


def leg_c1(sb, ctx, nthreads, cpu_seconds=12.0):
    """The CPU-runnable configuration: ER N=800 p=0.06 +-1 (J = -W), R=1,
    1000 bSB steps, dt=1.25, a0=1, c0 auto, keyed uniform init from seed 1.
    GPU: the thin R = 1 path (FP32), timed.  CPU: the oracle's FP32 loop (the
    reference's sequential-sum order) on one thread, repeated over seeds for a
    bounded sample, then on all threads (rows of each field split over the
    threads).  Both report H and the cut value of seed 1."""
    from oracle import oracle as O
    N, S, dt = 800, 1000, 1.25
    J = sb.generate(sb.GEN_ER, N, 0.06, seed=1)
    prob = ctx.problem(J)
    kw = dict(variant="bsb", precision="fp32", steps=S, replicas=1, init="uniform", seed=1, dt=dt)
    ctx.solve(prob, **kw)
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        g = ctx.solve(prob, **kw)
    gw = (time.perf_counter() - t0) / reps
    gpu = {"spin_updates_per_s": N * S / gw, "solve_ms": gw * 1e3, "avg_step_us": g.step_ms * 1e3 / S,
           "H": int(g.energy_i64[0]), "cut": float(g.cut[0]), "c0": g.c0}
    prob.close()
    c0 = O.c0_auto(O.BSB, J)
    W = -np.triu(J.astype(np.float64), 1)

    def cpu_solve(seed, threads):
        x, y = O.init_state("uniform", N, 1, seed=seed)
        O.run_f32(O.BSB, J, x, y, 0, S, S, dt, 1.0, c0, nthreads=threads)
        return x

    t0 = time.perf_counter()
    n1 = 0
    while time.perf_counter() - t0 < cpu_seconds or n1 < 3:
        x = cpu_solve(1 + n1, 1)
        if n1 == 0:
            s = O.spins(x, N)
            H1 = int(O.energy_i64(J, s)[0])
            cut1 = float(W[s[0][:, None] != s[0][None, :]].sum())
        n1 += 1
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    for i in range(n1):
        cpu_solve(1 + i, nthreads)
    tn = time.perf_counter() - t0
    cpu = {"value": N * S * n1 / t1, "unit": "spin-updates/s", "cores": 1, "kind": "port",
           "sample": f"C1 ER N={N} p=0.06 +-1, R=1, {S} bSB steps, dt={dt}: {n1} solves (seeds 1..{n1}) of the "
                     f"oracle's FP32 sequential-fmaf loop on 1 thread, {t1:.1f} s, {cpu_model()} "
                     f"(nproc={os.cpu_count()})",
           "H_seed1": H1, "cut_seed1": cut1,
           "all_cores": {"value": N * S * n1 / tn, "cores": nthreads, "seconds": tn,
                         "how": "the same solves, rows of each field over the threads (same sums)"}}
    log(f"C1: GPU {gpu['spin_updates_per_s']:.4g} spin-updates/s ({gpu['avg_step_us']:.2f} us/step), "
        f"CPU 1 thread {cpu['value']:.4g}, {nthreads} threads {cpu['all_cores']['value']:.4g}; "
        f"H {gpu['H']} / {H1}, cut {gpu['cut']:.0f} / {cut1:.0f}")
    return gpu, cpu


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed solves")
    ap.add_argument("--warmup", type=int, default=1, help="untimed solves")
    ap.add_argument("--sb-steps", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-precision", action="store_true", help="skip the precision-form leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 accuracy leg")
    ap.add_argument("--no-matvec", action="store_true", help="skip the R=1 matrix-vector leg")
    ap.add_argument("--no-sparse", action="store_true", help="skip the low-density (CSR) leg")
    ap.add_argument("--only-headline", action="store_true", help="skip every extra leg")
    ap.add_argument("--batch-instances", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=1500)
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="arrival window of the serving leg")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-drain-cap", type=float, default=300.0, help="max seconds to drain after arrivals")
    ap.add_argument("--serve-chunk", type=int, default=50, help="server steps between admission points")
    ap.add_argument("--serve-sparse-density", type=float, default=0.0,
                    help="requests at or below this density run the sparse family (0: never)")
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
        print(f"bench.py: one GPU only; --gpus {args.gpus} is not supported (replicas and instances would "
              "shard across GPUs with J replicated and no communication until the final gather)", file=sys.stderr)
        return 2
    if args.only_headline:
        args.no_cpu = args.no_precision = args.no_batched = args.no_serving = args.no_accuracy = True
        args.no_matvec = True
        args.no_sparse = True

    from annealing_amd import sb

    N, R, S = 2000, 1024, args.sb_steps
    ctx = sb.Context(0)
    J = sb.generate(sb.GEN_SK, N, 1.0, seed=2)
    prob = ctx.problem(J)
    kw = dict(variant="bsb", precision="fp32", steps=S, replicas=R, init="uniform", seed=2, dt=1.0)
    log(f"bench: SK N={N} R={R} steps={S} fp32 bSB, warmup={args.warmup} timed={args.steps}")
    wall, last = solve_timed(ctx, prob, args.steps, args.warmup, **kw)
    value = N * R * S * args.steps / wall
    avg_step_s = last.step_ms / 1e3 / S
    flops = 2.0 * N * N * R
    achieved = flops / avg_step_s / 1e12
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
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": measured_traffic("bSB SK N=2000 R=1024 fp32"),
                     "traffic_algorithmic": 4.0 * N * N + 16.0 * N * R,
                     "kernel": "k_step_fp32 (FP32 MFMA 32x32x2, fused bSB update); each step is two concurrent "
                               "launches over the replica halves (512 + 512) on two streams",
                     "launches_per_step": 2, "avg_step_us": avg_step_s * 1e6, "flops_per_step": flops},
        "best_energy": float(last.energy.min()),
        "mean_energy": float(last.energy.mean()),
    }
    log(f"fp32 leg: {value:.4g} spin-updates/s, {avg_step_s*1e6:.1f} us/step, {achieved:.1f} TFLOP/s")

    if not args.no_precision:
        out["precision_form"] = leg_precision(sb, ctx, S)
    if not args.no_matvec:
        out["matvec"] = leg_matvec(sb, ctx, 200)
    if not args.no_sparse:
        out["sparse"] = leg_sparse(sb, ctx, 200)
    if not args.no_batched:
        out["batched"] = leg_batched(sb, ctx, args.batch_instances, args.batch_steps)
    if not args.no_serving:
        out["serving"] = leg_serving(sb, ctx, args.serve_seconds, args.serve_rate, args.serve_drain_cap,
                                     args.serve_chunk, args.serve_sparse_density)
        out["p50_ms"] = out["serving"]["p50_ms"]
        out["p99_ms"] = out["serving"]["p99_ms"]
    threads = min(16, os.cpu_count() or 1)
    if not args.no_accuracy:
        out["accuracy"] = leg_accuracy(sb, ctx, S, threads)
    if not args.no_cpu:
        gpu_c1, cpu = leg_c1(sb, ctx, threads)
        out["c1"] = {"workload": "ER N=800 p=0.06 +-1 (J = -W), R=1, 1000 bSB steps, dt=1.25, a0=1, c0 auto, "
                                 "keyed uniform init seed 1; GPU thin R=1 FP32 path", "gpu": gpu_c1}
        out["cpu_baseline"] = cpu
    prob.close()
    ctx.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
