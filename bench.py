"""Benchmark of the global expectation step on one MI355X.

One step = one global E-step over the box-128 workload: 10,000 particles x
5,000 rotations x 121 translations (11 x 11 at 1 px), pixel cutoff 50
(K = 3,853), synthetic blob volume and particles resident in HBM.  The step
runs projection, particle-operand prep, norm GEMM, the MFMA cross-term
scorer and the FP64 marginal merge.

Prints one JSON line.  `value` is pose-evaluations per second of the mixed
precision path; the plain FP32 path is timed in the same run and reported
beside it, as is the CPU restatement of the reference loop on the host
cores (a bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "thunder-refactor-expectation_amd"))

import thx  # noqa: E402
import workloads as wl  # noqa: E402

PEAK = {"fp32": 157.3e12, "bf16": 2.5e15, "fp16": 2.5e15, "fp8": 5.0e15}


def run_mode(ctx, prec, steps, warmup, P, R, T, K):
    for _ in range(warmup):
        ctx.score_global(0, thx.PRECISIONS[prec])
    ctx.sync()
    per = []
    kern = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        ctx.score_global(0, thx.PRECISIONS[prec])
        ctx.sync()
        per.append(time.perf_counter() - s0)
        for k, v in ctx.last_timing().items():
            kern.setdefault(k, []).append(v)
    total = time.perf_counter() - t0
    ms = 1e3 * total / steps
    kern = {k: float(np.mean(v)) for k, v in kern.items()}
    evals = float(P) * R * T
    return dict(ms_per_step=ms, pose_evals_per_s=evals / (total / steps), kernels_ms=kern,
                score_flops=4.0 * K * evals)


def cpu_baseline(ctx, info, name, P_s=384, R_s=1000):
    """Reference CPU loop (oracle restatement, OpenMP) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    w = wl.WORKLOADS[name]
    V = w["box"] * w["pf"]
    vol = ctx.get_volume(0, V)
    dat, ctf, sig = ctx.get_particles()
    quat, trans = info["quat"][:R_s], info["trans"]
    T = len(trans)
    cores = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    oracle.global_scan(vol, V, w["pf"], info["iCol"], info["iRow"], w["box"], quat, trans,
                       dat[:P_s], ctf[:P_s], sig[:P_s], np.full(R_s, 1.0 / R_s),
                       np.full(T, 1.0 / T), f64=False, threads=cores)
    dt = time.perf_counter() - t0
    return dict(value=P_s * R_s * T / dt, unit="pose-evals/s", cores=cores, kind="port",
                sample=f"{P_s} particles x {R_s} rotations x {T} translations of {name}, "
                       f"FP32 reference loop order, {dt:.1f} s")


def traffic_bytes(name, prec):
    """HBM bytes per k_score launch from the committed rocprofv3 FETCH_SIZE +
    WRITE_SIZE summary of this workload (copy-kernel calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(path):
        return None
    d = json.load(open(path))
    tag = {"fp32": 0, "bf16": 1, "fp16": 2}[prec]
    for k, v in d["kernels"].items():
        if k.startswith(f"thx::k_score<{tag},"):
            return (v["read_GB"] + v["write_GB"]) * 1e9
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="box128")
    ap.add_argument("--mixed", default="bf16", choices=["bf16", "fp16"])
    ap.add_argument("--particles", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(particles shard across GPUs, see DESIGN.md)"}))
        sys.exit(2)
    name = args.workload
    w = wl.WORKLOADS[name]
    ctx = thx.Context(0)
    print(f"[bench] setting up {name}", file=sys.stderr, flush=True)
    info = wl.setup_context(ctx, name, n_particles=args.particles or None)
    P, R, T, K = ctx.P, ctx.R, ctx.T, ctx.K
    res = {}
    modes = [args.mixed] + ([] if args.no_fp32 else ["fp32"])
    for m in modes:
        print(f"[bench] mode {m}", file=sys.stderr, flush=True)
        res[m] = run_mode(ctx, m, args.steps, args.warmup, P, R, T, K)
        print(f"[bench] {m}: {res[m]['ms_per_step']:.1f} ms/step "
              f"{res[m]['kernels_ms']}", file=sys.stderr, flush=True)
    main_m = args.mixed
    r = res[main_m]
    t_score = r["kernels_ms"]["score"] * 1e-3
    achieved = r["score_flops"] / t_score
    out = {
        "metric": "pose-evals/sec (particles x rot x trans) at box 128",
        "value": r["pose_evals_per_s"],
        "unit": "pose-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": main_m,
        "data": "synthetic (seeded blob volume, projected particles with CTF, shifts, noise)",
        "config": {"workload": f"{name}: box {w['box']}, {P} particles x {R} rotations x "
                               f"{T} translations, cutoff {w['cutoff']} (K={K}), global scan, "
                               f"{main_m} cross term, no learned proposal",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK[main_m] / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK[main_m],
                     "traffic": traffic_bytes(name, main_m),
                     "kernel": "k_score", "kernel_ms": r["kernels_ms"]["score"],
                     "flops_per_launch": r["score_flops"],
                     "flops_definition": "4*K per scored (particle, rotation, translation)"},
        "kernels_ms": r["kernels_ms"],
    }
    if "fp32" in res:
        f = res["fp32"]
        tf = f["kernels_ms"]["score"] * 1e-3
        out["plain_fp32"] = {"value": f["pose_evals_per_s"], "ms_per_step": f["ms_per_step"],
                             "kernels_ms": f["kernels_ms"],
                             "score_tflops": f["score_flops"] / tf / 1e12,
                             "score_frac_of_fp32_peak": f["score_flops"] / tf / PEAK["fp32"]}
        out["speedup_vs_plain_fp32"] = r["pose_evals_per_s"] / f["pose_evals_per_s"]
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(ctx, info, name)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
