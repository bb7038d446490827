"""Benchmark: plain FP64 Kerr frame at 1024x1024, a=0.998, 250k stars.

One step = one pass of the hot path over one frame: the persistent tracer over
all 1,048,576 rays plus the shading pass (disc emission, star-cell lookup,
colour), inputs (scene, star catalogue, colour table) already resident on the
GPU.  Prints one JSON line; see DESIGN.md for the definitions of every field.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP64 operation counts of the reference's path (hand count of
# src/geodesic.h / src/kerr.h / src/render.cpp): per accepted DP5 step
# (6 RHS + stage/error/controller + dense output + per-step invariants and
# events) and per rejected attempt.
ALG_FLOPS_PER_ACCEPTED_STEP = 1503
ALG_FLOPS_PER_REJECTED_ATTEMPT = 1280
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector peak (AMD published)


def cpu_baseline(sc, threads):
    """The oracle (CPU restatement of the reference) on a strided quarter of
    the same frame, all host threads; rays/s scaled to Mrays/s."""
    import numpy as np
    import oracle
    ids = np.arange(0, sc.n_rays, 4)
    t0 = time.perf_counter()
    out = oracle.trace_list(sc, ids, threads=threads)
    dt = time.perf_counter() - t0
    steps = int(out["steps"].sum())
    att = int(out["attempts"].sum())
    return {
        "value": len(ids) / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
        "sample": f"every 4th ray of the 1024x1024 frame ({len(ids)} rays, {steps} accepted steps)"
                  f" traced+shaded by oracle/ (bitwise-equal CPU restatement) in {dt:.2f} s",
        "rhs_evals_per_s": (6 * att + len(ids)) / dt,
        "cpu_model": platform.processor() or platform.machine(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--width", type=int, default=1024)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; per-GPU sharding is described in "
              "DESIGN.md (no multi-GPU figure is measured)", file=sys.stderr)
        sys.exit(2)

    import torch
    from blackhole_amd import Tracer, scene as S

    sc = S.config2(args.width, args.width)
    dev = torch.device("cuda:0")
    tr = Tracer(0)
    rays = tr.alloc_rays(sc, full=False)
    out = {"xyz": torch.empty((sc.height, sc.width, 3), dtype=torch.float64, device=dev)}
    # warm-up: uploads the star catalogue, colour table and tile order
    for _ in range(max(1, args.warmup)):
        tr.trace_async(sc, rays)
        tr.shade_async(sc, rays, out)
    torch.cuda.synchronize()
    # one synchronous frame for the step / attempt counters
    _, st = tr.trace(sc, rays)
    _, st = tr.shade(sc, rays, out, stats=st)

    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        tr.trace_async(sc, rays)
        ev[k][1].record(stream)
        tr.shade_async(sc, rays, out)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    trace_ms = sum(a.elapsed_time(b) for a, b in ev) / args.steps

    ms_per_step = wall / args.steps * 1e3
    n_rays = sc.n_rays
    mrays = n_rays / (ms_per_step * 1e-3) / 1e6
    rejected = st.attempts - st.steps
    alg_flops = ALG_FLOPS_PER_ACCEPTED_STEP * st.steps + ALG_FLOPS_PER_REJECTED_ATTEMPT * rejected
    achieved = alg_flops / (trace_ms * 1e-3) / 1e12
    rhs = 6 * st.attempts + n_rays
    line = {
        "metric": "Mrays/s + DP5 RHS evals/s at 1024^2 a=0.998 (FP64); batched/plain; grad/fwd time",
        "value": round(mrays, 4),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain Kerr frame {sc.width}x{sc.height}, a=0.998, i=60 deg, "
                               f"r_cam=1000M, NT disc ISCO-50M @0.3 Edd, 250k stars "
                               f"(reference generator), rtol 1e-9 / atol 1e-11, 1 spp"},
        "rhs_evals_per_s": rhs / (ms_per_step * 1e-3),
        "accepted_steps_per_ray": st.steps / n_rays,
        "rejected_attempts": int(rejected),
        "trace_ms": round(trace_ms, 4),
        "shade_ms": round(st.shade_ms, 4),
        "batched_over_plain": None,
        "grad_over_fwd": None,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": None, "kernel": "k_trace",
                     "flops_per_launch": alg_flops},
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(sc, threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
