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
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector peak (AMD published; FP64 matrix peak is the same)
# HBM bytes per k_trace launch at this workload, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs scaled by a copy kernel of known size (the file says how).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_trace_starry_frame_1024_traffic.txt")


def traffic_per_launch():
    """HBM bytes per k_trace launch recorded in TRAFFIC_FILE, or None."""
    try:
        vals = {}
        with open(TRAFFIC_FILE) as f:
            for ln in f:
                if "=" in ln and not ln.lstrip().startswith("#"):
                    k, v = ln.split("=", 1)
                    vals[k.strip()] = float(v.split()[0])
        return vals["fetch_bytes_per_launch"] + vals["write_bytes_per_launch"]
    except (OSError, KeyError, ValueError, IndexError):
        return None


def copy_calibration(torch, dev, nbytes=1 << 28, reps=3):
    """A plain device copy of nbytes (read nbytes, write nbytes), for scaling
    the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    src = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(sc, threads):
    """The oracle (CPU restatement of the reference) on every ray of the same
    frame with `threads` host threads (the reference's row-pool model)."""
    import numpy as np
    import oracle
    ids = np.arange(sc.n_rays)
    t0 = time.perf_counter()
    out = oracle.trace_list(sc, ids, threads=threads)
    dt = time.perf_counter() - t0
    steps = int(out["steps"].sum())
    att = int(out["attempts"].sum())
    return {
        "value": len(ids) / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
        "sample": f"the whole 1024x1024 frame ({len(ids)} rays, {steps} accepted steps)"
                  f" traced+shaded by oracle/ (bitwise-equal CPU restatement) in {dt:.2f} s",
        "rhs_evals_per_s": (6 * att + len(ids)) / dt,
        "cpu_model": cpu_model(),
    }


def batched_leg(tr, torch, rounds=3):
    """64 uneven scenes (8 of them neutron stars): 64 sequential single-scene
    launches (trace + shade each) against one batched launch replayed as a
    hipGraph, after the batch's autotuner has picked its launch shape."""
    from blackhole_amd import Batch, scene as S
    scenes = S.survey_batch()
    n_rays = sum(sc.n_rays for sc in scenes)
    big = max(scenes, key=lambda sc: sc.n_rays)
    rays = tr.alloc_rays(big, full=False)
    out = {"xyz": torch.empty((big.height, big.width, 3), dtype=torch.float64, device="cuda:0")}

    def sequential():
        for sc in scenes:
            tr.trace_async(sc, rays)
            tr.shade_async(sc, rays, out)

    sequential()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(rounds):
        sequential()
    torch.cuda.synchronize()
    seq_ms = (time.perf_counter() - t0) / rounds * 1e3
    with Batch(tr, scenes) as b:
        b.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run()
        torch.cuda.synchronize()
        untuned_ms = (time.perf_counter() - t0) * 1e3
        shape, _ = b.autotune(reps=1)
        b.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(rounds):
            b.run()
        torch.cuda.synchronize()
        bat_ms = (time.perf_counter() - t0) / rounds * 1e3
        stats, _ = b.stats()
    steps = sum(st.steps for st in stats)
    attempts = sum(st.attempts for st in stats)
    return {
        "workload": "64 scenes, seed 20240921: spin U[0,0.998], inclination U[5,85] deg, r_cam "
                    "log-U[100,2000] M, side log-U[128,1024] px, disc log-U[0.01,1] Edd; 8 neutron "
                    "stars (Gamma=2 K=100 polytrope, 1.2-1.99 Msun)",
        "rays": n_rays, "accepted_steps": steps, "attempts": attempts,
        "plain_sequential_ms": round(seq_ms, 3), "batched_untuned_ms": round(untuned_ms, 3),
        "batched_graph_ms": round(bat_ms, 3), "batched_over_plain": round(seq_ms / bat_ms, 3),
        "batched_mrays_s": round(n_rays / (bat_ms * 1e-3) / 1e6, 3),
        "autotuned_launch": shape,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-launch leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 256 MiB device copy first (for HBM-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; per-GPU sharding is described in "
              "DESIGN.md (no multi-GPU figure is measured)", file=sys.stderr)
        sys.exit(2)

    import torch
    from blackhole_amd import Tracer, scene as S

    sc = S.starry_frame(args.width, args.width)
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        copy_calibration(torch, dev)
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
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": traffic_per_launch(), "kernel": "k_trace",
                     "pipe": "FP64 compute-bound on the vector ALUs: no MFMA (no dense "
                             "contraction on this path); 78.6 TFLOP/s is MI355X's FP64 vector "
                             "peak, equal to its FP64 matrix peak",
                     "flops_per_launch": alg_flops},
    }
    if not args.no_batch:
        line["batched"] = batched_leg(tr, torch)
        line["batched_over_plain"] = line["batched"]["batched_over_plain"]
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(sc, threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
