"""Benchmark: plain FP64 Kerr frame at 1024x1024, a=0.998, 250k-star sky.

One step = one pass of the hot path over one frame: the persistent tracer over
all 1,048,576 rays plus the shading pass (disc emission, star-cell lookup,
colour), inputs (scene, star catalogue, colour table) already resident on the
GPU.  The frame: a = 0.998, inclination 60 deg, camera at 1000 M, +-55 M image
plane, NT disc ISCO-50 M at 0.3 Eddington, 250,000 synthetic stars uniform on
the sphere (m <= 8, B-V ~ N(0.6, 0.4)), rtol 1e-10 / atol 1e-12.  Prints one
JSON line (the fields are defined where they are computed below).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP64 operation counts of the reference's path (src/geodesic.h,
# src/kerr.h, src/render.cpp), counted by instantiating the oracle's
# restatement with an operation-counting scalar (+ - * / sqrt and each
# transcendental = 1; tests/test_flop_counts.py checks these against it):
# per RHS evaluation; per accepted step (6 RHS + stages, error norm,
# controller, dense output + the loop's invariants, end point, hmax and disc
# sign test); per rejected attempt (6 RHS + stages, error norm, controller).
ALG_FLOPS_PER_RHS = 151
ALG_FLOPS_PER_ACCEPTED_STEP = 1659
ALG_FLOPS_PER_REJECTED_ATTEMPT = 1417
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector peak (AMD published; FP64 matrix peak is the same)
# HBM bytes per tracer launch at this workload (the frame has no jets and no
# Sun, so the library runs the tracer kernel without those branches,
# k_trace_lean), from rocprofv3 FETCH_SIZE and WRITE_SIZE runs scaled by a copy
# kernel of known size (the file says how).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_trace_sky_frame_1024_traffic.txt")
# Executed FP64 operations per k_trace_lean launch at this workload, from the
# rocprofv3 counters SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 (FMA = 2) scaled
# to lanes by the measured lane utilisation (the file says how).
COUNTERS_FILE = os.path.join(ROOT, "profiles", "k_trace_sky_frame_1024_counters.txt")


def profile_values(path):
    """key = value lines of a profile summary, as floats ({} if absent)."""
    vals = {}
    try:
        with open(path) as f:
            for ln in f:
                if "=" in ln and not ln.lstrip().startswith("#"):
                    k, v = ln.split("=", 1)
                    try:
                        vals[k.strip()] = float(v.split()[0])
                    except (ValueError, IndexError):
                        pass
    except OSError:
        pass
    return vals


def traffic_per_launch():
    """HBM bytes per tracer launch recorded in TRAFFIC_FILE, or None."""
    v = profile_values(TRAFFIC_FILE)
    if "fetch_bytes_per_launch" in v and "write_bytes_per_launch" in v:
        return v["fetch_bytes_per_launch"] + v["write_bytes_per_launch"]
    return None


def executed_flops_per_launch():
    """Counter-measured FP64 operations per tracer launch (COUNTERS_FILE), or None."""
    return profile_values(COUNTERS_FILE).get("executed_fp64_flops_per_launch")


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


def cpu_baseline(sc, threads, cat=None):
    """The oracle (CPU restatement of the reference) on every ray of the same
    frame with `threads` host threads (the reference's row-pool model); `cat`
    an oracle.Cat star catalogue (None: the reference's generator)."""
    import numpy as np
    import oracle
    ids = np.arange(sc.n_rays)
    t0 = time.perf_counter()
    out = oracle.trace_list(sc, ids, cat=cat, threads=threads)
    dt = time.perf_counter() - t0
    steps = int(out["steps"].sum())
    att = int(out["attempts"].sum())
    flops = ALG_FLOPS_PER_ACCEPTED_STEP * steps + ALG_FLOPS_PER_REJECTED_ATTEMPT * (att - steps)
    return {
        "value": len(ids) / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
        "sample": f"the whole {sc.width}x{sc.height} frame ({len(ids)} rays, {steps} accepted steps)"
                  f" traced+shaded by oracle/ (bitwise-equal CPU restatement) in {dt:.2f} s",
        "rhs_evals_per_s": (6 * att + len(ids)) / dt,
        "accepted_steps": steps, "rejected_attempts": att - steps,
        "alg_gflops": round(flops / dt / 1e9, 2),
        "nproc": os.cpu_count(), "cpu_model": cpu_model(),
    }


def small_frame_leg(tr, torch, threads, rounds=20, cpu=True):
    """The reference's CPU-runnable frame: 256x256, a = 0.9, camera at 500 M
    and 80 deg, +-20 M image plane, NT disc ISCO-20 M at 0.1 Eddington, no
    stars, rtol 1e-10 / atol 1e-12.  256^2 rays are fewer than the resident
    lanes, so the GPU frame is one ray lifetime long; reported beside the
    headline, not as it."""
    from blackhole_amd import scene as S
    sc = S.plain_frame()
    rays = tr.alloc_rays(sc, full=False)
    out = {"xyz": torch.empty((sc.height, sc.width, 3), dtype=torch.float64, device="cuda:0")}
    for _ in range(3):
        tr.trace_async(sc, rays)
        tr.shade_async(sc, rays, out)
    _, st = tr.trace(sc, rays)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(rounds):
        tr.trace_async(sc, rays)
        tr.shade_async(sc, rays, out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / rounds * 1e3
    res = {"workload": "plain Kerr frame 256x256, a=0.9, i=80 deg, r_cam=500M, +-20M, NT disc ISCO-20M "
                       "@0.1 Edd, no stars, rtol 1e-10 / atol 1e-12, 1 spp",
           "ms_per_frame": round(ms, 4), "mrays_s": round(sc.n_rays / (ms * 1e-3) / 1e6, 3),
           "accepted_steps_per_ray": st.steps / sc.n_rays}
    if cpu:
        res["cpu_baseline"] = cpu_baseline(sc, threads)
    return res


def cli_frame_leg(tr, torch, steps=20, warmup=5):
    """The same 1024^2 frame as the reference's CLI renders it: its default
    tolerances (rtol 1e-9 / atol 1e-11) and its own star generator
    (StarField(250000, seed): 45% of the stars near a galactic plane, m up to
    12).  Trace + shade per step; reported beside the headline."""
    from blackhole_amd import scene as S
    sc = S.starry_frame(1024, 1024)
    tr.set_catalogue()     # the reference's generator
    rays = tr.alloc_rays(sc, full=False)
    out = {"xyz": torch.empty((sc.height, sc.width, 3), dtype=torch.float64, device="cuda:0")}
    for _ in range(warmup):
        tr.trace_async(sc, rays)
        tr.shade_async(sc, rays, out)
    _, st = tr.trace(sc, rays)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tr.trace_async(sc, rays)
        tr.shade_async(sc, rays, out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    kern = [tr.trace(sc, rays)[1].trace_ms for _ in range(5)]
    rejected = st.attempts - st.steps
    flops = ALG_FLOPS_PER_ACCEPTED_STEP * st.steps + ALG_FLOPS_PER_REJECTED_ATTEMPT * rejected
    trace_ms = sum(kern) / len(kern)
    return {"workload": "plain Kerr frame 1024x1024, a=0.998, i=60 deg, r_cam=1000M, +-55M, NT disc ISCO-50M "
                        "@0.3 Edd, 250k stars from the reference's StarField generator, rtol 1e-9 / atol 1e-11 "
                        "(the reference CLI's defaults), 1 spp",
            "ms_per_frame": round(ms, 4), "mrays_s": round(sc.n_rays / (ms * 1e-3) / 1e6, 3),
            "trace_ms": round(trace_ms, 4), "accepted_steps_per_ray": st.steps / sc.n_rays,
            "k_trace_frac_fp64_peak": round(flops / (trace_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4)}


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


def adjoint_leg(tr, torch, width=512, Ks=(1, 16, 32, 64, 128, 256), rounds=3, fd=True, threads=16):
    """Differentiable disc image at width^2 (a=0.7, i=45 deg, 0.1 Edd, r=500M;
    target a=0.9, i=50 deg): loss = sum (G*I - G*I_target)^2 with a Gaussian
    sigma = 2 px.  Times a forward render (the plain path: trace + shade); the
    library's autotuner picks the checkpoint interval K among Ks and then the
    recording forward's waves per SIMD, and a gradient pass (recording
    forward + adjoint) is timed at that K and at each candidate; optionally checks the gradient against central differences
    (h = 1e-6 relative) of the oracle's images."""
    import math
    import numpy as np
    from blackhole_amd import scene as S
    from blackhole_amd.diff import DiscRender, gaussian_blur, with_params
    P0 = (0.7, math.radians(45.0), 0.1, 500.0)
    PT = (0.9, math.radians(50.0), 0.1, 500.0)
    sc = S.disc_frame(width, width)
    s0 = with_params(sc, P0)
    dev = "cuda:0"
    rays = tr.alloc_rays(s0, full=False)
    out = {"xyz": torch.empty((s0.height, s0.width, 3), dtype=torch.float64, device=dev)}
    for _ in range(2):
        tr.trace_async(s0, rays)
        tr.shade_async(s0, rays, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(rounds):
        tr.trace_async(s0, rays)
        tr.shade_async(s0, rays, out)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) / rounds * 1e3
    res = {"workload": f"{width}x{width} disc image, a=0.7, i=45 deg, 0.1 Edd, r_obs=500M; target "
                       f"a=0.9, i=50 deg; L = sum (G*I - G*I_t)^2, Gaussian sigma 2 px",
           "forward_render_ms": round(fwd_ms, 3), "per_K": {}}
    dr = DiscRender(tr, sc, K=64)
    It = dr.forward_raw(PT).view(width, width).clone()
    I0 = dr.forward_raw(P0).view(width, width).clone().requires_grad_(True)
    L = ((gaussian_blur(I0) - gaussian_blur(It)) ** 2).sum()
    L.backward()
    w = I0.grad.reshape(-1).contiguous()
    loss = float(L)
    dr.forward_raw(P0)
    tuned_K, tuned_ms = dr.autotune(w, Ks=Ks, reps=2)
    res["autotuned_K"] = tuned_K
    res["autotuned_forward_waves_per_simd"] = dr.info()["fwd_waves"]
    res["autotune_ms_per_pass"] = round(tuned_ms, 3)

    def grad_pass():
        dr.forward_raw(P0)
        g = dr.backward_raw(w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(rounds):
            dr.forward_raw(P0)
            g = dr.backward_raw(w)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / rounds * 1e3, g

    for K in Ks:   # every candidate, for the record
        dr.autotune(w, Ks=(K,), reps=1)
        ms, g = grad_pass()
        info = dr.info()
        res["per_K"][str(K)] = {"grad_pass_ms": round(ms, 3), "record_forward_ms": round(info["fwd_ms"], 3),
                                "adjoint_ms": round(info["bwd_ms"], 3), "checkpoint_bytes": info["ck_bytes"],
                                "replay_mismatches": info["mismatches"], "forward_waves_per_simd": info["fwd_waves"],
                                "grad_over_fwd": round(ms / fwd_ms, 3)}
    dr.autotune(w, Ks=(tuned_K,), reps=1)
    ms, g = grad_pass()
    best = (tuned_K, ms, g)
    res["forward_waves_per_simd"] = dr.info()["fwd_waves"]
    dr.close()
    res["best_K"] = best[0]
    res["grad_pass_ms"] = round(best[1], 3)
    res["grad_over_fwd"] = round(best[1] / fwd_ms, 3)
    res["loss"] = loss
    res["grad"] = dict(zip(("spin", "cam_theta", "eddington", "cam_r"), (float(v) for v in best[2])))
    if fd:
        # central differences of the oracle's images at h = 1e-6 relative (and
        # h/2 for a Richardson estimate: the photon-ring rays make the image a
        # strongly curved function of the camera, so the h = 1e-6 difference
        # quotient itself carries an O(h^2) error of ~1e-5 in the smallest
        # component), over the pixels whose class is the same in every render
        import oracle
        t0 = time.perf_counter()
        wn = w.cpu().numpy()
        I_c, c0 = oracle.disc_intensity(s0, threads=threads)
        dr = DiscRender(tr, sc, K=best[0])
        dr.forward_raw(P0)
        stable = c0["cls"] == dr.rays()["cls"]
        diffs = {}
        for hr in (1e-6, 5e-7):
            for k in range(4):
                h = hr * abs(P0[k])
                pp, pn = list(P0), list(P0)
                pp[k] += h
                pn[k] -= h
                Ip, cp = oracle.disc_intensity(with_params(sc, pp), threads=threads)
                In, cn = oracle.disc_intensity(with_params(sc, pn), threads=threads)
                stable &= (cp["cls"] == c0["cls"]) & (cn["cls"] == c0["cls"])
                diffs[(hr, k)] = (Ip - In) / (2 * h)
        fd1 = np.array([np.sum(np.where(stable, wn * diffs[(1e-6, k)], 0.0)) for k in range(4)])
        fd2 = np.array([np.sum(np.where(stable, wn * diffs[(5e-7, k)], 0.0)) for k in range(4)])
        fdr = (4.0 * fd2 - fd1) / 3.0
        wm = torch.from_numpy(np.where(stable, wn, 0.0)).to(dev)
        ga = dr.backward_raw(wm)
        dr.close()
        res["fd_check"] = {"h_rel": 1e-6, "stable_pixels": int(stable.sum()), "pixels": int(stable.size),
                           "adjoint": ga.tolist(), "oracle_fd": fd1.tolist(),
                           "rel_err": (np.abs(ga - fd1) / np.abs(fd1)).tolist(),
                           "oracle_fd_richardson_1e-6_5e-7": fdr.tolist(),
                           "rel_err_richardson": (np.abs(ga - fdr) / np.abs(fdr)).tolist(),
                           "seconds": round(time.perf_counter() - t0, 2)}
    return res


def sweep_leg(tr, torch, n_frames=120, width=1024, K=64, rounds=2, grad=True, fd_frames=(0, 40, 80, 119)):
    """The quasar sweep: n_frames frames at width^2 (a = 0.998, inclination
    0.5 -> 90 deg, NT disc ISCO-1000 M at 0.5 Edd, conical jets 5 deg Gamma 10
    (r0/r)^2 to 500 M, 250k stars, camera at 2500 M) as ONE batched
    persistent launch replayed as a hipGraph; then dL/dtheta of every frame
    from the discrete adjoint, L_k = sum (G*I_k - G*I_(k-1))^2 (G a Gaussian
    of sigma 2 px, I the disc + jet intensity, I_(-1) = 0), from one batched
    recording forward over all frames and one batched adjoint.  A few frames'
    dL/dtheta are checked against central differences (h = 1e-6 rad) of the
    same forward over the pixels whose class holds."""
    import numpy as np
    from blackhole_amd import Batch, scene as S
    from blackhole_amd.diff import DiscBatch, DiscRender, gaussian_blur
    frames = S.quasar_sweep(n_frames, width, width)
    n_rays = sum(sc.n_rays for sc in frames)
    res = {"workload": f"{n_frames} frames {width}x{width}: a=0.998, inclination 0.5->90 deg, NT disc "
                       f"ISCO-1000M @0.5 Edd, conical jets 5 deg Gamma 10 (r0/r)^2 6-500M, 250k stars "
                       f"uniform on the sphere (the headline frame's sky), r_cam 2500M, rtol 1e-9 / "
                       f"atol 1e-11; one batched launch",
           "frames": n_frames, "rays": n_rays}
    with Batch(tr, frames) as b:
        b.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(rounds):
            b.run()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / rounds * 1e3
        stats, dev_ms = b.stats()
    res.update({"sweep_ms": round(ms, 2), "sweep_device_ms": round(dev_ms, 2),
                "sweep_mrays_s": round(n_rays / (ms * 1e-3) / 1e6, 3),
                "accepted_steps": int(sum(st.steps for st in stats)),
                "attempts": int(sum(st.attempts for st in stats)),
                "jet_rays": int(sum(st.jet_rays for st in stats))})
    if not grad:
        return res
    # every frame's gradient from one batched adjoint: one recording forward
    # over all frames' rays, dL_k/dI_k per frame, one backward
    db = DiscBatch(tr, frames, K=K)
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    I = db.forward_raw()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter()
    # dL_k/dI_k, 8 frames at a time: L = sum_k L_k with each frame's target
    # (the blurred previous frame) held fixed
    w = torch.empty_like(I)
    prev = torch.zeros(width, width, dtype=torch.float64, device=I.device)
    for g0 in range(0, n_frames, 8):
        Ig = I[g0:g0 + 8].view(-1, width, width).clone().requires_grad_(True)
        Bg = gaussian_blur(Ig)
        Tg = torch.cat([prev[None], Bg[:-1].detach()])
        L = ((Bg - Tg) ** 2).sum()
        L.backward()
        w[g0:g0 + 8] = Ig.grad.view(-1, I.shape[1])
        prev = Bg[-1].detach()
    torch.cuda.synchronize()
    t_w = time.perf_counter()
    g = db.backward_raw(w)
    grad_s = time.perf_counter() - t_all
    info = db.info()
    res.update({"grad_K": K, "grad_total_s": round(grad_s, 3),
                "grad_forward_s": round(t_fwd - t_all, 3), "grad_loss_weights_s": round(t_w - t_fwd, 3),
                "grad_backward_s": round(time.perf_counter() - t_w, 3),
                "grad_frame_ms_mean": round(grad_s * 1e3 / n_frames, 2),
                "grad_mrays_s": round(n_rays / grad_s / 1e6, 3),
                "checkpoint_bytes": info["ck_bytes"], "replay_mismatches": info["mismatches"],
                "dL_dtheta": [float(f"{v:.6g}") for v in g[:, 1]]})
    # a few frames against central differences of the same forward over the
    # pixels whose class holds: the batch's backward with dL/dI zeroed
    # outside them (and outside the checked frames), and one frame through
    # the single-frame adjoint for comparison
    dr = DiscRender(tr, frames[0], K=K)
    wm_all = torch.zeros_like(w)
    checks = {}
    for k in fd_frames:
        sc = frames[k]
        p = (sc.spin, sc.cam_theta, sc.eddington, sc.cam_r)
        h = 1e-6
        masks, Is = [], []
        for sgn in (0, 1, -1):
            q = (p[0], p[1] + sgn * h, p[2], p[3])
            Is.append(dr.forward_raw(q).clone())
            masks.append(dr.rays()["cls"])
        stable = torch.from_numpy((masks[1] == masks[0]) & (masks[2] == masks[0])).to(w.device)
        wm_all[k] = torch.where(stable, w[k], torch.zeros_like(w[k]))
        fd = float(torch.sum(wm_all[k] * (Is[1] - Is[2])) / (2 * h))
        dr.forward_raw(p)
        single = float(dr.backward_raw(wm_all[k])[1])
        checks[k] = {"frame": k, "incl_deg": round(float(np.degrees(p[1])), 3), "fd": fd, "single_frame": single,
                     "stable_pixels": int(stable.sum())}
    gm = db.backward_raw(wm_all)
    for k, c in checks.items():
        c["batch"] = float(gm[k, 1])
        c["rel_err"] = abs(c["batch"] - c["fd"]) / abs(c["fd"]) if c["fd"] else None
        c["batch_vs_single_rel"] = abs(c["batch"] - c["single_frame"]) / abs(c["single_frame"]) \
            if c["single_frame"] else None
    res["fd_check_theta"] = [checks[k] for k in fd_frames]
    dr.close()
    db.close()
    return res


def png_leg(tr, torch, sc, xyz, rounds=5, cpu=True):
    """The frame's PNG (img::write_png, src/image.cpp:193-778) from device
    memory: auto exposure + develop on the device, then the GPU-assisted
    writer (scanline filters and the LZ77 match search on the device, parse
    and Huffman coding on the host).  Checked byte for byte against the
    oracle's restatement of the reference's writer, which is also timed as
    the single-threaded CPU baseline (the reference's writer is one thread)."""
    e = tr.auto_exposure(xyz)
    rgb = tr.develop(xyz, e).view(sc.height, sc.width, 3)
    b = tr.png_encode(rgb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(rounds):
        b = tr.png_encode(rgb)
    ms = (time.perf_counter() - t0) / rounds * 1e3
    res = {"workload": f"img::write_png of the developed {sc.width}x{sc.height} frame "
                       f"({3 * sc.width * sc.height} bytes of sRGB in device memory)",
           "png_bytes": len(b), "gpu_assisted_ms": round(ms, 3),
           "mb_per_s": round(3 * sc.width * sc.height / (ms * 1e-3) / 1e6, 1)}
    if cpu:
        import oracle
        host = rgb.cpu().numpy()
        t0 = time.perf_counter()
        ref = oracle.png_encode(host)
        cms = (time.perf_counter() - t0) * 1e3
        res["identical_to_oracle"] = ref == b
        res["cpu_baseline"] = {"value": round(cms, 3), "unit": "ms", "cores": 1, "kind": "port",
                               "sample": "the same frame through oracle/ (the reference writer restated, "
                                         "byte-identical to it), one thread like the reference"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-launch leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiation leg")
    ap.add_argument("--no-fd", action="store_true", help="skip the finite-difference check of the gradient")
    ap.add_argument("--no-sweep", action="store_true", help="skip the quasar-sweep leg")
    ap.add_argument("--no-png", action="store_true", help="skip the PNG-writer leg")
    ap.add_argument("--no-small-frame", action="store_true", help="skip the 256x256 plain-frame leg")
    ap.add_argument("--no-cli-frame", action="store_true",
                    help="skip the frame at the CLI's tolerances with the reference's star generator")
    ap.add_argument("--sweep-frames", type=int, default=120)
    ap.add_argument("--sweep-width", type=int, default=1024)
    ap.add_argument("--no-sweep-grad", action="store_true", help="sweep leg without the per-frame gradients")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 256 MiB device copy first (for HBM-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only (--gpus 1): the path shards by image tiles or by "
              "frames across GPUs, and no multi-GPU figure is measured", file=sys.stderr)
        sys.exit(2)

    import torch
    from blackhole_amd import Tracer, scene as S

    sc = S.sky_frame(args.width, args.width)
    sky = S.uniform_sky(sc.star_count, sc.seed)
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        copy_calibration(torch, dev)
    tr = Tracer(0)
    tr.set_catalogue(*sky)
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
    # device time of the whole trace call: counter fill, the cost-ordered
    # queue's pilot + sort + scan, k_raygen and the tracer kernel
    trace_call_ms = sum(a.elapsed_time(b) for a, b in ev) / args.steps
    # the tracer kernel alone (and the queue ordering alone), from the HIP events the
    # library records on the launch stream around each kernel; synchronous
    # calls, so kept out of the timed loop above
    reps = max(1, min(args.steps, 20))
    kern, order = [], []
    for _ in range(reps):
        _, sk = tr.trace(sc, rays)
        kern.append(sk.trace_ms)
        order.append(sk.order_ms)
    trace_ms = sum(kern) / reps
    order_ms = sum(order) / reps
    # end to end: the frame to 8-bit sRGB on the device (trace, shade, then
    # auto exposure - a radix selection - and develop, the exposure kept on
    # the device so the frames queue without a host round trip)
    e2e_steps = max(1, min(args.steps, 10))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(e2e_steps):
        tr.trace_async(sc, rays)
        tr.shade_async(sc, rays, out)
        rgb = tr.develop_auto(out["xyz"])
    torch.cuda.synchronize()
    e2e_ms = (time.perf_counter() - t0) / e2e_steps * 1e3
    del rgb

    ms_per_step = wall / args.steps * 1e3
    n_rays = sc.n_rays
    mrays = n_rays / (ms_per_step * 1e-3) / 1e6
    rejected = st.attempts - st.steps
    alg_flops = ALG_FLOPS_PER_ACCEPTED_STEP * st.steps + ALG_FLOPS_PER_REJECTED_ATTEMPT * rejected
    achieved = alg_flops / (trace_ms * 1e-3) / 1e12
    # RHS evaluations: 6 per attempt, the first stage of each ray (k_raygen)
    # and one at the end of each escaped ray (its escape direction, k_shade)
    rhs = 6 * st.attempts + n_rays + st.escaped
    ex_flops = executed_flops_per_launch()
    executed = ex_flops / (trace_ms * 1e-3) / 1e12 if ex_flops else None
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
                               f"r_cam=1000M, +-55M image plane, 10 Msun, NT disc ISCO-50M @0.3 Edd, "
                               f"250k stars uniform on the sphere (m<=8, B-V~N(0.6,0.4), seed "
                               f"{sc.seed}), rtol 1e-10 / atol 1e-12, 1 spp"},
        "rhs_evals_per_s": rhs / (ms_per_step * 1e-3),
        "end_to_end_ms": round(e2e_ms, 4),
        "end_to_end_mrays_s": round(n_rays / (e2e_ms * 1e-3) / 1e6, 3),
        "accepted_steps_per_ray": st.steps / n_rays,
        "rejected_attempts": int(rejected),
        "trace_ms": round(trace_ms, 4),
        "trace_call_ms": round(trace_call_ms, 4),
        "queue_order_ms": round(order_ms, 4),
        "queue_order": "cost" if order_ms > 0 else "centre-first",
        "shade_ms": round(st.shade_ms, 4),
        "batched_over_plain": None,
        "grad_over_fwd": None,
        "roofline": {"bound": "fp64-vector", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": traffic_per_launch(), "kernel": "k_trace_lean",
                     "pipe": "FP64 compute-bound on the vector ALUs: no MFMA (no dense "
                             "contraction on this path); 78.6 TFLOP/s is MI355X's FP64 vector "
                             "peak, equal to its FP64 matrix peak",
                     "flops_per_launch": alg_flops,
                     "executed": None if executed is None else round(executed, 3),
                     "executed_frac": None if executed is None else round(executed / FP64_PEAK_TFLOPS, 4),
                     "executed_flops_per_launch": ex_flops},
    }
    if not args.no_png:
        line["png"] = png_leg(tr, torch, sc, out["xyz"], cpu=not args.no_cpu)
    if not args.no_cli_frame:
        line["cli_frame"] = cli_frame_leg(tr, torch)
        tr.set_catalogue(*sky)
    if not args.no_small_frame:
        line["small_frame"] = small_frame_leg(tr, torch, min(16, os.cpu_count() or 1), cpu=not args.no_cpu)
    if not args.no_batch:
        line["batched"] = batched_leg(tr, torch)
        line["batched_over_plain"] = line["batched"]["batched_over_plain"]
    if not args.no_adjoint:
        line["adjoint"] = adjoint_leg(tr, torch, fd=not args.no_fd, threads=min(16, os.cpu_count() or 1))
        line["grad_over_fwd"] = line["adjoint"]["grad_over_fwd"]
    if not args.no_sweep:
        nf = args.sweep_frames
        line["sweep"] = sweep_leg(tr, torch, n_frames=nf, width=args.sweep_width, grad=not args.no_sweep_grad,
                                  fd_frames=tuple(sorted({0, nf // 3, (2 * nf) // 3, nf - 1})))
    if not args.no_cpu:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(sc, threads, cat=oracle.Cat(*oracle.uniform_sky(sc.star_count, sc.seed)))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
