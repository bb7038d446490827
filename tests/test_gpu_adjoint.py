"""Differentiable disc image on the MI355X (bhx_adjoint, blackhole_amd.diff).

* forward: the recording kernel's image equals the oracle's (class per ray,
  I = sigma (g T)^4 / pi within 1e-9 relative on disc hits);
* the replay reproduces every checkpoint bit for bit, so the gradient is
  independent of the checkpoint interval K, bit for bit;
* gradient of the blurred-intensity loss against central finite differences
  (h = 1e-6 relative) of the oracle's images, over the pixels whose class is
  the same in all nine renders (the smooth part of the image);
* the torch.autograd.Function gives the same gradient as the raw backward.
"""
import math

import numpy as np
import pytest

import oracle
from blackhole_amd import scene as S
from blackhole_amd.diff import DiscRender, gaussian_blur, gaussian_blur_np, with_params

from parity import compare

pytestmark = pytest.mark.gpu

P0 = (0.7, math.radians(45.0), 0.1, 500.0)
PT = (0.9, math.radians(50.0), 0.1, 500.0)


def _scene(n):
    return S.disc_frame(n, n)


def oracle_intensity(sc):
    """I per ray from the oracle: sigma (g T(r))^4 / pi on disc hits."""
    return oracle.disc_intensity(sc)


def test_forward_matches_oracle(tracer):
    sc = _scene(64)
    dr = DiscRender(tracer, sc, K=16)
    I = dr.forward_raw(P0).cpu().numpy()
    r = dr.rays()
    s0 = with_params(sc, P0)
    I_c, c = oracle_intensity(s0)
    ids = np.arange(sc.n_rays)
    g = {"cls": r["cls"], "y": r["y"], "g": c["g"], "dir": c["dir"]}
    rep = compare(s0, ids, g, c)
    print(f"[adjoint fwd] {rep}", flush=True)
    assert rep["fail"] == 0, rep
    same = (r["cls"] == c["cls"]) & (I_c > 0)
    rel = np.abs(I[same] - I_c[same]) / I_c[same]
    print(f"[adjoint fwd] disc rays {int(same.sum())}, max rel I {rel.max():.3e}", flush=True)
    assert rel.max() < 1e-9
    assert np.all(I[r["cls"] != 1] == 0.0)


def _loss_weights(tracer, dr, sc):
    import torch
    It = dr.forward_raw(PT).view(sc.height, sc.width).clone()
    p = torch.tensor(P0, dtype=torch.float64, requires_grad=True)
    I = dr(p)
    L = ((gaussian_blur(I) - gaussian_blur(It)) ** 2).sum()
    L.backward()
    return p.grad.numpy().copy(), It.cpu().numpy()


def test_gradient_independent_of_K(tracer):
    import torch
    sc = _scene(48)
    grads = []
    for K in (1, 7, 64):
        dr = DiscRender(tracer, sc, K=K)
        I = dr.forward_raw(P0)
        w = torch.linspace(-1.0, 1.0, sc.n_rays, dtype=torch.float64, device=I.device)
        grads.append(dr.backward_raw(w))
        info = dr.info()
        print(f"[adjoint K={K}] grad {grads[-1]} {info}", flush=True)
        assert info["mismatches"] == 0
        dr.close()
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


def test_gradient_against_oracle_fd(tracer):
    import torch
    n = 96
    sc = _scene(n)
    dr = DiscRender(tracer, sc, K=16)
    grad_fn, It = _loss_weights(tracer, dr, sc)
    # dL/dI at P0 (per pixel) from the loss, on the CPU
    I0 = dr.forward_raw(P0).view(n, n).cpu().numpy()
    It_t = torch.from_numpy(It)
    I0_t = torch.from_numpy(I0).requires_grad_(True)
    L = ((gaussian_blur(I0_t) - gaussian_blur(It_t)) ** 2).sum()
    L.backward()
    w = I0_t.grad.numpy().reshape(-1)
    # oracle renders at P0 and P0 +- h, h = 1e-6 relative
    s0 = with_params(sc, P0)
    _, c0 = oracle_intensity(s0)
    stable = c0["cls"] == dr.rays()["cls"]
    fd = np.zeros(4)
    Ipm = []
    for k in range(4):
        h = 1e-6 * abs(P0[k])
        pp, pm = list(P0), list(P0)
        pp[k] += h
        pm[k] -= h
        Ip, cp = oracle_intensity(with_params(sc, pp))
        Im, cm = oracle_intensity(with_params(sc, pm))
        stable &= (cp["cls"] == c0["cls"]) & (cm["cls"] == c0["cls"])
        Ipm.append((Ip, Im, h))
    for k, (Ip, Im, h) in enumerate(Ipm):
        fd[k] = np.sum(np.where(stable, w * (Ip - Im) / (2 * h), 0.0))
    wm = torch.from_numpy(np.where(stable, w, 0.0)).to(f"cuda:{tracer.device}")
    dr.forward_raw(P0)
    ga = dr.backward_raw(wm)
    rel = np.abs(ga - fd) / np.abs(fd)
    print(f"[adjoint fd] stable {int(stable.sum())}/{sc.n_rays} grad {ga} fd {fd} rel {rel}", flush=True)
    assert np.all(rel < 1e-5), rel
    # the autograd path gives the gradient of the full loss (all pixels)
    wf = torch.from_numpy(w).to(f"cuda:{tracer.device}")
    dr.forward_raw(P0)
    g_full = dr.backward_raw(wf)
    assert np.allclose(g_full, grad_fn, rtol=1e-10, atol=0.0), (g_full, grad_fn)


def test_autograd_backward_after_a_later_forward(tracer):
    """dr(p1), then dr(p2), then backward through the first image: the
    gradient is the one at p1 (the checkpoints of p1 are recorded again),
    and then through the second image, the one at p2; both bit-identical to
    a single forward/backward at each point."""
    import torch
    sc = _scene(32)
    dev = f"cuda:{tracer.device}"
    c = torch.linspace(0.5, 1.5, sc.n_rays, dtype=torch.float64, device=dev).view(sc.height, sc.width)

    def fresh(params):
        d = DiscRender(tracer, sc, K=16)
        q = torch.tensor(params, dtype=torch.float64, requires_grad=True)
        (d(q) * c).sum().backward()
        d.close()
        return q.grad.numpy().copy()

    g1, g2 = fresh(P0), fresh(PT)
    dr = DiscRender(tracer, sc, K=16)
    p1 = torch.tensor(P0, dtype=torch.float64, requires_grad=True)
    p2 = torch.tensor(PT, dtype=torch.float64, requires_grad=True)
    I1 = dr(p1)
    I2 = dr(p2)
    (I1 * c).sum().backward()
    (I2 * c).sum().backward()
    print(f"[adjoint stale] {p1.grad.numpy()} vs {g1}; {p2.grad.numpy()} vs {g2}", flush=True)
    assert not np.array_equal(g1, g2)
    assert np.array_equal(p1.grad.numpy(), g1)
    assert np.array_equal(p2.grad.numpy(), g2)
    dr.close()


def test_checkpoint_pool_grows_to_the_frame(tracer):
    """At rtol 1e-12 the rays take ~3x the steps the first guess of the
    checkpoint pool allows for: the first backward finds the forward short,
    grows the pool to the count the forward made and records it again; the
    gradient is the one of a forward and backward with the grown pool, bit
    for bit, and every replay reproduces its checkpoint."""
    import torch
    sc = S.disc_frame(40, 40, rtol=1e-12, atol=1e-14)
    dr = DiscRender(tracer, sc, K=16)
    I = dr.forward_raw(P0)
    before = dr.info()["ck_bytes"]
    w = torch.linspace(-1.0, 1.0, sc.n_rays, dtype=torch.float64, device=I.device)
    g1 = dr.backward_raw(w)
    info = dr.info()
    steps = dr.rays()["steps"]
    print(f"[pool] mean steps {steps.mean():.0f}, pool {before} -> {info['ck_bytes']} bytes", flush=True)
    assert info["ck_bytes"] > before and info["mismatches"] == 0
    I2 = dr.forward_raw(P0)
    g2 = dr.backward_raw(w)
    assert torch.equal(I, I2) and np.array_equal(g1, g2)
    assert dr.info()["mismatches"] == 0
    # the pool holds what the rays need: ceil(records / 8) chunks of 8 records
    # of 64 B plus a 4-byte link each; a ray of N accepted steps wrote
    # floor(N / K) + 1 records, or one fewer when it ended on a multiple of K
    taken = info["ck_bytes"] // (8 * 64 + 4)
    assert info["ck_bytes"] == taken * (8 * 64 + 4)
    lo = int(np.sum(((np.maximum(steps - 1, 0) // 16 + 1) + 7) // 8))
    hi = int(np.sum(((steps // 16 + 1) + 7) // 8))
    assert lo <= taken <= hi, (lo, taken, hi)
    dr.close()


def test_step_records_pool_grows_to_the_frame(tracer):
    """K = 1 at rtol 1e-12: the rays' step records outgrow the first guess of
    the pool (320 per ray); the first backward grows it to the forward's count
    and records again.  The gradient equals the replaying backward's (K = 16)
    bit for bit, and the pool holds exactly max(1, ceil(steps / 128)) chunks
    of 128 records of 48 B (plus a 4-byte link) per ray."""
    import torch
    sc = S.disc_frame(40, 40, rtol=1e-12, atol=1e-14)
    w = torch.linspace(-1.0, 1.0, sc.n_rays, dtype=torch.float64, device=f"cuda:{tracer.device}")
    ref = DiscRender(tracer, sc, K=16)
    ref.forward_raw(P0)
    g16 = ref.backward_raw(w)
    ref.close()
    dr = DiscRender(tracer, sc, K=1)
    dr.forward_raw(P0)
    before = dr.info()["ck_bytes"]
    g1 = dr.backward_raw(w)
    info = dr.info()
    steps = dr.rays()["steps"].astype(np.int64)
    print(f"[pool K=1] mean steps {steps.mean():.0f}, pool {before} -> {info['ck_bytes']} bytes", flush=True)
    assert info["ck_bytes"] > before and info["mismatches"] == 0
    assert np.array_equal(g1, g16), (g1, g16)
    chunk = 128 * 48 + 4
    assert info["ck_bytes"] % chunk == 0
    assert info["ck_bytes"] // chunk == int(np.sum(np.maximum(1, (steps + 127) // 128)))
    dr.close()


def test_full_frame_at_K16(tracer):
    """The 1024^2 differentiation frame at K = 16 (a fixed max_steps / K + 1
    slab per ray would be 84 GB): the pool is sized by the rays' steps; and
    at K = 1, the same gradient from the step records."""
    import torch
    sc = S.disc_frame(1024, 1024)
    dr = DiscRender(tracer, sc, K=16)
    I = dr.forward_raw(P0)
    w = torch.ones(sc.n_rays, dtype=torch.float64, device=I.device)
    g = dr.backward_raw(w)
    info = dr.info()
    steps = dr.rays()["steps"]
    print(f"[1024 K16] pool {info['ck_bytes'] / 1e9:.2f} GB, lit {info['lit_rays']}, mean steps "
          f"{steps.mean():.0f}, grad {g}", flush=True)
    assert info["mismatches"] == 0 and np.all(np.isfinite(g))
    assert info["ck_bytes"] < 8e9
    dr.close()
    # the same frame at K = 1 (a record of each of its ~260M accepted steps,
    # read back without replaying): the same gradient, bit for bit
    d1 = DiscRender(tracer, sc, K=1)
    d1.forward_raw(P0)
    g1 = d1.backward_raw(w)
    i1 = d1.info()
    print(f"[1024 K1] pool {i1['ck_bytes'] / 1e9:.2f} GB, grad {g1}", flush=True)
    assert i1["mismatches"] == 0 and np.array_equal(g1, g)
    d1.close()


def test_checkpoint_interval_autotuner(tracer):
    """The tuner leaves one of its candidates set; the gradient at the tuned
    K equals the gradient at K = 64 bit for bit."""
    import torch
    sc = _scene(48)
    w = torch.linspace(-1.0, 1.0, sc.n_rays, dtype=torch.float64, device=f"cuda:{tracer.device}")
    ref = DiscRender(tracer, sc, K=64)
    ref.forward_raw(P0)
    g64 = ref.backward_raw(w)
    ref.close()
    dr = DiscRender(tracer, sc, K=7)
    dr.forward_raw(P0)
    K, ms = dr.autotune(w, Ks=(1, 16, 32, 64, 128, 256), reps=1)
    print(f"[autotune K] {K} {ms:.3f} ms", flush=True)
    assert K in (1, 16, 32, 64, 128, 256) and ms > 0.0
    dr.forward_raw(P0)
    assert np.array_equal(dr.backward_raw(w), g64)
    assert dr.info()["mismatches"] == 0
    dr.close()


def _sweep_frames(n, width, picks, rtol=None):
    frames = S.quasar_sweep(n, width, width, stars=False)
    out = []
    for i in picks:
        sc = frames[i]
        if rtol is not None:
            sc.rtol, sc.atol = rtol, rtol * 1e-2
        out.append(sc)
    return out


def test_batch_matches_single_frames(tracer):
    """Three sweep frames (0.5, 45 and 90 deg; disc + conical jet) through one
    batched adjoint: each frame's intensities, classes and step counts equal
    its single-frame adjoint's bit for bit (the same rays through the same
    arithmetic), and so does each frame's gradient with the same dL/dI (the
    same per-ray terms, summed per frame in the same fixed order)."""
    import torch
    from blackhole_amd.diff import DiscBatch
    frames = _sweep_frames(9, 64, (0, 4, 8))
    n_per = frames[0].n_rays
    db = DiscBatch(tracer, frames, K=16)
    I = db.forward_raw()
    r = db.rays()
    w = torch.linspace(-1.0, 1.0, 3 * n_per, dtype=torch.float64, device=I.device).view(3, n_per)
    gb = db.backward_raw(w)
    info = db.info()
    print(f"[batch] {gb.tolist()} {info}", flush=True)
    assert info["mismatches"] == 0 and info["lit_rays"] > 0
    for f, sc in enumerate(frames):
        dr = DiscRender(tracer, sc, K=16)
        If = dr.forward_raw((sc.spin, sc.cam_theta, sc.eddington, sc.cam_r))
        rf = dr.rays()
        assert torch.equal(I[f], If), f
        assert np.array_equal(r["cls"][f], rf["cls"]) and np.array_equal(r["steps"][f], rf["steps"]), f
        gs = dr.backward_raw(w[f])
        print(f"[batch frame {f}] single {gs}", flush=True)
        assert np.array_equal(gb[f], gs), (f, gb[f], gs)
        assert dr.info()["mismatches"] == 0
        dr.close()
    # a frame whose dL/dI is zero gets a zero gradient; the others keep theirs
    w2 = w.clone()
    w2[1] = 0.0
    g2 = db.backward_raw(w2)
    assert np.all(g2[1] == 0.0) and np.array_equal(g2[0], gb[0]) and np.array_equal(g2[2], gb[2])
    db.close()


def test_batch_step_records(tracer):
    """K = 1 (a record of every accepted step, read back without replaying)
    against K = 16 (segments replayed from checkpoints) on three sweep frames
    with the conical jet: the same images and the same per-frame gradients,
    bit for bit."""
    import torch
    from blackhole_amd.diff import DiscBatch
    frames = _sweep_frames(9, 48, (0, 4, 8))
    out = {}
    for K in (16, 1):
        db = DiscBatch(tracer, frames, K=K)
        I = db.forward_raw()
        w = torch.linspace(-1.0, 1.0, I.numel(), dtype=torch.float64, device=I.device).view_as(I)
        g = db.backward_raw(w)
        info = db.info()
        print(f"[batch K={K}] {g.tolist()} {info}", flush=True)
        assert info["mismatches"] == 0 and info["lit_rays"] > 0
        out[K] = (I.clone(), g)
        db.close()
    assert torch.equal(out[1][0], out[16][0]) and np.array_equal(out[1][1], out[16][1])


def test_batch_pool_grows(tracer):
    """At rtol 1e-12 the batch's first pool guess is short: the backward grows
    it and records the same forward again; the gradients equal those of a
    second forward and backward with the grown pool, bit for bit."""
    import torch
    from blackhole_amd.diff import DiscBatch
    frames = _sweep_frames(5, 32, (1, 3), rtol=1e-12)
    db = DiscBatch(tracer, frames, K=16)
    I = db.forward_raw()
    before = db.info()["ck_bytes"]
    w = torch.linspace(-1.0, 1.0, I.numel(), dtype=torch.float64, device=I.device).view_as(I)
    g1 = db.backward_raw(w)
    info = db.info()
    print(f"[batch pool] {before} -> {info['ck_bytes']} bytes, mean steps {db.rays()['steps'].mean():.0f}",
          flush=True)
    assert info["ck_bytes"] > before and info["mismatches"] == 0
    I2 = db.forward_raw()
    g2 = db.backward_raw(w)
    assert torch.equal(I, I2) and np.array_equal(g1, g2)
    db.close()


def test_batch_refusals(tracer):
    """Frames that differ in more than their camera, or in size, are refused;
    a batch refuses a new scene and the single-frame backward."""
    import ctypes as C
    import torch
    from blackhole_amd import BhxError
    from blackhole_amd.diff import DiscBatch
    a, b = _sweep_frames(3, 32, (0, 2))
    b2 = with_params(b, (0.9, b.cam_theta, b.eddington, b.cam_r))
    with pytest.raises(BhxError, match="camera only"):
        DiscBatch(tracer, [a, b2])
    c = _sweep_frames(3, 40, (2,))[0]
    with pytest.raises(BhxError, match="one size"):
        DiscBatch(tracer, [a, c])
    db = DiscBatch(tracer, [a, b], K=16)
    I = db.forward_raw()
    assert tracer.lib.bhx_adjoint_set_scene(db.adj, C.byref(a)) != 0
    g = (C.c_double * 4)()
    w = torch.zeros_like(I)
    assert tracer.lib.bhx_adjoint_backward(db.adj, C.c_void_p(w.data_ptr()), g, None) != 0
    assert np.all(db.backward_raw(w) == 0.0)
    db.close()


def test_forward_queue_shape_leaves_results_unchanged(tracer):
    """The recording forward's queue chunk and refill threshold change the
    order lanes take rays in, not what any ray computes: intensities and
    gradients bit for bit the same; out-of-range shapes are refused."""
    import torch
    from blackhole_amd import BhxError
    sc = _scene(48)
    dr = DiscRender(tracer, sc, K=16)
    I1 = dr.forward_raw(P0).clone()
    w = torch.linspace(-1.0, 1.0, sc.n_rays, dtype=torch.float64, device=I1.device)
    g1 = dr.backward_raw(w)
    for q, r in ((1, 1), (7, 3), (64, 64), (4096, 2)):
        dr.set_launch(q, r)
        I2 = dr.forward_raw(P0)
        assert torch.equal(I1, I2), (q, r)
        assert np.array_equal(dr.backward_raw(w), g1), (q, r)
    with pytest.raises(BhxError):
        dr.set_launch(5000, 0)
    with pytest.raises(BhxError):
        dr.set_launch(0, 65)
    dr.close()
