"""Per-ray parity checker: HIP tracer output against the CPU oracle.

The bar (one ray = one sample):
  * termination class identical, except rays within EPS of a class boundary;
  * hit radius (disc, neutron-star surface, Sun: the root located on the
    dense output) and redshift g within REL relative, escape direction within
    REL (unit vectors), except rays whose oracle value itself moves by more
    than the GPU/CPU difference under an EPS displacement in the image plane
    (the ray is ill-conditioned at that scale: near-critical photon-ring rays).
    The radius where an escaping or captured ray stopped is not a hit: it is
    wherever the last step landed past r_escape / r_stop, and moves with the
    step sizes, which differ between any two implementations at ~1e-7
    relative because the embedded error estimate is a cancellation;
  * E = -p_t and L_z = p_phi equal to 1e-15 relative to the size of the
    terms that form them at the camera (the GPU forms the photon with FMA);
    along a ray both are carried unchanged, so their drift is exactly zero in
    both implementations.
"Within EPS of a class boundary" is made operational: the oracle re-traces the
ray at image-plane offsets (u +- d, v) and (u, v +- d), u, v in [-1, 1]
(src/render.cpp:447-448), for d = EPS, EPS/10, EPS/100 and EPS/1000; if any
of those rays ends in another class than the oracle's own, the ray is exempt
from the class check, and the value spread is the largest over all of them.
The smaller offsets sample the EPS neighbourhood where the result is not a
smooth function of the ray: a ray that grazes a pole of the Boyer-Lindquist
coordinates, where the reference clamps sin(theta) at 1e-9 (src/kerr.h:34),
can change its accepted-step count under a 1e-11 displacement, and with it
its escape direction by ~1e-8, while the four rays at exactly +-EPS happen to
land on the same side of that jump.
"""
import numpy as np

import oracle

REL = 1e-9
EPS = 1e-8


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def compare(scene, ids, gpu, cpu, cat=None, rel=REL, eps=EPS, threads=0):
    """gpu / cpu: dicts of numpy arrays over the rays `ids` (cls, y[8, n], g,
    dir[3, n]).  Returns a report dict; report['fail'] is the number of rays
    that break the bar without an exemption."""
    ids = np.asarray(ids)
    cg, cc = gpu["cls"].astype(int), cpu["cls"].astype(int)
    n = len(ids)
    cls_bad = cg != cc
    same = ~cls_bad
    hit_r = same & ((cc == 1) | (cc == 2) | (cc == 3))
    r_rel = np.where(hit_r, _rel(gpu["y"][1], cpu["y"][1]), 0.0)
    hit_g = same & ((cc == 1) | (cc == 3))
    g_rel = np.where(hit_g & (cpu["g"] != 0), _rel(gpu["g"], cpu["g"]), 0.0)
    sky = same & (cc == 4)
    d_err = np.where(sky, np.linalg.norm(gpu["dir"] - cpu["dir"], axis=0), 0.0)
    val_diff = np.maximum(np.maximum(r_rel, g_rel), d_err)
    val_bad = same & (val_diff > rel)
    # p_phi = g_tp p^t + g_pp p^phi at the camera: a sum of two terms of size
    # ~ r_cam |p_t|, which cancel for rays passing over a pole
    const_bad = (_rel(gpu["y"][4], cpu["y"][4]) > 1e-15) | (
        np.abs(gpu["y"][7] - cpu["y"][7]) > 1e-15 * (np.abs(cpu["y"][7]) + scene.cam_r * np.abs(cpu["y"][4])))

    suspects = np.nonzero(cls_bad | val_bad)[0]
    exempt = np.zeros(n, bool)
    if len(suspects):
        u, v = oracle.ray_uv(scene, ids[suspects])
        scales = eps * np.array([1.0, 1e-1, 1e-2, 1e-3])
        du = np.concatenate([[d, -d, 0.0, 0.0] for d in scales])
        dv = np.concatenate([[0.0, 0.0, d, -d] for d in scales])
        m = len(du)
        U = (u[:, None] + du[None, :]).ravel()
        V = (v[:, None] + dv[None, :]).ravel()
        probe = oracle.trace_uv(scene, U, V, cat=cat, threads=threads)
        pc = probe["cls"].reshape(-1, m).astype(int)
        centre = cc[suspects]
        boundary = (pc != centre[:, None]).any(axis=1)
        # conditioning of the value itself under the same displacements
        pr = probe["y"][1].reshape(-1, m)
        spread_r = np.where(np.isin(centre, (1, 2, 3)), np.abs(pr - cpu["y"][1][suspects, None]).max(axis=1) / np.maximum(
            np.abs(cpu["y"][1][suspects]), 1e-300), 0.0)
        pg = probe["g"].reshape(-1, m)
        spread_g = np.abs(pg - cpu["g"][suspects, None]).max(axis=1) / np.maximum(
            np.abs(cpu["g"][suspects]), 1e-300)
        pdir = probe["dir"].reshape(3, -1, m)
        spread_d = np.linalg.norm(pdir - cpu["dir"][:, suspects, None], axis=0).max(axis=1)
        spread = np.maximum(np.maximum(spread_r, np.where(cpu["g"][suspects] != 0, spread_g, 0)),
                            np.where(centre == 4, spread_d, 0))
        ill = spread >= val_diff[suspects]
        exempt[suspects] = np.where(cls_bad[suspects], boundary, ill | boundary)
    fail = (cls_bad | val_bad) & ~exempt
    good = same & ~val_bad
    return {
        "n": int(n),
        "class_mismatch": int(cls_bad.sum()),
        "value_mismatch": int(val_bad.sum()),
        "exempt": int(exempt.sum()),
        "fail": int(fail.sum()) + int(const_bad.sum()),
        "const_mismatch": int(const_bad.sum()),
        "max_r_rel": float(r_rel[good].max()) if good.any() else 0.0,
        "max_const_rel": float(max(_rel(gpu["y"][4], cpu["y"][4]).max(), 0.0)) if n else 0.0,
        "max_g_rel": float(g_rel[good].max()) if good.any() else 0.0,
        "max_dir_err": float(d_err[good].max()) if good.any() else 0.0,
        "fail_ids": ids[fail][:20].tolist(),
    }


def gpu_rays_numpy(rays, shaded):
    out = {k: rays[k].cpu().numpy() for k in ("cls", "y", "steps", "attempts", "norm_err", "carter")}
    out["g"] = shaded["g"].cpu().numpy()
    out["dir"] = shaded["dir"].cpu().numpy()
    out["ray_xyz"] = shaded["ray_xyz"].cpu().numpy()
    return out


def srgb_lsb_diff(xyz_gpu, xyz_cpu, exposure, tonemap=4, log_decades=16.0):
    """Max |sRGB8 difference| when both frames are developed with one exposure."""
    a = oracle.develop(xyz_gpu, exposure, tonemap, log_decades).astype(int)
    b = oracle.develop(xyz_cpu, exposure, tonemap, log_decades).astype(int)
    return int(np.abs(a - b).max()) if a.size else 0
