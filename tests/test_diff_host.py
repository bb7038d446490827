"""Host parts of the differentiable render (blackhole_amd.diff), CPU only.

* gaussian_blur: a constant image stays constant (normalised taps); a unit
  impulse away from the edges gives the outer product of the taps
  exp(-i^2 / 2 sigma^2) / sum (the separable kernel in closed form); at an
  edge the clamped taps pile onto the border pixel; a (n, H, W) batch is
  each frame blurred alone; the numpy twin gives the same image; the blur's
  gradient is its transpose (<G x, y> = <x, G^T y>, checked with an explicit
  matrix on a small image);
* with_params: parameters that keep spin and accretion rate keep the
  conical jet's j_nu0 bit for bit; a new spin carries j_nu0 over at the
  scene's phi_flux^2 eps_rad (the ratio to a fresh normalisation at the new
  spin is the ratio at the old one).
"""
import math

import numpy as np
import pytest
import torch

from blackhole_amd import scene as S
from blackhole_amd.diff import gaussian_blur, gaussian_blur_np, with_params


def _taps(sigma=2.0):
    r = int(math.ceil(3.0 * sigma))
    i = np.arange(-r, r + 1)
    k = np.exp(-0.5 * i * i / (sigma * sigma))
    return k / k.sum(), r


def test_constant_image_is_unchanged():
    img = torch.full((20, 17), 3.25, dtype=torch.float64)
    out = gaussian_blur(img)
    assert torch.allclose(out, img, rtol=1e-15, atol=0)


def test_impulse_gives_the_separable_kernel():
    k, r = _taps()
    img = torch.zeros(31, 29, dtype=torch.float64)
    img[15, 14] = 1.0
    out = gaussian_blur(img).numpy()
    want = np.zeros((31, 29))
    want[15 - r:15 + r + 1, 14 - r:14 + r + 1] = np.outer(k, k)
    assert np.allclose(out, want, rtol=0, atol=1e-15 * want.max())
    assert out.sum() == pytest.approx(1.0, abs=1e-15)


def test_edge_taps_pile_onto_the_border():
    k, r = _taps()
    img = torch.zeros(1, 25, dtype=torch.float64)
    img[0, 0] = 1.0
    row = gaussian_blur(img).numpy()[0]
    # output j collects the taps whose clamped source is column 0: j + i <= 0
    want = np.array([k[:r - j + 1].sum() if j <= r else 0.0 for j in range(25)])
    assert np.allclose(row, want, rtol=0, atol=1e-15)


def test_batch_equals_frames_and_numpy_twin():
    g = torch.Generator().manual_seed(5)
    imgs = torch.rand(3, 24, 19, generator=g, dtype=torch.float64)
    out = gaussian_blur(imgs)
    for f in range(3):
        one = gaussian_blur(imgs[f])
        assert torch.allclose(out[f], one, rtol=1e-15, atol=0)
        assert np.allclose(gaussian_blur_np(imgs[f].numpy()), one.numpy(), rtol=1e-15, atol=0)


def test_gradient_is_the_transpose():
    H, W = 9, 11
    x = torch.rand(H, W, dtype=torch.float64, generator=torch.Generator().manual_seed(1), requires_grad=True)
    y = torch.rand(H, W, dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    (gaussian_blur(x) * y).sum().backward()
    G = np.zeros((H * W, H * W))
    for j in range(H * W):
        e = torch.zeros(H * W, dtype=torch.float64)
        e[j] = 1.0
        G[:, j] = gaussian_blur(e.view(H, W)).reshape(-1).numpy()
    assert np.allclose(x.grad.reshape(-1).numpy(), G.T @ y.reshape(-1).numpy(), rtol=1e-13, atol=1e-15)


def test_with_params_keeps_or_carries_the_jet():
    sc = S.quasar_frame(30.0, 16, 16, stars=False)
    j0 = sc.jet_j_nu0
    same = with_params(sc, (sc.spin, math.radians(50.0), sc.eddington, 1800.0))
    assert same.jet_j_nu0 == j0
    assert same.cam_theta == math.radians(50.0) and same.cam_r == 1800.0
    moved = with_params(sc, (0.9, sc.cam_theta, 0.25, sc.cam_r))
    fresh_old = S.quasar_frame(30.0, 16, 16, stars=False)
    S.setup_conical_jet(fresh_old, phi_flux=1.0, eps_rad=1.0)
    fresh_new = with_params(S.quasar_frame(30.0, 16, 16, stars=False), (0.9, sc.cam_theta, 0.25, sc.cam_r))
    S.setup_conical_jet(fresh_new, phi_flux=1.0, eps_rad=1.0)
    assert moved.jet_j_nu0 != j0
    assert moved.jet_j_nu0 / fresh_new.jet_j_nu0 == pytest.approx(j0 / fresh_old.jet_j_nu0, rel=1e-14)
