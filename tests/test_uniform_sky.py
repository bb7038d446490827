"""The synthetic sky of the 1024^2 frame (bhx_uniform_sky): 250k stars
uniform on the sphere, magnitudes from log10 N(<m) = 0.45 m + c up to m = 8,
B-V ~ N(0.6, 0.4) mapped to colour temperature, irradiance F0 10^(-0.4 m).
CPU only.

* the library's table and the oracle's restatement are bit for bit the same
  table (what the GPU and the CPU baseline both trace against);
* directions: unit vectors; z and the azimuth uniform (Kolmogorov-Smirnov),
  z's mean and variance 0 and 1/3 within 4 standard errors, equal octants
  (chi-square);
* magnitudes, recovered from the irradiance: on [-1.5, 8], KS against
  F(m) = (10^(0.45 (m + 1.5)) - 1) / (10^(0.45 x 9.5) - 1); an m = 0 star has
  the reference's F0 = 3.64e-9 W m^-2 um^-1 x 0.089 um (src/scene.h:100-101);
* B-V, recovered from the temperature by inverting Ballesteros' formula
  (src/spectrum.cpp:158-162, a quadratic): KS against N(0.6, 0.4) with the
  formula's clamp to [-0.4, 2.0], whose two atoms hold Phi(-2.5) and
  1 - Phi(3.5) of the stars within 4 standard errors.
"""
import math

import numpy as np
import pytest
from scipy import stats

import oracle
from blackhole_amd import scene as S

N, SEED = 250000, 20240921
F0 = 3.64e-9 * 0.089


@pytest.fixture(scope="module")
def sky():
    return S.uniform_sky(N, SEED)


def test_library_and_oracle_tables_are_identical(sky):
    for seed, n in ((SEED, N), (7, 1000), (123456789, 0)):
        a = S.uniform_sky(n, seed) if (seed, n) != (SEED, N) else sky
        b = oracle.uniform_sky(n, seed)
        for u, v in zip(a, b):
            assert u.shape == (n,) and np.array_equal(u, v)
    other = S.uniform_sky(1000, SEED + 1)
    assert not np.array_equal(other[0], sky[0][:1000])


def test_directions_uniform_on_the_sphere(sky):
    x, y, z = sky[:3]
    assert np.max(np.abs(x * x + y * y + z * z - 1.0)) < 1e-15
    assert stats.kstest(z, "uniform", args=(-1.0, 2.0)).pvalue > 1e-3
    phi = np.arctan2(y, x)
    assert stats.kstest(phi, "uniform", args=(-math.pi, 2.0 * math.pi)).pvalue > 1e-3
    assert abs(z.mean()) < 4.0 * math.sqrt(1.0 / 3.0 / N)
    assert abs(z.var() - 1.0 / 3.0) < 4.0 * math.sqrt((1.0 / 5.0 - 1.0 / 9.0) / N)
    octant = (x > 0).astype(int) + 2 * (y > 0) + 4 * (z > 0)
    counts = np.bincount(octant, minlength=8)
    assert stats.chisquare(counts).pvalue > 1e-3


def test_magnitudes_follow_the_count_law(sky):
    irr = sky[4]
    m = -2.5 * np.log10(irr / F0)
    assert m.min() >= -1.5 - 1e-12 and m.max() <= 8.0 + 1e-12
    span = 10.0 ** (0.45 * 9.5) - 1.0
    cdf = lambda v: (10.0 ** (0.45 * (np.asarray(v) + 1.5)) - 1.0) / span
    assert stats.kstest(m, cdf).pvalue > 1e-3
    # log10 N(<m) = 0.45 m + c: the count ratio over one magnitude is 10^0.45
    n6, n7 = np.count_nonzero(m < 6.0), np.count_nonzero(m < 7.0)
    assert n7 / n6 == pytest.approx(10 ** 0.45, rel=0.05)
    # F0: the irradiance of an m = 0 star is the reference's 3.24e-10 W m^-2
    assert F0 == pytest.approx(3.2396e-10, rel=1e-12)


def _bv_from_temperature(T):
    """Inverse of T = 4600 (1/(0.92 x + 1.7) + 1/(0.92 x + 0.62)): with
    a = 0.92 x and t = T / 4600, t a^2 + (2.32 t - 2) a + (1.054 t - 2.32) = 0."""
    t = T / 4600.0
    b, c = 2.32 * t - 2.0, 1.054 * t - 2.32
    a = (-b + np.sqrt(b * b - 4.0 * t * c)) / (2.0 * t)
    return a / 0.92


def test_colour_index_distribution(sky):
    T = sky[3]
    bv = _bv_from_temperature(T)
    assert np.all(bv > -0.4 - 1e-9) and np.all(bv < 2.0 + 1e-9)
    lo, hi = np.abs(bv + 0.4) < 1e-9, np.abs(bv - 2.0) < 1e-9
    p_lo, p_hi = stats.norm.cdf(-2.5), stats.norm.sf(3.5)
    assert abs(lo.mean() - p_lo) < 4.0 * math.sqrt(p_lo * (1 - p_lo) / N)
    assert abs(hi.mean() - p_hi) < 4.0 * math.sqrt(p_hi * (1 - p_hi) / N) + 1.0 / N
    inner = bv[~lo & ~hi]
    # N(0.6, 0.4) restricted to the open interval (-0.4, 2.0)
    a0, b0 = stats.norm.cdf(-2.5), stats.norm.cdf(3.5)
    cdf = lambda v: (stats.norm.cdf(np.asarray(v), 0.6, 0.4) - a0) / (b0 - a0)
    assert stats.kstest(inner, cdf).pvalue > 1e-3
    assert inner.mean() == pytest.approx(stats.truncnorm.mean(-2.5, 3.5, 0.6, 0.4), abs=4 * 0.4 / math.sqrt(N))
